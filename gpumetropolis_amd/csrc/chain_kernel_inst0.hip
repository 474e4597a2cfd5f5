// chain_kernel_inst0.hip -- instantiates the interpreter's kernels of the
// variants (0, 16, 1), (0, 16, 2) (one group per file, so the library's
// translation units compile in parallel).
#include "chain_kernels.h"

namespace gpum {

GPUM_INSTANTIATE(0, 16, 1)
GPUM_INSTANTIATE(0, 16, 2)

}  // namespace gpum

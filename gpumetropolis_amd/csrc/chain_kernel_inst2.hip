// chain_kernel_inst2.hip -- instantiates the interpreter's kernels of the
// variants (1, 16, 1), (3, 16, 1) (one group per file, so the library's
// translation units compile in parallel).
#include "chain_kernels.h"

namespace gpum {

GPUM_INSTANTIATE(1, 16, 1)
GPUM_INSTANTIATE(3, 16, 1)

}  // namespace gpum

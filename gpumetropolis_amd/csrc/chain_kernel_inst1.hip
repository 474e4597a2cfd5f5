// chain_kernel_inst1.hip -- instantiates the interpreter's kernels of the
// variants (0, 8, 1), (0, 8, 2), (0, 4, 1), (0, 4, 2) (one group per file, so the library's
// translation units compile in parallel).
#include "chain_kernels.h"

namespace gpum {

GPUM_INSTANTIATE(0, 8, 1)
GPUM_INSTANTIATE(0, 8, 2)
GPUM_INSTANTIATE(0, 4, 1)
GPUM_INSTANTIATE(0, 4, 2)

}  // namespace gpum

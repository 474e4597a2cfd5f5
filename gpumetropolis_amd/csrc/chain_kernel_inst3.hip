// chain_kernel_inst3.hip -- instantiates the interpreter's kernels of the
// variants (7, 8, 1), (15, 4, 1), (31, 4, 1) (one group per file, so the library's
// translation units compile in parallel).
#include "chain_kernels.h"

namespace gpum {

GPUM_INSTANTIATE(7, 8, 1)
GPUM_INSTANTIATE(15, 4, 1)
GPUM_INSTANTIATE(31, 4, 1)

}  // namespace gpum

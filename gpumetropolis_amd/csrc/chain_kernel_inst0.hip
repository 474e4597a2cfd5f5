// chain_kernel_inst0.hip -- instantiates the affine-program interpreter's
// kernels (VmAffineEval; 16 observations per lane, 1, 2 and 4 chains per
// wave; one group per file, so the library's translation units compile in
// parallel).
#include "chain_kernels.h"

namespace gpum {

GPUM_AFFINE_VARIANTS(GPUM_AFFINE_INSTANTIATE)

}  // namespace gpum

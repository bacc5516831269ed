// Instantiation of the fused batched Newton kernels for padded size 40.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(40, 64)
}  // namespace vxm

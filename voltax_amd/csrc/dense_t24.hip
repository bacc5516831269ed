// Instantiation of the fused batched Newton kernels for padded size 24.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(24, 32)
}  // namespace vxm

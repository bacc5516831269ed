// Instantiation of the fused batched Newton kernels for padded size 56.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(56, 64)
}  // namespace vxm

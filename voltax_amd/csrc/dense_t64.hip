// Instantiation of the fused batched Newton kernels for padded size 64.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(64, 64)
}  // namespace vxm

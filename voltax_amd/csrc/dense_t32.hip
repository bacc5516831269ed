// Instantiation of the fused batched Newton kernels for padded size 32.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(32, 32)
}  // namespace vxm

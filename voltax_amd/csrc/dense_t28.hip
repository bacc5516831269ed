// Instantiation of the fused batched Newton kernels for padded size 28.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(28, 32)
}  // namespace vxm

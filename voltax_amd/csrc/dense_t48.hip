// Instantiation of the fused batched Newton kernels for padded size 48.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(48, 64)
}  // namespace vxm

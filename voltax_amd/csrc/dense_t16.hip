// Instantiation of the fused batched Newton kernels for padded size 16.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(16, 16)
}  // namespace vxm

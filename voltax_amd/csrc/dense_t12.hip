// Instantiation of the fused batched Newton kernels for padded size 12.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(12, 16)
}  // namespace vxm

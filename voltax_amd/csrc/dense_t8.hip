// Instantiation of the fused batched Newton kernels for padded size 8.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(8, 8)
}  // namespace vxm

// Instantiation of the fused batched Newton kernels for padded size 20.
#include "dense_newton.hpp"

namespace vxm {
VXM_DENSE_CLASS_DECL(20, 32)
}  // namespace vxm

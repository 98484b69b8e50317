// sx_kernels_p2.hip -- element-kernel instantiations for H1 order 2
// (3 nodes and 3 Gauss points per direction).
#include "sx_element.hpp"

namespace sxk {
int launch_element_p2(const ElemLaunch& P, hipError_t* err) { return dispatch_order<3>(P, err); }
}  // namespace sxk

// sx_kernels_p1.hip -- element-kernel instantiations for H1 order 1
// (2 nodes and 2 Gauss points per direction).
#include "sx_element.hpp"

namespace sxk {
int launch_element_p1(const ElemLaunch& P, hipError_t* err) { return dispatch_order<2>(P, err); }
}  // namespace sxk

// sx_kernels_p3.hip -- element-kernel instantiations for H1 order 3
// (4 nodes and 4 Gauss points per direction).
#include "sx_element.hpp"

namespace sxk {
int launch_element_p3(const ElemLaunch& P, hipError_t* err) { return dispatch_order<4>(P, err); }
}  // namespace sxk

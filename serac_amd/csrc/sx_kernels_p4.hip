// sx_kernels_p4.hip -- element-kernel instantiations for H1 order 4
// (5 nodes and 5 Gauss points per direction).
#include "sx_element.hpp"

namespace sxk {
int launch_element_p4(const ElemLaunch& P, hipError_t* err) { return dispatch_order<5>(P, err); }
}  // namespace sxk

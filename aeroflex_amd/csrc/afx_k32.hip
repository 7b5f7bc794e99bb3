// FP32 instantiation of the residual kernels (throughput form; built with
// FMA contraction allowed).
#define AFX_FP32_UNIT
#include "afx_device.hpp"

namespace afx {
AFX_INSTANTIATE_LAUNCHERS(float)
}  // namespace afx

// Shared between the C-ABI units of the vortex-lattice / beam path
// (afx_vlm.hip: device side; afx_structure.cpp: lattice generator, beam,
// load and displacement transfer, coupling loop).
#pragma once
#include <string>

namespace afx {
// sets the thread-local message returned by afx_last_error and returns code
int api_fail(int code, const std::string& msg);
}  // namespace afx

// Solution file of the reference's post-processing (save, post.h:58-180).
#pragma once
#include <string>

#include "afx_mesh.hpp"

namespace afx {

// Writes the reference's ASCII VTU (UnstructuredGrid) of the real cells:
// nodes, connectivity, offsets, types, then per cell the wall distance, Mach,
// density, pressure, temperature and velocity of q (q(4*i+k), reference cell
// order) with the reference's number formatting.  The text is formatted on
// up to `threads` threads (0: the hardware's) and written in order.
// Returns an empty string on success, else the error.
std::string write_vtu(const std::string& path, const Mesh& m, const double* q, const double* wall_dist, double gamma,
                      double R, unsigned threads = 0);

}  // namespace afx

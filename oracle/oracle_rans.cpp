// ============================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the AeroFLEX unstructured finite-volume residual
// (src/rans/include/rans/*.h).  Every function cites the reference lines it
// follows; the arithmetic keeps the reference's operation order (compile with
// -ffp-contract=off, see oracle/build_oracle.py) so that it reproduces the
// reference's FP64 results, quirks included.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker / CPU baseline -- never as the thing
// that is measured or shipped.
//
// Parity status (what pins each part; the reference itself cannot be built
// here because it needs Eigen 3, which this image lacks):
//   mesh reader / connectivity / metrics  pinned: file counts, closed-cell
//                                         normal sums, shoelace areas, hand
//                                         cases (tests/test_oracle_pins.py)
//   Roe flux, BC states                   pinned: consistency F(q,q)=Euler
//                                         flux, rotation/mirror invariance,
//                                         freestream preservation
//   GG gradient                           pinned: hand-computed meshes, exact
//                                         linear reproduction on consistently
//                                         owned patches
//   LSQ gradient                          pinned: -1/2 x exact gradient of a
//                                         linear field (interior cells)
//   limiter, dt                           pinned: scalar closed forms
//   wall profile (CL/CD/CM)               pinned: uniform cp -> 0, closed form
//                                         on a polygon
//   FD convective Jacobian                pinned: vs analytic-free central FD
//   SA source terms (own, no reference)   pinned: hand-evaluated terms and
//                                         nu~=0 reduction to laminar
//   SA implicit matrix (own)              pinned: its face function equals the
//                                         residual's first-order flux, blocks
//                                         and source row vs central FD, the
//                                         assembly vs a numpy assembly of the
//                                         blocks (tests/test_oracle_sa_implicit.py)
//   get_uniform_residual (stale qW, A8)   pinned: slip-wall channel closed
//                                         form 2 rho v sabs(c~) |L|; second
//                                         call doubles it
// ============================================================================
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>
#include <algorithm>

namespace orc {

typedef uint32_t u32;
static const u32 EDGE_NULL = 4294967295u;  // mesh.h:35

// ---------------------------------------------------------------- gas / bc --
// core.h:27-46
struct Gas {
    double R = 0.71428571428;
    double mu_L = 1e-5;
    double Pr_L = 0.72;
    double Pr_T = 0.9;
    double cp = 1.;
    double gamma = 1.4;
    double k() const { return cp * mu_L / Pr_L; }  // core.h:43-45
};

// core.h:61-84
struct BVars {
    double mach = 0.2, angle = 0, T = 1, p = 1.;
    void conservative(const Gas& g, double out[4]) const {
        double c = std::sqrt(g.gamma * g.R * T);
        double u = mach * c * std::cos(angle);
        double v = mach * c * std::sin(angle);
        double rho = p / (g.R * T);
        double rhoE = p / (g.gamma - 1) + 0.5 * rho * (u * u + v * v);
        out[0] = rho; out[1] = rho * u; out[2] = rho * v; out[3] = rhoE;
    }
};

// ------------------------------------------------------------- string utils --
// mesh.h:39-109 (cut_str / str_to_ints / str_to_floats): tokens are the
// substrings between single spaces; a trailing space adds no token.
static std::vector<std::string> split_sp(const std::string& s) {
    std::vector<std::string> out;
    std::vector<size_t> I;
    for (size_t i = 0; i < s.size(); ++i) if (s[i] == ' ') I.push_back(i);
    if (I.empty()) { out.push_back(s); return out; }
    out.push_back(s.substr(0, I[0]));
    for (size_t j = 0; j + 1 < I.size(); ++j) out.push_back(s.substr(I[j] + 1, I[j + 1] - I[j] - 1));
    if (I.back() != s.size() - 1) out.push_back(s.substr(I.back() + 1));
    return out;
}
static std::vector<u32> to_ints(const std::string& s) {
    std::vector<u32> o;
    for (auto& t : split_sp(s)) o.push_back((u32)std::stoi(t));
    return o;
}
static std::vector<double> to_floats(const std::string& s) {
    std::vector<double> o;
    for (auto& t : split_sp(s)) o.push_back(std::stod(t));
    return o;
}

// --------------------------------------------------------------------- mesh --
// mesh.h:198-273.  Arrays keep the reference names and layouts.
struct Mesh {
    std::map<u32, std::string> physicalNames;
    std::map<u32, u32> entityTagToPhysicalTag;
    std::vector<double> nodesX, nodesY;
    std::map<u32, u32> originalNodesRef;

    std::vector<u32> edgesNodes;  // 2 per face
    std::vector<u32> edgesCells;  // 2 per face
    std::vector<double> edgesLengths, edgesNormalsX, edgesNormalsY, edgesCentersX, edgesCentersY;

    std::vector<u32> boundaryEdges, boundaryEdges0, boundaryEdges1, boundaryEdgesIntTag;
    std::vector<std::string> boundaryEdgesPhysicals;

    std::map<std::tuple<u32, u32>, u32> edgesRef;

    std::vector<u32> cellsNodes;  // 4 per cell
    std::vector<u32> cellsEdges;  // 4 per real cell
    std::vector<char> cellsIsTriangle;
    std::vector<double> cellsAreas, cellsCentersX, cellsCentersY;
    std::vector<double> wall_dist;
    u32 nRealCells = 0;

    u32 nFaces() const { return (u32)edgesLengths.size(); }
    u32 nCells() const { return (u32)cellsAreas.size(); }

    // mesh.h:304-313
    bool edge_exists(u32 n0, u32 n1) const {
        return edgesRef.count(std::make_tuple(std::min(n0, n1), std::max(n0, n1))) > 0;
    }
    // mesh.h:290-301
    u32 find_edge(u32 n0, u32 n1) const {
        auto it = edgesRef.find(std::make_tuple(std::min(n0, n1), std::max(n0, n1)));
        return it == edgesRef.end() ? EDGE_NULL : it->second;
    }
    // mesh.h:317-341 -- faces are created on first sight, owner = creating cell
    void add_cell_edges(u32 c) {
        u32 size = cellsIsTriangle[c] ? 3 : 4;
        for (u32 i = 0; i < size; ++i) {
            u32 j = (i < size - 1) ? i + 1 : 0;
            u32 a = cellsNodes[4 * c + i], b = cellsNodes[4 * c + j];
            if (!edge_exists(a, b)) {
                edgesNormalsX.push_back(0.); edgesNormalsY.push_back(0.);
                edgesCentersX.push_back(0.); edgesCentersY.push_back(0.);
                edgesCells.push_back(c); edgesCells.push_back(0);
                edgesNodes.push_back(a); edgesNodes.push_back(b);
                edgesLengths.push_back(0.);
                edgesRef.insert({std::make_tuple(std::min(a, b), std::max(a, b)), (u32)edgesLengths.size() - 1});
            }
        }
    }
    // mesh.h:344-375
    void convert_node_face_info() {
        u32 ce[4];
        for (u32 i = 0; i < nCells(); ++i) {
            const u32* cn = &cellsNodes[4 * i];
            ce[0] = find_edge(cn[0], cn[1]);
            ce[1] = find_edge(cn[1], cn[2]);
            if (cellsIsTriangle[i]) { ce[2] = find_edge(cn[2], cn[0]); ce[3] = EDGE_NULL; }
            else { ce[2] = find_edge(cn[2], cn[3]); ce[3] = find_edge(cn[3], cn[0]); }
            for (int k = 0; k < 4; ++k) cellsEdges.push_back(ce[k]);
            for (int k = 0; k < 4; ++k) {
                u32 e = ce[k];
                if (e != EDGE_NULL && i != edgesCells[2 * e]) edgesCells[2 * e + 1] = i;
            }
        }
    }
    // mesh.h:378-453 (A11: vertex-average centroid; quad area = two triangles)
    void compute_mesh() {
        for (u32 i = 0; i < nCells(); ++i) {
            double cx = 0., cy = 0.;
            u32 n = cellsIsTriangle[i] ? 3 : 4;
            for (u32 j = 0; j < n; ++j) {
                cx += nodesX[cellsNodes[4 * i + j]] / ((double)n);
                cy += nodesY[cellsNodes[4 * i + j]] / ((double)n);
            }
            cellsCentersX[i] = cx; cellsCentersY[i] = cy;
        }
        for (u32 i = 0; i < nFaces(); ++i) {
            u32 a = edgesNodes[2 * i], b = edgesNodes[2 * i + 1];
            edgesCentersX[i] = (nodesX[b] + nodesX[a]) * 0.5;
            edgesCentersY[i] = (nodesY[b] + nodesY[a]) * 0.5;
            double dex = nodesX[b] - nodesX[a];
            double dey = nodesY[b] - nodesY[a];
            double l = std::sqrt(dex * dex + dey * dey);
            edgesLengths[i] = l;
            edgesNormalsX[i] = -dey / l;
            edgesNormalsY[i] = dex / l;
            u32 o = edgesCells[2 * i];
            double dot = edgesNormalsX[i] * (edgesCentersX[i] - cellsCentersX[o])
                       + edgesNormalsY[i] * (edgesCentersY[i] - cellsCentersY[o]);
            if (dot < 0) { edgesNormalsX[i] *= -1.; edgesNormalsY[i] *= -1.; }
        }
        for (u32 i = 0; i < nCells(); ++i) {
            const u32* cn = &cellsNodes[4 * i];
            double x1 = nodesX[cn[0]], x2 = nodesX[cn[1]], x3 = nodesX[cn[2]];
            double y1 = nodesY[cn[0]], y2 = nodesY[cn[1]], y3 = nodesY[cn[2]];
            if (cellsIsTriangle[i]) {
                cellsAreas[i] = 0.5 * std::fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
            } else {
                double x4 = nodesX[cn[3]], y4 = nodesY[cn[3]];
                cellsAreas[i] = 0.5 * std::fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2))
                              + 0.5 * std::fabs(x1 * (y3 - y4) + x3 * (y4 - y1) + x4 * (y1 - y3));
            }
        }
    }
    // mesh.h:744-787 (A12: mirrored ghost centroid, owner's area)
    void add_boundary_cells() {
        for (size_t i = 0; i < boundaryEdges0.size(); ++i) {
            u32 nmin = std::min(boundaryEdges0[i], boundaryEdges1[i]);
            u32 nmax = std::max(boundaryEdges0[i], boundaryEdges1[i]);
            auto it = edgesRef.find(std::make_tuple(nmin, nmax));
            if (it == edgesRef.end())
                throw std::invalid_argument("invalid edge ref (" + std::to_string(nmin) + ", " + std::to_string(nmax) + ")");
            u32 e = it->second;
            u32 cell = edgesCells[2 * e];
            double area = cellsAreas[cell];
            double dx = edgesCentersX[e] - cellsCentersX[cell];
            double dy = edgesCentersY[e] - cellsCentersY[cell];
            double dist = std::sqrt(dx * dx + dy * dy);
            double cx = edgesCentersX[e] + dist * edgesNormalsX[e];
            double cy = edgesCentersY[e] + dist * edgesNormalsY[e];
            cellsAreas.push_back(area);
            cellsCentersX.push_back(cx);
            cellsCentersY.push_back(cy);
            cellsIsTriangle.push_back(1);
            cellsNodes.push_back(nmin); cellsNodes.push_back(nmax);
            cellsNodes.push_back(0); cellsNodes.push_back(0);
            edgesCells[2 * e + 1] = nCells() - 1;
            boundaryEdges.push_back(e);
            boundaryEdgesPhysicals.push_back(physicalNames.at(boundaryEdgesIntTag[i]));
        }
    }
    // mesh.h:834-884 after the file has been parsed
    void finish() {
        for (u32 i = 0; i < nCells(); ++i) add_cell_edges(i);
        nRealCells = nCells();
        convert_node_face_info();
        compute_mesh();
        add_boundary_cells();
        wall_dist.assign(nCells(), 0.0);
    }

    // mesh.h:457-520 (read_entities)
    void read_entities(const std::string& fn) {
        std::ifstream in(fn);
        if (!in) throw std::runtime_error("cannot open " + fn);
        std::string line, section;
        u32 ns = 0, nss = 0;
        u32 npts = 0, ncur = 0;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (!line.empty() && line[0] == '$') { section = line.substr(1); ns = 0; }
            else if (section == "PhysicalNames") {
                if (ns > 1) {
                    auto t = split_sp(line);
                    std::string nm = t[2];
                    nm.erase(std::remove(nm.begin(), nm.end(), '"'), nm.end());
                    physicalNames[(u32)std::stoi(t[1]) + 1000u * (u32)std::stoi(t[0])] = nm;
                }
            } else if (section == "Entities") {
                if (ns == 1) {
                    auto l = to_ints(line);
                    npts = l[0]; ncur = l[1] + l[0];
                    nss = 0;
                } else if (nss <= npts) {
                } else if (nss <= ncur) {
                    auto l = to_ints(line);
                    entityTagToPhysicalTag[l[0]] = l[8] + 1000u;
                }
            } else if (section == "Nodes") break;
            ns += 1; nss += 1;
        }
    }
    // mesh.h:524-581 (read_nodes)
    void read_nodes(const std::string& fn) {
        std::ifstream in(fn);
        std::string line, section;
        u32 ns = 0, nss = 0, start = 0, nblk = 0;
        std::vector<u32> tags;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (!line.empty() && line[0] == '$') { section = line.substr(1); ns = 0; }
            else if (section == "Nodes") {
                if (ns == 1) { start = 1; nss = 0; }
                else if (ns > start + 2 * nblk) {
                    auto l = to_ints(line);
                    nblk = l[3]; start = ns; tags.resize(nblk); nss = 0;
                } else if (nss <= nblk) {
                    tags[nss - 1] = (u32)std::stoi(line) - 1;
                } else {
                    u32 t = tags[nss - 1 - nblk];
                    auto l = to_floats(line);
                    originalNodesRef[t] = (u32)nodesX.size();
                    nodesX.push_back(l[0]); nodesY.push_back(l[1]);
                }
            } else if (section == "Elements") break;
            ns += 1; nss += 1;
        }
    }
    // mesh.h:586-738 (read_boundaries + read_elements fused: same block walk)
    void read_elements(const std::string& fn) {
        std::ifstream in(fn);
        std::string line, section;
        u32 ns = 0, start = 0, nblk = 0, bdim = 0, btag = 0;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (!line.empty() && line[0] == '$') { section = line.substr(1); ns = 0; }
            else if (section == "Elements") {
                if (ns == 1) { start = 1; nblk = 0; }
                else if (ns > start + nblk) {
                    auto l = to_ints(line);
                    bdim = l[0];
                    if (bdim == 1) btag = entityTagToPhysicalTag.at(l[1]);
                    nblk = l[3]; start = ns;
                } else if (bdim == 1) {
                    auto li = to_ints(line);
                    if (li.size() == 3) {
                        boundaryEdges0.push_back(originalNodesRef.at(li[1] - 1));
                        boundaryEdges1.push_back(originalNodesRef.at(li[2] - 1));
                        boundaryEdgesIntTag.push_back(btag);
                    }
                } else if (bdim == 2) {
                    auto li = to_ints(line);
                    if (li.size() - 1 != 2) {
                        std::vector<u32> l;
                        for (size_t i = 1; i < li.size(); ++i) l.push_back(originalNodesRef.at(li[i] - 1));
                        if (l.size() == 3) { l.push_back(0); cellsIsTriangle.push_back(1); }
                        else cellsIsTriangle.push_back(0);
                        if (l.size() != 4) throw std::invalid_argument("cell with unsupported node count");
                        for (u32 v : l) cellsNodes.push_back(v);
                        cellsAreas.push_back(0.); cellsCentersX.push_back(0.); cellsCentersY.push_back(0.);
                    }
                }
            }
            ns += 1;
        }
    }
    void read_file(const std::string& fn) {
        read_entities(fn);
        read_nodes(fn);
        read_elements(fn);
        finish();
    }
};

// --------------------------------------------------------------- physics ----
enum { BC_INTERNAL = 0, BC_FARFIELD = 1, BC_SLIP = 2, BC_WALL = 3 };

struct FaceFlux {
    const Gas* g;
    double nx, ny;
    int kind;      // BC_*
    int viscous;   // 0 inviscid, 1 laminar (solver.h:203-208)
    bool two_sided() const { return kind == BC_INTERNAL; }
};

// physics.h:84-86
static inline double sabs(double x) { return std::sqrt(x * x + 1e-4); }
// physics.h:133-135
static inline double entropy_correction(double l, double d) { return l > d ? l : (l * l + d * d) / (2 * d); }

// physics.h:48-53
static inline double calc_p(const double* q, const Gas& g) {
    return (g.gamma - 1) * (q[3] - 0.5 / q[0] * (q[1] * q[1] + q[2] * q[2]));
}
// physics.h:33-46
static inline void calc_grad_p(double gp[2], const double* q, const double* gx, const double* gy, const Gas& g) {
    gp[0] = 2. * gx[3];
    gp[0] -= gx[1] * q[1] / q[0] + q[1] * (gx[1] * q[0] - gx[0] * q[1]) / (q[0] * q[0]);
    gp[0] -= gx[2] * q[2] / q[0] + q[2] * (gx[2] * q[0] - gx[0] * q[2]) / (q[0] * q[0]);
    gp[0] *= 0.5 * (g.gamma - 1);
    gp[1] = 2. * gy[3];
    gp[1] -= gy[1] * q[1] / q[0] + q[1] * (gy[1] * q[0] - gy[0] * q[1]) / (q[0] * q[0]);
    gp[1] -= gy[2] * q[2] / q[0] + q[2] * (gy[2] * q[0] - gy[0] * q[2]) / (q[0] * q[0]);
    gp[1] *= 0.5 * (g.gamma - 1);
}
// physics.h:55-68
static inline void calc_grad_temp(double gt[2], const double* q, const double* gx, const double* gy, const Gas& g) {
    double gp[2];
    const double p = calc_p(q, g);
    calc_grad_p(gp, q, gx, gy, g);
    gt[0] = (1. / g.R) * ((gp[0] * q[0] - gx[0] * p) / (q[0] * q[0]));
    gt[1] = (1. / g.R) * ((gp[1] * q[0] - gy[0] * p) / (q[0] * q[0]));
}

// physics.h:160-264 (internal_flux::operator()): Roe + Harten fix + laminar
static void roe_flux(const FaceFlux& F, const double* q_L, const double* q_R,
                     const double* gx, const double* gy, double f[4]) {
    const Gas& g = *F.g;
    const double nx = F.nx, ny = F.ny;
    const double rho_L = q_L[0], rho_u_L = q_L[1], rho_v_L = q_L[2], rho_e_L = q_L[3];
    const double rho_R = q_R[0], rho_u_R = q_R[1], rho_v_R = q_R[2], rho_e_R = q_R[3];
    const double V_L = (rho_u_L * nx + rho_v_L * ny) / rho_L;
    const double V_R = (rho_u_R * nx + rho_v_R * ny) / rho_R;
    const double p_L = (g.gamma - 1) * (rho_e_L - 0.5 / rho_L * (rho_u_L * rho_u_L + rho_v_L * rho_v_L));
    const double p_R = (g.gamma - 1) * (rho_e_R - 0.5 / rho_R * (rho_u_R * rho_u_R + rho_v_R * rho_v_R));
    f[0] = (V_L * rho_L + V_R * rho_R) * 0.5;
    f[1] = (V_L * rho_u_L + p_L * nx + V_R * rho_u_R + p_R * nx) * 0.5;
    f[2] = (V_L * rho_v_L + p_L * ny + V_R * rho_v_R + p_R * ny) * 0.5;
    f[3] = (V_L * (rho_e_L + p_L) + V_R * (rho_e_R + p_R)) * 0.5;
    const double pL = p_L, pR = p_R;
    const double uL = q_L[1] / q_L[0];
    const double uR = q_R[1] / q_R[0];
    const double vL = q_L[2] / q_L[0];
    const double vR = q_R[2] / q_R[0];
    const double srhoL = std::sqrt(q_L[0]);
    const double srhoR = std::sqrt(q_R[0]);
    const double rho = srhoR * srhoL;
    const double u = (uL * srhoL + uR * srhoR) / (srhoL + srhoR);
    const double v = (vL * srhoL + vR * srhoR) / (srhoL + srhoR);
    const double h = ((q_L[3] + pL) / q_L[0] * srhoL + (q_R[3] + pR) / q_R[0] * srhoR) / (srhoL + srhoR);
    const double q2 = u * u + v * v;
    const double c = std::sqrt((g.gamma - 1.) * (h - 0.5 * q2));
    const double V = u * nx + v * ny;
    const double VR = uR * nx + vR * ny;
    const double VL = uL * nx + vL * ny;
    const double lambda_cm = entropy_correction(sabs(V - c), 0.05 * c);
    const double lambda_c = entropy_correction(sabs(V), 0.05 * c);
    const double lambda_cp = entropy_correction(sabs(V + c), 0.05 * c);
    const double kF1 = lambda_cm * ((pR - pL) - rho * c * (VR - VL)) / (2. * c * c);
    const double kF234_0 = lambda_c * ((q_R[0] - q_L[0]) - (pR - pL) / (c * c));
    const double kF234_1 = lambda_c * rho;
    const double kF5 = lambda_cp * ((pR - pL) + rho * c * (VR - VL)) / (2 * c * c);
    f[0] -= 0.5 * (kF1 + kF234_0 + kF5);
    f[1] -= 0.5 * (kF1 * (u - c * nx) + kF234_0 * u + kF234_1 * (uR - uL - (VR - VL) * nx) + kF5 * (u + c * nx));
    f[2] -= 0.5 * (kF1 * (v - c * ny) + kF234_0 * v + kF234_1 * (vR - vL - (VR - VL) * ny) + kF5 * (v + c * ny));
    f[3] -= 0.5 * (kF1 * (h - c * V) + kF234_0 * q2 * 0.5 + kF234_1 * (u * (uR - uL) + v * (vR - vL) - V * (VR - VL)) + kF5 * (h + c * V));

    if (F.viscous == 1) {
        double qc[4];
        for (int k = 0; k < 4; ++k) qc[k] = 0.5 * (q_L[k] + q_R[k]);
        double gradT[2], gradu[2], gradv[2];
        calc_grad_temp(gradT, qc, gx, gy, g);
        gradu[0] = (qc[0] * gx[1] - qc[1] * gx[0]) / (qc[0] * qc[0]);   // physics.h:70-75
        gradu[1] = (qc[0] * gy[1] - qc[1] * gy[0]) / (qc[0] * qc[0]);
        gradv[0] = (qc[0] * gx[2] - qc[2] * gx[0]) / (qc[0] * qc[0]);   // physics.h:77-82
        gradv[1] = (qc[0] * gy[2] - qc[2] * gy[0]) / (qc[0] * qc[0]);
        const double div_v = gradu[0] + gradv[1];
        const double tau_xx = 2. * g.mu_L * (gradu[0] - div_v / 3.);
        const double tau_yy = 2. * g.mu_L * (gradv[1] - div_v / 3.);
        const double tau_xy = g.mu_L * (gradu[1] + gradv[0]);
        double phi0 = qc[1] / qc[0] * tau_xx + qc[2] / qc[0] * tau_xy + g.k() * gradT[0];
        double phi1 = qc[1] / qc[0] * tau_xy + qc[2] / qc[0] * tau_yy + g.k() * gradT[1];
        f[1] -= nx * tau_xx + ny * tau_xy;
        f[2] -= nx * tau_xy + ny * tau_yy;
        f[3] -= nx * phi0 + ny * phi1;
    }
}

// physics.h:267-276, 311-339, 377-405, 446-530 (vars() of each functor)
static void face_vars(const FaceFlux& F, const double* q_L, const double* q_bc, double q_R[4]) {
    const Gas& g = *F.g;
    const double nx = F.nx, ny = F.ny;
    if (F.kind == BC_INTERNAL) {
        for (int k = 0; k < 4; ++k) q_R[k] = (q_L[k] + q_bc[k]) * 0.5;   // A4
    } else if (F.kind == BC_SLIP) {
        const double rhoV = q_L[1] * nx + q_L[2] * ny;
        q_R[0] = q_L[0];
        q_R[1] = q_L[1] - 2. * rhoV * nx;
        q_R[2] = q_L[2] - 2. * rhoV * ny;
        q_R[3] = q_L[3];
    } else if (F.kind == BC_WALL) {
        q_R[0] = q_L[0]; q_R[1] = -q_L[1]; q_R[2] = -q_L[2]; q_R[3] = q_L[3];
    } else {
        const double rho = q_L[0], rho_u = q_L[1], rho_v = q_L[2], rho_e = q_L[3];
        const double bc_rho = q_bc[0], bc_rhou = q_bc[1], bc_rhov = q_bc[2], bc_rhoe = q_bc[3];
        const double bc_u = bc_rhou / bc_rho;
        const double bc_v = bc_rhov / bc_rho;
        const double bc_p = (g.gamma - 1) * (bc_rhoe - 0.5 / bc_rho * (bc_rhou * bc_rhou + bc_rhov * bc_rhov));
        const double p = (g.gamma - 1) * (rho_e - 0.5 / rho * (rho_u * rho_u + rho_v * rho_v));
        const double c = std::sqrt(g.gamma * p / rho);
        const double mach = std::sqrt(rho_u * rho_u + rho_v * rho_v) / (rho * c);
        const double io = rho_u * nx + rho_v * ny;
        if (mach > 1) {
            if (io < 0) {
                q_R[0] = bc_rho;
                q_R[1] = bc_rho * bc_u;
                q_R[2] = bc_rho * bc_v;
                q_R[3] = bc_p / (g.gamma - 1) + 0.5 * bc_rho * (bc_u * bc_u + bc_v * bc_v);
            } else {
                q_R[0] = rho; q_R[1] = rho_u; q_R[2] = rho_v; q_R[3] = rho_e;
            }
        } else {
            double pa = bc_p, rhoa = bc_rho, ua = bc_u, va = bc_v;
            double pd = p, rhod = rho, ud = rho_u / rho, vd = rho_v / rho;
            double rho0 = rho, c0 = c;
            if (io < 0) {
                double pb = 0.5 * (pa + pd - rho0 * c0 * (nx * (ua - ud) + ny * (va - vd)));
                q_R[0] = rhoa + (pb - pa) / (c0 * c0);
                q_R[1] = q_R[0] * (ua - nx * (pa - pb) / (rho0 * c0));
                q_R[2] = q_R[0] * (va - ny * (pa - pb) / (rho0 * c0));
                q_R[3] = pb / (g.gamma - 1) + 0.5 / q_R[0] * (q_R[1] * q_R[1] + q_R[2] * q_R[2]);
            } else {
                double pb = pa;
                q_R[0] = rhod + (pb - pd) / (c0 * c0);
                q_R[1] = q_R[0] * (ud + nx * (pd - pb) / (rho0 * c0));
                q_R[2] = q_R[0] * (va + ny * (pd - pb) / (rho0 * c0));   // A5: va
                q_R[3] = pb / (g.gamma - 1) + 0.5 / q_R[0] * (q_R[1] * q_R[1] + q_R[2] * q_R[2]);
            }
        }
    }
}

// operator() of each functor: internal -> Roe with gradients; BC functors ->
// Roe on (q_L, vars(q_L, q_bc)) WITHOUT gradients (physics.h:287-299,352-364,417-432; A6)
static void face_flux(const FaceFlux& F, const double* q_L, const double* q_R,
                      const double* gx, const double* gy, double f[4]) {
    static const double zero[4] = {0, 0, 0, 0};
    if (F.kind == BC_INTERNAL) { roe_flux(F, q_L, q_R, gx, gy, f); return; }
    double qb[4];
    face_vars(F, q_L, q_R, qb);
    FaceFlux inv = F; inv.kind = BC_INTERNAL;
    roe_flux(inv, q_L, qb, zero, zero, f);
}

// one (cell, face, variable) limiter value of calc_limiters (solver.h:545-584):
// the default build (solver.h:575-583) or RANS_MICHALAK_LIMITER
// (solver.h:557-573 with michalak_limiter, physics.h:580-592, RANS_YT 2)
static double limiter_value(bool michalak, double dqg, double dmax, double dmin, double K3a) {
    double lim = 1.0;
    if (michalak) {
        const double dMaxMin2 = (dmax - dmin) * (dmax - dmin);
        double sig;
        if (dMaxMin2 <= K3a) {
            sig = 1.;
        } else if (dMaxMin2 <= 2 * K3a) {
            const double y = (dMaxMin2 / K3a - 1.0);
            sig = 2.0 * y * y * y - 3.0 * y * y + 1.0;
        } else {
            sig = 0.;
        }
        if (sig < 1.0) {
            constexpr double YT = 2.0;
            constexpr double a = 1.0 / (YT * YT) - 2.0 / (YT * YT * YT);
            constexpr double b = -3.0 / 2.0 * a * YT - 0.5 / YT;
            auto mich = [&](double y) { return y >= YT ? 1.0 : a * y * y * y + b * y * y + y; };
            if (dqg > 1e-14) lim = mich(dmax / dqg);
            else if (dqg < -1e-14) lim = mich(dmin / dqg);
            else lim = 1.0;
        }
        return sig + (1.0 - sig) * lim;
    }
    if (dqg > 1e-16)
        lim = 1 / dqg * ((dmax * dmax + K3a) * dqg + 2 * dqg * dqg * dmax) / (dmax * dmax + 2 * dqg * dqg + dmax * dqg + K3a);
    else if (dqg < -1e-16)
        lim = 1 / dqg * ((dmin * dmin + K3a) * dqg + 2 * dqg * dqg * dmin) / (dmin * dmin + 2 * dqg * dqg + dmin * dqg + K3a);
    return lim;
}

// ---------------------------------------------------------------- solver ----
struct BcSpec { int type; BVars vars; };  // type: BC_FARFIELD / BC_SLIP / BC_WALL / -1 other

struct Solver {
    Gas g;
    Mesh m;
    int viscous = 0;              // 0 inviscid, 1 laminar
    bool viscous_model_named_non_inviscid = false;  // viscosity_model != "inviscid"
    bool second_order = true;
    bool michalak = false;        // RANS_MICHALAK_LIMITER build
    bool gg = true;               // green-gauss (else least-squares)
    double limiter_k = 5.;
    double cfl = 1;

    std::vector<double> q, qW, dt, gx, gy, limiters, q_min, q_max, RhoVector, qk;
    std::vector<FaceFlux> ff;
    std::vector<BVars> boundary_vars;
    std::vector<int> bface_type;  // per boundary face, BC_* or -1
    std::vector<double> lsq;      // 4 per real cell, (d^T d)^-1 row-major
    bool qW_touched = false;

    // solver.h:177-197
    void set_mesh_and_gas(const Mesh& mi, const Gas& gi) {
        m = mi; g = gi;
        size_t N = m.nCells();
        q.assign(4 * N, 0.0); qW.assign(4 * N, 0.0); dt.assign(N, 0.0);
        gx.assign(4 * N, 0.0); gy.assign(4 * N, 0.0); limiters.assign(4 * N, 0.0);
        q_min.assign(4 * N, 0.0); q_max.assign(4 * N, 0.0); RhoVector.assign(4 * N, 0.0);
        qk.assign(4 * N, 0.0);
    }
    // solver.h:200-247 ; bcs are given per physical-name key
    void set_bcs(const std::map<std::string, BcSpec>& bcs) {
        size_t nF = m.nFaces();
        std::vector<int> add(nF, BC_INTERNAL);
        boundary_vars.assign(m.boundaryEdges.size(), BVars());
        bface_type.assign(m.boundaryEdges.size(), -1);
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
            u32 e = m.boundaryEdges[b];
            const BcSpec& s = bcs.at(m.boundaryEdgesPhysicals[b]);
            BVars v;
            if (s.type == BC_FARFIELD) { add[e] = BC_FARFIELD; v = s.vars; }
            else if (s.type == BC_SLIP) add[e] = BC_SLIP;
            else if (s.type == BC_WALL) add[e] = BC_WALL;
            bface_type[b] = s.type;
            boundary_vars[b] = v;
        }
        ff.resize(nF);
        for (size_t e = 0; e < nF; ++e) {
            ff[e].g = &g; ff[e].nx = m.edgesNormalsX[e]; ff[e].ny = m.edgesNormalsY[e];
            ff[e].kind = add[e]; ff[e].viscous = viscous;
        }
    }
    // solver.h:259-274
    void refill_bcs() {
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
            u32 c1 = m.edgesCells[2 * m.boundaryEdges[b] + 1];
            double qc[4];
            boundary_vars[b].conservative(g, qc);
            for (int k = 0; k < 4; ++k) q[4 * c1 + k] = qc[k];
        }
    }
    // solver.h:276-287
    void bcs_from_internal() {
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
            u32 e = m.boundaryEdges[b];
            u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
            for (int k = 0; k < 4; ++k) q[4 * c1 + k] = q[4 * c0 + k];
        }
    }
    // solver.h:289-305
    void set_walls_from_internal(std::vector<double>& q_) {
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
            if (bface_type[b] == BC_WALL || bface_type[b] == BC_SLIP) {
                u32 e = m.boundaryEdges[b];
                u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
                for (int k = 0; k < 4; ++k) q_[4 * c1 + k] = q_[4 * c0 + k];
            }
        }
    }
    // solver.h:308-356
    void calc_dt() {
        std::fill(dt.begin(), dt.end(), 0.0);
        for (u32 e = 0; e < m.nFaces(); ++e) {
            u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
            u32 k0 = 4 * c0, k1 = 4 * c1;
            double n0 = m.edgesNormalsX[e], n1 = m.edgesNormalsY[e];
            double eig;
            if (ff[e].two_sided()) {
                const double V_L = (q[k0 + 1] * n0 + q[k0 + 2] * n1) / q[k0];
                const double V_R = (q[k1 + 1] * n0 + q[k1 + 2] * n1) / q[k1];
                const double p_L = (g.gamma - 1) * (q[k0 + 3] - 0.5 / q[k0] * (q[k0 + 1] * q[k0 + 1] + q[k0 + 2] * q[k0 + 2]));
                const double p_R = (g.gamma - 1) * (q[k1 + 3] - 0.5 / q[k1] * (q[k1 + 1] * q[k1 + 1] + q[k1 + 2] * q[k1 + 2]));
                const double eig_L = std::sqrt(p_L * g.gamma / q[k0]) + std::fabs(V_L);
                const double eig_R = std::sqrt(p_R * g.gamma / q[k1]) + std::fabs(V_R);
                eig = std::max(eig_L, eig_R);
            } else {
                const double V_L = (q[k0 + 1] * n0 + q[k0 + 2] * n1) / q[k0];
                const double p_L = (g.gamma - 1) * (q[k0 + 3] - 0.5 / q[k0] * (q[k0 + 1] * q[k0 + 1] + q[k0 + 2] * q[k0 + 2]));
                eig = std::sqrt(p_L * g.gamma / q[k0]) + std::fabs(V_L);
            }
            dt[c0] += eig * m.edgesLengths[e];
            if (ff[e].two_sided()) dt[c1] += eig * m.edgesLengths[e];
        }
        for (size_t i = 0; i < dt.size(); ++i) dt[i] = cfl * m.cellsAreas[i] / dt[i];
    }
    // solver.h:359-398 (A1: direction = normalised SUM of centroids; reads
    // member q).  a1_sum_direction = false: the textbook correction along the
    // cell connection, t = (x1 - x0)/|x1 - x0|, derivative (q1 - q0)/|x1 - x0|
    // (used by the Spalart-Allmaras model, not by the reference's laminar flux)
    bool a1_sum_direction = true;
    void average_gradients(double gxo[4], double gyo[4], u32 c0, u32 c1) const {
        if (c0 == c1) {
            for (int i = 0; i < 4; ++i) { gxo[i] = gx[4 * c0 + i]; gyo[i] = gy[4 * c0 + i]; }
            return;
        }
        double t0 = m.cellsCentersX[c0] + m.cellsCentersX[c1];
        double t1 = m.cellsCentersY[c0] + m.cellsCentersY[c1];
        if (!a1_sum_direction) {
            t0 = m.cellsCentersX[c1] - m.cellsCentersX[c0];
            t1 = m.cellsCentersY[c1] - m.cellsCentersY[c0];
        }
        const double l = std::sqrt(t0 * t0 + t1 * t1);
        t0 /= l; t1 /= l;
        for (int i = 0; i < 4; ++i) {
            double dir = a1_sum_direction ? (q[4 * c0 + i] - q[4 * c1 + i]) / l : (q[4 * c1 + i] - q[4 * c0 + i]) / l;
            double gxb = (gx[4 * c0 + i] + gx[4 * c1 + i]) * 0.5;
            double gyb = (gy[4 * c0 + i] + gy[4 * c1 + i]) * 0.5;
            const double dot = gxb * t0 + gyb * t1;
            gxo[i] = gxb - (dot - dir) * t0;
            gyo[i] = gyb - (dot - dir) * t1;
        }
    }
    // solver.h:402-422 ; the 2x2 inverse is the adjugate / determinant form
    // (Eigen's partial-pivot LU inverse differs at the ulp level: LSQ parity is
    // to tolerance, not bit-exact)
    void calc_least_squares_matrices() {
        lsq.assign(4 * m.nRealCells, 0.0);
        for (u32 i = 0; i < m.nRealCells; ++i) {
            u32 n = m.cellsIsTriangle[i] ? 3 : 4;
            double a = 0, b = 0, d = 0;
            for (u32 j = 0; j < n; ++j) {
                u32 e = m.cellsEdges[4 * i + j];
                u32 cn = m.edgesCells[2 * e] == i ? m.edgesCells[2 * e + 1] : m.edgesCells[2 * e];
                double dx = m.cellsCentersX[cn] - m.cellsCentersX[i];
                double dy = m.cellsCentersY[cn] - m.cellsCentersY[i];
                a += dx * dx; b += dx * dy; d += dy * dy;
            }
            double det = a * d - b * b;
            lsq[4 * i + 0] = d / det; lsq[4 * i + 1] = -b / det;
            lsq[4 * i + 2] = -b / det; lsq[4 * i + 3] = a / det;
        }
    }
    // solver.h:425-514 (A2, A3, A4). Uses member q regardless of argument.
    void calc_gradients() {
        if (gg) {
            std::fill(gx.begin(), gx.end(), 0.0);
            std::fill(gy.begin(), gy.end(), 0.0);
            for (u32 e = 0; e < m.nFaces(); ++e) {
                u32 i = m.edgesCells[2 * e], j = m.edgesCells[2 * e + 1];
                const double dxif = m.edgesCentersX[e] - m.cellsCentersX[i];
                const double dyif = m.edgesCentersY[e] - m.cellsCentersY[i];
                const double dif = std::sqrt(dxif * dxif + dyif * dyif);
                const double dxij = m.cellsCentersX[i] - m.cellsCentersX[j];
                const double dyij = m.cellsCentersY[i] - m.cellsCentersY[j];
                const double dij = std::sqrt(dxij * dxij + dyij * dyij);
                const double gf = dif / dij;
                double qR[4];
                face_vars(ff[e], &q[4 * i], &q[4 * j], qR);
                for (int k = 0; k < 4; ++k) {
                    const double fk = (q[4 * i + k] * (1.0 - gf) + qR[k] * gf) * m.edgesLengths[e];
                    gx[4 * i + k] += fk * m.edgesNormalsX[e];
                    gy[4 * i + k] += fk * m.edgesNormalsY[e];
                    gx[4 * j + k] -= fk * m.edgesNormalsX[e];
                    gy[4 * j + k] -= fk * m.edgesNormalsY[e];
                }
            }
            for (u32 i = 0; i < m.nRealCells; ++i)
                for (int k = 0; k < 4; ++k) { gx[4 * i + k] /= m.cellsAreas[i]; gy[4 * i + k] /= m.cellsAreas[i]; }
            for (u32 i = m.nRealCells; i < m.nCells(); ++i)
                for (int k = 0; k < 4; ++k) { gx[4 * i + k] *= 0; gy[4 * i + k] *= 0; }
        } else {
            if (lsq.empty()) calc_least_squares_matrices();
            for (u32 i = 0; i < m.nRealCells; ++i) {
                u32 n = m.cellsIsTriangle[i] ? 3 : 4;
                double dxs[4], dys[4], del[4][4];
                for (u32 j = 0; j < n; ++j) {
                    u32 e = m.cellsEdges[4 * i + j];
                    u32 cn = m.edgesCells[2 * e] == i ? m.edgesCells[2 * e + 1] : m.edgesCells[2 * e];
                    dxs[j] = m.cellsCentersX[cn] - m.cellsCentersX[i];
                    dys[j] = m.cellsCentersY[cn] - m.cellsCentersY[i];
                    double qR[4];
                    face_vars(ff[e], &q[4 * i], &q[4 * cn], qR);
                    for (int k = 0; k < 4; ++k) del[j][k] = q[4 * i + k] - qR[k];
                }
                for (int k = 0; k < 4; ++k) {
                    double rx = 0, ry = 0;   // d^T * delta
                    for (u32 j = 0; j < n; ++j) { rx += dxs[j] * del[j][k]; ry += dys[j] * del[j][k]; }
                    gx[4 * i + k] = lsq[4 * i + 0] * rx + lsq[4 * i + 1] * ry;
                    gy[4 * i + k] = lsq[4 * i + 2] * rx + lsq[4 * i + 3] * ry;
                }
            }
            for (u32 i = m.nRealCells; i < m.nCells(); ++i)
                for (int k = 0; k < 4; ++k) { gx[4 * i + k] *= 0; gy[4 * i + k] *= 0; }
        }
    }
    // solver.h:517-593, the default branch or the RANS_MICHALAK_LIMITER build (limiter_value)
    void calc_limiters(const std::vector<double>& q_) {
        for (size_t i = 0; i < limiters.size(); ++i) { limiters[i] = 1; q_min[i] = q_[i]; q_max[i] = q_[i]; }
        for (u32 e = 0; e < m.nFaces(); ++e) {
            u32 i = m.edgesCells[2 * e], j = m.edgesCells[2 * e + 1];
            for (int k = 0; k < 4; ++k) {
                q_min[4 * i + k] = std::min(q_min[4 * i + k], q_[4 * j + k]);
                q_min[4 * j + k] = std::min(q_min[4 * j + k], q_[4 * i + k]);
                q_max[4 * i + k] = std::max(q_max[4 * i + k], q_[4 * j + k]);
                q_max[4 * j + k] = std::max(q_max[4 * j + k], q_[4 * i + k]);
            }
        }
        for (u32 e = 0; e < m.nFaces(); ++e) {
            u32 ids[2] = {m.edgesCells[2 * e], m.edgesCells[2 * e + 1]};
            for (u32 id : ids) {
                if (id >= m.nRealCells) continue;
                const double dx = m.edgesCentersX[e] - m.cellsCentersX[id];
                const double dy = m.edgesCentersY[e] - m.cellsCentersY[id];
                const double sqrt_area = std::sqrt(m.cellsAreas[id]);
                for (int k = 0; k < 4; ++k) {
                    double dqg = gx[4 * id + k] * dx + gy[4 * id + k] * dy;
                    double dmax = q_max[4 * id + k] - q_[4 * id + k];
                    double dmin = q_min[4 * id + k] - q_[4 * id + k];
                    const double Ka = limiter_k * sqrt_area;
                    const double K3a = Ka * Ka * Ka;
                    const double lim = limiter_value(michalak, dqg, dmax, dmin, K3a);
                    limiters[4 * id + k] = std::min(limiters[4 * id + k], lim);
                }
            }
        }
    }
    // face loop shared by fillRhoRHS (solver.h:1096-1141) and
    // explicitSolver::calc_residual (solver.h:751-792); base = reconstruction state
    void face_loop(const std::vector<double>& base, std::vector<double>& out) {
        for (u32 e = 0; e < m.nFaces(); ++e) {
            u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
            u32 k0 = 4 * c0, k1 = 4 * c1;
            double gxa[4], gya[4];
            average_gradients(gxa, gya, c0, c1);
            double f[4];
            if (second_order) {
                const double d0x = m.edgesCentersX[e] - m.cellsCentersX[c0];
                const double d0y = m.edgesCentersY[e] - m.cellsCentersY[c0];
                const double d1x = m.edgesCentersX[e] - m.cellsCentersX[c1];
                const double d1y = m.edgesCentersY[e] - m.cellsCentersY[c1];
                double qL[4], qR[4];
                for (int k = 0; k < 4; ++k) {
                    qL[k] = base[k0 + k] + (gx[k0 + k] * d0x + gy[k0 + k] * d0y) * limiters[k0 + k];
                    qR[k] = base[k1 + k] + (gx[k1 + k] * d1x + gy[k1 + k] * d1y) * limiters[k1 + k];
                }
                face_flux(ff[e], qL, qR, gxa, gya, f);
            } else {
                face_flux(ff[e], &base[k0], &base[k1], gxa, gya, f);
            }
            for (int k = 0; k < 4; ++k) f[k] = f[k] * m.edgesLengths[e];
            for (int k = 0; k < 4; ++k) {
                out[k0 + k] -= f[k];
                if (ff[e].two_sided()) out[k1 + k] += f[k];
            }
        }
    }
    // solver.h:1079-1152
    void fillRhoRHS() {
        calc_dt();
        if (second_order || viscous_model_named_non_inviscid) {
            set_walls_from_internal(q);
            calc_gradients();
        }
        if (second_order) calc_limiters(q);
        std::fill(RhoVector.begin(), RhoVector.end(), 0.0);
        face_loop(q, RhoVector);
        for (u32 i = m.nRealCells; i < m.nCells(); ++i)
            for (int k = 0; k < 4; ++k) RhoVector[4 * i + k] = 0;
    }
    // solver.h:745-799
    void calc_residual(const std::vector<double>& q_) {
        std::fill(qW.begin(), qW.end(), 0.0);
        face_loop(q_, qW);
        for (u32 i = 0; i < m.nRealCells; ++i)
            for (int k = 0; k < 4; ++k) qW[4 * i + k] /= m.cellsAreas[i];
        qW_touched = true;
    }
    // solver.h:802-828 ; the returned norm is a plain serial sum (Eigen's
    // vectorised reduction order is not reproducible without Eigen)
    double explicit_solve(double relaxation) {
        static const double alpha[3] = {0.25, 0.5, 1.};
        calc_dt();
        qk = q;
        for (double ai : alpha) {
            if (viscous_model_named_non_inviscid || second_order) {
                set_walls_from_internal(qk);
                calc_gradients();
                if (second_order) calc_limiters(qk);
            }
            calc_residual(qk);
            for (u32 i = 0; i < m.nRealCells; ++i)
                for (int j = 0; j < 4; ++j)
                    qk[4 * i + j] = q[4 * i + j] + qW[4 * i + j] * dt[i] * ai * relaxation;
        }
        q = qk;
        double s = 0;
        for (double v : qW) s += v * v;
        return std::sqrt(s);
    }
    // solver.h:597-611
    BVars get_boundary_variables() const {
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b)
            if (bface_type[b] == BC_FARFIELD) return boundary_vars[b];
        return BVars();
    }
    // solver.h:615-631
    void init() {
        double qi[4];
        get_boundary_variables().conservative(g, qi);
        for (u32 i = 0; i < m.nRealCells; ++i)
            for (int k = 0; k < 4; ++k) q[4 * i + k] = qi[k];
    }
    // solver.h:636-690 (A8: real rows of qW keep their previous content)
    double get_uniform_residual() {
        for (u32 i = m.nRealCells; i < m.nCells(); ++i)
            for (int k = 0; k < 4; ++k) qW[4 * i + k] = 0;
        double qf[4];
        get_boundary_variables().conservative(g, qf);
        for (u32 e = 0; e < m.nFaces(); ++e) {
            u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
            double gxa[4], gya[4], f[4];
            average_gradients(gxa, gya, c0, c1);
            face_flux(ff[e], qf, qf, gxa, gya, f);
            for (int k = 0; k < 4; ++k) f[k] = f[k] * m.edgesLengths[e];
            for (int k = 0; k < 4; ++k) {
                qW[4 * c0 + k] -= f[k];
                if (ff[e].two_sided()) qW[4 * c1 + k] += f[k];
            }
        }
        double s = 0;
        for (double v : qW) s += v * v;
        return std::sqrt(s);
    }
    // post.h:301-387
    void wall_profile(const std::string& patch, double* cl, double* cd, double* cm) const {
        BVars vf = get_boundary_variables();
        const double p_inf = vf.p, mach_inf = vf.mach;
        double xmin = 0., xmax = 0., y_moment = 0.;
        u32 n_added = 0;
        bool first = true;
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
            if (m.boundaryEdgesPhysicals[b] != patch) continue;
            u32 e = m.boundaryEdges[b];
            if (first) { xmin = xmax = m.edgesCentersX[e]; y_moment = m.edgesCentersY[e]; first = false; }
            else { xmin = std::min(xmin, m.edgesCentersX[e]); xmax = std::max(xmax, m.edgesCentersX[e]); y_moment += m.edgesCentersY[e]; }
            n_added += 1;
        }
        y_moment /= (double)n_added;
        const double x_moment = (xmax - xmin) * 0.25 + xmin;
        double fx = 0., fy = 0., mm = 0.;
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
            if (m.boundaryEdgesPhysicals[b] != patch) continue;
            u32 e = m.boundaryEdges[b];
            u32 c = m.edgesCells[2 * e];
            const double* qc = &q[4 * c];
            const double p = (g.gamma - 1) * (qc[3] - 0.5 / qc[0] * (qc[1] * qc[1] + qc[2] * qc[2]));  // core.h:145-147
            const double cp = 2. / (g.gamma * mach_inf * mach_inf) * (p / p_inf - 1.);
            const double ecx = m.edgesCentersX[e], ecy = m.edgesCentersY[e];
            const double fxi = cp * m.edgesNormalsX[e] * m.edgesLengths[e] / (xmax - xmin);
            const double fyi = cp * m.edgesNormalsY[e] * m.edgesLengths[e] / (xmax - xmin);
            const double mi = (ecx - x_moment) / (xmax - xmin) * fyi - (ecy - y_moment) / (xmax - xmin) * fxi;
            fx += fxi; fy += fyi; mm -= mi;
        }
        const double aoa = vf.angle;
        *cd = fx * std::cos(aoa) + fy * std::sin(aoa);
        *cl = -fx * std::sin(aoa) + fy * std::cos(aoa);
        *cm = mm;
    }
    // solver.h:979-1075 (fillRhoLHS): the 4N x 4N matrix as COO triplets in
    // the order the reference touches them (duplicates are summed by the
    // caller); ghost rows become identity rows
    void fill_lhs(std::vector<int64_t>& rows, std::vector<int64_t>& cols, std::vector<double>& vals) {
        calc_dt();
        if (viscous_model_named_non_inviscid) {
            set_walls_from_internal(q);
            calc_gradients();
        }
        rows.clear(); cols.clear(); vals.clear();
        const int64_t nT = m.nCells(), nR = m.nRealCells;
        for (int64_t i = 0; i < nR; ++i)
            for (int j = 0; j < 4; ++j) { rows.push_back(4 * i + j); cols.push_back(4 * i + j); vals.push_back(m.cellsAreas[i] / dt[i]); }
        double J[64];
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const int64_t k0 = 4 * (int64_t)m.edgesCells[2 * e], k1 = 4 * (int64_t)m.edgesCells[2 * e + 1];
            face_jacobian(e, J);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    rows.push_back(k0 + i); cols.push_back(k0 + j); vals.push_back(J[8 * i + j]);
                    rows.push_back(k0 + i); cols.push_back(k1 + j); vals.push_back(J[8 * i + j + 4]);
                    if (ff[e].two_sided()) {
                        rows.push_back(k1 + i); cols.push_back(k0 + j); vals.push_back(J[8 * (i + 4) + j]);
                        rows.push_back(k1 + i); cols.push_back(k1 + j); vals.push_back(J[8 * (i + 4) + j + 4]);
                    }
                }
        }
        // ghost rows: identity (solver.h:1061-1071); ghost rows receive no
        // other entries because boundary faces are one-sided
        for (int64_t i = nR; i < nT; ++i)
            for (int j = 0; j < 4; ++j) { rows.push_back(4 * i + j); cols.push_back(4 * i + j); vals.push_back(1.0); }
    }

    // ransMatrix.h:138-178 (RansMatrix * x, RANS_MATRIX_FREE build) at the
    // member q and dt: real rows x * A / dt, ghost rows x; per edge the
    // two-argument flux (zero gradients), one step h = 1e-6 along the whole
    // perturbed 4-vector, no edge length
    void mf_apply(const double* x, double* y) const {
        const int64_t nT = m.nCells(), nR = m.nRealCells;
        for (int64_t i = 0; i < nR; ++i)
            for (int k = 0; k < 4; ++k) y[4 * i + k] = x[4 * i + k] * m.cellsAreas[i] / dt[i];
        for (int64_t i = nR; i < nT; ++i)
            for (int k = 0; k < 4; ++k) y[4 * i + k] = x[4 * i + k];
        const double h = 1e-6;
        const double zero[4] = {0, 0, 0, 0};
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const u32 i = m.edgesCells[2 * e], j = m.edgesCells[2 * e + 1];
            const double* qL = &q[4 * (size_t)i];
            const double* qR = &q[4 * (size_t)j];
            double FB[4], F1[4], qp[4];
            face_flux(ff[e], qL, qR, zero, zero, FB);
            for (int k = 0; k < 4; ++k) qp[k] = qL[k] + h * x[4 * (size_t)i + k];
            face_flux(ff[e], qp, qR, zero, zero, F1);
            for (int k = 0; k < 4; ++k) y[4 * (size_t)i + k] += (F1[k] - FB[k]) / h;
            if (ff[e].two_sided()) {
                for (int k = 0; k < 4; ++k) qp[k] = qR[k] + h * x[4 * (size_t)j + k];
                face_flux(ff[e], qL, qp, zero, zero, F1);
                for (int k = 0; k < 4; ++k) y[4 * (size_t)j + k] -= (F1[k] - FB[k]) / h;
            }
        }
    }

    // physics.h:533-578 (calc_convective_jacobian) for face e at first-order
    // states as in fillRhoLHS (solver.h:1030-1043), times the face length
    void face_jacobian(u32 e, double J[64]) {
        u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
        double qL[4], qR[4];
        for (int k = 0; k < 4; ++k) { qL[k] = q[4 * c0 + k]; qR[k] = q[4 * c1 + k]; }
        double gxa[4], gya[4];
        average_gradients(gxa, gya, c0, c1);
        double f[4], fp[4];
        face_flux(ff[e], qL, qR, gxa, gya, f);
        double Jm[8][8];
        for (int i = 0; i < 4; ++i) {
            {
                const double up = std::max(1e-6, std::fabs(qL[i]) * 1e-6);
                qL[i] += up;
                face_flux(ff[e], qL, qR, gxa, gya, fp);
                qL[i] -= up;
                for (int r = 0; r < 4; ++r) { Jm[r][i] = (fp[r] - f[r]) / up; Jm[r + 4][i] = -Jm[r][i]; }
            }
            {
                const double up = std::max(1e-6, std::fabs(qR[i]) * 1e-6);
                qR[i] += up;
                face_flux(ff[e], qL, qR, gxa, gya, fp);
                qR[i] -= up;
                for (int r = 0; r < 4; ++r) { Jm[r][i + 4] = (fp[r] - f[r]) / up; Jm[r + 4][i + 4] = -Jm[r][i + 4]; }
            }
        }
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 8; ++c) J[8 * r + c] = Jm[r][c] * m.edgesLengths[e];
    }
};


// ===================================================================== SA ==
// Spalart-Allmaras one-equation model.  The reference's SA branch is empty
// (physics.h:259-261; nu/nuW allocated but unused, solver.h:55,59,187-188), so
// there is nothing to restate: this is the published model,
//   SA-noft2  (Spalart & Allmaras 1994; NASA Turbulence Modeling Resource),
//   S-tilde limiter and negative-nu~ branch of SA-neg (Allmaras, Johnson &
//   Spalart 2012, ICCFD7-1902),
// in compressible conservative form with rho*nu~ as fifth variable:
//   d(rho nu~)/dt + div(rho u nu~) = rho (P - D)
//        + (1/sigma) [div((mu + rho nu~ f_n) grad nu~) + c_b2 rho |grad nu~|^2]
// discretised on the reference's machinery:
//   * gradients of rho nu~ by the same Green-Gauss / least-squares operator
//     as q (solver.h:425-514), limiter of solver.h:517-593, MUSCL reconstruction
//     of solver.h:1097-1120;
//   * face gradients of all five variables by average_gradients
//     (solver.h:359-398) with the textbook direction t = (x1 - x0)/|x1 - x0|
//     and derivative (q1 - q0)/|x1 - x0| in place of the sum of centroids
//     (A1), which makes the viscous operator non-dissipative once mu_t is
//     hundreds of mu_L;
//   * convection: rho nu~ upwinded by the sign of the Roe mass flux;
//   * mean-flow viscous terms of physics.h:230-257 with mu_L + mu_t,
//     mu_t = rho nu~ f_v1, k = cp (mu_L/Pr_L + mu_t/Pr_T);
//   * boundary faces carry no mean-flow gradient terms, like the reference's
//     boundary functors (A6): wall ghost rho nu~ = -rho nu~ (nu~ = 0 on the
//     wall face), slip wall = interior, farfield = ghost nu~ on inflow,
//     interior on outflow; the nu~ = 0 wall condition enters the fifth
//     equation as the diffusion flux (mu_L/sigma) (0 - nu~_c)/d_n through
//     no-slip wall faces (d_n: centroid-to-face normal distance);
//   * sources with the cell's unlimited gradients and wall distance
//     (mesh.h:794-830).
// At rho nu~ = 0 everywhere the mean-flow rows equal the laminar residual.
namespace sa {
const double cb1 = 0.1355, cb2 = 0.622, sigma = 2.0 / 3.0, kappa = 0.41, cv1 = 7.1, cw2 = 0.3, cw3 = 2.0;
const double cv2 = 0.7, cv3 = 0.9, ct3 = 1.2, cn1 = 16.0;
inline double cw1() { return cb1 / (kappa * kappa) + (1.0 + cb2) / sigma; }
// f_v1 (chi >= 0); SA-neg: mu_t = 0 for chi < 0
inline double fv1(double chi) {
    if (chi <= 0.0) return 0.0;
    const double c3 = chi * chi * chi;
    return c3 / (c3 + cv1 * cv1 * cv1);
}
// diffusion coefficient multiplier f_n (1 for chi >= 0)
inline double fn(double chi) {
    if (chi >= 0.0) return 1.0;
    const double c3 = chi * chi * chi;
    return (cn1 + c3) / (cn1 - c3);
}
// f_w(r) with g = r + c_w2 (r^6 - r)
inline double fw(double r) {
    const double r2 = r * r, g = r + cw2 * (r2 * r2 * r2 - r);
    const double g2 = g * g, c2 = cw3 * cw3;
    const double c6 = c2 * c2 * c2;
    return g * std::cbrt(std::sqrt((1.0 + c6) / (g2 * g2 * g2 + c6)));
}
// per-volume source of rho nu~: rho (P - D) + c_b2/sigma rho |grad nu~|^2
inline double source(double rho, double nut, double mu, double omega, double d, double gnx, double gny) {
    const double cross = cb2 / sigma * (gnx * gnx + gny * gny);
    if (nut < 0.0)   // SA-neg
        return rho * (cb1 * (1.0 - ct3) * omega * nut + cw1() * (nut / d) * (nut / d) + cross);
    const double chi = rho * nut / mu;
    const double f1 = fv1(chi);
    const double f2 = 1.0 - chi / (1.0 + chi * f1);
    const double kd2 = kappa * kappa * d * d;
    const double sbar = nut * f2 / kd2;
    const double st = (sbar >= -cv2 * omega)
                          ? omega + sbar
                          : omega + omega * (cv2 * cv2 * omega + cv3 * sbar) / ((cv3 - 2.0 * cv2) * omega - sbar);
    double r = 10.0;
    if (st > 0.0) { r = nut / (st * kd2); r = r < 10.0 ? r : 10.0; }
    const double P = cb1 * st * nut;
    const double D = cw1() * fw(r) * (nut / d) * (nut / d);
    return rho * (P - D + cross);
}
}  // namespace sa

// The SA model's wall distance: from each cell centroid to the nearest point
// of the wall and slip-wall face segments, the face set of compute_wall_dist
// (mesh.h:794-830), which measures to the face centres instead; on a wall
// cell of aspect ratio a that overestimates the distance up to a/2-fold, and
// the model's destruction ~ 1/d^2 with it.  1 when there is no such face, as
// there.  Point-to-segment: t = clamp(((p - a).(b - a)) / |b - a|^2, 0, 1),
// d = |p - (a + t (b - a))|, a and b the face's nodes in edge order.
static void model_wall_dist(const Mesh& m, const std::vector<int>& bface_type, std::vector<double>& dist) {
    dist.assign(m.nCells(), 1.0);
    for (u32 i = 0; i < m.nCells(); ++i) {
        const double px = m.cellsCentersX[i], py = m.cellsCentersY[i];
        double mind = 1;
        bool first = true;
        for (size_t j = 0; j < m.boundaryEdges.size(); ++j) {
            if (bface_type[j] != BC_WALL && bface_type[j] != BC_SLIP) continue;
            const u32 e = m.boundaryEdges[j];
            const double ax = m.nodesX[m.edgesNodes[2 * e]], ay = m.nodesY[m.edgesNodes[2 * e]];
            const double dx = m.nodesX[m.edgesNodes[2 * e + 1]] - ax, dy = m.nodesY[m.edgesNodes[2 * e + 1]] - ay;
            const double ll = dx * dx + dy * dy;
            double t = ((px - ax) * dx + (py - ay) * dy) / ll;
            t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
            const double ex = px - (ax + t * dx), ey = py - (ay + t * dy);
            const double d = std::sqrt(ex * ex + ey * ey);
            if (first) { mind = d; first = false; } else mind = std::min(mind, d);
        }
        dist[i] = mind;
    }
}

enum { SA_CONV = 1, SA_DIFF = 2, SA_SOURCE = 4, SA_MUT = 8, SA_WALL = 16, SA_ALL = 31 };

struct SASolver {
    Solver* s;
    std::vector<double> nu, gnx, gny, lnu, TurbVector;   // rho nu~ per cell (the reference's nu)
    std::vector<double> dist;                             // the model's wall distance (model_wall_dist)
    int terms = SA_ALL;                                   // for tests: which terms enter
    explicit SASolver(Solver* s_) : s(s_) {
        s->a1_sum_direction = false;   // textbook face gradients for every viscous term of the model
        const size_t N = s->m.nCells();
        nu.assign(N, 0.0); gnx.assign(N, 0.0); gny.assign(N, 0.0); lnu.assign(N, 1.0); TurbVector.assign(N, 0.0);
    }
    // rho nu~ on the far side of face e from the owner state (qL, nL) and the
    // ghost/neighbour state (qR, nR): gradient form (internal: average, A4)
    double vars_nu(u32 e, const double* qL, double nL, const double* qR, double nR) const {
        const FaceFlux& F = s->ff[e];
        if (F.kind == BC_INTERNAL) return (nL + nR) * 0.5;
        if (F.kind == BC_WALL) return -nL;
        if (F.kind == BC_SLIP) return nL;
        double qb[4];
        face_vars(F, qL, qR, qb);
        const double io = qL[1] * F.nx + qL[2] * F.ny;   // as physics.h:480
        return qb[0] * (io < 0 ? nR / qR[0] : nL / qL[0]);
    }
    void gradients() {
        const Solver& S = *s;
        const Mesh& m = S.m;
        const u32 nR = m.nRealCells, nT = m.nCells();
        std::fill(gnx.begin(), gnx.end(), 0.0);
        std::fill(gny.begin(), gny.end(), 0.0);
        if (S.gg) {   // solver.h:429-465 applied to rho nu~
            for (u32 e = 0; e < m.nFaces(); ++e) {
                const u32 i = m.edgesCells[2 * e], j = m.edgesCells[2 * e + 1];
                const double dxif = m.edgesCentersX[e] - m.cellsCentersX[i];
                const double dyif = m.edgesCentersY[e] - m.cellsCentersY[i];
                const double dif = std::sqrt(dxif * dxif + dyif * dyif);
                const double dxij = m.cellsCentersX[i] - m.cellsCentersX[j];
                const double dyij = m.cellsCentersY[i] - m.cellsCentersY[j];
                const double gf = dif / std::sqrt(dxij * dxij + dyij * dyij);
                const double vr = vars_nu(e, &S.q[4 * i], nu[i], &S.q[4 * j], nu[j]);
                const double fk = (nu[i] * (1.0 - gf) + vr * gf) * m.edgesLengths[e];
                gnx[i] += fk * m.edgesNormalsX[e]; gny[i] += fk * m.edgesNormalsY[e];
                gnx[j] -= fk * m.edgesNormalsX[e]; gny[j] -= fk * m.edgesNormalsY[e];
            }
            for (u32 i = 0; i < nR; ++i) { gnx[i] /= m.cellsAreas[i]; gny[i] /= m.cellsAreas[i]; }
        } else {      // solver.h:467-505 applied to rho nu~
            for (u32 i = 0; i < nR; ++i) {
                const u32 n = m.cellsIsTriangle[i] ? 3 : 4;
                double rx = 0, ry = 0;
                for (u32 jj = 0; jj < n; ++jj) {
                    const u32 e = m.cellsEdges[4 * i + jj];
                    const u32 cn = m.edgesCells[2 * e] == i ? m.edgesCells[2 * e + 1] : m.edgesCells[2 * e];
                    const double dx = m.cellsCentersX[cn] - m.cellsCentersX[i];
                    const double dy = m.cellsCentersY[cn] - m.cellsCentersY[i];
                    const double del = nu[i] - vars_nu(e, &S.q[4 * i], nu[i], &S.q[4 * cn], nu[cn]);
                    rx += dx * del; ry += dy * del;
                }
                gnx[i] = S.lsq[4 * i + 0] * rx + S.lsq[4 * i + 1] * ry;
                gny[i] = S.lsq[4 * i + 2] * rx + S.lsq[4 * i + 3] * ry;
            }
        }
        for (u32 i = nR; i < nT; ++i) { gnx[i] = 0; gny[i] = 0; }
    }
    void limiters() {   // solver.h:517-593 applied to rho nu~
        const Solver& S = *s;
        const Mesh& m = S.m;
        const u32 nR = m.nRealCells;
        std::fill(lnu.begin(), lnu.end(), 1.0);
        std::vector<double> mn(nu), mx(nu);
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const u32 i = m.edgesCells[2 * e], j = m.edgesCells[2 * e + 1];
            mn[i] = std::min(mn[i], nu[j]); mn[j] = std::min(mn[j], nu[i]);
            mx[i] = std::max(mx[i], nu[j]); mx[j] = std::max(mx[j], nu[i]);
        }
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const u32 ids[2] = {m.edgesCells[2 * e], m.edgesCells[2 * e + 1]};
            for (u32 id : ids) {
                if (id >= nR) continue;
                const double dx = m.edgesCentersX[e] - m.cellsCentersX[id];
                const double dy = m.edgesCentersY[e] - m.cellsCentersY[id];
                const double Ka = S.limiter_k * std::sqrt(m.cellsAreas[id]);
                const double K3a = Ka * Ka * Ka;
                const double dqg = gnx[id] * dx + gny[id] * dy;
                const double dmax = mx[id] - nu[id], dmin = mn[id] - nu[id];
                const double lim = limiter_value(S.michalak, dqg, dmax, dmin, K3a);
                lnu[id] = std::min(lnu[id], lim);
            }
        }
    }
    // flux of face e (times nothing: caller multiplies by the length) at the
    // face states (qL, nL), (qR, nR).  zL/znL, zR/znR: the cell-centre states
    // of the owner and the neighbour, which give the directional part of the
    // face gradients and the no-slip wall's cell-centre nu~ (null: the member
    // state, as in the residual)
    void face_flux5(u32 e, const double* qL, double nL, const double* qR, double nR, double f[5],
                    const double* zL = nullptr, double znL = 0, const double* zR = nullptr, double znR = 0) const {
        const Solver& S = *s;
        const FaceFlux& F = S.ff[e];
        const Gas& g = S.g;
        const u32 c0 = S.m.edgesCells[2 * e], c1 = S.m.edgesCells[2 * e + 1];
        if (!zL) { zL = &S.q[4 * c0]; znL = nu[c0]; }
        if (!zR) { zR = &S.q[4 * c1]; znR = nu[c1]; }
        static const double zero[4] = {0, 0, 0, 0};
        FaceFlux inv = F;
        inv.kind = BC_INTERNAL;
        inv.viscous = 0;
        if (F.kind != BC_INTERNAL) {   // boundary functors: Roe on (qL, vars(qL, q_ghost)), no gradients
            double qb[4];
            face_vars(F, qL, qR, qb);
            const double nb = vars_nu(e, qL, nL, qR, nR);
            roe_flux(inv, qL, qb, zero, zero, f);
            f[4] = (terms & SA_CONV) ? f[0] * (f[0] >= 0 ? nL / qL[0] : nb / qb[0]) : 0.0;
            if (F.kind == BC_WALL && (terms & SA_WALL)) {
                // nu~ = 0 on a no-slip wall: one-sided normal derivative from
                // the cell centre, (0 - nu~_c) / d_n, d_n = (x_f - x_c) . n
                const Mesh& m = S.m;
                const double dn = (m.edgesCentersX[e] - m.cellsCentersX[c0]) * F.nx +
                                  (m.edgesCentersY[e] - m.cellsCentersY[c0]) * F.ny;
                const double nut_c = znL / zL[0];
                f[4] -= (g.mu_L / sa::sigma) * ((0.0 - nut_c) / dn);
            }
            return;
        }
        roe_flux(inv, qL, qR, zero, zero, f);
        f[4] = (terms & SA_CONV) ? f[0] * (f[0] >= 0 ? nL / qL[0] : nR / qR[0]) : 0.0;
        // face gradients: S.average_gradients with the textbook direction
        // (a1_sum_direction is off for the model), written out here so the
        // directional part can come from the centre states zL, zR
        const Mesh& m = S.m;
        double t0 = m.cellsCentersX[c1] - m.cellsCentersX[c0], t1 = m.cellsCentersY[c1] - m.cellsCentersY[c0];
        const double l = std::sqrt(t0 * t0 + t1 * t1);
        t0 /= l; t1 /= l;
        double gxa[4], gya[4];
        for (int k = 0; k < 4; ++k) {
            const double dir = (zR[k] - zL[k]) / l;
            const double gxb = (S.gx[4 * c0 + k] + S.gx[4 * c1 + k]) * 0.5;
            const double gyb = (S.gy[4 * c0 + k] + S.gy[4 * c1 + k]) * 0.5;
            const double dot = gxb * t0 + gyb * t1;
            gxa[k] = gxb - (dot - dir) * t0;
            gya[k] = gyb - (dot - dir) * t1;
        }
        double qc[4];
        for (int k = 0; k < 4; ++k) qc[k] = 0.5 * (qL[k] + qR[k]);
        const double nc = 0.5 * (nL + nR);
        const double chi = nc / g.mu_L;
        const double mut = (terms & SA_MUT) ? nc * sa::fv1(chi) : 0.0;
        // physics.h:230-257 with mu_L + mu_t
        {
            const double mu = g.mu_L + mut;
            const double kk = g.cp * (g.mu_L / g.Pr_L + mut / g.Pr_T);
            double gradT[2];
            calc_grad_temp(gradT, qc, gxa, gya, g);
            const double r2 = qc[0] * qc[0];
            const double gu0 = (qc[0] * gxa[1] - qc[1] * gxa[0]) / r2, gu1 = (qc[0] * gya[1] - qc[1] * gya[0]) / r2;
            const double gv0 = (qc[0] * gxa[2] - qc[2] * gxa[0]) / r2, gv1 = (qc[0] * gya[2] - qc[2] * gya[0]) / r2;
            const double div_v = gu0 + gv1;
            const double txx = 2. * mu * (gu0 - div_v / 3.), tyy = 2. * mu * (gv1 - div_v / 3.), txy = mu * (gu1 + gv0);
            const double phi0 = qc[1] / qc[0] * txx + qc[2] / qc[0] * txy + kk * gradT[0];
            const double phi1 = qc[1] / qc[0] * txy + qc[2] / qc[0] * tyy + kk * gradT[1];
            f[1] -= F.nx * txx + F.ny * txy;
            f[2] -= F.nx * txy + F.ny * tyy;
            f[3] -= F.nx * phi0 + F.ny * phi1;
        }
        if (terms & SA_DIFF) {
            // face gradient of rho nu~ (average_gradients with the textbook
            // direction), then grad nu~ = (grad(rho nu~) - nu~ grad rho) / rho
            const double dir = (znR - znL) / l;
            const double gxb = (gnx[c0] + gnx[c1]) * 0.5, gyb = (gny[c0] + gny[c1]) * 0.5;
            const double dot = gxb * t0 + gyb * t1;
            const double gfx = gxb - (dot - dir) * t0, gfy = gyb - (dot - dir) * t1;
            const double nut = nc / qc[0];
            const double gvx = (gfx - nut * gxa[0]) / qc[0], gvy = (gfy - nut * gya[0]) / qc[0];
            const double dcoef = (g.mu_L + nc * sa::fn(chi)) / sa::sigma;
            f[4] -= dcoef * (F.nx * gvx + F.ny * gvy);
        }
    }
    // the part of fillRhoRHS before the limiters (solver.h:1083-1085): wall
    // and slip ghosts copy their owners (mean flow and rho nu~), then the
    // unlimited gradients of the five variables
    void prepare() {
        Solver& S = *s;
        const Mesh& m = S.m;
        if (dist.empty()) model_wall_dist(m, S.bface_type, dist);
        S.set_walls_from_internal(S.q);
        for (size_t b = 0; b < m.boundaryEdges.size(); ++b)
            if (S.bface_type[b] == BC_WALL || S.bface_type[b] == BC_SLIP) {
                const u32 e = m.boundaryEdges[b];
                nu[m.edgesCells[2 * e + 1]] = nu[m.edgesCells[2 * e]];
            }
        if (!S.gg && S.lsq.empty()) S.calc_least_squares_matrices();
        S.calc_gradients();
        gradients();
    }
    // fillRhoRHS (solver.h:1079-1152) with the fifth equation; RhoVector and
    // TurbVector are the un-divided residuals, ghost rows zero
    void fill_rhs() {
        Solver& S = *s;
        const Mesh& m = S.m;
        const Gas& g = S.g;
        const u32 nR = m.nRealCells, nT = m.nCells();
        S.calc_dt();
        prepare();
        if (S.second_order) { S.calc_limiters(S.q); limiters(); }
        std::fill(S.RhoVector.begin(), S.RhoVector.end(), 0.0);
        std::fill(TurbVector.begin(), TurbVector.end(), 0.0);
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
            double qL[4], qR[4], nL = nu[c0], nRv = nu[c1];
            for (int k = 0; k < 4; ++k) { qL[k] = S.q[4 * c0 + k]; qR[k] = S.q[4 * c1 + k]; }
            if (S.second_order) {
                const double d0x = m.edgesCentersX[e] - m.cellsCentersX[c0], d0y = m.edgesCentersY[e] - m.cellsCentersY[c0];
                const double d1x = m.edgesCentersX[e] - m.cellsCentersX[c1], d1y = m.edgesCentersY[e] - m.cellsCentersY[c1];
                for (int k = 0; k < 4; ++k) {
                    qL[k] = qL[k] + (S.gx[4 * c0 + k] * d0x + S.gy[4 * c0 + k] * d0y) * S.limiters[4 * c0 + k];
                    qR[k] = qR[k] + (S.gx[4 * c1 + k] * d1x + S.gy[4 * c1 + k] * d1y) * S.limiters[4 * c1 + k];
                }
                nL = nL + (gnx[c0] * d0x + gny[c0] * d0y) * lnu[c0];
                nRv = nRv + (gnx[c1] * d1x + gny[c1] * d1y) * lnu[c1];
            }
            double f[5];
            face_flux5(e, qL, nL, qR, nRv, f);
            for (int k = 0; k < 5; ++k) f[k] = f[k] * m.edgesLengths[e];
            for (int k = 0; k < 4; ++k) {
                S.RhoVector[4 * c0 + k] -= f[k];
                if (S.ff[e].two_sided()) S.RhoVector[4 * c1 + k] += f[k];
            }
            TurbVector[c0] -= f[4];
            if (S.ff[e].two_sided()) TurbVector[c1] += f[4];
        }
        if (terms & SA_SOURCE)
            for (u32 i = 0; i < nR; ++i) TurbVector[i] += m.cellsAreas[i] * cell_source(i);
        for (u32 i = nR; i < nT; ++i) {
            for (int k = 0; k < 4; ++k) S.RhoVector[4 * i + k] = 0;
            TurbVector[i] = 0;
        }
        (void)g;
    }
    // local time step of the SA explicit step: calc_dt (solver.h:308-356) plus
    // the viscous spectral radius 4 sum_e coef_e len_e^2 / A_i with
    // coef = max(4/3, gamma) (mu_L/Pr_L + mu_t/Pr_T) / rho at the face average
    // (owner state on boundary faces) - Blazek, Computational Fluid Dynamics,
    // eq. 6.22 with C = 4
    void calc_dt() {
        Solver& S = *s;
        const Mesh& m = S.m;
        const Gas& g = S.g;
        std::vector<double> sc(m.nCells(), 0.0), sv(m.nCells(), 0.0);
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const u32 c0 = m.edgesCells[2 * e], c1 = m.edgesCells[2 * e + 1];
            const double* q0 = &S.q[4 * c0];
            const double* q1 = &S.q[4 * c1];
            const double n0 = m.edgesNormalsX[e], n1 = m.edgesNormalsY[e];
            const bool two = S.ff[e].two_sided();
            const double V_L = (q0[1] * n0 + q0[2] * n1) / q0[0];
            const double p_L = (g.gamma - 1) * (q0[3] - 0.5 / q0[0] * (q0[1] * q0[1] + q0[2] * q0[2]));
            double eig = std::sqrt(p_L * g.gamma / q0[0]) + std::fabs(V_L);
            double rho = q0[0], nuf = nu[c0];
            if (two) {
                const double V_R = (q1[1] * n0 + q1[2] * n1) / q1[0];
                const double p_R = (g.gamma - 1) * (q1[3] - 0.5 / q1[0] * (q1[1] * q1[1] + q1[2] * q1[2]));
                eig = std::max(eig, std::sqrt(p_R * g.gamma / q1[0]) + std::fabs(V_R));
                rho = 0.5 * (rho + q1[0]);
                nuf = 0.5 * (nuf + nu[c1]);
            }
            const double mut = nuf * sa::fv1(nuf / g.mu_L);
            const double kmax = g.gamma > 4.0 / 3.0 ? g.gamma : 4.0 / 3.0;
            const double coef = kmax * (g.mu_L / g.Pr_L + mut / g.Pr_T) / rho;
            const double len = m.edgesLengths[e];
            sc[c0] += eig * len;
            sv[c0] += coef * len * len;
            if (two) { sc[c1] += eig * len; sv[c1] += coef * len * len; }
        }
        for (size_t i = 0; i < S.dt.size(); ++i)
            S.dt[i] = S.cfl * m.cellsAreas[i] / (sc[i] + 4.0 * (sv[i] / m.cellsAreas[i]));
    }
    // per-volume source of cell i from its unlimited gradients
    double cell_source(u32 i) const { return source_at(i, &s->q[4 * (size_t)i], nu[i]); }
    // the same at the cell state (qi, ni) with the cell's gradients
    double source_at(u32 i, const double* qi, double ni) const {
        const Solver& S = *s;
        const double* gxi = &S.gx[4 * i];
        const double* gyi = &S.gy[4 * i];
        const double rho = qi[0], nut = ni / rho;
        const double dudy = (gyi[1] - qi[1] / rho * gyi[0]) / rho;
        const double dvdx = (gxi[2] - qi[2] / rho * gxi[0]) / rho;
        const double omega = std::fabs(dvdx - dudy);
        const double gvx = (gnx[i] - nut * gxi[0]) / rho, gvy = (gny[i] - nut * gyi[0]) / rho;
        return sa::source(rho, nut, S.g.mu_L, omega, dist[i], gvx, gvy);
    }

    // ---- implicit step of the model.  The reference's SA branch is empty
    // (physics.h:259-261), so this Jacobian is the build's own, made by the
    // reference's rule for the mean flow (calc_convective_jacobian,
    // physics.h:533-578: forward differences, h = max(1e-6, 1e-6 |v|), the
    // perturbed component restored by subtraction, first-order cell states,
    // times the face length) applied to the five-variable face flux.  The
    // cell gradients stay frozen; the face gradients' directional parts
    // (v_n - v_o)/l and the no-slip wall's cell-centre nu~ follow the
    // perturbed states, so the viscous and nu~ diffusion couplings along the
    // cell connection enter the blocks.  The rho nu~ column's step floor is
    // 1e-6 mu_L instead of 1e-6: rho nu~ is of the order mu_L chi, so the
    // absolute floor would perturb it by O(1) wherever chi is O(1).
    // J: 10 x 10 row-major; rows 0-4 the owner's flux, rows 5-9 their
    // negatives; columns 0-4 the owner's variables, 5-9 the neighbour's.
    void face_jacobian5(u32 e, double J[100]) const {
        const Solver& S = *s;
        const u32 c0 = S.m.edgesCells[2 * e], c1 = S.m.edgesCells[2 * e + 1];
        double a[5], b[5], f0[5], fp[5];
        for (int k = 0; k < 4; ++k) { a[k] = S.q[4 * (size_t)c0 + k]; b[k] = S.q[4 * (size_t)c1 + k]; }
        a[4] = nu[c0];
        b[4] = nu[c1];
        auto flux = [&](double* out) { face_flux5(e, a, a[4], b, b[4], out, a, a[4], b, b[4]); };
        flux(f0);
        const double L = S.m.edgesLengths[e];
        for (int i = 0; i < 5; ++i) {
            for (int side = 0; side < 2; ++side) {
                double* v = side ? b : a;
                const double up = std::max(i == 4 ? 1e-6 * S.g.mu_L : 1e-6, std::fabs(v[i]) * 1e-6);
                v[i] += up;
                flux(fp);
                v[i] -= up;
                for (int r = 0; r < 5; ++r) {
                    const double d = (fp[r] - f0[r]) / up;
                    J[10 * r + 5 * side + i] = d * L;
                    J[10 * (r + 5) + 5 * side + i] = -d * L;
                }
            }
        }
    }
    // row 4 of cell i's own block from the source: -A dS/dv_k at frozen
    // gradients (same step rule and rho nu~ floor), the LHS of R_4 = ... + A S.  The diagonal
    // entry is kept only where it is >= 0 (the source's net derivative is
    // destruction): production stays explicit, so the source never lowers the
    // diagonal of the nu~ row.
    void source_row(u32 i, double out[5]) const {
        const Solver& S = *s;
        double v[5];
        for (int k = 0; k < 4; ++k) v[k] = S.q[4 * (size_t)i + k];
        v[4] = nu[i];
        const double A = S.m.cellsAreas[i];
        const double s0 = source_at(i, v, v[4]);
        for (int k = 0; k < 5; ++k) {
            const double up = std::max(k == 4 ? 1e-6 * S.g.mu_L : 1e-6, std::fabs(v[k]) * 1e-6);
            v[k] += up;
            const double sp = source_at(i, v, v[4]);
            v[k] -= up;
            out[k] = -A * (sp - s0) / up;
        }
        if (out[4] < 0) out[4] = 0;
    }
    // the 5N x 5N matrix of the implicit step in the AoS order [5 i + k] as
    // COO triplets (duplicates summed by the caller): A/dt (the model's time
    // step) on the real diagonal, the face blocks in face order as fillRhoLHS
    // places them (solver.h:1013-1056), the source rows, identity ghost rows
    void fill_lhs(std::vector<int64_t>& rows, std::vector<int64_t>& cols, std::vector<double>& vals) {
        const Solver& S = *s;
        const Mesh& m = S.m;
        calc_dt();
        prepare();
        rows.clear(); cols.clear(); vals.clear();
        const int64_t nT = m.nCells(), nR = m.nRealCells;
        auto put = [&](int64_t r, int64_t c, double v) { rows.push_back(r); cols.push_back(c); vals.push_back(v); };
        for (int64_t i = 0; i < nR; ++i)
            for (int j = 0; j < 5; ++j) put(5 * i + j, 5 * i + j, m.cellsAreas[i] / S.dt[i]);
        double J[100];
        for (u32 e = 0; e < m.nFaces(); ++e) {
            const int64_t k0 = 5 * (int64_t)m.edgesCells[2 * e], k1 = 5 * (int64_t)m.edgesCells[2 * e + 1];
            face_jacobian5(e, J);
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j) {
                    put(k0 + i, k0 + j, J[10 * i + j]);
                    put(k0 + i, k1 + j, J[10 * i + j + 5]);
                    if (S.ff[e].two_sided()) {
                        put(k1 + i, k0 + j, J[10 * (i + 5) + j]);
                        put(k1 + i, k1 + j, J[10 * (i + 5) + j + 5]);
                    }
                }
        }
        if (terms & SA_SOURCE) {
            double r[5];
            for (int64_t i = 0; i < nR; ++i) {
                source_row((u32)i, r);
                for (int k = 0; k < 5; ++k) put(5 * i + 4, 5 * i + k, r[k]);
            }
        }
        for (int64_t i = nR; i < nT; ++i)
            for (int j = 0; j < 5; ++j) put(5 * i + j, 5 * i + j, 1.0);
    }
};

// mesh.h:794-830
static void compute_wall_dist(Mesh& m, const std::vector<int>& bface_type) {
    for (u32 i = 0; i < m.nCells(); ++i) {
        double cx = m.cellsCentersX[i], cy = m.cellsCentersY[i];
        double mind = 1;
        bool first = true;
        for (size_t j = 0; j < m.boundaryEdges.size(); ++j) {
            if (bface_type[j] == BC_WALL || bface_type[j] == BC_SLIP) {
                u32 e = m.boundaryEdges[j];
                double dx = cx - m.edgesCentersX[e], dy = cy - m.edgesCentersY[e];
                double d = std::sqrt(dx * dx + dy * dy);
                if (first) { mind = d; first = false; } else mind = std::min(mind, d);
            }
        }
        m.wall_dist[i] = mind;
    }
}

}  // namespace orc

// ============================================================ C interface ===
using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_mesh_read(const char* path) {
    try { auto* m = new Mesh(); m->read_file(path); return m; }
    catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// Build from parsed entities (what read_file produces before add_cell_edges):
// cell_nodes: 4 per cell (triangles: 4th ignored), is_tri per cell;
// bnd_nodes: 2 per boundary segment; bnd_tag: physical key per segment;
// names: '\n'-separated "key=name" pairs.
void* orc_mesh_build(int nN, const double* X, const double* Y, int nC, const int* cell_nodes,
                     const int* is_tri, int nB, const int* bnd_nodes, const int* bnd_tag, const char* names) {
    try {
        auto* m = new Mesh();
        std::stringstream ss(names);
        std::string item;
        while (std::getline(ss, item, '\n')) {
            if (item.empty()) continue;
            auto p = item.find('=');
            m->physicalNames[(u32)std::stoul(item.substr(0, p))] = item.substr(p + 1);
        }
        m->nodesX.assign(X, X + nN); m->nodesY.assign(Y, Y + nN);
        for (int b = 0; b < nB; ++b) {
            m->boundaryEdges0.push_back((u32)bnd_nodes[2 * b]);
            m->boundaryEdges1.push_back((u32)bnd_nodes[2 * b + 1]);
            m->boundaryEdgesIntTag.push_back((u32)bnd_tag[b]);
        }
        for (int c = 0; c < nC; ++c) {
            m->cellsIsTriangle.push_back(is_tri[c] ? 1 : 0);
            for (int k = 0; k < 4; ++k) m->cellsNodes.push_back(is_tri[c] && k == 3 ? 0u : (u32)cell_nodes[4 * c + k]);
            m->cellsAreas.push_back(0.); m->cellsCentersX.push_back(0.); m->cellsCentersY.push_back(0.);
        }
        m->finish();
        return m;
    } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

void orc_mesh_free(void* h) { delete (Mesh*)h; }

// counts: nN, nC(real), nT, nF, nB
void orc_mesh_counts(void* h, int* out) {
    Mesh* m = (Mesh*)h;
    out[0] = (int)m->nodesX.size(); out[1] = (int)m->nRealCells; out[2] = (int)m->nCells();
    out[3] = (int)m->nFaces(); out[4] = (int)m->boundaryEdges.size();
}

// field getter (copies into caller buffer). Returns number of elements or -1.
long orc_mesh_get(void* h, const char* f, void* out) {
    Mesh* m = (Mesh*)h;
    std::string s(f);
    auto cpd = [&](const std::vector<double>& v) { if (out) std::memcpy(out, v.data(), v.size() * 8); return (long)v.size(); };
    auto cpu = [&](const std::vector<u32>& v) { if (out) std::memcpy(out, v.data(), v.size() * 4); return (long)v.size(); };
    if (s == "nodesX") return cpd(m->nodesX);
    if (s == "nodesY") return cpd(m->nodesY);
    if (s == "edgesNodes") return cpu(m->edgesNodes);
    if (s == "edgesCells") return cpu(m->edgesCells);
    if (s == "edgesLengths") return cpd(m->edgesLengths);
    if (s == "edgesNormalsX") return cpd(m->edgesNormalsX);
    if (s == "edgesNormalsY") return cpd(m->edgesNormalsY);
    if (s == "edgesCentersX") return cpd(m->edgesCentersX);
    if (s == "edgesCentersY") return cpd(m->edgesCentersY);
    if (s == "cellsNodes") return cpu(m->cellsNodes);
    if (s == "cellsEdges") return cpu(m->cellsEdges);
    if (s == "cellsAreas") return cpd(m->cellsAreas);
    if (s == "cellsCentersX") return cpd(m->cellsCentersX);
    if (s == "cellsCentersY") return cpd(m->cellsCentersY);
    if (s == "boundaryEdges") return cpu(m->boundaryEdges);
    if (s == "boundaryEdgesIntTag") return cpu(m->boundaryEdgesIntTag);
    if (s == "wall_dist") return cpd(m->wall_dist);
    if (s == "cellsIsTriangle") {
        if (out) for (size_t i = 0; i < m->cellsIsTriangle.size(); ++i) ((int32_t*)out)[i] = m->cellsIsTriangle[i];
        return (long)m->cellsIsTriangle.size();
    }
    return -1;
}

// physical name of boundary face b
const char* orc_mesh_bname(void* h, int b) { return ((Mesh*)h)->boundaryEdgesPhysicals[b].c_str(); }

// ---- solver ----
// viscous: 0 inviscid, 1 laminar, 2 "spallart-allmaras" selected through the
// reference Settings (A14: behaves as inviscid flux + gradients computed).
void* orc_solver_create(void* mh, double R, double mu_L, double Pr_L, double gamma, int viscous) {
    auto* s = new Solver();
    Gas g; g.R = R; g.mu_L = mu_L; g.Pr_L = Pr_L; g.gamma = gamma;
    s->set_mesh_and_gas(*(Mesh*)mh, g);
    s->viscous = viscous == 1 ? 1 : 0;   // 2 (SA): mean-flow viscous terms come from SASolver
    s->viscous_model_named_non_inviscid = viscous != 0;
    return s;
}
void orc_solver_set_prt(void* h, double Pr_T) { ((Solver*)h)->g.Pr_T = Pr_T; }
void orc_solver_set_a1(void* h, int on) { ((Solver*)h)->a1_sum_direction = on != 0; }
// average_gradients of face e from the current gx, gy, q: out = gx[4], gy[4]
void orc_solver_face_gradient(void* h, int e, double* out) {
    Solver* s = (Solver*)h;
    s->average_gradients(out, out + 4, s->m.edgesCells[2 * e], s->m.edgesCells[2 * e + 1]);
}
void orc_solver_free(void* h) { delete (Solver*)h; }

// bcs: nb entries; names '\n'-separated; types BC_*; per-entry farfield vars (mach, angle, T, p)
int orc_solver_set_bcs(void* h, int nb, const char* names, const int* types, const double* vars) {
    Solver* s = (Solver*)h;
    try {
        std::map<std::string, BcSpec> bcs;
        std::stringstream ss(names);
        std::string nm;
        int i = 0;
        while (std::getline(ss, nm, '\n') && i < nb) {
            BcSpec b; b.type = types[i];
            b.vars.mach = vars[4 * i]; b.vars.angle = vars[4 * i + 1]; b.vars.T = vars[4 * i + 2]; b.vars.p = vars[4 * i + 3];
            bcs[nm] = b; ++i;
        }
        s->set_bcs(bcs);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}
void orc_solver_options(void* h, int second_order, int green_gauss, double limiter_k, double cfl) {
    Solver* s = (Solver*)h;
    s->second_order = second_order != 0; s->gg = green_gauss != 0; s->limiter_k = limiter_k; s->cfl = cfl;
    if (!s->gg) s->calc_least_squares_matrices();
}
void orc_solver_init(void* h) { ((Solver*)h)->init(); }
void orc_solver_refill_bcs(void* h) { ((Solver*)h)->refill_bcs(); }
void orc_solver_bcs_from_internal(void* h) { ((Solver*)h)->bcs_from_internal(); }
void orc_solver_set_walls(void* h) { Solver* s = (Solver*)h; s->set_walls_from_internal(s->q); }
void orc_solver_calc_dt(void* h) { ((Solver*)h)->calc_dt(); }
void orc_solver_calc_gradients(void* h) { ((Solver*)h)->calc_gradients(); }
void orc_solver_calc_limiters(void* h) { Solver* s = (Solver*)h; s->calc_limiters(s->q); }
void orc_solver_fill_rhs(void* h) { ((Solver*)h)->fillRhoRHS(); }
void orc_solver_calc_residual(void* h, const double* qk) {
    Solver* s = (Solver*)h;
    std::vector<double> v(qk, qk + s->q.size());
    s->calc_residual(v);
}
double orc_solver_explicit_step(void* h, double relax) { return ((Solver*)h)->explicit_solve(relax); }

// ---- Spalart-Allmaras (the model of namespace sa) ----
void* orc_sa_create(void* solver_h) { return new SASolver((Solver*)solver_h); }
void orc_sa_free(void* h) { delete (SASolver*)h; }
void orc_sa_set_terms(void* h, int terms) { ((SASolver*)h)->terms = terms; }
void orc_sa_calc_dt(void* h) { ((SASolver*)h)->calc_dt(); }
static void sa_wall_dist(SASolver* a);
void orc_sa_fill_rhs(void* h) {
    SASolver* a = (SASolver*)h;
    sa_wall_dist(a);
    a->fill_rhs();
}
long orc_sa_vec(void* h, const char* name, double* out, const double* in) {
    SASolver* a = (SASolver*)h;
    std::string n(name);
    std::vector<double>* v = nullptr;
    if (n == "dist") sa_wall_dist(a);
    if (n == "nu") v = &a->nu; else if (n == "TurbVector") v = &a->TurbVector; else if (n == "gnx") v = &a->gnx;
    else if (n == "dist") v = &a->dist;
    else if (n == "gny") v = &a->gny; else if (n == "lnu") v = &a->lnu;
    if (!v) return -1;
    if (in) std::memcpy(v->data(), in, v->size() * 8);
    if (out) std::memcpy(out, v->data(), v->size() * 8);
    return (long)v->size();
}
static void sa_wall_dist(SASolver* a) {
    if (a->dist.empty()) model_wall_dist(a->s->m, a->s->bface_type, a->dist);
}
void orc_sa_prepare(void* h) { sa_wall_dist((SASolver*)h); ((SASolver*)h)->prepare(); }
void orc_sa_face_jacobian(void* h, int e, double* J) { ((SASolver*)h)->face_jacobian5((u32)e, J); }
void orc_sa_source_row(void* h, int i, double* out) { ((SASolver*)h)->source_row((u32)i, out); }
double orc_sa_cell_source(void* h, int i, const double* q5) {
    SASolver* a = (SASolver*)h;
    return q5 ? a->source_at((u32)i, q5, q5[4]) : a->cell_source((u32)i);
}
// the five-variable face flux at face states qL5, qR5 with centre states
// zL5, zR5 (null: the member states)
void orc_sa_face_flux(void* h, int e, const double* qL5, const double* qR5, const double* zL5, const double* zR5,
                      double* f) {
    SASolver* a = (SASolver*)h;
    a->face_flux5((u32)e, qL5, qL5[4], qR5, qR5[4], f, zL5, zL5 ? zL5[4] : 0.0, zR5, zR5 ? zR5[4] : 0.0);
}
double orc_sa_source(double rho, double nut, double mu, double omega, double d, double gnx, double gny) {
    return sa::source(rho, nut, mu, omega, d, gnx, gny);
}
double orc_sa_fw(double r) { return sa::fw(r); }
double orc_sa_fv1(double chi) { return sa::fv1(chi); }
double orc_solver_uniform_residual(void* h) { return ((Solver*)h)->get_uniform_residual(); }
void orc_solver_wall_profile(void* h, const char* patch, double* out3) {
    ((Solver*)h)->wall_profile(patch, &out3[0], &out3[1], &out3[2]);
}
void orc_solver_face_jacobian(void* h, int e, double* J) { ((Solver*)h)->face_jacobian((u32)e, J); }
void orc_solver_mf_apply(void* h, const double* x, double* y) { ((Solver*)h)->mf_apply(x, y); }
void orc_solver_set_michalak(void* h, int on) { ((Solver*)h)->michalak = on != 0; }

// multigrid<T>::gen_mapper (multigrid.h:96-163) on centroid / area arrays in
// reference order (ghosts included): first the row scales over all coarse
// cells in order, then the weights s_ij / scales_i.  CSR out; call with null
// buffers to get nnz.
static thread_local std::vector<int64_t> g_mrow;
static thread_local std::vector<int32_t> g_mcol;
static thread_local std::vector<double> g_mval;
long orc_gen_mapper(int nf, const double* fx, const double* fy, int nc, const double* cx, const double* cy,
                    const double* ca, int64_t* row_ptr, int32_t* cols, double* vals) {
    if (row_ptr) {
        std::memcpy(row_ptr, g_mrow.data(), g_mrow.size() * 8);
        std::memcpy(cols, g_mcol.data(), g_mcol.size() * 4);
        std::memcpy(vals, g_mval.data(), g_mval.size() * 8);
        return (long)g_mval.size();
    }
    std::vector<double> scales(nf);
    for (int i = 0; i < nf; ++i) {
        const double xi = fx[i], yi = fy[i];
        scales[i] = 0;
        for (int j = 0; j < nc; ++j) {
            const double d2 = (xi - cx[j]) * (xi - cx[j]) + (yi - cy[j]) * (yi - cy[j]);
            const double r2 = ca[j];
            if (d2 < 2 * r2) scales[i] += 1 / std::max(0.1 * std::sqrt(r2), std::sqrt(d2));
        }
    }
    g_mrow.assign(1, 0); g_mcol.clear(); g_mval.clear();
    for (int i = 0; i < nf; ++i) {
        const double xi = fx[i], yi = fy[i];
        for (int j = 0; j < nc; ++j) {
            const double d2 = (xi - cx[j]) * (xi - cx[j]) + (yi - cy[j]) * (yi - cy[j]);
            const double r2 = ca[j];
            if (d2 < 2 * r2) {
                const double si = 1 / std::max(0.1 * std::sqrt(r2), std::sqrt(d2));
                g_mcol.push_back(j);
                g_mval.push_back(si / scales[i]);
            }
        }
        g_mrow.push_back((int64_t)g_mval.size());
    }
    return (long)g_mval.size();
}

// fillRhoLHS triplets: returns the count; call with null buffers to size
static thread_local std::vector<int64_t> g_rows, g_cols;
static thread_local std::vector<double> g_vals;
long orc_solver_fill_lhs(void* h, int64_t* rows, int64_t* cols, double* vals) {
    if (!rows) { ((Solver*)h)->fill_lhs(g_rows, g_cols, g_vals); return (long)g_vals.size(); }
    std::memcpy(rows, g_rows.data(), g_rows.size() * 8);
    std::memcpy(cols, g_cols.data(), g_cols.size() * 8);
    std::memcpy(vals, g_vals.data(), g_vals.size() * 8);
    return (long)g_vals.size();
}
void orc_solver_wall_dist(void* h) { Solver* s = (Solver*)h; compute_wall_dist(s->m, s->bface_type); }
// the implicit step's matrix of the SA model, as orc_solver_fill_lhs
long orc_sa_fill_lhs(void* h, int64_t* rows, int64_t* cols, double* vals) {
    if (!rows) {
        sa_wall_dist((SASolver*)h);
        ((SASolver*)h)->fill_lhs(g_rows, g_cols, g_vals);
        return (long)g_vals.size();
    }
    std::memcpy(rows, g_rows.data(), g_rows.size() * 8);
    std::memcpy(cols, g_cols.data(), g_cols.size() * 8);
    std::memcpy(vals, g_vals.data(), g_vals.size() * 8);
    return (long)g_vals.size();
}
// CpProfile::calc_cp (post.h:246-290): per boundary face of the patch, in
// boundary order, the face centre and cp = 2/(gamma M_inf^2) (p/p_inf - 1)
// of the owner cell (sol.p, core.h:145-147); returns the face count (arrays
// may be null to count only)
int orc_cp_profile(void* h, const char* patch, double* x, double* y, double* cp) {
    const Solver& S = *(const Solver*)h;
    const Mesh& m = S.m;
    const BVars far = S.get_boundary_variables();
    int k = 0;
    for (size_t b = 0; b < m.boundaryEdges.size(); ++b) {
        if (m.boundaryEdgesPhysicals[b] != patch) continue;
        const u32 e = m.boundaryEdges[b];
        const u32 c = m.edgesCells[2 * e];
        const double* q = &S.q[4 * c];
        const double p = (S.g.gamma - 1) * (q[3] - 0.5 / q[0] * (q[1] * q[1] + q[2] * q[2]));
        if (x) {
            x[k] = m.edgesCentersX[e];
            y[k] = m.edgesCentersY[e];
            cp[k] = 2. / (S.g.gamma * far.mach * far.mach) * (p / far.p - 1.);
        }
        ++k;
    }
    return k;
}
// save (post.h:58-180): the reference's ASCII VTU of the real cells, with its
// number formats -- std::to_string (%f) for the node coordinates and indices,
// std::fixed with precision 16 (double2string, post.h:32-36) for cell data --
// and the solution accessors of core.h:132-163.  Returns 0, or 1 when the
// file cannot be written.
int orc_save_vtu(void* h, const char* path) {
    const Solver& S = *(const Solver*)h;
    const Mesh& m = S.m;
    const std::vector<double>& q = S.q;
    const double gm = S.g.gamma, R = S.g.R;
    auto fix16 = [](double x) { std::ostringstream o; o << std::fixed << std::setprecision(16) << x; return o.str(); };
    auto rho = [&](u32 i) { return q[4 * i]; };
    auto p = [&](u32 i) { return (gm - 1) * (q[4 * i + 3] - 0.5 / rho(i) * (q[4 * i + 1] * q[4 * i + 1] + q[4 * i + 2] * q[4 * i + 2])); };
    auto c = [&](u32 i) { return std::sqrt(gm * p(i) / rho(i)); };
    std::ostringstream o;
    o << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"BigEndian\">\n"
      << "  <UnstructuredGrid>\n"
      << "    <Piece NumberOfPoints=\"" << std::to_string(m.nodesX.size()) << "\" NumberOfCells=\""
      << std::to_string(m.nRealCells) << "\">\n"
      << "      <Points>\n"
      << "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
    for (size_t i = 0; i < m.nodesX.size(); ++i)
        o << "          " << std::to_string(m.nodesX[i]) << " " << std::to_string(m.nodesY[i]) << " 0.0\n";
    o << "        </DataArray>\n" << "      </Points>\n" << "      <Cells>\n"
      << "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n" << "          ";
    for (u32 i = 0; i < m.nRealCells; ++i) {
        const u32 nv = m.cellsIsTriangle[i] ? 3 : 4;
        for (u32 k = 0; k < nv; ++k) o << std::to_string(m.cellsNodes[4 * i + k]) << "  ";
        o << "\n          ";
    }
    o << "        </DataArray>\n" << "        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n" << "          ";
    u32 off = 0;
    for (u32 i = 0; i < m.nRealCells; ++i) {
        off += m.cellsIsTriangle[i] ? 3 : 4;
        o << std::to_string(off) << "  ";
    }
    o << "\n" << "        </DataArray>\n" << "        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n" << "          ";
    for (u32 i = 0; i < m.nRealCells; ++i) o << (m.cellsIsTriangle[i] ? "5  " : "9  ");
    o << "\n" << "        </DataArray>\n" << "      </Cells>\n" << "      <CellData Scalars=\"scalars\">\n";
    auto block = [&](const char* name, auto f) {
        o << "        <DataArray type=\"Float32\" Name=\"" << name << "\" Format=\"ascii\">\n";
        for (u32 i = 0; i < m.nRealCells; ++i) o << "          " << fix16(f(i)) << "\n";
        o << "        </DataArray>\n";
    };
    block("Wall Distance", [&](u32 i) { return m.wall_dist[i]; });
    block("Mach", [&](u32 i) { return std::sqrt(q[4 * i + 1] * q[4 * i + 1] + q[4 * i + 2] * q[4 * i + 2]) / (rho(i) * c(i)); });
    block("Density", [&](u32 i) { return rho(i); });
    block("Pressure", [&](u32 i) { return p(i); });
    block("Temperature", [&](u32 i) { return p(i) / (R * rho(i)); });
    o << "        <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
    for (u32 i = 0; i < m.nRealCells; ++i)
        o << "          " << fix16(q[4 * i + 1] / rho(i)) << " " << fix16(q[4 * i + 2] / rho(i)) << " 0.0\n";
    o << "        </DataArray>\n" << "      </CellData>\n" << "    </Piece>\n" << "  </UnstructuredGrid>\n" << "</VTKFile>";
    std::ofstream f(path, std::ios::binary);
    if (!f) return 1;
    const std::string txt = o.str();
    f.write(txt.data(), (std::streamsize)txt.size());
    return f ? 0 : 1;
}

// vector access: name in {q, qW, dt, gx, gy, limiters, q_min, q_max, RhoVector}
long orc_solver_vec(void* h, const char* name, double* out, const double* in) {
    Solver* s = (Solver*)h;
    std::string n(name);
    std::vector<double>* v = nullptr;
    if (n == "q") v = &s->q; else if (n == "qW") v = &s->qW; else if (n == "dt") v = &s->dt;
    else if (n == "gx") v = &s->gx; else if (n == "gy") v = &s->gy; else if (n == "limiters") v = &s->limiters;
    else if (n == "q_min") v = &s->q_min; else if (n == "q_max") v = &s->q_max; else if (n == "RhoVector") v = &s->RhoVector;
    else if (n == "wall_dist") v = &s->m.wall_dist;
    if (!v) return -1;
    if (in) std::memcpy(v->data(), in, v->size() * 8);
    if (out) std::memcpy(out, v->data(), v->size() * 8);
    return (long)v->size();
}

// single-face flux evaluation (tests): kind BC_*, viscous 0/1
void orc_face_flux(int kind, int viscous, double nx, double ny, double gamma, double R, double mu, double Pr,
                   const double* qL, const double* qR, const double* gx, const double* gy, double* f) {
    Gas g; g.gamma = gamma; g.R = R; g.mu_L = mu; g.Pr_L = Pr;
    FaceFlux F; F.g = &g; F.nx = nx; F.ny = ny; F.kind = kind; F.viscous = viscous;
    face_flux(F, qL, qR, gx, gy, f);
}
void orc_face_vars(int kind, double nx, double ny, double gamma, double R, const double* qL, const double* qbc, double* qR) {
    Gas g; g.gamma = gamma; g.R = R;
    FaceFlux F; F.g = &g; F.nx = nx; F.ny = ny; F.kind = kind; F.viscous = 0;
    face_vars(F, qL, qbc, qR);
}
void orc_freestream(double gamma, double R, double mach, double angle, double T, double p, double* out4) {
    Gas g; g.gamma = gamma; g.R = R;
    BVars v; v.mach = mach; v.angle = angle; v.T = T; v.p = p;
    v.conservative(g, out4);
}

}  // extern "C"

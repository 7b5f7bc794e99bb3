// Host side of the vortex-lattice / beam coupling path (include/afx_vlm.h):
// wing lattice generator, Euler-Bernoulli frame model with a banded
// Cholesky solve, load and displacement transfer and the coupling loop.
// The beam system is 6 dofs per node with a half bandwidth of 11 for a
// chain of elements (1542 dofs for 256 elements): its factorisation costs
// microseconds on the host, so it stays there; every vortex-lattice
// operation goes through the device entry points of afx_vlm.hip.
#include <algorithm>
#include <array>
#include <cmath>
#include <exception>
#include <string>
#include <vector>

#include "../../include/afx_vlm.h"
#include "afx_vlm.hpp"

using afx::api_fail;

namespace {
using V3 = std::array<double, 3>;
V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double norm(const V3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
V3 scale(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
}  // namespace

// ================================================================ beam ===
struct afx_beam {
    int n = 0, ne = 0, bw = 0;
    std::vector<V3> x;
    std::vector<int> en;          // 2 per element
    std::vector<double> props;    // 7 per element
    std::vector<char> fixed;      // per node
    std::vector<double> L;        // banded Cholesky factor, row i: L[i * (bw + 1) + (i - j)], j in [i - bw, i]
};

namespace {

// CBAR::set_K_local (element.hpp:77-109): 12 x 12 frame stiffness in the
// element axes, dofs (u, v, w, rx, ry, rz) of node 1 then node 2; bending in
// the local x-y plane uses EIz, in the x-z plane EIy.
void frame_local(double L, double EA, double EIy, double EIz, double GJ, double K[12][12]) {
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) K[i][j] = 0.0;
    const double X = EA / L, S = GJ / L;
    const double Y1 = 12.0 * EIz / (L * L * L), Y2 = 6.0 * EIz / (L * L), Y3 = 4.0 * EIz / L, Y4 = 2.0 * EIz / L;
    const double Z1 = 12.0 * EIy / (L * L * L), Z2 = 6.0 * EIy / (L * L), Z3 = 4.0 * EIy / L, Z4 = 2.0 * EIy / L;
    auto put = [&](int i, int j, double v) { K[i][j] = v; K[j][i] = v; };
    // axial and torsion: [k -k; -k k]
    put(0, 0, X); put(6, 6, X); put(0, 6, -X);
    put(3, 3, S); put(9, 9, S); put(3, 9, -S);
    // v / rz (x-y plane)
    put(1, 1, Y1); put(7, 7, Y1); put(1, 7, -Y1);
    put(1, 5, Y2); put(1, 11, Y2); put(5, 7, -Y2); put(7, 11, -Y2);
    put(5, 5, Y3); put(11, 11, Y3); put(5, 11, Y4);
    // w / ry (x-z plane): the rotation about y opposes dw/dx
    put(2, 2, Z1); put(8, 8, Z1); put(2, 8, -Z1);
    put(2, 4, -Z2); put(2, 10, -Z2); put(4, 8, Z2); put(8, 10, Z2);
    put(4, 4, Z3); put(10, 10, Z3); put(4, 10, Z4);
}

// CBAR::set_T_Rotation (element.hpp:111-150): rows x (node 1 -> node 2),
// y = -x cross z, z = unit(x cross V)
bool frame_axes(const V3& a, const V3& b, const V3& v, double lam[3][3], double& L) {
    V3 d = sub(b, a);
    L = norm(d);
    if (!(L > 0.0)) return false;
    V3 ex = scale(d, 1.0 / L);
    V3 ez = cross(ex, v);
    const double nz = norm(ez);
    if (!(nz > 0.0)) return false;
    ez = scale(ez, 1.0 / nz);
    V3 ey = cross(ex, ez);
    ey = scale(ey, -1.0);
    for (int k = 0; k < 3; ++k) { lam[0][k] = ex[k]; lam[1][k] = ey[k]; lam[2][k] = ez[k]; }
    return true;
}

// Assemble the banded global stiffness and factor it into Lout.  rot (nullable):
// per element the 3 x 3 rotation from element to global axes (row-major), as
// MODEL::set_K_global takes it from the element's current mid-node quaternion
// (model.hpp:103-131); null: the initial axes, Lambda^T.
int beam_assemble(const afx_beam* b, const std::vector<std::array<double, 9>>* rot, std::vector<double>& Lout) {
    const int nd = 6 * b->n, w = b->bw + 1;
    std::vector<double> K((size_t)nd * w, 0.0);   // lower band, same indexing as L
    auto addK = [&](int i, int j, double v) {
        if (j > i) return;   // symmetric: keep the lower band
        K[(size_t)i * w + (i - j)] += v;
    };
    for (int e = 0; e < b->ne; ++e) {
        const int n1 = b->en[2 * e], n2 = b->en[2 * e + 1];
        const double* p = &b->props[7 * e];
        double lam[3][3], L;
        if (!frame_axes(b->x[n1], b->x[n2], {p[4], p[5], p[6]}, lam, L))
            return api_fail(AFX_EARG, "beam element " + std::to_string(e) +
                                          ": zero length or orientation vector along the element");
        if (rot)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) lam[r][c] = (*rot)[e][3 * c + r];   // global -> element = R^T
        double Kl[12][12], T[12][12] = {}, Kg[12][12];
        frame_local(L, p[0], p[1], p[2], p[3], Kl);
        for (int blk = 0; blk < 4; ++blk)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) T[3 * blk + r][3 * blk + c] = lam[r][c];
        // Kg = T^T Kl T
        double KT[12][12];
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += Kl[i][k] * T[k][j];
                KT[i][j] = s;
            }
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += T[k][i] * KT[k][j];
                Kg[i][j] = s;
            }
        const int dof[2] = {6 * n1, 6 * n2};
        for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c)
                for (int i = 0; i < 6; ++i)
                    for (int j = 0; j < 6; ++j) addK(dof[a] + i, dof[c] + j, Kg[6 * a + i][6 * c + j]);
    }
    // SPC1 on clamped nodes: rows and columns zeroed, 1 on the diagonal
    for (int i = 0; i < nd; ++i)
        for (int j = std::max(0, i - b->bw); j <= i; ++j) {
            const bool fi = b->fixed[i / 6], fj = b->fixed[j / 6];
            if (fi || fj) K[(size_t)i * w + (i - j)] = (i == j) ? 1.0 : 0.0;
        }
    // banded Cholesky K = L L^T
    Lout.assign(K.size(), 0.0);
    auto Lr = [&](int i, int j) -> double& { return Lout[(size_t)i * w + (i - j)]; };
    for (int i = 0; i < nd; ++i) {
        const int j0 = std::max(0, i - b->bw);
        for (int j = j0; j <= i; ++j) {
            double s = K[(size_t)i * w + (i - j)];
            for (int k = std::max(j0, j - b->bw); k < j; ++k) s -= Lr(i, k) * Lr(j, k);
            if (i == j) {
                if (!(s > 0.0))
                    return api_fail(AFX_EARG, "beam stiffness is not positive definite at dof " + std::to_string(i) +
                                                  " (free mechanism: clamp a node)");
                Lr(i, i) = std::sqrt(s);
            } else {
                Lr(i, j) = s / Lr(j, j);
            }
        }
    }
    return AFX_OK;
}

void beam_solve_with(const afx_beam* b, const std::vector<double>& Lf, const double* f, double* u) {
    const int nd = 6 * b->n, w = b->bw + 1;
    auto L = [&](int i, int j) { return Lf[(size_t)i * w + (i - j)]; };
    std::vector<double> y(nd);
    for (int i = 0; i < nd; ++i) {
        double s = b->fixed[i / 6] ? 0.0 : f[i];   // apply_SPC1_Forces
        for (int k = std::max(0, i - b->bw); k < i; ++k) s -= L(i, k) * y[k];
        y[i] = s / L(i, i);
    }
    for (int i = nd - 1; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k <= std::min(nd - 1, i + b->bw); ++k) s -= L(k, i) * u[k];
        u[i] = s / L(i, i);
    }
}

void beam_solve_impl(const afx_beam* b, const double* f, double* u) { beam_solve_with(b, b->L, f, u); }

}  // namespace

// ========================================================= coupling =====
struct afx_aeroelastic {
    afx_vlm* v = nullptr;
    afx_beam* b = nullptr;
    int nn = 0, nr = 0, ns = 0;
    std::vector<double> x0;                 // undeformed lattice nodes
    std::vector<int> rn;                    // ring corner ids
    std::vector<int> na, nb;                // node -> beam nodes
    std::vector<double> nw;                 // weight of na
    std::vector<V3> fp0;                    // undeformed ring force points
    std::vector<int> ra, rb;
    std::vector<double> rw;
    std::vector<double> disp;               // beam displacement, 6 per node
    int tip = 0;
    bool have_factor = false;
};

namespace {

// the beam element bracketing y (first found), or the nearest beam node
void locate(const afx_beam* b, double y, int& a, int& c, double& wa) {
    for (int e = 0; e < b->ne; ++e) {
        const int n1 = b->en[2 * e], n2 = b->en[2 * e + 1];
        const double y1 = b->x[n1][1], y2 = b->x[n2][1];
        if (y1 == y2) continue;
        const double lo = std::min(y1, y2), hi = std::max(y1, y2);
        if (y >= lo && y <= hi) {
            a = n1;
            c = n2;
            wa = (y2 - y) / (y2 - y1);
            return;
        }
    }
    int best = 0;
    for (int k = 1; k < b->n; ++k)
        if (std::fabs(b->x[k][1] - y) < std::fabs(b->x[best][1] - y)) best = k;
    a = c = best;
    wa = 1.0;
}

// u(p) = sum_a w_a (u_a + r_a x (p - x_a))
V3 follow(const afx_beam* b, const std::vector<double>& disp, int a, int c, double wa, const V3& p) {
    V3 u{0.0, 0.0, 0.0};
    const int nodes[2] = {a, c};
    const double ws[2] = {wa, 1.0 - wa};
    for (int q = 0; q < 2; ++q) {
        if (ws[q] == 0.0) continue;
        const double* d = &disp[6 * nodes[q]];
        V3 rot{d[3], d[4], d[5]};
        V3 arm = sub(p, b->x[nodes[q]]);
        V3 ra = cross(rot, arm);
        for (int k = 0; k < 3; ++k) u[k] += ws[q] * (d[k] + ra[k]);
    }
    return u;
}

}  // namespace

extern "C" {

// ----------------------------------------------------------- lattice ----
int afx_wing_lattice(const afx_wing_spec* sp, int32_t* n_nodes, int32_t* n_rings, int32_t* n_stations, double* xyz,
                     int32_t* ring_nodes, int32_t* soff, double* ref) {
    if (!sp || !n_nodes || !n_rings || !n_stations) return api_fail(AFX_EARG, "null argument");
    const int nc = sp->n_chord, ns = sp->n_span;
    if (nc < 1 || ns < 2 || (ns & 1)) return api_fail(AFX_EARG, "need n_chord >= 1 and an even n_span >= 2");
    if (!(sp->aspect_ratio > 0.0) || !(sp->taper > 0.0) || !(sp->root_chord > 0.0))
        return api_fail(AFX_EARG, "aspect ratio, taper and root chord must be positive");
    *n_nodes = (nc + 1) * (ns + 1);
    *n_rings = nc * ns;
    *n_stations = ns;
    const double cr = sp->root_chord, lam = sp->taper;
    const double span = sp->aspect_ratio * cr * (1.0 + lam) / 2.0, semi = 0.5 * span;
    const double sref = span * cr * (1.0 + lam) / 2.0;
    const double mac = 2.0 / 3.0 * cr * (1.0 + lam + lam * lam) / (1.0 + lam);
    if (ref) { ref[0] = sref; ref[1] = mac; ref[2] = span; }
    if (!xyz && !ring_nodes && !soff) return AFX_OK;
    const double tq = std::tan(sp->sweep_deg * M_PI / 180.0);
    std::vector<double> xi(nc + 1), eta(nc + 1);
    for (int i = 0; i <= nc; ++i)
        xi[i] = sp->cosine ? 0.5 * (1.0 - std::cos(M_PI * i / nc)) : (double)i / nc;
    for (int i = 0; i < nc; ++i) eta[i] = xi[i] + 0.25 * (xi[i + 1] - xi[i]);
    eta[nc] = 1.0 + 0.25 * (xi[nc] - xi[nc - 1]);
    if (xyz)
        for (int j = 0; j <= ns; ++j) {
            double y = sp->cosine ? -semi * std::cos(M_PI * j / ns) : -semi + span * j / ns;
            if (2 * j == ns) y = 0.0;
            const double c = cr * (1.0 - (1.0 - lam) * std::fabs(y) / semi);
            const double xle = std::fabs(y) * tq + 0.25 * cr - 0.25 * c;
            for (int i = 0; i <= nc; ++i) {
                double* p = xyz + 3 * ((size_t)j * (nc + 1) + i);
                p[0] = xle + eta[i] * c;
                p[1] = y;
                p[2] = 0.0;
            }
        }
    if (ring_nodes)
        for (int j = 0; j < ns; ++j)
            for (int i = 0; i < nc; ++i) {
                int32_t* r = ring_nodes + 4 * ((size_t)j * nc + i);
                r[0] = j * (nc + 1) + i;             // leading edge, station line j
                r[1] = (j + 1) * (nc + 1) + i;       // leading edge, line j + 1
                r[2] = (j + 1) * (nc + 1) + i + 1;   // trailing edge, line j + 1
                r[3] = j * (nc + 1) + i + 1;         // trailing edge, line j
            }
    if (soff)
        for (int s = 0; s <= ns; ++s) soff[s] = s * nc;
    return AFX_OK;
}

// -------------------------------------------------------------- beam ----
int afx_beam_create(int32_t n_nodes, const double* xyz, int32_t n_elems, const int32_t* en, const double* props,
                    int32_t n_clamped, const int32_t* clamped, afx_beam** out) {
    if (!xyz || !en || !props || !out || (n_clamped > 0 && !clamped)) return api_fail(AFX_EARG, "null argument");
    if (n_nodes < 2 || n_elems < 1) return api_fail(AFX_EARG, "need at least one element");
    try {
        auto* b = new afx_beam();
        b->n = n_nodes;
        b->ne = n_elems;
        b->x.resize(n_nodes);
        for (int i = 0; i < n_nodes; ++i) b->x[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        b->en.assign(en, en + 2 * (size_t)n_elems);
        b->props.assign(props, props + 7 * (size_t)n_elems);
        b->fixed.assign(n_nodes, 0);
        int bw = 5;
        for (int e = 0; e < n_elems; ++e) {
            const int n1 = en[2 * e], n2 = en[2 * e + 1];
            if (n1 < 0 || n1 >= n_nodes || n2 < 0 || n2 >= n_nodes || n1 == n2) {
                delete b;
                return api_fail(AFX_EARG, "element " + std::to_string(e) + " has invalid node ids");
            }
            for (int k = 0; k < 4; ++k)
                if (!(props[7 * e + k] > 0.0)) {
                    delete b;
                    return api_fail(AFX_EARG, "element " + std::to_string(e) + ": EA, EIy, EIz and GJ must be > 0");
                }
            bw = std::max(bw, 6 * std::abs(n1 - n2) + 5);
        }
        b->bw = bw;
        for (int k = 0; k < n_clamped; ++k) {
            if (clamped[k] < 0 || clamped[k] >= n_nodes) {
                delete b;
                return api_fail(AFX_EARG, "clamped node id out of range");
            }
            b->fixed[clamped[k]] = 1;
        }
        int rc = beam_assemble(b, nullptr, b->L);
        if (rc) { delete b; return rc; }
        *out = b;
        return AFX_OK;
    } catch (std::exception& e) {
        return api_fail(AFX_EARG, e.what());
    }
}

int afx_beam_destroy(afx_beam* b) {
    delete b;
    return AFX_OK;
}

int afx_beam_solve(afx_beam* b, const double* f, double* u) {
    if (!b || !f || !u) return api_fail(AFX_EARG, "null argument");
    beam_solve_impl(b, f, u);
    return AFX_OK;
}

// ------------------------------------------------- nonlinear beam ----
// MODEL::get_NonLin_Solve (model.hpp:240-339) with the corotational CBAR of
// element.hpp:111-370.  Quaternions are (w, x, y, z) with Eigen's Hamilton
// product; rotation matrices and the conversions are those of the element.
namespace {
struct Quat { double w, x, y, z; };

Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
Quat qconj(const Quat& a) { return {a.w, -a.x, -a.y, -a.z}; }

// get_Rotation_Matrix_From_Quaternion (element.hpp:224-237), row-major
std::array<double, 9> qmat(const Quat& q) {
    const double s = q.w, x = q.x, y = q.y, z = q.z;
    return {1. - 2. * (y * y + z * z), 2. * (x * y - s * z), 2. * (x * z + s * y),
            2. * (x * y + s * z), 1. - 2. * (x * x + z * z), 2. * (y * z - s * x),
            2. * (x * z - s * y), 2. * (y * z + s * x), 1. - 2. * (x * x + y * y)};
}

// set_qe_From_Rotation_Matrix (element.hpp:159-211) of the element axes Lambda
Quat q_from_axes(const double m[3][3]) {
    const double tr = m[0][0] + m[1][1] + m[2][2];
    double c, s, vx, vy, vz;
    if (tr > 0.0) {
        c = 0.5 / std::sqrt(tr + 1.0);
        s = 0.25 / c;
        vx = (m[1][2] - m[2][1]) * c; vy = (m[2][0] - m[0][2]) * c; vz = (m[0][1] - m[1][0]) * c;
    } else if (m[0][0] > m[1][1] && m[0][0] > m[2][2]) {
        c = 2.0 * std::sqrt(1.0 + m[0][0] - m[1][1] - m[2][2]);
        s = (m[1][2] - m[2][1]) / c; vx = 0.25 * c; vy = (m[1][0] + m[0][1]) / c; vz = (m[2][0] + m[0][2]) / c;
    } else if (m[1][1] > m[2][2]) {
        c = 2.0 * std::sqrt(1.0 + m[1][1] - m[0][0] - m[2][2]);
        s = (m[2][0] - m[0][2]) / c; vx = (m[1][0] + m[0][1]) / c; vy = 0.25 * c; vz = (m[2][1] + m[1][2]) / c;
    } else {
        c = 2.0 * std::sqrt(1.0 + m[2][2] - m[0][0] - m[1][1]);
        s = (m[0][1] - m[1][0]) / c; vx = (m[2][0] + m[0][2]) / c; vy = (m[2][1] + m[1][2]) / c; vz = 0.25 * c;
    }
    return {s, vx, vy, vz};
}

// set_Quaternion_Map (model.hpp:341-366): increment rotations read as roll, pitch, yaw
Quat q_from_increment(double rx, double ry, double rz) {
    const double cr = std::cos(rx * 0.5), sr = std::sin(rx * 0.5);
    const double cp = std::cos(ry * 0.5), sp = std::sin(ry * 0.5);
    const double cy = std::cos(rz * 0.5), sy = std::sin(rz * 0.5);
    return {cr * cp * cy + sr * sp * sy, sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy,
            cr * cp * sy - sr * sp * cy};
}

// set_qmid_From_Interpolation (element.hpp:239-266): the half-way slerp
Quat q_half(const Quat& a, const Quat& b) {
    if (std::fabs(a.w - b.w) < 1e-12 && std::fabs(a.x - b.x) < 1e-12 && std::fabs(a.y - b.y) < 1e-12 &&
        std::fabs(a.z - b.z) < 1e-12)
        return a;
    const double na = std::sqrt(a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z);
    const double nb = std::sqrt(b.w * b.w + b.x * b.x + b.y * b.y + b.z * b.z);
    const double th = std::acos((a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z) / (na * nb));
    if (!(std::fabs(th) > 1e-14)) return a;
    const double c = std::sin(0.5 * th) / std::sin(th);
    return {c * a.w + c * b.w, c * a.x + c * b.x, c * a.y + c * b.y, c * a.z + c * b.z};
}

// get_Euler_Angles_From_Local_Rotation (element.hpp:273-287)
V3 euler_of(const Quat& q) {
    const double s = q.w, a = q.x, b = q.y, c = q.z;
    return {std::atan2(2 * s * a + 2 * b * c, 1 - 2 * (a * a + b * b)), std::asin(2 * s * b - 2 * c * a),
            std::atan2(2 * s * c + 2 * a * b, 1 - 2 * (b * b + c * c))};
}

V3 matvec(const std::array<double, 9>& R, const V3& v) {
    return {R[0] * v[0] + R[1] * v[1] + R[2] * v[2], R[3] * v[0] + R[4] * v[1] + R[5] * v[2],
            R[6] * v[0] + R[7] * v[1] + R[8] * v[2]};
}
V3 matTvec(const std::array<double, 9>& R, const V3& v) {
    return {R[0] * v[0] + R[3] * v[1] + R[6] * v[2], R[1] * v[0] + R[4] * v[1] + R[7] * v[2],
            R[2] * v[0] + R[5] * v[1] + R[8] * v[2]};
}

struct CoElem {
    int n1, n2;
    double Kl[12][12];
    V3 r1, r2;                  // node - midpoint in the initial configuration
    Quat qe, q1, q2, qmid;
    double u1[6], u2[6];        // accumulated global displacements of the end nodes
};

// CBAR::get_Force_In_GlobalRef (element.hpp:316-367)
void element_force(const CoElem& E, double F[12]) {
    const Quat q1p = qmul(qconj(E.qmid), E.q1), q2p = qmul(qconj(E.qmid), E.q2);
    const V3 t1 = euler_of(q1p), t2 = euler_of(q2p);
    const auto Rec = qmat(qmul(E.qmid, qconj(E.qe)));
    const auto Rgc = qmat(E.qmid);
    V3 dr1 = matvec(Rec, E.r1), dr2 = matvec(Rec, E.r2);
    for (int k = 0; k < 3; ++k) { dr1[k] -= E.r1[k]; dr2[k] -= E.r2[k]; }
    V3 a1, a2;
    for (int k = 0; k < 3; ++k) {
        const double um = 0.5 * (E.u1[k] + E.u2[k]);
        a1[k] = E.u1[k] - um - dr1[k];
        a2[k] = E.u2[k] - um - dr2[k];
    }
    const V3 d1 = matTvec(Rgc, a1), d2 = matTvec(Rgc, a2);
    const double D[12] = {d1[0], d1[1], d1[2], t1[0], t1[1], t1[2], d2[0], d2[1], d2[2], t2[0], t2[1], t2[2]};
    double Fl[12];
    for (int i = 0; i < 12; ++i) {
        double s = 0.0;
        for (int j = 0; j < 12; ++j) s += E.Kl[i][j] * D[j];
        Fl[i] = s;
    }
    for (int blk = 0; blk < 4; ++blk) {
        const V3 g = matvec(Rgc, {Fl[3 * blk], Fl[3 * blk + 1], Fl[3 * blk + 2]});
        for (int k = 0; k < 3; ++k) F[3 * blk + k] = g[k];
    }
}
}  // namespace

int afx_beam_solve_nonlinear(afx_beam* b, const double* f, int32_t n_steps, double tol, double damping,
                             int32_t max_iter, double* u, afx_beam_nl_info* info) {
    if (!b || !f || !u) return api_fail(AFX_EARG, "null argument");
    if (n_steps < 1 || !(tol > 0.0) || !(damping > 0.0) || max_iter < 1)
        return api_fail(AFX_EARG, "need n_steps >= 1, tol > 0, damping > 0 and max_iter >= 1");
    const int n = b->n, nd = 6 * n;
    std::vector<CoElem> el(b->ne);
    for (int e = 0; e < b->ne; ++e) {
        CoElem& E = el[e];
        E.n1 = b->en[2 * e];
        E.n2 = b->en[2 * e + 1];
        const double* p = &b->props[7 * e];
        double lam[3][3], L;
        frame_axes(b->x[E.n1], b->x[E.n2], {p[4], p[5], p[6]}, lam, L);
        frame_local(L, p[0], p[1], p[2], p[3], E.Kl);
        for (int k = 0; k < 3; ++k) {
            const double mid = 0.5 * (b->x[E.n1][k] + b->x[E.n2][k]);
            E.r1[k] = b->x[E.n1][k] - mid;
            E.r2[k] = b->x[E.n2][k] - mid;
        }
        E.qe = E.q1 = E.q2 = E.qmid = q_from_axes(lam);
        for (int k = 0; k < 6; ++k) E.u1[k] = E.u2[k] = 0.0;
    }
    // external loads with the clamped dofs zeroed (set_K_Final_sparse)
    std::vector<double> F(f, f + nd), Fint(nd, 0.0), Fd(nd), da(nd), df(nd);
    for (int i = 0; i < n; ++i)
        if (b->fixed[i])
            for (int k = 0; k < 6; ++k) F[6 * i + k] = 0.0;
    std::fill(u, u + nd, 0.0);
    std::vector<Quat> qmap(n, Quat{0, 0, 0, 0});
    std::vector<std::array<double, 9>> rot(b->ne);
    std::vector<double> Lf;
    auto factor = [&]() -> int {
        for (int e = 0; e < b->ne; ++e) rot[e] = qmat(el[e].qmid);
        return beam_assemble(b, &rot, Lf);
    };
    int iters = 0;
    double res = 0.0;
    bool ok = true;
    for (int step = 1; step <= n_steps && ok; ++step) {
        const double lf = (double)step / (double)n_steps;
        for (int i = 0; i < nd; ++i) Fd[i] = lf * F[i] - Fint[i];
        int rc = factor();
        if (rc) return rc;
        beam_solve_with(b, Lf, Fd.data(), da.data());
        do {
            for (int i = 0; i < nd; ++i) u[i] += da[i];
            for (int i = 0; i < n; ++i) qmap[i] = q_from_increment(da[6 * i + 3], da[6 * i + 4], da[6 * i + 5]);
            for (auto& E : el) {
                for (int k = 0; k < 6; ++k) { E.u1[k] += da[6 * E.n1 + k]; E.u2[k] += da[6 * E.n2 + k]; }
                E.q1 = qmul(qmap[E.n1], E.q1);
                E.q2 = qmul(qmap[E.n2], E.q2);
                E.qmid = q_half(E.q1, E.q2);
                double Fe[12];
                element_force(E, Fe);
                for (int k = 0; k < 6; ++k) { Fint[6 * E.n1 + k] += Fe[k]; Fint[6 * E.n2 + k] += Fe[6 + k]; }
            }
            for (int i = 0; i < n; ++i)
                if (b->fixed[i])
                    for (int k = 0; k < 6; ++k) Fint[6 * i + k] = 0.0;
            // Rotate_Ext_Loads (model.hpp:193-205): the loads follow each node's increment
            for (int i = 0; i < n; ++i) {
                const auto R = qmat(qmap[i]);
                const V3 fo = matvec(R, {F[6 * i], F[6 * i + 1], F[6 * i + 2]});
                const V3 mo = matvec(R, {F[6 * i + 3], F[6 * i + 4], F[6 * i + 5]});
                for (int k = 0; k < 3; ++k) { F[6 * i + k] = fo[k]; F[6 * i + 3 + k] = mo[k]; }
            }
            for (int i = 0; i < nd; ++i) Fd[i] = lf * F[i] - Fint[i];
            for (int i = 0; i < n; ++i)
                if (b->fixed[i])
                    for (int k = 0; k < 6; ++k) F[6 * i + k] = 0.0;
            rc = factor();
            if (rc) return rc;
            beam_solve_with(b, Lf, Fd.data(), df.data());
            double s2 = 0.0;
            for (int i = 0; i < nd; ++i) s2 += df[i] * df[i];
            res = std::sqrt(s2);
            std::fill(Fint.begin(), Fint.end(), 0.0);
            const double a = (res < 100. * tol) ? damping : 1.0;
            for (int i = 0; i < nd; ++i) da[i] = a * df[i];
            ++iters;
            if (!std::isfinite(res) || (iters >= max_iter && res > tol)) { ok = false; break; }
        } while (res > tol);
    }
    if (info) {
        info->iterations = iters;
        info->converged = ok && res <= tol ? 1 : 0;
        info->residual = res;
    }
    return AFX_OK;
}

// ------------------------------------------------------- aeroelastic ----
void afx_default_coupling_opts(afx_coupling_opts* o) {
    o->tol = 1e-8;
    o->max_iter = 50;
    o->relax = 1.0;
    o->solve_tol = 0.0;
    o->solve_max_iter = 30;
    o->refactor_above = 0.5;
    o->warm_start = 0;
    o->aitken = 1;
}

int afx_aeroelastic_create(afx_vlm* v, afx_beam* b, afx_aeroelastic** out) {
    if (!v || !b || !out) return api_fail(AFX_EARG, "null argument");
    auto* h = new afx_aeroelastic();
    h->v = v;
    h->b = b;
    int rc = afx_vlm_counts(v, &h->nn, &h->nr, &h->ns);
    if (rc) { delete h; return rc; }
    h->x0.resize(3 * (size_t)h->nn);
    afx_vlm_get_nodes(v, h->x0.data());
    h->na.resize(h->nn);
    h->nb.resize(h->nn);
    h->nw.resize(h->nn);
    for (int k = 0; k < h->nn; ++k) locate(b, h->x0[3 * k + 1], h->na[k], h->nb[k], h->nw[k]);
    // ring force points: leading-edge midpoints (vortexRing::forceActingPoint,
    // panel.cpp:96-99) of the undeformed lattice
    h->rn.resize(4 * (size_t)h->nr);
    afx_vlm_get_rings(v, h->rn.data(), nullptr);
    h->fp0.resize(h->nr);
    h->ra.resize(h->nr);
    h->rb.resize(h->nr);
    h->rw.resize(h->nr);
    for (int k = 0; k < h->nr; ++k) {
        const double* p0 = &h->x0[3 * (size_t)h->rn[4 * k]];
        const double* p1 = &h->x0[3 * (size_t)h->rn[4 * k + 1]];
        h->fp0[k] = {0.5 * (p0[0] + p1[0]), 0.5 * (p0[1] + p1[1]), 0.5 * (p0[2] + p1[2])};
        locate(b, h->fp0[k][1], h->ra[k], h->rb[k], h->rw[k]);
    }
    h->disp.assign(6 * (size_t)b->n, 0.0);
    h->tip = 0;
    for (int k = 1; k < b->n; ++k)
        if (b->x[k][1] > b->x[h->tip][1]) h->tip = k;
    *out = h;
    return AFX_OK;
}

int afx_aeroelastic_destroy(afx_aeroelastic* h) {
    delete h;
    return AFX_OK;
}

// Fixed-point coupling: lattice <- beam displacement, matrix and rhs rebuilt
// on the device, Gamma by FP64 refinement preconditioned with the FP32 LU of
// an earlier matrix (refactored when refinement contracts worse than
// refactor_above or does not converge), ring forces -> beam loads, K u = f,
// relaxed update (fixed or Aitken), until the tip deflection changes by less
// than tol (relative).
int afx_aeroelastic_solve(afx_aeroelastic* h, const afx_coupling_opts* o, afx_coupling_info* info) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    afx_coupling_opts opts;
    if (o) opts = *o; else afx_default_coupling_opts(&opts);
    if (opts.max_iter <= 0 || !(opts.relax > 0.0)) return api_fail(AFX_EARG, "max_iter and relax must be positive");
    afx_coupling_info inf{};
    if (!opts.warm_start) std::fill(h->disp.begin(), h->disp.end(), 0.0);
    const afx_beam* b = h->b;
    std::vector<double> xyz(3 * (size_t)h->nn), rf(3 * (size_t)h->nr), load(6 * (size_t)b->n), unew(6 * (size_t)b->n);
    double prev_tip = h->disp[6 * h->tip + 2];
    // inexact fixed point: with solve_tol <= 0 each Gamma solve is only as
    // accurate as the coupling needs, 1e-5 of the last relative tip change
    // (the matrix's condition amplifies it into Gamma), clamped to
    // [sqrt(N) eps64, 1e-9], and convergence is confirmed by one iteration
    // with the strict tolerance; solve_tol > 0 fixes it
    const double tol_floor = std::sqrt((double)h->nr) * 1.1102230246251565e-16;
    double last_change = 1.0;
    // Aitken dynamic relaxation (Irons-Tuck form, as in partitioned
    // fluid-structure coupling): omega_k = -omega_k-1 r_k-1.(r_k - r_k-1) / |r_k - r_k-1|^2
    std::vector<double> r_prev(unew.size(), 0.0);
    double omega = opts.relax;
    int rc = AFX_OK;
    for (int it = 1; it <= opts.max_iter; ++it) {
        const double stol =
            opts.solve_tol > 0.0 ? opts.solve_tol : std::min(1e-9, std::max(tol_floor, 1e-5 * last_change));
        for (int k = 0; k < h->nn; ++k) {
            const V3 p{h->x0[3 * k], h->x0[3 * k + 1], h->x0[3 * k + 2]};
            const V3 u = follow(b, h->disp, h->na[k], h->nb[k], h->nw[k], p);
            for (int d = 0; d < 3; ++d) xyz[3 * k + d] = p[d] + u[d];
        }
        if ((rc = afx_vlm_set_nodes(h->v, xyz.data()))) return rc;
        if ((rc = afx_vlm_build(h->v))) return rc;
        if (!h->have_factor) {
            if ((rc = afx_vlm_factor(h->v))) return rc;
            h->have_factor = true;
            ++inf.refactors;
        }
        afx_solve_info si{};
        const int warm = it > 1 ? 1 : 0;   // the previous iteration's Gamma starts the refinement
        if ((rc = afx_vlm_solve(h->v, stol, opts.solve_max_iter, warm, nullptr, &si))) return rc;
        inf.refine_steps += si.iterations;
        if (!si.converged || si.contraction > opts.refactor_above) {
            if ((rc = afx_vlm_factor(h->v))) return rc;
            ++inf.refactors;
            if ((rc = afx_vlm_solve(h->v, stol, opts.solve_max_iter, warm, nullptr, &si))) return rc;
            inf.refine_steps += si.iterations;
            if (!si.converged) return api_fail(AFX_EARG, "vortex-lattice refinement did not converge after refactoring");
        }
        inf.backward_error = si.backward_error;
        double coeffs[3];
        if ((rc = afx_vlm_forces(h->v, rf.data(), nullptr, coeffs))) return rc;
        inf.cl = coeffs[0];
        // loads: w_a F and w_a (p0 - x_a) x F on the two bracketing nodes
        std::fill(load.begin(), load.end(), 0.0);
        for (int k = 0; k < h->nr; ++k) {
            const V3 F{rf[3 * k], rf[3 * k + 1], rf[3 * k + 2]};
            const int nodes[2] = {h->ra[k], h->rb[k]};
            const double ws[2] = {h->rw[k], 1.0 - h->rw[k]};
            for (int q = 0; q < 2; ++q) {
                if (ws[q] == 0.0) continue;
                const V3 M = cross(sub(h->fp0[k], b->x[nodes[q]]), F);
                double* l = &load[6 * nodes[q]];
                for (int d = 0; d < 3; ++d) {
                    l[d] += ws[q] * F[d];
                    l[3 + d] += ws[q] * M[d];
                }
            }
        }
        beam_solve_impl(b, load.data(), unew.data());
        if (opts.aitken && it > 1) {
            double num = 0.0, den = 0.0;
            for (size_t i = 0; i < unew.size(); ++i) {
                const double r = unew[i] - h->disp[i], dr = r - r_prev[i];
                num += r_prev[i] * dr;
                den += dr * dr;
            }
            if (den > 0.0) omega = -omega * num / den;
        }
        for (size_t i = 0; i < unew.size(); ++i) {
            const double r = unew[i] - h->disp[i];
            r_prev[i] = r;
            h->disp[i] += (opts.aitken ? omega : opts.relax) * r;
        }
        const double tip = h->disp[6 * h->tip + 2];
        inf.iterations = it;
        inf.tip_deflection = tip;
        if (!std::isfinite(tip)) return api_fail(AFX_EARG, "coupling diverged (non-finite tip deflection)");
        last_change = std::fabs(tip - prev_tip) / std::max(std::fabs(tip), 1e-300);
        if (std::fabs(tip - prev_tip) <= opts.tol * std::max(std::fabs(tip), 1e-300)) {
            // accepted only after an iteration whose solve met the strict tolerance
            if (opts.solve_tol > 0.0 || stol <= tol_floor) {
                inf.converged = 1;
                break;
            }
            last_change = 0.0;
        }
        prev_tip = tip;
    }
    if (info) *info = inf;
    return AFX_OK;
}

int afx_aeroelastic_get_disp(afx_aeroelastic* h, double* disp) {
    if (!h || !disp) return api_fail(AFX_EARG, "null argument");
    std::copy(h->disp.begin(), h->disp.end(), disp);
    return AFX_OK;
}

}  // extern "C"

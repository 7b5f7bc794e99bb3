// Oracle (test infrastructure only): CPU FP64 restatement of the reference's
// vortex-lattice path (src/vlm/src) and of its linear Euler-Bernoulli frame
// model (src/structure/include/structure), plus the load / displacement
// transfer that src/aeroelasticite intends (aerotools.cpp:25-217 does not
// compile, so the transfer follows its stated design: nearest two beam nodes
// along the span, lever-arm moments).  Only tests, the smoke check and the
// CPU baseline of the benchmark call it.
//
// Each function follows the reference line by line where one exists:
//   ovl_segment   potential.cpp:36-58 (vortexLine::influence, distToPoint,
//                 inducedVelOrientation)
//   ovl_ring      panel.cpp:122-128 (vortexRing::influence)
//   ovl_geometry  panel.cpp:46-72 (centre, normal, area), 111-120 (collocation)
//   ovl_build     solver.cpp:139-254 (buildLHS), 256-290 (buildRHS),
//                 surface.cpp:35-58 (generateWake), solver.cpp:33-37 (wake
//                 length 100 cref)
//   ovl_solve     the exact solution that solver.cpp:41-50 approximates with
//                 BiCGSTAB: dense LU with partial pivoting in FP64
//   ovl_forces    surface.cpp:164-201 (wingStation::computeForces), solver.cpp:
//                 112-128 and surface.cpp:262-277 (CL summed over stations)
//   ovl_nonlinear nonlinear::steady::solve / iterate / computeForces
//                 (solver.cpp:330-476), database.cpp:192-242 (polars)
//   ovl_beam_*    element.hpp:77-160 (set_K_local, set_T_Rotation,
//                 set_K_elem_global), model.hpp:103-131 (assembly),
//                 225-238 (linear solve) with SPC1 as row/column zeroing
//   ovl_beam_nonlinear  model.hpp:240-366 (get_NonLin_Solve, Rotate_Ext_Loads,
//                 set_Quaternion_Map), element.hpp:111-370 (corotational CBAR)
//
// Pinning (tests/test_oracle_vlm.py): segment and ring closed forms of the
// Biot-Savart law, Helmbold's lifting-surface lift slope, symmetry, Kutta-
// Joukowski total lift against the circulation sum, cantilever closed forms
// (examples/structure/Moment.txt, tip force, torsion, axial), rotation
// invariance of the frame, and conservation of force and moment by the load
// transfer.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct V3 {
    double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double nrm(V3 a) { return std::sqrt(dot(a, a)); }
V3 load3(const double* p) { return {p[0], p[1], p[2]}; }
void store3(double* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// potential.cpp:36-58 with gamma = 1 (vortexRing::initialize, panel.cpp:87-93)
V3 segment(V3 P, V3 a, V3 b, double rc) {
    V3 r1 = P - a, r2 = P - b;
    V3 r1r2 = cross(r1, r2);
    double norm = nrm(r1r2);
    V3 dl = b - a;
    double dlr1 = dot(dl, r1), dlr2 = dot(dl, r2);
    double h = norm / (dlr1 / nrm(r1) - dlr2 / nrm(r2));
    V3 dir = cross(dl, r1);
    dir = (1.0 / nrm(dir)) * dir;
    double lo = rc > 0.0 ? 1.0 - std::exp(-(h / rc) * (h / rc)) : 1.0;   // rc = 0: exp(-inf) = 0
    return (1.0 / (4.0 * M_PI * h) * lo) * dir;
}

V3 ring(V3 P, const V3 q[4], double rc) {
    V3 v{0, 0, 0};
    for (int i = 0; i < 4; ++i) v = v + segment(P, q[i], q[(i + 1) % 4], rc);
    return v;
}

double kfac(double aoa) {
    if (aoa < 1e-6 && aoa > -1e6) return 1.0;
    return aoa * M_PI / (180.0 * std::sin(aoa * M_PI / 180.0));
}

struct Lattice {
    int nn, nr, ns;
    const double* xyz;
    const int32_t* soff;
    const int32_t* rn;
    V3 node(int ring, int c) const { return load3(xyz + 3 * rn[4 * ring + c]); }
};

void geometry(const Lattice& L, int k, double aoa, V3& colloc, V3& normal, double& area) {
    V3 p[4];
    for (int c = 0; c < 4; ++c) p[c] = L.node(k, c);
    V3 center{0, 0, 0};
    for (int c = 0; c < 4; ++c) center = center + p[c];
    center = (1.0 / 4.0) * center;
    V3 dl[4];
    for (int c = 0; c < 4; ++c) dl[c] = p[(c + 1) % 4] - p[c];
    V3 n{0, 0, 0};
    for (int c = 0; c < 3; ++c) n = n - cross(dl[c], dl[c + 1]);
    n = n - cross(dl[3], dl[0]);
    normal = (1.0 / nrm(n)) * n;
    area = 0.0;
    for (int c = 0; c < 4; ++c) area += 0.5 * nrm(cross(dl[c], center - p[c]));
    double kk = kfac(aoa);
    colloc = kk * center + (1 - kk) * (0.5 * (p[0] + p[1]));
}

V3 freestream(double aoa, double vinf) {
    double a = aoa * M_PI / 180.0;
    return {vinf * std::cos(a), 0.0, vinf * std::sin(a)};
}

// dense LU with partial pivoting, FP64, in place on a copy
bool lu_solve(int n, std::vector<double> A, std::vector<double> b, double* x) {
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        if (A[(size_t)p * n + k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            A[(size_t)i * n + k] = l;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            b[i] -= l * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * x[j];
        x[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

// ------------------------------------------------------------------ beam --
struct Beam {
    int n, ne;
    const double* xyz;
    const int32_t* en;
    const double* props;
    std::vector<char> fixed;
};

void element_local(const Beam& B, int e, double Kout[12][12], double lout[3][3]) {
    V3 N1 = load3(B.xyz + 3 * B.en[2 * e]), N2 = load3(B.xyz + 3 * B.en[2 * e + 1]);
    const double* p = B.props + 7 * e;
    double L = nrm(N2 - N1);
    double Li = 1 / L;
    // set_K_local: X, Y1..Y4 (Iz), Z1..Z4 (Iy), S
    double X = p[0] * Li;
    double Y2 = 6. * p[2] * Li * Li, Y1 = Y2 * 2. * Li, Y3 = 4. * p[2] * Li, Y4 = Y3 / 2.;
    double Z2 = 6. * p[1] * Li * Li, Z1 = Z2 * 2. * Li, Z3 = 4. * p[1] * Li, Z4 = Z3 / 2.;
    double S = p[3] * Li;
    const double Kl[12][12] = {
        {X, 0, 0, 0, 0, 0, -X, 0, 0, 0, 0, 0},       {0, Y1, 0, 0, 0, Y2, 0, -Y1, 0, 0, 0, Y2},
        {0, 0, Z1, 0, -Z2, 0, 0, 0, -Z1, 0, -Z2, 0}, {0, 0, 0, S, 0, 0, 0, 0, 0, -S, 0, 0},
        {0, 0, -Z2, 0, Z3, 0, 0, 0, Z2, 0, Z4, 0},   {0, Y2, 0, 0, 0, Y3, 0, -Y2, 0, 0, 0, Y4},
        {-X, 0, 0, 0, 0, 0, X, 0, 0, 0, 0, 0},       {0, -Y1, 0, 0, 0, -Y2, 0, Y1, 0, 0, 0, -Y2},
        {0, 0, -Z1, 0, Z2, 0, 0, 0, Z1, 0, Z2, 0},   {0, 0, 0, -S, 0, 0, 0, 0, 0, S, 0, 0},
        {0, 0, -Z2, 0, Z4, 0, 0, 0, Z2, 0, Z3, 0},   {0, Y2, 0, 0, 0, Y4, 0, -Y2, 0, 0, 0, Y3}};
    // set_T_Rotation
    V3 ex = Li * (N2 - N1);
    V3 v = load3(p + 4);
    v = (1.0 / nrm(v)) * v;
    V3 ez = cross(ex, v);
    ez = (1.0 / nrm(ez)) * ez;
    V3 ey = -1.0 * cross(ex, ez);
    double lam[3][3] = {{ex.x, ex.y, ex.z}, {ey.x, ey.y, ey.z}, {ez.x, ez.y, ez.z}};
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) Kout[i][j] = Kl[i][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) lout[i][j] = lam[i][j];
}

void element_global(const Beam& B, int e, double Kg[12][12]) {
    double Kl[12][12], lam[3][3];
    element_local(B, e, Kl, lam);
    double T[12][12] = {};
    for (int b = 0; b < 4; ++b)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) T[3 * b + r][3 * b + c] = lam[r][c];
    // set_K_elem_global: T^T K T
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double s = 0;
            for (int a = 0; a < 12; ++a)
                for (int b = 0; b < 12; ++b) s += T[a][i] * Kl[a][b] * T[b][j];
            Kg[i][j] = s;
        }
}

void beam_solve(const Beam& B, const double* f, double* u) {
    const int nd = 6 * B.n;
    std::vector<double> K((size_t)nd * nd, 0.0), rhs(f, f + nd);
    for (int e = 0; e < B.ne; ++e) {
        double Kg[12][12];
        element_global(B, e, Kg);
        int d[2] = {6 * B.en[2 * e], 6 * B.en[2 * e + 1]};
        for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c)
                for (int i = 0; i < 6; ++i)
                    for (int j = 0; j < 6; ++j) K[(size_t)(d[a] + i) * nd + d[c] + j] += Kg[6 * a + i][6 * c + j];
    }
    for (int i = 0; i < nd; ++i)
        if (B.fixed[i / 6]) {
            for (int j = 0; j < nd; ++j) K[(size_t)i * nd + j] = K[(size_t)j * nd + i] = 0.0;
            K[(size_t)i * nd + i] = 1.0;
            rhs[i] = 0.0;
        }
    lu_solve(nd, K, rhs, u);
}

// nearest two beam nodes bracketing y, else the nearest node
void bracket(const Beam& B, double y, int& a, int& c, double& wa) {
    for (int e = 0; e < B.ne; ++e) {
        int n1 = B.en[2 * e], n2 = B.en[2 * e + 1];
        double y1 = B.xyz[3 * n1 + 1], y2 = B.xyz[3 * n2 + 1];
        if (y1 == y2) continue;
        if (y >= std::min(y1, y2) && y <= std::max(y1, y2)) {
            a = n1;
            c = n2;
            wa = (y2 - y) / (y2 - y1);
            return;
        }
    }
    int best = 0;
    for (int k = 1; k < B.n; ++k)
        if (std::fabs(B.xyz[3 * k + 1] - y) < std::fabs(B.xyz[3 * best + 1] - y)) best = k;
    a = c = best;
    wa = 1.0;
}

}  // namespace

extern "C" {

void ovl_segment(const double* P, const double* a, const double* b, double rc, double* v) {
    store3(v, segment(load3(P), load3(a), load3(b), rc));
}

void ovl_ring(const double* P, const double* q12, double rc, double* v) {
    V3 q[4];
    for (int c = 0; c < 4; ++c) q[c] = load3(q12 + 3 * c);
    store3(v, ring(load3(P), q, rc));
}

void ovl_geometry(int nn, const double* xyz, int ns, const int32_t* soff, const int32_t* rn, double aoa,
                  double* colloc, double* normal, double* area) {
    Lattice L{nn, soff[ns], ns, xyz, soff, rn};
    for (int k = 0; k < L.nr; ++k) {
        V3 c, n;
        double a;
        geometry(L, k, aoa, c, n, a);
        if (colloc) store3(colloc + 3 * k, c);
        if (normal) store3(normal + 3 * k, n);
        if (area) area[k] = a;
    }
}

// lhs (row-major N x N) and rhs of the linear steady solver
void ovl_build(int nn, const double* xyz, int ns, const int32_t* soff, const int32_t* rn, double aoa, double vinf,
               double wake_length, double rc, double* A, double* rhs) {
    Lattice L{nn, soff[ns], ns, xyz, soff, rn};
    const int N = L.nr;
    std::vector<V3> P(N), n(N);
    for (int k = 0; k < N; ++k) {
        double a;
        geometry(L, k, aoa, P[k], n[k], a);
    }
    V3 Vf = freestream(aoa, vinf);
    V3 w = (1.0 / nrm(Vf)) * Vf;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            V3 q[4];
            for (int c = 0; c < 4; ++c) q[c] = L.node(j, c);
            A[(size_t)i * N + j] = dot(ring(P[i], q, rc), n[i]);
        }
    for (int s = 0; s < ns; ++s) {
        int te = soff[s + 1] - 1;
        V3 p1 = L.node(te, 3), p2 = L.node(te, 2);
        V3 q[4] = {p1, p2, p2 + wake_length * w, p1 + wake_length * w};
        for (int i = 0; i < N; ++i) A[(size_t)i * N + te] += dot(ring(P[i], q, rc), n[i]);
    }
    for (int i = 0; i < N; ++i) rhs[i] = -dot(Vf, n[i]);
}

int ovl_solve(int n, const double* A, const double* b, double* x) {
    return lu_solve(n, std::vector<double>(A, A + (size_t)n * n), std::vector<double>(b, b + n), x) ? 0 : -1;
}

// ring forces (3 per ring, nullable) and {CL, CY, CD_kj} over (q sref)
void ovl_forces(int nn, const double* xyz, int ns, const int32_t* soff, const int32_t* rn, const double* gamma,
                double aoa, double vinf, double rho, double sref, double* rforce, double* coeffs) {
    Lattice L{nn, soff[ns], ns, xyz, soff, rn};
    V3 Vf = freestream(aoa, vinf);
    double a = aoa * M_PI / 180.0;
    V3 lift{-std::sin(a), 0.0, std::cos(a)};
    V3 stream = (1.0 / nrm(Vf)) * Vf;
    double q = 0.5 * rho * vinf * vinf;
    double cl = 0, cy = 0, cd = 0;
    for (int s = 0; s < ns; ++s) {
        double prev = 0.0;
        V3 F{0, 0, 0};
        for (int k = soff[s]; k < soff[s + 1]; ++k) {
            V3 dl = L.node(k, 1) - L.node(k, 0);
            V3 f = (rho * (gamma[k] - prev)) * cross(Vf, dl);
            if (rforce) store3(rforce + 3 * k, f);
            F = F + f;
            prev = gamma[k];
        }
        cl += dot(F, lift) / (q * sref);
        cy += F.y / (q * sref);
        cd += dot(F, stream) / (q * sref);
    }
    if (coeffs) { coeffs[0] = cl; coeffs[1] = cy; coeffs[2] = cd; }
}

// linear frame model: K u = f with SPC1 (all six dofs) on the clamped nodes
void ovl_beam_solve(int n, const double* xyz, int ne, const int32_t* en, const double* props, int nclamped,
                    const int32_t* clamped, const double* f, double* u) {
    Beam B{n, ne, xyz, en, props, std::vector<char>(n, 0)};
    for (int k = 0; k < nclamped; ++k) B.fixed[clamped[k]] = 1;
    beam_solve(B, f, u);
}

// beam loads from ring forces at the undeformed force points, and lattice
// node displacements from a beam displacement (the transfer's two halves)
void ovl_transfer_loads(int nr, const double* fpoints, const double* rforce, int n, const double* bxyz, int ne,
                        const int32_t* en, double* load) {
    Beam B{n, ne, bxyz, en, nullptr, {}};
    std::fill(load, load + 6 * n, 0.0);
    for (int k = 0; k < nr; ++k) {
        V3 p = load3(fpoints + 3 * k), F = load3(rforce + 3 * k);
        int a, c;
        double wa;
        bracket(B, p.y, a, c, wa);
        int nodes[2] = {a, c};
        double ws[2] = {wa, 1.0 - wa};
        for (int q = 0; q < 2; ++q) {
            V3 M = cross(p - load3(bxyz + 3 * nodes[q]), F);
            double* l = load + 6 * nodes[q];
            l[0] += ws[q] * F.x; l[1] += ws[q] * F.y; l[2] += ws[q] * F.z;
            l[3] += ws[q] * M.x; l[4] += ws[q] * M.y; l[5] += ws[q] * M.z;
        }
    }
}

void ovl_transfer_disp(int nn, const double* x0, int n, const double* bxyz, int ne, const int32_t* en,
                       const double* disp, double* xyz) {
    Beam B{n, ne, bxyz, en, nullptr, {}};
    for (int k = 0; k < nn; ++k) {
        V3 p = load3(x0 + 3 * k);
        int a, c;
        double wa;
        bracket(B, p.y, a, c, wa);
        int nodes[2] = {a, c};
        double ws[2] = {wa, 1.0 - wa};
        V3 u{0, 0, 0};
        for (int q = 0; q < 2; ++q) {
            const double* d = disp + 6 * nodes[q];
            V3 rot{d[3], d[4], d[5]};
            u = u + ws[q] * (load3(d) + cross(rot, p - load3(bxyz + 3 * nodes[q])));
        }
        store3(xyz + 3 * k, p + u);
    }
}

// nonlinear::steady (vlm/src/solver.cpp:330-476) with exact FP64 solves and
// the polar interpolation of database.cpp:192-242 (mlinterp: linear, end
// values outside).  Polar rows: per airfoil [off[a], off[a+1]) of
// (alpha_deg, cl, cd, cmy); sections (airfoil, span 0..1).  out = {iterations,
// residual, CL, CD, CY}; alpha_s and station cl (ns each) returned.
static double lin1(const double* xd, const double* yd, int n, double x) {
    if (n == 1 || x <= xd[0]) return yd[0];
    if (x >= xd[n - 1]) return yd[n - 1];
    int m = 0;
    while (!(x >= xd[m] && x < xd[m + 1])) ++m;   // linear scan (the reference bisects; same interval)
    double w = (xd[m + 1] - x) / (xd[m + 1] - xd[m]);
    return w * yd[m] + (1 - w) * yd[m + 1];
}

static void polar(int na, const int32_t* off, const double* pa, const double* pcl, const double* pcd, int nsec,
                  const int32_t* sa, const double* sspan, double alpha, double span, double* out) {
    (void)na;
    std::vector<double> c0, c1;
    for (int q = 0; q < nsec; ++q) {
        int a = sa[q], o = off[a], n = off[a + 1] - o;
        c0.push_back(lin1(pa + o, pcl + o, n, alpha));
        c1.push_back(lin1(pa + o, pcd + o, n, alpha));
    }
    out[0] = lin1(sspan, c0.data(), nsec, span);
    out[1] = lin1(sspan, c1.data(), nsec, span);
}

void ovl_nonlinear(int nn, const double* xyz, int ns, const int32_t* soff, const int32_t* rn, double aoa, double vinf,
                   double rho, double sref, double cref, int na, const int32_t* off, const double* pa, const double* pcl,
                   const double* pcd, int nsec, const int32_t* sa, const double* sspan, double relax, double tol,
                   int max_iter, double* alpha_s, double* cl_s, double* out) {
    Lattice L{nn, soff[ns], ns, xyz, soff, rn};
    const int N = L.nr;
    const double CLA = 0.109662271123215082635482531259185634553432464599609375;
    std::vector<double> A((size_t)N * N), rhs(N), g(N);
    ovl_build(nn, xyz, ns, soff, rn, aoa, vinf, 100.0 * cref, 0.0, A.data(), rhs.data());
    std::vector<V3> nor(N);
    for (int k = 0; k < N; ++k) {
        V3 c;
        double a;
        geometry(L, k, aoa, c, nor[k], a);
    }
    std::vector<double> area(ns, 0.0), spanloc(ns), clinv(ns);
    std::vector<V3> le(ns);
    double span = 0;
    for (int s = 0; s < ns; ++s) {
        for (int k = soff[s]; k < soff[s + 1]; ++k) {
            V3 c, n;
            double a;
            geometry(L, k, aoa, c, n, a);
            area[s] += a;
        }
        le[s] = L.node(soff[s], 1) - L.node(soff[s], 0);
        span += nrm(le[s]);
        spanloc[s] = span - 0.5 * nrm(le[s]);
    }
    const double root = spanloc[0];
    for (int s = 0; s < ns; ++s) spanloc[s] = (spanloc[s] - root) / span;
    for (int s = 0; s < ns; ++s) alpha_s[s] = aoa;
    const double q = 0.5 * rho * vinf * vinf;
    auto axis = [&](int s) { V3 v = cross(freestream(alpha_s[s], 1.0), le[s]); return (1.0 / nrm(v)) * v; };
    int iter = 0;
    double res = 0;
    do {
        for (int s = 0; s < ns; ++s)
            for (int k = soff[s]; k < soff[s + 1]; ++k) rhs[k] = -dot(freestream(alpha_s[s], vinf), nor[k]);
        ovl_solve(N, A.data(), rhs.data(), g.data());
        res = 0;
        for (int s = 0; s < ns; ++s) {
            V3 V = freestream(alpha_s[s], vinf), F{0, 0, 0};
            double prev = 0;
            for (int k = soff[s]; k < soff[s + 1]; ++k) {
                F = F + (rho * (g[k] - prev)) * cross(V, L.node(k, 1) - L.node(k, 0));
                prev = g[k];
            }
            clinv[s] = dot(F, axis(s)) / (q * area[s]);
            double c2[2];
            polar(na, off, pa, pcl, pcd, nsec, sa, sspan, clinv[s] / CLA - alpha_s[s] + aoa, spanloc[s], c2);
            alpha_s[s] += relax * (c2[0] - clinv[s]) / CLA;
            res += ((c2[0] - clinv[s]) / c2[0]) * ((c2[0] - clinv[s]) / c2[0]);
        }
        res = std::sqrt(res);
        ++iter;
    } while (res > tol && iter <= max_iter);
    double a = aoa * M_PI / 180.0;
    V3 lift{-std::sin(a), 0.0, std::cos(a)};
    double cl = 0, cd = 0, cy = 0;
    for (int s = 0; s < ns; ++s) {
        double c2[2];
        polar(na, off, pa, pcl, pcd, nsec, sa, sspan, clinv[s] / CLA - alpha_s[s] + aoa, spanloc[s], c2);
        V3 ax = axis(s);
        cl_s[s] = c2[0] * dot(ax, lift);
        cl += cl_s[s] * area[s] / sref;
        cy += c2[0] * ax.y * area[s] / sref;
        cd += c2[1] * area[s] / sref;
    }
    out[0] = iter;
    out[1] = res;
    out[2] = cl;
    out[3] = cd;
    out[4] = cy;
}

// fixed-point coupling with exact FP64 solves; out = {iterations, tip, CL}
// MODEL::get_NonLin_Solve (model.hpp:240-339) with the corotational CBAR
// (element.hpp:111-370): quaternion helpers follow element.hpp:159-287 and
// model.hpp:193-220, 341-366; Eigen's quaternion product is the Hamilton
// product; the stiffness is re-assembled as Rot K_local Rot^T with Rot the
// element's mid-node rotation (set_K_global) and solved densely with the
// clamped rows and columns replaced by identity (set_K_Final_sparse).
// Dep starts at zero (its setZero is commented out in the reference).
// out3: iterations, converged (0/1), last residual.
namespace {
struct Q4 { double w, x, y, z; };
Q4 qprod(Q4 a, Q4 b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
Q4 qconjugate(Q4 a) { return {a.w, -a.x, -a.y, -a.z}; }
void qrotation(Q4 q, double M[3][3]) {
    const double s = q.w, x = q.x, y = q.y, z = q.z;
    M[0][0] = 1. - 2. * (y * y + z * z); M[0][1] = 2. * (x * y - s * z); M[0][2] = 2. * (x * z + s * y);
    M[1][0] = 2. * (x * y + s * z); M[1][1] = 1. - 2. * (x * x + z * z); M[1][2] = 2. * (y * z - s * x);
    M[2][0] = 2. * (x * z - s * y); M[2][1] = 2. * (y * z + s * x); M[2][2] = 1. - 2. * (x * x + y * y);
}
Q4 qe_from(const double m[3][3]) {
    const double trace = m[0][0] + m[1][1] + m[2][2];
    double cst, s, vx, vy, vz;
    if (trace > 0.0) {
        cst = 0.5 / std::sqrt(trace + 1.0);
        s = 0.25 / cst;
        vx = (m[1][2] - m[2][1]) * cst;
        vy = (m[2][0] - m[0][2]) * cst;
        vz = (m[0][1] - m[1][0]) * cst;
    } else if (m[0][0] > m[1][1] && m[0][0] > m[2][2]) {
        cst = 2.0 * std::sqrt(1.0 + m[0][0] - m[1][1] - m[2][2]);
        s = (m[1][2] - m[2][1]) / cst;
        vx = 0.25 * cst;
        vy = (m[1][0] + m[0][1]) / cst;
        vz = (m[2][0] + m[0][2]) / cst;
    } else if (m[1][1] > m[2][2]) {
        cst = 2.0 * std::sqrt(1.0 + m[1][1] - m[0][0] - m[2][2]);
        s = (m[2][0] - m[0][2]) / cst;
        vx = (m[1][0] + m[0][1]) / cst;
        vy = 0.25 * cst;
        vz = (m[2][1] + m[1][2]) / cst;
    } else {
        cst = 2.0 * std::sqrt(1.0 + m[2][2] - m[0][0] - m[1][1]);
        s = (m[0][1] - m[1][0]) / cst;
        vx = (m[2][0] + m[0][2]) / cst;
        vy = (m[2][1] + m[1][2]) / cst;
        vz = 0.25 * cst;
    }
    return {s, vx, vy, vz};
}
struct CoBar {
    int n1, n2;
    double Kl[12][12];
    V3 r1, r2;
    Q4 qe, q1, q2, qmid;
    double u1[6], u2[6], umid[6];
};
}  // namespace

void ovl_beam_nonlinear(int n, const double* xyz, int ne, const int32_t* en, const double* props, int nclamped,
                        const int32_t* clamped, const double* f, int max_load_step, double tol, double amor,
                        int max_iter, double* dep, double* out3) {
    Beam B{n, ne, xyz, en, props, std::vector<char>(n, 0)};
    for (int k = 0; k < nclamped; ++k) B.fixed[clamped[k]] = 1;
    const int nd = 6 * n;
    std::vector<CoBar> bars(ne);
    for (int e = 0; e < ne; ++e) {
        CoBar& c = bars[e];
        c.n1 = en[2 * e];
        c.n2 = en[2 * e + 1];
        double lam[3][3];
        element_local(B, e, c.Kl, lam);
        V3 N1 = load3(xyz + 3 * c.n1), N2 = load3(xyz + 3 * c.n2);
        V3 Nmid = 0.5 * (N1 + N2);
        c.r1 = N1 - Nmid;
        c.r2 = N2 - Nmid;
        c.qe = qe_from(lam);
        c.q1 = c.q2 = c.qmid = c.qe;
        for (int k = 0; k < 6; ++k) c.u1[k] = c.u2[k] = c.umid[k] = 0;
    }
    std::vector<double> Forces(f, f + nd), Fint(nd, 0.0), Fdiff(nd), damor(nd), dfull(nd);
    for (int i = 0; i < nd; ++i)
        if (B.fixed[i / 6]) Forces[i] = 0;
    for (int i = 0; i < nd; ++i) dep[i] = 0;
    std::vector<Q4> qmap(n, Q4{0, 0, 0, 0});
    // set_K_global + set_K_Final_sparse + get_Solve
    auto solve = [&](const std::vector<double>& rhs, std::vector<double>& x) {
        std::vector<double> K((size_t)nd * nd, 0.0);
        for (auto& c : bars) {
            double D[3][3];
            qrotation(c.qmid, D);
            double Rot[12][12] = {};
            for (int b = 0; b < 4; ++b)
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc) Rot[3 * b + r][3 * b + cc] = D[r][cc];
            double T[12][12];
            for (int i = 0; i < 12; ++i)
                for (int j = 0; j < 12; ++j) {
                    double s = 0;
                    for (int a = 0; a < 12; ++a)
                        for (int b = 0; b < 12; ++b) s += Rot[i][a] * c.Kl[a][b] * Rot[j][b];
                    T[i][j] = s;
                }
            for (int j = 0; j < 6; ++j)
                for (int k = 0; k < 6; ++k) {
                    K[(size_t)(6 * c.n1 + j) * nd + 6 * c.n1 + k] += T[j][k];
                    K[(size_t)(6 * c.n2 + j) * nd + 6 * c.n2 + k] += T[6 + j][6 + k];
                    K[(size_t)(6 * c.n1 + j) * nd + 6 * c.n2 + k] += T[j][6 + k];
                    K[(size_t)(6 * c.n2 + j) * nd + 6 * c.n1 + k] += T[6 + j][k];
                }
        }
        for (int i = 0; i < nd; ++i)
            if (B.fixed[i / 6]) {
                for (int j = 0; j < nd; ++j) K[(size_t)i * nd + j] = K[(size_t)j * nd + i] = 0.0;
                K[(size_t)i * nd + i] = 1.0;
            }
        x.assign(nd, 0.0);
        lu_solve(nd, K, rhs, x.data());
    };
    int iters = 0;
    double residu = 0;
    bool ok = true;
    for (double step = 1.; step <= (double)max_load_step && ok; step++) {
        for (int i = 0; i < nd; ++i) Fdiff[i] = (step / max_load_step) * Forces[i] - Fint[i];
        solve(Fdiff, damor);
        do {
            for (int i = 0; i < nd; ++i) dep[i] += damor[i];
            for (int i = 0; i < n; ++i) {   // set_Quaternion_Map
                const double rx = damor[6 * i + 3], ry = damor[6 * i + 4], rz = damor[6 * i + 5];
                const double cr = std::cos(rx * 0.5), sr = std::sin(rx * 0.5);
                const double cp = std::cos(ry * 0.5), sp = std::sin(ry * 0.5);
                const double cy = std::cos(rz * 0.5), sy = std::sin(rz * 0.5);
                qmap[i] = {cr * cp * cy + sr * sp * sy, sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy,
                           cr * cp * sy - sr * sp * cy};
            }
            for (auto& c : bars) {
                // set_u_i
                for (int k = 0; k < 6; ++k) {
                    c.u1[k] += damor[6 * c.n1 + k];
                    c.u2[k] += damor[6 * c.n2 + k];
                    c.umid[k] = 0.5 * (c.u1[k] + c.u2[k]);
                }
                // set_q1_And_q2, set_qmid_From_Interpolation
                c.q1 = qprod(qmap[c.n1], c.q1);
                c.q2 = qprod(qmap[c.n2], c.q2);
                if (std::fabs(c.q1.w - c.q2.w) < 1e-12 && std::fabs(c.q1.x - c.q2.x) < 1e-12 &&
                    std::fabs(c.q1.y - c.q2.y) < 1e-12 && std::fabs(c.q1.z - c.q2.z) < 1e-12) {
                    c.qmid = c.q1;
                } else {
                    const double n1 = std::sqrt(c.q1.w * c.q1.w + c.q1.x * c.q1.x + c.q1.y * c.q1.y + c.q1.z * c.q1.z);
                    const double n2 = std::sqrt(c.q2.w * c.q2.w + c.q2.x * c.q2.x + c.q2.y * c.q2.y + c.q2.z * c.q2.z);
                    const double theta =
                        std::acos((c.q1.w * c.q2.w + c.q1.x * c.q2.x + c.q1.y * c.q2.y + c.q1.z * c.q2.z) / (n1 * n2));
                    if (std::fabs(theta) > 1e-14) {
                        const double cst = std::sin(0.5 * theta) / std::sin(theta);
                        c.qmid = {cst * c.q1.w + cst * c.q2.w, cst * c.q1.x + cst * c.q2.x,
                                  cst * c.q1.y + cst * c.q2.y, cst * c.q1.z + cst * c.q2.z};
                    } else {
                        c.qmid = c.q1;
                    }
                }
                // set_Quaternion_Local_Rotations, get_Force_In_GlobalRef
                Q4 qp[2] = {qprod(qconjugate(c.qmid), c.q1), qprod(qconjugate(c.qmid), c.q2)};
                double th[2][3];
                for (int a = 0; a < 2; ++a) {
                    const double s = qp[a].w, aa = qp[a].x, bb = qp[a].y, cc = qp[a].z;
                    th[a][0] = std::atan2(2 * s * aa + 2 * bb * cc, 1 - 2 * (aa * aa + bb * bb));
                    th[a][1] = std::asin(2 * s * bb - 2 * cc * aa);
                    th[a][2] = std::atan2(2 * s * cc + 2 * aa * bb, 1 - 2 * (bb * bb + cc * cc));
                }
                double Rec[3][3], Rgc[3][3];
                qrotation(qprod(c.qmid, qconjugate(c.qe)), Rec);
                qrotation(c.qmid, Rgc);
                const double r[2][3] = {{c.r1.x, c.r1.y, c.r1.z}, {c.r2.x, c.r2.y, c.r2.z}};
                const double* u[2] = {c.u1, c.u2};
                double Dl[12];
                for (int a = 0; a < 2; ++a) {
                    double w[3];
                    for (int i = 0; i < 3; ++i) {
                        const double dr = Rec[i][0] * r[a][0] + Rec[i][1] * r[a][1] + Rec[i][2] * r[a][2] - r[a][i];
                        w[i] = u[a][i] - c.umid[i] - dr;
                    }
                    for (int i = 0; i < 3; ++i) Dl[6 * a + i] = Rgc[0][i] * w[0] + Rgc[1][i] * w[1] + Rgc[2][i] * w[2];
                    for (int i = 0; i < 3; ++i) Dl[6 * a + 3 + i] = th[a][i];
                }
                double Fl[12], Fg[12];
                for (int i = 0; i < 12; ++i) {
                    double s = 0;
                    for (int j = 0; j < 12; ++j) s += c.Kl[i][j] * Dl[j];
                    Fl[i] = s;
                }
                for (int b = 0; b < 4; ++b)
                    for (int i = 0; i < 3; ++i)
                        Fg[3 * b + i] = Rgc[i][0] * Fl[3 * b] + Rgc[i][1] * Fl[3 * b + 1] + Rgc[i][2] * Fl[3 * b + 2];
                for (int k = 0; k < 6; ++k) {
                    Fint[6 * c.n1 + k] += Fg[k];
                    Fint[6 * c.n2 + k] += Fg[6 + k];
                }
            }
            for (int i = 0; i < nd; ++i)   // apply_SPC1_Forces
                if (B.fixed[i / 6]) Fint[i] = 0;
            for (int i = 0; i < n; ++i) {  // Rotate_Ext_Loads
                double D[3][3];
                qrotation(qmap[i], D);
                double tmp[6];
                for (int r = 0; r < 3; ++r) {
                    tmp[r] = D[r][0] * Forces[6 * i] + D[r][1] * Forces[6 * i + 1] + D[r][2] * Forces[6 * i + 2];
                    tmp[3 + r] = D[r][0] * Forces[6 * i + 3] + D[r][1] * Forces[6 * i + 4] + D[r][2] * Forces[6 * i + 5];
                }
                for (int k = 0; k < 6; ++k) Forces[6 * i + k] = tmp[k];
            }
            for (int i = 0; i < nd; ++i) Fdiff[i] = (step / max_load_step) * Forces[i] - Fint[i];
            for (int i = 0; i < nd; ++i)
                if (B.fixed[i / 6]) Forces[i] = 0;
            solve(Fdiff, dfull);
            double s2 = 0;
            for (int i = 0; i < nd; ++i) s2 += dfull[i] * dfull[i];
            residu = std::sqrt(s2);
            std::fill(Fint.begin(), Fint.end(), 0.0);
            for (int i = 0; i < nd; ++i) damor[i] = residu < 100. * tol ? amor * dfull[i] : dfull[i];
            ++iters;
            if (!std::isfinite(residu) || (iters >= max_iter && residu > tol)) { ok = false; break; }
        } while (residu > tol);
    }
    out3[0] = iters;
    out3[1] = ok && residu <= tol ? 1 : 0;
    out3[2] = residu;
}

void ovl_aeroelastic(int nn, const double* x0, int ns, const int32_t* soff, const int32_t* rn, double aoa,
                     double vinf, double rho, double sref, double cref, int n, const double* bxyz, int ne,
                     const int32_t* en, const double* props, int nclamped, const int32_t* clamped, double tol,
                     int max_iter, double relax, double* disp, double* out) {
    const int N = soff[ns];
    int tip = 0;
    for (int k = 1; k < n; ++k)
        if (bxyz[3 * k + 1] > bxyz[3 * tip + 1]) tip = k;
    std::vector<double> xyz(3 * (size_t)nn), A((size_t)N * N), rhs(N), g(N), rf(3 * (size_t)N), fp(3 * (size_t)N);
    std::vector<double> load(6 * (size_t)n), unew(6 * (size_t)n);
    for (int k = 0; k < N; ++k)
        for (int d = 0; d < 3; ++d) fp[3 * k + d] = 0.5 * (x0[3 * rn[4 * k] + d] + x0[3 * rn[4 * k + 1] + d]);
    std::fill(disp, disp + 6 * n, 0.0);
    double prev = 0.0, cl = 0.0;
    int it = 1;
    for (; it <= max_iter; ++it) {
        ovl_transfer_disp(nn, x0, n, bxyz, ne, en, disp, xyz.data());
        ovl_build(nn, xyz.data(), ns, soff, rn, aoa, vinf, 100.0 * cref, 0.0, A.data(), rhs.data());
        ovl_solve(N, A.data(), rhs.data(), g.data());
        double c[3];
        ovl_forces(nn, xyz.data(), ns, soff, rn, g.data(), aoa, vinf, rho, sref, rf.data(), c);
        cl = c[0];
        ovl_transfer_loads(N, fp.data(), rf.data(), n, bxyz, ne, en, load.data());
        ovl_beam_solve(n, bxyz, ne, en, props, nclamped, clamped, load.data(), unew.data());
        for (int i = 0; i < 6 * n; ++i) disp[i] += relax * (unew[i] - disp[i]);
        double t = disp[6 * tip + 2];
        if (std::fabs(t - prev) <= tol * std::max(std::fabs(t), 1e-300)) break;
        prev = t;
    }
    out[0] = std::min(it, max_iter);
    out[1] = disp[6 * tip + 2];
    out[2] = cl;
}

}  // extern "C"

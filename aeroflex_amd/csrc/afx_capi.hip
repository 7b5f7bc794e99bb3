// C ABI of the MI355X residual (include/afx_rans.h).
//
// Host side of the drop-in boundary: owns the reference-ordered mesh copy
// (like rans::solver owns its mesh, solver.h:179), builds the device layout
// (RCM cell order, per-cell face lists in reference face order), keeps all
// state resident in HBM and translates between the reference's AoS
// Eigen layout and the device's SoA layout only when the caller exchanges
// vectors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/afx_rans.h"
#include "afx_kernels.hpp"
#include "afx_mesh.hpp"
#include "afx_tiles.hpp"
#include "afx_implicit.hpp"
#include "afx_adjoint_kernels.hpp"
#include "afx_mapper.hpp"
#include "afx_post.hpp"

using namespace afx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(AFX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t count) {
        n = count;
        return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    }
    hipError_t upload(const T* h, size_t count, hipStream_t s) {
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
    hipError_t upload(const std::vector<T>& h, hipStream_t s) { return upload(h.data(), h.size(), s); }
};

constexpr int NV = 4;       // mean-flow variables exchanged with the caller (q(4*i+k))
constexpr int NV_MAX = 5;   // device planes: + rho nu~ of the Spalart-Allmaras model

}  // namespace

namespace afx {
// error channel shared by the other C-ABI units (afx_vlm.hip, afx_beam.cpp)
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace afx

struct afx_mesh {
    Mesh m;
};

struct afx_rans {
    Mesh m;
    afx_gas gas{};
    afx_opts opts{};
    hipStream_t stream = nullptr;
    int32_t nC = 0, nT = 0, nF = 0, nB = 0;
    double cfl = 1.0;

    std::vector<int32_t> c_new2old, c_old2new, f_new2old, f_old2new;

    // geometry (FP64 and FP32 copies)
    DBuf<double> xc, yc, area, fnx, fny, flen, fxf, fyf, fgw, lsq;
    DBuf<float> xc32, yc32, area32, fnx32, fny32, flen32, fxf32, fyf32, fgw32, lsq32;
    DBuf<int32_t> c0, c1;
    DBuf<int8_t> fkind;
    DBuf<uint32_t> cface;

    // state and work arrays
    DBuf<double> q, qk, qW, dt, gx, gy, lim, R, flux;
    DBuf<float> q32, gx32, gy32, lim32, R32, flux32;
    DBuf<double> partial, sum1, uni64, wall_d, states;
    DBuf<double> sa_wd;   // the Spalart-Allmaras model's wall distance (compute_wall_distance)
    DBuf<float> uni32;
    DBuf<int32_t> w_owner, w_ghost, b_owner, b_ghost, patch_cells;
    int32_t n_walls = 0;
    int nblk_red = 0;

    // fused tile residual
    TileLayout tl;
    DBuf<int32_t> t_cell0, t_loff, t_ng, t_foff, t_nft, t_goff, t_Lglob, t_nbf;
    DBuf<uint64_t> t_slots;
    DBuf<uint32_t> t_fcell;
    DBuf<int8_t> t_fkind;
    DBuf<double> t_geo64;
    DBuf<float> t_geo32, iarea32, a15_32;
    DBuf<float4> q32v, R32v, t_geo4;   // FP32 fused form: AoS float4 per cell / two per tile face
    DBuf<uint32_t> t_fck;              // c0 | c1 << 12 | kind << 24 per tile face
    bool r32_aos = false;              // last FP32 residual is in R32v (fused) rather than R32
    int last_precision = -1;           // precision of the last residual (for its rho nu~ row)

    // Spalart-Allmaras: variables per cell (4, or 5 with rho nu~ in plane 4 of
    // q / R), freestream rho nu~ = nu_ratio * mu_L, FP32 rho nu~ planes and
    // (area, wall distance) per cell
    int nv = 4;
    double nu_ratio = 3.0;
    bool wd_ready = false;
    DBuf<float> q32n, R32n;
    DBuf<float2> sa_ad32;
    DBuf<float> dt32;   // FP32 local time step of the batched explicit form
    // state adjoint work arrays
    DBuf<double> a_w, a_lam, a_bL, a_bR, a_lgx, a_lgy, a_nbv, a_co, a_cn;
    DBuf<int8_t> a_act;
    DBuf<int32_t> a_jmax, a_jmin, a_nbj;
    DBuf<int32_t> a_bface;   // device indices of the boundary faces, ascending
    DBuf<double> a_vgx, a_vgy, a_vdq;   // laminar terms' face-gradient / difference-quotient adjoints
    // adjoint with respect to the mesh nodes: metric adjoints, node data in
    // device face / cell order, node incidence lists (CSR)
    DBuf<double> g_fbar, g_abar, g_nodex, g_nodey, g_fend, g_fcen, g_cvert, g_xbar;
    DBuf<int32_t> g_fnode, g_cnode, g_noff, g_nent;
    int32_t nN = 0;
    DBuf<float4> q32v_b, q32v_c;   // further FP32 stage states of the batched form
    DBuf<float> q32n_b, q32n_c;

    // implicit step: Jacobian blocks, diagonal blocks and their inverses,
    // GMRES work space and state (kg: the GmDev arrays, kflags its flags)
    DBuf<double> Jf, Dblk, Dinv, kV, kw, kz, kr, ktmp, kpart, kg;
    DBuf<double> ellE, ellD;   // ELL copy of the matrix for the product (launch_assemble_diag)
    DBuf<int32_t> ellN;
    DBuf<double> dilE, dilD;   // colour-ordered copy for the DILU sweeps (launch_dilu_setup)
    DBuf<int32_t> dilN;
    DBuf<int> kbad, kflags;
    int32_t gm_restart = 30, gm_maxit = 500;
    // Arnoldi steps enqueued before the host first reads the flags of a
    // cycle: the column count of the last cycle that converged (its solves
    // tend to need about as many), 8 at first
    int32_t gm_chunk = 8;
    double gm_tol = 1e-2;
    // preconditioner: 0 block Jacobi, 1 multicolour block DILU (colour per
    // real cell, cells by colour, colour offsets on the host, Dt^-1 blocks)
    int precond = 1;
    DBuf<uint8_t> colour;
    DBuf<int32_t> col_cells, col_pos;   // real cells in colour order, and each cell's position in it
    DBuf<double> kdil;                   // the DILU sweeps' work vector, by colour position (nv planes of nC)
    std::vector<int32_t> col_off;
    DBuf<double> Dti;
    GmDev gm_dev() const {
        const int m = gm_restart;
        double* b = kg.p;
        GmDev G;
        G.hdr = b;
        G.H = b + 8;
        G.cs = G.H + (size_t)(m + 1) * m;
        G.sn = G.cs + m;
        G.g = G.sn + m;
        G.y = G.g + (m + 1);
        G.h1 = G.y + m;
        G.h2 = G.h1 + (m + 1);
        G.flags = kflags.p;
        G.tol = gm_tol;
        G.m = m;
        G.maxit = gm_maxit;
        return G;
    }
    static size_t gm_doubles(int m) { return 8 + (size_t)(m + 1) * m + 2 * (size_t)m + (m + 1) + m + 2 * (size_t)(m + 1); }
    bool lhs_ready = false;
    int lhs_bad = 0;
    bool mf = false;             // matrix-free operator (RANS_MATRIX_FREE build)
    DBuf<double> mfFb;           // its unperturbed face fluxes at the current state
    // tangent (Dual1 tile kernel): geometry copies, state, residual, direction
    DBuf<Dual1> tg_geo, tg_area, tg_lsq, tg_xc, tg_yc, tg_wd, tg_q, tg_R;
    DBuf<double> tg_v, tg_out;
    int64_t lin_iters = 0;
    int kblk = 0;
    std::vector<int32_t> t_lg;   // device cell of each local cell (unflagged)
    bool tile64_ok = false, tile32_ok = false;

    // boundary conditions
    bool bcs_set = false;
    std::vector<int32_t> bface_type;       // per boundary face
    std::vector<double> bface_vars;        // 4 per boundary face: mach angle T p
    bool q32_valid = false;

    ~afx_rans() {
        if (stream) (void)hipStreamDestroy(stream);
    }

    CellsT<double> cells64() const {
        return CellsT<double>{xc.p, yc.p, area.p, cface.p, lsq.p, nC, nT};
    }
    CellsT<float> cells32() const {
        return CellsT<float>{xc32.p, yc32.p, area32.p, cface.p, lsq32.p, nC, nT};
    }
    FacesT<double> faces64() const {
        return FacesT<double>{c0.p, c1.p, fnx.p, fny.p, flen.p, fxf.p, fyf.p, fgw.p, fkind.p, nF};
    }
    FacesT<float> faces32() const {
        return FacesT<float>{c0.p, c1.p, fnx32.p, fny32.p, flen32.p, fxf32.p, fyf32.p, fgw32.p, fkind.p, nF};
    }
    GasT<double> gas64() const {
        return GasT<double>{gas.gamma, gas.R, gas.mu_L, 1.0 * gas.mu_L / gas.Pr_L, 1.0, gas.Pr_L, gas.Pr_T};
    }
    GasT<float> gas32() const {
        return GasT<float>{(float)gas.gamma, (float)gas.R, (float)gas.mu_L, (float)(1.0 * gas.mu_L / gas.Pr_L), 1.0f,
                           (float)gas.Pr_L, (float)gas.Pr_T};
    }
    bool sa() const { return nv == 5; }
    bool needs_gradients() const { return opts.second_order || opts.viscosity != 0; }
    bool michalak() const { return opts.second_order && opts.limiter == 1; }
    // second-order flag as the kernels take it: bit 0 second order, bit 1 the
    // Michalak limiter build
    int so() const { return opts.second_order ? (opts.limiter == 1 ? 3 : 1) : 0; }
    TilesF32 tiles32() const {
        return TilesF32{t_cell0.p, t_loff.p, t_ng.p, t_foff.p, t_nft.p, t_goff.p, t_Lglob.p, t_slots.p,
                        t_fck.p, t_geo4.p, t_nbf.p, tl.ntiles, tl.lmax, tl.gmax, tl.ftmax};
    }
    template <typename T>
    TilesT<T> tiles(const T* geo) const {
        return TilesT<T>{t_cell0.p, t_loff.p, t_ng.p, t_foff.p, t_nft.p, t_goff.p, t_Lglob.p, t_slots.p,
                         t_fcell.p, t_fkind.p, geo, t_nbf.p, (int64_t)tl.fref.size(), tl.ntiles, tl.lmax, tl.gmax,
                         tl.ftmax};
    }
};

namespace {

// core.h:73-83
void conservative(const afx_gas& g, const double v[4], double out[4]) {
    const double mach = v[0], angle = v[1], T = v[2], p = v[3];
    double c = std::sqrt(g.gamma * g.R * T);
    double u = mach * c * std::cos(angle);
    double vv = mach * c * std::sin(angle);
    double rho = p / (g.R * T);
    double rhoE = p / (g.gamma - 1) + 0.5 * rho * (u * u + vv * vv);
    out[0] = rho; out[1] = rho * u; out[2] = rho * vv; out[3] = rhoE;
}

// solver.h:597-611: first farfield boundary face's variables (defaults otherwise)
void boundary_variables(const afx_rans* h, double v[4]) {
    v[0] = 0.2; v[1] = 0; v[2] = 1; v[3] = 1.;
    for (int b = 0; b < h->nB; ++b)
        if (h->bface_type[b] == AFX_BC_FARFIELD) {
            for (int k = 0; k < 4; ++k) v[k] = h->bface_vars[4 * b + k];
            return;
        }
}

template <typename T>
std::vector<T> to_device_order_cells(const afx_rans* h, const std::vector<double>& ref) {
    std::vector<T> out(h->nT);
    for (int32_t i = 0; i < h->nT; ++i) out[i] = (T)ref[h->c_new2old[i]];
    return out;
}
template <typename T>
std::vector<T> to_device_order_faces(const afx_rans* h, const std::vector<double>& ref) {
    std::vector<T> out(h->nF);
    for (int32_t i = 0; i < h->nF; ++i) out[i] = (T)ref[h->f_new2old[i]];
    return out;
}

// device SoA [NV][nT] -> reference AoS q(4*i+k)
int download_aos(afx_rans* h, const double* d, double* out, int nv = NV) {
    std::vector<double> tmp((size_t)nv * h->nT);
    HIPCHK(hipMemcpyAsync(tmp.data(), d, tmp.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int32_t i = 0; i < h->nT; ++i) {
        int32_t o = h->c_new2old[i];
        for (int k = 0; k < nv; ++k) out[(size_t)nv * o + k] = tmp[(size_t)k * h->nT + i];
    }
    return AFX_OK;
}
int download_aos32(afx_rans* h, const float* d, double* out) {
    std::vector<float> tmp((size_t)NV * h->nT);
    HIPCHK(hipMemcpyAsync(tmp.data(), d, tmp.size() * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int32_t i = 0; i < h->nT; ++i) {
        int32_t o = h->c_new2old[i];
        for (int k = 0; k < NV; ++k) out[(size_t)NV * o + k] = (double)tmp[(size_t)k * h->nT + i];
    }
    return AFX_OK;
}
int upload_aos(afx_rans* h, const double* in, double* d) {
    std::vector<double> tmp((size_t)NV * h->nT);
    for (int32_t i = 0; i < h->nT; ++i) {
        int32_t o = h->c_new2old[i];
        for (int k = 0; k < NV; ++k) tmp[(size_t)k * h->nT + i] = in[(size_t)NV * o + k];
    }
    HIPCHK(hipMemcpyAsync(d, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}
int download_cells(afx_rans* h, const double* d, double* out) {
    std::vector<double> tmp(h->nT);
    HIPCHK(hipMemcpyAsync(tmp.data(), d, tmp.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int32_t i = 0; i < h->nT; ++i) out[h->c_new2old[i]] = tmp[i];
    return AFX_OK;
}

int norm_of(afx_rans* h, const double* x, int64_t ld, int nrows, double* out) {
    HIPCHK(launch_sumsq(h->stream, x, ld, ld, nrows, h->partial.p, h->nblk_red));
    HIPCHK(launch_final_sum(h->stream, h->partial.p, h->nblk_red, h->sum1.p));
    double s = 0;
    HIPCHK(hipMemcpyAsync(&s, h->sum1.p, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = std::sqrt(s);
    return AFX_OK;
}
int norm_of32(afx_rans* h, const float* x, int64_t ld, int nrows, double* out) {
    HIPCHK(launch_sumsq_f32(h->stream, x, ld, ld, nrows, h->partial.p, h->nblk_red));
    HIPCHK(launch_final_sum(h->stream, h->partial.p, h->nblk_red, h->sum1.p));
    double s = 0;
    HIPCHK(hipMemcpyAsync(&s, h->sum1.p, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = std::sqrt(s);
    return AFX_OK;
}

int ensure_q32(afx_rans* h) {
    if (!h->q32_valid) {
        HIPCHK(launch_to_f32(h->stream, h->q.p, h->q32.p, (int64_t)NV * h->nT));
        if (h->q32v.p) HIPCHK(launch_to_f32x4(h->stream, h->q.p, h->nT, h->q32v.p));
        if (h->q32n.p) HIPCHK(launch_to_f32(h->stream, h->q.p + (size_t)4 * h->nT, h->q32n.p, h->nT));
        h->q32_valid = true;
    }
    return AFX_OK;
}

// set_walls_from_internal + calc_gradients + calc_limiters (solver.h:1083-1087)
int run_gradients64(afx_rans* h, const double* qg, const double* ql) {
    HIPCHK(launch_set_walls<double>(h->stream, h->w_owner.p, h->w_ghost.p, h->n_walls, const_cast<double*>(ql), h->nT, NV));
    HIPCHK(launch_grad_lim<double>(h->stream, h->cells64(), h->faces64(), qg, ql, h->gas64(), h->opts.gradient,
                                   h->so(), h->opts.limiter_k, h->gx.p, h->gy.p, h->lim.p));
    return AFX_OK;
}
int run_gradients32(afx_rans* h) {
    HIPCHK(launch_set_walls<float>(h->stream, h->w_owner.p, h->w_ghost.p, h->n_walls, h->q32.p, h->nT, NV));
    HIPCHK(launch_grad_lim<float>(h->stream, h->cells32(), h->faces32(), h->q32.p, h->q32.p, h->gas32(),
                                  h->opts.gradient, h->so(), (float)h->opts.limiter_k, h->gx32.p,
                                  h->gy32.p, h->lim32.p));
    return AFX_OK;
}

// mesh::compute_wall_dist (mesh.h:794-830) on the device: distance from each
// cell centroid to the nearest wall / slip-wall face centre, capped at 1.
// With the SA model also the model's distance (sa_wd): to the nearest point
// of the same faces' segments, which the face centres overestimate up to a/2
// times on wall cells of aspect ratio a (and the model's destruction ~ 1/d^2
// with it); and the FP32 (area, model distance) pairs.
int compute_wall_distance(afx_rans* h) {
    const Mesh& m = h->m;
    std::vector<double> wx, wy;
    std::vector<double4> seg;
    for (int32_t b = 0; b < h->nB; ++b)
        if (h->bface_type[b] == AFX_BC_WALL || h->bface_type[b] == AFX_BC_SLIP) {
            const uint32_t e = m.boundaryEdges[b];
            wx.push_back(m.edgesCentersX[e]);
            wy.push_back(m.edgesCentersY[e]);
            const uint32_t na = m.edgesNodes[2 * (size_t)e], nb = m.edgesNodes[2 * (size_t)e + 1];
            seg.push_back(make_double4(m.nodesX[na], m.nodesY[na], m.nodesX[nb] - m.nodesX[na],
                                       m.nodesY[nb] - m.nodesY[na]));
        }
    DBuf<double> dwx, dwy;
    HIPCHK(dwx.alloc(wx.size())); HIPCHK(dwx.upload(wx, h->stream));
    HIPCHK(dwy.alloc(wy.size())); HIPCHK(dwy.upload(wy, h->stream));
    HIPCHK(launch_wall_dist(h->stream, h->xc.p, h->yc.p, h->nT, dwx.p, dwy.p, (int32_t)wx.size(), h->wall_d.p));
    DBuf<double4> dseg;
    if (h->sa()) {
        if (!h->sa_wd.p) HIPCHK(h->sa_wd.alloc(h->nT));
        HIPCHK(dseg.alloc(seg.size())); HIPCHK(dseg.upload(seg, h->stream));
        HIPCHK(launch_wall_dist_seg(h->stream, h->xc.p, h->yc.p, h->nT, dseg.p, (int32_t)seg.size(), h->sa_wd.p));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->sa_ad32.p) {
        std::vector<double> d(h->nT);
        HIPCHK(hipMemcpyAsync(d.data(), h->sa_wd.p, d.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        std::vector<float2> ad(h->nT);
        for (int32_t i = 0; i < h->nT; ++i) ad[i] = make_float2((float)m.cellsAreas[h->c_new2old[i]], (float)d[i]);
        HIPCHK(h->sa_ad32.upload(ad, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    h->wd_ready = true;
    return AFX_OK;
}

// enqueue one residual evaluation (no synchronisation)
int enqueue_residual(afx_rans* h, int precision) {
    const bool fused = h->opts.fused != 0;
    h->last_precision = precision;
    if (precision == AFX_FP64 && fused && h->tile64_ok) {
        HIPCHK(launch_residual_tile<double>(h->stream, h->tiles<double>(h->t_geo64.p), h->q.p, h->nT, h->nC, h->area.p,
                                           h->lsq.p, h->xc.p, h->yc.p, h->sa_wd.p, h->gas64(), h->opts.gradient,
                                           h->so(), h->opts.limiter_k, h->opts.viscosity, h->R.p, h->q.p));
        return AFX_OK;
    }
    if (precision == AFX_FP32 && fused && h->tile32_ok && !h->michalak()) {
        int rc = ensure_q32(h);
        if (rc) return rc;
        const double k = h->opts.limiter_k;
        SAFields32 sa{h->q32n.p, h->sa_ad32.p, h->R32n.p, h->q32n.p};
        HIPCHK(launch_residual_tile_f32(h->stream, h->tiles32(), h->q32v.p, h->nC, h->iarea32.p, h->a15_32.p,
                                        h->lsq32.p, h->xc32.p, h->yc32.p, h->gas32(), h->opts.gradient,
                                        h->opts.second_order, (float)(k * k * k), h->opts.viscosity, h->R32v.p,
                                        h->q32v.p, sa));
        h->r32_aos = true;
        return AFX_OK;
    }
    if (h->sa() && precision == AFX_FP32 && h->michalak())
        return fail(AFX_EUNSUP, "the FP32 Spalart-Allmaras residual has no Michalak-limiter form");
    if (h->sa())
        return fail(AFX_EUNSUP, "the Spalart-Allmaras residual runs in the fused tile kernels only "
                                "(renumber = 1, fused = 1, tiles within the LDS limit)");
    if (precision == AFX_FP64) {
        if (h->needs_gradients()) {
            int rc = run_gradients64(h, h->q.p, h->q.p);
            if (rc) return rc;
        }
        HIPCHK(launch_face_flux<double>(h->stream, h->cells64(), h->faces64(), h->q.p, h->q.p, h->gx.p, h->gy.p,
                                        h->lim.p, h->gas64(), h->opts.viscosity, h->opts.second_order, nullptr,
                                        h->flux.p));
        HIPCHK(launch_gather<double>(h->stream, h->cells64(), h->flux.p, h->nF, 0, h->R.p));
    } else {
        int rc = ensure_q32(h);
        if (rc) return rc;
        if (h->needs_gradients()) {
            rc = run_gradients32(h);
            if (rc) return rc;
        }
        HIPCHK(launch_face_flux<float>(h->stream, h->cells32(), h->faces32(), h->q32.p, h->q32.p, h->gx32.p,
                                       h->gy32.p, h->lim32.p, h->gas32(), h->opts.viscosity, h->opts.second_order,
                                       nullptr, h->flux32.p));
        HIPCHK(launch_gather<float>(h->stream, h->cells32(), h->flux32.p, h->nF, 0, h->R32.p));
        h->r32_aos = false;
    }
    return AFX_OK;
}

int check_handle(afx_rans* h, bool need_bcs = true) {
    if (!h) return fail(AFX_EARG, "null solver handle");
    if (need_bcs && !h->bcs_set) return fail(AFX_ESTATE, "boundary conditions not set (call afx_rans_set_bcs)");
    return AFX_OK;
}


// ------------------------------------------------------------ implicit step --
// the Spalart-Allmaras implicit step: the assembled operator (no matrix-free
// form), Green-Gauss gradients (those its Jacobian freezes), the fused
// residual, the wall distance of the source
int sa_implicit_check(afx_rans* h) {
    if (h->mf) return fail(AFX_EUNSUP, "the matrix-free operator covers the four-variable residual");
    if (h->opts.gradient != 1) return fail(AFX_EUNSUP, "the Spalart-Allmaras implicit step covers Green-Gauss gradients");
    if (!h->tile64_ok) return fail(AFX_EUNSUP, "the Spalart-Allmaras residual needs the fused tile kernels");
    if (!h->wd_ready) return compute_wall_distance(h);
    return AFX_OK;
}
// an operand of the implicit operator into ktmp: x AoS (4 per cell), x_nu the
// rho nu~ row (the Spalart-Allmaras model only, where it is required)
int upload_operand(afx_rans* h, const double* x, const double* x_nu, const double* y_nu) {
    if (h->sa() != (x_nu != nullptr) || h->sa() != (y_nu != nullptr))
        return fail(AFX_EARG, h->sa() ? "the Spalart-Allmaras operator needs the rho nu~ rows x_nu and y_nu"
                                      : "rho nu~ rows only with the Spalart-Allmaras model");
    int rc = upload_aos(h, x, h->ktmp.p);
    if (rc || !x_nu) return rc;
    std::vector<double> tmp(h->nT);
    for (int32_t i = 0; i < h->nT; ++i) tmp[i] = x_nu[h->c_new2old[i]];
    HIPCHK(hipMemcpyAsync(h->ktmp.p + 4 * (size_t)h->nT, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}
int download_operand(afx_rans* h, double* y, double* y_nu) {
    int rc = download_aos(h, h->kw.p, y);
    if (rc || !y_nu) return rc;
    return download_cells(h, h->kw.p + 4 * (size_t)h->nT, y_nu);
}

int ensure_implicit(afx_rans* h) {
    if (h->mf && !h->mfFb.p) HIPCHK(h->mfFb.alloc((size_t)4 * h->nF));
    if (h->Jf.p) return AFX_OK;
    // blocks of nv x nv (five with the Spalart-Allmaras model), vectors of nv planes
    const int nv = h->nv;
    const size_t nn = (size_t)nv * nv;
    const size_t n = (size_t)nv * h->nT;
    const int m = h->gm_restart;
    h->kblk = 2048;   // at most eight waves per SIMD in the Gram-Schmidt passes
    HIPCHK(h->Jf.alloc(2 * nn * h->nF));
    HIPCHK(h->Dblk.alloc(nn * h->nC));
    HIPCHK(h->Dinv.alloc(nn * h->nC));
    HIPCHK(h->kV.alloc((size_t)(m + 1) * n));
    HIPCHK(h->kw.alloc(n));
    HIPCHK(h->kz.alloc(n));
    HIPCHK(h->kr.alloc(n));
    HIPCHK(h->ktmp.alloc(n));
    HIPCHK(h->kpart.alloc((size_t)(m + 2) * h->kblk));
    HIPCHK(h->kg.alloc(afx_rans::gm_doubles(m)));
    HIPCHK(h->kflags.alloc(8));
    HIPCHK(h->kbad.alloc(1));
    HIPCHK(h->Dti.alloc(nn * h->nC));
    HIPCHK(h->ellE.alloc(4 * nn * h->nC));
    HIPCHK(h->ellD.alloc(nn * h->nC));
    HIPCHK(h->ellN.alloc(4 * (size_t)h->nC));
    HIPCHK(h->dilE.alloc(4 * nn * h->nC));
    HIPCHK(h->dilD.alloc(nn * h->nC));
    HIPCHK(h->dilN.alloc(4 * (size_t)h->nC));
    HIPCHK(hipMemsetAsync(h->Jf.p, 0, 2 * nn * h->nF * 8, h->stream));
    HIPCHK(hipMemsetAsync(h->kflags.p, 0, 8 * sizeof(int), h->stream));
    // greedy colouring of the real cells in device order over their
    // real-cell face neighbours (at most four), for the DILU sweeps
    {
        const int32_t nC = h->nC;
        std::vector<uint32_t> cf((size_t)4 * nC);
        std::vector<int32_t> c0(h->nF), c1(h->nF);
        HIPCHK(hipMemcpyAsync(cf.data(), h->cface.p, cf.size() * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(c0.data(), h->c0.p, c0.size() * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(c1.data(), h->c1.p, c1.size() * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        std::vector<uint8_t> col(nC, 0xFF);
        int ncol = 0;
        for (int32_t i = 0; i < nC; ++i) {
            unsigned used = 0;
            for (int sl = 0; sl < 4; ++sl) {
                const uint32_t e = cf[(size_t)sl * nC + i];
                if (e == 0xFFFFFFFFu) break;
                const int32_t f = (int32_t)(e & 0x7FFFFFFFu);
                const int32_t j = (e >> 31) ? c0[f] : c1[f];
                if (j < nC && col[j] != 0xFF) used |= 1u << col[j];
            }
            int c = 0;
            while (used & (1u << c)) ++c;   // at most 4 neighbours: c <= 4
            col[i] = (uint8_t)c;
            ncol = std::max(ncol, c + 1);
        }
        h->col_off.assign(ncol + 1, 0);
        for (int32_t i = 0; i < nC; ++i) h->col_off[col[i] + 1]++;
        for (int c = 0; c < ncol; ++c) h->col_off[c + 1] += h->col_off[c];
        std::vector<int32_t> cells(nC), pos(h->col_off.begin(), h->col_off.end() - 1);
        for (int32_t i = 0; i < nC; ++i) cells[pos[col[i]]++] = i;
        HIPCHK(h->colour.alloc(std::max<size_t>(nC, 1)));
        HIPCHK(h->colour.upload(col, h->stream));
        HIPCHK(h->col_cells.alloc(std::max<size_t>(nC, 1)));
        HIPCHK(h->col_cells.upload(cells, h->stream));
        std::vector<int32_t> where(nC);
        for (int32_t p = 0; p < nC; ++p) where[cells[p]] = p;
        HIPCHK(h->col_pos.alloc(std::max<size_t>(nC, 1)));
        HIPCHK(h->col_pos.upload(where, h->stream));
        HIPCHK(h->kdil.alloc(std::max<size_t>((size_t)nv * nC, 1)));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return AFX_OK;
}

// Right-preconditioned restarted GMRES (classical Gram-Schmidt applied twice)
// for LHS x = b; b and x are device vectors of nv planes of nT.  The Arnoldi
// state stays on the device (GmDev): the steps of a restart cycle are
// enqueued in chunks, the kernels of a step do nothing once the iteration
// has converged or broken down, and between chunks the host reads the flags
// and ends the cycle early once the Arnoldi process is inactive (the steps
// it skips would do nothing, so the result is the same as enqueueing all m).
// The first chunk has as many steps as the last converged cycle took, later
// ones four.  Returns AFX_OK and *converged, or a negative error code.
int gmres(afx_rans* h, const double* b, double* x, bool* converged) {
    const int64_t n = (int64_t)h->nv * h->nT;
    const int m = h->gm_restart;
    // Gram-Schmidt blocks: about two entries per thread, at most kblk
    const int nb = (int)std::min<int64_t>(h->kblk, std::max<int64_t>(8, (n + 511) / 512));
    hipStream_t s = h->stream;
    double* V = h->kV.p;
    double* part = h->kpart.p;
    const GmDev G = h->gm_dev();
    const int* active = G.flags;
    *converged = false;
    // the operator and the right preconditioner: the assembled matrix with
    // block Jacobi or block DILU, or the matrix-free product with none
    // (solver.h:868)
    auto apply = [&](const double* in, double* out, const int* act) -> hipError_t {
        if (h->mf)
            return launch_mf_apply(s, h->cells64(), h->faces64(), h->q.p, h->dt.p, h->mfFb.p, h->gas64(),
                                   h->opts.viscosity, in, out, act);
        return launch_lhs_apply(s, h->nv, h->nC, h->nT, h->ellD.p, h->ellE.p, h->ellN.p, in, out, act);
    };
    auto precond = [&](const double* in, double* out, const double** res, const int* act) -> hipError_t {
        if (h->mf) { *res = in; return hipSuccess; }
        *res = out;
        if (h->precond == 1)
            return launch_dilu_apply(s, h->nv, h->nC, h->nT, h->col_cells.p, h->col_pos.p, h->col_off.data(),
                                     (int)h->col_off.size() - 1, h->dilE.p, h->dilN.p, h->dilD.p, in, out, h->kdil.p,
                                     act);
        return launch_block_jacobi(s, h->nv, h->nC, h->nT, h->Dinv.p, in, out, act);
    };
    // |r|^2 of the cycle's start vector into hdr[3], then the cycle's set-up
    auto begin = [&](int first) -> hipError_t {
        hipError_t e = launch_gs_pass(s, nullptr, n, 0, nullptr, h->kr.p, n, false, part, nb, nullptr);
        if (e == hipSuccess) e = launch_gm_reduce(s, part, nb, 0, 1, G.hdr + 3, nullptr);
        if (e == hipSuccess) e = launch_gm_begin(s, G, first);
        if (e == hipSuccess) e = launch_scale_dev(s, G.hdr + 2, h->kr.p, V, n, active);
        return e;
    };
    const double* pz = nullptr;
    HIPCHK(hipMemsetAsync(x, 0, n * 8, s));
    HIPCHK(launch_copy(s, b, h->kr.p, n));
    HIPCHK(begin(1));
    int flags[4] = {0, 0, 0, 0};
    int total_before = 0;
    while (true) {
        int chunk = std::max(1, std::min<int>(h->gm_chunk, m));
        for (int j = 0; j < m; ++j) {
            if (j == chunk) {   // a chunk is enqueued: is the Arnoldi process still active?
                HIPCHK(hipMemcpyAsync(flags, G.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if (!flags[0]) break;
                chunk += 4;
            }
            double* vj = V + (size_t)j * n;
            HIPCHK(precond(vj, h->kz.p, &pz, active));
            HIPCHK(apply(pz, h->kw.p, active));
            HIPCHK(launch_gs_pass(s, V, n, j + 1, nullptr, h->kw.p, n, true, part, nb, active));
            HIPCHK(launch_gm_reduce(s, part, nb, 0, j + 1, G.h1, active));
            HIPCHK(launch_gs_pass(s, V, n, j + 1, G.h1, h->kw.p, n, true, part, nb, active));
            HIPCHK(launch_gm_reduce(s, part, nb, 0, j + 1, G.h2, active));
            HIPCHK(launch_gs_pass(s, V, n, j + 1, G.h2, h->kw.p, n, false, part, nb, active));
            HIPCHK(launch_gm_reduce(s, part, nb, j + 1, 1, G.hdr + 3, active));
            HIPCHK(launch_gm_step(s, G, j));
            if (j + 1 < m) HIPCHK(launch_scale_dev(s, G.hdr + 2, h->kw.p, V + (size_t)(j + 1) * n, n, active));
        }
        // y = H^-1 g, x += M^-1 (V y)
        HIPCHK(launch_gm_solve(s, G));
        HIPCHK(launch_lincomb_dev(s, V, n, G.flags + 3, G.y, h->ktmp.p, n));
        HIPCHK(precond(h->ktmp.p, h->kz.p, &pz, nullptr));
        HIPCHK(launch_axpy(s, 1.0, pz, x, n));
        HIPCHK(hipMemcpyAsync(flags, G.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        h->lin_iters += flags[2] - total_before;
        total_before = flags[2];
        if (flags[1]) {
            h->gm_chunk = std::max(1, flags[3]);
            *converged = true;
            return AFX_OK;
        }
        if (flags[2] >= h->gm_maxit) return AFX_OK;
        // restart from the true residual
        HIPCHK(apply(x, h->kw.p, nullptr));
        HIPCHK(launch_sub(s, b, h->kw.p, h->kr.p, n));
        HIPCHK(begin(0));   // converged here: the next cycle's kernels do nothing
    }
}

}  // namespace

extern "C" {

const char* afx_last_error(void) { return g_err.c_str(); }

void afx_default_gas(afx_gas* g) {
    g->R = 0.71428571428; g->mu_L = 1e-5; g->Pr_L = 0.72; g->Pr_T = 0.9; g->gamma = 1.4;
}
void afx_default_opts(afx_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->viscosity = 0; o->second_order = 1; o->gradient = 1; o->limiter_k = 5.0; o->renumber = 1;
    o->tile_cells = 0; o->fused = 1;
}

// ------------------------------------------------------------------ mesh ---
int afx_mesh_read_msh(const char* path, afx_mesh** out) {
    if (!path || !out) return fail(AFX_EARG, "null argument");
    try {
        Elements el = read_msh41(path);
        auto* m = new afx_mesh();
        m->m = build_mesh(el);
        *out = m;
        return AFX_OK;
    } catch (std::exception& e) {
        return fail(AFX_EIO, e.what());
    }
}

int afx_mesh_from_elements(int64_t n_nodes, const double* x, const double* y, int64_t n_cells, const int32_t* cell_nodes,
                           const int32_t* cell_nverts, int64_t n_bsegs, const int32_t* bseg_nodes,
                           const int32_t* bseg_keys, int32_t n_names, const int32_t* name_keys,
                           const char* const* names, afx_mesh** out) {
    if (!out || n_nodes <= 0 || n_cells <= 0 || !x || !y || !cell_nodes || !cell_nverts) return fail(AFX_EARG, "bad mesh arrays");
    try {
        Elements el;
        el.x.assign(x, x + n_nodes);
        el.y.assign(y, y + n_nodes);
        for (int64_t c = 0; c < n_cells; ++c) {
            int nv = cell_nverts[c];
            if (nv != 3 && nv != 4) return fail(AFX_EARG, "cells must have 3 or 4 nodes");
            for (int k = 0; k < 4; ++k) {
                int32_t v = (k < nv) ? cell_nodes[4 * c + k] : 0;
                if (v < 0 || v >= n_nodes) return fail(AFX_EARG, "cell node index out of range");
                el.cell_nodes.push_back((uint32_t)v);
            }
            el.cell_nv.push_back((uint8_t)nv);
        }
        for (int64_t b = 0; b < n_bsegs; ++b) {
            for (int k = 0; k < 2; ++k) {
                int32_t v = bseg_nodes[2 * b + k];
                if (v < 0 || v >= n_nodes) return fail(AFX_EARG, "boundary node index out of range");
                el.bseg_nodes.push_back((uint32_t)v);
            }
            el.bseg_key.push_back((uint32_t)bseg_keys[b]);
        }
        for (int32_t i = 0; i < n_names; ++i) {
            el.name_keys.push_back((uint32_t)name_keys[i]);
            el.names.push_back(names[i]);
        }
        auto* m = new afx_mesh();
        m->m = build_mesh(el);
        *out = m;
        return AFX_OK;
    } catch (std::exception& e) {
        return fail(AFX_EARG, e.what());
    }
}

int afx_mesh_gen_naca_ogrid(int32_t ni, int32_t nj, double r_far, double thickness, double first_h, afx_mesh** out) {
    if (!out || ni < 8 || nj < 3 || r_far <= 1.0 || first_h <= 0) return fail(AFX_EARG, "bad O-grid parameters");
    try {
        auto* m = new afx_mesh();
        m->m = build_mesh(gen_naca_ogrid(ni, nj, r_far, thickness, first_h));
        *out = m;
        return AFX_OK;
    } catch (std::exception& e) {
        return fail(AFX_EARG, e.what());
    }
}

int afx_mesh_gen_jittered(int32_t ni, int32_t nj, double r_far, double thickness, double first_h, double jitter,
                          double quad_frac, uint64_t seed, afx_mesh** out) {
    if (!out || ni < 8 || nj < 3 || r_far <= 1.0 || first_h <= 0 || jitter < 0 || jitter > 0.2 || quad_frac < 0 ||
        quad_frac > 1)
        return fail(AFX_EARG, "bad jittered-mesh parameters");
    try {
        auto* m = new afx_mesh();
        m->m = build_mesh(gen_jittered_airfoil(ni, nj, r_far, thickness, first_h, jitter, quad_frac, seed));
        *out = m;
        return AFX_OK;
    } catch (std::exception& e) {
        return fail(AFX_EARG, e.what());
    }
}

int afx_mesh_write_msh(const afx_mesh* mh, const char* path) {
    if (!mh || !path) return fail(AFX_EARG, "null argument");
    try {
        const Mesh& m = mh->m;
        Elements el;
        el.x = m.nodesX; el.y = m.nodesY;
        el.name_keys = m.name_keys; el.names = m.names;
        for (uint32_t c = 0; c < m.nRealCells; ++c) {
            for (int k = 0; k < 4; ++k) el.cell_nodes.push_back(m.cellsNodes[4 * (size_t)c + k]);
            el.cell_nv.push_back(m.cellsIsTriangle[c] ? 3 : 4);
        }
        // boundary segments in their original node order are not kept; the
        // face's node pair (creation order) spans the same edge
        for (uint32_t b = 0; b < m.n_bfaces(); ++b) {
            uint32_t e = m.boundaryEdges[b];
            el.bseg_nodes.push_back(m.edgesNodes[2 * (size_t)e]);
            el.bseg_nodes.push_back(m.edgesNodes[2 * (size_t)e + 1]);
            el.bseg_key.push_back(m.boundaryEdgesIntTag[b]);
        }
        write_msh41(el, path);
        return AFX_OK;
    } catch (std::exception& e) {
        return fail(AFX_EIO, e.what());
    }
}

int afx_mesh_counts(const afx_mesh* mh, int64_t c[5]) {
    if (!mh || !c) return fail(AFX_EARG, "null argument");
    const Mesh& m = mh->m;
    c[0] = m.n_nodes(); c[1] = m.nRealCells; c[2] = m.n_cells(); c[3] = m.n_faces(); c[4] = m.n_bfaces();
    return AFX_OK;
}

int afx_mesh_get(const afx_mesh* mh, const char* field, void* out, int64_t* n_out) {
    if (!mh || !field) return fail(AFX_EARG, "null argument");
    const Mesh& m = mh->m;
    std::string f(field);
    const void* src = nullptr;
    size_t n = 0, esz = 8;
    auto D = [&](const std::vector<double>& v) { src = v.data(); n = v.size(); esz = 8; };
    auto U = [&](const std::vector<uint32_t>& v) { src = v.data(); n = v.size(); esz = 4; };
    if (f == "nodesX") D(m.nodesX);
    else if (f == "nodesY") D(m.nodesY);
    else if (f == "edgesNodes") U(m.edgesNodes);
    else if (f == "edgesCells") U(m.edgesCells);
    else if (f == "edgesLengths") D(m.edgesLengths);
    else if (f == "edgesNormalsX") D(m.edgesNormalsX);
    else if (f == "edgesNormalsY") D(m.edgesNormalsY);
    else if (f == "edgesCentersX") D(m.edgesCentersX);
    else if (f == "edgesCentersY") D(m.edgesCentersY);
    else if (f == "cellsNodes") U(m.cellsNodes);
    else if (f == "cellsEdges") U(m.cellsEdges);
    else if (f == "cellsAreas") D(m.cellsAreas);
    else if (f == "cellsCentersX") D(m.cellsCentersX);
    else if (f == "cellsCentersY") D(m.cellsCentersY);
    else if (f == "boundaryEdges") U(m.boundaryEdges);
    else if (f == "boundaryEdgesIntTag") U(m.boundaryEdgesIntTag);
    else if (f == "cellsIsTriangle") { src = m.cellsIsTriangle.data(); n = m.cellsIsTriangle.size(); esz = 4; }
    else return fail(AFX_EARG, "unknown mesh field: " + f);
    if (n_out) *n_out = (int64_t)n;
    if (out && n) std::memcpy(out, src, n * esz);
    return AFX_OK;
}

int afx_mesh_boundary_name(const afx_mesh* mh, int64_t b, char* buf, int32_t buflen) {
    if (!mh || !buf || buflen <= 0) return fail(AFX_EARG, "null argument");
    if (b < 0 || b >= (int64_t)mh->m.n_bfaces()) return fail(AFX_EARG, "boundary index out of range");
    const std::string& s = mh->m.names[mh->m.boundaryNameIdx[b]];
    if ((int32_t)s.size() + 1 > buflen) return fail(AFX_EARG, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return AFX_OK;
}

int afx_mesh_rcm(const afx_mesh* mh, int32_t* perm) {
    if (!mh || !perm) return fail(AFX_EARG, "null argument");
    auto p = rcm_order(mh->m);
    for (size_t i = 0; i < p.size(); ++i) perm[i] = (int32_t)p[i];
    return AFX_OK;
}

int afx_mesh_tile_stats(const afx_mesh* mh, int32_t tile_cells, int64_t st[8]) {
    if (!mh || !st) return fail(AFX_EARG, "null argument");
    try {
        TileLayout tl = build_tiles(mh->m, rcm_order(mh->m), tile_cells);
        st[0] = tl.ntiles; st[1] = tl.lmax; st[2] = tl.gmax; st[3] = tl.fmax; st[4] = tl.ftmax;
        st[5] = (int64_t)tl.L_ref.size(); st[6] = (int64_t)tl.slots.size(); st[7] = (int64_t)tl.fref.size();
        return AFX_OK;
    } catch (std::exception& e) {
        return fail(AFX_EARG, e.what());
    }
}

void afx_mesh_destroy(afx_mesh* m) { delete m; }

// ---------------------------------------------------------------- solver ---
namespace {
int create_impl(const Mesh& mesh_in, const afx_gas* g, const afx_opts* o, const afx_rans* like, afx_rans** out);
}

int afx_rans_create(const afx_mesh* mh, const afx_gas* g, const afx_opts* o, afx_rans** out) {
    if (!mh || !out) return fail(AFX_EARG, "null argument");
    return create_impl(mh->m, g, o, nullptr, out);
}

int afx_rans_create_like(const afx_rans* base, const afx_gas* g, afx_rans** out) {
    if (!base || !out) return fail(AFX_EARG, "null argument");
    return create_impl(base->m, g, &base->opts, base, out);
}

namespace {
int create_impl(const Mesh& mesh_in, const afx_gas* g, const afx_opts* o, const afx_rans* like, afx_rans** out) {
    afx_opts opts;
    afx_default_opts(&opts);
    if (o) opts = *o;
    if (opts.viscosity < 0 || opts.viscosity > 2)
        return fail(AFX_EUNSUP, "viscosity model must be 0 (inviscid), 1 (laminar) or 2 (Spalart-Allmaras)");
    if (opts.gradient != 0 && opts.gradient != 1) return fail(AFX_EARG, "gradient must be 0 (least-squares) or 1 (green-gauss)");
    if (opts.limiter != 0 && opts.limiter != 1) return fail(AFX_EARG, "limiter must be 0 (default) or 1 (Michalak)");
    auto h = new afx_rans();
    try {
        h->m = mesh_in;
    } catch (std::exception& e) {
        delete h;
        return fail(AFX_EARG, e.what());
    }
    if (g) h->gas = *g; else afx_default_gas(&h->gas);
    h->opts = opts;
    h->nv = opts.viscosity == 2 ? 5 : 4;
    const Mesh& m = h->m;
    h->nC = (int32_t)m.nRealCells; h->nT = (int32_t)m.n_cells(); h->nF = (int32_t)m.n_faces(); h->nB = (int32_t)m.n_bfaces();
    const int32_t nC = h->nC, nT = h->nT, nF = h->nF;

    auto bail = [&](int rc) { delete h; return rc; };
#define HIPCHK_C(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return bail(fail(AFX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_))); \
    } while (0)

    HIPCHK_C(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));

    // cell order: tiles grown from RCM seeds (or file order), ghosts keep
    // reference order
    h->c_new2old.resize(nT);
    if (like) {
        h->tl = like->tl;   // same mesh and options: same tiles and cell order
        for (int32_t i = 0; i < nC; ++i) h->c_new2old[i] = like->c_new2old[i];
    } else {
        std::vector<uint32_t> seeds;
        if (opts.renumber) seeds = rcm_order(m);
        else { seeds.resize(nC); std::iota(seeds.begin(), seeds.end(), 0u); }
        // automatic size: the largest tile (from 256 cells down) whose FP64
        // working set fits the 64 KiB of LDS one workgroup may use
        int32_t target = opts.tile_cells > 0 ? opts.tile_cells : 256;
        try {
            h->tl = build_tiles(m, seeds, target);
            while (opts.tile_cells <= 0 && target > 64 &&
                   tile_lds_bytes<double>(h->tl.lmax, h->tl.gmax, h->tl.ftmax, h->nv) > 65536) {
                target -= 32;
                h->tl = build_tiles(m, seeds, target);
            }
        } catch (std::exception& e) {
            return bail(fail(AFX_EARG, e.what()));
        }
        if (opts.renumber)
            for (int32_t i = 0; i < nC; ++i) h->c_new2old[i] = h->tl.cell_new2old[i];
        else
            for (int32_t i = 0; i < nC; ++i) h->c_new2old[i] = i;
    }
    for (int32_t i = nC; i < nT; ++i) h->c_new2old[i] = i;
    h->c_old2new.resize(nT);
    for (int32_t i = 0; i < nT; ++i) h->c_old2new[h->c_new2old[i]] = i;

    // face order: by the lowest device index of the cells it touches
    std::vector<int64_t> key(nF);
    for (int32_t f = 0; f < nF; ++f) {
        int32_t a = h->c_old2new[m.edgesCells[2 * (size_t)f]], b = h->c_old2new[m.edgesCells[2 * (size_t)f + 1]];
        key[f] = (b < nC) ? std::min(a, b) : a;
    }
    h->f_new2old.resize(nF);
    std::iota(h->f_new2old.begin(), h->f_new2old.end(), 0);
    std::stable_sort(h->f_new2old.begin(), h->f_new2old.end(), [&](int32_t x, int32_t y) { return key[x] < key[y]; });
    h->f_old2new.resize(nF);
    for (int32_t i = 0; i < nF; ++i) h->f_old2new[h->f_new2old[i]] = i;

    // per-face arrays in device order
    std::vector<int32_t> hc0(nF), hc1(nF);
    std::vector<double> hgw(nF);
    for (int32_t nf = 0; nf < nF; ++nf) {
        int32_t f = h->f_new2old[nf];
        uint32_t i = m.edgesCells[2 * (size_t)f], j = m.edgesCells[2 * (size_t)f + 1];
        hc0[nf] = h->c_old2new[i];
        hc1[nf] = h->c_old2new[j];
        // solver.h:436-444
        const double dxif = m.edgesCentersX[f] - m.cellsCentersX[i];
        const double dyif = m.edgesCentersY[f] - m.cellsCentersY[i];
        const double dif = std::sqrt(dxif * dxif + dyif * dyif);
        const double dxij = m.cellsCentersX[i] - m.cellsCentersX[j];
        const double dyij = m.cellsCentersY[i] - m.cellsCentersY[j];
        const double dij = std::sqrt(dxij * dxij + dyij * dyij);
        hgw[nf] = dif / dij;
    }
    // per-cell face lists sorted by reference face index
    std::vector<uint32_t> hcface((size_t)4 * nC, 0xFFFFFFFFu);
    for (int32_t ci = 0; ci < nC; ++ci) {
        int32_t old = h->c_new2old[ci];
        uint32_t fs[4];
        int n = 0;
        for (int s = 0; s < 4; ++s) {
            uint32_t e = m.cellsEdges[4 * (size_t)old + s];
            if (e != kNoEdge) fs[n++] = e;
        }
        std::sort(fs, fs + n);
        for (int s = 0; s < n; ++s) {
            uint32_t f = fs[s];
            uint32_t side = (m.edgesCells[2 * (size_t)f] == (uint32_t)old) ? 0u : 1u;
            hcface[(size_t)s * nC + ci] = (uint32_t)h->f_old2new[f] | (side << 31);
        }
    }
    // least-squares matrices (solver.h:402-422), cellsEdges order
    std::vector<double> hlsq((size_t)4 * nC, 0.0);
    for (int32_t ci = 0; ci < nC; ++ci) {
        int32_t i = h->c_new2old[ci];
        int n = m.cellsIsTriangle[i] ? 3 : 4;
        double a = 0, b = 0, d = 0;
        for (int j = 0; j < n; ++j) {
            uint32_t e = m.cellsEdges[4 * (size_t)i + j];
            uint32_t cn = m.edgesCells[2 * (size_t)e] == (uint32_t)i ? m.edgesCells[2 * (size_t)e + 1] : m.edgesCells[2 * (size_t)e];
            double dx = m.cellsCentersX[cn] - m.cellsCentersX[i];
            double dy = m.cellsCentersY[cn] - m.cellsCentersY[i];
            a += dx * dx; b += dx * dy; d += dy * dy;
        }
        double det = a * d - b * b;
        hlsq[ci] = d / det; hlsq[(size_t)nC + ci] = -b / det;
        hlsq[2 * (size_t)nC + ci] = -b / det; hlsq[3 * (size_t)nC + ci] = a / det;
    }

    hipStream_t s = h->stream;
    // geometry
    HIPCHK_C(h->xc.alloc(nT)); HIPCHK_C(h->xc.upload(to_device_order_cells<double>(h, m.cellsCentersX), s));
    HIPCHK_C(h->yc.alloc(nT)); HIPCHK_C(h->yc.upload(to_device_order_cells<double>(h, m.cellsCentersY), s));
    HIPCHK_C(h->area.alloc(nT)); HIPCHK_C(h->area.upload(to_device_order_cells<double>(h, m.cellsAreas), s));
    HIPCHK_C(h->xc32.alloc(nT)); HIPCHK_C(h->xc32.upload(to_device_order_cells<float>(h, m.cellsCentersX), s));
    HIPCHK_C(h->yc32.alloc(nT)); HIPCHK_C(h->yc32.upload(to_device_order_cells<float>(h, m.cellsCentersY), s));
    HIPCHK_C(h->area32.alloc(nT)); HIPCHK_C(h->area32.upload(to_device_order_cells<float>(h, m.cellsAreas), s));
    HIPCHK_C(h->fnx.alloc(nF)); HIPCHK_C(h->fnx.upload(to_device_order_faces<double>(h, m.edgesNormalsX), s));
    HIPCHK_C(h->fny.alloc(nF)); HIPCHK_C(h->fny.upload(to_device_order_faces<double>(h, m.edgesNormalsY), s));
    HIPCHK_C(h->flen.alloc(nF)); HIPCHK_C(h->flen.upload(to_device_order_faces<double>(h, m.edgesLengths), s));
    HIPCHK_C(h->fxf.alloc(nF)); HIPCHK_C(h->fxf.upload(to_device_order_faces<double>(h, m.edgesCentersX), s));
    HIPCHK_C(h->fyf.alloc(nF)); HIPCHK_C(h->fyf.upload(to_device_order_faces<double>(h, m.edgesCentersY), s));
    HIPCHK_C(h->fnx32.alloc(nF)); HIPCHK_C(h->fnx32.upload(to_device_order_faces<float>(h, m.edgesNormalsX), s));
    HIPCHK_C(h->fny32.alloc(nF)); HIPCHK_C(h->fny32.upload(to_device_order_faces<float>(h, m.edgesNormalsY), s));
    HIPCHK_C(h->flen32.alloc(nF)); HIPCHK_C(h->flen32.upload(to_device_order_faces<float>(h, m.edgesLengths), s));
    HIPCHK_C(h->fxf32.alloc(nF)); HIPCHK_C(h->fxf32.upload(to_device_order_faces<float>(h, m.edgesCentersX), s));
    HIPCHK_C(h->fyf32.alloc(nF)); HIPCHK_C(h->fyf32.upload(to_device_order_faces<float>(h, m.edgesCentersY), s));
    HIPCHK_C(h->fgw.alloc(nF)); HIPCHK_C(h->fgw.upload(hgw, s));
    {
        std::vector<float> g32(hgw.begin(), hgw.end());
        HIPCHK_C(h->fgw32.alloc(nF)); HIPCHK_C(h->fgw32.upload(g32, s));
        std::vector<float> l32(hlsq.begin(), hlsq.end());
        HIPCHK_C(h->lsq.alloc(hlsq.size())); HIPCHK_C(h->lsq.upload(hlsq, s));
        HIPCHK_C(h->lsq32.alloc(l32.size())); HIPCHK_C(h->lsq32.upload(l32, s));
        HIPCHK_C(hipStreamSynchronize(s));
    }
    HIPCHK_C(h->c0.alloc(nF)); HIPCHK_C(h->c0.upload(hc0, s));
    HIPCHK_C(h->c1.alloc(nF)); HIPCHK_C(h->c1.upload(hc1, s));
    HIPCHK_C(h->cface.alloc(hcface.size())); HIPCHK_C(h->cface.upload(hcface, s));
    {
        std::vector<int8_t> kind(nF, K_INTERNAL);
        HIPCHK_C(h->fkind.alloc(nF)); HIPCHK_C(h->fkind.upload(kind, s));
        HIPCHK_C(hipStreamSynchronize(s));
    }
    // state arrays (NV_MAX planes; the fifth holds rho nu~ with the SA model)
    const size_t nq = (size_t)NV_MAX * nT;
    for (DBuf<double>* b : {&h->q, &h->qk, &h->qW, &h->gx, &h->gy, &h->lim, &h->R}) {
        HIPCHK_C(b->alloc(nq));
        HIPCHK_C(hipMemsetAsync(b->p, 0, nq * 8, s));
    }
    for (DBuf<float>* b : {&h->q32, &h->gx32, &h->gy32, &h->lim32, &h->R32}) {
        HIPCHK_C(b->alloc(nq));
        HIPCHK_C(hipMemsetAsync(b->p, 0, nq * 4, s));
    }
    HIPCHK_C(h->dt.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->dt.p, 0, (size_t)nT * 8, s));
    HIPCHK_C(h->wall_d.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->wall_d.p, 0, (size_t)nT * 8, s));
    HIPCHK_C(h->flux.alloc((size_t)NV * nF));
    HIPCHK_C(h->flux32.alloc((size_t)NV * nF));
    // ghost limiters are 1 and ghost gradients 0 (solver.h:466-469,519,540)
    {
        std::vector<double> ones(nq, 1.0);
        std::vector<float> ones32(nq, 1.0f);
        HIPCHK_C(h->lim.upload(ones, s));
        HIPCHK_C(h->lim32.upload(ones32, s));
        HIPCHK_C(hipStreamSynchronize(s));
    }
    // tile tables (only usable when the cell order is the tile order)
    if (opts.renumber) {
        const TileLayout& tl = h->tl;
        std::vector<int32_t> lg(tl.L_ref.size());
        for (size_t i = 0; i < lg.size(); ++i) lg[i] = h->c_old2new[tl.L_ref[i]];
        h->t_lg = lg;
        {
            // FP32 form: 1/A and A^1.5 per cell (K3a = K^3 A^1.5, solver.h:551-552)
            std::vector<float> ia(nT), a15(nT);
            for (int32_t i = 0; i < nT; ++i) {
                double A = m.cellsAreas[h->c_new2old[i]];
                ia[i] = (float)(1.0 / A);
                a15[i] = (float)(A * std::sqrt(A));
            }
            HIPCHK_C(h->iarea32.alloc(nT)); HIPCHK_C(h->iarea32.upload(ia, s));
            HIPCHK_C(h->a15_32.alloc(nT)); HIPCHK_C(h->a15_32.upload(a15, s));
            HIPCHK_C(hipStreamSynchronize(s));
        }
        const size_t nft = tl.fref.size();
        std::vector<double> geo(8 * nft);
        std::vector<float> geo32(8 * nft);
        for (size_t i = 0; i < nft; ++i) {
            int32_t f = tl.fref[i];
            uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
            double v[8];
            v[0] = m.edgesNormalsX[f]; v[1] = m.edgesNormalsY[f]; v[2] = m.edgesLengths[f];
            v[3] = hgw[h->f_old2new[f]];
            v[4] = m.edgesCentersX[f] - m.cellsCentersX[a];   // solver.h:1122-1123
            v[5] = m.edgesCentersY[f] - m.cellsCentersY[a];
            v[6] = m.edgesCentersX[f] - m.cellsCentersX[b];   // solver.h:1125-1126
            v[7] = m.edgesCentersY[f] - m.cellsCentersY[b];
            for (int k = 0; k < 8; ++k) { geo[k * nft + i] = v[k]; geo32[k * nft + i] = (float)v[k]; }
        }
        HIPCHK_C(h->t_cell0.alloc(tl.cell0.size())); HIPCHK_C(h->t_cell0.upload(tl.cell0, s));
        HIPCHK_C(h->t_loff.alloc(tl.loff.size())); HIPCHK_C(h->t_loff.upload(tl.loff, s));
        HIPCHK_C(h->t_ng.alloc(tl.ng.size())); HIPCHK_C(h->t_ng.upload(tl.ng, s));
        HIPCHK_C(h->t_foff.alloc(tl.foff.size())); HIPCHK_C(h->t_foff.upload(tl.foff, s));
        HIPCHK_C(h->t_nft.alloc(tl.nft.size())); HIPCHK_C(h->t_nft.upload(tl.nft, s));
        HIPCHK_C(h->t_nbf.alloc(tl.nbf.size())); HIPCHK_C(h->t_nbf.upload(tl.nbf, s));
        HIPCHK_C(h->t_goff.alloc(tl.goff.size())); HIPCHK_C(h->t_goff.upload(tl.goff, s));
        HIPCHK_C(h->t_Lglob.alloc(lg.size())); HIPCHK_C(h->t_Lglob.upload(lg, s));
        HIPCHK_C(h->t_slots.alloc(tl.slots.size())); HIPCHK_C(h->t_slots.upload(tl.slots, s));
        HIPCHK_C(h->t_fcell.alloc(tl.fcell.size())); HIPCHK_C(h->t_fcell.upload(tl.fcell, s));
        HIPCHK_C(h->t_fkind.alloc(nft));
        HIPCHK_C(hipMemsetAsync(h->t_fkind.p, 0, std::max<size_t>(nft, 1), s));
        HIPCHK_C(h->t_geo64.alloc(geo.size())); HIPCHK_C(h->t_geo64.upload(geo, s));
        HIPCHK_C(h->t_geo32.alloc(geo32.size())); HIPCHK_C(h->t_geo32.upload(geo32, s));
        {
            std::vector<float4> g4(2 * nft);
            std::vector<uint32_t> fck(nft);
            for (size_t i = 0; i < nft; ++i) {
                g4[2 * i] = make_float4(geo32[i], geo32[nft + i], geo32[2 * nft + i], geo32[3 * nft + i]);
                g4[2 * i + 1] = make_float4(geo32[4 * nft + i], geo32[5 * nft + i], geo32[6 * nft + i], geo32[7 * nft + i]);
                uint32_t c = tl.fcell[i];
                fck[i] = (c & 0xFFFFu) | ((c >> 16) << 12);   // kinds are added by set_bcs
            }
            HIPCHK_C(h->t_geo4.alloc(g4.size())); HIPCHK_C(h->t_geo4.upload(g4, s));
            HIPCHK_C(h->t_fck.alloc(nft)); HIPCHK_C(h->t_fck.upload(fck, s));
            HIPCHK_C(h->q32v.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->q32v.p, 0, (size_t)nT * 16, s));
            HIPCHK_C(h->R32v.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->R32v.p, 0, (size_t)nT * 16, s));
            HIPCHK_C(h->dt32.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->dt32.p, 0, (size_t)nT * 4, s));
            if (h->sa()) {
                HIPCHK_C(h->q32n.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->q32n.p, 0, (size_t)nT * 4, s));
                HIPCHK_C(h->R32n.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->R32n.p, 0, (size_t)nT * 4, s));
                HIPCHK_C(h->sa_ad32.alloc(nT)); HIPCHK_C(hipMemsetAsync(h->sa_ad32.p, 0, (size_t)nT * 8, s));
            }
            HIPCHK_C(hipStreamSynchronize(s));
        }
        HIPCHK_C(hipStreamSynchronize(s));
        h->tile64_ok = tile_lds_bytes<double>(tl.lmax, tl.gmax, tl.ftmax, h->nv) <= 65536;
        h->tile32_ok = tile_lds_bytes<float>(tl.lmax, tl.gmax, tl.ftmax, h->nv) <= 65536;
    }
    h->nblk_red = 1024;
    HIPCHK_C(h->partial.alloc(h->nblk_red));
    HIPCHK_C(h->sum1.alloc(1));
    HIPCHK_C(h->uni64.alloc(NV_MAX));
    HIPCHK_C(h->uni32.alloc(NV_MAX));
    HIPCHK_C(hipStreamSynchronize(s));
#undef HIPCHK_C
    *out = h;
    return AFX_OK;
}
}  // namespace

int afx_rans_set_bcs(afx_rans* h, int32_t n, const char* const* names, const int32_t* types, const double* vars) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!names || !types || !vars))) return fail(AFX_EARG, "bad BC arrays");
    const Mesh& m = h->m;
    std::vector<int32_t> type_of_name(m.names.size(), -1);
    std::vector<int32_t> entry_of_name(m.names.size(), -1);
    for (int32_t i = 0; i < n; ++i) {
        if (types[i] < AFX_BC_FARFIELD || types[i] > AFX_BC_WALL)
            return fail(AFX_EUNSUP, std::string("unsupported boundary type for '") + names[i] + "'");
        for (size_t k = 0; k < m.names.size(); ++k)
            if (m.names[k] == names[i]) { type_of_name[k] = types[i]; entry_of_name[k] = i; }
    }
    std::vector<int32_t> bt(h->nB);
    std::vector<double> bv((size_t)4 * h->nB);
    std::vector<int8_t> kind(h->nF, K_INTERNAL);
    std::vector<int32_t> wo, wg, bo, bg;
    for (int32_t b = 0; b < h->nB; ++b) {
        int ni = m.boundaryNameIdx[b];
        if (type_of_name[ni] < 0) return fail(AFX_EARG, "no boundary condition for patch '" + m.names[ni] + "'");
        bt[b] = type_of_name[ni];
        double v[4] = {0.2, 0, 1, 1.};   // boundary_variables() defaults (core.h:62-65)
        if (bt[b] == AFX_BC_FARFIELD)
            for (int k = 0; k < 4; ++k) v[k] = vars[4 * entry_of_name[ni] + k];
        for (int k = 0; k < 4; ++k) bv[4 * (size_t)b + k] = v[k];
        uint32_t e = m.boundaryEdges[b];
        kind[h->f_old2new[e]] = (int8_t)bt[b];   // AFX_BC_* == K_* for the three BC functors
        int32_t own = h->c_old2new[m.edgesCells[2 * (size_t)e]], gh = h->c_old2new[m.edgesCells[2 * (size_t)e + 1]];
        bo.push_back(own); bg.push_back(gh);
        if (bt[b] == AFX_BC_WALL || bt[b] == AFX_BC_SLIP) { wo.push_back(own); wg.push_back(gh); }
    }
    hipStream_t s = h->stream;
    HIPCHK(h->fkind.upload(kind, s));
    if (h->t_fkind.p && !h->tl.fref.empty()) {
        std::vector<int8_t> tk(h->tl.fref.size());
        for (size_t i = 0; i < tk.size(); ++i) tk[i] = kind[h->f_old2new[h->tl.fref[i]]];
        HIPCHK(h->t_fkind.upload(tk, s));
        // wall/slip ghosts are never read by the fused kernels: flag them
        std::vector<char> wall_ghost(h->nT, 0);
        for (size_t i = 0; i < wg.size(); ++i) wall_ghost[wg[i]] = 1;
        std::vector<int32_t> lg = h->t_lg;
        for (auto& v : lg) if (wall_ghost[v]) v = -1 - v;
        HIPCHK(h->t_Lglob.upload(lg, s));
        std::vector<uint32_t> fck(h->tl.fref.size());
        for (size_t i = 0; i < fck.size(); ++i) {
            uint32_t c = h->tl.fcell[i];
            fck[i] = (c & 0xFFFFu) | ((c >> 16) << 12) | ((uint32_t)tk[i] << 24);
        }
        HIPCHK(h->t_fck.upload(fck, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    h->n_walls = (int32_t)wo.size();
    HIPCHK(h->w_owner.alloc(wo.size())); HIPCHK(h->w_owner.upload(wo, s));
    HIPCHK(h->w_ghost.alloc(wg.size())); HIPCHK(h->w_ghost.upload(wg, s));
    HIPCHK(h->b_owner.alloc(bo.size())); HIPCHK(h->b_owner.upload(bo, s));
    HIPCHK(h->b_ghost.alloc(bg.size())); HIPCHK(h->b_ghost.upload(bg, s));
    HIPCHK(hipStreamSynchronize(s));
    h->bface_type = bt;
    h->bface_vars = bv;
    h->bcs_set = true;
    h->wd_ready = false;
    if (h->sa()) {
        int r2 = compute_wall_distance(h);
        if (r2) return r2;
    }
    // freestream state used by the uniform residual
    double v[4], qf[4];
    boundary_variables(h, v);
    conservative(h->gas, v, qf);
    float qf32[4] = {(float)qf[0], (float)qf[1], (float)qf[2], (float)qf[3]};
    HIPCHK(h->uni64.upload(qf, 4, s));
    HIPCHK(h->uni32.upload(qf32, 4, s));
    HIPCHK(hipStreamSynchronize(s));
    return AFX_OK;
}

int afx_rans_set_options(afx_rans* h, const afx_opts* o) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!o) return fail(AFX_EARG, "null options");
    if (o->gradient != 0 && o->gradient != 1) return fail(AFX_EARG, "gradient must be 0 (least-squares) or 1 (green-gauss)");
    if (o->limiter != 0 && o->limiter != 1) return fail(AFX_EARG, "limiter must be 0 (default) or 1 (Michalak)");
    h->opts.second_order = o->second_order;
    h->opts.gradient = o->gradient;
    h->opts.limiter_k = o->limiter_k;
    h->opts.limiter = o->limiter;
    return AFX_OK;
}

int afx_rans_init(afx_rans* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    double v[4], qi[4];
    boundary_variables(h, v);
    conservative(h->gas, v, qi);
    std::vector<double> tmp((size_t)NV * h->nT);
    HIPCHK(hipMemcpyAsync(tmp.data(), h->q.p, tmp.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < NV; ++k)
        for (int32_t i = 0; i < h->nC; ++i) tmp[(size_t)k * h->nT + i] = qi[k];
    HIPCHK(hipMemcpyAsync(h->q.p, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, h->stream));
    if (h->sa()) {   // freestream rho nu~ in the real cells
        std::vector<double> nu(h->nC, h->nu_ratio * h->gas.mu_L);
        HIPCHK(hipMemcpyAsync(h->q.p + (size_t)4 * h->nT, nu.data(), nu.size() * 8, hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->q32_valid = false;
    return AFX_OK;
}

int afx_rans_refill_bcs(afx_rans* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    // ghost b is device cell nC + b (ghosts keep reference order)
    std::vector<double> g((size_t)NV * h->nB);
    for (int32_t b = 0; b < h->nB; ++b) {
        double qc[4];
        conservative(h->gas, &h->bface_vars[4 * (size_t)b], qc);
        for (int k = 0; k < NV; ++k) g[(size_t)k * h->nB + b] = qc[k];
    }
    for (int k = 0; k < NV; ++k)
        HIPCHK(hipMemcpyAsync(h->q.p + (size_t)k * h->nT + h->nC, g.data() + (size_t)k * h->nB, (size_t)h->nB * 8,
                              hipMemcpyHostToDevice, h->stream));
    std::vector<double> gn(h->sa() ? h->nB : 0, h->nu_ratio * h->gas.mu_L);
    if (h->sa())
        HIPCHK(hipMemcpyAsync(h->q.p + (size_t)4 * h->nT + h->nC, gn.data(), gn.size() * 8, hipMemcpyHostToDevice,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->q32_valid = false;
    return AFX_OK;
}

int afx_rans_bcs_from_internal(afx_rans* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    HIPCHK(launch_set_walls<double>(h->stream, h->b_owner.p, h->b_ghost.p, h->nB, h->q.p, h->nT, h->nv));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->q32_valid = false;
    return AFX_OK;
}

int afx_rans_set_walls_from_internal(afx_rans* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    HIPCHK(launch_set_walls<double>(h->stream, h->w_owner.p, h->w_ghost.p, h->n_walls, h->q.p, h->nT, h->nv));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->q32_valid = false;
    return AFX_OK;
}

int afx_rans_set_state(afx_rans* h, const double* q) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!q) return fail(AFX_EARG, "null state");
    h->q32_valid = false;
    return upload_aos(h, q, h->q.p);
}

int afx_rans_get_state(afx_rans* h, double* q) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!q) return fail(AFX_EARG, "null state");
    return download_aos(h, h->q.p, q);
}

int afx_rans_set_cfl(afx_rans* h, double cfl) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    h->cfl = cfl;
    return AFX_OK;
}

int afx_rans_calc_dt(afx_rans* h, double* dt_out) {
    int rc = check_handle(h);
    if (rc) return rc;
    HIPCHK(launch_dt<double>(h->stream, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->cfl, h->dt.p,
                             h->sa() ? 1 : 0));
    if (dt_out) return download_cells(h, h->dt.p, dt_out);
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

int afx_rans_gradients(afx_rans* h, int32_t precision, double* gx, double* gy, double* limiters) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (precision == AFX_FP64) {
        rc = run_gradients64(h, h->q.p, h->q.p);
        if (rc) return rc;
        if (gx && (rc = download_aos(h, h->gx.p, gx))) return rc;
        if (gy && (rc = download_aos(h, h->gy.p, gy))) return rc;
        if (limiters && (rc = download_aos(h, h->lim.p, limiters))) return rc;
    } else if (precision == AFX_FP32) {
        if ((rc = ensure_q32(h))) return rc;
        if ((rc = run_gradients32(h))) return rc;
        if (gx && (rc = download_aos32(h, h->gx32.p, gx))) return rc;
        if (gy && (rc = download_aos32(h, h->gy32.p, gy))) return rc;
        if (limiters && (rc = download_aos32(h, h->lim32.p, limiters))) return rc;
    } else {
        return fail(AFX_EARG, "precision must be AFX_FP64 or AFX_FP32");
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

int afx_rans_residual(afx_rans* h, int32_t precision, double* R, double* norm2) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (precision != AFX_FP64 && precision != AFX_FP32) return fail(AFX_EARG, "precision must be AFX_FP64 or AFX_FP32");
    if ((rc = enqueue_residual(h, precision))) return rc;
    if (precision == AFX_FP64) {
        if (R && (rc = download_aos(h, h->R.p, R))) return rc;
        if (norm2 && (rc = norm_of(h, h->R.p, h->nT, h->nv, norm2))) return rc;
    } else if (h->r32_aos) {
        if (R) {
            std::vector<float> tmp((size_t)NV * h->nT);
            HIPCHK(hipMemcpyAsync(tmp.data(), h->R32v.p, tmp.size() * 4, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (int32_t i = 0; i < h->nT; ++i)
                for (int k = 0; k < NV; ++k) R[(size_t)NV * h->c_new2old[i] + k] = (double)tmp[(size_t)NV * i + k];
        }
        if (norm2) {
            if ((rc = norm_of32(h, reinterpret_cast<const float*>(h->R32v.p), (int64_t)NV * h->nT, 1, norm2))) return rc;
            if (h->sa()) {
                double n5 = 0;
                if ((rc = norm_of32(h, h->R32n.p, h->nT, 1, &n5))) return rc;
                *norm2 = std::sqrt(*norm2 * *norm2 + n5 * n5);
            }
        }
    } else {
        if (R && (rc = download_aos32(h, h->R32.p, R))) return rc;
        if (norm2 && (rc = norm_of32(h, h->R32.p, h->nT, NV, norm2))) return rc;
    }
    return AFX_OK;
}

// ------------------------------------------ allocation-free residual norm --
// afx_rans_residual_norm_dev keeps the partial sums of its norm reduction in
// this file-scope device array instead of the handle's buffers.  Every GPU
// holds its own copy for the life of the process.  Layout per call: the
// partial sums of up to two reductions (kNormParts x kNormBlocks), then their
// kNormParts totals.  Every slot a call reads is written earlier in the same
// call (one partial per block of k_sumsq, one total per k_final_sum), so
// nothing needs clearing; overlapping calls on one GPU share the array and
// race.
constexpr int kNormParts = 2;       // FP32 Spalart-Allmaras: float4 rows + the rho nu~ row
constexpr int kNormBlocks = 1024;   // blocks of one norm reduction (the handles' nblk_red)
constexpr int kScratchDoubles = kNormParts * kNormBlocks + kNormParts;
static __device__ __align__(256) double scratch[kScratchDoubles];   // 16 KiB + 16 B per GPU

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxScratchDevices] = {};

// device address of `scratch` on the current device (looked up once per
// device; the symbol's host-side name is not a device address)
int scratch_of_current_device(double** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return fail(AFX_EUNSUP, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(AFX_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return AFX_OK;
}

// out = sqrt(sum) for one part; with two parts sqrt(n0^2 + n1^2) of
// n_k = sqrt(sum_k), the host arithmetic of afx_rans_residual
__global__ void k_norm_from_sums(const double* __restrict__ sums, int parts, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double n0 = sqrt(sums[0]);
        if (parts == 1) {
            out[0] = n0;
        } else {
            const double n1 = sqrt(sums[1]);
            out[0] = sqrt(n0 * n0 + n1 * n1);
        }
    }
}
}  // namespace

int afx_rans_residual_norm_dev(afx_rans* h, int32_t precision, double* norm2_dev) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (precision != AFX_FP64 && precision != AFX_FP32) return fail(AFX_EARG, "precision must be AFX_FP64 or AFX_FP32");
    if (!norm2_dev) return fail(AFX_EARG, "null device output");
    if (h->nblk_red > kNormBlocks)
        return fail(AFX_EUNSUP, "the norm reduction needs " + std::to_string(h->nblk_red) +
                                    " partial sums; the allocation-free form holds " + std::to_string(kNormBlocks));
    double* sc = nullptr;
    if ((rc = scratch_of_current_device(&sc))) return rc;
    double* sums = sc + kNormParts * kNormBlocks;
    const int nb = h->nblk_red;
    hipStream_t s = h->stream;
    if ((rc = enqueue_residual(h, precision))) return rc;
    int parts = 1;
    if (precision == AFX_FP64) {
        HIPCHK(launch_sumsq(s, h->R.p, h->nT, h->nT, h->nv, sc, nb));
    } else if (h->r32_aos) {
        HIPCHK(launch_sumsq_f32(s, reinterpret_cast<const float*>(h->R32v.p), (int64_t)NV * h->nT,
                                (int64_t)NV * h->nT, 1, sc, nb));
        if (h->sa()) {
            HIPCHK(launch_sumsq_f32(s, h->R32n.p, h->nT, h->nT, 1, sc + kNormBlocks, nb));
            parts = 2;
        }
    } else {
        HIPCHK(launch_sumsq_f32(s, h->R32.p, h->nT, h->nT, NV, sc, nb));
    }
    for (int k = 0; k < parts; ++k) HIPCHK(launch_final_sum(s, sc + k * kNormBlocks, nb, sums + k));
    hipLaunchKernelGGL(k_norm_from_sums, dim3(1), dim3(64), 0, s, sums, parts, norm2_dev);
    HIPCHK(hipGetLastError());
    return AFX_OK;
}

int afx_rans_explicit_residual(afx_rans* h, const double* qk, double* qW) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (h->sa()) return fail(AFX_EUNSUP, "explicit stages are not available with the Spalart-Allmaras model");
    if (!qk) return fail(AFX_EARG, "null stage state");
    if ((rc = upload_aos(h, qk, h->qk.p))) return rc;
    if (h->needs_gradients() && (rc = run_gradients64(h, h->q.p, h->qk.p))) return rc;
    HIPCHK(launch_face_flux<double>(h->stream, h->cells64(), h->faces64(), h->qk.p, h->q.p, h->gx.p, h->gy.p, h->lim.p,
                                    h->gas64(), h->opts.viscosity, h->opts.second_order, nullptr, h->flux.p));
    HIPCHK(launch_gather<double>(h->stream, h->cells64(), h->flux.p, h->nF, 1, h->qW.p));
    if (qW) return download_aos(h, h->qW.p, qW);
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

}  // extern "C"

namespace {
// explicitSolver::solve (solver.h:802-828) for four variables, from qin to
// qout (they may alias): dt from qin, stage states qk, gradients from qin
// (A3), limiters and fluxes from qk, qk = qin + qW dt alpha_s relax.  With
// keep non-null the stage residuals go to keep[0..2] and the stage states
// qk_1, qk_2 to keep[3], keep[4] (for the reverse sweep).
int explicit_step64(afx_rans* h, const double* qin, double* qout, double relax, double* const* keep) {
    static const double alpha[3] = {0.25, 0.5, 1.};
    hipStream_t s = h->stream;
    const int64_t nq = (int64_t)NV * h->nT;
    int rc;
    HIPCHK(launch_dt<double>(s, h->cells64(), h->faces64(), qin, h->gas64(), h->cfl, h->dt.p));
    HIPCHK(launch_copy(s, qin, h->qk.p, nq));
    double* qs = h->qk.p;
    for (int st = 0; st < 3; ++st) {
        if (h->needs_gradients() && (rc = run_gradients64(h, qin, qs))) return rc;
        double* W = keep ? keep[st] : h->qW.p;
        HIPCHK(launch_face_flux<double>(s, h->cells64(), h->faces64(), qs, qin, h->gx.p, h->gy.p, h->lim.p,
                                        h->gas64(), h->opts.viscosity, h->opts.second_order, nullptr, h->flux.p));
        HIPCHK(launch_gather<double>(s, h->cells64(), h->flux.p, h->nF, 1, W));
        double* next = qs;
        if (keep) {
            next = st < 2 ? keep[3 + st] : h->qk.p;
            if (st < 2) HIPCHK(launch_copy(s, qs, next, nq));   // ghost rows of the stage state
        }
        HIPCHK(launch_rk_update(s, qin, W, h->dt.p, alpha[st], relax, h->nC, h->nT, NV, next));
        qs = next;
    }
    if (qs != qout) HIPCHK(launch_copy(s, qs, qout, nq));
    return AFX_OK;
}
}  // namespace

extern "C" {

int afx_rans_explicit_step(afx_rans* h, double relaxation, double* err) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (h->sa()) {
        // explicitSolver::solve (solver.h:802-828) with the five-variable
        // fused residual: gradients and limiters from the stage state
        if (!h->tile64_ok) return fail(AFX_EUNSUP, "the Spalart-Allmaras residual needs the fused tile kernels");
        static const double alpha5[3] = {0.25, 0.5, 1.};
        hipStream_t s = h->stream;
        const int64_t nq = (int64_t)h->nv * h->nT;
        HIPCHK(launch_dt<double>(s, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->cfl, h->dt.p, 1));
        HIPCHK(launch_copy(s, h->q.p, h->qk.p, nq));
        for (double ai : alpha5) {
            HIPCHK(launch_residual_tile<double>(s, h->tiles<double>(h->t_geo64.p), h->qk.p, h->nT, h->nC, h->area.p,
                                                h->lsq.p, h->xc.p, h->yc.p, h->sa_wd.p, h->gas64(), h->opts.gradient,
                                                h->so(), h->opts.limiter_k, h->opts.viscosity, h->R.p,
                                                h->qk.p));
            HIPCHK(launch_rk_update_area(s, h->q.p, h->R.p, h->area.p, h->dt.p, ai * relaxation, h->nC, h->nT, h->nv,
                                         h->qk.p, h->qW.p));
        }
        HIPCHK(launch_copy(s, h->qk.p, h->q.p, nq));
        h->q32_valid = false;
        h->last_precision = AFX_FP64;
        if (err) return norm_of(h, h->qW.p, h->nT, h->nv, err);
        HIPCHK(hipStreamSynchronize(s));
        return AFX_OK;
    }
    if ((rc = explicit_step64(h, h->q.p, h->q.p, relaxation, nullptr))) return rc;
    h->q32_valid = false;
    if (err) return norm_of(h, h->qW.p, h->nT, NV, err);
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

int afx_rans_set_linear_solver(afx_rans* h, int32_t restart, double tol, int32_t maxit) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (restart < 1 || restart > 32 || !(tol > 0) || maxit < 1)
        return fail(AFX_EARG, "bad linear solver settings (restart 1..32, tol > 0, maxit >= 1)");
    if (h->Jf.p && restart != h->gm_restart) return fail(AFX_ESTATE, "restart length is fixed after the first fill");
    h->gm_restart = restart;
    h->gm_tol = tol;
    h->gm_maxit = maxit;
    return AFX_OK;
}

int afx_rans_set_preconditioner(afx_rans* h, int32_t kind) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (kind != 0 && kind != 1) return fail(AFX_EARG, "preconditioner: 0 block Jacobi, 1 multicolour block DILU");
    h->precond = kind;
    h->lhs_ready = false;
    return AFX_OK;
}

int afx_rans_implicit_fill(afx_rans* h) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (h->sa() && (rc = sa_implicit_check(h))) return rc;
    if ((rc = ensure_implicit(h))) return rc;
    hipStream_t s = h->stream;
    HIPCHK(launch_dt<double>(s, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->cfl, h->dt.p, h->sa() ? 1 : 0));
    if (h->mf) {
        // fillRhoLHS of the matrix-free build assembles nothing: the operator
        // reads q and dt when applied.  (Its gradient evaluation feeds no
        // output of the two-argument fluxes, so it is skipped.)
        HIPCHK(hipStreamSynchronize(s));
        h->lhs_ready = true;
        h->lhs_bad = 0;
        h->q32_valid = false;
        return AFX_OK;
    }
    SaSource src{nullptr, nullptr, nullptr, nullptr, h->gas64()};
    if (h->sa()) {
        // wall / slip ghosts copy their owners (all five planes), then the
        // unlimited Green-Gauss gradients of the five variables
        HIPCHK(launch_set_walls<double>(s, h->w_owner.p, h->w_ghost.p, h->n_walls, h->q.p, h->nT, h->nv));
        HIPCHK(launch_sa_gradients(s, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->gx.p, h->gy.p));
        HIPCHK(launch_face_jacobian_sa(s, h->cells64(), h->faces64(), h->q.p, h->gx.p, h->gy.p, h->gas64(), h->Jf.p));
        src = SaSource{h->q.p, h->gx.p, h->gy.p, h->sa_wd.p, h->gas64()};
    } else {
        if (h->opts.viscosity != 0 && (rc = run_gradients64(h, h->q.p, h->q.p))) return rc;
        HIPCHK(launch_face_jacobian(s, h->cells64(), h->faces64(), h->q.p, h->gx.p, h->gy.p, h->gas64(),
                                    h->opts.viscosity, h->Jf.p));
    }
    HIPCHK(hipMemsetAsync(h->kbad.p, 0, sizeof(int), s));
    HIPCHK(launch_assemble_diag(s, h->nv, h->cells64(), h->faces64(), h->dt.p, h->Jf.p, src, h->Dblk.p, h->Dinv.p,
                                h->ellE.p, h->ellD.p, h->ellN.p, h->kbad.p));
    if (h->precond == 1)
        HIPCHK(launch_dilu_setup(s, h->nv, h->cells64(), h->faces64(), h->colour.p, h->col_cells.p, h->col_pos.p,
                                 h->col_off.data(), (int)h->col_off.size() - 1, h->Dblk.p, h->Jf.p, h->Dti.p,
                                 h->dilE.p, h->dilN.p, h->dilD.p, h->kbad.p));
    HIPCHK(hipMemcpyAsync(&h->lhs_bad, h->kbad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    h->lhs_ready = true;
    h->q32_valid = false;
    return AFX_OK;
}

int afx_rans_implicit_compute(afx_rans* h, int32_t* status) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->lhs_ready) return fail(AFX_ESTATE, "call afx_rans_implicit_fill first");
    if (status) *status = h->lhs_bad ? -1 : 0;
    return AFX_OK;
}

int afx_rans_implicit_solve(afx_rans* h, double relaxation, double tol, int32_t rhs_iterations, int32_t precision,
                            double* err_out) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->lhs_ready) return fail(AFX_ESTATE, "call afx_rans_implicit_fill first");
    if (precision != AFX_FP64) return fail(AFX_EUNSUP, "the implicit step uses the FP64 residual");
    hipStream_t s = h->stream;
    const int64_t n = (int64_t)h->nv * h->nT;
    h->lin_iters = 0;
    auto residual_norm = [&](double* e) -> int {
        // fillRhoRHS starts with calc_dt (solver.h:1081); only the
        // matrix-free operator reads dt after the fill
        if (h->mf) HIPCHK(launch_dt<double>(s, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->cfl, h->dt.p));
        int r2 = enqueue_residual(h, AFX_FP64);
        if (r2) return r2;
        return norm_of(h, h->R.p, h->nT, h->nv, e);
    };
    auto step = [&](bool* ok) -> int {
        if (h->mf)
            HIPCHK(launch_mf_base(s, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->opts.viscosity, h->mfFb.p));
        int r2 = gmres(h, h->R.p, h->qW.p, ok);
        if (r2) return r2;
        if (*ok) {
            HIPCHK(launch_axpy(s, relaxation, h->qW.p, h->q.p, n));
            h->q32_valid = false;
        }
        return AFX_OK;
    };
    // Deviation from solver.h:1170-1213: a non-finite residual norm is the
    // failure value -1 (the reference would return NaN, which its loop reads
    // as converged); the state is left as it was when the norm was taken.
    double err = 0;
    bool ok = false;
    auto failed = [&](const char*) {
        if (err_out) *err_out = -1;
        return AFX_OK;
    };
    if ((rc = residual_norm(&err))) return rc;
    if (!std::isfinite(err)) return failed("non-finite residual norm");
    if (err < tol) { if (err_out) *err_out = err; return AFX_OK; }
    const double err_ini = err;
    if ((rc = step(&ok))) return rc;
    if (!ok) return failed("GMRES did not converge");
    for (int i = 0; i < rhs_iterations; ++i) {
        if ((rc = residual_norm(&err))) return rc;
        if (!std::isfinite(err)) return failed("non-finite residual norm");
        if (err < tol || err > 10 * err_ini) { if (err_out) *err_out = err; return AFX_OK; }
        if ((rc = step(&ok))) return rc;
        if (!ok) return failed("GMRES did not converge");
    }
    if ((rc = residual_norm(&err))) return rc;
    if (!std::isfinite(err)) return failed("non-finite residual norm");
    if (err_out) *err_out = err;
    return AFX_OK;
}

int afx_rans_lhs_apply_nu(afx_rans* h, const double* x, const double* x_nu, double* y, double* y_nu) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!x || !y) return fail(AFX_EARG, "null vector");
    if (!h->lhs_ready) return fail(AFX_ESTATE, "call afx_rans_implicit_fill first");
    if (h->mf) return fail(AFX_ESTATE, "no assembled matrix in matrix-free mode (afx_rans_matrix_free_apply)");
    if ((rc = upload_operand(h, x, x_nu, y_nu))) return rc;
    HIPCHK(launch_lhs_apply(h->stream, h->nv, h->nC, h->nT, h->ellD.p, h->ellE.p, h->ellN.p, h->ktmp.p, h->kw.p));
    return download_operand(h, y, y_nu);
}

int afx_rans_lhs_apply(afx_rans* h, const double* x, double* y) {
    return afx_rans_lhs_apply_nu(h, x, nullptr, y, nullptr);
}

int afx_rans_precond_apply_nu(afx_rans* h, const double* x, const double* x_nu, double* y, double* y_nu) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!x || !y) return fail(AFX_EARG, "null vector");
    if (!h->lhs_ready) return fail(AFX_ESTATE, "call afx_rans_implicit_fill first");
    if (h->mf) return fail(AFX_ESTATE, "the matrix-free mode has no preconditioner");
    if ((rc = upload_operand(h, x, x_nu, y_nu))) return rc;
    if (h->precond == 1)
        HIPCHK(launch_dilu_apply(h->stream, h->nv, h->nC, h->nT, h->col_cells.p, h->col_pos.p, h->col_off.data(),
                                 (int)h->col_off.size() - 1, h->dilE.p, h->dilN.p, h->dilD.p, h->ktmp.p, h->kw.p,
                                 h->kdil.p));
    else
        HIPCHK(launch_block_jacobi(h->stream, h->nv, h->nC, h->nT, h->Dinv.p, h->ktmp.p, h->kw.p));
    return download_operand(h, y, y_nu);
}

int afx_rans_precond_apply(afx_rans* h, const double* x, double* y) {
    return afx_rans_precond_apply_nu(h, x, nullptr, y, nullptr);
}

int afx_rans_cell_colours(afx_rans* h, int32_t* colours, int32_t* ncolours) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!h->colour.p) return fail(AFX_ESTATE, "no colouring before the first afx_rans_implicit_fill");
    std::vector<uint8_t> col(h->nC);
    HIPCHK(hipMemcpyAsync(col.data(), h->colour.p, col.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (colours)
        for (int32_t i = 0; i < h->nC; ++i) colours[h->c_new2old[i]] = col[i];
    if (ncolours) *ncolours = (int32_t)h->col_off.size() - 1;
    return AFX_OK;
}

int afx_rans_set_matrix_free(afx_rans* h, int32_t enable) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (enable != 0 && enable != 1) return fail(AFX_EARG, "enable must be 0 or 1");
    if (h->sa() && enable)
        return fail(AFX_EUNSUP, "the matrix-free operator covers the four-variable residual");
    h->mf = enable != 0;
    // the GMRES settings each build of implicitSolver sets (solver.h:906-914)
    h->gm_tol = h->mf ? 1e-3 : 1e-2;
    h->gm_maxit = h->mf ? 1000 : 500;
    h->lhs_ready = false;
    return AFX_OK;
}

int afx_rans_matrix_free_apply(afx_rans* h, const double* x, double* y) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!x || !y) return fail(AFX_EARG, "null vector");
    if (!h->mf || !h->lhs_ready) return fail(AFX_ESTATE, "call afx_rans_set_matrix_free(h, 1) and afx_rans_implicit_fill first");
    if ((rc = upload_aos(h, x, h->ktmp.p))) return rc;
    HIPCHK(launch_mf_base(h->stream, h->cells64(), h->faces64(), h->q.p, h->gas64(), h->opts.viscosity, h->mfFb.p));
    HIPCHK(launch_mf_apply(h->stream, h->cells64(), h->faces64(), h->q.p, h->dt.p, h->mfFb.p, h->gas64(),
                           h->opts.viscosity, h->ktmp.p, h->kw.p));
    return download_aos(h, h->kw.p, y);
}

int afx_rans_linear_iterations(afx_rans* h, int64_t* iters) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    *iters = h->lin_iters;
    return AFX_OK;
}

int afx_rans_uniform_residual(afx_rans* h, double* err) {
    int rc = check_handle(h);
    if (rc) return rc;
    HIPCHK(launch_face_flux<double>(h->stream, h->cells64(), h->faces64(), h->q.p, h->q.p, h->gx.p, h->gy.p, h->lim.p,
                                    h->gas64(), std::min(h->opts.viscosity, 1), h->opts.second_order, h->uni64.p, h->flux.p));
    HIPCHK(launch_gather<double>(h->stream, h->cells64(), h->flux.p, h->nF, 2, h->qW.p));
    double e = 0;
    if ((rc = norm_of(h, h->qW.p, h->nT, NV, &e))) return rc;
    if (err) *err = e;
    return AFX_OK;
}

// boundary segments of a patch (reference order) and their owners' states
// (NV per face), gathered on the device
static int gather_patch(afx_rans* h, const char* patch, std::vector<int32_t>& bsel, std::vector<double>& st) {
    if (!patch) return fail(AFX_EARG, "null patch name");
    const Mesh& m = h->m;
    std::vector<int32_t> cells;
    bsel.clear();
    for (int32_t b = 0; b < h->nB; ++b)
        if (m.names[m.boundaryNameIdx[b]] == patch) {
            bsel.push_back(b);
            cells.push_back(h->c_old2new[m.edgesCells[2 * (size_t)m.boundaryEdges[b]]]);
        }
    if (bsel.empty()) return fail(AFX_EARG, std::string("no boundary faces on patch '") + patch + "'");
    if (h->patch_cells.n < cells.size()) {
        if (h->patch_cells.p) { (void)hipFree(h->patch_cells.p); h->patch_cells.p = nullptr; }
        if (h->states.p) { (void)hipFree(h->states.p); h->states.p = nullptr; }
        HIPCHK(h->patch_cells.alloc(cells.size()));
        HIPCHK(h->states.alloc(cells.size() * NV));
    }
    HIPCHK(h->patch_cells.upload(cells, h->stream));
    HIPCHK(launch_gather_states(h->stream, h->q.p, h->nT, NV, h->patch_cells.p, (int32_t)cells.size(), h->states.p));
    st.assign(cells.size() * NV, 0.0);
    HIPCHK(hipMemcpyAsync(st.data(), h->states.p, st.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

int afx_rans_wall_forces(afx_rans* h, const char* patch, double* cl, double* cd, double* cm) {
    int rc = check_handle(h);
    if (rc) return rc;
    const Mesh& m = h->m;
    std::vector<int32_t> bsel;
    std::vector<double> st;
    if ((rc = gather_patch(h, patch, bsel, st))) return rc;
    // post.h:301-387, in the reference's boundary order
    double v[4];
    boundary_variables(h, v);
    const double p_inf = v[3], mach_inf = v[0], aoa = v[1];
    double xmin = 0., xmax = 0., y_moment = 0.;
    uint32_t n_added = 0;
    for (size_t k = 0; k < bsel.size(); ++k) {
        uint32_t e = m.boundaryEdges[bsel[k]];
        if (k == 0) { xmin = xmax = m.edgesCentersX[e]; y_moment = m.edgesCentersY[e]; }
        else { xmin = std::min(xmin, m.edgesCentersX[e]); xmax = std::max(xmax, m.edgesCentersX[e]); y_moment += m.edgesCentersY[e]; }
        n_added += 1;
    }
    y_moment /= (double)n_added;
    const double x_moment = (xmax - xmin) * 0.25 + xmin;
    const double gm = h->gas.gamma;
    double fx = 0., fy = 0., mm = 0.;
    for (size_t k = 0; k < bsel.size(); ++k) {
        uint32_t e = m.boundaryEdges[bsel[k]];
        const double* qc = &st[k * NV];
        const double p = (gm - 1) * (qc[3] - 0.5 / qc[0] * (qc[1] * qc[1] + qc[2] * qc[2]));
        const double cp = 2. / (gm * mach_inf * mach_inf) * (p / p_inf - 1.);
        const double ecx = m.edgesCentersX[e], ecy = m.edgesCentersY[e];
        const double fxi = cp * m.edgesNormalsX[e] * m.edgesLengths[e] / (xmax - xmin);
        const double fyi = cp * m.edgesNormalsY[e] * m.edgesLengths[e] / (xmax - xmin);
        const double mi = (ecx - x_moment) / (xmax - xmin) * fyi - (ecy - y_moment) / (xmax - xmin) * fxi;
        fx += fxi; fy += fyi; mm -= mi;
    }
    if (cd) *cd = fx * std::cos(aoa) + fy * std::sin(aoa);
    if (cl) *cl = -fx * std::sin(aoa) + fy * std::cos(aoa);
    if (cm) *cm = mm;
    return AFX_OK;
}

int afx_rans_cp_profile(afx_rans* h, const char* patch, int64_t* n, double* x, double* y, double* cp) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!n) return fail(AFX_EARG, "n is null");
    std::vector<int32_t> bsel;
    std::vector<double> st;
    if ((rc = gather_patch(h, patch, bsel, st))) return rc;
    const int64_t cap = *n;
    *n = (int64_t)bsel.size();
    if (!x && !y && !cp) return AFX_OK;
    if (!x || !y || !cp) return fail(AFX_EARG, "x, y and cp are all null (query) or all set");
    if (cap < *n) return fail(AFX_EARG, "arrays hold " + std::to_string(cap) + " faces, the patch has " +
                                             std::to_string(*n));
    // CpProfile::calc_cp (post.h:246-290): the owner's pressure, first order
    const Mesh& m = h->m;
    double v[4];
    boundary_variables(h, v);
    const double p_inf = v[3], mach_inf = v[0], gm = h->gas.gamma;
    for (size_t k = 0; k < bsel.size(); ++k) {
        const uint32_t e = m.boundaryEdges[bsel[k]];
        const double* q = &st[k * NV];
        const double p = (gm - 1) * (q[3] - 0.5 / q[0] * (q[1] * q[1] + q[2] * q[2]));
        x[k] = m.edgesCentersX[e];
        y[k] = m.edgesCentersY[e];
        cp[k] = 2. / (gm * mach_inf * mach_inf) * (p / p_inf - 1.);
    }
    return AFX_OK;
}

int afx_rans_wall_dist(afx_rans* h, double* wall_dist) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->wd_ready && (rc = compute_wall_distance(h))) return rc;
    if (wall_dist) return download_cells(h, h->wall_d.p, wall_dist);
    return AFX_OK;
}

int afx_rans_sa_wall_dist(afx_rans* h, double* dist) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->sa()) return fail(AFX_EUNSUP, "the model's wall distance belongs to the Spalart-Allmaras model");
    if (!h->wd_ready && (rc = compute_wall_distance(h))) return rc;
    if (dist) return download_cells(h, h->sa_wd.p, dist);
    return AFX_OK;
}

int afx_rans_save_vtu(afx_rans* h, const char* path) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!path) return fail(AFX_EARG, "path is null");
    if (!h->wd_ready && (rc = compute_wall_distance(h))) return rc;
    std::vector<double> q((size_t)4 * h->nT), d(h->nT);
    if ((rc = afx_rans_get_state(h, q.data()))) return rc;
    if ((rc = download_cells(h, h->wall_d.p, d.data()))) return rc;
    const std::string err = write_vtu(path, h->m, q.data(), d.data(), h->gas.gamma, h->gas.R);
    if (!err.empty()) return fail(AFX_EIO, err);
    return AFX_OK;
}

int afx_rans_set_turbulence(afx_rans* h, double nu_ratio) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!h->sa()) return fail(AFX_ESTATE, "the solver was not created with the Spalart-Allmaras model");
    if (!(nu_ratio >= 0)) return fail(AFX_EARG, "nu_ratio must be >= 0");
    h->nu_ratio = nu_ratio;
    return AFX_OK;
}

int afx_rans_set_nu(afx_rans* h, const double* nu) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!h->sa()) return fail(AFX_ESTATE, "the solver was not created with the Spalart-Allmaras model");
    if (!nu) return fail(AFX_EARG, "null rho nu~ vector");
    std::vector<double> tmp(h->nT);
    for (int32_t i = 0; i < h->nT; ++i) tmp[i] = nu[h->c_new2old[i]];
    HIPCHK(hipMemcpyAsync(h->q.p + (size_t)4 * h->nT, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->q32_valid = false;
    return AFX_OK;
}

int afx_rans_get_nu(afx_rans* h, double* nu) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!h->sa()) return fail(AFX_ESTATE, "the solver was not created with the Spalart-Allmaras model");
    if (!nu) return fail(AFX_EARG, "null rho nu~ vector");
    return download_cells(h, h->q.p + (size_t)4 * h->nT, nu);
}

int afx_rans_nu_residual(afx_rans* h, double* Rnu) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    if (!h->sa()) return fail(AFX_ESTATE, "the solver was not created with the Spalart-Allmaras model");
    if (!Rnu) return fail(AFX_EARG, "null residual vector");
    if (h->last_precision == AFX_FP64) return download_cells(h, h->R.p + (size_t)4 * h->nT, Rnu);
    if (h->last_precision != AFX_FP32) return fail(AFX_ESTATE, "no residual has been evaluated");
    std::vector<float> tmp(h->nT);
    HIPCHK(hipMemcpyAsync(tmp.data(), h->R32n.p, tmp.size() * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int32_t i = 0; i < h->nT; ++i) Rnu[h->c_new2old[i]] = (double)tmp[i];
    return AFX_OK;
}

int afx_rans_counts(afx_rans* h, int64_t c[5]) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    c[0] = h->m.n_nodes(); c[1] = h->nC; c[2] = h->nT; c[3] = h->nF; c[4] = h->nB;
    return AFX_OK;
}

int afx_rans_stream(afx_rans* h, void** stream) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    *stream = (void*)h->stream;
    return AFX_OK;
}

int afx_rans_launch_stage(afx_rans* h, int32_t precision, int32_t stage, int32_t n) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (precision != AFX_FP64 && precision != AFX_FP32) return fail(AFX_EARG, "bad precision");
    if (stage < 0 || stage > 3 || n < 0) return fail(AFX_EARG, "bad stage");
    if (precision == AFX_FP32 && (rc = ensure_q32(h))) return rc;
    for (int32_t it = 0; it < n; ++it) {
        if (stage == 3) { if ((rc = enqueue_residual(h, precision))) return rc; continue; }
        if (precision == AFX_FP64) {
            if (stage == 0) { if ((rc = run_gradients64(h, h->q.p, h->q.p))) return rc; }
            else if (stage == 1) {
                HIPCHK(launch_face_flux<double>(h->stream, h->cells64(), h->faces64(), h->q.p, h->q.p, h->gx.p, h->gy.p,
                                                h->lim.p, h->gas64(), h->opts.viscosity, h->opts.second_order, nullptr,
                                                h->flux.p));
            } else {
                HIPCHK(launch_gather<double>(h->stream, h->cells64(), h->flux.p, h->nF, 0, h->R.p));
            }
        } else {
            if (stage == 0) { if ((rc = run_gradients32(h))) return rc; }
            else if (stage == 1) {
                HIPCHK(launch_face_flux<float>(h->stream, h->cells32(), h->faces32(), h->q32.p, h->q32.p, h->gx32.p,
                                               h->gy32.p, h->lim32.p, h->gas32(), h->opts.viscosity,
                                               h->opts.second_order, nullptr, h->flux32.p));
            } else {
                HIPCHK(launch_gather<float>(h->stream, h->cells32(), h->flux32.p, h->nF, 0, h->R32.p));
            }
        }
    }
    return AFX_OK;
}

namespace {
// laminar_ok: the state adjoint covers the laminar residual too; the adjoints
// with respect to the mesh nodes and through pseudo-time cover the inviscid one
// state_adjoint: the state adjoint (every model); otherwise the adjoints with
// respect to the mesh nodes and through pseudo-time (inviscid only)
int adjoint_check(afx_rans* h, bool state_adjoint = false) {
    if (h->opts.viscosity != 0 && !state_adjoint)
        return fail(AFX_EUNSUP, "the adjoints with respect to the mesh nodes and through pseudo-time cover the "
                                "inviscid residual");
    if (h->opts.gradient != 1) return fail(AFX_EUNSUP, "the state adjoint covers Green-Gauss gradients");
    if (h->michalak()) return fail(AFX_EUNSUP, "the state adjoint covers the default limiter, not the Michalak build");
    if (h->sa() && !h->wd_ready) {
        int rc = compute_wall_distance(h);
        if (rc) return rc;
    }
    if (!h->a_w.p) {
        // planes: one per variable (five with the Spalart-Allmaras model)
        const size_t c4 = (size_t)h->nv * h->nC, t4 = (size_t)h->nv * h->nT, f4 = (size_t)h->nv * h->nF;
        for (DBuf<double>* b : {&h->a_w, &h->a_lam, &h->a_lgx, &h->a_lgy}) HIPCHK(b->alloc(t4));
        for (DBuf<double>* b : {&h->a_bL, &h->a_bR, &h->a_co, &h->a_cn}) HIPCHK(b->alloc(f4));
        HIPCHK(h->a_nbv.alloc(c4));
        HIPCHK(h->a_act.alloc(c4));
        HIPCHK(h->a_jmax.alloc(c4)); HIPCHK(h->a_jmin.alloc(c4)); HIPCHK(h->a_nbj.alloc(c4));
        HIPCHK(hipMemsetAsync(h->a_lam.p, 0, t4 * 8, h->stream));
        std::vector<int32_t> bf;
        bf.reserve(h->m.boundaryEdges.size());
        for (uint32_t e : h->m.boundaryEdges) bf.push_back(h->f_old2new[e]);
        std::sort(bf.begin(), bf.end());
        if (!bf.empty()) { HIPCHK(h->a_bface.alloc(bf.size())); HIPCHK(h->a_bface.upload(bf, h->stream)); }
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    if (h->opts.viscosity != 0 && !h->a_vgx.p) {
        const size_t f4 = (size_t)h->nv * h->nF;
        for (DBuf<double>* b : {&h->a_vgx, &h->a_vgy, &h->a_vdq}) HIPCHK(b->alloc(f4));
    }
    return AFX_OK;
}
AdjBufs adj_bufs(const afx_rans* h) {
    const bool lam = h->opts.viscosity != 0;
    return AdjBufs{h->gx.p, h->gy.p, h->lim.p, h->a_act.p, h->a_jmax.p, h->a_jmin.p, h->a_bL.p, h->a_bR.p,
                   h->a_lgx.p, h->a_lgy.p, h->a_nbv.p, h->a_nbj.p, h->a_co.p, h->a_cn.p,
                   h->a_bface.p, (int32_t)h->m.boundaryEdges.size(), lam ? h->a_vgx.p : nullptr,
                   lam ? h->a_vgy.p : nullptr, lam ? h->a_vdq.p : nullptr};
}
int geo_check(afx_rans* h) {
    if (h->g_fbar.p) return AFX_OK;
    const Mesh& m = h->m;
    const int32_t nC = h->nC, nF = h->nF;
    const int32_t nN = (int32_t)m.nodesX.size();
    h->nN = nN;
    std::vector<int32_t> fnode(2 * (size_t)nF), cnode(4 * (size_t)nC);
    for (int32_t nf = 0; nf < nF; ++nf) {
        const int32_t f = h->f_new2old[nf];
        fnode[nf] = (int32_t)m.edgesNodes[2 * (size_t)f];
        fnode[nF + nf] = (int32_t)m.edgesNodes[2 * (size_t)f + 1];
    }
    for (int32_t ci = 0; ci < nC; ++ci) {
        const int32_t c = h->c_new2old[ci];
        for (int s = 0; s < 4; ++s)
            cnode[(size_t)s * nC + ci] = (s == 3 && m.cellsIsTriangle[c]) ? -1 : (int32_t)m.cellsNodes[4 * (size_t)c + s];
    }
    // incidences: face ends first (face order, end a then b), then cell vertices
    std::vector<int32_t> off(nN + 1, 0);
    for (int32_t f = 0; f < nF; ++f) { ++off[fnode[f] + 1]; ++off[fnode[nF + f] + 1]; }
    for (int32_t c = 0; c < nC; ++c)
        for (int s = 0; s < 4; ++s)
            if (cnode[(size_t)s * nC + c] >= 0) ++off[cnode[(size_t)s * nC + c] + 1];
    for (int32_t v = 0; v < nN; ++v) off[v + 1] += off[v];
    std::vector<int32_t> ent(off[nN]), pos(off.begin(), off.end() - 1);
    for (int32_t f = 0; f < nF; ++f) {
        ent[pos[fnode[f]]++] = 2 * f;
        ent[pos[fnode[nF + f]]++] = 2 * f + 1;
    }
    for (int32_t c = 0; c < nC; ++c)
        for (int s = 0; s < 4; ++s) {
            const int32_t v = cnode[(size_t)s * nC + c];
            if (v >= 0) ent[pos[v]++] = -(1 + 4 * c + s);
        }
    HIPCHK(h->g_fbar.alloc(8 * (size_t)nF));
    HIPCHK(h->g_abar.alloc(nC));
    HIPCHK(h->g_fend.alloc(4 * (size_t)nF));
    HIPCHK(h->g_fcen.alloc(4 * (size_t)nF));
    HIPCHK(h->g_cvert.alloc(8 * (size_t)nC));
    HIPCHK(h->g_xbar.alloc(2 * (size_t)nN));
    HIPCHK(h->g_nodex.alloc(nN)); HIPCHK(h->g_nodex.upload(m.nodesX, h->stream));
    HIPCHK(h->g_nodey.alloc(nN)); HIPCHK(h->g_nodey.upload(m.nodesY, h->stream));
    HIPCHK(h->g_fnode.alloc(fnode.size())); HIPCHK(h->g_fnode.upload(fnode, h->stream));
    HIPCHK(h->g_cnode.alloc(cnode.size())); HIPCHK(h->g_cnode.upload(cnode, h->stream));
    HIPCHK(h->g_noff.alloc(off.size())); HIPCHK(h->g_noff.upload(off, h->stream));
    HIPCHK(h->g_nent.alloc(ent.size())); HIPCHK(h->g_nent.upload(ent, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}
int enqueue_adjoint(afx_rans* h) {
    const AdjBufs A = adj_bufs(h);
    if (h->sa())
        HIPCHK(launch_adjoint_state_sa(h->stream, h->cells64(), h->faces64(), h->q.p, h->a_w.p, h->sa_wd.p,
                                       h->gas64(), h->opts.second_order, h->opts.limiter_k, A, h->a_lam.p));
    else
        HIPCHK(launch_adjoint_state(h->stream, h->cells64(), h->faces64(), h->q.p, h->a_w.p, h->gas64(),
                                    h->opts.second_order, h->opts.limiter_k, A, h->a_lam.p));
    return AFX_OK;
}
}  // namespace

int afx_rans_adjoint_nu(afx_rans* h, const double* w, const double* w_nu, double* lambda, double* lambda_nu) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!w || !lambda) return fail(AFX_EARG, "null vector");
    if (!h->sa() && (w_nu || lambda_nu))
        return fail(AFX_EARG, "rho nu~ seed / output only with the Spalart-Allmaras model");
    if ((rc = adjoint_check(h, true))) return rc;
    if ((rc = upload_aos(h, w, h->a_w.p))) return rc;
    if (h->sa()) {
        std::vector<double> tmp(h->nT, 0.0);
        if (w_nu)
            for (int32_t i = 0; i < h->nT; ++i) tmp[i] = w_nu[h->c_new2old[i]];
        HIPCHK(hipMemcpyAsync(h->a_w.p + 4 * (size_t)h->nT, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice,
                              h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    // wall/slip ghosts as fillRhoRHS sees them (they enter no flux, but the
    // state is then the one the residual reads)
    HIPCHK(launch_set_walls<double>(h->stream, h->w_owner.p, h->w_ghost.p, h->n_walls, h->q.p, h->nT, h->nv));
    if ((rc = enqueue_adjoint(h))) return rc;
    h->q32_valid = false;
    if ((rc = download_aos(h, h->a_lam.p, lambda))) return rc;
    if (lambda_nu && (rc = download_cells(h, h->a_lam.p + 4 * (size_t)h->nT, lambda_nu))) return rc;
    return AFX_OK;
}

int afx_rans_adjoint(afx_rans* h, const double* w, double* lambda) {
    return afx_rans_adjoint_nu(h, w, nullptr, lambda, nullptr);
}

int afx_rans_adjoint_nodes(afx_rans* h, const double* w, double* lambda, double* xbar) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!w || !xbar) return fail(AFX_EARG, "null vector");
    if ((rc = adjoint_check(h))) return rc;
    if ((rc = geo_check(h))) return rc;
    if ((rc = upload_aos(h, w, h->a_w.p))) return rc;
    HIPCHK(launch_set_walls<double>(h->stream, h->w_owner.p, h->w_ghost.p, h->n_walls, h->q.p, h->nT, NV));
    const AdjBufs A = adj_bufs(h);
    GeoBufs G{h->g_fbar.p, h->g_abar.p, h->g_fnode.p, h->g_cnode.p, h->g_nodex.p, h->g_nodey.p, h->g_fend.p,
              h->g_fcen.p, h->g_cvert.p, h->g_noff.p, h->g_nent.p, h->nN, h->g_xbar.p};
    HIPCHK(launch_adjoint_state(h->stream, h->cells64(), h->faces64(), h->q.p, h->a_w.p, h->gas64(),
                                h->opts.second_order, h->opts.limiter_k, A, h->a_lam.p, &G));
    h->q32_valid = false;
    if (lambda && (rc = download_aos(h, h->a_lam.p, lambda))) return rc;
    HIPCHK(hipMemcpyAsync(xbar, h->g_xbar.p, 2 * (size_t)h->nN * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

}  // extern "C"

namespace {
// get_wall_profile's CL and CD (post.h:301-387) at the device state q and the
// gradient dCL/dq in the device layout (owner cells of the patch's faces)
int wall_cl_seed(afx_rans* h, const char* patch, const double* qdev, double* cl, double* cd, std::vector<double>& seed,
                 std::vector<std::pair<uint32_t, double>>* face_cp = nullptr) {
    const Mesh& m = h->m;
    std::vector<int32_t> bsel, cells;
    for (int32_t b = 0; b < h->nB; ++b)
        if (m.names[m.boundaryNameIdx[b]] == patch) {
            bsel.push_back(b);
            cells.push_back(h->c_old2new[m.edgesCells[2 * (size_t)m.boundaryEdges[b]]]);
        }
    if (bsel.empty()) return fail(AFX_EARG, std::string("no boundary faces on patch '") + patch + "'");
    DBuf<int32_t> dc;
    DBuf<double> st;
    HIPCHK(dc.alloc(cells.size()));
    HIPCHK(dc.upload(cells, h->stream));
    HIPCHK(st.alloc(cells.size() * NV));
    HIPCHK(launch_gather_states(h->stream, qdev, h->nT, NV, dc.p, (int32_t)cells.size(), st.p));
    std::vector<double> q(cells.size() * NV);
    HIPCHK(hipMemcpyAsync(q.data(), st.p, q.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double v[4];
    boundary_variables(h, v);
    const double p_inf = v[3], mach = v[0], aoa = v[1], gm = h->gas.gamma;
    double xmin = 0., xmax = 0.;
    for (size_t k = 0; k < bsel.size(); ++k) {
        const double x = m.edgesCentersX[m.boundaryEdges[bsel[k]]];
        if (k == 0) xmin = xmax = x;
        else { xmin = std::min(xmin, x); xmax = std::max(xmax, x); }
    }
    const double ext = xmax - xmin, sa = std::sin(aoa), ca = std::cos(aoa);
    double fx = 0., fy = 0.;
    seed.assign((size_t)NV * h->nT, 0.0);
    for (size_t k = 0; k < bsel.size(); ++k) {
        const uint32_t e = m.boundaryEdges[bsel[k]];
        const double* qc = &q[k * NV];
        const double p = (gm - 1) * (qc[3] - 0.5 / qc[0] * (qc[1] * qc[1] + qc[2] * qc[2]));
        const double cp = 2. / (gm * mach * mach) * (p / p_inf - 1.);
        if (face_cp) face_cp->push_back({e, cp});
        const double nx = m.edgesNormalsX[e], ny = m.edgesNormalsY[e], L = m.edgesLengths[e];
        fx += cp * nx * L / ext;
        fy += cp * ny * L / ext;
        const double dcl_dp = (-nx * sa + ny * ca) * L / ext * 2. / (gm * mach * mach * p_inf);
        const double dp[4] = {(gm - 1) * 0.5 * (qc[1] * qc[1] + qc[2] * qc[2]) / (qc[0] * qc[0]),
                              -(gm - 1) * qc[1] / qc[0], -(gm - 1) * qc[2] / qc[0], gm - 1};
        for (int j = 0; j < 4; ++j) seed[(size_t)j * h->nT + cells[k]] += dcl_dp * dp[j];
    }
    *cl = -fx * sa + fy * ca;
    *cd = fx * ca + fy * sa;
    return AFX_OK;
}
}  // namespace

extern "C" {

int afx_rans_dcl_dalpha(afx_rans* h, const char* patch, int32_t steps, int32_t every, double relaxation, double* cl,
                        double* dcl) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!patch || !cl || !dcl) return fail(AFX_EARG, "null argument");
    if (steps < 1 || every < 1) return fail(AFX_EARG, "steps and the checkpoint interval must be >= 1");
    if (h->sa()) return fail(AFX_EUNSUP, "dCL/dalpha covers the four-variable inviscid residual");
    if ((rc = adjoint_check(h))) return rc;
    hipStream_t s = h->stream;
    const int32_t nC = h->nC, nT = h->nT;
    const size_t nq = (size_t)NV * nT;
    const int nck = (steps + every - 1) / every;
    // farfield ghosts (their state is the alpha-dependent boundary data)
    const Mesh& m = h->m;
    std::vector<int32_t> far;
    std::vector<double> dfar;   // d(rho u, rho v) / d angle per farfield ghost
    for (int32_t b = 0; b < h->nB; ++b)
        if (h->bface_type[b] == AFX_BC_FARFIELD) {
            far.push_back(h->c_old2new[m.edgesCells[2 * (size_t)m.boundaryEdges[b] + 1]]);
            const double* v = &h->bface_vars[4 * (size_t)b];   // mach, angle, T, p (core.h:61-84)
            const double c = std::sqrt(h->gas.gamma * h->gas.R * v[2]), rho = v[3] / (h->gas.R * v[2]);
            dfar.push_back(-rho * v[0] * c * std::sin(v[1]));
            dfar.push_back(rho * v[0] * c * std::cos(v[1]));
        }
    const int32_t nfar = (int32_t)far.size();
    DBuf<double> ck, seg, keepb, g, qb, lam, lamg, wseed, dtb, eb, gh, gacc;
    DBuf<int32_t> dfarc;
    HIPCHK(ck.alloc(nck * nq));
    HIPCHK(seg.alloc((size_t)std::min(every, steps) * nq));
    HIPCHK(keepb.alloc(5 * nq));
    for (DBuf<double>* b : {&g, &qb, &lam, &lamg, &wseed, &gh}) HIPCHK(b->alloc(nq));
    HIPCHK(dtb.alloc(nT));
    HIPCHK(eb.alloc(h->nF));
    HIPCHK(gacc.alloc(4 * (size_t)std::max(nfar, 1)));
    HIPCHK(dfarc.alloc(std::max(nfar, 1)));
    if (nfar) HIPCHK(dfarc.upload(far, s));
    HIPCHK(hipMemsetAsync(wseed.p, 0, nq * 8, s));
    HIPCHK(hipMemsetAsync(gacc.p, 0, 4 * (size_t)std::max(nfar, 1) * 8, s));
    double* keep[5];
    for (int k = 0; k < 5; ++k) keep[k] = keepb.p + k * nq;

    // forward: steps explicit steps of the handle's state, checkpoints every `every`
    for (int32_t n = 0; n < steps; ++n) {
        if (n % every == 0) HIPCHK(launch_copy(s, h->q.p, ck.p + (size_t)(n / every) * nq, (int64_t)nq));
        if ((rc = explicit_step64(h, h->q.p, h->q.p, relaxation, nullptr))) return rc;
    }
    h->q32_valid = false;
    std::vector<double> seed;
    double cd = 0.0;
    if ((rc = wall_cl_seed(h, patch, h->q.p, cl, &cd, seed))) return rc;
    HIPCHK(g.upload(seed, s));

    // reverse sweep, one checkpoint segment at a time (states recomputed)
    static const double alpha[3] = {0.25, 0.5, 1.};
    const AdjBufs A = adj_bufs(h);
    for (int k = nck - 1; k >= 0; --k) {
        const int32_t start = k * every, end = std::min(steps, start + every);
        HIPCHK(launch_copy(s, ck.p + (size_t)k * nq, seg.p, (int64_t)nq));
        for (int32_t n = start; n < end - 1; ++n)
            if ((rc = explicit_step64(h, seg.p + (size_t)(n - start) * nq, seg.p + (size_t)(n - start + 1) * nq,
                                      relaxation, nullptr)))
                return rc;
        for (int32_t n = end - 1; n >= start; --n) {
            double* qn = seg.p + (size_t)(n - start) * nq;
            if ((rc = explicit_step64(h, qn, h->qk.p, relaxation, keep))) return rc;   // dt, W_s, qk_1, qk_2
            HIPCHK(hipMemsetAsync(qb.p, 0, nq * 8, s));
            HIPCHK(hipMemsetAsync(dtb.p, 0, (size_t)nT * 8, s));
            for (int st = 2; st >= 0; --st) {
                HIPCHK(launch_rev_stage(s, g.p, keep[st], h->dt.p, h->area.p, alpha[st] * relaxation, nC, nT,
                                        wseed.p, dtb.p, qb.p));
                const double* qs = st == 0 ? qn : keep[2 + st];
                HIPCHK(launch_adjoint_stage(s, h->cells64(), h->faces64(), qn, qs, wseed.p, h->gas64(),
                                            h->opts.second_order, h->opts.limiter_k, A, lam.p, lamg.p, gh.p));
                HIPCHK(launch_ghost_acc(s, gh.p, dfarc.p, nfar, nT, gacc.p));
                HIPCHK(launch_axpy4(s, lamg.p, 1.0, nC, nT, 0, qb.p));
                HIPCHK(launch_axpy4(s, lam.p, 1.0, nC, nT, st > 0 ? 1 : 0, st > 0 ? g.p : qb.p));
            }
            HIPCHK(launch_dt_vjp(s, h->cells64(), h->faces64(), qn, h->gas64(), h->dt.p, dtb.p, h->cfl, eb.p, qb.p));
            HIPCHK(launch_copy(s, qb.p, g.p, (int64_t)nq));
        }
    }
    std::vector<double> ga(4 * (size_t)std::max(nfar, 1));
    HIPCHK(hipMemcpyAsync(ga.data(), gacc.p, ga.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // CL = -fx sin a + fy cos a: explicit part -CD, plus the boundary data
    double d = -cd;
    for (int32_t j = 0; j < nfar; ++j) d += ga[(size_t)1 * nfar + j] * dfar[2 * j] + ga[(size_t)2 * nfar + j] * dfar[2 * j + 1];
    *dcl = d;
    return AFX_OK;
}

int afx_rans_dcl_dnodes(afx_rans* h, const char* patch, int32_t steps, int32_t every, double relaxation, double* cl,
                        double* xbar) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!patch || !cl || !xbar) return fail(AFX_EARG, "null argument");
    if (steps < 1 || every < 1) return fail(AFX_EARG, "steps and the checkpoint interval must be >= 1");
    if (h->sa()) return fail(AFX_EUNSUP, "dCL/dX covers the four-variable inviscid residual");
    if ((rc = adjoint_check(h))) return rc;
    if ((rc = geo_check(h))) return rc;
    hipStream_t s = h->stream;
    const int32_t nC = h->nC, nT = h->nT, nF = h->nF;
    const size_t nq = (size_t)NV * nT;
    const int nck = (steps + every - 1) / every;
    DBuf<double> ck, seg, keepb, g, qb, lam, lamg, wseed, dtb, eb, facc, aacc;
    HIPCHK(ck.alloc(nck * nq));
    HIPCHK(seg.alloc((size_t)std::min(every, steps) * nq));
    HIPCHK(keepb.alloc(5 * nq));
    for (DBuf<double>* b : {&g, &qb, &lam, &lamg, &wseed}) HIPCHK(b->alloc(nq));
    HIPCHK(dtb.alloc(nT));
    HIPCHK(eb.alloc(nF));
    HIPCHK(facc.alloc(8 * (size_t)nF));
    HIPCHK(aacc.alloc(nC));
    HIPCHK(hipMemsetAsync(wseed.p, 0, nq * 8, s));
    HIPCHK(hipMemsetAsync(aacc.p, 0, (size_t)nC * 8, s));
    double* keep[5];
    for (int k = 0; k < 5; ++k) keep[k] = keepb.p + k * nq;

    // forward: steps explicit steps of the handle's state, checkpoints every `every`
    for (int32_t n = 0; n < steps; ++n) {
        if (n % every == 0) HIPCHK(launch_copy(s, h->q.p, ck.p + (size_t)(n / every) * nq, (int64_t)nq));
        if ((rc = explicit_step64(h, h->q.p, h->q.p, relaxation, nullptr))) return rc;
    }
    h->q32_valid = false;
    std::vector<double> seed;
    std::vector<std::pair<uint32_t, double>> fcp;
    double cd = 0.0;
    if ((rc = wall_cl_seed(h, patch, h->q.p, cl, &cd, seed, &fcp))) return rc;
    HIPCHK(g.upload(seed, s));
    // CL's own metric dependence (post.h:301-387): fx, fy = sum cp (n_x, n_y) len / ext
    const Mesh& m = h->m;
    double v4[4];
    boundary_variables(h, v4);
    const double sa = std::sin(v4[1]), ca = std::cos(v4[1]);
    double xmin = 0, xmax = 0;
    for (size_t k = 0; k < fcp.size(); ++k) {
        const double x = m.edgesCentersX[fcp[k].first];
        if (k == 0 || x < xmin) xmin = x;
        if (k == 0 || x > xmax) xmax = x;
    }
    // faces at the extremes; a symmetric section has two at each (upper and
    // lower surface), where the extent's subgradient is split evenly - the
    // derivative central differences see
    std::vector<uint32_t> emin, emax;
    for (auto& fc : fcp) {
        if (m.edgesCentersX[fc.first] == xmin) emin.push_back(fc.first);
        if (m.edgesCentersX[fc.first] == xmax) emax.push_back(fc.first);
    }
    const double ext = xmax - xmin;
    {
        std::vector<double> f0(8 * (size_t)nF, 0.0);
        for (auto& [e, cp] : fcp) {
            const size_t f = (size_t)h->f_old2new[e];
            const double L = m.edgesLengths[e], nx = m.edgesNormalsX[e], ny = m.edgesNormalsY[e];
            f0[0 * (size_t)nF + f] += -sa * cp * L / ext;
            f0[1 * (size_t)nF + f] += ca * cp * L / ext;
            f0[2 * (size_t)nF + f] += (-sa * nx + ca * ny) * cp / ext;
        }
        HIPCHK(facc.upload(f0, s));
        HIPCHK(hipStreamSynchronize(s));
    }

    // reverse sweep as in afx_rans_dcl_dalpha, with each stage's metric
    // adjoints and the time step's accumulated over all stages and steps
    static const double alpha[3] = {0.25, 0.5, 1.};
    const AdjBufs A = adj_bufs(h);
    GeoBufs G{h->g_fbar.p, h->g_abar.p, h->g_fnode.p, h->g_cnode.p, h->g_nodex.p, h->g_nodey.p, h->g_fend.p,
              h->g_fcen.p, h->g_cvert.p, h->g_noff.p, h->g_nent.p, h->nN, h->g_xbar.p};
    for (int k = nck - 1; k >= 0; --k) {
        const int32_t start = k * every, end = std::min(steps, start + every);
        HIPCHK(launch_copy(s, ck.p + (size_t)k * nq, seg.p, (int64_t)nq));
        for (int32_t n = start; n < end - 1; ++n)
            if ((rc = explicit_step64(h, seg.p + (size_t)(n - start) * nq, seg.p + (size_t)(n - start + 1) * nq,
                                      relaxation, nullptr)))
                return rc;
        for (int32_t n = end - 1; n >= start; --n) {
            double* qn = seg.p + (size_t)(n - start) * nq;
            if ((rc = explicit_step64(h, qn, h->qk.p, relaxation, keep))) return rc;   // dt, W_s, qk_1, qk_2
            HIPCHK(hipMemsetAsync(qb.p, 0, nq * 8, s));
            HIPCHK(hipMemsetAsync(dtb.p, 0, (size_t)nT * 8, s));
            for (int st = 2; st >= 0; --st) {
                HIPCHK(launch_rev_stage(s, g.p, keep[st], h->dt.p, h->area.p, alpha[st] * relaxation, nC, nT,
                                        wseed.p, dtb.p, qb.p));
                const double* qs = st == 0 ? qn : keep[2 + st];
                HIPCHK(launch_adjoint_stage_metrics(s, h->cells64(), h->faces64(), qn, qs, wseed.p, h->gas64(),
                                                    h->opts.second_order, h->opts.limiter_k, A, lam.p, lamg.p, G));
                HIPCHK(launch_acc_metrics(s, G, nF, nC, facc.p, aacc.p));
                HIPCHK(launch_axpy4(s, lamg.p, 1.0, nC, nT, 0, qb.p));
                HIPCHK(launch_axpy4(s, lam.p, 1.0, nC, nT, st > 0 ? 1 : 0, st > 0 ? g.p : qb.p));
            }
            HIPCHK(launch_dt_vjp(s, h->cells64(), h->faces64(), qn, h->gas64(), h->dt.p, dtb.p, h->cfl, eb.p, qb.p));
            HIPCHK(launch_dt_vjp_geo(s, h->cells64(), h->faces64(), qn, h->gas64(), h->dt.p, dtb.p, h->cfl,
                                     facc.p));
            HIPCHK(launch_copy(s, qb.p, g.p, (int64_t)nq));
        }
    }
    HIPCHK(launch_copy(s, facc.p, G.fbar, 8 * (int64_t)nF));
    HIPCHK(launch_copy(s, aacc.p, G.abar, nC));
    HIPCHK(launch_geo_to_nodes(s, h->cells64(), h->faces64(), G));
    HIPCHK(hipMemcpyAsync(xbar, G.xbar, 2 * (size_t)h->nN * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // the chord extent: x of the extreme wall-face midpoints, CL proportional to 1 / ext
    const double eb_ = -(*cl) / ext;
    for (uint32_t e : emax)
        for (int end = 0; end < 2; ++end) xbar[2 * (size_t)m.edgesNodes[2 * (size_t)e + end]] += 0.5 * eb_ / emax.size();
    for (uint32_t e : emin)
        for (int end = 0; end < 2; ++end) xbar[2 * (size_t)m.edgesNodes[2 * (size_t)e + end]] -= 0.5 * eb_ / emin.size();
    return AFX_OK;
}

namespace {
// buffers and geometry copies of the tangent (first call), the wall distance
// it reads (every call: it follows the BCs)
int tangent_prepare(afx_rans* h) {
    if (!h->opts.fused || !h->tile64_ok)
        return fail(AFX_EUNSUP, "the tangent runs in the fused tile kernel (renumber = 1, fused = 1, tiles within the "
                                "LDS limit)");
    const int nv = h->nv;
    const size_t lds = tile_lds_bytes_tangent(h->tl.lmax, h->tl.gmax, h->tl.ftmax, nv);
    int dev = 0, maxl = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&maxl, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if ((size_t)maxl < lds)
        return fail(AFX_EUNSUP, "the tangent's tiles need " + std::to_string(lds) + " bytes of LDS per workgroup; the "
                                "device allows " + std::to_string(maxl) + " (use smaller tiles)");
    int rc;
    if (h->sa() && !h->wd_ready && (rc = compute_wall_distance(h))) return rc;
    hipStream_t s = h->stream;
    const int64_t nT = h->nT;
    if (!h->tg_q.p) {
        HIPCHK(h->tg_geo.alloc(h->t_geo64.n));
        HIPCHK(h->tg_area.alloc(h->area.n));
        HIPCHK(h->tg_lsq.alloc(h->lsq.n));
        HIPCHK(h->tg_xc.alloc(h->xc.n));
        HIPCHK(h->tg_yc.alloc(h->yc.n));
        HIPCHK(h->tg_wd.alloc(h->wall_d.n));   // the model's distance (sa_wd) with SA, else unused
        HIPCHK(h->tg_q.alloc((size_t)nv * nT));
        HIPCHK(h->tg_R.alloc((size_t)nv * nT));
        HIPCHK(h->tg_v.alloc((size_t)nv * nT));
        HIPCHK(h->tg_out.alloc((size_t)nv * nT));
        HIPCHK(launch_to_dual(s, h->t_geo64.p, (int64_t)h->t_geo64.n, h->tg_geo.p));
        HIPCHK(launch_to_dual(s, h->area.p, (int64_t)h->area.n, h->tg_area.p));
        HIPCHK(launch_to_dual(s, h->lsq.p, (int64_t)h->lsq.n, h->tg_lsq.p));
        HIPCHK(launch_to_dual(s, h->xc.p, (int64_t)h->xc.n, h->tg_xc.p));
        HIPCHK(launch_to_dual(s, h->yc.p, (int64_t)h->yc.n, h->tg_yc.p));
        HIPCHK(hipMemsetAsync(h->tg_v.p, 0, (size_t)nv * nT * 8, s));
        // the tile kernel writes the real rows only: ghost rows of J v stay 0
        HIPCHK(hipMemsetAsync(h->tg_R.p, 0, (size_t)nv * nT * sizeof(Dual1), s));
    }
    if (h->sa())   // follows the BCs
        HIPCHK(launch_to_dual(s, h->sa_wd.p, (int64_t)h->sa_wd.n, h->tg_wd.p));
    return AFX_OK;
}
// (q, v in tg_v) -> tg_out = (dR/dq) v, on the handle's stream
int enqueue_tangent(afx_rans* h) {
    hipStream_t s = h->stream;
    const int nv = h->nv;
    const int64_t nT = h->nT;
    HIPCHK(launch_dual_state(s, h->q.p, h->tg_v.p, h->nT, h->nC, nv, h->tg_q.p));
    HIPCHK(launch_residual_tile_tangent(s, h->tiles<Dual1>(h->tg_geo.p), h->tg_q.p, h->nT, h->nC, h->tg_area.p,
                                        h->tg_lsq.p, h->tg_xc.p, h->tg_yc.p, h->tg_wd.p, h->gas64(), h->opts.gradient,
                                        h->so(), h->opts.limiter_k, h->opts.viscosity, h->tg_R.p));
    HIPCHK(launch_dual_parts(s, h->tg_R.p, (int64_t)nv * nT, nullptr, h->tg_out.p));
    return AFX_OK;
}
}  // namespace

int afx_rans_tangent(afx_rans* h, const double* v, const double* v_nu, double* jv, double* jv_nu) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (!v || !jv) return fail(AFX_EARG, "null direction or output");
    if (!h->sa() && (v_nu || jv_nu)) return fail(AFX_EARG, "rho nu~ direction / output only with the Spalart-Allmaras model");
    if ((rc = tangent_prepare(h))) return rc;
    hipStream_t s = h->stream;
    const int nv = h->nv;
    const int64_t nT = h->nT;
    if ((rc = upload_aos(h, v, h->tg_v.p))) return rc;
    if (nv == 5) {
        std::vector<double> tmp(h->nT, 0.0);
        if (v_nu)
            for (int32_t i = 0; i < h->nT; ++i) tmp[i] = v_nu[h->c_new2old[i]];
        HIPCHK(hipMemcpyAsync(h->tg_v.p + 4 * nT, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if ((rc = enqueue_tangent(h))) return rc;
    if ((rc = download_aos(h, h->tg_out.p, jv))) return rc;
    if (jv_nu && (rc = download_cells(h, h->tg_out.p + 4 * nT, jv_nu))) return rc;
    return AFX_OK;
}

int afx_rans_time_tangent(afx_rans* h, int32_t reps, double* ms) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (reps < 1 || !ms) return fail(AFX_EARG, "reps must be >= 1 and ms non-null");
    if ((rc = tangent_prepare(h))) return rc;
    if ((rc = enqueue_tangent(h))) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, h->stream));
    for (int32_t i = 0; i < reps; ++i)
        if ((rc = enqueue_tangent(h))) return rc;
    HIPCHK(hipEventRecord(b, h->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    *ms = (double)t / reps;
    return AFX_OK;
}

int afx_rans_time_adjoint(afx_rans* h, int32_t reps, double* ms) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (reps < 1 || !ms) return fail(AFX_EARG, "reps must be >= 1 and ms non-null");
    if ((rc = adjoint_check(h, true))) return rc;
    if ((rc = enqueue_adjoint(h))) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, h->stream));
    for (int32_t i = 0; i < reps; ++i)
        if ((rc = enqueue_adjoint(h))) return rc;
    HIPCHK(hipEventRecord(b, h->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    *ms = (double)t / reps;
    return AFX_OK;
}

int afx_rans_time_residual(afx_rans* h, int32_t precision, int32_t reps, double* ms) {
    int rc = check_handle(h);
    if (rc) return rc;
    if (reps < 1 || !ms) return fail(AFX_EARG, "reps must be >= 1 and ms non-null");
    if ((rc = enqueue_residual(h, precision))) return rc;   // warm-up (and FP32 state copy)
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, h->stream));
    for (int32_t i = 0; i < reps; ++i)
        if ((rc = enqueue_residual(h, precision))) return rc;
    HIPCHK(hipEventRecord(b, h->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    *ms = (double)t / reps;
    return AFX_OK;
}

// ------------------------------------------------------------------ batch --
// Many flow conditions / meshes advanced together: each pseudo-time step is
// the explicit three-stage scheme of explicitSolver::solve (solver.h:802-828);
// each stage is ONE launch of the FP32 tile kernel over the (member, tile)
// work items of all members, which also applies the FP64 stage update of its
// cells (and, at stage 0, computes the local time step).  The FP32 stage
// states rotate through three buffers per member (A -> B -> C -> A), so a
// launch never rewrites the states it reads.  Steps are replayed from a
// captured hipGraph.
struct afx_batch {
    std::vector<afx_rans*> mem;
    hipStream_t stream = nullptr;
    DBuf<TileArgsF32> stage[3];
    DBuf<int2> items;
    int32_t nitems = 0;
    size_t lds = 0, lds_dt = 0;
    bool sa = false;
    int viscous = 0, gg = 1, so = 1;   // the members' common model, gradient scheme and order
    double relax = 1.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t graph_steps = 0;
    ~afx_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
int batch_step(afx_batch* b) {
    for (int st = 0; st < 3; ++st)
        HIPCHK(launch_residual_tile_f32_batch(b->stream, b->stage[st].p, b->items.p, b->nitems,
                                              st == 0 ? b->lds_dt : b->lds, b->viscous, b->gg, b->so));
    return AFX_OK;
}
int batch_tables(afx_batch* b) {
    static const double alpha[3] = {0.25, 0.5, 1.};
    const size_t n = b->mem.size();
    std::vector<TileArgsF32> ta[3];
    for (auto& v : ta) v.resize(n);
    std::vector<int2> items;
    b->lds = b->lds_dt = 0;
    for (size_t i = 0; i < n; ++i) {
        afx_rans* h = b->mem[i];
        const double k = h->opts.limiter_k;
        float4* X[3] = {h->q32v.p, h->q32v_b.p, h->q32v_c.p};
        float* Xn[3] = {h->q32n.p, h->q32n_b.p, h->q32n_c.p};
        for (int st = 0; st < 3; ++st) {
            SAFields32 saf{Xn[st], h->sa_ad32.p, nullptr, nullptr};
            TileArgsF32 A{h->tiles32(), X[st], h->iarea32.p, h->a15_32.p, h->lsq32.p, h->xc32.p, h->yc32.p,
                          nullptr, nullptr, saf, st == 0 ? h->dt32.p : nullptr, h->gas32(), (float)(k * k * k),
                          (float)h->cfl, h->nC, h->opts.gradient, h->opts.second_order, h->opts.viscosity,
                          h->q.p, st == 2 ? h->q.p : h->qk.p, X[(st + 1) % 3], Xn[(st + 1) % 3],
                          st == 0 ? nullptr : h->dt32.p, h->area.p, alpha[st] * b->relax, h->nT};
            ta[st][i] = A;
        }
        const int nv = h->nv;
        b->lds = std::max(b->lds, tile_lds_bytes_f32(ta[0][i].S, nv, false));
        b->lds_dt = std::max(b->lds_dt, tile_lds_bytes_f32(ta[0][i].S, nv, true));
        for (int32_t t = 0; t < h->tl.ntiles; ++t) items.push_back(make_int2((int)i, t));
    }
    while (items.size() % 8) items.push_back(make_int2(-1, -1));
    if (b->lds_dt > 65536) return fail(AFX_EUNSUP, "a member's tiles exceed 64 KiB of LDS");
    b->nitems = (int32_t)items.size();
    hipStream_t s = b->stream;
    if (!b->items.p) {
        for (auto& d : b->stage) HIPCHK(d.alloc(n));
        HIPCHK(b->items.alloc(items.size()));
    }
    for (int st = 0; st < 3; ++st) HIPCHK(b->stage[st].upload(ta[st], s));
    HIPCHK(b->items.upload(items, s));
    HIPCHK(hipStreamSynchronize(s));
    return AFX_OK;
}
}  // namespace

int afx_batch_create(int32_t n, afx_rans* const* members, double relaxation, afx_batch** out) {
    if (n < 1 || !members || !out) return fail(AFX_EARG, "need at least one member and an output pointer");
    auto b = new afx_batch();
    b->relax = relaxation;
    for (int32_t i = 0; i < n; ++i) {
        afx_rans* h = members[i];
        int rc = check_handle(h);
        if (!rc && !(h->tile32_ok && h->opts.fused && h->iarea32.p))
            rc = fail(AFX_EUNSUP, "batch members need the fused tile kernels (renumber = 1, fused = 1)");
        if (!rc && h->michalak())
            rc = fail(AFX_EUNSUP, "batched stepping runs the default limiter (the FP32 fused kernel), not the Michalak build");
        if (!rc && i > 0 && (h->opts.viscosity != members[0]->opts.viscosity ||
                             h->opts.gradient != members[0]->opts.gradient ||
                             h->opts.second_order != members[0]->opts.second_order))
            rc = fail(AFX_EARG, "batch members must share the viscosity model, gradient scheme and order "
                                "(one kernel instantiation per launch)");
        if (!rc && h->sa() && !h->wd_ready) rc = compute_wall_distance(h);
        if (!rc && !h->q32v_b.p) {   // the two further FP32 stage buffers
            if (h->q32v_b.alloc(h->nT) != hipSuccess || h->q32v_c.alloc(h->nT) != hipSuccess ||
                (h->sa() && (h->q32n_b.alloc(h->nT) != hipSuccess || h->q32n_c.alloc(h->nT) != hipSuccess)))
                rc = fail(AFX_EHIP, "hipMalloc of the batch stage buffers failed");
        }
        if (rc) { delete b; return rc; }
        b->mem.push_back(h);
    }
    b->sa = members[0]->sa();
    b->viscous = members[0]->opts.viscosity;
    b->gg = members[0]->opts.gradient;
    b->so = members[0]->opts.second_order;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(AFX_EHIP, "hipStreamCreateWithFlags failed");
    }
    int rc = batch_tables(b);
    if (rc) { delete b; return rc; }
    *out = b;
    return AFX_OK;
}

int afx_batch_run(afx_batch* b, int32_t steps, int32_t graph_steps) {
    if (!b) return fail(AFX_EARG, "null batch");
    if (steps < 0 || graph_steps < 0) return fail(AFX_EARG, "steps and graph_steps must be >= 0");
    // current CFL numbers, then the FP32 copies of every member's state
    // (ghosts included) before the first stage; the tables keep their device
    // addresses, so a captured graph reads the refreshed contents
    int rt = batch_tables(b);
    if (rt) return rt;
    for (afx_rans* h : b->mem) {
        int rc = ensure_q32(h);
        if (rc) return rc;
        // ghost rows of every stage buffer (farfield states stay constant)
        HIPCHK(hipMemcpyAsync(h->q32v_b.p, h->q32v.p, (size_t)h->nT * 16, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->q32v_c.p, h->q32v.p, (size_t)h->nT * 16, hipMemcpyDeviceToDevice, h->stream));
        if (h->sa()) {
            HIPCHK(hipMemcpyAsync(h->q32n_b.p, h->q32n.p, (size_t)h->nT * 4, hipMemcpyDeviceToDevice, h->stream));
            HIPCHK(hipMemcpyAsync(h->q32n_c.p, h->q32n.p, (size_t)h->nT * 4, hipMemcpyDeviceToDevice, h->stream));
        }
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    int32_t done = 0;
    if (graph_steps > 0 && steps >= graph_steps) {
        if (b->graph_steps != graph_steps) {
            if (b->exec) { HIPCHK(hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
            if (b->graph) { HIPCHK(hipGraphDestroy(b->graph)); b->graph = nullptr; }
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            for (int32_t i = 0; i < graph_steps; ++i) {
                int rc = batch_step(b);
                if (rc) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(b->stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
            }
            HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
            HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_steps = graph_steps;
        }
        for (; done + graph_steps <= steps; done += graph_steps) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    }
    for (; done < steps; ++done) {
        int rc = batch_step(b);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    for (afx_rans* h : b->mem) h->q32_valid = true;   // q32v / q32n hold the final state
    return AFX_OK;
}

int afx_batch_residual_norms(afx_batch* b, int32_t precision, double* norms) {
    if (!b || !norms) return fail(AFX_EARG, "null argument");
    for (size_t i = 0; i < b->mem.size(); ++i) {
        int rc = afx_rans_residual(b->mem[i], precision, nullptr, &norms[i]);
        if (rc) return rc;
    }
    return AFX_OK;
}

int afx_batch_stream(afx_batch* b, void** stream) {
    if (!b || !stream) return fail(AFX_EARG, "null argument");
    *stream = (void*)b->stream;
    return AFX_OK;
}

int afx_batch_destroy(afx_batch* b) {
    if (b) {
        (void)hipStreamSynchronize(b->stream);
        delete b;
    }
    return AFX_OK;
}

int afx_rans_synchronize(afx_rans* h) {
    int rc = check_handle(h, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return AFX_OK;
}

int afx_rans_destroy(afx_rans* h) {
    if (!h) return AFX_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
    return AFX_OK;
}

// ------------------------------------------------- multigrid prolongation --
struct afx_mapper {
    const afx_rans* coarse = nullptr;
    const afx_rans* fine = nullptr;
    int32_t nc = 0, nf = 0;
    int64_t nnz = 0;
    DBuf<int64_t> rowptr;           // reference fine order, nf + 1
    DBuf<int32_t> cols, cols_dev;   // coarse reference index / coarse device index
    DBuf<double> vals;
    DBuf<int32_t> fnew2old;         // fine device cell -> reference cell
};

int afx_mapper_create(const afx_rans* coarse, const afx_rans* fine, afx_mapper** out) {
    if (!out) return fail(AFX_EARG, "null output");
    *out = nullptr;
    if (!coarse || !fine) return fail(AFX_EARG, "null solver handle");
    if (coarse->nv != fine->nv) return fail(AFX_EARG, "the two levels carry different numbers of variables");
    auto* mp = new afx_mapper();
    auto bail = [&](int rc) { delete mp; return rc; };
#define HIPCHK_M(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return bail(fail(AFX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_))); \
    } while (0)
    mp->coarse = coarse;
    mp->fine = fine;
    mp->nc = coarse->nT;
    mp->nf = fine->nT;
    hipStream_t s = fine->stream;
    // centroids and areas in reference order, ghosts included (multigrid.h:101-102
    // loop over cellsAreas.size())
    const Mesh& mc = coarse->m;
    const Mesh& mf = fine->m;
    DBuf<double> cx, cy, ca, fx, fy, box;
    DBuf<int32_t> cnt, o2n;
    const int32_t ntiles = (mp->nc + 255) / 256;
    HIPCHK_M(cx.alloc(mp->nc)); HIPCHK_M(cx.upload(mc.cellsCentersX.data(), mp->nc, s));
    HIPCHK_M(cy.alloc(mp->nc)); HIPCHK_M(cy.upload(mc.cellsCentersY.data(), mp->nc, s));
    HIPCHK_M(ca.alloc(mp->nc)); HIPCHK_M(ca.upload(mc.cellsAreas.data(), mp->nc, s));
    HIPCHK_M(fx.alloc(mp->nf)); HIPCHK_M(fx.upload(mf.cellsCentersX.data(), mp->nf, s));
    HIPCHK_M(fy.alloc(mp->nf)); HIPCHK_M(fy.upload(mf.cellsCentersY.data(), mp->nf, s));
    HIPCHK_M(box.alloc((size_t)4 * ntiles));
    HIPCHK_M(cnt.alloc(mp->nf));
    HIPCHK_M(launch_map_tile_boxes(s, cx.p, cy.p, ca.p, mp->nc, box.p));
    HIPCHK_M(launch_map_count(s, fx.p, fy.p, mp->nf, cx.p, cy.p, ca.p, mp->nc, box.p, cnt.p));
    std::vector<int32_t> hc(mp->nf);
    HIPCHK_M(hipMemcpyAsync(hc.data(), cnt.p, (size_t)mp->nf * 4, hipMemcpyDeviceToHost, s));
    HIPCHK_M(hipStreamSynchronize(s));
    std::vector<int64_t> rp((size_t)mp->nf + 1, 0);
    for (int32_t i = 0; i < mp->nf; ++i) rp[i + 1] = rp[i] + hc[i];
    mp->nnz = rp.back();
    HIPCHK_M(mp->rowptr.alloc(rp.size())); HIPCHK_M(mp->rowptr.upload(rp, s));
    HIPCHK_M(mp->cols.alloc(mp->nnz));
    HIPCHK_M(mp->cols_dev.alloc(mp->nnz));
    HIPCHK_M(mp->vals.alloc(mp->nnz));
    HIPCHK_M(launch_map_fill(s, fx.p, fy.p, mp->nf, cx.p, cy.p, ca.p, mp->nc, box.p, mp->rowptr.p, mp->cols.p,
                             mp->vals.p));
    HIPCHK_M(o2n.alloc(mp->nc)); HIPCHK_M(o2n.upload(coarse->c_old2new, s));
    HIPCHK_M(launch_map_cols(s, mp->cols.p, o2n.p, mp->nnz, mp->cols_dev.p));
    HIPCHK_M(mp->fnew2old.alloc(mp->nf)); HIPCHK_M(mp->fnew2old.upload(fine->c_new2old, s));
    HIPCHK_M(hipStreamSynchronize(s));
#undef HIPCHK_M
    *out = mp;
    return AFX_OK;
}

int afx_mapper_destroy(afx_mapper* mp) {
    if (!mp) return AFX_OK;
    if (mp->fine) (void)hipStreamSynchronize(mp->fine->stream);
    delete mp;
    return AFX_OK;
}

int afx_mapper_get(const afx_mapper* mp, int64_t* nnz, int64_t* row_ptr, int32_t* cols, double* weights) {
    if (!mp) return fail(AFX_EARG, "null mapper");
    if (nnz) *nnz = mp->nnz;
    hipStream_t s = mp->fine->stream;
    if (row_ptr) HIPCHK(hipMemcpyAsync(row_ptr, mp->rowptr.p, ((size_t)mp->nf + 1) * 8, hipMemcpyDeviceToHost, s));
    if (cols) HIPCHK(hipMemcpyAsync(cols, mp->cols.p, (size_t)mp->nnz * 4, hipMemcpyDeviceToHost, s));
    if (weights) HIPCHK(hipMemcpyAsync(weights, mp->vals.p, (size_t)mp->nnz * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return AFX_OK;
}

int afx_mapper_prolongate(const afx_mapper* mp, afx_rans* coarse, afx_rans* fine) {
    if (!mp) return fail(AFX_EARG, "null mapper");
    if (coarse != mp->coarse || fine != mp->fine)
        return fail(AFX_EARG, "the mapper was generated for another pair of solvers");
    // the coarse state must be complete before the fine stream reads it
    HIPCHK(hipStreamSynchronize(coarse->stream));
    HIPCHK(launch_map_prolong(fine->stream, mp->rowptr.p, mp->fnew2old.p, mp->cols_dev.p, mp->vals.p, coarse->q.p,
                              coarse->nT, fine->q.p, fine->nT, fine->nv));
    HIPCHK(hipStreamSynchronize(fine->stream));
    fine->q32_valid = false;
    return AFX_OK;
}

}  // extern "C"

"""Residual-throughput benchmark (one GPU).

One step = one evaluation of the unstructured finite-volume RANS residual
with the Spalart-Allmaras model (gradients + limiters of the five variables,
Roe + viscous + nu~ face fluxes, SA sources, signed cell sums) on a 2.0M-cell
jittered mixed triangle/quad mesh around a NACA 0012 section, Re 1e6 on the
chord, state = freestream (M 0.7, alpha 1.5 deg, rho nu~ = 3 mu) with 1%
seeded Gaussian perturbations of rho, u, v, p and nu~.  The timed form is FP32
(FP64-accumulated cell sums); the plain FP64 form is timed beside it.  Inputs
are resident in HBM before timing.  --model inviscid runs the four-variable
Euler residual instead.

Prints one JSON line (see the keys below).  Kernel durations come from HIP
events recorded on the library's own stream around back-to-back launches.
The CPU baseline is the oracle (serial FP64 restatement of the reference's
fillRhoRHS with the same SA terms, calc_dt included) on a bounded sample
mesh from the same generator.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells N] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E peak, GB/s
TILE_CELLS = 200        # cells per tile of the fused residual (tile-size sweep on MI355X: 160-256)
FP64_TILE_CELLS = 160   # the plain FP64 form's (sweep 64-200: 0.82, 0.62, 0.51, 0.51, 0.55 ms at 2.0M cells)


class HipEvents:
    """hipEvent timing on an arbitrary stream (the library's).  The HIP
    functions are looked up through the library's handle, i.e. in the HIP
    runtime the library is linked against."""

    def __init__(self):
        from aeroflex_amd.rans import LIB_PATH
        self.hip = C.CDLL(LIB_PATH)
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = C.c_void_p()
        rc = self.hip.hipEventCreate(C.byref(e))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")
        return e

    def record(self, e, stream):
        rc = self.hip.hipEventRecord(e, C.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = C.c_float()
        rc = self.hip.hipEventElapsedTime(C.byref(ms), a, b)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return ms.value

    def device_sync(self):
        rc = self.hip.hipDeviceSynchronize()
        if rc != 0:
            raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")

    def copy_probe(self, mib, reps=2):
        """Device-to-device copies of a known size (a plain copy kernel whose
        FETCH_SIZE / WRITE_SIZE readings calibrate the traffic counters)."""
        h = self.hip
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        n = mib << 20
        a, b = C.c_void_p(), C.c_void_p()
        if h.hipMalloc(C.byref(a), n) or h.hipMalloc(C.byref(b), n):
            raise RuntimeError("hipMalloc failed")
        h.hipMemset(a, 1, n)
        for _ in range(reps):
            if h.hipMemcpy(b, a, n, 3):   # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        self.device_sync()
        h.hipFree(a)
        h.hipFree(b)


def mesh_dims(target_cells, quad_frac=0.25):
    """O-grid size (ni around, nj radial) for about target_cells cells."""
    per_quad = 2.0 - quad_frac
    nq = target_cells / per_quad
    ni = max(64, int(round(np.sqrt(nq / 1.09) / 64.0)) * 64)
    nj = int(round(nq / ni)) + 1
    return ni, nj


MACH, ALPHA_DEG, REYNOLDS = 0.7, 1.5, 1e6


def freestream_mu(rans):
    """mu_L for Re 1e6 on the unit chord at the freestream state."""
    g = rans.Gas()
    q = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG)).get_conservative(g)
    return float(np.hypot(q[1], q[2]) / REYNOLDS)   # rho |u| c / Re with c = 1


def perturbed(q, nu, n, qinf, gamma, rng):
    """Freestream with 1% seeded Gaussian perturbations of rho, u, v, p, nu~."""
    rho = qinf[0] * (1 + 0.01 * rng.standard_normal(n))
    u = qinf[1] / qinf[0] * (1 + 0.01 * rng.standard_normal(n))
    v = qinf[2] / qinf[0] * (1 + 0.01 * rng.standard_normal(n))
    p0 = (gamma - 1) * (qinf[3] - 0.5 * (qinf[1] ** 2 + qinf[2] ** 2) / qinf[0])
    p = p0 * (1 + 0.01 * rng.standard_normal(n))
    q[0:4 * n:4] = rho
    q[1:4 * n:4] = rho * u
    q[2:4 * n:4] = rho * v
    q[3:4 * n:4] = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    if nu is not None:
        nut_inf = nu[0] / qinf[0]
        nu[:n] = rho * nut_inf * (1 + 0.01 * rng.standard_normal(n))
    return q, nu


def build_case(rans, target_cells, model="spalart-allmaras", seed_mesh=1, seed_state=2, tile_cells=0):
    ni, nj = mesh_dims(target_cells)
    t0 = time.perf_counter()
    m = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25,
                           seed=seed_mesh)
    t_mesh = time.perf_counter() - t0
    s, mu = case_solver(rans, m, model, seed_state, tile_cells)
    return m, s, dict(ni=ni, nj=nj, mesh_seconds=t_mesh, mu_L=mu)


def case_solver(rans, m, model="spalart-allmaras", seed_state=2, tile_cells=0):
    """The headline workload's solver and seeded state on mesh m."""
    sa = model == "spalart-allmaras"
    gas = rans.Gas(mu_L=freestream_mu(rans)) if sa else rans.Gas()
    s = rans.Solver(m, gas=gas, viscosity_model=model, second_order=True, gradient_scheme="green-gauss",
                    limiter_k=5.0, tile_cells=tile_cells)
    far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far),
               "wall": rans.BoundaryCondition("wall" if sa else "slip-wall")})
    s.init()
    s.refill_bcs()
    qinf = far.get_conservative(s.gas)
    rng = np.random.default_rng(seed_state)
    q, nu = perturbed(s.get_q(), s.get_nu() if sa else None, s.n_real, qinf, s.gas.gamma, rng)
    s.set_q(q)
    if sa:
        s.set_nu(nu)
    return s, gas.mu_L


def alg_bytes(nc, nf, nb, word, V=4, index_bytes=4):
    """Algorithmic HBM bytes of one residual evaluation: per real cell state V,
    geometry 4 (area, x_c, y_c, d), gradients 2V written + 2V read, limiter V
    written + V read, residual V written; per face two cell indices and
    n_x, n_y, length, x_f, y_f; per boundary face a BC tag and a ghost state."""
    return word * nc * (8 * V + 4) + nf * (2 * index_bytes + 5 * word) + word * nb * (V + 1)


def time_stage(ev, s, stream, prec, stage, n):
    a, b = ev.event(), ev.event()
    s.synchronize()
    ev.record(a, stream)
    s.launch_stage(prec, stage, n)
    ev.record(b, stream)
    return ev.elapsed_ms(a, b) / n


def fp32_accuracy(s, nc, V):
    """The FP32 residual against the plain FP64 form on the bench mesh and
    state (the FP64 form equals the oracle's residual to 1e-12 in the test
    suite): relative L2 norm of the difference over the real cells' V rows,
    and its max norm relative to max |R|."""
    from aeroflex_amd import rans
    R64, _ = s.fillRhoRHS(rans.AFX_FP64)
    rows64 = [R64.reshape(-1, 4)[:nc]]
    if V == 5:
        rows64.append(s.nu_residual()[:nc, None])
    R32, _ = s.fillRhoRHS(rans.AFX_FP32)
    rows32 = [R32.reshape(-1, 4)[:nc]]
    if V == 5:
        rows32.append(s.nu_residual()[:nc, None])
    a, b = np.hstack(rows64), np.hstack(rows32)
    d = b - a
    return {"rel_l2": float(f"{np.linalg.norm(d) / np.linalg.norm(a):.3e}"),
            "max_abs_over_max_R": float(f"{np.max(np.abs(d)) / np.max(np.abs(a)):.3e}"),
            "rows": "real cells, all V variables"}


def fp32_forces(rans, s, steps=20, cfl=0.3, relax=0.8):
    """CL and CD (get_wall_profile, post.h:301-387) after `steps` explicit
    three-stage pseudo-time steps from the bench state, with FP64 residuals
    and FP64 updates (the plain form) and with FP32 residuals and FP64 updates
    (the batched launch with one member), on the headline mesh.  The device's
    FP64 steps match the oracle's to 1e-10 in CL/CD in the test suite.  The
    solver's state is restored afterwards."""
    q0, nu0 = s.get_q(), s.get_nu()
    s32 = s.clone()
    s.set_cfl(cfl)
    s32.set_cfl(cfl)
    try:
        for _ in range(steps):
            s.explicit_solve(relax)
        cl64, cd64, _ = s.get_wall_profile("wall")
        with rans.Batch([s32], relaxation=relax) as b:
            b.run(steps, 0)
        cl32, cd32, _ = s32.get_wall_profile("wall")
    finally:
        s32.close()
        s.set_q(q0)
        s.set_nu(nu0)
    return {"steps": steps, "cfl": cfl, "relaxation": relax, "cl_fp64": cl64, "cd_fp64": cd64, "cl_fp32": cl32,
            "cd_fp32": cd32, "cl_rel_err": float(f"{abs(cl32 - cl64) / abs(cl64):.3e}"),
            "cd_rel_err": float(f"{abs(cd32 - cd64) / abs(cd64):.3e}"),
            "note": "FP32 residual + FP64 update vs FP64 residual + FP64 update, same start state and steps"}


def read_traffic(path):
    """HBM bytes per residual from a committed rocprofv3 PMC summary, if any."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_residual"])
    except (KeyError, ValueError):
        return None


def _oracle_case(path, model):
    """The oracle's solver on the sample mesh file with the bench state;
    returns (fill, real cells)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    om = orc.OMesh.read(path)
    sa = model == "spalart-allmaras"
    if sa:   # mu_L for Re 1e6 on the unit chord, from the oracle's freestream state
        qf = orc.freestream(MACH, np.deg2rad(ALPHA_DEG))
        mu = float(np.hypot(qf[1], qf[2]) / REYNOLDS)
    else:
        mu = 1e-5
    o = orc.OSolver(om, mu_L=mu, viscous=2 if sa else 0)
    o.set_bcs({"farfield": ("farfield", (MACH, np.deg2rad(ALPHA_DEG), 1.0, 1.0)),
               "wall": ("wall" if sa else "slip-wall", (0.2, 0, 1, 1))})
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=1.0)
    o.init()
    o.refill_bcs()
    nc = om.counts()["real_cells"]
    rng = np.random.default_rng(2)
    qinf = orc.freestream(MACH, np.deg2rad(ALPHA_DEG))
    nu = np.full(o.n_cells, 3.0 * mu) if sa else None
    q, nu = perturbed(o.vec("q"), nu, nc, qinf, 1.4, rng)
    o.vec("q", q)
    if sa:
        o.sa_vec("nu", nu)
    fill = o.sa_fill_rhs if sa else o.fill_rhs
    return o, fill, nc


def _timed_fills(fill, budget_s, cap=2000):
    fill()
    n, t0 = 0, time.perf_counter()
    while True:
        fill()
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= cap:
            return n, el


def _cpu_worker(args):
    """One core of the all-cores CPU leg (a fresh interpreter: no GPU)."""
    path, model, budget_s = args
    o, fill, nc = _oracle_case(path, model)
    n, el = _timed_fills(fill, budget_s)
    return nc, n, el


def vtu_bench(rans, s, model, cpu=True):
    """save (post.h:58-180) of the headline case's state: the reference's
    ASCII VTU written from the device state (wall distance computed before
    the timed call), and beside it the oracle's one-thread restatement on the
    120k-cell CPU sample, scaled by cells."""
    import tempfile
    s.wall_dist()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "solution.vtu")
        t0 = time.perf_counter()
        s.save_vtu(path)
        t = time.perf_counter() - t0
        size = os.path.getsize(path)
        out = {"vtu_cells": s.n_real, "vtu_s": round(t, 3), "vtu_mb": round(size / 1e6, 1),
               "vtu_threads": max(1, min(16, os.cpu_count() or 1))}
        # CpProfile::calc_cp (post.h:246-290), refreshed every iteration by
        # the reference's loop: a gather of the wall faces on the device
        s.cp_profile("wall")
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            x, _, cp = s.cp_profile("wall")
        out["cp_profile_ms"] = round((time.perf_counter() - t0) * 1e3 / reps, 3)
        out["cp_profile_faces"] = int(len(cp))
        if cpu:
            ni, nj = mesh_dims(120_000)
            m = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25,
                                   seed=1)
            mp = os.path.join(d, "sample.msh")
            m.write_msh(mp)
            o, _, nc = _oracle_case(mp, model)
            o.wall_dist()
            t0 = time.perf_counter()
            o.save_vtu(os.path.join(d, "oracle.vtu"))
            to = time.perf_counter() - t0
            out["cpu_baseline"] = {"kind": "port", "cores": 1, "sample": f"oracle save on {nc} cells, {to:.2f} s",
                                   "vtu_s_at_cells": round(to * s.n_real / nc, 2)}
    return out


def cpu_baseline(rans, model, budget_s=12.0):
    """Oracle (serial FP64 fillRhoRHS with the same model, calc_dt included) on
    a 120k-cell sample mesh of the same generator and state: one thread, and
    beside it the same serial oracle in one process per core of the box's
    share, summed (independent residual evaluations, the CPU's best use of
    its cores without changing the oracle's serial face loops)."""
    import multiprocessing
    import tempfile
    ni, nj = mesh_dims(120_000)
    m = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25, seed=1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "sample.msh")
        m.write_msh(path)
        o, fill, nc = _oracle_case(path, model)
        n, el = _timed_fills(fill, budget_s)
        out = {"value": round(nc * n / el / 1e6, 4), "unit": "Mcell-residuals/s", "cores": 1, "kind": "port",
               "sample": f"oracle fillRhoRHS ({model}, FP64, serial, calc_dt included) x{n} on a {nc}-cell mesh "
                         f"of the same generator and state, {el:.1f} s"}
        cores = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16")), os.cpu_count() or 1, 16))
        with multiprocessing.get_context("spawn").Pool(cores) as pool:
            res = pool.map(_cpu_worker, [(path, model, budget_s)] * cores)
    rate = sum(r[0] * r[1] / r[2] for r in res) / 1e6
    out["all_cores"] = {"value": round(rate, 4), "unit": "Mcell-residuals/s", "cores": cores, "kind": "port",
                        "sample": f"{cores} processes, each the serial oracle on the same sample for "
                                  f"{budget_s:.0f} s, rates summed"}
    out["cpu_model"] = cpu_model()
    return out


def cpu_model():
    """The host CPU's model name (first entry of /proc/cpuinfo), if readable."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


SWEEP_CFL = 0.3


def sweep_solver(rans, mesh, mach, alpha_deg, reynolds, tile_cells=0, base=None):
    """SA solver for one condition (mu_L from Re on the unit chord), state =
    freestream with the 1% perturbations of the headline workload."""
    g = rans.Gas()
    far = rans.BoundaryVariables(mach=mach, angle=np.deg2rad(alpha_deg))
    qinf = far.get_conservative(g)
    gas = rans.Gas(mu_L=float(np.hypot(qinf[1], qinf[2]) / reynolds))
    if base is None:
        s = rans.Solver(mesh, gas=gas, viscosity_model="spalart-allmaras", tile_cells=tile_cells)
    else:
        s = base.like(gas)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("wall")})
    s.init()
    s.refill_bcs()
    q, nu = perturbed(s.get_q(), s.get_nu(), s.n_real, qinf, g.gamma, np.random.default_rng(2))
    s.set_q(q)
    s.set_nu(nu)
    s.set_cfl(SWEEP_CFL)
    return s


def batched_sweep(rans, steps, graph_steps=25, n_cond=64, n_mesh=24, cond_cells=250_000):
    """64 conditions on one 250k-cell mesh (alpha ~ U[-4, 8] deg, M ~ U[0.3,
    0.75], Re log-uniform in [1e6, 1e7], seed 4) plus 24 meshes of log-uniform
    size in [2e4, 5e5] cells (seed 5; M 0.5, alpha 2 deg, Re 1e6; first-layer
    height 2e-3 sqrt(16384 / cells)), every mesh's tile size autotuned; all
    advanced together (FP32 fused residuals of all members in one launch per
    stage, FP64 updates, hipGraph replay) against the plain form: one FP64
    explicit solve per condition, in sequence, for the same number of steps."""
    t_setup = time.perf_counter()
    rng = np.random.default_rng(4)
    alpha = rng.uniform(-4.0, 8.0, n_cond)
    mach = rng.uniform(0.3, 0.75, n_cond)
    reyn = 10 ** rng.uniform(6.0, 7.0, n_cond)
    cands = (128, 160, 192, 224)
    ni, nj = mesh_dims(cond_cells)
    m0 = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25, seed=1)
    tile0, _ = rans.autotune_tile(lambda t: sweep_solver(rans, m0, mach[0], alpha[0], reyn[0], t), cands,
                                  reps=4, replicas=8)
    conds = [sweep_solver(rans, m0, mach[0], alpha[0], reyn[0], tile0)]
    conds += [sweep_solver(rans, m0, mach[i], alpha[i], reyn[i], base=conds[0]) for i in range(1, n_cond)]
    rng5 = np.random.default_rng(5)
    sizes = 10 ** rng5.uniform(np.log10(2e4), np.log10(5e5), n_mesh)
    uneven, tiles = [], []
    for k, nc in enumerate(sizes):
        a, b = mesh_dims(int(nc))
        # first-layer height scaled with the resolution (2e-3 at 16k cells, the
        # O-grid's), so that coarse meshes keep explicit-stable aspect ratios
        fh = 2e-3 * np.sqrt(16384.0 / nc)
        mk = rans.Mesh.jittered(a, b, r_far=20.0, thickness=0.12, first_h=fh, jitter=0.1, quad_frac=0.25,
                                seed=100 + k)
        tk, _ = rans.autotune_tile(lambda t: sweep_solver(rans, mk, 0.5, 2.0, 1e6, t), (128, 192), reps=4,
                                   replicas=4)
        tiles.append(tk)
        uneven.append(sweep_solver(rans, mk, 0.5, 2.0, 1e6, tk))
    t_setup = time.perf_counter() - t_setup

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    # every timed phase starts from the same freestream states
    init = [(m.get_q(), m.get_nu()) for m in conds + uneven]

    def reset():
        for m, (q, nu) in zip(conds + uneven, init):
            m.set_q(q)
            m.set_nu(nu)

    everyone = rans.Batch(conds + uneven, relaxation=0.8)
    everyone.run(graph_steps, graph_steps)   # warm-up + graph capture
    reset()
    t_all = timed(lambda: everyone.run(steps, graph_steps))
    norms = everyone.residual_norms(rans.AFX_FP64)
    only = rans.Batch(conds, relaxation=0.8)
    only.run(graph_steps, graph_steps)
    reset()
    t_cond = timed(lambda: only.run(steps, graph_steps))
    conds[0].explicit_solve(0.8)             # warm-up of the plain form
    reset()
    t_plain = timed(lambda: [c.explicit_solve(0.8) for c in conds for _ in range(steps)])
    # the same FP32 form one condition at a time (a batch of one member
    # each, replayed from its own graph): what batching alone buys
    singles = [rans.Batch([c], relaxation=0.8) for c in conds]
    for b in singles:
        b.run(graph_steps, graph_steps)
    reset()
    t_single = timed(lambda: [b.run(steps, graph_steps) for b in singles])
    for b in singles:
        b.close()
    cells_all = sum(m.n_real for m in conds + uneven)
    cells_cond = sum(m.n_real for m in conds)
    return {
        "members": len(conds) + len(uneven), "conditions": n_cond, "condition_mesh_cells": conds[0].n_real,
        "uneven_mesh_cells": [m.n_real for m in uneven], "tiles": {"condition_mesh": tile0, "uneven": tiles},
        "steps": steps, "graph_steps": graph_steps, "stages_per_step": 3,
        "batched_all_s": round(t_all, 4),
        "batched_all_Mcell_res_per_s": round(3 * steps * cells_all / t_all / 1e6, 1),
        "batched_64_conditions_s": round(t_cond, 4),
        "plain_64_sequential_s": round(t_plain, 4),
        "speedup_vs_plain_sequential": round(t_plain / t_cond, 2),
        "fp32_64_sequential_s": round(t_single, 4),
        "speedup_vs_fp32_sequential": round(t_single / t_cond, 2),
        "plain_Mcell_res_per_s": round(3 * steps * cells_cond / t_plain / 1e6, 1),
        "batched_64_Mcell_res_per_s": round(3 * steps * cells_cond / t_cond / 1e6, 1),
        "finite_final_residuals": bool(np.all(np.isfinite(norms))),
        "nonfinite_members": [int(i) for i in np.where(~np.isfinite(norms))[0]],
        "setup_s": round(t_setup, 1),
        "note": f"explicit three-stage pseudo-time steps (SA, CFL {SWEEP_CFL}, relaxation 0.8) from freestream; "
                "plain = FP64 fused residual "
                "+ FP64 update per condition with a norm per step, batched = one launch per stage over all "
                "members (FP32 residual, FP64 stage update), replayed from a hipGraph; fp32 sequential = the "
                "batched kernel on one condition at a time, each from its own graph; tiles autotuned on the "
                "batched step",
    }


def implicit_bench(rans, cells=2_000_000, cfls=(40.0, 10.0, 2.0), relaxation=0.9, rhs_iterations=5):
    """One implicit pseudo-time iteration of the reference loop
    (multigrid.h:262-268: fillRhoLHS, then implicitSolver::solve with the
    reference's rhs_iterations 5 and relaxation 0.9, solver.h:973-1213) at its
    start CFL 40 and at 10 and 2, on the bench's 2M-cell jittered mesh,
    inviscid with a slip wall, from the perturbed freestream state; then the
    SA model's step (5x5 blocks, no-slip wall, Re 1e6, DILU) at CFL 10 and 2.
    Per preconditioner and CFL: fill and solve wall times, GMRES iterations,
    time per GMRES iteration, and the returned residual norm (-1: the linear
    solve failed or the state went non-finite, as the reference's step does
    from this state at CFL 40)."""
    ni, nj = mesh_dims(cells)
    m = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25, seed=1)
    far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))
    out = {"cells": None, "relaxation": relaxation, "rhs_iterations": rhs_iterations,
           "gmres": "restart 30, tol 1e-2, maxit 500 (solver.h:906-910), device-resident Arnoldi state"}
    runs = [("", p, c) for p in ("dilu", "block-jacobi") for c in cfls] + [("sa_", "dilu", c) for c in (10.0, 2.0)]
    for model, precond, cfl in runs:
        sa = model == "sa_"
        if sa:
            s = rans.Solver(m, gas=rans.Gas(mu_L=freestream_mu(rans)), viscosity_model="spalart-allmaras",
                            second_order=True, gradient_scheme="green-gauss", limiter_k=5.0)
            s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("wall")})
        else:
            s = rans.Solver(m, viscosity_model="inviscid", second_order=True, gradient_scheme="green-gauss",
                            limiter_k=5.0)
            s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far),
                       "wall": rans.BoundaryCondition("slip-wall")})
        s.init()
        s.refill_bcs()
        q, nu = perturbed(s.get_q(), s.get_nu() if sa else None, s.n_real, far.get_conservative(s.gas),
                          s.gas.gamma, np.random.default_rng(2))
        out["cells"] = s.n_real
        s.set_preconditioner(precond)
        s.set_cfl(cfl)

        def reset():
            s.set_q(q)
            if sa:
                s.set_nu(nu)
        # untimed: one fill and a single linear solve (allocations, first launches)
        reset()
        s.fill()
        s.solve(relaxation, 0.0, 0)
        reset()
        s.synchronize()
        t0 = time.perf_counter()
        s.fill()
        s.synchronize()
        t_fill = time.perf_counter() - t0
        t0 = time.perf_counter()
        err = s.solve(relaxation, 0.0, rhs_iterations)
        s.synchronize()
        t_solve = time.perf_counter() - t0
        its = s.linear_iterations()
        out[f"{model}{precond}_cfl{cfl:g}"] = {"fill_ms": round(t_fill * 1e3, 2), "solve_ms": round(t_solve * 1e3, 2),
                                         "iteration_ms": round((t_fill + t_solve) * 1e3, 2),
                                         "gmres_iterations": int(its),
                                         "ms_per_gmres_iteration": round(t_solve * 1e3 / max(its, 1), 3),
                                         "residual_norm": err}
        s.close()
    m.close()
    return out


def sa_config1(rans, iterations=500):
    """The reference's plain FP64 case with the SA model: NACA 0012 O-grid of
    129 x 65 nodes (16 384 triangles), farfield at 20 chords, M 0.5, alpha
    2 deg (the reference's degree factor 0.01745), Re 1e6 (mu_L = 7e-7),
    no-slip wall, freestream start with rho nu~ = 3 mu_L, 500 iterations of
    the reference loop (multigrid.h:240-293: CFL 40 + 50 i up to 100,
    relaxation 0.9, rhs_iterations 5, GMRES 1e-2); then CFL 5 + i up to 20
    with relaxation 1 and rhs_iterations 0 (stalls), and with relaxation 0.5
    and rhs_iterations 9 to a residual of 1e-10 (the converged state the
    oracle's fixture holds).  Per run: loop state, iterations, the residual
    history's ends, CL/CD and the largest chi = nu~/nu_L every 10
    iterations, time per iteration."""
    out = {"mesh": "NACA 0012 O-grid 128 x 65 (16 384 triangles), r_far 20, first cell 2e-3",
           "flow": "M 0.5, alpha 2 x 0.01745 rad, Re 1e6 (mu_L 7e-7), no-slip wall, rho nu~ = 3 mu_L"}
    mesh = rans.Mesh.naca_ogrid(128, 65, r_far=20.0, first_h=2e-3)
    far = rans.BoundaryVariables(mach=0.5, angle=2 * 0.01745)
    mu = 7e-7
    for name, st in (("reference_loop", rans.Settings(max_iterations=iterations, tolerance=1e-13)),
                     ("cfl5_to_20", rans.Settings(start_cfl=5.0, slope_cfl=1.0, max_cfl=20.0, rhs_iterations=0,
                                                  max_iterations=iterations, tolerance=1e-13)),
                     ("cfl5_to_20_relaxation_0.5", rans.Settings(start_cfl=5.0, slope_cfl=1.0, max_cfl=20.0,
                                                                 relaxation=0.5, rhs_iterations=9,
                                                                 max_iterations=2 * iterations, tolerance=1e-10))):
        s = rans.Solver(mesh, gas=rans.Gas(mu_L=mu), viscosity_model="spalart-allmaras")
        s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("wall")})
        s.init()
        s.refill_bcs()
        chi, forces = [], []

        class Watch:
            updates = 0

            def calc_cp(self, solver, patch):
                self.updates += 1
                nu = solver.get_nu()[:solver.n_real]
                chi.append(float(f"{nu.max() / mu:.4g}"))
                cl, cd, _ = solver.get_wall_profile("wall")
                forces.append((float(f"{cl:.6g}"), float(f"{cd:.6g}")))

        s.synchronize()
        t0 = time.perf_counter()
        state, res = rans.run_solver(s, st, profile=Watch(), profile_every=10)
        s.synchronize()
        dt = time.perf_counter() - t0
        n = len(res)
        cl, cd, _ = s.get_wall_profile("wall")
        out[name] = {"state": state, "iterations": n, "seconds": round(dt, 2), "cl": cl, "cd": cd,
                     "ms_per_iteration": round(dt * 1e3 / max(n, 1), 2),
                     "residuals_first": [float(f"{r:.4g}") for r in res[:5]],
                     "residuals_last": [float(f"{r:.4g}") for r in res[-5:]],
                     "chi_max_every_10": chi, "cl_cd_every_10": forces}
        s.close()
    mesh.close()
    return out


def sa_adjoint_check(rans, cells=500_000, n_pairs=3):
    """State adjoint of the SA residual (the headline model) on a 500k-cell
    mesh of the same generator (Green-Gauss, second order, M 0.7 alpha 1.5
    deg Re 1e6 + 1% perturbations, no-slip wall): time against the FP64
    residual and the tangent, and the tangent/adjoint dot-product test over the
    five rows, |<w, Jv> - <J^T w, v>| / (|w| |Jv|)."""
    ni, nj = mesh_dims(cells)
    m = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25, seed=1)
    s = rans.Solver(m, gas=rans.Gas(mu_L=freestream_mu(rans)), viscosity_model="spalart-allmaras",
                    second_order=True, gradient_scheme="green-gauss", limiter_k=5.0)
    far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("wall")})
    s.init()
    s.refill_bcs()
    q, nu = perturbed(s.get_q(), s.get_nu(), s.n_real, far.get_conservative(s.gas), s.gas.gamma,
                      np.random.default_rng(2))
    s.set_q(q)
    s.set_nu(nu)
    t_primal = s.time_residual(rans.AFX_FP64, 50)
    t_adj = s.time_adjoint(20)
    t_tan = s.time_tangent(20)
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(4)

    def seed(nv):
        x = np.zeros(nv * nt)
        x[:nv * nr] = rng.standard_normal(nv * nr)
        return x

    worst = 0.0
    for _ in range(n_pairs):
        v, vn, w, wn = seed(4), seed(1), seed(4), seed(1)
        s.set_q(q)
        s.set_nu(nu)
        jv, jn = s.tangent(v, vn)
        s.set_q(q)
        s.set_nu(nu)
        lam, ln = s.adjoint(w, w_nu=wn)
        lhs, rhs = float(w @ jv) + float(wn @ jn), float(lam @ v) + float(ln @ vn)
        scale = float(np.sqrt(w @ w + wn @ wn) * np.sqrt(jv @ jv + jn @ jn))
        worst = max(worst, abs(lhs - rhs) / scale)
    s.close()
    m.close()
    return {"cells": nr, "model": "Spalart-Allmaras, Green-Gauss, second order, no-slip wall",
            "adjoint_ms": round(t_adj, 4), "primal_fp64_ms": round(t_primal, 4),
            "adjoint_over_primal": round(t_adj / t_primal, 2), "tangent_ms": round(t_tan, 4),
            "dot_product_max_rel": float(f"{worst:.3e}"), "dot_product_tests": n_pairs,
            "kernels": "5 passes: afx::k_sadj_forward, k_sadj_face, k_sadj_cell1, k_sadj_face2, k_sadj_cell2"}


def adjoint_check(rans, cells=500_000, n_dir=64, h=1e-7, node_fd_dirs=2):
    """State adjoint of the FP64 residual on a 500k-cell mesh of the same
    generator (inviscid, Green-Gauss, second order, M 0.7 alpha 1.5 deg + 1%
    perturbations): time against the FP64 residual, and central differences
    of the residual along 64 random unit directions (seed 3) against
    lambda . v for a random unit seed w."""
    ni, nj = mesh_dims(cells)
    m = rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25, seed=1)
    s = rans.Solver(m, viscosity_model="inviscid", second_order=True, gradient_scheme="green-gauss", limiter_k=5.0)
    far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
    s.init()
    s.refill_bcs()
    q, _ = perturbed(s.get_q(), None, s.n_real, far.get_conservative(s.gas), s.gas.gamma, np.random.default_rng(2))
    s.set_q(q)
    t_primal = s.time_residual(rans.AFX_FP64, 50)
    t_adj = s.time_adjoint(20)
    t_tan = s.time_tangent(20)
    rng = np.random.default_rng(3)

    def unit():
        x = np.zeros(4 * s.n_cells)
        x[:4 * s.n_real] = rng.standard_normal(4 * s.n_real)
        return x / np.linalg.norm(x)

    w = unit()
    s.set_q(q)
    lam = s.adjoint(w)
    scale = max(1.0, float(np.max(np.abs(q))))
    worst = 0.0
    for _ in range(n_dir):
        v = unit()
        s.set_q(q + h * scale * v)
        rp = s.fillRhoRHS(rans.AFX_FP64)[0]
        s.set_q(q - h * scale * v)
        rm = s.fillRhoRHS(rans.AFX_FP64)[0]
        fd = float(w @ (rp - rm)) / (2 * h * scale)
        ad = float(lam @ v)
        worst = max(worst, abs(ad - fd) / max(abs(fd), 1e-3 * float(np.linalg.norm(lam))))
    # tangent/adjoint dot-product test on the same state: <w, J v> = <J^T w, v>
    dots = []
    for _ in range(4):
        v = unit()
        s.set_q(q)
        jv = s.tangent(v)
        dots.append(abs(float(w @ jv) - float(lam @ v)) / (float(np.linalg.norm(w)) * float(np.linalg.norm(jv))))
    out = {"cells": s.n_real, "model": "inviscid, Green-Gauss, second order (active-branch limiter)",
           "adjoint_ms": round(t_adj, 4), "primal_fp64_ms": round(t_primal, 4),
           "adjoint_over_primal": round(t_adj / t_primal, 2), "fd_directions": n_dir, "fd_step": h,
           "fd_max_rel_err": float(f"{worst:.3e}"),
           "kernels": "5 passes: afx::k_adj_forward, k_adj_face, k_adj_cell1, k_adj_face2, k_adj_cell2",
           "tangent_ms": round(t_tan, 4), "tangent_over_primal": round(t_tan / t_primal, 2),
           "tangent_kernel": "afx::k_residual_tile<afx::Dual1, 4> (forward mode of the FP64 tile residual)",
           "dot_product_max_rel": float(f"{max(dots):.3e}"), "dot_product_tests": len(dots)}
    out["node_adjoint"] = node_adjoint_check(rans, m, s, q, w, n_dir=node_fd_dirs)
    out["dcl_dalpha"] = dcl_dalpha_check(rans, s, q)
    out["dcl_dnodes"] = dcl_dnodes_check(rans, m, s, q)
    return out


def node_adjoint_check(rans, m, s, q, w, n_dir=2, h=1e-6):
    """(dR/dX)^T w over the node coordinates on the same mesh: wall time of one
    evaluation (with its state adjoint and the host copy of xbar), and central
    differences of w . R along random node displacements scaled by the local
    cell size, each on a mesh rebuilt from the moved nodes."""
    s.set_q(q)
    s.adjoint_nodes(w)
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        s.set_q(q)
        _, xb = s.adjoint_nodes(w)
    t_nodes = (time.perf_counter() - t0) / reps
    c = m.counts()
    nr = c["real_cells"]
    x, y = m.get("nodesX"), m.get("nodesY")
    cells = m.get("cellsNodes").reshape(-1, 4)[:nr].astype(np.int32)
    nverts = np.where(m.get("cellsIsTriangle")[:nr] != 0, 3, 4).astype(np.int32)
    en = m.get("edgesNodes").reshape(-1, 2)
    bsegs = en[m.get("boundaryEdges")].astype(np.int32)
    bkeys = m.get("boundaryEdgesIntTag").astype(np.int32)
    names = dict(zip((int(k) for k in bkeys), m.boundary_names()))
    far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))

    def residual(dxy):
        mm = rans.Mesh.from_elements(x + dxy[:, 0], y + dxy[:, 1], cells, nverts, bsegs, bkeys, names)
        ss = rans.Solver(mm, viscosity_model="inviscid", second_order=True, gradient_scheme="green-gauss",
                         limiter_k=5.0)
        ss.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
        ss.init()
        ss.set_q(q)
        return ss.fillRhoRHS(rans.AFX_FP64)[0]

    # each node moves by h times its local length (sqrt of the smallest
    # incident cell area): a uniform relative perturbation, so the 1e-4 wall
    # cells stay in the linear range of their limiter branches
    area = m.get("cellsAreas")[:nr]
    scale = np.full(len(x), np.inf)
    for k in range(4):
        idx = np.nonzero(nverts > k)[0]
        np.minimum.at(scale, cells[idx, k], np.sqrt(area[idx]))
    rng = np.random.default_rng(5)
    worst = 0.0
    t0 = time.perf_counter()
    for i in range(n_dir):
        V = rng.standard_normal(xb.shape) * scale[:, None]
        fd = float(w @ (residual(h * V) - residual(-h * V))) / (2 * h)
        ad = float(np.sum(xb * V))
        worst = max(worst, abs(ad - fd) / max(abs(fd), 1e-3 * float(np.linalg.norm(xb) * np.linalg.norm(V))))
        if n_dir > 2:   # a long check: progress on stderr
            print(f"node FD direction {i + 1}/{n_dir}: worst relative error so far {worst:.3e} "
                  f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    return {"nodes": int(c["nodes"]), "wall_ms_per_eval": round(t_nodes * 1e3, 3), "fd_directions": n_dir,
            "fd_step": h, "fd_max_rel_err": float(f"{worst:.3e}"),
            "kernels": "metric adjoints in passes 2-4, then afx::k_geo_face, k_geo_cell, k_geo_node"}


def dcl_dalpha_check(rans, s, q, steps=300, every=25, h=1e-6):
    """dCL/d(angle) through `steps` explicit steps (CFL 1) from the perturbed
    state, checkpoints every `every` steps, against central differences of
    CL after the same steps at angle +- h."""
    s.set_cfl(1.0)
    angle = float(np.deg2rad(ALPHA_DEG))
    s.set_q(q)
    t0 = time.perf_counter()
    cl, d = s.dcl_dalpha(steps=steps, checkpoint_every=every, patch="wall")
    t_rev = time.perf_counter() - t0

    def cl_at(a):
        far = rans.BoundaryVariables(mach=MACH, angle=a)
        s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
        s.set_q(q)
        s.refill_bcs()
        for _ in range(steps):
            s.explicit_solve(1.0)
        return s.get_wall_profile("wall")[0]

    t0 = time.perf_counter()
    cp = cl_at(angle + h)
    t_fwd = time.perf_counter() - t0
    fd = (cp - cl_at(angle - h)) / (2 * h)
    return {"steps": steps, "checkpoint_every": every, "cl": cl, "dcl_dangle": d, "fd_dcl_dangle": fd,
            "fd_step": h, "rel_err": float(f"{abs(d - fd) / abs(fd):.3e}"),
            "wall_s_forward_and_reverse": round(t_rev, 3), "wall_s_forward_only": round(t_fwd, 3)}


def dcl_dnodes_check(rans, m, s, q, steps=300, every=25):
    """dCL/dX through `steps` explicit steps (CFL 1, checkpoints every `every`
    steps) on the adjoint section's mesh: wall time of the second-order form,
    and central differences of CL after the same steps on meshes rebuilt from
    nodes moved along a random direction scaled by the local cell size (nodes
    of wall faces within 1e-6 of the chord extremes, where the extent is not
    differentiable at this scale, stay).  The first-order form is checked over
    all `steps`; the second-order form over 20 steps, because over longer
    horizons the limiter's min / max switches of this M 0.7 state lie within
    any usable difference step (on 20k cells after 100 steps: 11% at h = 1e-6,
    2e-4 at h = 1e-8), while the adjoint differentiates the active branches."""
    far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))
    c = m.counts()
    nr = c["real_cells"]
    x, y = m.get("nodesX"), m.get("nodesY")
    cells = m.get("cellsNodes").reshape(-1, 4)[:nr].astype(np.int32)
    nverts = np.where(m.get("cellsIsTriangle")[:nr] != 0, 3, 4).astype(np.int32)
    en = m.get("edgesNodes").reshape(-1, 2)
    bsegs = en[m.get("boundaryEdges")].astype(np.int32)
    bkeys = m.get("boundaryEdgesIntTag").astype(np.int32)
    bnames = np.array(m.boundary_names())
    names = dict(zip((int(k) for k in bkeys), bnames))

    def solver(mm, so):
        ss = rans.Solver(mm, viscosity_model="inviscid", second_order=so, gradient_scheme="green-gauss",
                         limiter_k=5.0)
        ss.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
        ss.init()
        ss.set_cfl(1.0)
        ss.set_q(q)
        ss.refill_bcs()
        return ss

    def cl_moved(dxy, so, n):
        ss = solver(rans.Mesh.from_elements(x + dxy[:, 0], y + dxy[:, 1], cells, nverts, bsegs, bkeys, names), so)
        for _ in range(n):
            ss.explicit_solve(1.0)
        return ss.get_wall_profile("wall")[0]

    area = m.get("cellsAreas")[:nr]
    scale = np.full(len(x), np.inf)
    for k in range(4):
        idx = np.nonzero(nverts > k)[0]
        np.minimum.at(scale, cells[idx, k], np.sqrt(area[idx]))
    on_wall = bnames == "wall"
    xc = m.get("edgesCentersX")[m.get("boundaryEdges")[on_wall]]
    near = np.array([(0 < abs(v - xc.max()) < 1e-6) or (0 < abs(v - xc.min()) < 1e-6) for v in xc])
    frozen = np.unique(bsegs[on_wall][near | (((xc == xc.max()) | (xc == xc.min())) & near.any())].ravel())
    rng = np.random.default_rng(6)
    V = rng.standard_normal((len(x), 2)) * scale[:, None]
    V[frozen] = 0.0

    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
    s.set_cfl(1.0)
    s.set_q(q)
    s.refill_bcs()
    t0 = time.perf_counter()
    cl, xb = s.dcl_dnodes(steps=steps, checkpoint_every=every, patch="wall")
    t_rev = time.perf_counter() - t0
    out = {"steps": steps, "checkpoint_every": every, "second_order": True, "cl": cl,
           "wall_s_forward_and_reverse": round(t_rev, 3), "checks": []}
    for so, n, h in ((False, steps, 1e-6), (True, 20, 1e-8)):
        sv = s if so else solver(m, False)
        sv.set_q(q)
        sv.refill_bcs()
        c0, xbn = sv.dcl_dnodes(steps=n, checkpoint_every=every, patch="wall")
        ad = float(np.sum(xbn * V))
        fd = (cl_moved(h * V, so, n) - cl_moved(-h * V, so, n)) / (2 * h)
        out["checks"].append({"second_order": so, "steps": n, "fd_step": h, "dcl_dV": ad, "fd_dcl_dV": fd,
                              "rel_err": float(f"{abs(ad - fd) / abs(fd):.3e}"),
                              "fd_rounding_bound_rel": float(f"{50 * 2.2e-16 * abs(c0) / (2 * h) / abs(fd):.1e}")})
    return out


def aeroelastic_bench(n_chord=64, n_span=256, n_cases=32, alpha_max=8.0, cpu=True):
    """Vortex-lattice / beam coupling: 64 x 256 rings (N = 16384; AR 10, taper
    0.4, quarter-chord sweep 25 deg, cosine spacing) on a 256-element beam
    (EI, GJ linear from root to tip), n_cases load cases alpha 0..alpha_max to
    a converged tip deflection.  Timings are wall clock around synchronous
    calls; the factor runs on the alpha = 0 matrix (timed warm, the first
    call reported apart) and preconditions the
    FP64 refinement of every later matrix (refactored when refinement
    contracts worse than 0.5).  Each load case starts from the previous one's
    displacement (an alpha sweep) and relaxes with Aitken's factor.  Parity
    against the FP64 oracle (exact dense solves) at 16 x 64 rings, one load
    case at alpha_max, in the CPU leg."""
    from aeroflex_amd import vlm as V
    xyz, rn, so, ref = V.wing_lattice(n_chord, n_span)
    N = n_chord * n_span
    h = V.VLM(xyz, rn, so)
    h.set_flow(0.0, sref=ref["sref"], cref=ref["cref"])
    h.build()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        h.build()
    t_build = (time.perf_counter() - t0) / reps
    # the first factorisation also pays one-time costs of the process
    # (rocBLAS loads its GEMM code objects on first use, the first
    # cooperative launch is set up): reported apart from the warm figure
    t0 = time.perf_counter()
    h.factor()
    t_factor_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    h.factor()
    t_factor = time.perf_counter() - t0
    h.set_flow(alpha_max, sref=ref["sref"], cref=ref["cref"])
    h.build()
    t0 = time.perf_counter()
    _, sinfo = h.solve()
    t_solve = time.perf_counter() - t0
    beam = V.Beam(*V.wing_beam(xyz, so, n_chord))
    ae = V.Aeroelastic(h, beam)
    alphas = np.linspace(0.0, alpha_max, n_cases)
    cases = []
    t0 = time.perf_counter()
    for a in alphas:
        h.set_flow(float(a), sref=ref["sref"], cref=ref["cref"])
        info = ae.solve(tol=1e-8, max_iter=60, warm_start=1)
        cases.append(info)
    t_cases = time.perf_counter() - t0
    last = cases[-1]
    # release the coupling, beam and lattice (device matrices, streams and
    # rocBLAS handles) before the CPU leg, dependents first
    ae.close()
    beam.close()
    h.close()
    out = {
        "workload": f"VLM {n_chord}x{n_span} rings (N={N}), AR 10, taper 0.4, sweep 25 deg, cosine spacing; "
                    f"{n_span}-element Euler-Bernoulli beam; {n_cases} load cases alpha 0..{alpha_max} deg",
        "aic_build_ms": round(t_build * 1e3, 3),
        "aic_entries_per_s": round(N * N / t_build / 1e9, 2),
        "lu_factor_ms": round(t_factor * 1e3, 2),
        "lu_factor_first_call_ms": round(t_factor_first * 1e3, 2),
        "lu_tflops": round(2.0 / 3.0 * N ** 3 / t_factor / 1e12, 2),
        "lu_dtype": "f32 (hand-written MFMA GEMM for the trailing update) + f64 refinement",
        "solve_ms": round(t_solve * 1e3, 3),
        "solve_refinement_steps": sinfo.iterations,
        "solve_backward_error": sinfo.backward_error,
        "coupling_total_s": round(t_cases, 3),
        "coupling_ms_per_case": round(t_cases * 1e3 / n_cases, 2),
        "coupling_iterations": [c.iterations for c in cases],
        "refactors": int(sum(c.refactors for c in cases)),
        "refine_steps_total": int(sum(c.refine_steps for c in cases)),
        "all_converged": all(bool(c.converged) for c in cases),
        "tip_deflection_at_alpha_max": last.tip_deflection,
        "cl_at_alpha_max": last.cl,
        "semispan": ref["span"] / 2,
    }
    if cpu:
        # CPU leg: the oracle (one core) as baseline and as the checker of the
        # FP32-LU + FP64-refinement coupling against exact FP64 solves, on a
        # 16 x 64 lattice (N = 1024), one load case at alpha_max; build and
        # dense solve times scaled to N = 16384 by N^2 and N^3
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc_vlm as ov
        nc2, ns2 = 16, 64
        x2, r2, s2, ref2 = V.wing_lattice(nc2, ns2)
        b2 = V.wing_beam(x2, s2, nc2)
        with V.VLM(x2, r2, s2) as h2, V.Beam(*b2) as bm2:
            h2.set_flow(alpha_max, sref=ref2["sref"], cref=ref2["cref"])
            with V.Aeroelastic(h2, bm2) as ae2:
                i2 = ae2.solve(tol=1e-11, max_iter=80)
        t0 = time.perf_counter()
        o2 = ov.aeroelastic(x2, r2, s2, alpha_max, b2, sref=ref2["sref"], cref=ref2["cref"], tol=1e-11,
                            max_iter=80)
        t_oracle = time.perf_counter() - t0
        out["vs_fp64_oracle"] = {"lattice": f"{nc2}x{ns2}", "alpha": alpha_max,
                                 "tip_rel_err": abs(i2.tip_deflection - o2["tip"]) / abs(o2["tip"]),
                                 "cl_rel_err": abs(i2.cl - o2["cl"]) / abs(o2["cl"])}
        A, rhs = ov.build(x2, r2, s2, 2.0, wake_length=100 * ref2["cref"])
        t0 = time.perf_counter()
        A, rhs = ov.build(x2, r2, s2, 2.0, wake_length=100 * ref2["cref"])
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        ov.solve(A, rhs)
        ts = time.perf_counter() - t0
        n2 = nc2 * ns2
        out["cpu_baseline"] = {"kind": "port", "cores": 1, "sample": f"oracle build + dense solve at N={n2}",
                               "aic_build_ms_at_N": round(tb * (N / n2) ** 2 * 1e3, 1),
                               "dense_solve_ms_at_N": round(ts * (N / n2) ** 3 * 1e3, 1),
                               "coupled_case_s_at_sample": round(t_oracle, 3)}
    return out


def multigrid_setup_bench(rans, coarse_cells=500_000, fine_cells=2_000_000, cpu=True, cpu_rows=1000):
    """Full-multigrid set-up at scale: multigrid::gen_mapper (multigrid.h:96-163)
    between a 500k-cell and a 2M-cell jittered mesh of the benchmark generator
    (device build, wall clock around the synchronous call), and one
    prolongation q_fine = M q_coarse.  CPU leg: the oracle's restatement
    (all N_fine x N_coarse pairs, one core) on the first cpu_rows fine cells,
    scaled to all of them."""
    def mesh(n, seed):
        ni, nj = mesh_dims(n)
        return rans.Mesh.jittered(ni, nj, r_far=20.0, thickness=0.12, first_h=1e-4, jitter=0.1, quad_frac=0.25,
                                  seed=seed)

    def solver(m):
        s = rans.Solver(m, viscosity_model="inviscid")
        far = rans.BoundaryVariables(mach=MACH, angle=np.deg2rad(ALPHA_DEG))
        s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far),
                   "wall": rans.BoundaryCondition("slip-wall")})
        s.init()
        return s

    mc, mf = mesh(coarse_cells, 1), mesh(fine_cells, 5)
    sc, sf = solver(mc), solver(mf)
    rans.Mapper(sc, sf)    # first build: module load and allocator warm-up
    t0 = time.perf_counter()
    mp = rans.Mapper(sc, sf)
    t_build = time.perf_counter() - t0
    rp, cols, w = mp.csr()
    mp.prolongate()
    t0 = time.perf_counter()
    for _ in range(10):
        mp.prolongate()
    t_prol = (time.perf_counter() - t0) / 10
    out = {"coarse_cells": sc.n_cells, "fine_cells": sf.n_cells, "nnz": int(len(w)),
           "empty_rows": int(np.sum(np.diff(rp) == 0)), "gen_mapper_ms": round(t_build * 1e3, 2),
           "prolongate_ms": round(t_prol * 1e3, 3)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc
        fx, fy = mf.get("cellsCentersX"), mf.get("cellsCentersY")
        cx, cy, ca = mc.get("cellsCentersX"), mc.get("cellsCentersY"), mc.get("cellsAreas")
        t0 = time.perf_counter()
        ref = orc.gen_mapper(fx[:cpu_rows], fy[:cpu_rows], cx, cy, ca)
        t_cpu = time.perf_counter() - t0
        k = int(rp[cpu_rows])
        out["vs_oracle_rows"] = {"rows": cpu_rows, "cols_equal": bool(np.array_equal(cols[:k], ref.indices)),
                                 "max_weight_diff": float(np.max(np.abs(w[:k] - ref.data))) if k else 0.0}
        out["cpu_baseline"] = {"kind": "port", "cores": 1,
                               "sample": f"oracle gen_mapper on {cpu_rows} fine rows x all coarse cells, "
                                         f"{t_cpu:.1f} s, scaled by fine rows",
                               "gen_mapper_s": round(t_cpu * sf.n_cells / cpu_rows, 1)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--tile", type=int, default=TILE_CELLS, help="cells per tile of the fused kernel")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run D2D copies of this size (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--only-fp32", action="store_true", help="skip per-kernel and FP64 timings (profiling runs)")
    ap.add_argument("--model", default="spalart-allmaras", choices=["spalart-allmaras", "inviscid"])
    ap.add_argument("--sweep-steps", type=int, default=50,
                    help="pseudo-time steps of the batched-sweep comparison (0: skip it)")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the adjoint timing and finite-difference check")
    ap.add_argument("--no-aeroelastic", dest="aeroelastic", action="store_false",
                    help="skip the vortex-lattice / beam coupling section")
    ap.add_argument("--only-aeroelastic", action="store_true",
                    help="run only the vortex-lattice / beam coupling section (profiling runs)")
    ap.add_argument("--no-multigrid", dest="multigrid", action="store_false",
                    help="skip the multigrid set-up section (level map at 500k -> 2M cells)")
    ap.add_argument("--only-multigrid", action="store_true", help="run only the multigrid set-up section")
    ap.add_argument("--no-implicit", dest="implicit", action="store_false",
                    help="skip the implicit-iteration timing")
    ap.add_argument("--only-implicit", action="store_true", help="run only the implicit-iteration timing")
    ap.add_argument("--no-sa-config1", dest="sa_config1", action="store_false",
                    help="skip the SA run of the 129 x 65 O-grid case in the implicit section (profiling runs)")
    ap.add_argument("--node-fd-dirs", type=int, default=64,
                    help="random node displacements of the node adjoint's central-difference check (each rebuilds "
                         "the 500k-cell mesh and solver twice)")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run only the adjoint section (state, node coordinates, dCL/dalpha)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the residual's multi-GPU sharding is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    if args.only_aeroelastic:
        print(json.dumps({"aeroelastic": aeroelastic_bench(cpu=not args.no_cpu)}), flush=True)
        return
    if args.only_multigrid:
        from aeroflex_amd import rans
        print(json.dumps({"multigrid_setup": multigrid_setup_bench(rans, cpu=not args.no_cpu)}), flush=True)
        return
    if args.only_adjoint:
        from aeroflex_amd import rans
        adj = adjoint_check(rans, node_fd_dirs=args.node_fd_dirs)
        adj["spalart_allmaras"] = sa_adjoint_check(rans)
        print(json.dumps({"adjoint": adj}), flush=True)
        return
    if args.only_implicit:
        from aeroflex_amd import rans
        imp = implicit_bench(rans)
        if args.sa_config1:
            imp["sa_config1"] = sa_config1(rans)
        print(json.dumps({"implicit": imp}), flush=True)
        return
    from aeroflex_amd import rans
    ev = HipEvents()
    if args.copy_probe_mib:
        ev.copy_probe(args.copy_probe_mib)
    m, s, info = build_case(rans, args.cells, model=args.model, tile_cells=args.tile)
    V = 5 if args.model == "spalart-allmaras" else 4
    nc, nf, nb = s.n_real, s.n_faces, s.n_bfaces
    stream = s.stream()
    FP32, FP64 = rans.AFX_FP32, rans.AFX_FP64

    # warmup (also converts the FP64 state to the FP32 copy)
    s.launch_stage(FP32, 3, max(args.warmup, 1))
    s.synchronize()
    ev.device_sync()

    # timed region: exactly K residual evaluations
    a, b = ev.event(), ev.event()
    s.synchronize()
    t0 = time.perf_counter()
    ev.record(a, stream)
    s.launch_stage(FP32, 3, args.steps)
    ev.record(b, stream)
    s.synchronize()
    ev.device_sync()
    wall = time.perf_counter() - t0
    ev_ms = ev.elapsed_ms(a, b)
    ms_per_step = wall * 1e3 / args.steps
    value = nc * args.steps / wall / 1e6
    res_ms_timed = ev_ms / args.steps

    if args.only_fp32:
        print(json.dumps({"value": round(value, 2), "ms_per_step": round(ms_per_step, 5), "cells": nc}), flush=True)
        return

    # the dominant (and only) kernel of the FP32 residual: its average launch
    # duration from events around back-to-back launches on the library's stream
    nk = max(50, min(args.steps, 500))
    res32_ms = time_stage(ev, s, stream, FP32, 3, nk)
    b32 = alg_bytes(nc, nf, nb, 4, V=V)
    achieved = b32 / (res32_ms * 1e-3) / 1e9
    # three-kernel FP32 path (gradients+limiters, face fluxes, cell sums; four variables only)
    unf = None
    if V == 4:
        unf = {"grad_lim": time_stage(ev, s, stream, FP32, 0, nk // 5),
               "face_flux": time_stage(ev, s, stream, FP32, 1, nk // 5),
               "gather": time_stage(ev, s, stream, FP32, 2, nk // 5)}

    # plain FP64 form on the same mesh and state (fused kernel, reference
    # rounding), on its own tile size: its LDS per cell is twice the FP32
    # form's, so smaller tiles keep more of them resident
    s64, _ = case_solver(rans, m, args.model, tile_cells=FP64_TILE_CELLS)
    s64.launch_stage(FP64, 3, 5)
    s64.synchronize()
    t64 = time_stage(ev, s64, s64.stream(), FP64, 3, args.fp64_steps)
    s64.close()
    b64 = alg_bytes(nc, nf, nb, 8, V=V)
    acc32 = fp32_accuracy(s, nc, V)
    forces32 = fp32_forces(rans, s) if V == 5 else None
    post = vtu_bench(rans, s, args.model, cpu=not args.no_cpu)

    prof = "residual_fp32_sa_2M_traffic.txt" if V == 5 else "residual_fp32_2M_traffic.txt"
    traffic = read_traffic(os.path.join(ROOT, "profiles", prof))
    out = {
        "metric": "Mcell-residuals/s + % HBM roofline; FP32/batched vs plain FP64; adjoint/primal",
        "value": round(value, 2),
        "unit": "Mcell-residuals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("FP32 SA-RANS residual evaluation (GG gradients + limiter of 5 variables, Roe + "
                                "viscous + nu~ fluxes, SA sources, cell sums)" if V == 5 else
                                "FP32 Euler residual evaluation (GG gradients + limiter, Roe fluxes, cell sums)")
                               + " on a jittered mixed tri/quad mesh around NACA 0012, M0.7 a1.5 deg Re 1e6 + 1% "
                                 "seeded perturbations",
                   "cells": nc, "faces": nf, "boundary_faces": nb, "ogrid": [info["ni"], info["nj"]],
                   "variables": V, "model": args.model, "mu_L": info["mu_L"], "tile_cells": args.tile,
                   "seeds": {"mesh": 1, "state": 2}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"afx::k_residual_tile_f32<{2 if V == 5 else 0}, true, true>",
                     "alg_bytes_per_launch": b32,
                     "kernel_ms": round(res32_ms, 5), "timed_region_ms_per_step": round(ev_ms / args.steps, 5)},
        "fp64_plain": {"value": round(nc / (t64 * 1e-3) / 1e6, 2), "kernel_ms": round(t64, 5),
                       "tile_cells": FP64_TILE_CELLS,
                       "kernel": f"afx::k_residual_tile<double, {V}>", "alg_bytes_per_launch": b64,
                       "roofline_frac": round(b64 / (t64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "fp32_vs_fp64": acc32,
        "fp32_vs_fp64_forces": forces32,
        "fp32_three_kernel_ms": None if unf is None else {k: round(v, 5) for k, v in unf.items()},
    }
    if args.sweep_steps > 0 or args.adjoint or args.implicit:
        s.close()
        m.close()
        del s, m
    if args.adjoint:
        out["adjoint"] = adjoint_check(rans, node_fd_dirs=args.node_fd_dirs)
        out["adjoint"]["spalart_allmaras"] = sa_adjoint_check(rans)
    if args.implicit:
        out["implicit"] = implicit_bench(rans)
        if args.sa_config1:
            out["implicit"]["sa_config1"] = sa_config1(rans)
    if args.sweep_steps > 0:
        out["batched_sweep"] = batched_sweep(rans, args.sweep_steps)
    if args.aeroelastic:
        out["aeroelastic"] = aeroelastic_bench(cpu=not args.no_cpu)
    if args.multigrid:
        out["multigrid_setup"] = multigrid_setup_bench(rans, cpu=not args.no_cpu)
    out["post"] = post
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(rans, args.model)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    try:
        main()
    finally:
        # every device handle is released here, before interpreter shutdown
        if "aeroflex_amd.rans" in sys.modules:
            sys.modules["aeroflex_amd.rans"].close_all()

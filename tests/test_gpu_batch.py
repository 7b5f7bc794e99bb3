"""Explicit pseudo-time stepping: the five-variable (SA) explicit step against
the oracle's SA residual driven through the reference's three-stage scheme
(solver.h:802-828), and the batched FP32/FP64 form against single-member
runs (bit for bit), eager launches (bit for bit on every graph replay) and
the FP64 step (FP32 rounding)."""
import numpy as np
import pytest

import orc
from cases import QUAD, TRI, perturbed_state, rel_max

pytestmark = pytest.mark.gpu


def sa_solver(path, mach=0.5, angle=2 * 0.01745, mu_L=5e-4, cfl=0.5, base=None, tile_cells=0):
    from aeroflex_amd import rans
    gas = rans.Gas(mu_L=mu_L)
    if base is None:
        s = rans.Solver(rans.Mesh.read_file(path), gas=gas, viscosity_model="spalart-allmaras",
                        tile_cells=tile_cells)
    else:
        s = base.like(gas)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", rans.BoundaryVariables(mach, angle)),
               "wall": rans.BoundaryCondition("wall")})
    s.init()
    s.refill_bcs()
    q = perturbed_state(s.n_real, s.n_cells, s.get_q(), orc.freestream(mach, angle), seed=3)
    s.set_q(q)
    nu = s.get_nu()
    rng = np.random.default_rng(9)
    nu[:s.n_real] *= 1 + 0.01 * rng.standard_normal(s.n_real)
    s.set_nu(nu)
    s.set_cfl(cfl)
    return s


def test_sa_explicit_step_matches_oracle():
    s = sa_solver(TRI)
    q0, nu0 = s.get_q(), s.get_nu()
    e = s.explicit_solve(0.8)
    om = orc.OMesh.read(TRI)
    o = orc.OSolver(om, mu_L=5e-4, viscous=2)
    o.set_bcs({"farfield": ("farfield", (0.5, 2 * 0.01745, 1.0, 1.0)), "wall": ("wall", (0.2, 0.0, 1.0, 1.0))})
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=0.5)
    area = om.get("cellsAreas")
    nr = o.n_real
    q, nu = q0.copy(), nu0.copy()
    o.vec("q", q)
    o.sa_vec("nu", nu)
    dt = o.sa_calc_dt()                # local time step of the stage-0 state
    for ai in (0.25, 0.5, 1.0):
        o.vec("q", q)
        o.sa_vec("nu", nu)
        R, T = o.sa_fill_rhs()
        qs, nus = o.vec("q"), o.sa_vec("nu")   # wall ghosts refreshed by the residual
        q, nu = qs.copy(), nus.copy()
        for k in range(4):
            q[k:4 * nr:4] = q0[k:4 * nr:4] + R[k:4 * nr:4] / area[:nr] * dt[:nr] * ai * 0.8
        nu[:nr] = nu0[:nr] + T[:nr] / area[:nr] * dt[:nr] * ai * 0.8
    assert rel_max(s.get_q()[:4 * nr], q[:4 * nr]) <= 1e-12
    assert rel_max(s.get_nu()[:nr], nu[:nr]) <= 1e-12
    qW = np.concatenate([R[:4 * nr] / np.repeat(area[:nr], 4), T[:nr] / area[:nr]])
    assert e == pytest.approx(np.linalg.norm(qW), rel=1e-10)


def test_sa_time_step_matches_oracle():
    s = sa_solver(QUAD)
    dt = s.calc_dt()
    o = orc.OSolver(orc.OMesh.read(QUAD), mu_L=5e-4, viscous=2)
    o.set_bcs({"farfield": ("farfield", (0.5, 0.02, 1.0, 1.0)), "wall": ("wall", (0.2, 0.0, 1.0, 1.0))})
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=0.5)
    o.vec("q", s.get_q())
    o.sa_vec("nu", s.get_nu())
    nr = s.n_real
    assert rel_max(dt[:nr], o.sa_calc_dt()[:nr]) <= 1e-14


def members(tile_cells=0):
    a = sa_solver(QUAD, 0.5, 0.02, tile_cells=tile_cells)
    b = sa_solver(None, 0.6, 0.05, mu_L=2e-4, base=a)
    c = sa_solver(TRI, 0.4, -0.03, mu_L=1e-4)
    return [a, b, c]


def test_batch_equals_single_member_runs():
    from aeroflex_amd import rans
    together = members()
    rans.Batch(together, relaxation=0.8).run(6, graph_steps=3)
    alone = members()
    for m in alone:
        rans.Batch([m], relaxation=0.8).run(6, graph_steps=3)
    for x, y in zip(together, alone):
        assert np.array_equal(x.get_q(), y.get_q())
        assert np.array_equal(x.get_nu(), y.get_nu())
    assert not np.array_equal(together[0].get_q(), members()[0].get_q())   # the state moved


def test_batch_graph_replays_match_eager_launches():
    from aeroflex_amd import rans
    g, e = members(), members()
    bg, be = rans.Batch(g, relaxation=0.8), rans.Batch(e, relaxation=0.8)
    for rep in range(3):
        bg.run(4, graph_steps=2)          # two replays of a 2-step graph
        be.run(4, graph_steps=0)
        ng, ne = bg.residual_norms(rans.AFX_FP64), be.residual_norms(rans.AFX_FP64)   # eager work in between
        assert np.array_equal(ng, ne), rep
        for x, y in zip(g, e):
            assert np.array_equal(x.get_q(), y.get_q()), rep


def test_batch_fp32_tracks_fp64_explicit():
    from aeroflex_amd import rans
    b32 = members()
    rans.Batch(b32, relaxation=0.8).run(5, graph_steps=0)
    for m, ref in zip(b32, members()):
        for _ in range(5):
            ref.explicit_solve(0.8)
        nr = ref.n_real
        dq = m.get_q()[:4 * nr] - ref.get_q()[:4 * nr]
        assert np.max(np.abs(dq)) <= 1e-5 * np.max(np.abs(ref.get_q()[:4 * nr]))
        assert np.max(np.abs(m.get_nu()[:nr] - ref.get_nu()[:nr])) <= 1e-4 * np.max(np.abs(ref.get_nu()[:nr]))


def test_autotune_returns_fastest_candidate():
    from aeroflex_amd import rans
    best, times = rans.autotune_tile(lambda t: sa_solver(QUAD, tile_cells=t), candidates=(64, 128, 4096))
    assert set(times) <= {64, 128} and best in times and times[best] == min(times.values())

"""Host-side parts of the product (no GPU): the mesh layer against the oracle,
the benchmark mesh generators, MSH round trips, RCM ordering, tile layout
statistics, and the C ABI's exported symbols and error codes."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import orc
from cases import elements_of
from cases import QUAD, TRI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ["nodesX", "nodesY", "edgesNodes", "edgesCells", "edgesLengths", "edgesNormalsX", "edgesNormalsY",
          "edgesCentersX", "edgesCentersY", "cellsNodes", "cellsEdges", "cellsAreas", "cellsCentersX",
          "cellsCentersY", "boundaryEdges", "boundaryEdgesIntTag", "cellsIsTriangle"]


def rans():
    from aeroflex_amd import rans as r
    return r


def assert_same_mesh(a, o):
    assert a.counts() == o.counts()
    for f in FIELDS:
        x, y = a.get(f), o.get(f)
        assert x.dtype == y.dtype and x.shape == y.shape, f
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{f} differs"
    assert a.boundary_names() == o.boundary_names()


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_reader_bitexact_vs_oracle(path):
    assert_same_mesh(rans().Mesh.read_file(path), orc.OMesh.read(path))


@pytest.mark.parametrize("kind", ["ogrid", "jittered"])
def test_generated_mesh_bitexact_vs_oracle(kind):
    r = rans()
    m = r.Mesh.naca_ogrid(128, 65) if kind == "ogrid" else r.Mesh.jittered(96, 40, seed=3)
    x, y, cells, nv, bsegs, bkeys = elements_of(m)
    names = {1001: "farfield", 1002: "wall", 2003: "internal"}
    o = orc.OMesh.build(x, y, cells, nv, bsegs, bkeys, names)
    assert_same_mesh(m, o)


def test_naca_ogrid_counts():
    """129 x 65 nodes with the seam merged: 128 x 64 quads split in two."""
    c = rans().Mesh.naca_ogrid(128, 65, r_far=20.0).counts()
    assert c == {"nodes": 8320, "real_cells": 16384, "cells": 16640, "faces": 24704, "bfaces": 256}


def test_jittered_mesh_is_valid():
    m = rans().Mesh.jittered(128, 60, jitter=0.1, quad_frac=0.25, seed=1)
    c = m.counts()
    nr = c["real_cells"]
    A = m.get("cellsAreas")[:nr]
    assert np.all(A > 0)
    tri = m.get("cellsIsTriangle")[:nr]
    nq = 128 * 59
    # quads kept with probability 1/4
    assert abs((nr - tri.sum()) / nq - 0.25) < 0.03
    assert c["nodes"] - c["faces"] + nr == 0
    # same seed, same mesh; other seed, other mesh
    m2 = rans().Mesh.jittered(128, 60, jitter=0.1, quad_frac=0.25, seed=1)
    assert np.array_equal(m.get("nodesX"), m2.get("nodesX"))
    m3 = rans().Mesh.jittered(128, 60, jitter=0.1, quad_frac=0.25, seed=2)
    assert not np.array_equal(m.get("nodesX"), m3.get("nodesX"))


def test_jittered_triangles_are_delaunay_pairs():
    """Each split quad's diagonal satisfies the empty-circle test: the fourth
    vertex is not inside the circumcircle of either triangle."""
    m = rans().Mesh.jittered(64, 30, jitter=0.1, quad_frac=0.0, seed=7)
    X, Y = m.get("nodesX"), m.get("nodesY")
    nr = m.counts()["real_cells"]
    cn = m.get("cellsNodes").reshape(-1, 4)[:nr]
    bad = 0
    for k in range(0, nr, 2):
        t1, t2 = cn[k, :3], cn[k + 1, :3]
        d = list(set(t2) - set(t1))[0]
        a, b, c = t1
        M = np.array([[X[v] - X[d], Y[v] - Y[d], (X[v] - X[d]) ** 2 + (Y[v] - Y[d]) ** 2] for v in (a, b, c)])
        orient = (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a])
        if np.sign(orient) * np.linalg.det(M) > 1e-12 * np.abs(M).max() ** 2:
            bad += 1
    assert bad == 0


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_msh_round_trip(path, tmp_path):
    r = rans()
    a = r.Mesh.read_file(path)
    out = tmp_path / "rt.msh"
    a.write_msh(out)
    b = r.Mesh.read_file(out)
    assert_same_mesh(b, orc.OMesh.read(str(out)))
    for f in FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f


def test_generated_mesh_round_trip_through_oracle_reader(tmp_path):
    m = rans().Mesh.jittered(80, 30, seed=4)
    out = tmp_path / "j.msh"
    m.write_msh(out)
    assert_same_mesh(m, orc.OMesh.read(str(out)))


def bandwidth(m, perm):
    nr = m.counts()["real_cells"]
    ec = m.get("edgesCells").reshape(-1, 2).astype(np.int64)
    ec = ec[(ec[:, 1] < nr)]
    inv = np.empty(nr, np.int64)
    inv[perm] = np.arange(nr)
    return int(np.max(np.abs(inv[ec[:, 0]] - inv[ec[:, 1]])))


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_rcm_is_a_valid_bandwidth_reducing_order(path):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    m = rans().Mesh.read_file(path)
    nr = m.counts()["real_cells"]
    p = m.rcm()
    assert np.array_equal(np.sort(p), np.arange(nr))
    ec = m.get("edgesCells").reshape(-1, 2).astype(np.int64)
    ec = ec[ec[:, 1] < nr]
    A = coo_matrix((np.ones(len(ec)), (ec[:, 0], ec[:, 1])), shape=(nr, nr)).tocsr()
    A = A + A.T
    ps = reverse_cuthill_mckee(A, symmetric_mode=True)
    ident = np.arange(nr)
    assert bandwidth(m, p) <= 1.25 * bandwidth(m, ps)
    assert bandwidth(m, p) <= bandwidth(m, ident) or bandwidth(m, ident) < 2 * bandwidth(m, ps)


@pytest.mark.parametrize("tile", [16, 64, 192, 512])
def test_tile_layout_statistics(tile):
    m = rans().Mesh.read_file(TRI)
    st = m.tile_stats(tile)
    nr = m.counts()["real_cells"]
    # tiles hold about `tile` cells: pockets left by the BFS growth are merged
    # into a neighbour, so no tile count beyond ceil(nr / tile) by more than
    # a few percent, and the average tile is within 10% of the target
    assert (nr + tile - 1) // tile <= st["tiles"] * 1.1 + 1
    assert 0.9 * tile <= nr / st["tiles"] <= 1.15 * tile or st["tiles"] <= 2
    assert st["gmax"] <= st["lmax"] < 4096
    # every real cell computes its gradient in its own tile; halos add more
    assert nr <= st["sum_grad"] <= st["sum_local"]
    assert st["ftmax"] <= st["fmax"]


def header_symbols():
    txt = ""
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt += open(os.path.join(ROOT, "include", h)).read()
    return sorted(set(re.findall(r"\b(afx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "aeroflex_amd", "libafx_rans.so")
    lib = C.CDLL(lib_path)
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (afx_[a-z0-9_]+)", out))
    assert set(syms) <= exported


def test_mesh_errors_are_codes_not_exceptions(tmp_path):
    r = rans()
    h = C.c_void_p()
    rc = r.lib.afx_mesh_read_msh(str(tmp_path / "missing.msh").encode(), C.byref(h))
    assert rc == -3 and b"cannot open" in r.lib.afx_last_error()
    bad = tmp_path / "bad.msh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    assert r.lib.afx_mesh_read_msh(str(bad).encode(), C.byref(h)) == -3
    with pytest.raises(r.RansError):
        r.Mesh.naca_ogrid(4, 2)
    # boundary segment that is not an edge of any cell (mesh.h:757)
    x = np.array([0.0, 1.0, 0.0, 5.0])
    y = np.array([0.0, 0.0, 1.0, 5.0])
    with pytest.raises(r.RansError, match="invalid edge ref"):
        r.Mesh.from_elements(x, y, [[0, 1, 2, 0]], [3], [[0, 3]], [1001], {1001: "wall"})


def test_python_mirror_rejects_unknown_names():
    r = rans()
    with pytest.raises(r.RansError):
        r.BoundaryCondition("inlet-outlet")
    with pytest.raises(r.RansError):
        r.Solver.__init__(object.__new__(r.Solver), None, viscosity_model="spalart")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_jittered_mesh_has_no_folded_cells(seed):
    """Every corner cross product of every cell has the same sign (no
    inverted triangle, no non-convex quad), also in the skewed trailing-edge
    cells of a 1e-4 first layer."""
    m = rans().Mesh.jittered(256, 120, first_h=1e-4, jitter=0.1, quad_frac=0.3, seed=seed)
    nr = m.counts()["real_cells"]
    cn = m.get("cellsNodes").reshape(-1, 4)[:nr].astype(np.int64)
    tri = m.get("cellsIsTriangle")[:nr] == 1
    x, y = m.get("nodesX"), m.get("nodesY")
    nv = np.where(tri, 3, 4)
    signs = []
    for k in range(4):
        a = cn[np.arange(nr), k % nv]
        b = cn[np.arange(nr), (k + 1) % nv]
        c = cn[np.arange(nr), (k + 2) % nv]
        signs.append(np.sign((x[b] - x[a]) * (y[c] - y[a]) - (y[b] - y[a]) * (x[c] - x[a])))
    signs = np.array(signs)
    assert np.all(signs == signs[0, 0])

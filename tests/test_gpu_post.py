"""Post-processing from the device state against the oracle's restatements:
the solution file of save (post.h:58-180, afx_rans_save_vtu), the same bytes,
and the wall Cp profile of CpProfile::calc_cp (post.h:246-290)."""
import numpy as np
import pytest

import orc
from cases import QUAD, TRI, make_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wall", ["slip-wall", "wall"])
@pytest.mark.parametrize("path", [QUAD, TRI])
def test_vtu_matches_oracle_bytes(path, wall, tmp_path):
    print(f"[post] vtu {path} {wall}", flush=True)
    s, o, q = make_pair(path, wall=wall)
    o.wall_dist()
    s.save_vtu(tmp_path / "gpu.vtu")
    o.save_vtu(tmp_path / "orc.vtu")
    a, b = (tmp_path / "gpu.vtu").read_bytes(), (tmp_path / "orc.vtu").read_bytes()
    assert len(a) == len(b) and a == b


def test_vtu_matches_oracle_bytes_ogrid(tmp_path):
    """A generated 128 x 65 O-grid (8192 quads), long enough for the writer to
    split every section over several threads."""
    import tempfile
    from aeroflex_amd import rans
    m = rans.Mesh.naca_ogrid(128, 65, r_far=20.0, first_h=2e-3)
    with tempfile.TemporaryDirectory() as d:
        p = f"{d}/og.msh"
        m.write_msh(p)
        om = orc.OMesh.read(p)
    s, o, q = make_pair(mesh=m, omesh=om)
    o.wall_dist()
    s.save_vtu(tmp_path / "gpu.vtu")
    o.save_vtu(tmp_path / "orc.vtu")
    assert (tmp_path / "gpu.vtu").read_bytes() == (tmp_path / "orc.vtu").read_bytes()


def test_vtu_rejects_unwritable_path(tmp_path):
    from aeroflex_amd import rans
    s, _, _ = make_pair(TRI)
    with pytest.raises(rans.RansError):
        s.save_vtu(tmp_path / "no_such_dir" / "x.vtu")


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_cp_profile_matches_oracle(path):
    s, o, q = make_pair(path)
    x, y, cp = s.cp_profile("wall")
    xo, yo, cpo = o.cp_profile("wall")
    assert np.array_equal(x, xo) and np.array_equal(y, yo)
    assert np.max(np.abs(cp - cpo)) <= 1e-15 * max(1.0, np.max(np.abs(cpo)))


def test_cp_profile_rejects_unknown_patch():
    from aeroflex_amd import rans
    s, _, _ = make_pair(TRI)
    with pytest.raises(rans.RansError):
        s.cp_profile("no-such-patch")

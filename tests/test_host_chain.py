"""Host side of the RANS -> polar -> vortex-lattice chain (app.cpp:168-184),
no GPU: the alpha list of Rans::compute_alphas (rans.h:53-57), the
database::airfoil record, and the polar table the nonlinear vortex lattice
reads (database::table, common/database/database.cpp:192-242) built from
computed airfoil polars."""
import numpy as np
import pytest


def test_compute_alphas_follows_the_reference_loop():
    from aeroflex_amd import rans
    # the reference's Settings defaults (core.h:190-192): 1, 7, 3
    assert rans.compute_alphas(1.0, 7.0, 3.0) == [1.0, 4.0, 7.0]
    assert rans.compute_alphas(-2.0, 7.0, 3.0) == [-2.0, 1.0, 4.0, 7.0]
    assert rans.compute_alphas(2.0, 1.0, 1.0) == []
    assert rans.compute_alphas(3.0, 3.0, 1.0) == [3.0]
    # accumulated in double as `for (double i = start; i <= end; i += step)`:
    # ten additions of 0.1 stop just below 1, so 1.0 itself is not reached
    # exactly but 0.9999999999999999 is still <= 1
    a = rans.compute_alphas(0.0, 1.0, 0.1)
    assert len(a) == 11 and a[3] == 0.30000000000000004 and a[-1] == 0.9999999999999999
    # an end just below the accumulated value drops it
    assert len(rans.compute_alphas(0.0, 0.9999999999999998, 0.1)) == 10


def test_airfoil_record():
    from aeroflex_amd import rans
    db = rans.Airfoil([1.0, 4.0])
    assert db.alpha == [1.0, 4.0] and db.cl == [] and db.cd == [] and db.cmy == []
    assert rans.DEG_TO_RAD == 0.01745   # rans.h:94, the reference's degree factor


def polar_pair():
    from aeroflex_amd import rans
    a = rans.Airfoil([-2.0, 1.0, 4.0])
    a.cl, a.cd, a.cmy = [-0.2, 0.1, 0.4], [0.011, 0.010, 0.012], [0.0, -0.01, -0.02]
    b = rans.Airfoil([0.0, 5.0])
    b.cl, b.cd, b.cmy = [0.0, 0.5], [0.02, 0.03], [0.1, 0.2]
    return a, b


def test_polar_table_from_airfoils():
    from aeroflex_amd import vlm as V
    a, b = polar_pair()
    p = V.polar_from_airfoils({"root": a, "tip": b}, [("root", 0.0), ("tip", 0.6), ("root", 1.0)])
    assert p["offsets"].tolist() == [0, 3, 5]
    assert p["alpha_deg"].tolist() == [-2.0, 1.0, 4.0, 0.0, 5.0]
    assert p["cl"].tolist() == [-0.2, 0.1, 0.4, 0.0, 0.5]
    assert p["cd"].tolist() == [0.011, 0.010, 0.012, 0.02, 0.03]
    assert p["cmy"].tolist() == [0.0, -0.01, -0.02, 0.1, 0.2]
    assert p["section_airfoil"].tolist() == [0, 1, 0]
    assert p["section_span"].tolist() == [0.0, 0.6, 1.0]
    assert p["offsets"].dtype == np.int32 and p["section_airfoil"].dtype == np.int32


def test_polar_table_rejects_incomplete_or_unordered_polars():
    from aeroflex_amd import vlm as V
    a, b = polar_pair()
    a.cd = a.cd[:2]
    with pytest.raises(V.VlmError):
        V.polar_from_airfoils({"root": a}, [("root", 0.0)])
    a, b = polar_pair()
    b.alpha = [5.0, 0.0]
    with pytest.raises(V.VlmError):
        V.polar_from_airfoils({"tip": b}, [("tip", 0.0)])
    from aeroflex_amd import rans
    with pytest.raises(V.VlmError):
        V.polar_from_airfoils({"empty": rans.Airfoil([])}, [("empty", 0.0)])

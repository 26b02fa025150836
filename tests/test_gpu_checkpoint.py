"""Binary checkpoint and bitwise resume (tbl_checkpoint_save / _load): N
steps straight equal k steps + save + a fresh context + load + N - k steps,
bit for bit, in every field and in the statistics, recycling, wall-shear
window and clock.  The run crosses a statistics rescale, the start of the
recycled inflow and of the KE guard (freqTimAvg = 3), with the wall-shear
window on and, in one case, the extrapolated pressure guess (which keeps a
second PCorr)."""
import numpy as np
import pytest

from tests.case_util import C1_GRID, RAGGED, gpu_from_case, make_case

pytestmark = pytest.mark.gpu


def _state(sv):
    Ut, P, PC = sv.download()
    st = sv.get_state()
    return dict(Ut=Ut, P=P, PC=PC, **{k: st[k] for k in ("sums", "avg", "Wt", "newk", "Ufluc_rec", "bltkns",
                                                         "deltastar", "theta", "lamda", "lamda03", "time",
                                                         "nstep", "cnt")})


@pytest.mark.parametrize("precision,grid,extra", [("fp64", RAGGED, {}), ("mixed", C1_GRID, {}),
                                                  ("fp64", C1_GRID, dict(pc_extrap=1))])
def test_resume_from_checkpoint_is_bitwise(tmp_path, precision, grid, extra):
    g, s = make_case(grid, freq=3, nstep0=1, seed=7)
    s.cnt = 1
    kw = dict(kmax=30, errtol=1e-10, wall_shear_window=2, **extra)
    if precision == "mixed":
        from turbulentbl_amd import tbl
        kw.update(tol_mode=tbl.TBL_TOL_TRUE, inner_tol=1e-4)
    n, k = 6, 3
    with gpu_from_case(g, s, precision=precision, **kw) as sv:
        d_all = sv.step(n)
        ref = _state(sv)
        tref = sv.wall_shear_mean()
    path = str(tmp_path / "tbl.ckpt")
    with gpu_from_case(g, s, precision=precision, **kw) as sv:
        d1 = sv.step(k)
        sv.checkpoint_save(path)
    from turbulentbl_amd import tbl
    with tbl.Solver(g.nX, g.nY, g.nZ, i_rec=g.i_rec, planes=g.planes, freq_tim_avg=g.freq_tim_avg,
                    precision=precision, **kw) as sv:      # a fresh context: nothing uploaded
        sv.checkpoint_load(path)
        d2 = sv.step(n - k)
        got = _state(sv)
        tgot = sv.wall_shear_mean()
    for key in ref:
        a, b = ref[key], got[key]
        assert np.array_equal(np.asarray(a), np.asarray(b)), key
    for x, y in zip(tref, tgot):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    for a, b in zip(d_all, d1 + d2):
        for key in ("div", "ke", "gm_iters", "gm_true_res", "gs_sweeps", "mass", "rescaled", "cf_ready", "nstep"):
            assert a[key] == b[key], key
    assert [d["rescaled"] for d in d_all] == [0, 1, 0, 0, 1, 0]


def test_checkpoint_rejects_another_grid(tmp_path):
    from turbulentbl_amd.tbl import TblError
    g, s = make_case(RAGGED)
    path = str(tmp_path / "a.ckpt")
    with gpu_from_case(g, s) as sv:
        sv.checkpoint_save(path)
    g2, s2 = make_case((21, 12, 17))
    with gpu_from_case(g2, s2) as sv:
        with pytest.raises(TblError) as e:
            sv.checkpoint_load(path)
        assert e.value.code == -1

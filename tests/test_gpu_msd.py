"""selm_msd (GPU) against the oracle's restatement of the reference's MSD
script, on periodic random-walk trajectories whose steps cross the box
seams; and the script-compatible msd() on a DCD."""
import numpy as np
import pytest

import host_lib as H
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("F, n, nlags, mult", [(200, 37, 41, 1), (500, 300, 21, 7), (12, 3, 15, 1)])
def test_msd_gpu_matches_oracle(gpu_lib, F, n, nlags, mult):
    from monomers2026_amd import analysis
    rng = np.random.default_rng(F + n)
    L = 6.0
    steps = rng.normal(0.0, 0.3, (F, n, 3))
    x = np.mod(np.cumsum(steps, axis=0), L)
    fr = np.ascontiguousarray(x.transpose(0, 2, 1)).astype(np.float32)
    g = analysis.msd_frames(fr, L, nlags, mult)
    o = O.msd(fr, L, nlags, mult)
    assert np.array_equal(np.isnan(g), np.isnan(o))
    ok = ~np.isnan(o)
    err = np.abs(g[ok] - o[ok]).max() / max(np.abs(o[ok]).max(), 1e-300)
    print("msd gpu vs oracle F=%d n=%d: %.2e" % (F, n, err), flush=True)
    assert err < 1e-12


def test_msd_script_on_a_dcd(gpu_lib, tmp_path, monkeypatch):
    from monomers2026_amd import analysis
    rng = np.random.default_rng(5)
    L, n, F = 6.0, 4, 101
    x = np.mod(np.cumsum(rng.normal(0.0, 0.2, (F, n, 3)), axis=0), L)
    path = str(tmp_path / "walk.dcd")
    H.write_dcd(path, [(k * 4000, x[k]) for k in range(F)], 4000, 0.005, (L, L, L))
    monkeypatch.chdir(tmp_path)
    t, m = analysis.msd(0.005, 11, 1, 1, path)
    fr = np.ascontiguousarray(x.transpose(0, 2, 1)).astype(np.float32)
    o = O.msd(fr, L, 11, 1)
    assert t[1] == 1 * 0.005 * 4000 and len(m) == 11
    np.testing.assert_allclose(m[1:], o[1:], rtol=1e-12)
    rows = open(tmp_path / "walk.txt").read().split("\n")
    assert rows[0] == "0 0" and len([r for r in rows if r]) == 11

"""Mean-square displacement with the reference script's semantics
(py/Diffusion Coefficient/MSD.py:15-69): the FP64-summing oracle against
trajectories whose MSD is known exactly, and the DCD reader against the
reference's trajectory layout."""
import json
import os

import numpy as np

import host_lib as H
import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _traj(F, n, L, pos):
    fr = np.zeros((F, 3, n), dtype=np.float32)
    for f in range(F):
        fr[f] = pos(f).T
    return fr


def test_ballistic_msd_is_v2_t2():
    """Straight lines x(t) = x0 + v t (no boundary crossed): MSD(lag) =
    |v|^2 lag^2 for every origin, so the average is that too."""
    rng = np.random.default_rng(0)
    n, F, L = 7, 60, 1000.0
    x0 = rng.uniform(100, 200, (n, 3))
    v = rng.uniform(-0.5, 0.5, (n, 3))
    fr = _traj(F, n, L, lambda f: x0 + v * f)
    m = O.msd(fr, L, 20)
    lags = np.arange(20)
    # the script's per-bead average of |v_a|^2 lag^2
    want = (v ** 2).sum(axis=1).mean() * lags ** 2
    np.testing.assert_allclose(m, want, rtol=2e-5, atol=1e-9)


def test_origins_start_at_one_fifth_and_normalisation():
    """Frames 0..9 of one bead whose x is the frame index squared: origins are
    frames 2..9 (floor(10/5) = 2), so MSD(1) = mean over n in [2, 9) of
    ((n+1)^2 - n^2)^2 = mean of (2n+1)^2."""
    F = 10
    fr = _traj(F, 1, 1e6, lambda f: np.array([[float(f * f), 0.0, 0.0]]))
    m = O.msd(fr, 1e6, 4)
    for lag in (1, 2, 3):
        ns = np.arange(2, F - lag)
        want = np.mean(((ns + lag) ** 2 - ns ** 2) ** 2.0)
        assert abs(m[lag] - want) <= 1e-12 * want
    assert m[0] == 0.0
    # no origin left for a lag of F - 2 frames: the script would divide by zero
    assert np.isnan(O.msd(fr, 1e6, 9)[8])


def test_minimum_image_with_the_x_half_box_in_every_dimension():
    """A bead hops +9 in x across a 10-wide box: the image displacement is -1.
    A hop of +3 in y stays +3 even when the y edge is only 4: the script uses
    half the x edge (5) as the threshold in every dimension."""
    fr = np.zeros((10, 3, 1), dtype=np.float32)
    fr[:, 0, 0] = [0.5, 9.5] * 5
    fr[:, 1, 0] = [0.0, 3.0] * 5
    m = O.msd(fr, 10.0, 2)
    assert abs(m[1] - (1.0 + 9.0)) < 1e-6


def test_read_dcd_matches_the_reference_header(tmp_path):
    """The reader parses a DCD written with the reference trajectory's layout:
    control words, cell, and float32 coordinates."""
    from monomers2026_amd import analysis
    g = json.load(open(os.path.join(GOLD, "dcd_1base_head.json")))
    frame0 = bytes.fromhex(g["frame0_hex"])
    n = g["natoms"]
    xyz = np.stack([np.frombuffer(frame0, "<f4", n, 56 + k * (8 + 4 * n) + 4) for k in range(3)], axis=1)
    path = str(tmp_path / "r.dcd")
    H.write_dcd(path, [(k * 4000, xyz.astype(np.float64) + k) for k in range(5)], 4000, 0.005, (6.0, 6.0, 6.0))
    d = analysis.read_dcd(path)
    assert d["natoms"] == n and d["nsavc"] == 4000 and abs(d["delta"] - 0.005) < 1e-9
    assert d["frames"].shape == (5, 3, n)
    np.testing.assert_array_equal(d["frames"][0].T, xyz)
    assert tuple(d["cells"][0]) == (6.0, 0.0, 6.0, 0.0, 0.0, 6.0)


DCD_1BASE = "/root/reference/data/cases/Diffusion Coefficient/1base.dcd"


def _msd_float32_sequential(frames, box_x, nlags):
    """The script's loop with NumPy >= 2 scalar promotion (NEP 50): a Python
    int 0 plus a float32 term stays float32, so msd[d][lag] is a float32 sum
    taken sequentially over atoms (outer) and origins (inner); the terms are
    formed in float32 (difference, the +-2X shift, the square) and MSD(lag) =
    (m_x + m_y + m_z) / (Fd * natoms) in float32.  np.cumsum adds in order."""
    F, _, n = frames.shape
    X = np.float32(np.float32(box_x) / np.float32(2))
    start = F // 5
    fr = frames[start:]
    Fn = fr.shape[0]
    out = np.zeros(nlags, np.float64)
    for lag in range(1, nlags):
        Fd = Fn - lag
        m = []
        for d in range(3):
            diff = (fr[lag:, d, :] - fr[:Fd, d, :]).astype(np.float32)  # [origin][atom]
            t = np.where(diff >= X, (diff - np.float32(2) * X) ** 2,
                         np.where(diff <= -X, (diff + np.float32(2) * X) ** 2, diff ** 2)).astype(np.float32)
            m.append(np.cumsum(t.T.ravel(), dtype=np.float32)[-1])  # atoms outer, origins inner
        out[lag] = np.float32(np.float32(m[0] + m[1]) + m[2]) / np.float32(Fd * n)
    return out


def test_msd_accumulation_precision_on_the_reference_trajectory():
    """The reference's trajectory (cases/Diffusion Coefficient/1base.dcd:
    5001 frames of 4 beads): the oracle's MSD (float32 terms, FP64 sums, the
    script's result under NumPy 1.x, where 0 + float32 promotes to float64)
    against the script's float32 accumulation under NumPy >= 2, for the
    script's 40 lags.  The difference is the float32 rounding of ~16,000
    sequential additions per lag; it is printed and bounded here."""
    import pytest
    if not os.path.exists(DCD_1BASE):
        pytest.skip("the reference trajectory is not present")
    _, _, nat, frames = H.read_dcd(DCD_1BASE)
    fr = np.stack([xyz.T for _, xyz in frames]).astype(np.float32)  # [F][3][natoms]
    box = frames[0][0][0]
    m64 = O.msd(fr, box, 41)
    m32 = _msd_float32_sequential(fr, box, 41)
    rel = np.abs(m32[1:] - m64[1:]) / m64[1:]
    print("1base.dcd MSD, float32 sequential (NumPy >= 2) vs FP64 sums: max rel diff %.2e at lag %d; "
          "MSD(40) %.6g vs %.6g" % (rel.max(), 1 + int(rel.argmax()), m32[40], m64[40]))
    assert nat == 4 and fr.shape[0] == 5001
    assert rel.max() < 1e-4

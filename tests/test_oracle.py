"""Pins the FP64 CPU oracle against mathematics and independent libraries.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, an index off by one, a transposed operand) fails it:
  * Peskin 4-point factor: the kernel's moment conditions (sum 1, first
    moment 0, sum of squares 3/8) and the reference's r >= 2 branch quirk.
  * gamma/lambda: interpolation of constant and linear fields on the
    staggered grid, and adjointness <Gamma u, F> = -<u, Lambda F>.
  * FFTs: numpy's pocketfft (rfftn/irfftn, including irfftn's handling of a
    non-Hermitian x-DC plane), Parseval.
  * Philox4x32-10: the published Random123 known-answer vectors.
  * the whole step: discrete divergence and mean of u^{n+1}, the closed-form
    first step of a free bead, momentum bookkeeping, and the analytic
    harmonic-bond force.
"""
import numpy as np
import pytest

import oracle_lib as O


# --------------------------------------------------------------------------- Peskin
def _stencil(r):
    p = np.floor(r) - 1
    pts = p + np.arange(4)
    return pts, np.array([O.phi(abs(r - q)) for q in pts])


@pytest.mark.parametrize("r", [0.1, 0.37, 0.5, 0.93, 3.25, 17.999])
def test_peskin_moments(r):
    pts, w = _stencil(r)
    assert abs(w.sum() - 1.0) < 1e-15
    assert abs(((r - pts) * w).sum()) < 1e-14
    assert abs((w ** 2).sum() - 3.0 / 8.0) < 1e-15


def test_peskin_integer_quirk():
    # grid coordinate exactly integer: distances 1, 0, 1, 2 -> (1/4, 1/2, 1/4, 1)
    pts, w = _stencil(16.0)
    assert np.allclose(w, [0.25, 0.5, 0.25, 1.0], rtol=0, atol=1e-16)
    assert O.phi(2.0, 1) == 1.0 and O.phi(2.0, 0) == 0.0
    assert O.phi(2.5, 1) == 1.0


def test_peskin_closed_form_values():
    # phi(r) = (3 - 2r + sqrt(1 + 4r - 4r^2)) / 8 on [0,1): phi(0)=1/2, phi(1/2)=(2+sqrt2)/8
    assert O.phi(0.0) == 0.5
    assert abs(O.phi(0.5) - (2 + np.sqrt(2)) / 8) < 1e-16
    assert abs(O.phi(1.5) - (2 - np.sqrt(2)) / 8) < 1e-16
    assert abs(O.phi(1.0) - 0.25) < 1e-16


# --------------------------------------------------------------------------- gamma / lambda
def test_gamma_constant_and_linear():
    nz, ny, nx = 9, 10, 11
    u = np.full((nz, ny, nx), 2.5)
    assert abs(O.gamma(3.3, 4.6, 2.2, u) - 2.5) < 1e-14
    # linear in x (away from the periodic seam): Gamma reproduces u(r)
    xs = np.arange(nx, dtype=float)
    u = np.broadcast_to(0.5 + 0.25 * xs, (nz, ny, nx)).copy()
    for rx in (3.3, 4.5, 6.71):
        assert abs(O.gamma(rx, 4.2, 5.9, u) - (0.5 + 0.25 * rx)) < 1e-14
    # linear in z
    zs = np.arange(nz, dtype=float)
    u = np.broadcast_to((1.0 - 0.1 * zs)[:, None, None], (nz, ny, nx)).copy()
    assert abs(O.gamma(5.5, 4.4, 3.7, u) - (1.0 - 0.1 * 3.7)) < 1e-14


def test_gamma_lambda_adjoint():
    rng = np.random.default_rng(0)
    shape = (7, 8, 9)
    for _ in range(5):
        u = rng.standard_normal(shape)
        r = rng.uniform(0, 7, 3)
        F = rng.standard_normal()
        g = O.gamma(r[0], r[1], r[2], u)
        uf = np.zeros(shape)
        O.lam(r[0], r[1], r[2], F, uf)
        assert abs(g * F + (u * uf).sum()) < 1e-12


def test_lambda_brute_force_small_grid():
    # spread on a 5x6x7 grid against a direct triple loop with numpy weights
    shape = (5, 6, 7)
    r = (2.3, 4.8, 0.4)
    F = 1.7
    uf = np.zeros(shape)
    O.lam(r[0], r[1], r[2], F, uf)
    ref = np.zeros(shape)
    nz, ny, nx = shape
    px, py, pz = (int(np.floor(c)) - 1 for c in r)
    for k in range(4):
        for j in range(4):
            for i in range(4):
                w = O.phi(abs(r[2] - (pz + k))) * O.phi(abs(r[1] - (py + j))) * O.phi(abs(r[0] - (px + i)))
                ref[(pz + k) % nz, (py + j) % ny, (px + i) % nx] -= F * w
    assert np.abs(uf - ref).max() < 1e-15
    assert abs(uf.sum() + F) < 1e-14


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 15, 16, 24, 30, 32, 48, 64, 96, 128])
def test_fft_1d(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = np.abs(np.fft.fft(a)).max()
    assert np.abs(O.c2c1(a, -1) - np.fft.fft(a)).max() < 1e-14 * scale * np.log2(n + 1)
    assert np.abs(O.c2c1(a, +1) - n * np.fft.ifft(a)).max() < 1e-14 * scale * np.log2(n + 1)


@pytest.mark.parametrize("shape", [(4, 6, 8), (5, 6, 10), (8, 8, 8), (6, 5, 9), (10, 12, 30)])
def test_fft_3d_vs_pocketfft(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    h = O.r2c3(a)
    ref = np.fft.rfftn(a)
    assert np.abs(h - ref).max() < 1e-13 * np.abs(ref).max()
    back = O.c2r3(h, shape[2])
    assert np.abs(back - a * a.size).max() < 1e-12 * a.size
    # Parseval
    full = np.fft.fftn(a)
    assert abs((np.abs(full) ** 2).sum() / a.size - (a ** 2).sum()) < 1e-10 * (a ** 2).sum()


@pytest.mark.parametrize("shape", [(6, 4, 8), (5, 6, 9)])
def test_c2r_non_hermitian_semantics(shape):
    """c2r of a half spectrum that is NOT Hermitian in the x-DC / Nyquist
    planes (what the reference's thermal-noise generator feeds it) equals
    pocketfft's irfftn (complex inverse over z, y; real inverse over x that
    keeps only the real part of DC and Nyquist), times N."""
    rng = np.random.default_rng(3)
    nz, ny, nx = shape
    X = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    out = O.c2r3(X, nx)
    ref = np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * (nx * ny * nz)
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
    # equivalently: the DC plane enters as (X(j,k) + conj X(-j,-k)) / 2
    Xs = X.copy()
    jm = (-np.arange(ny)) % ny
    km = (-np.arange(nz)) % nz
    planes = [0] + ([nx // 2] if nx % 2 == 0 else [])
    for i in planes:
        P = X[:, :, i]
        Xs[:, :, i] = 0.5 * (P + np.conj(P[km][:, jm]))
    full = np.zeros((nz, ny, nx), complex)
    full[:, :, : nx // 2 + 1] = Xs
    for i in range(1, (nx + 1) // 2):
        full[:, :, nx - i] = np.conj(Xs[km][:, jm][:, :, i])
    ref2 = np.fft.ifftn(full).real * (nx * ny * nz)
    assert np.abs(out - ref2).max() < 1e-12 * np.abs(ref2).max()


# --------------------------------------------------------------------------- RNG
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normals_moments():
    z = np.array([O.normal_pair(42, 0, 1, i, 0) for i in range(40000)]).ravel()
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.02
    assert abs((z ** 4).mean() - 3.0) < 0.1
    # pure function of its counter
    assert np.array_equal(O.normal_pair(42, 0, 1, 7, 2), O.normal_pair(42, 0, 1, 7, 2))
    assert not np.array_equal(O.normal_pair(42, 0, 1, 7, 2), O.normal_pair(42, 1, 1, 7, 2))


def test_mode_normals():
    """The six normals of a thermal-forcing mode: each of the six is a
    standard normal on its own (moments over 30000 modes), the six are
    uncorrelated, and the pairs come from the Philox words they are assigned
    to: with key (0, 0) and counter (0, 0, 0, 0) Random123's known answer is
    [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8], so component 0's pair
    has radius sqrt(-2 ln((w0 + 1) 2^-32)) and angle 2 pi (w1 - 2^31) 2^-32,
    and component 1's the same from (w2, w3)."""
    z = np.array([O.mode_normals(42, 0, 3, i) for i in range(30000)])
    for q in range(6):
        assert abs(z[:, q].mean()) < 0.03, q
        assert abs(z[:, q].var() - 1.0) < 0.04, q
        assert abs((z[:, q] ** 4).mean() - 3.0) < 0.25, q
    c = np.corrcoef(z.T)
    assert np.abs(c - np.eye(6)).max() < 0.03
    # tails: P(|z| > 3) = 0.0027 for a standard normal
    assert abs((np.abs(z) > 3).mean() - 0.0027) < 0.001
    w = [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    z0 = O.mode_normals(0, 0, 0, 0)
    for d in range(2):
        wa, wb = w[2 * d], w[2 * d + 1]
        rad = np.sqrt(-2.0 * np.log((wa + 1) * 2.0 ** -32))
        th = 2 * np.pi * (wb - 2 ** 31) * 2.0 ** -32
        assert abs(z0[2 * d] - rad * np.cos(th)) < 1e-14 and abs(z0[2 * d + 1] - rad * np.sin(th)) < 1e-14, d
    # the third component comes from the second block (counter slot 1)
    w1 = O.philox([0, 0, 1, 0], [0, 0])
    rad = np.sqrt(-2.0 * np.log((w1[0] + 1) * 2.0 ** -32))
    th = 2 * np.pi * (w1[1] - 2 ** 31) * 2.0 ** -32
    assert abs(z0[4] - rad * np.cos(th)) < 1e-14 and abs(z0[5] - rad * np.sin(th)) < 1e-14
    assert np.array_equal(O.mode_normals(42, 0, 1, 7), O.mode_normals(42, 0, 1, 7))
    assert not np.array_equal(O.mode_normals(42, 0, 1, 7), O.mode_normals(42, 0, 2, 7))


# --------------------------------------------------------------------------- whole step
def _div(u):
    ux, uy, uz = u
    return (np.roll(ux, -1, 2) - ux) + (np.roll(uy, -1, 1) - uy) + (np.roll(uz, -1, 0) - uz)


def _system(dims, noise=False, seed=5, mu=0.8, rho=1.1, dt=0.02):
    L = float(dims[0])
    return O.Oracle(*dims, (0.0, 0.0, 0.0), (L, float(dims[1]), float(dims[2])), rho=rho, mu=mu, kB=1.0,
                    T=1.5, dt=dt, seed=seed, noise=noise)


@pytest.mark.parametrize("noise", [False, True])
def test_step_projection_invariants(noise):
    rng = np.random.default_rng(1)
    dims = (12, 10, 8)
    o = _system(dims, noise)
    M = 6
    x = rng.uniform(0, 8, (M, 3))
    o.set_beads(x, rng.standard_normal((M, 3)), np.ones(M))
    o.set_bonds(np.array([(i, i + 1) for i in range(M - 1)]), np.full(M - 1, 20.0), np.full(M - 1, 1.0))
    o.init()
    o.set_field(0, rng.standard_normal((3, dims[2], dims[1], dims[0])))
    o.step(3)
    u = o.field(0)
    assert np.abs(u.sum(axis=(1, 2, 3))).max() < 1e-11
    assert np.abs(_div(u)).max() < 1e-11 * np.abs(u).max()


def test_free_bead_first_step_closed_form():
    # quiescent fluid, no bonds: v1 = v0 (1 - ups dt / m), x1 = x0 + v0 dt
    o = _system((16, 16, 16), noise=False, mu=1.0, rho=1.0, dt=0.01)
    x0 = np.array([[5.3, 7.1, 8.9]])
    v0 = np.array([[0.3, -0.2, 0.5]])
    m = 2.0
    o.set_beads(x0, v0, np.array([m]))
    o.init()
    o.step(1)
    x, v, _, ft = o.beads()
    ups = 6 * np.pi * 2 * 1.0 * 1.0
    assert np.allclose(v, v0 * (1 - ups * 0.01 / m), rtol=1e-15, atol=1e-16)
    assert np.allclose(x, x0 + v0 * 0.01, rtol=1e-15)
    assert np.allclose(ft, -ups * v0, rtol=1e-14)


def test_momentum_bookkeeping_first_step():
    """From u = 0, the fluid gains exactly the bead impulse: after step 1,
    rho * sum(u*) = -sum(t_f) before the mean is removed.  Observable here via
    the second step's gather being zero-mean-free; instead check the spread
    directly: sum of the spread field equals -sum of impulses (sum phi = 1)."""
    rng = np.random.default_rng(9)
    shape = (8, 9, 10)
    uf = np.zeros(shape)
    tot = 0.0
    for _ in range(10):
        r = rng.uniform(0.1, 7.9, 3) + 0.013
        F = rng.standard_normal()
        O.lam(r[0], r[1], r[2], F, uf)
        tot += F
    assert abs(uf.sum() + tot) < 1e-12


def test_bond_force_two_beads():
    o = _system((16, 16, 16))
    x = np.array([[1.0, 1.0, 1.0], [15.5, 1.0, 1.0]])  # 1.5 apart through the periodic seam
    o.set_beads(x, np.zeros((2, 3)), np.ones(2))
    o.set_bonds(np.array([[0, 1]]), np.array([100.0]), np.array([1.0]))
    f = o.bond_forces(x)
    # E = K (r - r0)^2 -> |F| = 2 K (r - r0) = 100, attractive: bead 0 pulled in -x
    assert np.allclose(f[0], [-100.0, 0, 0], atol=1e-12)
    assert np.allclose(f[1], [100.0, 0, 0], atol=1e-12)

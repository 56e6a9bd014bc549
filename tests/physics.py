"""Closed forms of the discrete SELM scheme for the statistical and
steady-state checks (numpy; test infrastructure).

Conventions (dx = 1 is assumed by the helpers that say so):
- grid fields are [z][y][x];
- component d of the velocity lives on the MAC face, so its Peskin stencil
  sits at the bead's grid coordinate shifted by -1/2 on the other two axes
  (cuda/SELM_Kernel.cu:375-422);
- the Peskin factor keeps the reference's value 1.0 at r >= 2
  (cuda/SELM_Kernel.cu:387-395);
- numpy's fftn uses exp(-i k x), the forward convention of the step's r2c.

Symbols of the discrete operators at mode k (theta_a = 2 pi k_a / n_a):

| symbol | definition |
|---|---|
| D_a (forward difference) | exp(i theta_a) - 1 |
| lambda_k (7-point Laplacian) | sum_a 2 (cos theta_a - 1) = -sum_a abs(D_a)^2 |
| pi_d(k) (Helmholtz projector, diagonal) | 1 - abs(D_d)^2 / sum_a abs(D_a)^2 |
"""
import numpy as np


def phi(r, f4=True):
    r = np.abs(np.asarray(r, dtype=np.float64))
    out = np.where(r < 1.0, (3.0 - 2.0 * r + np.sqrt(np.maximum(1.0 + 4.0 * r * (1.0 - r), 0.0))) / 8.0,
                   np.where(r < 2.0, (5.0 - 2.0 * r - np.sqrt(np.maximum(-7.0 + 4.0 * r * (3.0 - r), 0.0))) / 8.0,
                            1.0 if f4 else 0.0))
    return out


def stencil_field(X, d, dims, lo=(0.0, 0.0, 0.0), dx=1.0, f4=True):
    """Peskin weights of component d's stencil for a bead at X, as a dense
    [z][y][x] array (the field one unit of impulse spreads to)."""
    nx, ny, nz = dims
    W = np.zeros((nz, ny, nx))
    s = [(X[a] - lo[a]) / dx for a in range(3)]
    r = [s[a] - (0.0 if a == d else 0.5) for a in range(3)]
    p = [int(np.floor(r[a])) - 1 for a in range(3)]
    w = [phi(r[a] - (p[a] + np.arange(4)), f4) for a in range(3)]
    for k in range(4):
        for j in range(4):
            for i in range(4):
                W[(p[2] + k) % nz, (p[1] + j) % ny, (p[0] + i) % nx] += w[2][k] * w[1][j] * w[0][i]
    return W


def symbols(dims):
    """|D_a|^2 per axis and lambda on the full [z][y][x] mode grid."""
    nx, ny, nz = dims
    th = [2.0 * np.pi * np.fft.fftfreq(n) for n in (nx, ny, nz)]
    d2x = (2.0 - 2.0 * np.cos(th[0]))[None, None, :]
    d2y = (2.0 - 2.0 * np.cos(th[1]))[None, :, None]
    d2z = (2.0 - 2.0 * np.cos(th[2]))[:, None, None]
    d2 = [np.broadcast_to(a, (nz, ny, nx)) for a in (d2x, d2y, d2z)]
    lam = -(d2[0] + d2[1] + d2[2])
    return d2, lam


def projector_diag(dims):
    """pi_d(k) for d = 0, 1, 2 (0 at k = 0)."""
    d2, lam = symbols(dims)
    tot = -lam
    safe = np.where(tot > 0, tot, 1.0)
    return [np.where(tot > 0, 1.0 - d2[d] / safe, 0.0) for d in range(3)]


def half_strength_planes(dims):
    """h_k of the reference's thermal forcing: 1/2 in the x-frequency planes
    kx = 0 and kx = nx/2 (where the c2r of per-mode draws keeps only the
    Hermitian part), except at self-conjugate modes; 1 elsewhere."""
    nx, ny, nz = dims
    kx = np.arange(nx)[None, None, :]
    ky = np.arange(ny)[None, :, None]
    kz = np.arange(nz)[:, None, None]
    plane = (kx == 0) | (2 * kx == nx)
    selfc = ((2 * ky) % ny == 0) & ((2 * kz) % nz == 0)
    h = np.where(plane & ~selfc, 0.5, 1.0)
    return np.broadcast_to(h, (nz, ny, nx))


def fluid_mode_variance(dims, mu, rho, kT, dt):
    """Stationary E|U_{k,d}|^2 / N of the noise-driven fluid (no beads, dx = 1).

    Per step the mode evolves as U <- P[(1 - a_k) U + Xi], with a_k =
    mu dt |lambda_k| / rho.  The forcing has E|Xi_{k,d}|^2 = 2 N kT dt mu
    |lambda_k| h_k / rho^2, so the stationary variance is E|PXi|^2 / (a (2 - a)):

        E|U_{k,d}|^2 / N = 2 kT h_k pi_d(k) / (rho (2 - a_k)).
    """
    _, lam = symbols(dims)
    a = mu * dt * (-lam) / rho
    h = half_strength_planes(dims)
    P = projector_diag(dims)
    return [np.where(lam < 0, 2.0 * kT * h * P[d] / (rho * (2.0 - a)), 0.0) for d in range(3)]


def fluid_mobility(X, dims, mu, f4=True, weights=None):
    """K_dd = (1/(mu N)) sum_{k != 0} w_k |S_d(k)|^2 pi_d(k) / |lambda_k| (dx = 1).

    This is the fluid-mediated part of the steady velocity of a bead at X per
    unit force along d.  At steady state mu L u = -P S F; the interpolation is
    the adjoint of the spread.  weights = None gives the mobility;
    weights = h_k gives the fluid-noise part of the self-diffusivity with the
    reference's half-strength planes."""
    nx, ny, nz = dims
    N = nx * ny * nz
    _, lam = symbols(dims)
    P = projector_diag(dims)
    out = np.zeros(3)
    for d in range(3):
        S = np.fft.fftn(stencil_field(X, d, dims, f4=f4))
        term = np.abs(S) ** 2 * P[d]
        if weights is not None:
            term = term * weights
        out[d] = np.sum(np.where(lam < 0, term / np.where(lam < 0, -lam, 1.0), 0.0)) / (mu * N)
    return out


def upsilon(mu, dx=1.0):
    """Bead drag coefficient 6 pi (2 dx) mu (cuda/SELM_Lagrangian_LAMMPS.cu:138)."""
    return 6.0 * np.pi * 2.0 * dx * mu


def unwrap(xs, L):
    """Unwrap a [T][M][3] sequence of periodic positions (steps < L/2)."""
    d = np.diff(xs, axis=0)
    d -= L * np.round(d / L)
    return np.concatenate([xs[:1], xs[:1] + np.cumsum(d, axis=0)], axis=0)

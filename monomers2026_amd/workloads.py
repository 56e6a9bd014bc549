"""Seeded synthetic bead-spring workloads.

Chains are random walks of unit steps with isotropic directions, wrapped into
the periodic box [0, L)^3 (dx = 1); bonds are LAMMPS `bond_style harmonic`
K = 100, r0 = 1 between consecutive beads; bead mass 1; v0 = 0.  The named
workloads are the SELM configurations this package is measured on:

  chain64_32      32^3 grid, one 64-bead chain, bead 0 at the box centre
  chains1000_64   64^3 grid, 10 chains x 100 beads at uniform-random centres
  dilute20k_128   128^3 grid, 200 chains x 100 beads, 1000 seeded beads carry
                  a constant external force (0.1, 0, 0)
  ensemble256     256 independent systems, edge in {16,24,32,48,64}, one chain
                  of 32-512 beads (log-uniform), mu ~ U[0.5, 2]
"""
import numpy as np

K_BOND = 100.0
R0_BOND = 1.0


def random_walk_chain(rng, n, start, L):
    """n beads, unit steps with isotropic directions, wrapped into [0, L)."""
    pos = np.empty((n, 3))
    pos[0] = start
    for i in range(1, n):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        pos[i] = pos[i - 1] + d
    return np.mod(pos, L)


def chains(rng, nchains, nper, L, centre=None):
    xs, bonds = [], []
    for c in range(nchains):
        if centre is not None and c == 0:
            start = np.asarray(centre, dtype=np.float64)
        else:
            start = rng.uniform(0.0, L, 3)
        xs.append(random_walk_chain(rng, nper, start, L))
        base = c * nper
        bonds.extend((base + i, base + i + 1) for i in range(nper - 1))
    x = np.concatenate(xs, axis=0)
    ij = np.asarray(bonds, dtype=np.int32).reshape(-1, 2)
    return x, ij


class Workload:
    def __init__(self, name, edge, x, ij, ext_idx=None, ext_F=None, mu=1.0, rho=1.0, kB=1.0, T=1.0,
                 dt=0.01, steps=1000):
        self.name = name
        self.edge = edge
        self.dims = (edge, edge, edge)
        self.lo = (0.0, 0.0, 0.0)
        self.hi = (float(edge),) * 3
        self.x = x
        self.M = x.shape[0]
        self.v = np.zeros_like(x)
        self.m = np.ones(self.M)
        self.ij = ij
        self.K = np.full(ij.shape[0], K_BOND)
        self.r0 = np.full(ij.shape[0], R0_BOND)
        self.ext_idx = ext_idx
        self.ext_F = ext_F
        self.mu, self.rho, self.kB, self.T, self.dt = mu, rho, kB, T, dt
        self.steps = steps

    def params(self, precision, stochastic, seed=123456, f4_quirk=True):
        from . import selm
        return selm.make_params(self.dims, self.lo, self.hi, rho=self.rho, mu=self.mu, kB=self.kB, T=self.T,
                                dt=self.dt, seed=seed, stochastic=stochastic, precision=precision,
                                f4_quirk=f4_quirk)

    def build_system(self, precision, stochastic, seed=123456, f4_quirk=True):
        from . import selm
        s = selm.System(self.params(precision, stochastic, seed, f4_quirk))
        s.set_beads(self.x, self.v, self.m)
        s.set_bonds(self.ij, self.K, self.r0)
        if self.ext_idx is not None:
            s.set_external_force(self.ext_idx, self.ext_F)
        s.init()
        return s


def chain64_32(seed=1):
    rng = np.random.default_rng(seed)
    L = 32
    x, ij = chains(rng, 1, 64, L, centre=(L / 2.0,) * 3)
    return Workload("chain64_32", L, x, ij, steps=1000)


def chains1000_64(seed=2):
    rng = np.random.default_rng(seed)
    L = 64
    x, ij = chains(rng, 10, 100, L)
    return Workload("chains1000_64", L, x, ij, steps=1000)


def dilute20k_128(seed=3):
    rng = np.random.default_rng(seed)
    L = 128
    x, ij = chains(rng, 200, 100, L)
    ext_idx = np.sort(rng.choice(x.shape[0], 1000, replace=False)).astype(np.int32)
    ext_F = np.tile(np.array([0.1, 0.0, 0.0]), (ext_idx.size, 1))
    return Workload("dilute20k_128", L, x, ij, ext_idx=ext_idx, ext_F=ext_F, steps=1000)


def stretched(edge=48, nchains=6, nper=40, seed=12, stretch=(0.7, 1.3), fext=0.5):
    """Chains with bond steps drawn in [0.7, 1.3] (off the rest length 1) and a
    constant force on every fifth bead: non-trivial deterministic dynamics."""
    rng = np.random.default_rng(seed)
    xs, bonds = [], []
    for c in range(nchains):
        pos = np.empty((nper, 3))
        pos[0] = rng.uniform(0.0, edge, 3)
        for i in range(1, nper):
            d = rng.standard_normal(3)
            d *= rng.uniform(*stretch) / np.linalg.norm(d)
            pos[i] = pos[i - 1] + d
        xs.append(np.mod(pos, edge))
        bonds.extend((c * nper + i, c * nper + i + 1) for i in range(nper - 1))
    x = np.concatenate(xs)
    ij = np.asarray(bonds, np.int32)
    idx = np.arange(0, x.shape[0], 5, dtype=np.int32)
    F = rng.uniform(-fext, fext, (idx.size, 3))
    return Workload("stretched%d" % edge, edge, x, ij, ext_idx=idx, ext_F=F, steps=30)


def ensemble(n=256, seed=4, edges=(16, 24, 32, 48, 64), mmin=32, mmax=512):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        L = int(rng.choice(edges))
        M = int(np.floor(np.exp(rng.uniform(np.log(mmin), np.log(mmax)))))
        mu = float(rng.uniform(0.5, 2.0))
        x, ij = chains(rng, 1, M, L)
        out.append(Workload("ensemble_%d" % r, L, x, ij, mu=mu, steps=500))
    return out

"""Device status flags and dense tiles, through the C ABI.

- 600 beads packed into one 8^3 tile, FP64 against the oracle to 1e-12 with
  nothing flagged: on a 16^3 grid the tile's bucket (sized to the bead
  density: 2432 slots for 600 beads over 8 tiles) holds them all; on a 64^3
  grid it has the minimum 256 slots and the rest go through the overflow
  pairs.
- More than 16384 beads in one tile's reach is outside the fixed-point range
  of the spread: the kernel exits normally and selm_sync returns
  SELM_ERR_DEVICE with SELM_STATUS_TILE_OVERFLOW raised.
- In a pool, a system whose bead moves more than half the box in one step
  and whose velocity then overflows is flagged (DISPLACEMENT | NONFINITE)
  and evicted by the serving loop; the other systems' results stay bitwise
  equal to running them alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _packed(n, lo, hi, edge, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, (n, 3))
    v = rng.standard_normal((n, 3))
    m = rng.uniform(0.5, 2.0, n)
    return x, v, m


@pytest.mark.parametrize("noise,edge", [(False, 16), (True, 16), (True, 64)])
def test_dense_tile_beyond_bucket_fp64_vs_oracle(gpu_lib, noise, edge):
    import oracle_lib as O
    from monomers2026_amd import selm
    dims, edge = (edge,) * 3, float(edge)
    x, v, m = _packed(600, 2.0, 6.0, edge, 3)
    ij = np.array([(i, i + 1) for i in range(599)], np.int32)
    K, r0 = np.full(599, 20.0), np.full(599, 0.5)
    p = selm.make_params(dims, (0, 0, 0), (edge,) * 3, dt=0.005, seed=17, stochastic=noise)
    s = selm.System(p)
    s.set_beads(x, v, m)
    s.set_bonds(ij, K, r0)
    s.init()
    o = O.Oracle(*dims, (0, 0, 0), (edge,) * 3, dt=0.005, seed=17, noise=noise)
    o.set_beads(x, v, m)
    o.set_bonds(ij, K, r0)
    o.init()
    for n in (1, 4):
        s.step(n)
        o.step(n)
        xg, vg, _, ftg = s.beads()
        xo, vo, _, fto = o.beads()
        ug, uo = s.field(), o.field(0)
        ex = np.abs(xg - xo).max() / edge
        ev = np.abs(vg - vo).max() / max(np.abs(vo).max(), 1.0)
        ef = np.abs(ftg - fto).max() / max(np.abs(fto).max(), 1.0)
        eu = np.abs(ug - uo).max() / max(np.abs(uo).max(), 1.0)
        print("dense tile (600 beads) noise=%d step %d: x %.2e v %.2e f %.2e u %.2e"
              % (noise, o.timestep, ex, ev, ef, eu), flush=True)
        assert max(ex, ev, ef, eu) <= 1e-12
    assert s.status() == 0
    s.sync()
    s.close()


def test_dense_tile_fp32_bitwise_repeatable(gpu_lib):
    from monomers2026_amd import selm
    dims = (16, 16, 16)
    x, v, m = _packed(600, 2.0, 6.0, 16.0, 4)
    out = []
    for _ in range(2):
        p = selm.make_params(dims, (0, 0, 0), (16.0,) * 3, dt=0.005, seed=5, stochastic=True,
                             precision=selm.PREC_FP32_BULK)
        s = selm.System(p)
        s.set_beads(x, v, m)
        s.init()
        s.step(5)
        out.append((s.beads()[0], s.field()))
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_tile_overflow_raises_status(gpu_lib):
    from monomers2026_amd import selm
    dims = (16, 16, 16)
    x, v, m = _packed(17000, 2.5, 5.5, 16.0, 6)
    p = selm.make_params(dims, (0, 0, 0), (16.0,) * 3, dt=0.001, seed=5, precision=selm.PREC_FP32_BULK)
    s = selm.System(p)
    s.set_beads(x, np.zeros_like(v), m)
    s.init()
    s.step(1)
    with pytest.raises(selm.SelmError) as e:
        s.sync()
    assert e.value.code == selm.SELM_ERR_DEVICE and "tile-overflow" in str(e.value)
    f = s.status(clear=True)
    assert f & selm.STATUS_TILE_OVERFLOW
    assert s.status() == 0
    s.close()


def test_pool_evicts_flagged_system(gpu_lib):
    from monomers2026_amd import selm, serve, workloads
    prec = selm.PREC_FP32_BULK
    rng = np.random.default_rng(8)
    reqs = []
    for i, L in enumerate((24, 32, 16)):
        xx, ij = workloads.chains(rng, 3, 20, L)
        w = workloads.Workload("s%d" % i, L, xx, ij, mu=1.0)
        if i == 1:
            # bead 7 moves 1e306 box lengths in the first step and its
            # velocity update overflows (finite inputs, checked by the host)
            w.v = w.v.copy()
            w.m = w.m.copy()
            w.v[7] = (1e308, 0.0, 0.0)
            w.m[7] = 1e-10
        reqs.append(serve.Request(i, w, 12, seed=40 + i, arrival=0))
    srv = serve.Server(1 << 30, 4, precision=prec, status_every=1)
    for r in reqs:
        srv.submit(r)
    t = 0
    while not srv.idle():
        srv.tick(t)
        t += 1
        assert t < 1000
    srv.close()
    assert [r.rid for r in srv.evicted] == [1]
    f = reqs[1].failed
    assert f & selm.STATUS_DISPLACEMENT and f & selm.STATUS_NONFINITE, f
    assert sorted(r.rid for r in srv.finished) == [0, 2]
    for r in (reqs[0], reqs[2]):
        s = r.w.build_system(prec, True, seed=r.seed)
        s.step(r.steps)
        xa, va, _, _ = s.beads()
        assert s.status() == 0
        s.close()
        assert np.array_equal(r.result[0], xa) and np.array_equal(r.result[1], va), r.rid


def test_pool_sync_names_flagged_slot(gpu_lib):
    from monomers2026_amd import selm, workloads
    prec = selm.PREC_FP64
    rng = np.random.default_rng(9)
    pool = selm.Pool(1 << 30, 4, precision=prec)
    slots = []
    for i in range(3):
        xx, ij = workloads.chains(rng, 2, 10, 16)
        w = workloads.Workload("p%d" % i, 16, xx, ij)
        v = np.zeros_like(xx)
        if i == 2:
            v[3] = (0.0, 0.6 * 16 / w.dt, 0.0)  # moves 0.6 of the box in one step: finite, flagged
        slots.append(pool.admit(w.params(prec, False, seed=i), xx, v, w.m, w.ij, w.K, w.r0))
    pool.step(1)
    with pytest.raises(selm.SelmError) as e:
        pool.sync()
    assert e.value.code == selm.SELM_ERR_DEVICE and ("slot %d" % slots[2]) in str(e.value)
    assert pool.status(slots[0]) == 0 and pool.status(slots[1]) == 0
    assert pool.status(slots[2], clear=True) == selm.STATUS_DISPLACEMENT
    pool.sync()  # cleared
    pool.close()

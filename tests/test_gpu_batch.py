"""The ensemble form (lulesh_batch_*): every member of a batch must be bit
for bit the run of the oracle (project cube root) on the same configuration
and energy factor, whatever the launch shape, the persistent grid sizes,
graph replay or plain launches, and an autotuning pass in between.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "ql", "qq", "v",
          "ss", "delv", "vdov", "arealg")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def member_state(batch, m):
    rc, st = batch.stats(m)
    return rc, st, {f: batch.get(m, f) for f in FIELDS}


def assert_same(sa, sb, label):
    rca, sta, fa = sa
    rcb, stb, fb = sb
    assert rca == rcb, label
    assert (sta.cycle, sta.time, sta.deltatime) == (stb.cycle, stb.time, stb.deltatime), label
    bad = [f for f in FIELDS if not np.array_equal(bits(fa[f]), bits(fb[f]))]
    assert not bad, (label, bad)


def test_members_match_oracle(lulesh_gpu, oracle):
    """Small members run to t = 1e-2 against the oracle, full state."""
    nx = [4, 7, 5, 10, 6]
    ef = [1.0, 0.85, 1.17, 0.93, 1.2]
    b = lulesh_gpu.Batch(nx, ef)
    b.set_shape(16, 8, 4, 2, sweep_blocks=0, elem_blocks=0, cycles_per_graph=8)
    assert b.run() == 0
    for m, (n, f) in enumerate(zip(nx, ef)):
        o = oracle.Oracle(n, mode=2, cbrt_mode=1, threads=8, efactor=f)
        assert o.run() == 0
        rc, st = b.stats(m)
        assert rc == 0 and st.done == 1
        assert st.cycle == o.cycle, (n, f)
        assert st.time == o.scalar("time")
        for fld in FIELDS:
            g, r = b.get(m, fld), o.get(fld)
            assert np.array_equal(bits(g), bits(r)), (n, f, fld)


def test_member_equals_single_domain(lulesh_gpu):
    """With factor 1 a member is the plain single-domain run."""
    b = lulesh_gpu.Batch([12, 9], max_cycles=150)
    assert b.run() == 0
    for m, n in enumerate([12, 9]):
        d = lulesh_gpu.Domain(n, max_cycles=150)
        d.run()
        rc, st = b.stats(m)
        assert st.cycle == d.stats().cycle == 150
        for fld in FIELDS:
            assert np.array_equal(bits(b.get(m, fld)), bits(d.get(fld))), (n, fld)


@pytest.mark.parametrize("shape", [
    (32, 8, 16, 2, 0, 0, 16),
    (16, 16, 8, 2, 512, 1024, 16),
    (16, 8, 32, 2, 256, 0, 0),     # plain launches, no graph
    (32, 4, 8, 2, 0, 4096, 5),
])
def test_shapes_and_graph_are_bitwise_neutral(lulesh_gpu, shape):
    nx, ef = lulesh_gpu.draw_ensemble(12, 8, 40, seed=7)
    ref = lulesh_gpu.Batch(nx, ef, max_cycles=60)
    ref.set_shape(32, 8, 16, 2, 0, 0, 0)
    ref.run()
    b = lulesh_gpu.Batch(nx, ef, max_cycles=60)
    b.set_shape(*shape)
    b.run()
    for m in range(len(nx)):
        assert_same(member_state(ref, m), member_state(b, m), (shape, m))


def test_autotune_restores_state(lulesh_gpu):
    nx, ef = lulesh_gpu.draw_ensemble(10, 16, 48, seed=3)
    a = lulesh_gpu.Batch(nx, ef, max_cycles=80)
    a.step(20)
    ms = a.autotune(trial_cycles=8)
    assert ms > 0.0
    a.run()
    b = lulesh_gpu.Batch(nx, ef, max_cycles=80)
    b.step(20)
    b.run()
    for m in range(len(nx)):
        assert_same(member_state(a, m), member_state(b, m), m)


def test_members_stop_on_their_own(lulesh_gpu):
    """Members of different sizes need different cycle counts to reach the
    stop time; each stops at its own, the others go on."""
    nx = [4, 6, 9]
    b = lulesh_gpu.Batch(nx, stoptime=2.0e-3)
    pending = b.step(1)
    assert pending == 3
    b.run()
    cycles = []
    for m, n in enumerate(nx):
        rc, st = b.stats(m)
        d = lulesh_gpu.Domain(n, stoptime=2.0e-3)
        d.run()
        assert rc == 0 and st.done == 1
        assert st.cycle == d.stats().cycle and st.time == d.stats().time
        cycles.append(st.cycle)
    assert len(set(cycles)) == 3


def test_default_ensemble_members(lulesh_gpu, oracle):
    """The default 64-member ensemble (nx in [16, 64], factors in
    [0.8, 1.2], seed 0) to t = 1e-2; its smallest member against the oracle
    and every member's stop time reached."""
    nx, ef = lulesh_gpu.draw_ensemble()
    assert len(nx) == 64 and nx.min() >= 16 and nx.max() <= 64
    b = lulesh_gpu.Batch(nx, ef)
    assert b.run() == 0
    for m in range(64):
        rc, st = b.stats(m)
        assert rc == 0 and st.done == 1 and st.time >= 1.0e-2 * (1 - 1e-12)
    m = int(np.argmin(nx))
    o = oracle.Oracle(int(nx[m]), mode=2, cbrt_mode=1, threads=16,
                      efactor=float(ef[m]))
    o.run()
    rc, st = b.stats(m)
    assert st.cycle == o.cycle
    for fld in ("e", "p", "xd", "x", "v", "ss"):
        assert np.array_equal(bits(b.get(m, fld)), bits(o.get(fld))), fld


def test_fused_members_equal_fused_single_domains(lulesh_gpu):
    """With fused multiply-adds (lulesh_batch_set_fused_multiply_add) every
    member is bit for bit the fused single-domain run of its size and
    energy factor, to t = 1e-2, and differs from the bit-identical one."""
    nx = [9, 14, 11]
    ef = [1.0, 0.87, 1.12]
    b = lulesh_gpu.Batch(nx, ef, fused=True)
    b.set_shape(16, 8, 4, 2, sweep_blocks=0, elem_blocks=0, cycles_per_graph=8)
    assert b.run() == 0
    e = lulesh_gpu.Batch(nx[:1], ef[:1])
    assert e.run() == 0
    assert not np.array_equal(bits(b.get(0, "e")), bits(e.get(0, "e")))
    e.close()
    for m, (n, f) in enumerate(zip(nx, ef)):
        d = lulesh_gpu.Domain(n, efactor=f, fused=True)
        d.run()
        rc, st = b.stats(m)
        assert rc == 0 and st.cycle == d.stats().cycle and st.time == d.stats().time
        for fld in FIELDS:
            assert np.array_equal(bits(b.get(m, fld)), bits(d.get(fld))), (n, f, fld)
        d.close()
    b.close()

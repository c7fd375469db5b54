"""Cycles replayed from a captured graph (lulesh_step replays graphs of 32
cycles, then queues the rest) against the same cycles queued one launch at
a time (kernel timing on, or lulesh_set_cycle_graphs off): bit for bit,
across remainders, a parameter change that must force a new capture, the
other arithmetic modes and an unstructured mesh."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def same(a, b):
    assert a.stats().cycle == b.stats().cycle
    assert a.stats().time == b.stats().time
    for f in FIELDS:
        assert np.array_equal(bits(a.get(f)), bits(b.get(f))), f


@pytest.mark.parametrize("kw", [dict(), dict(fused=True), dict(mixed=True)],
                         ids=["exact", "fused", "mixed"])
def test_graph_replay_is_direct_queueing(lulesh_gpu, kw):
    g = lulesh_gpu.Domain(24, **kw)
    q = lulesh_gpu.Domain(24, **kw)
    q.set_kernel_timing(True)
    for n in (5, 32, 45, 64, 1):
        assert g.step(n) == q.step(n)
        same(g, q)
    # the parameters are kernel arguments: the held graph no longer matches
    for d in (g, q):
        d.set_parameter("hgcoef", 2.0)
    for n in (40, 33):
        assert g.step(n) == q.step(n)
        same(g, q)
    # so is the sweep's launch shape
    for d in (g, q):
        d.set_sweep_shape(32, 16, 4, 2)
    assert g.step(70) == q.step(70)
    same(g, q)
    g.close()
    q.close()


def test_graphs_off_is_the_same(lulesh_gpu):
    """lulesh_set_cycle_graphs(d, 0) queues every cycle one launch at a
    time, with the same bits; turned back on, the graph is captured anew."""
    g = lulesh_gpu.Domain(20)
    q = lulesh_gpu.Domain(20)
    q.set_cycle_graphs(False)
    for n in (64, 37):
        assert g.step(n) == q.step(n)
        same(g, q)
    q.set_cycle_graphs(True)
    assert g.step(96) == q.step(96)
    same(g, q)
    g.close()
    q.close()


def test_graph_replay_runs_to_the_end(lulesh_gpu, oracle):
    """A full run ends where the oracle's does: the graph's cycles after
    the stop time leave the state alone."""
    d = lulesh_gpu.Domain(20, ic_mode=2)
    rc, n = d.run()
    o = oracle.Oracle(20, mode=2, cbrt_mode=1, threads=8)
    assert rc == 0 and o.run() == 0 and n == o.cycle
    for f in FIELDS:
        assert np.array_equal(bits(d.get(f)), bits(o.get(f))), f
    d.close()


def test_graph_replay_unstructured(lulesh_gpu, oracle):
    mesh = oracle.cube_mesh(12)
    g = lulesh_gpu.Domain.from_mesh(mesh, max_cycles=120)
    q = lulesh_gpu.Domain.from_mesh(mesh, max_cycles=120)
    q.set_kernel_timing(True)
    for n in (40, 64, 16):
        assert g.step(n) == q.step(n)
        same(g, q)
    g.close()
    q.close()

"""Size-independent properties of the HIP path at the benchmark size
(150^3) and failure handling forced from the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_runs_are_bit_reproducible(lulesh_gpu):
    """No atomics touch floating-point data: two runs give identical bits."""
    out = []
    for _ in range(2):
        d = lulesh_gpu.Domain(150, ic_mode=2, max_cycles=20)
        assert d.run() == (0, 20)
        out.append({f: d.get(f) for f in ("x", "xd", "e", "p", "q", "v", "ss")})
        d.close()
    for f in out[0]:
        assert np.array_equal(bits(out[0][f]), bits(out[1][f])), f


def test_sedov_symmetry_at_150(lulesh_gpu):
    """The Sedov problem is symmetric under any permutation of x, y, z;
    the discrete solution keeps it up to rounding."""
    d = lulesh_gpu.Domain(150, ic_mode=2, max_cycles=60)
    assert d.run() == (0, 60)
    n = 150
    e = d.get("e").reshape(n, n, n)  # [z][y][x]
    scale = np.abs(e).max()
    assert scale > 0
    for perm in ((0, 2, 1), (1, 0, 2), (2, 1, 0)):
        assert np.max(np.abs(e - e.transpose(perm))) <= 1e-10 * scale
    x = d.get("x").reshape(n + 1, n + 1, n + 1)
    y = d.get("y").reshape(n + 1, n + 1, n + 1)
    assert np.max(np.abs(x - y.transpose(0, 2, 1))) <= 1e-12
    d.close()


def test_checkpoint_resume_is_exact(lulesh_gpu):
    """Downloading the state and scalars and uploading them into a fresh
    domain continues the run bit for bit."""
    a = lulesh_gpu.Domain(24, ic_mode=2, max_cycles=80)
    assert a.step(40) == (0, 40)
    st = a.download_state()
    sc = a.stats()
    extra = {f: a.get(f) for f in ("ql", "qq")}
    assert a.step(40) == (0, 40)
    b = lulesh_gpu.Domain(24, ic_mode=2, max_cycles=80)
    b.upload_state(st)
    for f, v in extra.items():
        b.set(f, v)
    b.set_scalars(dtcourant=sc.dtcourant, dthydro=sc.dthydro, cycle=sc.cycle)
    assert b.step(40) == (0, 40)
    for f in ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss"):
        assert np.array_equal(bits(a.get(f)), bits(b.get(f))), f
    a.close()
    b.close()


def test_qstop_error(lulesh_gpu, oracle):
    """q above qstop stops the run with QStopError, as the reference exits
    with QStopError (reference/C/LULESH-OMP.cc:2368-2381); the check reads the
    q of the previous cycle.  The oracle stops at the same cycle."""
    d = lulesh_gpu.Domain(8, ic_mode=2)
    assert d.step(10) == (0, 10)
    qmax = d.get("q").max()
    assert qmax > 0
    d.set_parameter("qstop", qmax / 2)
    rc, n = d.step(5, raise_on_error=False)
    assert rc == lulesh_gpu.QSTOP_ERROR
    assert n == 1
    # the domain stays in the error state
    assert d.step(1, raise_on_error=False)[0] == lulesh_gpu.QSTOP_ERROR
    o = oracle.Oracle(8, mode=2)
    assert o.run(10) == 0
    o.set_scalar("qstop", qmax / 2)
    assert o.run(15) == -2
    assert o.cycle == d.stats().cycle == 11
    d.close()


def test_volume_error(lulesh_gpu):
    """A non-positive relative volume stops the run with VolumeError
    (reference/C/LULESH-OMP.cc:1541-1543)."""
    d = lulesh_gpu.Domain(8, ic_mode=1)
    v = d.get("v")
    v[3] = -1.0
    d.set("v", v)
    rc, _ = d.step(2, raise_on_error=False)
    assert rc == lulesh_gpu.VOLUME_ERROR
    d.close()


def test_tangled_mesh_volume_error(lulesh_gpu):
    """Folding a layer of elements inside out makes their Jacobian volume
    negative (reference/C/LULESH-OMP.cc:1579-1584)."""
    d = lulesh_gpu.Domain(6, ic_mode=2)
    x = d.get("x").reshape(7, 7, 7)
    x[:, :, 2] += 0.4  # the i = 2 node plane moves past i = 3 (spacing 0.1875)
    d.set("x", x.ravel())
    rc, _ = d.step(1, raise_on_error=False)
    assert rc == lulesh_gpu.VOLUME_ERROR
    d.close()


@pytest.mark.parametrize("form", ["fp64", "mixed", "lists"])
def test_nonpositive_volume_without_hourglass(lulesh_gpu, oracle, form):
    """v <= 0 raises VolumeError also when hgcoef = 0: the reference tests v
    of every element before it looks at hgcoef
    (reference/C/LULESH-OMP.cc:1540-1548).  The EOS clamps v to eosvmin
    before its own test, so without this check the run would go on."""
    nx = 6
    if form == "lists":
        d = lulesh_gpu.Domain.from_mesh(oracle.cube_mesh(nx), max_cycles=10)
        o = oracle.Oracle.from_mesh(oracle.cube_mesh(nx), cbrt_mode=1)
    else:
        d = lulesh_gpu.Domain(nx, ic_mode=2, mixed=(form == "mixed"), max_cycles=10)
        o = oracle.Oracle(nx, mode=2, cbrt_mode=1)
    for dom in (d, o):
        if dom is d:
            dom.set_parameter("hgcoef", 0.0)
        else:
            dom.set_scalar("hgcoef", 0.0)
        v = dom.get("v")
        v[40] = -0.25
        dom.set("v", v)
    rc, _ = d.step(1, raise_on_error=False)
    assert o.run(1) == lulesh_gpu.VOLUME_ERROR
    assert rc == lulesh_gpu.VOLUME_ERROR
    d.close()


def test_launch_schedule_does_not_change_results(lulesh_gpu):
    """The sweep's launch shape and the Q/EOS band interleave only move work
    between blocks: every choice gives the same bits."""
    runs = {}
    for shape, bands in (((32, 8, 16, 2), 0), ((32, 8, 16, 2), 3), ((32, 8, 5, 2), 1),
                         ((16, 16, 7, 2), 4), ((64, 4, 3, 1), 7), ((32, 16, 9, 2), 2)):
        d = lulesh_gpu.Domain(26, ic_mode=2, num_reg=11, cost=2, max_cycles=60)
        d.set_sweep_shape(*shape)
        d.set_eos_bands(bands)
        assert d.run() == (0, 60)
        runs[(shape, bands)] = {f: d.get(f) for f in ("x", "xd", "e", "p", "q", "v", "ss")}
        d.close()
    ref = next(iter(runs.values()))
    for key, r in runs.items():
        for f in r:
            assert np.array_equal(bits(r[f]), bits(ref[f])), (key, f)
    d = lulesh_gpu.Domain(5)
    with pytest.raises(lulesh_gpu.LuleshError):
        d.set_eos_bands(-1)


def test_largest_accepted_domain(lulesh_gpu):
    """nx = 600, the largest edge the library accepts (601^3 nodes keep every
    index in 32 bits): a few cycles run, stay finite and keep the Sedov
    symmetry; 601 is refused."""
    d = lulesh_gpu.Domain(600, ic_mode=2, max_cycles=3)
    assert d.run() == (0, 3)
    assert np.isfinite(d.get("e")).all() and np.isfinite(d.get("xd")).all()
    # the published criterion (reference/lulesh-ports.pdf section 6.2)
    assert d.symmetry()["max_rel_diff"] <= 1e-8
    d.close()
    with pytest.raises(lulesh_gpu.LuleshError):
        lulesh_gpu.Domain(601)

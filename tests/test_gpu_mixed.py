"""The mixed-precision form (lulesh_create_mixed): float32 element-local
arithmetic with float64 nodes, volumes, energy update and dt reductions.
Measured against the float64 step (itself bit-identical to the oracle) on the
quantities the form promises: origin energy within 1e-5 relative at
t = 1e-2, total energy, shock position, and the cycle count.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORIGIN_E_TOL = 1e-5


@pytest.mark.parametrize("nx", [30, 45])
def test_mixed_against_fp64_to_stoptime(lulesh_gpu, nx):
    ref = lulesh_gpu.Domain(nx)
    ref.run()
    mix = lulesh_gpu.Domain(nx, mixed=True)
    assert mix.mixed and not ref.mixed
    mix.run()
    sr, sm = ref.stats(), mix.stats()
    assert sm.done == 1 and sm.time >= 1.0e-2 * (1 - 1e-12)
    assert abs(sm.cycle - sr.cycle) <= max(2, sr.cycle // 200), (sm.cycle, sr.cycle)
    rel_e0 = abs(sm.e_origin - sr.e_origin) / abs(sr.e_origin)
    assert rel_e0 <= ORIGIN_E_TOL, rel_e0
    et_r, et_m = ref.total_energy(), mix.total_energy()
    assert abs(et_m - et_r) / et_r <= 1e-4, (et_m, et_r)
    h = 1.125 / nx
    rr, rm = ref.shock_radius(), mix.shock_radius()
    assert abs(rm["radius_x_axis"] - rr["radius_x_axis"]) <= 1.01 * h, (rr, rm)
    # the fields stay close where they matter
    for f in ("e", "p", "v"):
        a, b = mix.get(f), ref.get(f)
        assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(b)), f


def test_mixed_field_roundtrip(lulesh_gpu):
    d = lulesh_gpu.Domain(8, mixed=True, max_cycles=5)
    d.run()
    p = d.get("p")
    d.set("p", p)
    assert np.array_equal(d.get("p"), p)          # float32 values survive
    st = d.download_state()
    d.upload_state(st)
    st2 = d.download_state()
    for k in ("x", "xd", "e", "p", "q", "v", "ss"):
        assert np.array_equal(st[k], st2[k]), k


def test_mixed_volume_error(lulesh_gpu):
    """The float64 Jacobian check still raises VolumeError on a tangled
    mesh."""
    d = lulesh_gpu.Domain(6, mixed=True, max_cycles=10)
    x = d.get("x").reshape(7, 7, 7)
    x[:, :, 2] += 0.4
    d.set("x", x.reshape(-1))
    rc, _ = d.step(1, raise_on_error=False)
    assert rc == lulesh_gpu.VOLUME_ERROR

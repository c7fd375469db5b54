"""The benchmark's byte models and the measured traffic it reports, on the
CPU: the float64 phase table sums to 520 B per element and 296 B per node,
the mixed form's float32 arrays take 128 B per element off it, and the
cycle's measured traffic comes from the committed request counts."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# per phase: (element bytes, node bytes) of the float64 model, and the
# element bytes that are float32 arrays in the mixed form (counted at 8 B in
# the float64 model): force p, q, ss; kinematics delv, arealg, vdov; Q
# gradients the six delv_* / delx_*; Q region delv_* (3), delx_* (3), vdov
# and ql, qq written; EOS delv, p, q, ql, qq read and p, q, ss written; dt
# ss, vdov, arealg.
PHASES = {
    "force": ((80, 72), 3),
    "nodal": ((0, 128), 0),
    "kinematics": ((80, 48), 3),
    "q_gradients": ((96, 48), 6),
    "q_region": ((128, 0), 9),
    "eos": ((108, 0), 8),
    "dt": ((28, 0), 3),
}


def test_float64_model_sums():
    import bench
    e = sum(b[0] for b, _ in PHASES.values())
    n = sum(b[1] for b, _ in PHASES.values())
    assert (e, n) == (520, 296)
    # the kernels' shares of the same model
    assert sum(v[0] for v in bench.KERNEL_BYTES.values()) == 520
    assert sum(v[1] for v in bench.KERNEL_BYTES.values()) == 296


def test_mixed_model_halves_the_float32_arrays():
    import bench
    saved = sum(4 * n32 for _, n32 in PHASES.values())
    assert saved == 128
    assert bench.MIXED_BYTES == (520 - saved, 296)


def test_cycle_traffic_from_committed_counts():
    import bench
    t = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    keys = ("k_force_sweep@150", "k_kinematics_q@150", "k_q_eos_dt@150")
    total = sum(t[k]["bytes_per_launch"] for k in keys)
    r = bench.cycle_traffic(150, 0.5e-3)
    assert r["bytes_per_cycle"] == total
    assert abs(r["achieved_gbs"] - total / 0.5e-3 / 1e9) < 1e-9 * r["achieved_gbs"]
    assert r["bytes_per_zone_cycle"] == total / 150 ** 3
    # the measured cycle moves less than the model's 2.774 GB (the kernels
    # skip the nodelist, the force arrays and the second node read)
    assert total < 520 * 150 ** 3 + 296 * 151 ** 3
    assert bench.cycle_traffic(97, 1e-3) is None

"""Host-side pieces of the learned-Q pipeline: the equivalent-resolution
estimate inverts an exact power law, and the restriction of a state."""
import math
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "lulesh-accelerate_amd"))


def test_equivalent_resolution_inverts_a_power_law():
    import learned_q
    orders = {"density_l1": 1.7, "energy_l1": 1.2, "pressure_l1": 2.0}
    err = lambda k, n: 3.0 * n ** -orders[k]
    plain = {k: err(k, 40) for k in orders}
    mid = {k: err(k, 80) for k in orders}
    target = {k: err(k, 55.0) for k in orders}
    eq = learned_q.equivalent_resolution(40, plain, 80, mid, target)
    for k, s in orders.items():
        assert math.isclose(eq[k]["order"], s, rel_tol=1e-12)
        assert math.isclose(eq[k]["plain_nx"], 55.0, rel_tol=1e-12)
        assert math.isclose(eq[k]["cost_ratio"], (55.0 / 40) ** 4, rel_tol=1e-12)


def test_errors_of_a_state_dict():
    import numpy as np
    import learned_q
    n = 4
    ref = {"v": np.full(n ** 3, 0.5), "e": np.full(n ** 3, 2.0), "p": np.full(n ** 3, 4.0)}
    same = {k: a.copy() for k, a in ref.items()}
    d = learned_q.errors(same, ref)
    assert d["density_l1"] == d["energy_l1"] == d["pressure_l1"] == 0.0
    off = dict(same, e=np.full(n ** 3, 3.0))
    assert math.isclose(learned_q.errors(off, ref)["energy_l1"], 0.5)

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


class FakeDomain:
    """A domain-like object over numpy arrays (nx, get)."""

    def __init__(self, nx, fields):
        self.nx = nx
        self.fields = fields

    def get(self, k):
        return self.fields[k]

    def stats(self):
        class S:
            time, deltatime = 0.0, 0.0
        return S()


def test_restrict_to_is_the_block_restriction_where_it_divides():
    import numpy as np
    import learned_q
    rng = np.random.default_rng(3)
    nf, r = 12, 3
    ne = nf ** 3
    f = {"v": rng.uniform(0.5, 2.0, ne), "e": rng.uniform(0, 5, ne), "p": rng.uniform(0, 3, ne),
         "q": rng.uniform(0, 1, ne), "ss": rng.uniform(0, 1, ne), "ql": rng.uniform(0, 1, ne),
         "qq": rng.uniform(0, 1, ne)}
    for k in learned_q.NODE:
        f[k] = rng.uniform(size=(nf + 1) ** 3)
    fine = FakeDomain(nf, f)
    a = learned_q.restrict(fine, r)
    b = learned_q.restrict_to(fine, nf // r)
    for k in ("v", "e", "p"):
        assert np.allclose(a[k], b[k], rtol=1e-13, atol=0), k


def test_restrict_to_any_size_keeps_means_and_material():
    """On a grid that does not divide the fine one, every fine element goes
    to exactly one coarse element (the one holding its initial centroid),
    constants stay constant and the mass-weighted means are conserved."""
    import numpy as np
    import learned_q
    nf, nc = 12, 5
    ne = nf ** 3
    rng = np.random.default_rng(4)
    const = FakeDomain(nf, {"v": np.full(ne, 0.7), "e": np.full(ne, 2.5), "p": np.full(ne, 4.0)})
    c = learned_q.restrict_to(const, nc)
    assert np.allclose(c["v"], 0.7) and np.allclose(c["e"], 2.5) and np.allclose(c["p"], 4.0)
    e = rng.uniform(0, 1, ne)
    d = learned_q.restrict_to(FakeDomain(nf, {"v": np.ones(ne), "e": e, "p": e}), nc)
    m = ((np.arange(nf) + 0.5) * nc / nf).astype(int)
    cnt = np.bincount(m, minlength=nc)
    counts = (cnt[:, None, None] * cnt[None, :, None] * cnt[None, None, :]).reshape(-1)
    assert counts.sum() == ne
    assert np.isclose((d["e"] * counts).sum(), e.sum())
    # a brute-force check of one coarse element
    I = J = K = 2
    sel = [(k, j, i) for k in range(nf) for j in range(nf) for i in range(nf)
           if m[i] == I and m[j] == J and m[k] == K]
    want = np.mean([e[(k * nf + j) * nf + i] for k, j, i in sel])
    assert np.isclose(d["e"][(K * nc + J) * nc + I], want)


def test_measured_equivalent_interpolates_between_measured_runs():
    import learned_q
    sweep = []
    for n in (40, 48, 60):
        sweep.append({"nx": n, "density_l1": 3.0 * n ** -1.5, "energy_l1": 2.0 * n ** -1.0,
                      "pressure_l1": 1.0 * n ** -2.0, "wall_s": 1e-3 * n ** 4,
                      "zone_cycles": n ** 4})
    learned = {"density_l1": 3.0 * 44.0 ** -1.5, "energy_l1": 2.0 * 70.0 ** -1.0,
               "pressure_l1": 1.0 * 40.0 ** -2.0 * 1.5, "zone_cycles": 40 ** 4}
    eq = learned_q.measured_equivalent(learned, 1e-3 * 40 ** 4, sweep)
    d = eq["density_l1"]
    assert d["bound"] == "measured" and d["bracket"] == [40, 48]
    assert math.isclose(d["plain_nx"], 44.0, rel_tol=1e-12)
    assert math.isclose(d["wall_speedup"], (44.0 / 40) ** 4, rel_tol=1e-12)
    assert math.isclose(d["zone_cycle_ratio"], (44.0 / 40) ** 4, rel_tol=1e-12)
    assert eq["energy_l1"]["bound"] == "above" and eq["energy_l1"]["plain_nx"] == 60
    assert eq["pressure_l1"]["bound"] == "below" and eq["pressure_l1"]["plain_nx"] == 40

"""The learned-Q closure in the Q phase (lulesh_set_learned_q,
lulesh_q_features) and the trained network shipped with the package."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def net_or_skip(lulesh_gpu):
    import learned_q
    if not os.path.exists(learned_q.NET_FILE):
        pytest.fail(f"{learned_q.NET_FILE} missing: run learned_q.py train")
    return learned_q.load()


def test_fallback_everywhere_is_the_plain_step(lulesh_gpu):
    """With every input out of range the learned kernel keeps the plain
    limiter for every element: the run is bit-identical to the plain one."""
    net, _ = net_or_skip(lulesh_gpu)
    a = lulesh_gpu.Domain(20)
    a.run()
    b = lulesh_gpu.Domain(20)
    b.set_learned_q(net, z_max=1e-30)
    b.run()
    assert a.stats().cycle == b.stats().cycle
    for f in FIELDS:
        assert np.array_equal(bits(a.get(f)), bits(b.get(f))), f


def test_features_match_their_definition(lulesh_gpu):
    """lulesh_q_features against a numpy restatement of the inputs from the
    fields it leaves behind."""
    d = lulesh_gpu.Domain(16, max_cycles=200)
    d.run()
    feat, q = d.q_features()
    n = d.nx
    dv = [d.get(k).reshape(n, n, n) for k in ("delv_xi", "delv_eta", "delv_zeta")]
    dx = [d.get(k) for k in ("delx_xi", "delx_eta", "delx_zeta")]
    vdov, vnew, ss = d.get("vdov"), d.get("vnew"), d.get("ss")
    rho = d.get("elemMass") / (d.get("volo") * vnew)

    def phi(a, axis):
        # neighbours: symmetry (self) at the low face, free (0) at the high face
        lo = np.concatenate([np.take(a, [0], axis), np.take(a, range(n - 1), axis)], axis)
        hi = np.concatenate([np.take(a, range(1, n), axis),
                             np.zeros_like(np.take(a, [0], axis))], axis)
        norm = 1.0 / (a + 1e-36)
        m, p = lo * norm, hi * norm
        ph = 0.5 * (m + p)
        ph = np.minimum(ph, 2 * m)
        ph = np.minimum(ph, 2 * p)
        return np.clip(ph, 0.0, 1.0).reshape(-1)

    phis = [phi(dv[0], 2), phi(dv[1], 1), phi(dv[2], 0)]
    comp = ~(vdov > 0)
    du = [np.where(comp, np.minimum(dv[i].reshape(-1) * dx[i], 0.0), 0.0) for i in range(3)]
    hbar = (dx[0] + dx[1] + dx[2]) / 3.0
    want = np.stack(
        phis + [du[i] / (ss + np.abs(du[i])) for i in range(3)] +
        [dx[i] / hbar for i in range(3)] +
        [vdov * hbar / (ss + np.abs(vdov) * hbar), np.log(vnew), np.log(ss)], axis=1)
    assert np.allclose(feat, want.astype(np.float32), rtol=2e-6, atol=1e-7)
    sdu = du[0] + du[1] + du[2]
    usable = comp & (sdu < 0)
    assert np.allclose(q[:, 0], np.where(usable, -rho * sdu, 0.0), rtol=1e-14, atol=0)
    assert np.allclose(q[:, 1], np.where(usable, rho * (du[0]**2 + du[1]**2 + du[2]**2), 0.0),
                       rtol=1e-14, atol=0)


def test_trained_network_beats_the_plain_limiter(lulesh_gpu):
    """At 40^3 against the 120^3 solution (restricted) at t = 1e-2, for an
    origin-energy factor outside the training set: the learned closure's
    density and energy errors are below the plain limiter's."""
    import learned_q
    net, meta = net_or_skip(lulesh_gpu)
    f = 1.0
    assert all(abs(f - g) > 1e-3 for g in meta["efactors"])
    res = learned_q.evaluate(net, [f])[f"{f:.4f}"]
    plain, learned = res["plain"], res["learned"]
    assert learned["rc"] == 0 and plain["rc"] == 0
    assert learned["density_l1"] < plain["density_l1"], res
    assert learned["energy_l1"] < plain["energy_l1"], res
    assert abs(learned["shock_cell_x"] - learned["shock_cell_x_fine"]) <= 1, res

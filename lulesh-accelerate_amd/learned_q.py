"""The learned-Q closure: training data from fine runs, training on the GPU,
and evaluation on a coarse mesh against the fine solution.

Data.  Fine float64 Sedov runs at nx_f^3 (default 120^3, LULESH 2.0 ICs, 11
regions) with seeded origin-energy factors are restricted to the coarse
mesh nx_c = nx_f / r (default 40^3, r = 3) every `every` cycles:
  - coarse nodes are the fine nodes (r I, r J, r K): positions and
    velocities are injected;
  - coarse element values are means over the r^3 fine elements, mass
    weighted for the energy (equal fine masses: the plain mean) and volume
    weighted for the pressure, q and its parts and the sound speed; the
    relative volume is the fine volumes' sum over the coarse reference
    volume.
The restricted state is loaded into a coarse domain, whose kinematics and Q
gradients give the network inputs (lulesh_q_features).  The network's two
outputs multiply the monotonic limiter's linear and quadratic Q; their
targets are chosen against the fine solution itself.  From each restricted
snapshot the coarse solver runs to the next snapshot's time once per
candidate pair (c_l, c_q) of scales of the linear and quadratic Q (qlc and
qqc multiplied; (1, 1) is the plain limiter), and each compressing element
takes the pair whose coarse density, energy and pressure there, after the
interval, are closest to the restricted fine ones (a weighted sum of the
three errors, SEARCH_WEIGHTS, each over its field's mean size, averaged
over the element's 3 x 3 x 3 neighbourhood; a pair must beat (1, 1) by 2%
to be taken).
(Two simpler targets were tried and rejected: the fine solution's own
restricted Q, which is far smaller than a coarse mesh needs -- its shock is
three times thinner -- and leaves the coarse shock under-damped; and one
scale for both terms, which misses that the coarse mesh wants more linear
and less quadratic Q.)

Training.  Inputs standardised with the training mean and deviation; the
12-64-64-2 network trained by the library's bf16 MFMA kernels with Adam.

Evaluation.  Coarse runs (plain limiter, learned closure) and the fine run
of the same factor to t = 1e-2; each coarse solution is compared with the
restricted fine one: L1 errors of density, energy and pressure over the
coarse elements, relative to the fine solution's own L1 size, and the
shock position.

    python learned_q.py train [--out FILE] [--out-standin FILE] [--runs 16]
                              [--fine 120]
    python learned_q.py evaluate [--net FILE] [--efactor 1.0 ...]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, PKG)

import lulesh_amd  # noqa: E402

NET_FILE = os.path.join(PKG, "learned_q_net.npz")
QLC, QQC = 0.5, 2.0 / 3.0
NODE = ("x", "y", "z", "xd", "yd", "zd")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def block_mean(a: np.ndarray, n: int, r: int, w: np.ndarray | None = None) -> np.ndarray:
    """Mean over r^3 blocks of an n^3 element array (k, j, i order),
    weighted by w when given."""
    c = n // r
    a = a.reshape(c, r, c, r, c, r)
    if w is None:
        return a.mean(axis=(1, 3, 5)).reshape(-1)
    w = w.reshape(c, r, c, r, c, r)
    return ((a * w).sum(axis=(1, 3, 5)) / w.sum(axis=(1, 3, 5))).reshape(-1)


def restrict(fine: "lulesh_amd.Domain", r: int) -> dict:
    """The coarse state (nx_f / r)^3 of a fine domain (see module text)."""
    n = fine.nx
    if n % r:
        raise ValueError("fine size must be a multiple of the ratio")
    en = n + 1
    out = {}
    for k in NODE:
        out[k] = np.ascontiguousarray(
            fine.get(k).reshape(en, en, en)[::r, ::r, ::r]).reshape(-1)
    v = fine.get("v")
    out["v"] = block_mean(v, n, r)
    out["e"] = block_mean(fine.get("e"), n, r)
    for k in ("p", "q", "ss", "ql", "qq"):
        out[k] = block_mean(fine.get(k), n, r, v)
    st = fine.stats()
    out["time"], out["deltatime"] = st.time, st.deltatime
    return out


def restrict_to(fine: "lulesh_amd.Domain", nc: int) -> dict:
    """The fine solution's v, e and p on an nc^3 grid of any size, by
    material: the mesh is Lagrangian, so an element's logical index labels
    its material, and coarse element (I, J, K) takes the fine elements whose
    initial centroid lies in it, I = floor((i + 1/2) nc / nf).  v is the
    mean relative volume (the aggregate's density is 1 / v, the initial
    density being 1), e the mean energy (equal masses), p the volume-weighted
    pressure.  For nc dividing nf this is the block restriction of
    `restrict`."""
    nf = fine.nx
    m = ((np.arange(nf) + 0.5) * nc / nf).astype(np.int64)
    cid = ((m[:, None, None] * nc + m[None, :, None]) * nc + m[None, None, :]).reshape(-1)
    cnt = np.bincount(cid, minlength=nc ** 3).astype(np.float64)
    v = fine.get("v")
    sv = np.bincount(cid, weights=v, minlength=nc ** 3)
    return {"v": sv / cnt,
            "e": np.bincount(cid, weights=fine.get("e"), minlength=nc ** 3) / cnt,
            "p": np.bincount(cid, weights=fine.get("p") * v, minlength=nc ** 3) / sv}


def load_coarse(coarse: "lulesh_amd.Domain", st: dict) -> None:
    coarse.upload_state(st)


SCALES_L = (0.8, 1.0, 1.25, 1.5)
SCALES_Q = (0.5, 0.65, 0.8, 1.0)
# Weights of the density, energy and pressure errors in the target search.
# Networks trained with other weights, measured on held-out factors (0.85,
# 0.9, 1.0, 1.05, 1.175): weighting pressure up, e.g. (0.5, 1, 2), lowers the
# density and pressure L1 errors further (equal-error wall-time gain 3.4x and
# 2.5x against 2.6x and 1.4x) but leaves the 40^3 shock two cells behind the
# fine one at 0.9 and 1.05, where (1, 1, 1) keeps it within one (as the
# plain limiter does): the lower L1 comes from a smeared, lagging shock.  The
# energy error barely moves with any weighting (1.34x-1.43x).
SEARCH_WEIGHTS = (1.0, 1.0, 1.0)


def smooth3(a: np.ndarray, n: int) -> np.ndarray:
    """Mean over the 3 x 3 x 3 neighbourhood (clipped at the faces)."""
    a = a.reshape(n, n, n)
    p = np.pad(a, 1, mode="edge")
    s = np.zeros_like(a)
    for dk in range(3):
        for dj in range(3):
            for di in range(3):
                s += p[dk:dk + n, dj:dj + n, di:di + n]
    return (s / 27.0).reshape(-1)


FORMS = (0, 1)  # lulesh_amd.Domain.LEARNED_FACTORS, LEARNED_STANDIN


def search_targets(coarse, st0: dict, st1: dict, tol: float = 0.02,
                   weights=SEARCH_WEIGHTS):
    """Per form, (inputs, targets, plain-limiter outputs) of the compressing
    elements of the restricted state st0, targets from the scale search over
    the interval to st1 (see module text).  Factors form: the targets are the
    best scales (c_l, c_q), the plain limiter (1, 1).  Stand-in form: the
    targets are the limiting factors that give the same Q,
    c_l * ql_plain / (qlc * ql_unlimited) and the same for qq, the plain
    limiter's own factors the plain outputs."""
    load_coarse(coarse, st0)
    feats = {}
    for form in FORMS:
        coarse.set_learned_q_form(form)
        feats[form], q = coarse.q_features()
    coarse.set_learned_q_form(FORMS[0])
    feat = feats[FORMS[0]]
    keep = (q[:, 0] > 0) & ((q[:, 2] > 0) | (q[:, 3] > 0))
    if not keep.any():
        return None
    keep &= q[:, 0] > 1e-4 * q[keep, 0].max()
    n = coarse.nx
    ref = {"rho": 1.0 / st1["v"], "e": st1["e"], "p": st1["p"]}
    scale = {"rho": np.abs(ref["rho"] - 1.0).mean() + 1e-30,
             "e": np.abs(ref["e"]).mean() + 1e-30, "p": np.abs(ref["p"]).mean() + 1e-30}
    pairs = [(cl, cq) for cl in SCALES_L for cq in SCALES_Q]
    errs = []
    for cl, cq in pairs:
        load_coarse(coarse, st0)
        coarse.set_parameter("qlc_monoq", QLC * cl)
        coarse.set_parameter("qqc_monoq", QQC * cq)
        coarse.set_parameter("stoptime", st1["time"])
        rc, _ = coarse.run(raise_on_error=False)
        if rc != 0:
            errs.append(np.full(n ** 3, np.inf))
            coarse.set_scalars(time=st0["time"])  # clears the error state
            continue
        got = {"rho": 1.0 / coarse.get("v"), "e": coarse.get("e"), "p": coarse.get("p")}
        err = sum(w * np.abs(got[k] - ref[k]) / scale[k]
                  for w, k in zip(weights, ("rho", "e", "p")))
        errs.append(smooth3(err, n))
    coarse.set_parameter("qlc_monoq", QLC)
    coarse.set_parameter("qqc_monoq", QQC)
    coarse.set_parameter("stoptime", 1.0e-2)
    errs = np.array(errs)
    i_plain = pairs.index((1.0, 1.0))
    best = np.argmin(errs, axis=0)
    gain = errs[i_plain] - errs[best, np.arange(n ** 3)]
    best = np.where(gain > tol * (errs[i_plain] + 1e-12), best, i_plain)
    t = np.asarray(pairs)[best]
    ok = keep & np.all(np.isfinite(feats[0]), axis=1) & np.all(np.isfinite(feats[1]), axis=1)
    ok &= q[:, 1] > 0
    out = {0: (feats[0][ok].astype(np.float32), t[ok].astype(np.float32),
               np.ones((int(ok.sum()), 2), np.float32))}
    plain_l = q[ok, 2] / (QLC * q[ok, 0])
    plain_q = q[ok, 3] / (QQC * q[ok, 1])
    out[1] = (feats[1][ok].astype(np.float32),
              np.stack([t[ok, 0] * plain_l, t[ok, 1] * plain_q], axis=1).astype(np.float32),
              np.stack([plain_l, plain_q], axis=1).astype(np.float32))
    return out


def run_efactors(seed: int, n: int) -> np.ndarray:
    """n origin-energy factors in [0.8, 1.2] from the library's stream."""
    _, ef = lulesh_amd.draw_ensemble(n, 16, 16, 0.8, 1.2, seed=seed)
    return ef


def generate(efactors, fine_nx: int = 120, ratio: int = 3, every: int = 200) -> dict:
    """Training data of both forms: form -> (inputs, targets, plain outputs)."""
    data = {form: ([], [], []) for form in FORMS}
    coarse = lulesh_amd.Domain(fine_nx // ratio)
    for f in efactors:
        t0 = time.perf_counter()
        fine = lulesh_amd.Domain(fine_nx, efactor=float(f))
        nsnap = 0
        prev = None
        while True:
            rc, done = fine.step(every)
            if rc != 0:
                break
            cur = restrict(fine, ratio)
            if prev is not None:
                s = search_targets(coarse, prev, cur)
                if s is not None:
                    for form in FORMS:
                        for lst, a in zip(data[form], s[form]):
                            lst.append(a)
                    nsnap += 1
            prev = cur
            if done < every:
                break
        st = fine.stats()
        fine.close()
        log(f"learned_q: fine run efactor={f:.4f}: {st.cycle} cycles, {nsnap} "
            f"snapshots, {sum(len(x) for x in data[0][0])} samples so far, "
            f"{time.perf_counter() - t0:.1f} s")
    coarse.close()
    return {form: tuple(np.concatenate(lst) for lst in data[form]) for form in FORMS}


def init_params(seed: int) -> np.ndarray:
    """He-scaled normal weights, zero hidden biases, output biases 1 (the
    plain limiter's factors in the factors form, full limiting factors in
    the stand-in form)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, np.sqrt(2 / 12), (12, 64))
    w2 = rng.normal(0, np.sqrt(2 / 64), (64, 64))
    w3 = rng.normal(0, np.sqrt(1 / 64), (64, 2)) * 0.1
    b3 = np.array([1.0, 1.0])  # the plain limiter
    return np.concatenate([w1.ravel(), np.zeros(64), w2.ravel(), np.zeros(64),
                           w3.ravel(), b3]).astype(np.float32)


def train(x, t, epochs=40, batch=4096, lr=2e-3, seed=0):
    mean = x.mean(axis=0).astype(np.float32)
    std = (x.std(axis=0) + 1e-6).astype(np.float32)
    net = lulesh_amd.Mlp(init_params(seed), mean, std)
    _, loss0 = net.gradient(x, t)
    t0 = time.perf_counter()
    loss = net.train(x, t, epochs=epochs, batch=batch, lr=lr, seed=seed)
    log(f"learned_q: trained {epochs} epochs on {len(x)} samples in "
        f"{time.perf_counter() - t0:.1f} s, loss {loss0:.4g} -> {loss:.4g}")
    return net, loss0, loss


def save(path: str, net, meta: dict) -> None:
    np.savez(path, params=net.params(), in_mean=net.in_mean, in_std=net.in_std,
             meta=np.array(json.dumps(meta)))


NET_FILE_STANDIN = os.path.join(PKG, "learned_q_standin_net.npz")


def load(path: str = NET_FILE):
    """(network, meta) of a saved network; meta["form"] is its form (0 for
    files saved without one)."""
    with np.load(path, allow_pickle=False) as z:
        net = lulesh_amd.Mlp(z["params"], z["in_mean"], z["in_std"])
        meta = json.loads(str(z["meta"]))
    meta.setdefault("form", 0)
    return net, meta


METRICS = ("density_l1", "energy_l1", "pressure_l1")


def errors(coarse, ref: dict) -> dict:
    """L1 errors of a coarse solution (a Domain, or a restricted state dict
    with v, e, p) against a restricted fine one."""
    get = coarse.get if hasattr(coarse, "nx") else coarse.__getitem__
    rho_c, rho_f = 1.0 / get("v"), 1.0 / ref["v"]
    e_c, e_f = get("e"), ref["e"]
    p_c, p_f = get("p"), ref["p"]
    n = round(len(e_f) ** (1.0 / 3.0))
    d = {
        "density_l1": float(np.abs(rho_c - rho_f).sum() / np.abs(rho_f - 1.0).sum()),
        "energy_l1": float(np.abs(e_c - e_f).sum() / np.abs(e_f).sum()),
        "pressure_l1": float(np.abs(p_c - p_f).sum() / np.abs(p_f).sum()),
    }
    # shock position along x: the element of largest pressure on the axis
    d["shock_cell_x"] = int(np.argmax(p_c.reshape(n, n, n)[0, 0, :]))
    d["shock_cell_x_fine"] = int(np.argmax(p_f.reshape(n, n, n)[0, 0, :]))
    return d


def equivalent_resolution(nc: int, plain: dict, mid: int, plain_mid: dict,
                          target: dict) -> dict:
    """The plain-limiter resolution that would reach the target errors: a
    power law err(N) = a N^-s through the plain errors at nc and mid (both
    measured on the nc^3 grid), solved for the target's error, per metric,
    with the zone*cycle cost of that run relative to nc^3 (cycles grow
    with N, so the cost goes as N^4).  An estimate from two points."""
    out = {}
    for k in METRICS:
        s_k = np.log(plain[k] / plain_mid[k]) / np.log(mid / nc)
        n_eq = nc * (plain[k] / target[k]) ** (1.0 / s_k)
        out[k] = {"order": float(s_k), "plain_nx": float(n_eq),
                  "cost_ratio": float((n_eq / nc) ** 4)}
    return out


def measured_equivalent(learned: dict, learned_wall: float, sweep: list[dict]) -> dict:
    """Per metric, the plain resolution whose measured error equals the
    learned run's, interpolated (log-log) between the two measured plain
    resolutions that bracket it, and the plain run's wall time and
    zone*cycles there (interpolated the same way between the two measured
    runs) over the learned run's: the speed-up at equal error.  When the
    learned error lies beyond the measured range the nearest run bounds it
    and "bound" says which way."""
    sweep = sorted(sweep, key=lambda r: r["nx"])
    out = {}
    for k in METRICS:
        target = learned[k]
        errs = [r[k] for r in sweep]
        row = None
        for a, b in zip(sweep, sweep[1:]):
            ea, eb = a[k], b[k]
            if (ea - target) * (eb - target) <= 0 and ea != eb:
                w = np.log(target / ea) / np.log(eb / ea)

                def interp(key):
                    return float(np.exp(np.log(a[key]) + w * (np.log(b[key]) - np.log(a[key]))))

                n_eq = float(np.exp(np.log(a["nx"]) + w * (np.log(b["nx"]) - np.log(a["nx"]))))
                row = {"plain_nx": n_eq, "bracket": [a["nx"], b["nx"]], "bound": "measured",
                       "plain_wall_s": interp("wall_s"),
                       "plain_zone_cycles": interp("zone_cycles")}
                break
        if row is None:
            last = sweep[-1] if target < min(errs) else sweep[0]
            row = {"plain_nx": float(last["nx"]), "bracket": [last["nx"], last["nx"]],
                   "bound": "above" if target < min(errs) else "below",
                   "plain_wall_s": last["wall_s"], "plain_zone_cycles": last["zone_cycles"]}
        row["wall_speedup"] = row["plain_wall_s"] / learned_wall
        row["zone_cycle_ratio"] = row["plain_zone_cycles"] / learned["zone_cycles"]
        out[k] = row
    return out


PLAIN_SWEEP = (40, 44, 48, 52, 56, 60, 64, 68, 72)


def evaluate(net, efactors, fine_nx=120, ratio=3, z_max=4.0, mid_ratio=2,
             sweep=PLAIN_SWEEP, form: int = 0) -> dict:
    """Errors of the plain and learned limiters at fine_nx / ratio against
    the fine run restricted to that grid, and of the plain limiter at
    mid_ratio times that resolution (restricted the same way), from which
    the plain resolution matching the learned errors is estimated by a power
    law; and of the plain limiter at every resolution of `sweep`, each
    against the fine run restricted to its own grid (restrict_to), from
    which the plain resolution of equal error and its cost are measured
    (measured_equivalent)."""
    out = {}
    nc = fine_nx // ratio
    for f in efactors:
        t0 = time.perf_counter()
        fine = lulesh_amd.Domain(fine_nx, efactor=float(f))
        fine.run()
        ref = restrict(fine, ratio)
        refs = {n: restrict_to(fine, n) for n in sweep}
        fine_cycles = fine.stats().cycle
        fine.close()
        row = {"fine_cycles": fine_cycles}
        for name in ("plain", "learned"):
            d = lulesh_amd.Domain(nc, efactor=float(f))
            if name == "learned":
                d.set_learned_q(net, z_max, form)
            d.step(1)
            t1 = time.perf_counter()
            rc, n = d.run()
            wall = time.perf_counter() - t1
            st = d.stats()
            row[name] = errors(d, ref)
            row[name].update({"cycles": st.cycle, "rc": rc, "wall_s": wall,
                              "zone_cycles": nc ** 3 * n,
                              "zones_cycles_per_s": nc ** 3 * n / wall})
            d.close()
        if mid_ratio:
            nm = nc * mid_ratio
            d = lulesh_amd.Domain(nm, efactor=float(f))
            d.step(1)
            t1 = time.perf_counter()
            rc, n = d.run()
            wall = time.perf_counter() - t1
            row[f"plain_{nm}"] = errors(restrict(d, mid_ratio), ref)
            row[f"plain_{nm}"].update({"cycles": d.stats().cycle, "rc": rc,
                                      "zones_cycles_per_s": nm ** 3 * n / wall})
            d.close()
            row["learned_equivalent_plain"] = equivalent_resolution(
                nc, row["plain"], nm, row[f"plain_{nm}"], row["learned"])
        if sweep:
            runs = []
            for n_s in sweep:
                d = lulesh_amd.Domain(n_s, efactor=float(f))
                d.step(1)
                t1 = time.perf_counter()
                rc, n = d.run()
                wall = time.perf_counter() - t1
                r = errors(d, refs[n_s])
                r.update({"nx": n_s, "cycles": d.stats().cycle, "rc": rc, "wall_s": wall,
                          "zone_cycles": n_s ** 3 * n})
                d.close()
                runs.append(r)
            row["plain_sweep"] = runs
            row["learned_measured_equivalent"] = measured_equivalent(
                row["learned"], row["learned"]["wall_s"], runs)
        out[f"{float(f):.4f}"] = row
        log(f"learned_q: efactor {f:.4f}: plain {row['plain']['density_l1']:.4f} "
            f"learned {row['learned']['density_l1']:.4f} (density L1), "
            f"{time.perf_counter() - t0:.1f} s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("train", "evaluate"))
    ap.add_argument("--out", default=NET_FILE)
    ap.add_argument("--out-standin", default=NET_FILE_STANDIN)
    ap.add_argument("--net", default=NET_FILE)
    ap.add_argument("--runs", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--fine", type=int, default=120)
    ap.add_argument("--ratio", type=int, default=3)
    ap.add_argument("--every", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--efactor", type=float, nargs="*", default=[0.9, 1.0, 1.1])
    ap.add_argument("--z-max", type=float, default=4.0)
    args = ap.parse_args()
    if args.mode == "train":
        ef = run_efactors(args.seed, args.runs)
        data = generate(ef, args.fine, args.ratio, args.every)
        metas = {}
        for form, path in ((0, args.out), (1, args.out_standin)):
            x, t, plain = data[form]
            net, loss0, loss = train(x, t, epochs=args.epochs, lr=args.lr, seed=args.seed)
            plain_loss = float(0.5 * np.mean(np.sum((plain - t) ** 2, axis=1)))
            meta = {"form": form, "fine": args.fine, "ratio": args.ratio,
                    "every": args.every,
                    "criterion": "density+energy+pressure" if SEARCH_WEIGHTS == (1.0, 1.0, 1.0)
                                 else "%g density + %g energy + %g pressure" % SEARCH_WEIGHTS,
                    "efactors": [float(v) for v in ef], "samples": int(len(x)),
                    "epochs": args.epochs, "lr": args.lr, "loss_init": loss0,
                    "loss_final": loss, "loss_of_plain_limiter": plain_loss}
            save(path, net, meta)
            log(f"learned_q: saved {path}: {json.dumps(meta)}")
            metas[form] = meta
        print(json.dumps(metas))
    else:
        net, meta = load(args.net)
        res = evaluate(net, args.efactor, args.fine, args.ratio, args.z_max,
                       form=meta["form"])
        print(json.dumps({"net": meta, "evaluation": res}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Benchmark of the MI355X LULESH step.

A step is one LULESH cycle (TimeIncrement + LagrangeLeapFrog) of the
benchmark workload: the Sedov blast problem at 150^3 elements (3.375 M
zones, 151^3 nodes) with the LULESH 2.0 multi-region layout (11 regions,
balance 1, cost 1, seeded region assignment) and LULESH 2.0 initial
conditions, in float64 with fused multiply-adds (results within the
reference's tolerances; the bit-identical step, every value rounded as the
reference rounds it, is timed the same way and reported beside it).  The
domain is built once and stays resident on the GPU; the timed region is K
consecutive cycles after W warm-up cycles.

Prints one JSON line: zones*cycles/s of the timed cycles (queued back to
back, no events between kernels), the roofline of the dominant kernel
(algorithmic bytes per launch over its mean duration, from HIP events
around every launch on the domain's stream during cycles 26-50, which run
before the warm-up and the timed cycles) and
the reference LULESH 2.0.3 (oracle/_ref/lulesh2, built from the reference
sources) timed on this host's cores on a 20-cycle sample of the same
workload.

The "forms" object holds the other forms of the step on their own
workloads: the ensemble (64 Sedov members of uneven size, each to t=1e-2)
the mixed-precision step (120^3 to t=1e-2, with its errors against the
float64 run), the learned-Q closure (40^3 against the 120^3 solution) and
the unstructured step (the 150^3 cube given as explicit mesh lists).

    python bench.py [--steps K] [--warmup W] [--nx N] [--no-cpu-baseline]
                    [--no-forms | --forms batched,mixed,learned_q,unstructured]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lulesh-accelerate_amd"))

METRIC = ("zones·cycles/s, plain FP64 vs batched/mixed/learned-Q at "
          "stated error; HBM %")
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per element and per node of each kernel: the per-phase
# figures of the reference path's byte model (each array a phase reads or
# writes counted once, float64 = 8 B, int32 = 4 B; 2.0 adds the 4-byte
# region index to the Q-region, EOS and dt phases), assigned to the kernel
# that performs the phase.  Sum: 520 B/element + 296 B/node per cycle.
KERNEL_BYTES = {
    # force phase: p,q,v,volo,ss,elemMass + nodelist (80), read x..zd and
    # write fx,fy,fz (72 per node); nodal phase: read f, m, xd, x and write
    # xd, x (128 per node)
    "k_force_sweep": (80, 200),
    # kinematics (80 + 48 node) + monotonic-Q gradients (96 + 48 node)
    "k_kinematics_q": (176, 96),
    # Q region 124+4, EOS 104+4, dt constraints 24+4
    "k_q_eos_dt": (264, 0),
    "k_time_increment": (0, 0),
    "k_dt_finish": (0, 0),
}


def host_cores() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS where the environment
    sets it (the host's CPU share for this GPU), else the cores this
    process may run on."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    if n < 1:
        try:
            n = len(os.sched_getaffinity(0))
        except AttributeError:
            n = os.cpu_count() or 1
    return max(1, n)


def host_cpu_model() -> str:
    """The host CPU's model name (the lscpu "Model name" line)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nx: int, cycles: int) -> dict | None:
    """The reference LULESH 2.0.3 on this host's cores, same workload,
    first `cycles` cycles; the reference times its own cycle loop."""
    exe = os.path.join(ROOT, "oracle", "_ref", "lulesh2")
    cores = host_cores()
    if os.path.exists(exe):
        env = dict(os.environ, OMP_NUM_THREADS=str(cores),
                   OMP_PROC_BIND="close", OMP_PLACES="cores")
        out = subprocess.run([exe, "-s", str(nx), "-i", str(cycles)],
                             capture_output=True, text=True, env=env,
                             timeout=600)
        m = re.search(r"Grind time \(us/z/c\)\s*=\s*([0-9.eE+-]+)", out.stdout)
        c = re.search(r"Iteration count\s*=\s*(\d+)", out.stdout)
        if out.returncode == 0 and m and c:
            grind_us = float(m.group(1))
            return {"value": 1.0e6 / grind_us, "unit": "zones·cycles/s",
                    "cores": cores, "cpu": host_cpu_model(), "kind": "reference",
                    "sample": (f"lulesh2 (reference/C/lulesh-2.0.3, g++ -O3 "
                               f"-fopenmp) -s {nx} -i {c.group(1)}, "
                               f"OMP_NUM_THREADS={cores}; reference's own "
                               "loop timer")}
    # the oracle restatement when the reference binary is unavailable
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lulesh_oracle
    o = lulesh_oracle.Oracle(nx, mode=2, threads=cores)
    t0 = time.perf_counter()
    o.run(cycles)
    dt = time.perf_counter() - t0
    return {"value": nx ** 3 * o.cycle / dt, "unit": "zones·cycles/s",
            "cores": cores, "cpu": host_cpu_model(), "kind": "port",
            "sample": f"oracle restatement, {nx}^3, {o.cycle} cycles, "
                      f"{cores} OpenMP threads"}


# The mixed-precision form's algorithmic bytes in the same accounting: the
# element arrays it keeps in float32 (p, q, ql, qq, ss, delv, arealg, vdov
# and the six Q gradients) count 4 B, which takes 128 B off the element
# side of every cycle (force 12, kinematics 12, Q gradients 24, Q region
# 36, EOS 32, dt 12); the node side stays float64.
MIXED_BYTES = (392, 296)


def load_traffic(kernel: str, nx: int) -> float | None:
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc (the
    L2-to-fabric read and write request counts by request size, an
    accounting a 1 GiB copy reproduces exactly), committed under profiles/
    by the profiling runs; None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        return float(data[f"{kernel}@{nx}"]["bytes_per_launch"])
    except (KeyError, ValueError, TypeError):
        return None


def cycle_traffic(nx: int, s_per_cycle: float) -> dict | None:
    """The HBM bytes of a whole cycle as measured (the three phase kernels'
    request-count traffic from profiles/traffic.json, fused build), per zone
    and as a rate over the measured cycle time; None when not measured."""
    per = [load_traffic(k, nx) for k in ("k_force_sweep", "k_kinematics_q", "k_q_eos_dt")]
    if any(t is None for t in per):
        return None
    total = sum(per)
    gbs = total / s_per_cycle / 1e9
    return {"bytes_per_cycle": total, "bytes_per_zone_cycle": total / nx ** 3,
            "achieved_gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}


def bench_batched(lulesh_amd, plain_value: float) -> dict:
    """The ensemble form on its workload: 64 Sedov members, nx uniform in
    [16, 64] and origin-energy factors uniform in [0.8, 1.2] (splitmix64,
    seed 0), LULESH 2.0 ICs with 11 regions, every member run to t = 1e-2
    with its own dt.  zones*cycles counts each member's own cycles only.
    Beside it: the same members run one after another through the
    single-domain path, the way N separate runs would."""
    nx, ef = lulesh_amd.draw_ensemble()
    t0 = time.perf_counter()
    b = lulesh_amd.Batch(nx, ef)
    t_build = time.perf_counter() - t0
    print(f"bench: ensemble built in {t_build:.1f} s", file=sys.stderr, flush=True)
    # the default launch shape (the autotuner's pick on this ensemble, set
    # in the library), no tuning at run time
    shape = b.shape()
    t0 = time.perf_counter()
    rc = b.run()
    wall = time.perf_counter() - t0
    zc = b.zone_cycles()
    cycles = [b.stats(i)[1].cycle for i in range(b.n)]
    b.close()
    print(f"bench: ensemble ran in {wall:.2f} s", file=sys.stderr, flush=True)
    # the autotuner on a fresh copy of the ensemble: its time and its pick
    b = lulesh_amd.Batch(nx, ef)
    t0 = time.perf_counter()
    tuned_ms = b.autotune(4)
    t_tune = time.perf_counter() - t0
    tuned_shape = b.shape()
    b.close()
    print(f"bench: autotuned in {t_tune:.2f} s: {tuned_shape}", file=sys.stderr, flush=True)
    zc_plain, wall_plain = 0, 0.0
    for n, f in zip(nx, ef):
        d = lulesh_amd.Domain(int(n), efactor=float(f))
        t0 = time.perf_counter()
        d.run()
        wall_plain += time.perf_counter() - t0
        zc_plain += int(n) ** 3 * d.stats().cycle
        d.close()
    print(f"bench: single-domain runs took {wall_plain:.2f} s", file=sys.stderr,
          flush=True)
    # the same in the headline's arithmetic (fused multiply-adds and refined
    # reciprocals), ensemble and single-domain path alike
    b = lulesh_amd.Batch(nx, ef, fused=True)
    t0 = time.perf_counter()
    rc_f = b.run()
    wall_f = time.perf_counter() - t0
    zc_f = b.zone_cycles()
    b.close()
    wall_plain_f = 0.0
    for n, f in zip(nx, ef):
        d = lulesh_amd.Domain(int(n), efactor=float(f), fused=True)
        t0 = time.perf_counter()
        d.run()
        wall_plain_f += time.perf_counter() - t0
        d.close()
    print(f"bench: fused ensemble ran in {wall_f:.2f} s, its single-domain runs "
          f"took {wall_plain_f:.2f} s", file=sys.stderr, flush=True)
    value = zc / wall
    return {
        "value": value, "unit": "zones·cycles/s", "rc": rc,
        "workload": "ensemble of 64 Sedov members, nx~U{16..64}, "
                    "efactor~U[0.8,1.2], splitmix64 seed 0, LULESH 2.0 ICs, "
                    "11 regions, each to t=1e-2",
        "members": int(len(nx)), "zones_total": int(sum(int(n) ** 3 for n in nx)),
        "zone_cycles": zc, "member_cycles": [int(min(cycles)), int(max(cycles))],
        "wall_s": wall, "build_s": t_build, "shape": shape,
        "autotune": {"s": t_tune, "ms_per_cycle": tuned_ms, "shape": tuned_shape,
                     "same_as_default": tuned_shape == shape},
        "same_sizes_single_domain_path": {"value": zc_plain / wall_plain,
                                          "wall_s": wall_plain},
        "vs_single_domain_path": value / (zc_plain / wall_plain),
        # as if the autotuner had run before the ensemble
        "vs_single_domain_path_with_autotune": wall_plain / (wall + t_tune),
        "vs_plain_150": value / plain_value,
        "fused": {"value": zc_f / wall_f, "rc": rc_f, "wall_s": wall_f,
                  "zone_cycles": zc_f,
                  "same_sizes_single_domain_path": {"value": zc_f / wall_plain_f,
                                                    "wall_s": wall_plain_f},
                  "vs_single_domain_path": wall_plain_f / wall_f,
                  "vs_plain_150": (zc_f / wall_f) / plain_value},
    }


# Algorithmic bytes of the unstructured step's kernels (lulesh_create_mesh)
# per cycle, from how the force kernel cuts the mesh (Domain.force_blocks):
# - k_forces_u_blocks: per element the nodelist 32, its corners' slots 32,
#   p, q, v, volo, ss, elemMass 48; x..zd 48 per node; stress and hourglass
#   forces 48 per boundary corner written (each padded to two 32-byte
#   sectors: 64 move); per interior node its corner-list start 4, its local
#   corner list 2 per corner, mass 8, symmetry byte 1, x..zd written 48;
# - k_nodes_u: per boundary node its corner range 4, mass 8, symmetry byte
#   1, x..zd read 48 and written 48; its corners' forces 48 each;
# - the kinematics with the nodelist (176 + 32 per element, 96 per node) and
#   Q/EOS/dt with the face neighbours, face masks and EOS schedule instead of
#   the region index (264 - 12 + 24 + 4 + 4 + 2 per element).
# With every node a boundary node this is the corner-array scheme's 1374 B
# per element + 253 B per node.
def unstructured_bytes(ne: int, nn: int, fb: dict) -> dict:
    ni, nb, cb = fb["interior_nodes"], fb["boundary_nodes"], fb["boundary_corners"]
    ci = 8 * ne - cb  # corners of interior nodes
    return {
        "k_forces_u_blocks": 112 * ne + 48 * nn + 48 * cb + 61 * ni + 2 * ci,
        "k_nodes_u": 109 * nb + 48 * cb,
        "k_kinematics_q": 208 * ne + 96 * nn,
        "k_q_eos_dt": 286 * ne,
    }


def cube_lists(nx: int) -> dict:
    """The Sedov cube nx^3 as explicit mesh lists (node coordinates,
    nodelist, face neighbours, symmetry and free-surface node sets), the
    form a .lmesh file holds."""
    import numpy as np
    en = nx + 1
    coord = np.array([0.0] + [1.125 * c / nx for c in range(1, en)])
    k, j, i = (a.ravel() for a in np.meshgrid(np.arange(en), np.arange(en), np.arange(en),
                                                indexing="ij"))
    ek, ej, ei = (a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx),
                                                   indexing="ij"))
    n0 = (ek * en + ej) * en + ei
    pl = en * en
    nodelist = np.stack([n0, n0 + 1, n0 + en + 1, n0 + en, n0 + pl, n0 + pl + 1,
                         n0 + pl + en + 1, n0 + pl + en], axis=1).astype(np.int32)
    el = np.arange(nx ** 3)
    nb = np.stack([np.where(ei > 0, el - 1, el), np.where(ei < nx - 1, el + 1, el),
                   np.where(ej > 0, el - nx, el), np.where(ej < nx - 1, el + nx, el),
                   np.where(ek > 0, el - nx * nx, el),
                   np.where(ek < nx - 1, el + nx * nx, el)], axis=1).astype(np.int32)
    nid = np.arange(en ** 3, dtype=np.int32)
    return {"x": coord[i], "y": coord[j], "z": coord[k], "nodelist": nodelist,
            "neighbors": nb, "symm_x": nid[i == 0], "symm_y": nid[j == 0],
            "symm_z": nid[k == 0], "free_surf": nid[(i == nx) | (j == nx) | (k == nx)]}


def bench_unstructured(lulesh_amd, nx: int = 150, steps: int = 50, warmup: int = 5) -> dict:
    """The unstructured step (explicit connectivity, element blocks with
    on-chip corner forces for their interior nodes) on the Sedov cube nx^3
    given as mesh lists, LULESH 1.0 ICs,
    beside the structured step on the same cube: zones*cycles/s over
    `steps` cycles and the roofline of each unstructured kernel; then both
    steps again with fused multiply-adds."""
    ne, nn = nx ** 3, (nx + 1) ** 3
    kt = 20
    out = {}
    lists = cube_lists(nx)
    for name in ("unstructured", "structured", "unstructured_fused", "structured_fused"):
        if name.startswith("unstructured"):
            d = lulesh_amd.Domain.from_mesh(lists, max_cycles=warmup + steps + kt)
        else:
            d = lulesh_amd.Domain(nx, ic_mode=1, max_cycles=warmup + steps + kt)
        d.set_fused_multiply_add(name.endswith("_fused"))
        d.step(warmup)
        t0 = time.perf_counter()
        rc, n = d.step(steps)
        wall = time.perf_counter() - t0
        if rc != 0 or n != steps:
            raise RuntimeError(f"{name} step returned rc={rc}, {n} of {steps} cycles")
        out[name] = {"value": ne * steps / wall, "ms_per_cycle": wall * 1e3 / steps}
        if name == "unstructured":
            d.set_kernel_timing(True)
            d.step(kt)
            times = d.kernel_times()
            fb = d.force_blocks()
            kbytes = unstructured_bytes(ne, nn, fb)
            out[name]["force_blocks"] = fb
            kern = {}
            for k, (ms, cnt) in times.items():
                if not cnt or k not in kbytes:
                    continue
                mean = ms / cnt
                b = kbytes[k]
                kern[k] = {"mean_ms": mean, "alg_bytes": b,
                           "achieved_gbs": b / (mean * 1e-3) / 1e9 if b else 0.0,
                           "frac": b / (mean * 1e-3) / 1e9 / HBM_PEAK_GBS if b else 0.0}
            out[name]["kernels"] = kern
        d.close()
        print(f"bench: {name} {nx}^3: {out[name]['ms_per_cycle']:.3f} ms/cycle",
              file=sys.stderr, flush=True)
    u = out["unstructured"]
    b_cycle = sum(unstructured_bytes(ne, nn, u["force_blocks"]).values())
    roof = HBM_PEAK_GBS * 1e9 / b_cycle * ne
    return {
        "value": u["value"], "unit": "zones·cycles/s",
        "workload": f"sedov_{nx}^3 as explicit mesh lists, LULESH 1.0 ICs, cycles "
                    f"{warmup + 1}-{warmup + steps}",
        "unstructured": u, "structured_same_cube": out["structured"],
        "vs_structured": u["value"] / out["structured"]["value"],
        "fused": {"arithmetic": "float64 with fused multiply-adds, both steps",
                  "unstructured": out["unstructured_fused"],
                  "structured_same_cube": out["structured_fused"],
                  "vs_structured": out["unstructured_fused"]["value"] /
                  out["structured_fused"]["value"]},
        "cycle_roofline": {"b_alg_per_cycle": b_cycle, "roofline_zones_cycles_per_s": roof,
                           "frac": u["value"] / roof},
    }


def bench_mixed(lulesh_amd, nx: int = 120) -> dict:
    """The mixed-precision form on its workload: Sedov at nx^3 (LULESH 2.0
    ICs, 11 regions), run to t = 1e-2 in float64 (bit-identical), in mixed
    precision and in float64 with fused multiply-adds; throughput of each
    over the whole run and the mixed run's errors against the bit-identical
    float64 one (origin energy, total energy, shock radius)."""
    out = {}
    doms = {}
    for name, mixed, fused in (("fp64", False, False), ("mixed", True, False),
                               ("fp64_fused", False, True)):
        d = lulesh_amd.Domain(nx, mixed=mixed, fused=fused)
        d.step(1)  # first launch of each kernel outside the timed run
        t0 = time.perf_counter()
        rc, n = d.run()
        wall = time.perf_counter() - t0
        st = d.stats()
        out[name] = {"value": nx ** 3 * n / wall, "cycles": st.cycle, "wall_s": wall}
        doms[name] = d
        print(f"bench: {name} {nx}^3 to t=1e-2: {st.cycle} cycles in {wall:.2f} s",
              file=sys.stderr, flush=True)
    doms.pop("fp64_fused").close()
    r, m = doms["fp64"], doms["mixed"]
    e_r, e_m = r.stats().e_origin, m.stats().e_origin
    t_r, t_m = r.total_energy(), m.total_energy()
    s_r, s_m = r.shock_radius(), m.shock_radius()
    for d in doms.values():
        d.close()
    ne, nn = nx ** 3, (nx + 1) ** 3
    b_mixed = MIXED_BYTES[0] * ne + MIXED_BYTES[1] * nn
    b_fp64 = 520 * ne + 296 * nn
    roof_mixed = HBM_PEAK_GBS * 1e9 / b_mixed * ne
    roof_fp64 = HBM_PEAK_GBS * 1e9 / b_fp64 * ne
    return {
        "value": out["mixed"]["value"], "unit": "zones·cycles/s",
        "workload": f"sedov_{nx}^3_lulesh2.0_regions11 to t=1e-2",
        "fp64": out["fp64"], "mixed": out["mixed"], "fp64_fused": out["fp64_fused"],
        # each against its own algorithmic bytes (float32 arrays at 4 B)
        "roofline": {"b_alg_per_cycle": b_mixed, "roofline_zones_cycles_per_s": roof_mixed,
                     "frac": out["mixed"]["value"] / roof_mixed,
                     "fp64_b_alg_per_cycle": b_fp64,
                     "fp64_frac": out["fp64"]["value"] / roof_fp64,
                     "fp64_fused_frac": out["fp64_fused"]["value"] / roof_fp64},
        # against the bit-identical float64 step (the errors' reference) and
        # the float64 step with fused multiply-adds
        "vs_fp64_same_workload": out["mixed"]["value"] / out["fp64"]["value"],
        "vs_fp64_fused_same_workload": out["mixed"]["value"] / out["fp64_fused"]["value"],
        "errors_vs_fp64": {
            "origin_energy_rel": abs(e_m - e_r) / abs(e_r),
            "total_energy_rel": abs(t_m - t_r) / abs(t_r),
            "shock_radius_abs": abs(s_m["radius"] - s_r["radius"]),
            "shock_radius_x_axis_abs": abs(s_m["radius_x_axis"] - s_r["radius_x_axis"]),
            "element_size": 1.125 / nx,
            "cycles": [out["fp64"]["cycles"], out["mixed"]["cycles"]],
        },
    }


def bench_learned_q(lulesh_amd, efactor: float = 1.0) -> dict:
    """The learned-Q form: the network shipped with the package
    (lulesh-accelerate_amd/learned_q_net.npz, trained by learned_q.py on
    fine 120^3 runs of other energy factors) at 40^3, beside the plain
    limiter at 40^3, both measured against the 120^3 run of the same factor
    restricted to 40^3 at t = 1e-2; and the plain limiter at 40^3 to 64^3,
    each against the 120^3 run restricted to its own grid, from which the
    plain resolution reaching the learned run's errors and its wall time are
    interpolated between measured runs (learned_q.measured_equivalent)."""
    import learned_q
    net, meta = learned_q.load()
    res = learned_q.evaluate(net, [efactor], mid_ratio=0, form=meta["form"])[f"{efactor:.4f}"]
    plain, learned = res["plain"], res["learned"]
    eq = res["learned_measured_equivalent"]
    # the stand-in form (the network in the limiter's place) against the same
    # plain runs
    snet, smeta = learned_q.load(learned_q.NET_FILE_STANDIN)
    sres = learned_q.evaluate(snet, [efactor], mid_ratio=0, sweep=(),
                              form=smeta["form"])[f"{efactor:.4f}"]
    standin = sres["learned"]
    seq = learned_q.measured_equivalent(standin, standin["wall_s"], res["plain_sweep"])
    speed = [eq[k]["wall_speedup"] for k in learned_q.METRICS]
    keep = ("density_l1", "energy_l1", "pressure_l1", "cycles", "wall_s", "zone_cycles")
    return {
        "value": learned["zones_cycles_per_s"], "unit": "zones·cycles/s",
        "workload": "sedov 40^3 (LULESH 2.0 ICs, 11 regions) to t=1e-2, learned "
                    "closure vs plain limiter, errors against the 120^3 run "
                    "restricted to each grid",
        "plain_40": plain, "learned_40": learned,
        # the 40^3 domain's arrays (about 53 MB at 520 B per element and
        # 296 B per node) stay in the GPU's 256 MB last-level cache, so an
        # HBM fraction says nothing at this size; reported for completeness
        "roofline": {"b_alg_per_cycle": 520 * 40 ** 3 + 296 * 41 ** 3,
                     "frac": learned["zones_cycles_per_s"] / (HBM_PEAK_GBS * 1e9 / (
                         520 * 40 ** 3 + 296 * 41 ** 3) * 40 ** 3),
                     "note": "working set in last-level cache: not HBM-bound"},
        "fine_120_cycles": res["fine_cycles"],
        "plain_sweep": [{k: r[k] for k in ("nx",) + keep} for r in res["plain_sweep"]],
        # per metric: the plain resolution of equal error (interpolated
        # between the two measured runs that bracket it) and that run's wall
        # time and zone*cycles over the learned 40^3 run's
        "learned_measured_equivalent": eq,
        "wall_speedup_at_equal_error": {k: eq[k]["wall_speedup"] for k in learned_q.METRICS},
        "wall_speedup_geomean": float(np.exp(np.mean(np.log(speed)))),
        "error_ratio_learned_over_plain": {
            k: learned[k] / plain[k] for k in learned_q.METRICS},
        "training": {k: meta[k] for k in ("samples", "epochs", "loss_init", "loss_final",
                                          "loss_of_plain_limiter")},
        "standin_40": {
            "errors": {k: standin[k] for k in learned_q.METRICS},
            "wall_speedup_at_equal_error": {k: seq[k]["wall_speedup"]
                                            for k in learned_q.METRICS},
            "plain_nx_at_equal_error": {k: seq[k]["plain_nx"] for k in learned_q.METRICS},
            "zones_cycles_per_s": standin["zones_cycles_per_s"],
        },
    }


def time_plain(lulesh_amd, nx: int, K: int, W: int, fused: bool) -> dict | None:
    """The plain float64 step of the benchmark workload: KR cycles that take
    the GPU from idle to its sustained clock, the per-kernel breakdown (HIP
    events around every launch on the domain's stream) over KT cycles, W
    warm-up cycles, then K timed cycles queued back to back with no events
    between the kernels.  From idle, the first 40 cycles of this domain run
    up to 10% slower (the sweep 346 us in cycles 1-10, 291 us from cycle 40
    on); a GPU kept busy beforehand runs cycles 6-25 at the steady-state
    rate.  The events cost about 1% of a cycle, so they stay out of the
    timed cycles."""
    KR, KT = 25, 25
    d = lulesh_amd.Domain(nx, ic_mode=2, num_reg=11, balance=1, cost=1,
                          max_cycles=KR + KT + W + K, fused=fused)
    rc, done = d.step(KR)
    if rc != 0 or done != KR:
        print(f"bench: step returned rc={rc}, {done} of {KR} cycles", file=sys.stderr)
        return None
    d.set_kernel_timing(True)
    rc, done = d.step(KT)
    if rc != 0 or done != KT:
        print(f"bench: breakdown step returned rc={rc}, {done} of {KT} cycles",
              file=sys.stderr)
        return None
    times = d.kernel_times()
    d.set_kernel_timing(False)
    if W:
        d.step(W)
    # the timed region: lulesh_step is synchronous, so the device is idle
    # before and after
    t0 = time.perf_counter()
    rc, done = d.step(K)
    wall = time.perf_counter() - t0
    if rc != 0 or done != K:
        print(f"bench: step returned rc={rc}, {done} of {K} cycles", file=sys.stderr)
        return None
    st = d.stats()
    d.close()
    print(f"bench: {nx}^3 {'fused' if fused else 'exact'}: {wall * 1e3 / K:.4f} ms/cycle",
          file=sys.stderr, flush=True)
    return {"wall": wall, "times": times, "stats": st, "kt": KT,
            "first_timed": KR + KT + W + 1, "breakdown": [KR + 1, KR + KT]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=150)
    ap.add_argument("--cpu-cycles", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the other forms of the step")
    ap.add_argument("--forms", default="batched,mixed,learned_q,unstructured",
                    help="comma-separated forms to run after the plain step")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not "
              "supported (the LULESH step here is a single-GPU build)",
              file=sys.stderr)
        return 2

    import lulesh_amd

    nx, K, W = args.nx, args.steps, args.warmup
    ne, nn = nx ** 3, (nx + 1) ** 3
    # the float64 step with fused multiply-adds (results within the
    # reference's tolerances, tests/test_gpu_fused.py), then the same with
    # every value rounded as the reference rounds it (bit-identical)
    runs = {}
    for fused in (True, False):
        r = time_plain(lulesh_amd, nx, K, W, fused)
        if r is None:
            return 1
        runs[fused] = r
    fz, ex = runs[True], runs[False]
    wall, times, st = fz["wall"], fz["times"], fz["stats"]
    KT = fz["kt"]
    zones_cycles = ne * K
    value = zones_cycles / wall
    # dominant kernel and its roofline
    per_kernel = {k: (ms / n if n else 0.0) for k, (ms, n) in times.items()}
    dom = max(per_kernel, key=lambda k: times[k][0])
    be, bn = KERNEL_BYTES[dom]
    alg_bytes = be * ne + bn * nn
    achieved = alg_bytes / (per_kernel[dom] * 1e-3) / 1e9
    b_cycle = 520 * ne + 296 * nn
    roof_value = HBM_PEAK_GBS * 1e9 / b_cycle * ne
    ex_value = ne * K / ex["wall"]
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "zones·cycles/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: deterministic Sedov blast initial conditions",
        "config": {"workload": f"sedov_{nx}^3_lulesh2.0_regions11_balance1_cost1",
                   "nx": nx, "zones": ne, "nodes": nn, "ic": "LULESH 2.0.3",
                   "num_reg": 11, "balance": 1, "cost": 1,
                   "arithmetic": "float64 with fused multiply-adds "
                                 "(lulesh_set_fused_multiply_add): within the "
                                 "reference's tolerances, not bit-identical",
                   "cycles_timed": K, "first_timed_cycle": fz["first_timed"],
                   "final_cycle": st.cycle},
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(dom, nx),
                     "kernel": dom, "alg_bytes_per_launch": alg_bytes,
                     "mean_ms": per_kernel[dom]},
        "cycle_roofline": {"b_alg_per_cycle": b_cycle,
                           "roofline_zones_cycles_per_s": roof_value,
                           "frac": value / roof_value},
        "cycle_traffic": cycle_traffic(nx, wall / K),
        "grind_us_per_zone_cycle": 1e6 / value,
        "kernels_mean_ms": {k: round(v, 6) for k, v in per_kernel.items()},
        "gpu_event_ms_per_step": sum(ms for ms, _ in times.values()) / KT,
        "breakdown_cycles": fz["breakdown"],
        # the same workload with every value rounded as the reference rounds
        # it: bit-identical to the reference (tests/test_gpu_parity.py)
        "exact_rounding": {
            "value": ex_value, "ms_per_step": ex["wall"] * 1e3 / K,
            "cycle_roofline_frac": ex_value / roof_value,
            "kernels_mean_ms": {k: round(ms / n, 6) for k, (ms, n) in ex["times"].items()
                                if n},
        },
    }
    forms = [] if args.no_forms else [f for f in args.forms.split(",") if f]
    unknown = set(forms) - {"batched", "mixed", "learned_q", "unstructured"}
    if unknown:
        print(f"bench: unknown forms {sorted(unknown)}", file=sys.stderr)
        return 2
    if forms:
        result["forms"] = {}
    if "batched" in forms:
        result["forms"]["batched"] = bench_batched(lulesh_amd, value)
    if "mixed" in forms:
        result["forms"]["mixed"] = bench_mixed(lulesh_amd)
        result["forms"]["mixed"]["vs_plain_150"] = (
            result["forms"]["mixed"]["value"] / value)
    if "learned_q" in forms:
        result["forms"]["learned_q"] = bench_learned_q(lulesh_amd)
    if "unstructured" in forms:
        result["forms"]["unstructured"] = bench_unstructured(lulesh_amd)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(nx, args.cpu_cycles)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())

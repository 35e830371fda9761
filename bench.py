#!/usr/bin/env python3
"""Benchmark of the MI355X Laplace FMM evaluator on one GPU.

Metric: particles/s per FMM solve (N / wall time of one device evaluation:
upward pass, P2P, M2L, downward pass) with bodies, tree and interaction lists
already resident in HBM; p50/p99 per-solve latency, rel-L2 error of potential
and acceleration vs a direct sum on sampled targets, the roofline of the
dominant kernel and a CPU baseline (exafmm-beta's own operators, one core, on
a bounded sample of the same lists).

Default workload: N = 2^23 bodies uniform in [-1,1]^3, zero-mean charges in
[-1,1], P = 12, theta = 0.4, ncrit = 64.  `value` is the FP64 solve (the
reference's precision); the mixed-precision solve of the same system (same
tree and lists) is reported beside it (`mixed_same_system`,
`mixed_over_fp64`, their rel-L2 distance), both with their errors against
the FP64 direct sum, and the same system at theta = 0.35
(`accuracy_point`), where both reach rel-L2 <= 1e-6.

  python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]
  python bench.py --copy-probe MIB   # one device copy of MIB MiB (HBM counter calibration)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "exafmm-beta_amd"))

METRIC = "particles/s per FMM solve + p50/p99 latency, mixed vs FP64 at equal L2 error"
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 matrix (MFMA)
PEAK_FP32_VALU_TFLOPS = 157.3  # MI355X FP32 vector (the packed-FP32 rate)
PEAK_FP16_TFLOPS = 2500.0  # MI355X FP16 dense MFMA
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # N = 2^23 uniform cube, P = 12: the FP64 solve (`value`) and the mixed
    # solve of the same system (same tree and lists) beside it, plus the same
    # system at theta = 0.35, where both reach rel-L2 <= 1e-6 against the
    # direct sum
    "cube8m-p12": dict(n=8388608, dist="cube1", P=12, theta=0.4, ncrit=64, nspawn=5000,
                       ntarget=1000, precision="fp64", compare_mixed=True, accurate_theta=0.35,
                       desc="N=8388608 cube [-1,1]^3, zero-mean q in [-1,1], P=12, theta=0.4, ncrit=64"),
    "cube8m-p8": dict(n=8388608, dist="cube1", P=8, theta=0.4, ncrit=64, nspawn=5000,
                      ntarget=1000, precision="fp64", compare_mixed=True,
                      desc="N=8388608 cube [-1,1]^3, zero-mean q in [-1,1], P=8, theta=0.4, ncrit=64"),
    # N = 2^20 Plummer, FP64, P = 10
    "plummer1m-p10-fp64": dict(n=1 << 20, dist="plummer1", P=10, theta=0.4, ncrit=64, nspawn=5000,
                               ntarget=1000, precision="fp64",
                               desc="N=2^20 Plummer (a=1, r<10), q=1/N, P=10, theta=0.4, ncrit=64"),
    # N = 32768 uniform cube [-1,1]^3, zero-mean q in [-1,1]: CPU-runnable case
    "cube32k-p10-fp64": dict(n=32768, dist="cube1", P=10, theta=0.4, ncrit=64, nspawn=5000,
                             ntarget=100, precision="fp64",
                             desc="N=32768 cube [-1,1]^3, zero-mean q in [-1,1], P=10, theta=0.4"),
}
DEFAULT_WORKLOAD = "cube8m-p12"

# Serving stream: independent systems, N = floor(2^U[12,18]), uniform cube or
# Plummer with equal probability, zero-mean uniform charges, P = 8, theta =
# 0.4, mixed precision, Poisson arrivals at `rate` requests/s.
SERVE = {
    "serve-200": dict(rate=200.0, requests=2048, P=8, theta=0.4, ncrit=64, nspawn=5000,
                      precision="mixed", log2n=(12, 18)),
    "serve-1000": dict(rate=1000.0, requests=2048, P=8, theta=0.4, ncrit=64, nspawn=5000,
                       precision="mixed", log2n=(12, 18)),
}

# Periodic MD stream: a cubic box of side 32 holding 32768 three-site water
# molecules (O -0.834, H +0.417; O-H 0.308 box units = 0.9572 A with the box
# taken as 99.4 A, the side of 32768 waters at liquid density; H-O-H 104.52
# deg; oxygen positions uniform, orientations random), images = 3, P = 10,
# theta = 0.35; successive steps displace every site by N(0, (0.002 L)^2).
MD = {
    "md-water98k-p10": dict(molecules=32768, box=32.0, bond=0.308, angle_deg=104.52, images=3, P=10,
                            theta=0.35, ncrit=64, nspawn=5000, sigma=0.002, requests=1000),
}


def water_box(w, seed=0):
    rng = np.random.default_rng(seed)
    m, L = w["molecules"], w["box"]
    o = (rng.random((m, 3)) - 0.5) * L
    u1 = rng.normal(size=(m, 3))
    u1 /= np.linalg.norm(u1, axis=1, keepdims=True)
    v = rng.normal(size=(m, 3))
    v -= (v * u1).sum(1, keepdims=True) * u1
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    a = math.radians(w["angle_deg"])
    u2 = math.cos(a) * u1 + math.sin(a) * v
    x = np.stack([o, o + w["bond"] * u1, o + w["bond"] * u2], 1).reshape(-1, 3)
    q = np.tile([-0.834, 0.417, 0.417], m)
    return np.ascontiguousarray(x), q


def md_bench(F, name, nreq, log):
    """Successive MD steps through fmm_md_evaluate (wrap, tree, lists, periodic
    FMM, dipole correction; results in the caller's order): per-step latency
    includes the host plan, as every step moves every site."""
    w = MD[name]
    nreq = nreq or w["requests"]
    x0, q = water_box(w)
    n = len(q)
    rng = np.random.default_rng(1)
    out = {}
    for prec in ("fp64", "mixed"):
        f = F.FMM(P=w["P"], theta=w["theta"], ncrit=w["ncrit"], nspawn=w["nspawn"], images=w["images"],
                  cycle=w["box"], precision=F.FMM_MIXED if prec == "mixed" else F.FMM_FP64)
        x = x0.copy()
        f.md_evaluate(x, q)  # warm-up
        lat, dev, host = [], [], []
        first = None
        t0 = time.perf_counter()
        for k in range(nreq):
            a = time.perf_counter()
            p, fo = f.md_evaluate(x, q)
            lat.append(time.perf_counter() - a)
            s = f.stats()
            dev.append(s["dev_total_ms"])
            host.append(s["host_bounds_ms"] + s["host_tree_ms"] + s["host_lists_ms"] + s["host_upload_ms"])
            if first is None:
                first = (p.copy(), fo.copy())
            x += rng.normal(scale=w["sigma"] * w["box"], size=x.shape)
        wall = time.perf_counter() - t0
        f.close()
        lat = np.array(lat)
        out[prec] = {"particles_per_s": n * nreq / wall, "ms_per_step": 1e3 * wall / nreq,
                     "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                    "p99": 1e3 * float(np.percentile(lat, 99))},
                     "device_ms_mean": float(np.mean(dev)), "host_plan_ms_mean": float(np.mean(host)),
                     "energy_first_step": float(q @ first[0]), "_first": first}
        log("[md] %s %s: %d steps, %.2f ms/step (device %.2f ms, host plan %.2f ms), p99 %.2f ms"
            % (name, prec, nreq, out[prec]["ms_per_step"], out[prec]["device_ms_mean"],
               out[prec]["host_plan_ms_mean"], out[prec]["latency_ms"]["p99"]))
    p64, f64 = out["fp64"].pop("_first")
    p32, f32 = out["mixed"].pop("_first")
    # validation of the first step against the Ewald sum (device), alpha and
    # ksize chosen so that the real and reciprocal truncations are below 1e-7:
    # erfc(alpha * cutoff) < 1e-7 with cutoff = L/2, exp(-K^2 / 4 alpha^2) < 1e-7
    Lb = w["box"]
    alpha = 3.9 / (Lb / 2)
    ksize = int(math.ceil(2 * alpha * math.sqrt(math.log(1e7)) * Lb / (2 * math.pi)))
    f = F.FMM(P=w["P"], images=w["images"], cycle=Lb)
    t0 = time.perf_counter()
    pe, fe = f.ewald(x0, q, ksize, alpha, 0.25 / math.pi, Lb / 2)
    t_ew = time.perf_counter() - t0
    f.close()
    ee = float(q @ pe)
    out["ewald_first_step"] = {
        "ksize": ksize, "alpha": alpha, "cutoff": Lb / 2, "seconds": t_ew, "energy": ee,
        "fp64": {"energy_rel": abs(q @ p64 - ee) / abs(ee),
                 "acc_rel_l2": float(np.linalg.norm(f64 - fe) / np.linalg.norm(fe))},
        "mixed": {"energy_rel": abs(q @ p32 - ee) / abs(ee),
                  "acc_rel_l2": float(np.linalg.norm(f32 - fe) / np.linalg.norm(fe))},
        "note": "the reference's periodic FMM leaves the image shell at distance 2 out at theta <= 0.5"}
    log("[md] Ewald (ksize %d, alpha %.4f, %.2f s): FMM fp64 energy %.2e acc %.2e"
        % (ksize, alpha, t_ew, out["ewald_first_step"]["fp64"]["energy_rel"],
           out["ewald_first_step"]["fp64"]["acc_rel_l2"]))
    out["mixed_vs_fp64_first_step"] = {
        "energy_rel": abs(q @ p32 - q @ p64) / abs(q @ p64),
        "acc_rel_l2": float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))}
    out["config"] = {"workload": name, "sites": n, **{k: v for k, v in w.items() if k != "requests"},
                     "requests": nreq}
    return out


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB through a copy kernel: the known
    byte count that FETCH_SIZE / WRITE_SIZE of a profiled run are scaled by."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("[copy-probe] copied %d MiB (read %d B, write %d B)" % (mib, 4 * n, 4 * n), flush=True)


def traffic_from_profile(workload, kernel):
    """HBM bytes per launch of `kernel` for `workload`, from the committed
    rocprofv3 --pmc summary (profiles/<kernel>_<workload>_traffic.md, a line
    'bytes_per_launch: <fetch> <write>'), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.md" % (kernel, workload))
    try:
        for line in open(path):
            if line.startswith("bytes_per_launch:"):
                fetch, write = (float(v) for v in line.split(":", 1)[1].split())
                return fetch + write
    except (OSError, ValueError):
        pass
    return None


def m2l_flops(P, pairs):
    """Algorithmic M2L work: one dense complex NTERM x P^2 translation per pair."""
    return 8.0 * (P * (P + 1) // 2) * P * P * pairs


def m2l_exec_flops(P, pairs):
    """M2L work as the kernels execute it: one dense real P^2 x P^2 operator
    applied per pair (2 P^4 flop), without the padding of partly filled
    slices or chunks."""
    return 2.0 * P ** 4 * pairs


def p2p_flops(body_pairs):
    """Algorithmic P2P work: 20 flops per body pair (potential + gradient)."""
    return 20.0 * body_pairs


def rel_l2(ref, got):
    """Verify::getDif*/getNrm* ratios (include/verify.h:42-92)."""
    pd = np.sum((ref[:, 0] - got[:, 0]) ** 2)
    pn = np.sum(ref[:, 0] ** 2)
    ad = np.sum((ref[:, 1:] - got[:, 1:]) ** 2)
    an = np.sum(ref[:, 1:] ** 2)
    return math.sqrt(pd / pn), math.sqrt(ad / an)


def cpu_baseline(x, q, w, plan, budget_s):
    """exafmm-beta's own Kernel operators (oracle/_ref, built from the
    reference sources) timed on one host core on a stride sample of every
    pass of the same solve -- P2M and L2P of every s-th leaf, M2M of every
    s-th internal cell, L2L of every s-th cell, and every k-th pair of the
    M2L and P2P lists -- each phase scaled by its share of the whole work.
    The tree build and the traversal, which the reference's "Total FMM"
    timer also covers, are timed whole on one core with the oracle's C
    restatement (the reference's tree and traversal headers need its
    generated config.h; the restatement makes the same tree and pairs).
    Falls back to the oracle's C restatement (whole upward/downward passes)
    when _ref is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    tr = plan.tree()
    xs, qs = x[tr["perm"]], q[tr["perm"]]
    mp, ms, pp, ps = plan.lists()
    nc = len(tr["nbody"])
    mi = np.repeat(np.arange(nc, dtype=np.int32), np.diff(mp))
    pi = np.repeat(np.arange(nc, dtype=np.int32), np.diff(pp))
    body_pairs = tr["nbody"][pi].astype(np.int64) * tr["nbody"][ps]
    P = w["P"]
    n = len(x)
    # per-unit cost estimates on one core (M2L ~17 us at P=10 scaling as P^4,
    # P2P ~3.4 ns per body pair, a pass ~2 us per body at P=10 scaling as P^2)
    est_m2l = 17e-6 * (P / 10.0) ** 4 * len(mi)
    est_p2p = 3.4e-9 * float(body_pairs.sum())
    est_pass = 2 * 2e-6 * (P / 10.0) ** 2 * n + 2 * 20e-6 * (P / 10.0) ** 4 * nc
    s_m2l = max(1, int(math.ceil(est_m2l / (0.5 * budget_s))))
    s_p2p = max(1, int(math.ceil(est_p2p / (0.25 * budget_s))))
    s_pass = max(1, int(math.ceil(est_pass / (0.25 * budget_s))))
    sm, sp = np.arange(0, len(mi), s_m2l), np.arange(0, len(pi), s_p2p)
    if O.ref_available():
        t, work = O.ref_time_sampled(O.Kernel(P, use_ref=True), tr, xs, qs, s_pass, (mi[sm], ms[sm]),
                                     (pi[sp], ps[sp]))
        full = [float(n), float(tr["nchild"].sum()), float(len(mi)), float(body_pairs.sum()),
                float(nc - 1), float(n)]
        phase = {k: t[i] * full[i] / max(work[i], 1.0)
                 for i, k in enumerate(("p2m", "m2m", "m2l", "p2p", "l2l", "l2p"))}
        t0 = time.perf_counter()
        otree = O.Tree(x, w["ncrit"])
        t1 = time.perf_counter()
        otree.traverse(w["theta"], w["nspawn"])
        t2 = time.perf_counter()
        del otree
        phase.update(tree=t1 - t0, lists=t2 - t1)
        total = sum(phase.values())
        kind, sampled = "reference", float(t.sum()) + (t2 - t0)
    else:
        tree = O.Tree(x, w["ncrit"])
        k = O.Kernel(P)
        t0 = time.perf_counter()
        M = tree.upward(k, xs, qs)
        t1 = time.perf_counter()
        trg, Lc = tree.eval_lists(k, xs, qs, mi[sm], ms[sm], np.zeros(len(sm), np.int32), M)
        t2 = time.perf_counter()
        tree.eval_lists(k, xs, qs, pi[sp], ps[sp], np.ones(len(sp), np.int32), M, trg=trg)
        t3 = time.perf_counter()
        tree.downward(k, xs, Lc, trg)
        t4 = time.perf_counter()
        phase = {"upward": t1 - t0, "m2l": (t2 - t1) * len(mi) / max(len(sm), 1),
                 "p2p": (t3 - t2) * float(body_pairs.sum()) / max(float(body_pairs[sp].sum()), 1.0),
                 "downward": t4 - t3}
        total = sum(phase.values())
        kind, sampled = "port", t4 - t0
    return {
        "value": n / total,
        "unit": "particles/s",
        "cores": 1,
        "kind": kind,
        "sample": ("%s: every %d-th cell of each pass (P2M, M2M, L2L, L2P), every %d-th of %d M2L "
                   "pairs and every %d-th of %d P2P pairs timed on one core and each phase "
                   "scaled to the whole solve%s (%.1f s timed; est. %.1f s per solve at N=%d)"
                   % (w["desc"], s_pass, s_m2l, len(mi), s_p2p, len(pi),
                      "; the tree build and traversal timed whole with the oracle's C restatement"
                      if "tree" in phase else "", sampled, total, n)),
        "n": n,
        "phase_s": phase,
    }


def solve_and_time(F, w, plan, x, q, precision, steps, warmup, log, m2l_mode=0, theta=None):
    """Upload one system (host plan shared), run `warmup` + `steps` solves, and
    return timings, kernel figures and the accuracy on sampled targets."""
    n, P = w["n"], w["P"]
    prec = {"fp64": F.FMM_FP64, "mixed": F.FMM_MIXED, "mixed16": F.FMM_MIXED16}[precision]
    f = F.FMM(P=P, theta=theta or w["theta"], ncrit=w["ncrit"], nspawn=w["nspawn"], m2l_mode=m2l_mode,
              precision=prec)
    f.set_bodies_plan(plan, q)
    for _ in range(warmup):
        f.evaluate()
    f.synchronize()
    lat, dev, km, kp = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        f.evaluate()           # returns after the last kernel completes
        lat.append(time.perf_counter() - a)
        s = f.stats()
        dev.append(s["dev_total_ms"])
        km.append(s["dev_m2l_kernel_ms"])
        kp.append(s["dev_p2p_kernel_ms"])
    f.synchronize()
    wall = time.perf_counter() - t0
    st = f.stats()
    log("[bench] %s %s: %d steps, %.3f ms/step (device %.3f ms: up %.3f p2p %.3f m2l %.3f down %.3f)"
        % (w["desc"].split(",")[0], precision, steps, 1e3 * wall / steps, np.mean(dev),
           st["dev_upward_ms"], st["dev_p2p_ms"], st["dev_m2l_ms"], st["dev_downward_ms"]))
    # accuracy: Traversal::direct (FP64, IEEE sqrt/div) on the GPU at
    # Dataset::sampleBodies-style tree-order targets
    trg = f.targets(input_order=False)
    perm = f.tree()["perm"]
    idx = np.arange(w["ntarget"]) * (n // w["ntarget"])
    d = f.direct(x[perm][idx])
    pot_rel, acc_rel = rel_l2(d, trg[idx])
    f.close()
    m2l_ms, p2p_ms = float(np.mean(km)), float(np.mean(kp))
    fm = m2l_flops(P, st["m2l_pairs"])
    fe = m2l_exec_flops(P, st["m2l_pairs"])
    fp = p2p_flops(st["p2p_body_pairs"])
    # mixed16: three FP16 products (hi x hi, hi x lo, lo x hi) per dense term
    peak = {"fp64": PEAK_FP64_TFLOPS, "mixed": PEAK_FP32_TFLOPS, "mixed16": PEAK_FP16_TFLOPS}[precision]
    if precision == "mixed16":
        fe = 3 * fe
    return {
        "wall_s": wall, "lat": lat, "dev": dev, "st": st, "trg": trg, "perm": perm,
        "accuracy": {"pot_rel_l2": pot_rel, "acc_rel_l2": acc_rel, "targets": int(w["ntarget"]),
                     "reference": "FP64 direct sum (Traversal::direct) on the GPU, IEEE sqrt/div"},
        "kernels": {
            "m2l": {"avg_ms": m2l_ms, "alg_flops": fm, "tflops": fm / m2l_ms * 1e-9,
                    "peak_tflops": peak, "frac": fm / m2l_ms * 1e-9 / peak,
                    "exec_flops": fe, "exec_tflops": fe / m2l_ms * 1e-9,
                    "frac_executed": fe / m2l_ms * 1e-9 / peak},
            "p2p": {"avg_ms": p2p_ms, "alg_flops": fp, "tflops": fp / p2p_ms * 1e-9,
                    "peak_tflops": PEAK_FP64_TFLOPS if precision == "fp64" else PEAK_FP32_VALU_TFLOPS,
                    "frac": fp / p2p_ms * 1e-9 / (PEAK_FP64_TFLOPS if precision == "fp64"
                                                   else PEAK_FP32_VALU_TFLOPS)},
        },
    }


def graph_small(F, log, steps=50):
    """Serving-sized systems (N = 2^14 and 2^16 uniform cube, P = 8, mixed):
    per-solve latency of the eager launch sequence against one replay of the
    solve captured as a HIP graph (fmm_evaluate_graph), same system."""
    out = {}
    for lg in (14, 16):
        n = 1 << lg
        x, q = F.make_bodies(n, "cube1", lg)
        f = F.FMM(P=8, precision=F.FMM_MIXED)
        f.set_bodies(x, q)
        res = {}
        for mode in ("eager", "graph"):
            for _ in range(3):
                f.evaluate(graph=mode == "graph")
            lat = []
            for _ in range(steps):
                a = time.perf_counter()
                f.evaluate(graph=mode == "graph")
                lat.append(time.perf_counter() - a)
            res[mode] = {"p50_ms": 1e3 * float(np.percentile(lat, 50)),
                         "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                         "device_ms": f.stats()["dev_total_ms"]}
            res[mode + "_targets"] = f.targets(input_order=False)
        d = rel_l2(res.pop("eager_targets"), res.pop("graph_targets"))
        res["graph_vs_eager_rel_l2"] = {"pot": d[0], "acc": d[1]}
        f.close()
        log("[bench] graph N=2^%d: eager p50 %.3f ms, graph p50 %.3f ms" % (lg, res["eager"]["p50_ms"],
                                                                            res["graph"]["p50_ms"]))
        out["n=2^%d" % lg] = res
    return out


def serve_bench(F, name, nreq, check_every, log, plan_on_device=0, m2l_mode=0, slots=8, host_threads=0,
                batch_requests=1):
    """Poisson arrivals into the request loop; latency = submit -> result on
    the server's clock; throughput over the whole stream; rel-L2 vs the FP64
    direct sum on 100 sampled targets of every `check_every`-th request.
    plan_on_device: the builders build trees and lists on the device."""
    w = SERVE[name]
    nreq = nreq or w["requests"]
    rng = np.random.default_rng(0)
    lo, hi = w["log2n"]
    sizes = np.floor(2.0 ** rng.uniform(lo, hi, nreq)).astype(np.int64)
    kinds = rng.integers(0, 2, nreq)
    gaps = rng.exponential(1.0 / w["rate"], nreq)
    systems = [F.make_bodies(int(n), "cube1" if k == 0 else "plummer1", 1000 + i)
               for i, (n, k) in enumerate(zip(sizes, kinds))]
    log("[serve] %s: %d requests, %d bodies, mean N %.0f, rate %.0f/s"
        % (name, nreq, int(sizes.sum()), sizes.mean(), w["rate"]))
    prec = F.FMM_MIXED if w["precision"] == "mixed" else F.FMM_FP64
    srv = F.Server(P=w["P"], theta=w["theta"], ncrit=w["ncrit"], nspawn=w["nspawn"], precision=prec,
                   m2l_mode=m2l_mode, slots=slots, host_threads=host_threads, batch_bodies=1 << 20,
                   batch_requests=batch_requests, plan_on_device=plan_on_device)
    for x, q in systems[:8]:   # warm-up: contexts allocate, kernels load
        srv.wait(srv.submit(x, q))
    base = srv.stats()
    ids = []
    t0 = time.perf_counter()
    due = t0
    for (x, q), gap in zip(systems, gaps):
        due += gap
        d = due - time.perf_counter()
        if d > 0:
            time.sleep(d)
        ids.append(srv.submit(x, q))
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    infos = [r[1] for r in res]
    lat = np.array([i["t_done_ms"] - i["t_submit_ms"] for i in infos])
    span_s = 1e-3 * (max(i["t_done_ms"] for i in infos) - min(i["t_submit_ms"] for i in infos))
    errs = []
    for k in range(0, nreq, check_every):
        x, q = systems[k]
        f = F.FMM(P=w["P"], theta=w["theta"], ncrit=w["ncrit"], nspawn=w["nspawn"])
        f.set_bodies(x, q)
        idx = np.arange(100) * (len(x) // 100)
        d = f.direct(x[idx])
        f.close()
        errs.append(rel_l2(d, res[k][0][idx]))
    errs = np.array(errs)
    batches = st["batches"] - base["batches"]
    out = {
        "value": float(sizes.sum()) / span_s, "unit": "particles/s",
        "requests_per_s": nreq / span_s, "offered_rate": w["rate"], "requests": nreq,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "plan_wait_ms_p50": float(np.percentile([i["t_planned_ms"] - i["t_submit_ms"] for i in infos], 50)),
        "batches": batches, "mean_batch_requests": nreq / max(batches, 1),
        "max_batch_requests": st["max_batch_requests"],
        "device_busy_frac": (st["device_ms"] - base["device_ms"]) / (1e3 * span_s),
        "accuracy": {"checked_requests": len(errs), "pot_rel_l2_max": float(errs[:, 0].max()),
                     "acc_rel_l2_max": float(errs[:, 1].max()),
                     "pot_rel_l2_median": float(np.median(errs[:, 0])),
                     "acc_rel_l2_median": float(np.median(errs[:, 1])),
                     "reference": "FP64 direct sum on 100 sampled targets per checked request"},
        "config": {"workload": name, "P": w["P"], "theta": w["theta"], "ncrit": w["ncrit"],
                   "precision": w["precision"], "log2n": list(w["log2n"]),
                   "inputs": "N=floor(2^U[12,18]), cube [-1,1]^3 or Plummer (p=1/2), "
                             "zero-mean q; Poisson arrivals"},
    }
    log("[serve] %.1f req/s, %.3g particles/s, latency p50 %.1f ms p99 %.1f ms, %d batches"
        % (out["requests_per_s"], out["value"], out["latency_ms"]["p50"], out["latency_ms"]["p99"], batches))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD,
                    choices=sorted(WORKLOADS) + sorted(SERVE) + sorted(MD))
    ap.add_argument("--requests", type=int, default=0,
                    help="serve-*: number of requests (default 2048); md-*: number of steps (default 1000)")
    ap.add_argument("--m2l-mode", type=int, default=0, choices=(0, 1, 2, 3),
                    help="M2L execution: 0 auto, 1 per target, 2 grouped chunks, 3 target tiles")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy-point", action="store_true",
                    help="skip the smaller-theta solve that reaches rel-L2 <= 1e-6")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the 2048-request serving streams at 200 and 1000 req/s")
    ap.add_argument("--no-device-plan", action="store_true",
                    help="skip the plan whose tree and lists are built on the device (fmm_plan_build_device)")
    ap.add_argument("--no-graph-small", action="store_true",
                    help="skip the eager-vs-graph latency of serving-sized systems")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build evaluates one system on one GPU; multi-GPU runs are not "
              "supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    import exafmm_amd as F
    if args.workload in MD:
        r = md_bench(F, args.workload, args.requests, log)
        out = {"metric": METRIC, "value": r["fp64"]["particles_per_s"], "unit": "particles/s",
               "n_gpus": 1, "steps": r["config"]["requests"], "warmup": 1,
               "ms_per_step": r["fp64"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": r.pop("config"), **r}
        print(json.dumps(out))
        return
    if args.workload in SERVE:
        r = serve_bench(F, args.workload, args.requests, 64, log)
        out = {"metric": METRIC, "value": r.pop("value"), "unit": r.pop("unit"), "n_gpus": 1,
               "steps": r["requests"], "warmup": 8, "ms_per_step": r["latency_ms"]["mean"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32+f64" if SERVE[args.workload]["precision"] == "mixed" else "f64",
               "data": "synthetic", "config": r.pop("config"), **r}
        print(json.dumps(out))
        return
    w = WORKLOADS[args.workload]
    n, P = w["n"], w["P"]
    x, q = F.make_bodies(n, w["dist"], 0)
    t_plan = time.perf_counter()
    plan = F.Plan(x, ncrit=w["ncrit"], theta=w["theta"], nspawn=w["nspawn"])
    t_plan = time.perf_counter() - t_plan
    st0 = plan.stats()
    log("[bench] %s: ncell=%d depth=%d m2l=%d p2p=%d body_pairs=%d; host plan %.2f s"
        % (args.workload, st0["ncell"], st0["depth"], st0["m2l_pairs"], st0["p2p_pairs"],
           st0["p2p_body_pairs"], t_plan))
    dplan_ms = None
    if not args.no_device_plan:
        # the same plan with the tree, the lists and (FP64) the M2L groups and
        # tiles built on the device: the same tree, pairs and groups
        # (tests/test_gpu_device_plan.py), each target's list in source
        # order; then uploaded into an FP64 context, timed as a whole.  The
        # first build in a process also loads the kernels' code (first
        # launches): it is timed apart, the second build is the figure
        fd = F.FMM(P=P, theta=w["theta"], ncrit=w["ncrit"], nspawn=w["nspawn"])
        t0 = time.perf_counter()
        fd.set_bodies_device(x, q)
        fd.synchronize()
        first_ms = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        fd.set_bodies_device(x, q)   # fmm_set_bodies_device: plan on the device, uploaded without a copy
        fd.synchronize()
        t1 = time.perf_counter()
        sd = fd.stats()
        fd.close()
        dplan_ms = {"bounds_ms": sd["host_bounds_ms"], "tree_ms": sd["host_tree_ms"],
                    "lists_ms": sd["host_lists_ms"], "groups_tiles_ms": sd["plan_groups_ms"],
                    "upload_ms": sd["host_upload_ms"], "total_ms": 1e3 * (t1 - t0), "first_build_total_ms": first_ms,
                    "host_tree_ms": st0["host_tree_ms"], "host_lists_ms": st0["host_lists_ms"],
                    "same_counts": all(sd[k] == st0[k] for k in ("ncell", "m2l_pairs", "p2p_pairs",
                                                                 "p2p_body_pairs"))}
        log("[bench] plan on the device: tree %.1f ms (host %.1f), lists %.1f ms (host %.1f), groups and tiles "
            "%.1f ms, upload %.1f ms; %.0f ms in all (first build %.0f ms); same counts %s"
            % (sd["host_tree_ms"], st0["host_tree_ms"], sd["host_lists_ms"], st0["host_lists_ms"],
               sd["plan_groups_ms"], sd["host_upload_ms"], dplan_ms["total_ms"], first_ms, dplan_ms["same_counts"]))
    r = solve_and_time(F, w, plan, x, q, w["precision"], args.steps, args.warmup, log, args.m2l_mode)
    st = r["st"]
    kernels = r["kernels"]
    dom = "m2l" if kernels["m2l"]["avg_ms"] >= kernels["p2p"]["avg_ms"] else "p2p"
    kd = kernels[dom]
    # M2L runs on MFMA (FP64 MFMA, or FP32 MFMA when mixed); achieved/frac use
    # the algorithmic count, *_executed the dense real form the MFMAs execute
    roofline = {"bound": "mfma", "kernel": dom, "achieved": kd["tflops"], "peak": kd["peak_tflops"],
                "unit": "TFLOP/s", "frac": kd["frac"],
                "traffic": traffic_from_profile(args.workload, dom)}
    if dom == "m2l":
        roofline.update(achieved_executed=kd["exec_tflops"], frac_executed=kd["frac_executed"],
                        count="achieved: 8*NTERM*P^2 flop per M2L pair (the reference's complex "
                              "NTERM x P^2 product); achieved_executed: 2*P^4 per pair (the dense "
                              "real operator applied on MFMA, padding not counted)")
    host_ms = st["host_bounds_ms"] + st["host_tree_ms"] + st["host_lists_ms"] + st["host_upload_ms"]
    value = n * args.steps / r["wall_s"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if w["precision"] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "P": P, "theta": w["theta"],
                   "ncrit": w["ncrit"], "nspawn": w["nspawn"], "inputs": w["desc"],
                   "precision": w["precision"], "m2l_mode": ("auto", "target", "grouped", "tiles")[args.m2l_mode]},
        "latency_ms": {"p50": 1e3 * float(np.percentile(r["lat"], 50)),
                       "p99": 1e3 * float(np.percentile(r["lat"], 99))},
        "device_ms": {"total": float(np.mean(r["dev"])), "upward": st["dev_upward_ms"],
                      "p2p": st["dev_p2p_ms"], "m2l": st["dev_m2l_ms"],
                      "downward": st["dev_downward_ms"]},
        "host_ms": {"bounds": st["host_bounds_ms"], "tree": st["host_tree_ms"],
                    "lists": st["host_lists_ms"], "upload": st["host_upload_ms"]},
        "end_to_end_particles_per_s": n / (1e-3 * (host_ms + float(np.mean(r["dev"])))),
        "device_plan": None if dplan_ms is None else dict(
            dplan_ms, end_to_end_particles_per_s=n / (1e-3 * (dplan_ms["total_ms"] + float(np.mean(r["dev"]))))),
        "accuracy": r["accuracy"],
        "counts": {"cells": st["ncell"], "leaves": st["nleaf"], "depth": st["depth"],
                   "m2l_pairs": st["m2l_pairs"], "p2p_pairs": st["p2p_pairs"],
                   "p2p_body_pairs": st["p2p_body_pairs"], "m2l_groups": st["m2l_groups"]},
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if w.get("compare_mixed"):
        # the mixed-precision solve of the same system (same tree and lists),
        # for "mixed vs FP64 at equal L2 error"
        rm = solve_and_time(F, w, plan, x, q, "mixed", args.steps, args.warmup, log, args.m2l_mode)
        vm = n * args.steps / rm["wall_s"]
        out["mixed_same_system"] = {
            "value": vm, "ms_per_step": 1e3 * rm["wall_s"] / args.steps,
            "latency_ms": {"p50": 1e3 * float(np.percentile(rm["lat"], 50)),
                           "p99": 1e3 * float(np.percentile(rm["lat"], 99))},
            "device_ms": float(np.mean(rm["dev"])), "accuracy": rm["accuracy"], "kernels": rm["kernels"],
            "roofline": {"bound": "mfma", "kernel": "m2l", "achieved": rm["kernels"]["m2l"]["tflops"],
                         "peak": rm["kernels"]["m2l"]["peak_tflops"], "unit": "TFLOP/s",
                         "frac": rm["kernels"]["m2l"]["frac"],
                         "achieved_executed": rm["kernels"]["m2l"]["exec_tflops"],
                         "frac_executed": rm["kernels"]["m2l"]["frac_executed"]}}
        out["mixed_over_fp64"] = vm / value
        out["mixed_vs_fp64_rel_l2"] = dict(zip(("pot", "acc"), rel_l2(r["trg"], rm["trg"])))
        # the same with the M2L on split-FP16 operands (FMM_MIXED16)
        rh = solve_and_time(F, w, plan, x, q, "mixed16", args.steps, args.warmup, log, args.m2l_mode)
        vh = n * args.steps / rh["wall_s"]
        out["mixed16_same_system"] = {
            "value": vh, "ms_per_step": 1e3 * rh["wall_s"] / args.steps,
            "latency_ms": {"p50": 1e3 * float(np.percentile(rh["lat"], 50)),
                           "p99": 1e3 * float(np.percentile(rh["lat"], 99))},
            "device_ms": float(np.mean(rh["dev"])), "accuracy": rh["accuracy"], "kernels": rh["kernels"],
            "roofline": {"bound": "mfma", "kernel": "m2l", "achieved": rh["kernels"]["m2l"]["tflops"],
                         "peak": rh["kernels"]["m2l"]["peak_tflops"], "unit": "TFLOP/s",
                         "frac": rh["kernels"]["m2l"]["frac"],
                         "achieved_executed": rh["kernels"]["m2l"]["exec_tflops"],
                         "frac_executed": rh["kernels"]["m2l"]["frac_executed"]},
            "vs_fp64_rel_l2": dict(zip(("pot", "acc"), rel_l2(r["trg"], rh["trg"])))}
        out["mixed16_over_fp64"] = vh / value
    at = w.get("accurate_theta")
    if at and not args.no_accuracy_point:
        # the same system at a smaller theta, where both precisions reach
        # rel-L2 <= 1e-6 against the direct sum
        plan2 = F.Plan(x, ncrit=w["ncrit"], theta=at, nspawn=w["nspawn"])
        k = max(2, args.steps // 3)
        pts = {}
        for prec in ("fp64", "mixed", "mixed16"):
            ra = solve_and_time(F, w, plan2, x, q, prec, k, 1, log, args.m2l_mode, theta=at)
            pts[prec] = {"value": n * k / ra["wall_s"], "ms_per_step": 1e3 * ra["wall_s"] / k,
                         "accuracy": ra["accuracy"], "m2l_pairs": ra["st"]["m2l_pairs"],
                         "p2p_body_pairs": ra["st"]["p2p_body_pairs"]}
        pts["theta"] = at
        pts["mixed_over_fp64"] = pts["mixed"]["value"] / pts["fp64"]["value"]
        pts["mixed16_over_fp64"] = pts["mixed16"]["value"] / pts["fp64"]["value"]
        out["accuracy_point"] = pts
        del plan2
    if not args.no_graph_small:
        out["graph_small"] = graph_small(F, log)
    if not args.no_serving:
        # the serving stream at its stated size: 2048 requests, N = 2^U[12,18]
        # cube or Plummer, P = 8, mixed, Poisson arrivals at 200 and 1000 req/s
        out["serving"] = {name: serve_bench(F, name, 0, 64, log) for name in ("serve-200", "serve-1000")}
        # the same 200 req/s stream with every request planned on the device
        # (tree, lists, M2L groups and tiles) and solved alone: the plan's
        # arrays go device to device, no host merge or grouping
        out["serving"]["serve-200-device-plan"] = dict(
            serve_bench(F, "serve-200", 0, 64, log, plan_on_device=1, m2l_mode=F.M2L_TILES, batch_requests=1),
            loop={"plan_on_device": 1, "m2l_mode": "tiles", "batch_requests": 1})
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(x, q, w, plan, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

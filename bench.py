"""Sweep throughput on one MI355X: parameter points per second.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads (every input generated on the device from the grid rule and the
seeded initial-state rule, so nothing crosses PCIe in the timed region):

  rossler-2048-surrogate (default)  Rossler (b=0.2) over a 2048x2048 (a, c) plane,
      adaptive DVERK 6(5) rtol=atol=1e-10, transient 200, averaging 2000, MGS every
      0.5.  A step is the chosen form end to end: a 256x256 training map in the
      mixed form, surrogate training (BF16 MFMA, Adam, 80 epochs of 2048-point
      batches), inference over the plane, mixed-form reruns where the surrogate
      is uncertain or |mu1| < 0.02, and the FP64 judge on a seeded 1 %
      verification subsample.
  rossler-512   the same system on a 512x512 grid; a step is one mixed-form sweep.
  shimizu-200-ref  the reference's shipped configuration (shimizu 200x200,
      alpha in [0.2,2] x lambda in [0,1.2], h=1e-2, skips 1e3 + 1e3, total 1e4,
      renormalisation every step, FD mode, FP64): one sweep of the grid.
  rossler-2048  the 2048x2048 plane in the mixed form alone (no surrogate): one
      mixed-form sweep of every point, the same kernel as the surrogate's rerun.
  lorenz-64     Lorenz (sigma=10) 64x64, Dopri5 fixed h=0.01; one mixed-form sweep.
  lorenz96-16   Lorenz-96 N=16 on a 256x256 (F, d) grid, adaptive Dopri5 rtol=1e-9; the
      tile shape (waves per point or points per wave) picked by the autotuner first.
  batch-mix     2^20 points (--batch-points) drawn as 40% Lorenz, 30% Rossler, 20%
      Lorenz-96 N=8, 8% N=16, 2% N=32, adaptive Dopri5 rtol=1e-9, mixed form; a
      step is one replay of the captured persistent batched launch, tuned first
      by the autotuner on a 1/16 sample.  Parameters are uploaded once, before
      the timed region.
  hbm-copy      a plain 1 GiB device copy: the reference for the HBM byte counters.

The plain FP64 sweep of the same grid is timed once for the speed-up.  Prints
one JSON line.

Roofline: the sweep is bound by the vector ALU; `achieved` is the algorithmic
flop count of the dominant kernel launch (per-point accepted/rejected step
counters x per-step flops, chaos_oneapi_amd/flops.py) over that launch's
duration from HIP events on its stream; `peak` is the published MI355X vector
peak weighted by the launch's binary64/binary32 flop mix (78.6 / 157.3
TFLOP/s); the issue rate the valu_peak probe measures on the box (binary64
and unpacked binary32 FMAs) is reported beside it.  `traffic` comes from the
rocprofv3 byte-counter summary of the same workload in profiles/ when one is
there (FETCH_SIZE and WRITE_SIZE scaled by their readings for a plain copy).

cpu_baseline: the CPU FP64 oracle (oracle/lyap_oracle.c, OpenMP) timed on a
bounded seeded subsample of the same grid; its values double as a parity check.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("param points/sec on 1 GPU at |Δλ1|≤1e-3 vs FP64 oracle; "
          "chosen form vs plain FP64")
CHAOS_THRESHOLD = 0.01
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def valu_probe():
    """The valu_peak probe's measured FMA issue rates (None if not built)."""
    import subprocess

    exe = os.path.join(ROOT, "chaos_oneapi_amd", "valu_peak")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe, "20000", "3"], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise RuntimeError(f"valu_peak failed: {r.stderr}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def measured_traffic(workload):
    """HBM bytes per launch of the workload's dominant kernel from the
    committed byte-counter summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return None if rec is None else rec.get("bytes_per_launch")


def error_map_stats(lam_a, lam_b, ok, n1, n2, blocks=4):
    """Per-exponent |lam_a - lam_b| over the points where ok: max, 99.9th
    percentile, median, points over 1e-3, and a blocks x blocks map of the
    block maxima over the (p1, p2) grid (p1 outer)."""
    import numpy as np

    d = np.where(ok[:, None], np.abs(lam_a - lam_b), np.nan)
    k = d.shape[1]
    out = {"max": [], "p999": [], "median": [], "over_1e-3": [], "block_max": []}
    grid = d.reshape(n1, n2, k)
    e1 = np.linspace(0, n1, blocks + 1).astype(int)
    e2 = np.linspace(0, n2, blocks + 1).astype(int)
    for j in range(k):
        v = d[ok, j]
        out["max"].append(float(v.max()) if v.size else 0.0)
        out["p999"].append(float(np.quantile(v, 0.999)) if v.size else 0.0)
        out["median"].append(float(np.median(v)) if v.size else 0.0)
        out["over_1e-3"].append(int(np.sum(v > 1e-3)))
        bm = [[float(np.nanmax(grid[e1[a]:e1[a + 1], e2[b]:e2[b + 1], j]))
               if np.isfinite(grid[e1[a]:e1[a + 1], e2[b]:e2[b + 1], j]).any() else None
               for b in range(blocks)] for a in range(blocks)]
        out["block_max"].append(bm)
    return out


def cpu_baseline(desc, lam_chosen, lam_fp64, budget_s: float, threads: int, idx_pool=None):
    """Time the CPU oracle on a seeded subsample; returns (record, parity dict)."""
    import numpy as np

    from oracle import oracle as O

    od = O.make_desc(desc.to_dict())
    od.precision = 0
    if od.method == O.METHOD_VAR_TRACE:
        od.method = O.METHOD_VAR
    n = desc.n1 * desc.n2
    rng = np.random.default_rng(12345)
    order = rng.permutation(n) if idx_pool is None else rng.permutation(idx_pool)
    done, t_cpu, chunk = 0, 0.0, max(threads, 16)
    idx_all, lam_all = [], []
    while t_cpu < budget_s and done < len(order):
        idx = order[done:done + chunk]
        t0 = time.perf_counter()
        lam_o, _, _, _ = O.sweep(od, idx_list=idx, n_threads=threads)
        t_cpu += time.perf_counter() - t0
        idx_all.append(idx)
        lam_all.append(lam_o)
        done += len(idx)
        chunk = min(4 * chunk, 4096)
    idx = np.concatenate(idx_all)
    lam_o = np.concatenate(lam_all)
    # one core on a smaller seeded sample (about a quarter of the budget)
    t1, n1c = 0.0, 0
    one = rng.permutation(order)[:4096]
    while t1 < budget_s / 4 and n1c < len(one):
        t0 = time.perf_counter()
        O.sweep(od, idx_list=one[n1c:n1c + 4], n_threads=1)
        t1 += time.perf_counter() - t0
        n1c += 4
    rec = {"value": done / t_cpu, "unit": "points/s", "cores": threads, "kind": "port",
           "sample": f"{done} seeded random points of the same grid, FP64 oracle, "
                     f"OpenMP {threads} threads, {t_cpu:.1f} s",
           "single_core_value": n1c / t1 if t1 > 0 else None,
           "single_core_sample": f"{n1c} seeded random points, 1 thread, {t1:.1f} s",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    parity = {"oracle_sample": int(done)}
    if lam_fp64 is not None:
        ok = np.isfinite(lam_o[:, 0]) & np.isfinite(lam_fp64[idx, 0])
        parity["max_dlambda_fp64_vs_oracle"] = float(np.max(np.abs(lam_fp64[idx][ok] - lam_o[ok]))) if ok.any() else 0.0
    if lam_chosen is not None:
        ok = np.isfinite(lam_o[:, 0]) & np.isfinite(lam_chosen[idx, 0])
        parity["max_dlambda1_chosen_vs_oracle"] = float(np.max(np.abs(lam_chosen[idx][ok, 0] - lam_o[ok, 0]))) if ok.any() else 0.0
    return rec, parity


def chosen_form_name(desc):
    from chaos_oneapi_amd import lyap as L

    name = "mixed" if desc.precision == L.PREC_MIXED else "fp64"
    if desc.precision == L.PREC_MIXED and desc.adaptive and desc.integrator == L.INT_DVERK and desc.dim == 3:
        name += " (binary32 tangent on DVERK's fifth-order weights)"
    if desc.method == L.METHOD_VAR_TRACE:
        name += (f", {desc.dim - 1} deviation vectors + trace closure (lambda_{desc.dim} = "
                 "<tr J> - the others)")
    return name


def plain_desc_of(desc):
    """The plain method of a chosen form: FP64, every deviation vector."""
    from chaos_oneapi_amd import lyap as L

    plain = L.SweepDesc.from_buffer_copy(desc)
    plain.precision = L.PREC_FP64
    plain.method = L.METHOD_VAR if desc.method == L.METHOD_VAR_TRACE else desc.method
    return plain


def set_lane_shape(ctx, shape):
    """(lane_max_blocks, lane_gather, lane_quad) on ctx."""
    for key, v in zip(("lane_max_blocks", "lane_gather", "lane_quad"), shape):
        ctx.set_option(key, int(v))


def plain_fp64(ctx, desc, lane_shape=None, resident=None):
    """One plain FP64 sweep of the grid: (lam, status, ms, flops, order tuning);
    lane_shape: the 3-D flows' launch shape to run it with (its own tuning);
    resident: the launch's resident points, to tune its own heavy-first order
    (None: the order set on ctx)."""
    import torch

    from chaos_oneapi_amd import autotune as A
    from chaos_oneapi_amd import flops as F

    plain = plain_desc_of(desc)
    if lane_shape is not None:
        set_lane_shape(ctx, lane_shape)
    order = None
    if resident is not None:
        order = A.tune_order(ctx, plain, resident, cache=A.Cache(None))
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    lam, st, na, nr = ctx.run(plain)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    return (lam.cpu().numpy(), st.cpu().numpy(), ms,
            F.sweep_flops(plain, na.cpu().numpy(), nr.cpu().numpy()), order)


def bench_sweep(args, ctx, spec, wl_name):
    import numpy as np
    import torch

    from chaos_oneapi_amd import flops as F
    from chaos_oneapi_amd import lyap as L

    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    from chaos_oneapi_amd import autotune as A

    tuned = tuned_lanes = None
    if args.no_tune:
        # fixed shapes (profiling runs): the surrogate rerun's gather for the
        # 3-D flows, the default Lorenz-96 kernel, heavy-first order by
        # --order-probe for adaptive grids
        if desc.dim == 3:
            ctx.set_option("lane_gather", 4)
        if desc.adaptive:
            ctx.set_option("order_probe", args.order_probe)
    elif desc.model == L.MODEL_L96:
        tuned = A.tune_l96(ctx, desc, cache=A.Cache(None), time_scale=0.1)  # sets the tile shape on ctx
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    # an adaptive 3-D grid of at most a few points per lane: the heavy-first
    # order and the launch shape set its length together, so the order is
    # tuned first and the shape on full runs in that order
    few_rounds = (not args.no_tune and desc.dim == 3 and desc.n1 * desc.n2 <= 4 * A.resident_points(desc, cus)
                  and len(A.order_candidates(desc, n, A.resident_points(desc, cus))) > 1)
    tuned_order = None
    if few_rounds:
        tuned_order = A.tune_order(ctx, desc, A.resident_points(desc, cus), cache=A.Cache(None))
        tuned_lanes = A.tune_lanes(ctx, desc, cache=A.Cache(None), time_scale=1.0)
    elif desc.dim == 3 and not args.no_tune:
        # persistent grid cap and renormalisation gather, set on ctx for the
        # chosen form and the plain FP64 sweep alike
        tuned_lanes = A.tune_lanes(ctx, desc, cache=A.Cache(None), time_scale=0.1)
    # heavy-first order of adaptive grids that start their points in a few
    # rounds of the launch, timed on full runs (probe, sort and scatter included)
    resident = A.resident_points(desc, cus, int(tuned.best[1]) if tuned is not None and tuned.best[0] == "ppw"
                                 else 0)
    if tuned_order is None and not args.no_tune:
        tuned_order = A.tune_order(ctx, desc, resident, cache=A.Cache(None))
    stream = torch.cuda.current_stream()
    dev = torch.device("cuda", 0)
    bufs = (torch.empty((n, desc.n_exps), dtype=torch.float64, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.int32, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev))
    for _ in range(args.warmup):
        ctx.run(desc, out=bufs)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        ctx.run(desc, out=bufs)
        ev[i][1].record(stream)
        torch.cuda.synchronize()
        kms.append(ctx.kernel_ms())  # the sweep's own launch (an ordered run's probe excluded)
    wall = time.perf_counter() - t0
    kern_avg = sum(kms) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    lam_c, st_c = bufs[0].cpu().numpy().copy(), bufs[1].cpu().numpy().copy()
    na_c, nr_c = bufs[2].cpu().numpy().copy(), bufs[3].cpu().numpy().copy()
    f64_c, f32_c = F.sweep_flops_split(desc, na_c, nr_c)
    kname = "k_var_lane" if desc.model != L.MODEL_L96 else (
        "k_l96_wave" if tuned is not None and tuned.best[0] == "ppw" else "k_l96")
    if args.no_plain:
        # profiling runs: the chosen form's launches only
        rec = base_record(args, n / (ms_per_step * 1e-3), ms_per_step, desc)
        rec["config"] = {"workload": wl_name, "points": n, "profiling_run": True,
                         "order_probe": args.order_probe if args.no_tune and desc.adaptive else None,
                         "mean_steps_per_point": float(na_c.mean())}
        rec["roofline"] = roofline(f64_c, f32_c, kern_avg, kname + " (mixed) over the grid", args.probe,
                                   measured_traffic(wl_name))
        return rec
    # the plain FP64 sweep with a launch shape tuned for it (its own method and
    # precision), not the chosen form's
    tuned_plain = None
    if tuned_lanes is not None and few_rounds:
        A.tune_order(ctx, plain_desc_of(desc), resident, cache=A.Cache(None))
        tuned_plain = A.tune_lanes(ctx, plain_desc_of(desc), cache=A.Cache(None), time_scale=1.0, apply=False)
    elif tuned_lanes is not None:
        tuned_plain = A.tune_lanes(ctx, plain_desc_of(desc), cache=A.Cache(None), time_scale=0.1,
                                   apply=False)
    lam_p, st_p, plain_ms, flops_p, order_plain = plain_fp64(
        ctx, desc, None if tuned_plain is None else tuned_plain.best, None if args.no_tune else resident)
    ok = (st_c == 0) & (st_p == 0)
    d1 = float(np.max(np.abs(lam_c[ok, 0] - lam_p[ok, 0]))) if ok.any() else 0.0
    cls = float(np.mean((lam_c[ok, 0] > CHAOS_THRESHOLD) == (lam_p[ok, 0] > CHAOS_THRESHOLD)))
    rec = base_record(args, n / (ms_per_step * 1e-3), ms_per_step, desc)
    rec["config"] = {
        "workload": wl_name, "points": n,
        "chosen_form": chosen_form_name(desc),
        "plain_fp64_points_per_s": n / (plain_ms * 1e-3),
        "speedup_vs_plain_fp64": plain_ms / ms_per_step,
        "max_dlambda1_chosen_vs_plain_fp64": d1,
        "points_over_1e-3": int(np.sum(np.abs(lam_c[ok, 0] - lam_p[ok, 0]) > 1e-3)),
        "dlambda1_chosen_vs_plain_fp64_p999": float(np.quantile(np.abs(lam_c[ok, 0] - lam_p[ok, 0]), 0.999)) if ok.any() else 0.0,
        "dlambda1_chosen_vs_plain_fp64_median": float(np.median(np.abs(lam_c[ok, 0] - lam_p[ok, 0]))) if ok.any() else 0.0,
        "classification_agreement": cls, "points_status_ok": int(ok.sum()),
        "mean_steps_per_point": float(na_c.mean()), "mean_rejects_per_point": float(nr_c.mean()),
        "plain_fp64_achieved_tflops": flops_p / (plain_ms * 1e-3) / 1e12,
        # per-exponent error maps of the chosen form against the plain FP64 sweep
        "dlambda_chosen_vs_plain_fp64_per_exponent": error_map_stats(lam_c, lam_p, ok, desc.n1, desc.n2),
    }
    if args.error_maps:
        np.savez_compressed(args.error_maps, lam_chosen=lam_c, lam_fp64=lam_p, status_chosen=st_c,
                            status_fp64=st_p, n1=desc.n1, n2=desc.n2)
    if tuned_lanes is not None:
        rec["config"]["lane_shape"] = {"lane_max_blocks": int(tuned_lanes.best[0]),
                                       "lane_gather": int(tuned_lanes.best[1]),
                                       "lane_quad": int(tuned_lanes.best[2])}
        rec["config"]["lane_tuning_ms_scaled"] = {str(k): round(v, 1) for k, v in tuned_lanes.timings.items()}
        rec["config"]["lane_tuning_time_scale"] = 1.0 if few_rounds else 0.1
        rec["config"]["plain_fp64_lane_shape"] = {"lane_max_blocks": int(tuned_plain.best[0]),
                                                  "lane_gather": int(tuned_plain.best[1]),
                                                  "lane_quad": int(tuned_plain.best[2])}
    if tuned_order is not None:
        rec["config"]["order_probe"] = int(tuned_order.best[0])
        rec["config"]["order_tuning_ms"] = {str(k[0]): round(v, 1) for k, v in tuned_order.timings.items()}
        rec["config"]["plain_fp64_order_probe"] = int(order_plain.best[0])
        rec["config"]["plain_fp64_order_tuning_ms"] = {str(k[0]): round(v, 1)
                                                       for k, v in order_plain.timings.items()}
    if tuned is not None:
        rec["config"]["l96_tile"] = list(tuned.best)
        rec["config"]["l96_tuning_ms_scaled"] = {str(k): round(v, 1) for k, v in tuned.timings.items()}
    rec["roofline"] = roofline(f64_c, f32_c, kern_avg, kname + " (mixed) over the grid", args.probe,
                               measured_traffic(wl_name))
    if not args.no_cpu:
        cpu, parity = cpu_baseline(desc, lam_c, lam_p, args.cpu_budget, args.cpu_threads)
        rec["cpu_baseline"] = cpu
        rec["config"].update(parity)
    return rec


def bench_reference_mode(args, ctx, spec, wl_name):
    """The reference's own configuration through its own mode (FP64, fixed-step
    DVERK, the FD default): one sweep of the grid a step; the CPU oracle of the
    same mode beside it.  There is no other precision to compare with."""
    import numpy as np
    import torch

    from chaos_oneapi_amd import flops as F
    from chaos_oneapi_amd import lyap as L

    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    bufs = (torch.empty((n, desc.n_exps), dtype=torch.float64, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.int32, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev))
    for _ in range(args.warmup):
        ctx.run(desc, out=bufs)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        ctx.run(desc, out=bufs)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    kern_avg = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    lam = bufs[0].cpu().numpy()
    rec = base_record(args, n / (ms_per_step * 1e-3), ms_per_step, desc)
    rec["config"] = {"workload": wl_name, "points": n,
                     "chosen_form": "the reference's FD mode, FP64 (its shipped default)",
                     "points_finite": int(np.isfinite(lam).all(axis=1).sum())}
    rec["roofline"] = roofline(F.ref_fd_point_flops(desc) * n, 0.0, kern_avg,
                               "k_fdref_lane (FP64) over the grid", args.probe, measured_traffic(wl_name))
    if not args.no_cpu:
        cpu, parity = cpu_baseline(desc, None, lam, args.cpu_budget, args.cpu_threads)
        rec["cpu_baseline"] = cpu
        rec["config"].update(parity)
    return rec


def bench_surrogate(args, ctx, spec, wl_name):
    import numpy as np
    import torch

    from chaos_oneapi_amd import flops as F
    from chaos_oneapi_amd import lyap as L
    from chaos_oneapi_amd import refine as R

    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    kw = dict(train_n=args.train_n, epochs=args.epochs, fourier_sigma=args.fourier_sigma,
              one_launch=not args.split_schedule, order_probe=args.order_probe)
    for _ in range(args.warmup):
        R.refine(ctx, spec, **kw)
    torch.cuda.synchronize()
    times, res = [], None
    for k in range(args.steps):
        t0 = time.perf_counter()
        res = R.refine(ctx, spec, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        print(f"step {k}: {times[-1]:.3f} s, phases {res.timings}", file=sys.stderr, flush=True)
    ms_per_step = 1e3 * sum(times) / len(times)
    st = res.stats
    # the comparison sweeps take their points in the same heavy-first order as
    # the refinement's launch (the probe's time included in theirs)
    ctx.set_option("order_probe", args.order_probe)
    try:
        lam_p, st_p, plain_ms, flops_p, _ = plain_fp64(ctx, desc)
    finally:
        ctx.set_option("order_probe", 0)
    print(f"plain FP64 sweep {plain_ms:.0f} ms", file=sys.stderr, flush=True)
    # the mixed form alone over the whole plane (no surrogate), with the
    # rerun's renormalisation gather: what the surrogate adds to the mixed form
    stream = torch.cuda.current_stream()
    ctx.set_option("lane_gather", 4)
    ctx.set_option("order_probe", args.order_probe)
    try:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        lam_m, st_m, _, _ = ctx.run(desc)
        e1.record(stream)
        torch.cuda.synchronize()
        mixed_ms = e0.elapsed_time(e1)
    finally:
        ctx.set_option("lane_gather", 0)
        ctx.set_option("order_probe", 0)
    lam_m, st_m = lam_m.cpu().numpy(), st_m.cpu().numpy()
    print(f"mixed-only sweep {mixed_ms:.0f} ms", file=sys.stderr, flush=True)
    # the final map against the plain FP64 map over the whole plane
    okp = st_p == 0
    cls_final = R.classify(np.nan_to_num(res.lam[:, 0], nan=0.0), res.status)
    cls_plain = R.classify(lam_p[:, 0], st_p)
    computed = res.rerun & okp & (res.status == 0)
    answered = ~res.rerun
    ans_val = answered & okp & (res.status == 0)
    dl_s = np.abs(res.lam[ans_val, 0] - lam_p[ans_val, 0])
    # the step's mixed-form launches of k_var_lane (one watched launch over the
    # plane, or the split schedule's three): total algorithmic flops over total
    # duration = the average flops per launch over the average launch duration
    acc_fl, rej_fl, f32_acc, n_vec = F.desc_step(desc)
    per = (acc_fl, rej_fl)
    renorms = F.mgs(desc.dim, n_vec) * (desc.n_renorm_transient + desc.n_renorm_average)
    acc = sum(x["acc_steps"] for x in st["launches"])
    rej = sum(x["rej_steps"] for x in st["launches"])
    fl = per[0] * acc + per[1] * rej + renorms * sum(x["points"] for x in st["launches"])
    f32 = f32_acc * acc
    launch_ms = [x["kernel_ms"] for x in st["launches"]]
    rec = base_record(args, n / (ms_per_step * 1e-3), ms_per_step, desc)
    dl_c = np.abs(res.lam[computed, 0] - lam_p[computed, 0])
    rec["config"] = {
        "workload": wl_name, "points": n, "chosen_form": chosen_form_name(desc) + " + surrogate refinement",
        "plain_fp64_points_per_s": n / (plain_ms * 1e-3),
        "speedup_vs_plain_fp64": plain_ms / ms_per_step,
        "mixed_only_points_per_s": n / (mixed_ms * 1e-3),
        "speedup_mixed_only_vs_plain_fp64": plain_ms / mixed_ms,
        "surrogate_gain_vs_mixed_only": mixed_ms / ms_per_step,
        "order_probe": args.order_probe,
        "rerun_fraction": st["rerun_fraction"], "surrogate_answered_fraction": 1.0 - st["rerun_fraction"],
        "train_points": st["train_points"],
        "verify_points": st["verify"], "calibration_points": st["calib"],
        "std_multiplier_kappa": st["kappa"] if math.isfinite(st["kappa"]) else None,
        "escape_answers": st["escape_answers"], "calibration_agreement": st["calib_agreement"],
        "calibration_table": [[k if math.isfinite(k) else None, e, a, w, d]
                              for k, e, a, w, d in st["calib_table"]],
        # the gates on the held-out verification subsample: the chosen form's
        # answer at each point (the surrogate's where the rule answers it, the
        # method's elsewhere) against the FP64 judge
        "gate_met_verify": st.get("gate_met"),
        "max_dlambda1_chosen_vs_fp64_verify": st.get("max_dlambda1_chosen_vs_fp64"),
        "dlambda1_chosen_vs_fp64_verify_p999": st.get("dlambda1_chosen_vs_fp64_p999"),
        "chosen_points_over_1e-3_verify": st.get("chosen_points_over_gate"),
        "class_agreement_verify": st.get("class_agreement"),
        "surrogate_answered_verify": st.get("surrogate_answered_verify"),
        "surrogate_valued_answers_verify": st.get("surrogate_valued_answers_verify"),
        "max_dlambda1_surrogate_answers_verify": st.get("max_dlambda1_surrogate_answers"),
        "surrogate_class_agreement_verify": st.get("surrogate_class_agreement"),
        "max_dlambda1_computed_vs_fp64_verify": st.get("max_dlambda1_computed_vs_fp64"),
        "max_dlambda_method_vs_fp64_verify_per_exponent": st.get("max_dlambda_method_vs_fp64_per_exponent"),
        # over the whole plane, against the plain FP64 sweep
        "surrogate_valued_answers_full_plane": int(ans_val.sum()),
        "max_dlambda1_surrogate_answers_full_plane": float(dl_s.max()) if dl_s.size else 0.0,
        "surrogate_answers_over_1e-3_full_plane": int(np.sum(dl_s > 1e-3)),
        "max_dlambda1_computed_vs_plain_fp64": float(np.max(dl_c)) if dl_c.size else 0.0,
        "computed_points_over_1e-3_vs_plain_fp64": int(np.sum(dl_c > 1e-3)),
        "dlambda1_computed_vs_plain_fp64_p999": float(np.quantile(dl_c, 0.999)) if dl_c.size else 0.0,
        "class_agreement_full_plane": float(np.mean(cls_final == cls_plain)),
        "class_agreement_answered_full_plane": float(np.mean(cls_final[answered] == cls_plain[answered])) if answered.any() else 1.0,
        # per-exponent error maps of the mixed form against FP64 over the plane
        "dlambda_mixed_vs_plain_fp64_per_exponent": error_map_stats(lam_m, lam_p, (st_m == 0) & okp, desc.n1, desc.n2),
        "rerun_lane_gather": st.get("rerun_gather"),
        "launch_free_blocks": st.get("train_blocks"),
        "one_launch": st.get("one_launch"),
        "computed_points": st.get("computed_points"),
        "speculative_points": st.get("speculative_points"),
        "timings_s": {k: round(v, 3) for k, v in res.timings.items()},
        "plain_fp64_achieved_tflops": flops_p / (plain_ms * 1e-3) / 1e12,
    }
    if args.error_maps:
        np.savez_compressed(args.error_maps, lam_chosen=res.lam, lam_mixed=lam_m, lam_fp64=lam_p,
                            status_chosen=res.status, status_fp64=st_p, answered=answered,
                            n1=desc.n1, n2=desc.n2)
    parts = [x["name"] for x in st["launches"]]
    # the launches' flops over their summed durations (each charged its own)
    rec["roofline"] = roofline(fl - f32, f32, sum(launch_ms), "k_var_lane (mixed): " + ", ".join(parts),
                               args.probe, measured_traffic(wl_name))
    rec["roofline"]["launch_ms"] = launch_ms
    rec["roofline"]["avg_launch_ms"] = sum(launch_ms) / len(launch_ms)
    if not args.no_cpu:
        cpu, parity = cpu_baseline(desc, None, lam_p, args.cpu_budget, args.cpu_threads)
        rec["cpu_baseline"] = cpu
        rec["config"].update(parity)
    return rec


def bench_batch(args, ctx, n_points, wl_name):
    """The batched uneven sweep: one persistent launch over the five classes,
    replayed from its hipGraph; tuned on a 1/16 sample by the autotuner."""
    import numpy as np
    import torch

    from chaos_oneapi_amd import autotune as A
    from chaos_oneapi_amd import flops as F
    from chaos_oneapi_amd import lyap as L
    from chaos_oneapi_amd import workloads as W

    dev = torch.device("cuda", 0)
    mix = W.batch_mix(n_points)

    def device_classes(precision, trace=False):
        out = []
        for c in mix:
            spec = dict(c["spec"], precision=precision)
            if trace and spec["dim"] == 3:
                spec["method"] = L.METHOD_VAR_TRACE
            out.append(dict(desc=L.finalize(L.make_desc(spec)), count=c["count"],
                            params=torch.from_numpy(c["params"]).to(dev).contiguous()))
        return out

    def progress(msg):
        print(f"[batch-mix] {msg}", file=sys.stderr, flush=True)

    classes = device_classes(L.PREC_MIXED, trace=not args.full_vectors)
    progress("classes " + ", ".join(f"{c['name']} {c['count']}" for c in mix))
    t0 = time.perf_counter()
    tune = A.tune_batch(ctx, classes, cache=A.Cache(None), time_scale=0.1)
    tune_s = time.perf_counter() - t0
    progress(f"tuned in {tune_s:.1f} s: {tune.best} {tune.timings}")
    b = L.Batch(ctx, classes, **A.batch_kwargs(tune.best))
    for _ in range(args.warmup):
        b.launch()
    torch.cuda.synchronize()
    kern, t0 = [], time.perf_counter()
    for _ in range(args.steps):
        b.launch()
        kern.append(b.last_ms())
        progress(f"batched launch {kern[-1]:.0f} ms")
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    outs = [tuple(t.cpu().numpy() for t in o) for o in b.outputs]
    b.close()
    split = [F.sweep_flops_split(c["desc"], o[2], o[3]) for c, o in zip(classes, outs)]
    f64_b, f32_b = sum(x[0] for x in split), sum(x[1] for x in split)
    # the same classes as separate single-class launches of the same kernels
    # (the batch layout's Lorenz-96 points per wave)
    single_ms = 0.0
    layout = A.batch_kwargs(tune.best)["layout"]
    for c, m in zip(classes, mix):
        dim = c["desc"].dim
        if c["desc"].model == L.MODEL_L96:
            ctx.set_option(f"l96_ppw_{dim}", L.batch_layout_ppw(layout, dim))
        try:
            ctx.run(c["desc"], count=c["count"], params=c["params"])
        finally:
            if c["desc"].model == L.MODEL_L96:
                ctx.set_option(f"l96_ppw_{dim}", 0)
        single_ms += ctx.kernel_ms()
        progress(f"single-class launch {m['name']}: {ctx.kernel_ms():.0f} ms")
    # plain FP64: the same batch in FP64 (same tuning)
    c64 = device_classes(L.PREC_FP64)
    try:
        b64 = L.Batch(ctx, c64, **A.batch_kwargs(tune.best))
    except L.LyapError:
        b64 = L.Batch(ctx, c64)
    b64.launch()
    plain_ms = b64.last_ms()
    progress(f"FP64 batched launch {plain_ms:.0f} ms")
    outs64 = [tuple(t.cpu().numpy() for t in o) for o in b64.outputs]
    b64.close()
    per_class, worst = {}, 0.0
    for c, o, o64 in zip(mix, outs, outs64):
        ok = (o[1] == 0) & (o64[1] == 0)
        d1 = float(np.max(np.abs(o[0][ok, 0] - o64[0][ok, 0]))) if ok.any() else 0.0
        dk = np.abs(o[0][ok] - o64[0][ok])
        worst = max(worst, d1)
        per_class[c["name"]] = {"points": c["count"], "status_ok": int(ok.sum()),
                                "mean_steps": float(o[2].mean()), "max_steps": int(o[2].max()),
                                "max_dlambda1_vs_fp64": d1,
                                "points_over_1e-3_lambda1": int(np.sum(dk[:, 0] > 1e-3)) if dk.size else 0,
                                "max_dlambda_vs_fp64_per_exponent": [float(x) for x in dk.max(axis=0)] if dk.size else [],
                                "dlambda_vs_fp64_p999_per_exponent": [float(x) for x in np.quantile(dk, 0.999, axis=0)] if dk.size else []}
    kern_avg = sum(kern) / len(kern)
    rec = base_record(args, n_points / (ms_per_step * 1e-3), ms_per_step, classes[0]["desc"])
    rec["config"] = {
        "workload": wl_name, "points": n_points,
        "chosen_form": "mixed, one persistent batched launch"
                       + ("" if args.full_vectors else ", 3-D classes with the trace closure"),
        "tuning": A.batch_kwargs(tune.best), "tuning_s": round(tune_s, 2),
        "tuning_timings_ms_sample": {str(k): round(v, 2) for k, v in tune.timings.items()},
        "plain_fp64_points_per_s": n_points / (plain_ms * 1e-3),
        "speedup_vs_plain_fp64": plain_ms / ms_per_step,
        "single_class_launches_ms": single_ms, "batched_kernel_ms": kern_avg,
        "max_dlambda1_chosen_vs_plain_fp64": worst, "classes": per_class,
    }
    rec["roofline"] = roofline(f64_b, f32_b, kern_avg, "k_batch (mixed) persistent launch", args.probe,
                               measured_traffic(wl_name))
    if not args.no_cpu:
        cpu, parity = cpu_baseline_batch(mix, outs64, args.cpu_budget, args.cpu_threads)
        rec["cpu_baseline"] = cpu
        rec["config"].update(parity)
    return rec


def cpu_baseline_batch(mix, outs64, budget_s, threads):
    """The CPU FP64 oracle on a seeded sample of every class of the batch:
    each class gets a share of the budget in proportion to its share of the
    points; the mix's rate is 1 / sum_c share_c * (seconds per point of c).
    The sample's values are compared with the FP64 batch's."""
    import numpy as np

    from chaos_oneapi_amd import lyap as L
    from oracle import oracle as O

    rng = np.random.default_rng(12346)
    total = sum(c["count"] for c in mix)
    sec_per_point, sampled, worst = 0.0, {}, 0.0
    sec_per_point_1, sampled_1 = 0.0, {}
    t_all = 0.0
    for c, o64 in zip(mix, outs64):
        share = c["count"] / total
        fd = L.finalize(L.make_desc(dict(c["spec"], precision=L.PREC_FP64)))
        od = O.make_desc(fd.to_dict())
        order = rng.permutation(c["count"])
        done, t_c, chunk = 0, 0.0, max(threads, 16)
        params = np.ascontiguousarray(c["params"])
        while t_c < budget_s * max(share, 0.05) and done < len(order):
            # the points keep their class index: it keys the seeded initial state
            idx = np.sort(order[done:done + chunk])
            t0 = time.perf_counter()
            lam_o, st_o, _, _ = O.sweep(od, idx_list=idx, params_in=params, n_threads=threads)
            t_c += time.perf_counter() - t0
            ok = (st_o == 0) & (o64[1][idx] == 0)
            if ok.any():
                worst = max(worst, float(np.max(np.abs(lam_o[ok] - o64[0][idx][ok]))))
            done += len(idx)
            chunk = min(2 * chunk, 1024)
        sec_per_point += share * t_c / done
        t_all += t_c
        sampled[c["name"]] = done
        # one core: a few points of the class (Lorenz-96 points cost seconds each)
        k1 = 8 if c["spec"]["dim"] == 3 else (2 if c["spec"]["dim"] == 8 else 1)
        idx1 = np.sort(order[:k1])
        t0 = time.perf_counter()
        O.sweep(od, idx_list=idx1, params_in=params, n_threads=1)
        sec_per_point_1 += share * (time.perf_counter() - t0) / len(idx1)
        sampled_1[c["name"]] = len(idx1)
    rec = {"value": 1.0 / sec_per_point, "unit": "points/s", "cores": threads, "kind": "port",
           "sample": "seeded samples per class " + ", ".join(f"{k} {v}" for k, v in sampled.items())
                     + f", FP64 oracle, OpenMP {threads} threads, {t_all:.1f} s; rate of the 40/30/20/8/2 mix",
           "single_core_value": 1.0 / sec_per_point_1,
           "single_core_sample": "1 thread on " + ", ".join(f"{k} {v}" for k, v in sampled_1.items())
                                 + " points; rate of the mix",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    return rec, {"oracle_sample": int(sum(sampled.values())), "max_dlambda_fp64_vs_oracle": worst}


def bench_hbm_copy(args):
    """A plain device copy of known size: the reference reading for the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) and the copy bandwidth."""
    import torch

    n = 1 << 27  # 128 Mi doubles = 1 GiB
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"metric": "device copy bandwidth (read + write)", "value": 2 * n * 8 / dt / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": dt * 1e3,
            "higher_is_better": True, "bytes_read_per_copy": n * 8, "bytes_written_per_copy": n * 8}


def base_record(args, value, ms_per_step, desc):
    from chaos_oneapi_amd import lyap as L

    return {
        "metric": METRIC, "value": value, "unit": "points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 stages + f64 state" if desc.precision == L.PREC_MIXED else "f64",
        "data": "synthetic (seeded grid and initial states generated on the device)",
    }


def roofline(f64, f32, ms, kernel, probe=None, traffic=None):
    """Roofline record of a launch (or launches) doing f64 + f32 algorithmic
    flops in ms milliseconds: against the precision-weighted published peak,
    and against the binary64 / unpacked binary32 FMA issue rate measured by
    the valu_peak probe (the rate this build's instructions can reach)."""
    from chaos_oneapi_amd import flops as F

    achieved = (f64 + f32) / (ms * 1e-3) / 1e12
    peak = F.weighted_peak(f64, f32)
    rec = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
           "frac": achieved / peak, "traffic": traffic, "kernel": kernel,
           "flops_fp64": f64, "flops_fp32": f32, "launch_ms": ms,
           "peak_basis": "published vector peaks 78.6 (fp64) / 157.3 (fp32) TFLOP/s weighted by the flop mix"}
    if probe is not None:
        tot = f64 + f32
        meas = tot / (f64 / probe["fp64_fma_tflops"] + f32 / probe["fp32_unpacked_fma_tflops"]) if tot else 0.0
        rec["measured_issue_peak"] = meas
        rec["frac_of_measured_issue_peak"] = achieved / meas if meas else None
        rec["valu_probe"] = probe
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="rossler-2048-surrogate")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-n", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--fourier-sigma", type=float, default=6.0)
    ap.add_argument("--order-probe", type=int, default=10,
                    help="heavy-first order by a probe over this many thousandths of the times (0: off): "
                         "the surrogate workload's launch and comparison sweeps, and the single-grid "
                         "workloads under --no-tune (the tuner picks it otherwise)")
    ap.add_argument("--split-schedule", action="store_true",
                    help="surrogate workload: the split schedule (separate first launches, then a rerun)")
    ap.add_argument("--batch-points", type=int, default=1 << 20)
    ap.add_argument("--error-maps", default=None,
                    help="write the full per-point lambda maps (.npz) to this path")
    ap.add_argument("--no-probe", action="store_true", help="skip the valu_peak issue-rate probe")
    ap.add_argument("--no-tune", action="store_true",
                    help="single-grid workloads: fixed launch shapes instead of the autotuner (profiling)")
    ap.add_argument("--no-plain", action="store_true",
                    help="single-grid workloads: skip the plain FP64 sweep and comparisons (profiling)")
    ap.add_argument("--full-vectors", action="store_true",
                    help="3-D flows: the chosen form integrates all N deviation vectors instead of "
                         "N - 1 and the trace closure")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every point is independent and this build runs on a "
              "single MI355X; it launches no multi-GPU sharding (the grid would split into one "
              "index range per GPU, each run by this same single-GPU path).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "hbm-copy":
        args.steps = 5 if args.steps is None else args.steps
        print(json.dumps(bench_hbm_copy(args)), flush=True)
        return

    from chaos_oneapi_amd import lyap as L
    from chaos_oneapi_amd import workloads as W

    # the issue-rate probe runs before this process touches the GPU (a child
    # process of its own)
    args.probe = None if args.no_probe else valu_probe()
    ctx = L.Context(0)
    if args.workload == "rossler-2048-surrogate":
        # three steps: single steps vary by up to ~0.6 s (host stalls, clocks)
        args.steps = 3 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        spec = W.rossler_dverk(2048, precision=L.PREC_MIXED)
        if not args.full_vectors:
            spec = dict(spec, method=L.METHOD_VAR_TRACE)
        rec = bench_surrogate(args, ctx, spec, "rossler-2048x2048-dverk65-adaptive-surrogate")
    else:
        if args.workload != "batch-mix":
            args.steps = 3 if args.steps is None else args.steps
            args.warmup = 1 if args.warmup is None else args.warmup
        if args.workload == "rossler-512":
            spec, name = W.rossler_dverk(512, precision=L.PREC_MIXED), "rossler-512x512-dverk65-adaptive"
        elif args.workload == "rossler-2048":
            spec, name = W.rossler_dverk(2048, precision=L.PREC_MIXED), "rossler-2048x2048-dverk65-adaptive"
        elif args.workload == "shimizu-200-ref":
            spec, name = W.shimizu_reference(), "shimizu-200x200-reference-fd"
        elif args.workload == "lorenz-64":
            spec, name = W.lorenz_fixed(64, precision=L.PREC_MIXED), "lorenz-64x64-dopri5-fixed"
        elif args.workload == "lorenz96-16":
            spec, name = W.lorenz96_dopri(256, 16, precision=L.PREC_MIXED), "lorenz96-16-256x256-dopri5-adaptive"
        elif args.workload == "batch-mix":
            spec, name = None, f"batch-mix-{args.batch_points}-dopri5-adaptive-1e-9"
        else:
            raise SystemExit(f"unknown workload {args.workload}")
        if spec is not None and spec["method"] == L.METHOD_VAR and spec["dim"] == 3 \
                and not args.full_vectors:
            spec = dict(spec, method=L.METHOD_VAR_TRACE)
        if spec is None:
            args.steps = 1 if args.steps is None else args.steps
            args.warmup = 0 if args.warmup is None else args.warmup
            rec = bench_batch(args, ctx, args.batch_points, name)
        elif spec["method"] not in (L.METHOD_VAR, L.METHOD_VAR_TRACE):
            rec = bench_reference_mode(args, ctx, spec, name)
        else:
            rec = bench_sweep(args, ctx, spec, name)
    print(json.dumps(rec), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

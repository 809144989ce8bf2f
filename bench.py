"""Sweep throughput on one MI355X: parameter points per second.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full sweep of the chosen form over the workload's grid, with
every input generated on the device (parameters from the grid rule, initial
states from the seeded generator), so nothing crosses PCIe in the timed
region.  The same grid is also swept once in plain FP64 (the form that
matches the CPU FP64 oracle bit for bit) for the speed-up and the |dlambda1|
check.  Prints one JSON line.

Roofline: the sweep is bound by the vector ALU (FP64 for the plain form);
`achieved` is the algorithmic flop count of the kernel launch (from the
per-point accepted/rejected step counters the kernel writes, see
chaos_oneapi_amd/flops.py) divided by the launch's duration measured with HIP
events on the launch stream.

cpu_baseline: the CPU FP64 oracle (oracle/lyap_oracle.c, OpenMP, one point
per iteration) timed on a bounded seeded subsample of the same grid; its
values double as the parity check of that subsample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("param points/sec on 1 GPU at |Δλ1|≤1e-3 vs FP64 oracle; "
          "chosen form vs plain FP64")
PEAK_FP64_VALU_TFLOPS = 78.6  # MI355X vector FP64 (also the FP32 rate without packed ops)
CHAOS_THRESHOLD = 0.01  # lambda1 above it: chaotic


def workload(name: str):
    from chaos_oneapi_amd import lyap as L
    from chaos_oneapi_amd import workloads as W

    if name == "rossler-512":
        return W.rossler_dverk(512, precision=L.PREC_MIXED), "rossler-512x512-dverk65-adaptive"
    if name == "lorenz-64":
        return W.lorenz_fixed(64, precision=L.PREC_MIXED), "lorenz-64x64-dopri5-fixed"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(desc, lam_chosen, lam_fp64, budget_s: float, threads: int):
    """Time the CPU oracle on a seeded subsample; returns (record, parity dict)."""
    import numpy as np

    from oracle import oracle as O

    od = O.make_desc(desc.to_dict())
    od.precision = 0
    n = desc.n1 * desc.n2
    rng = np.random.default_rng(12345)
    order = rng.permutation(n)
    done, t_cpu, chunk = 0, 0.0, max(threads, 16)
    idx_all, lam_all = [], []
    while t_cpu < budget_s and done < n:
        idx = order[done:done + chunk]
        t0 = time.perf_counter()
        lam_o, st_o, _, _ = O.sweep(od, idx_list=idx, n_threads=threads)
        t_cpu += time.perf_counter() - t0
        idx_all.append(idx)
        lam_all.append(lam_o)
        done += len(idx)
        chunk = min(4 * chunk, 4096)
    idx = np.concatenate(idx_all)
    lam_o = np.concatenate(lam_all)
    ok = np.isfinite(lam_o[:, 0]) & np.isfinite(lam_chosen[idx, 0])
    d_chosen = float(np.max(np.abs(lam_chosen[idx][ok, 0] - lam_o[ok, 0]))) if ok.any() else 0.0
    d_fp64 = float(np.max(np.abs(lam_fp64[idx][ok] - lam_o[ok]))) if ok.any() else 0.0
    rec = {"value": done / t_cpu, "unit": "points/s", "cores": threads, "kind": "port",
           "sample": f"{done} seeded random points of the same grid, FP64 oracle, "
                     f"OpenMP {threads} threads, {t_cpu:.1f} s"}
    parity = {"oracle_sample": int(done), "max_dlambda1_chosen_vs_oracle": d_chosen,
              "max_dlambda_fp64_vs_oracle": d_fp64}
    return rec, parity


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rossler-512")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every point is independent and this build runs on a "
              "single MI355X; multi-GPU sharding is described in DESIGN.md, not launched.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from chaos_oneapi_amd import flops as F
    from chaos_oneapi_amd import lyap as L

    spec, wl_name = workload(args.workload)
    desc = L.finalize(L.make_desc(spec))
    plain = L.SweepDesc.from_buffer_copy(desc)
    plain.precision = L.PREC_FP64
    n = desc.n1 * desc.n2
    ctx = L.Context(0)
    stream = torch.cuda.current_stream()
    dev = torch.device("cuda", 0)
    bufs = (torch.empty((n, desc.n_exps), dtype=torch.float64, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev),
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
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    ms_per_step = wall * 1e3 / args.steps
    lam_c = bufs[0].cpu().numpy().copy()
    st_c = bufs[1].cpu().numpy().copy()
    na_c, nr_c = bufs[2].cpu().numpy().copy(), bufs[3].cpu().numpy().copy()
    flops_c = F.sweep_flops(desc, na_c, nr_c)

    # plain FP64 on the same grid
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    ctx.run(plain, out=bufs)
    e1.record(stream)
    torch.cuda.synchronize()
    plain_ms = e0.elapsed_time(e1)
    lam_p = bufs[0].cpu().numpy().copy()
    st_p = bufs[1].cpu().numpy().copy()
    flops_p = F.sweep_flops(plain, bufs[2].cpu().numpy(), bufs[3].cpu().numpy())

    ok = (st_c == 0) & (st_p == 0)
    d1 = float(np.max(np.abs(lam_c[ok, 0] - lam_p[ok, 0]))) if ok.any() else 0.0
    cls = float(np.mean((lam_c[ok, 0] > CHAOS_THRESHOLD) == (lam_p[ok, 0] > CHAOS_THRESHOLD)))
    kern_avg = sum(kern_ms) / len(kern_ms)
    achieved = flops_c / (kern_avg * 1e-3) / 1e12

    rec = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stages + f64 state" if desc.precision == L.PREC_MIXED else "f64",
        "data": "synthetic (seeded grid and initial states generated on the device)",
        "config": {
            "workload": wl_name,
            "points": n,
            "chosen_form": "mixed" if desc.precision == L.PREC_MIXED else "fp64",
            "plain_fp64_points_per_s": n / (plain_ms * 1e-3),
            "speedup_vs_plain_fp64": plain_ms / kern_avg,
            "max_dlambda1_chosen_vs_plain_fp64": d1,
            "classification_agreement": cls,
            "points_status_ok": int(ok.sum()),
            "mean_steps_per_point": float(na_c.mean()),
            "mean_rejects_per_point": float(nr_c.mean()),
            "plain_fp64_achieved_tflops": flops_p / (plain_ms * 1e-3) / 1e12,
        },
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_VALU_TFLOPS,
                     "traffic": None},
    }
    if not args.no_cpu:
        cpu, parity = cpu_baseline(desc, lam_c, lam_p, args.cpu_budget, args.cpu_threads)
        rec["cpu_baseline"] = cpu
        rec["config"].update(parity)
    print(json.dumps(rec), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

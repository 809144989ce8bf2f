"""Launch-shape autotuner for the sweep kernels.

Knobs, all exposed by liblyap.so:
  - batched launch (lyap_batch_tuning): Lorenz-96 points per wavefront
    (layout), the register budget / waves per SIMD (occupancy3), the
    persistent grid size (waves_per_cu), the queue order of the classes and
    how many 3-D lanes gather before a renormalisation (gather);
  - single Lorenz-96 sweeps (lyap_ctx_set_option): the LDS tile shape, i.e.
    one point per block of W waves (l96_waves_N) or P points per wave
    (l96_ppw_N);
  - single 3-D sweeps: the persistent grid size (lane_max_blocks), the
    renormalisation gather (lane_gather) and four lanes per point (lane_quad);
  - adaptive single sweeps: heavy-first order (order_probe), timed on full
    runs of the grid (tune_order).

Each candidate is timed on a sample of the workload — a seeded subset of the
points with the full integration times (per-point costs keep their
distribution), or, for the batched launch, every point at a fraction of the
integration times (the launch fills the GPU as the full one does and is not
timed by a few slow points) — the fastest is kept, and the choice is cached in a JSON file
keyed by the device name and a signature of the workload.  Every candidate
computes the same results bit for bit (the layouts change which lanes do the
work, not the arithmetic), so tuning never changes the output.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

from . import lyap as L

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "chaos_oneapi_amd", "autotune.json")

# batched-launch candidates: (layout, occupancy3, waves_per_cu, order, gather)
BATCH_CANDIDATES = (
    (0, False, 0, 0, 0), (0, True, 0, 0, 0), (1, False, 0, 0, 0), (1, True, 0, 0, 0),
    (1, True, 0, 0, 8), (1, True, 0, 0, 32), (1, True, 0, 2, 0),
)

# single Lorenz-96 sweep candidates per N: ("waves", W) or ("ppw", P)
L96_CANDIDATES = {
    8: (("waves", 1), ("ppw", 8), ("ppw", 4), ("ppw", 2)),
    16: (("waves", 2), ("waves", 4), ("ppw", 4), ("ppw", 2), ("ppw", 1)),
    32: (("waves", 4), ("waves", 2), ("ppw", 1), ("ppw", 2)),
}


@dataclass
class TuneResult:
    best: tuple
    timings: dict = field(default_factory=dict)  # candidate -> ms
    cached: bool = False


class Cache:
    """A small JSON store {key: {"best": [...], "timings": {...}}}."""

    def __init__(self, path: str | None = DEFAULT_CACHE):
        self.path = path
        self.data = {}
        if path and os.path.exists(path):
            try:
                with open(path) as f:
                    self.data = json.load(f)
            except (OSError, ValueError):
                self.data = {}

    def get(self, key: str):
        return self.data.get(key)

    def put(self, key: str, best, timings: dict) -> None:
        self.data[key] = {"best": list(best), "timings": {str(k): v for k, v in timings.items()}}
        if not self.path:
            return
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def device_name(device: int = 0) -> str:
    import torch

    return torch.cuda.get_device_properties(device).name


def _bucket(n: int) -> int:
    """Counts enter the key rounded to a power of two (tuning transfers)."""
    return 1 << max(0, int(n).bit_length() - 1)


def desc_signature(desc) -> str:
    return (f"m{desc.model}d{desc.dim}i{desc.integrator}a{int(desc.adaptive)}p{desc.precision}"
            f"r{desc.rtol:.0e}t{desc.t_transient + desc.t_slave:g}+{desc.t_average:g}/{desc.t_renorm:g}")


def batch_key(device: str, classes) -> str:
    parts = [f"{desc_signature(c['desc'])}x{_bucket(c['count'])}" for c in classes]
    return "batch|" + device + "|" + ";".join(parts)


def pick_best(timings: dict):
    """Fastest candidate; ties go to the earlier one (the default comes first)."""
    best, best_ms = None, float("inf")
    for cand, ms in timings.items():
        if ms < best_ms:
            best, best_ms = cand, ms
    return best


def sample_indices(count: int, frac: float, min_points: int, seed: int):
    import numpy as np

    k = min(count, max(min_points, int(round(count * frac))))
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(count, size=k, replace=False)) if k < count else np.arange(count)


MIN_SAMPLE_TOTAL = 1 << 16


def sample_classes(classes, frac: float = 1 / 16, min_points: int = 256, seed: int = 5,
                   min_total: int = MIN_SAMPLE_TOTAL):
    """A seeded subset of every class's points (explicit parameter rows).

    The sample keeps at least min_total points overall (the whole batch when it
    is smaller): a sample that does not fill the GPU several times over is
    timed by its few slowest points, not by the launch shape."""
    import torch

    total = sum(int(c["count"]) for c in classes)
    if total and frac * total < min_total:
        frac = min(1.0, min_total / total)
    out = []
    for j, c in enumerate(classes):
        if c.get("params") is None:
            raise ValueError("batch tuning samples classes with explicit parameters")
        idx = sample_indices(c["count"], frac, min_points, seed + j)
        t = torch.as_tensor(idx, device=c["params"].device)
        s = dict(c)
        s["params"] = c["params"][t].contiguous()
        s["count"] = len(idx)
        if c.get("inits") is not None:
            s["inits"] = c["inits"][t].contiguous()
        out.append(s)
    return out


def scaled_classes(classes, time_scale: float):
    """The same points with the transient and averaging times scaled by
    time_scale (the descriptors re-finalised: renormalisation counts follow,
    t_renorm and the step control stay).  Per-step work and its spread over the
    points are unchanged, so a launch shape that wins on the scaled batch wins
    on the full one, while the scaled batch still fills the GPU as the full one
    does — a subset of the points would be timed by its few slowest points."""
    out = []
    for c in classes:
        s = dict(c)
        s["desc"] = scaled_desc(c["desc"], time_scale)
        out.append(s)
    return out


def tune_batch(ctx, classes, frac: float = 1 / 16, min_points: int = 256, repeats: int = 1,
               cache: Cache | None = None, candidates=BATCH_CANDIDATES, force: bool = False,
               timer=None, time_scale: float | None = None) -> TuneResult:
    """Pick the batched-launch tuning for `classes` (the Batch class list).

    With time_scale the candidates run on every point at time_scale times the
    integration time (scaled_classes); without it, on a seeded subset of the
    points (sample_classes: frac of each class, at least MIN_SAMPLE_TOTAL).
    timer(cand, sample) -> ms may replace the GPU timing (host-side tests)."""
    cache = cache if cache is not None else Cache()
    key = batch_key(device_name(ctx.device) if timer is None else "test", classes)
    hit = cache.get(key)
    if hit and not force:
        return TuneResult(best=tuple(hit["best"]), timings=hit["timings"], cached=True)
    sample = scaled_classes(classes, time_scale) if time_scale else sample_classes(classes, frac, min_points)
    timings = {}
    for cand in candidates:
        try:
            timings[cand] = timer(cand, sample) if timer else _time_batch(ctx, sample, cand, repeats)
        except L.LyapError:
            continue  # a variant that is not built for this integrator/precision
    if not timings:
        raise RuntimeError("no batch tuning candidate could run")
    best = pick_best(timings)
    cache.put(key, best, timings)
    return TuneResult(best=best, timings=timings)


def _time_batch(ctx, sample, cand, repeats: int) -> float:
    b = L.Batch(ctx, sample, **batch_kwargs(cand))
    try:
        ms = []
        for _ in range(repeats):
            b.launch()
            ms.append(b.last_ms())
        return min(ms)
    finally:
        b.close()


def batch_kwargs(best) -> dict:
    layout, occ3, wpc, order = best[:4]
    gather = best[4] if len(best) > 4 else 0
    return dict(layout=layout, occupancy3=bool(occ3), waves_per_cu=wpc, order=order, gather=gather)


def scaled_desc(desc, time_scale: float):
    """desc with the transient and averaging times scaled (re-finalised)."""
    d = L.SweepDesc.from_buffer_copy(desc)
    d.t_transient *= time_scale
    d.t_slave *= time_scale
    d.t_average *= time_scale
    return L.finalize(d)


def tune_l96(ctx, desc, frac: float = 1 / 16, min_points: int = 16384, cache: Cache | None = None,
             force: bool = False, apply: bool = True, timer=None,
             time_scale: float | None = None) -> TuneResult:
    """Pick the Lorenz-96 tile shape of a single sweep over `desc`'s grid and
    (apply=True) set it on ctx.  With time_scale the candidates run over the
    whole grid at that fraction of the integration times (the launch fills
    the GPU as the full one does); without it, on a sample of at least
    min_points points (the whole grid when smaller): fewer than a few per
    wave slot and the launch is timed by its slowest points, not by the shape."""
    import torch

    cache = cache if cache is not None else Cache()
    n_total = desc.n1 * desc.n2
    key = f"l96|{device_name(ctx.device) if timer is None else 'test'}|{desc_signature(desc)}x{_bucket(n_total)}"
    hit = cache.get(key)
    if hit and not force:
        res = TuneResult(best=tuple(hit["best"]), timings=hit["timings"], cached=True)
    else:
        if time_scale:
            run_desc, idx = scaled_desc(desc, time_scale), None
        else:
            run_desc = desc
            idx = torch.as_tensor(sample_indices(n_total, frac, min_points, 3), dtype=torch.int64)
        timings = {}
        for cand in L96_CANDIDATES[desc.dim]:
            if timer:
                timings[cand] = timer(cand, idx)
                continue
            _set_l96(ctx, desc.dim, cand)
            try:
                ctx.run(run_desc, idx_list=None if idx is None else idx.to(torch.device("cuda", ctx.device)))
                timings[cand] = ctx.kernel_ms()
            finally:
                _set_l96(ctx, desc.dim, None)
        res = TuneResult(best=pick_best(timings), timings=timings)
        cache.put(key, res.best, timings)
    if apply and timer is None:
        _set_l96(ctx, desc.dim, res.best)
    return res


def _set_l96(ctx, dim: int, cand) -> None:
    ctx.set_option(f"l96_ppw_{dim}", 0)
    ctx.set_option(f"l96_waves_{dim}", 0)
    if cand is None:
        return
    kind, v = cand
    ctx.set_option(f"l96_{kind}_{dim}", int(v))


def tune_lanes(ctx, desc, frac: float = 1 / 16, min_points: int = 1 << 16,
               cache: Cache | None = None, force: bool = False, apply: bool = True,
               timer=None, time_scale: float | None = None) -> TuneResult:
    """Pick the launch shape of a single 3-D sweep and (apply=True) set it on
    ctx: (lane_max_blocks, lane_gather, lane_quad) — the persistent grid cap
    (0 = every block that fits, else that many 256-lane blocks), how many
    lanes of a wave gather before renormalising together, and whether a point
    takes four lanes (small grids, whose launches last as long as their slowest
    point).  With time_scale the candidates run over the whole grid at that
    fraction of the integration times (1.0: the full run, in the heavy-first
    order set on ctx, for grids of a few points per lane, whose launch length
    the order and the shape set together); without it, on a sample of the
    points."""
    import torch

    cache = cache if cache is not None else Cache()
    n_total = desc.n1 * desc.n2
    name = device_name(ctx.device) if timer is None else "test"
    cus = torch.cuda.get_device_properties(ctx.device).multi_processor_count if timer is None else 256
    key = f"lanes4|{name}|{desc_signature(desc)}x{_bucket(n_total)}|t{time_scale}"
    hit = cache.get(key)
    if hit and not force:
        res = TuneResult(best=tuple(hit["best"]), timings=hit["timings"], cached=True)
    else:
        if time_scale:
            run_desc, idx = scaled_desc(desc, time_scale), None
        else:
            run_desc = desc
            idx = torch.as_tensor(sample_indices(n_total, frac, min_points, 4), dtype=torch.int64)
        timings = {}
        for cand in ((0, 0, 0), (0, 4, 0), (0, 8, 0), (0, 16, 0), (2 * cus, 0, 0), (2 * cus, 8, 0), (0, 0, 1)):
            if timer:
                timings[cand] = timer(cand, idx)
                continue
            ctx.set_option("lane_max_blocks", cand[0])
            ctx.set_option("lane_gather", cand[1])
            ctx.set_option("lane_quad", cand[2])
            try:
                ctx.run(run_desc, idx_list=None if idx is None else idx.to(torch.device("cuda", ctx.device)))
                timings[cand] = ctx.kernel_ms()
            finally:
                ctx.set_option("lane_max_blocks", 0)
                ctx.set_option("lane_gather", 0)
                ctx.set_option("lane_quad", 0)
        res = TuneResult(best=pick_best(timings), timings=timings)
        cache.put(key, res.best, timings)
    if apply and timer is None:
        ctx.set_option("lane_max_blocks", int(res.best[0]))
        ctx.set_option("lane_gather", int(res.best[1]))
        ctx.set_option("lane_quad", int(res.best[2]))
    return res


ORDER_CANDIDATES = (0, 10, 30)


def order_candidates(desc, n_points: int, resident_points: int, max_rounds: int = 32) -> tuple:
    """order_probe values worth timing for a sweep: heavy-first order pays where
    an adaptive grid's points start in a few rounds of the launch's resident
    points (a heavy point of a later round sets the launch's length); with more
    rounds than max_rounds, or fixed steps (every point alike), it stays off."""
    if not desc.adaptive or n_points > max_rounds * max(1, resident_points):
        return (0,)
    return ORDER_CANDIDATES


def resident_points(desc, cus: int, l96_ppw: int = 0, waves_per_simd: int = 3) -> int:
    """Points a persistent launch of the sweep holds at once (about): one per
    lane for the 3-D flows and maps, l96_ppw per wave for the Lorenz-96 wave
    drain, one per block of waves otherwise."""
    waves = cus * 4 * waves_per_simd
    if desc.model == L.MODEL_L96:
        return waves * l96_ppw if l96_ppw > 0 else cus * waves_per_simd
    return waves * 64


def tune_order(ctx, desc, resident: int, cache: Cache | None = None, force: bool = False,
               apply: bool = True, timer=None) -> TuneResult:
    """Pick order_probe for a sweep over desc's whole grid (and, apply=True, set
    it on ctx): each candidate is one full run timed on the stream, probe, sort
    and scatter included — the gain is the schedule of the whole launch, which
    a shortened or sampled run does not show."""
    import torch

    cache = cache if cache is not None else Cache()
    n_total = desc.n1 * desc.n2
    cands = order_candidates(desc, n_total, resident)
    name = device_name(ctx.device) if timer is None else "test"
    key = f"order|{name}|{desc_signature(desc)}x{_bucket(n_total)}r{_bucket(resident)}"
    hit = cache.get(key)
    if len(cands) == 1:
        res = TuneResult(best=(cands[0],), timings={})
    elif hit and not force:
        res = TuneResult(best=tuple(hit["best"]), timings=hit["timings"], cached=True)
    else:
        timings = {}
        for c in cands:
            if timer:
                timings[(c,)] = timer(c)
                continue
            ctx.set_option("order_probe", c)
            try:
                stream = torch.cuda.current_stream()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                ctx.run(desc)
                e1.record(stream)
                torch.cuda.synchronize()
                timings[(c,)] = e0.elapsed_time(e1)
            finally:
                ctx.set_option("order_probe", 0)
        res = TuneResult(best=pick_best(timings), timings=timings)
        cache.put(key, res.best, timings)
    if apply and timer is None:
        ctx.set_option("order_probe", int(res.best[0]))
    return res

"""Surrogate-guided refinement of a sweep (the learned part of the sweep).

Pipeline, all on the device through liblyap.so:
  1. a coarse training map (train_n x train_n over the same parameter ranges,
     the stride-s lattice of the fine grid when it has one) computed by the
     method in the chosen precision;
  2. the surrogate network (lyap_mlp_*) trained on it with Adam;
  3. inference over the fine grid;
  4. the method is rerun where the prediction is uncertain (kappa * std of
     lambda1 above std_thr), near the chaos boundary (|mu1| < lam_thr), or
     where the escape probability is not decided, plus on seeded random
     calibration and verification subsamples (calib_frac, verify_frac of the
     grid, disjoint); the answer rule (std multiplier kappa >= 1, escape
     answers on or off) is the one that answers the most calibration points
     while the map's class agreement there stays >= calib_target and every
     lambda1 it answers there is within dlam_cal (half the 1e-3 gate) of the
     method's value (the predicted std alone is not calibrated);
  5. the plain FP64 method judges the verification subsample.
With longest_first the rerun takes its points longest first: a point's
expected cost is the step count of its nearest training-lattice point, and
the kernel takes the sorted list in order (context option item_order).  Off
by default: on the 2048^2 Rossler plane (3.6 M rerun points, ~18 per lane)
it measured no gain (10.40 s against 10.36 s spread over the plane).
rerun_gather: the lane_gather option of the rerun (2^20 points of that
plane: 2,537 ms at 4 against 2,564 at 0, 2,579 at 2, 2,561 at 6, 2,591-2,620
at 8 and 2,685 at 12; bit-identical).  With the renormalisation's Newton
inverse square root and 64-renormalisation log windows the block a wave runs
for its gathered lanes is cheap, so little gathering pays.
The training map and the calibration and verification runs do not depend on
the surrogate: they are one launch; the FP64 judge runs on a side stream (with
a context of its own) while the surrogate trains; the rerun of step 4 skips
every point already computed.  Optionally (overlap_points > 0), while steps 2-4 run, a
seeded subset of the other points is computed on another side stream with
its persistent grid capped so the training kernels keep room.  Off by
default: on the 2048^2 Rossler plane it saved at most 0.16 s of 9.3 s and
made training time erratic (0.99 s, once 1.98 s, against 0.75 s alone).
The returned map takes the method's values where it ran and the surrogate's
elsewhere.  Gates reported on the held-out verification subsample
(verify_gates): max |dlambda1| between the chosen form's answer (the
surrogate's where the rule answers the point, the method's elsewhere) and
FP64, and agreement of the chaotic/regular/escaped class between the same
answers and FP64.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import lyap as L
from . import workloads as W

CHAOS_THRESHOLD = 0.01


def grid_params(desc, idx):
    """Device tensor [len(idx), 2] of the two swept parameters (p1 outer)."""
    import torch

    i1 = torch.div(idx, desc.n2, rounding_mode="floor").to(torch.float64)
    i2 = torch.remainder(idx, desc.n2).to(torch.float64)
    return torch.stack([desc.start1 + i1 * desc.delta1, desc.start2 + i2 * desc.delta2], dim=1).contiguous()


def classify(lam1, status):
    """0 regular, 1 chaotic, 2 escaped/failed."""
    out = np.where(lam1 > CHAOS_THRESHOLD, 1, 0)
    return np.where(status != 0, 2, out)


@dataclass
class RefineResult:
    lam: np.ndarray
    status: np.ndarray
    rerun: np.ndarray
    timings: dict = field(default_factory=dict)
    stats: dict = field(default_factory=dict)
    # the held-out verification subsample: grid indices, the chosen precision's
    # values there, and the FP64 judge's (None without the judge)
    verify_idx: np.ndarray | None = None
    verify_lam: np.ndarray | None = None
    verify_fp64: np.ndarray | None = None
    verify_fp64_status: np.ndarray | None = None


# Side streams of refine(), created once per device: a fresh torch stream per
# call comes from torch's pool in rotation, and with four hardware queues per
# process (GPU_MAX_HW_QUEUES) consecutive pool streams land on different
# queues from call to call, sometimes the main stream's, which serialises a
# side launch behind the main one (measured: 0.6-0.8 s stalls in every other
# step).  The same three streams every call keep one placement.
_SIDE_STREAMS = {}


def spread_order(n: int, device):
    """Positions (i * s) mod n for i < n, s ~ n/phi coprime to n: the order in
    which the sweep kernels take a list's items by default (item_order 0).
    Applied to an index list once, on the device, it lets the kernel take the
    list in order (item_order 1) with the same spread over the grid: the list
    is then read sequentially instead of one scattered 8-byte entry a point
    (on the 2048^2 rerun the scattered reads fetched 491 MB for a 32 MB list)."""
    import math

    import torch

    if n <= 2:
        return torch.arange(n, device=device, dtype=torch.int64)
    s = max(1, int(n * 0.6180339887498949))
    while math.gcd(s, n) != 1:
        s += 1
    return (torch.arange(n, device=device, dtype=torch.int64) * s) % n


def side_streams(dev):
    import torch

    key = dev.index if dev.index is not None else torch.cuda.current_device()
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = tuple(torch.cuda.Stream(dev) for _ in range(3))
    return _SIDE_STREAMS[key]


# std multipliers the calibration tries.  The answer rule kappa * std <= std_thr
# (std_thr = 0.01) admits predicted stds up to std_thr / kappa; an answered
# lambda1 must also be within the accuracy gate (1e-3) of the method's value,
# which needs stds of a few 1e-4, hence the multipliers up to 128 (std <= 7.8e-5).
# inf: the surrogate answers no lambda (only confident escapes, when allowed).
KAPPAS = (1.0, 1.05, 1.1, 1.15, 1.2, 1.25, 1.3, 1.35, 1.4, 1.45, 1.5, 1.6, 1.75, 2.0, 2.5, 3.0,
          4.0, 5.0, 6.0, 8.0, 12.0, 16.0, 24.0, 32.0, 48.0, 64.0, 96.0, 128.0, float("inf"))

# accuracy gate of the chosen form: max |dlambda1| against the FP64 method
DLAMBDA1_GATE = 1e-3


def answerable(mu1, std1, pbad, kappa, std_thr, lam_thr, bad_lo, bad_hi, escapes=True):
    """Points the surrogate answers at std multiplier kappa: confidently
    escaping (when escapes is set), or decided (not escaping), away from
    lambda1 = 0 and with kappa * std below the threshold (kappa inf: no
    lambda answers).  Works on numpy arrays and torch tensors."""
    if kappa == float("inf"):
        settled = pbad < -1.0  # no lambda answers (probabilities are >= 0)
    else:
        settled = (kappa * std1 <= std_thr) & (abs(mu1) >= lam_thr) & (pbad <= bad_lo)
    return ((pbad >= bad_hi) | settled) if escapes else settled


def answer_errors(pred, lam, status, answered):
    """|mu1 - lambda1| at the answered points where the surrogate answers a
    value (escape probability < 0.5) and the method's point is OK; the other
    answered points are judged by their class only."""
    import numpy as np

    both = answered & (pred[:, 4] < 0.5) & (status == 0) & np.isfinite(lam[:, 0])
    return np.abs(pred[both, 0].astype(np.float64) - lam[both, 0])


# a rule that answers lambda values needs at least this many valued answers on
# the calibration subsample (all within dlam_cal): a handful passing is no
# evidence that its values hold elsewhere (on the 2048^2 Rossler plane a rule
# with a few valued calibration answers missed the gate at 37 of its 50 valued
# answers over the plane)
MIN_VALUED_CALIBRATION = 1000


def calibrate_kappa(pred, lam, status, std_thr, lam_thr, bad_lo, bad_hi, target,
                    dlam_cal=0.5 * DLAMBDA1_GATE, min_valued=MIN_VALUED_CALIBRATION,
                    plane_valued=None):
    """The answer rule that answers the most calibration points while (1) the
    final map's class agreement on the calibration subsample stays at or above
    `target` (the method's own points agree by construction, so this bounds
    the wrong answers: wrong <= (1 - target) * n), (2) every lambda1 it
    answers there is within dlam_cal of the method's value (half the 1e-3
    accuracy gate, so the held-out verification subsample, which sees other
    points of the same error distribution, keeps a margin) and (3) it answers
    lambda values at no calibration point or at min_valued of them at least
    (the evidence that (2) holds beyond the points seen); plane_valued(kappa,
    escapes), when given, counts the points of the whole plane at which the
    rule would answer a value, and a rule with no valued calibration answer
    must answer no value on the plane either.  Rules: kappa in
    KAPPAS, with or without the escape answers (which kappa does not scale).
    Returns (kappa, escapes, agreement, answered, table); kappa inf and no
    escapes: the surrogate answers nothing.  pred [n,5] numpy, lam/status the
    method's values at the same points; table lists (kappa, escapes, answered,
    wrong, max |dlambda1| of the answered values)."""
    import numpy as np

    n = len(pred)
    computed = classify(lam[:, 0], status)
    sur_status = np.where(pred[:, 4] >= 0.5, 2, 0)
    sur = classify(pred[:, 0], sur_status)
    finite = np.isfinite(pred[:, 0])
    budget = (1.0 - target) * n
    best = (float("inf"), False, 1.0, 0)
    table = []
    for escapes in (True, False):
        for kappa in KAPPAS:
            a = answerable(pred[:, 0], pred[:, 3], pred[:, 4], kappa, std_thr, lam_thr, bad_lo,
                           bad_hi, escapes) & finite
            n_a = int(a.sum())
            wrong = int(np.sum(sur[a] != computed[a]))
            err = answer_errors(pred, lam, status, a)
            dmax = float(err.max()) if err.size else 0.0
            table.append((kappa, escapes, n_a, wrong, dmax))
            if n_a > 0 and wrong <= budget and dmax <= dlam_cal and n_a > best[3]:
                if err.size >= min_valued:
                    evidence = True
                elif err.size == 0:
                    evidence = plane_valued is None or plane_valued(kappa, escapes) == 0
                else:
                    evidence = False
                if not evidence:
                    continue
                best = (kappa, escapes, 1.0 - wrong / n, n_a)
    return best + (table,)


def verify_gates(flagged, pred, lam_m, st_m, lam_j, st_j):
    """The gates of the chosen form on the held-out verification points.

    flagged [n] bool: the answer rule does not answer the point (the method
    computes it); pred [n,5] the surrogate's outputs; lam_m/st_m the method's
    values in the chosen precision (every verification point is computed, so
    these are what the rerun would give); lam_j/st_j the FP64 judge.  The
    chosen form's answer at a point is the surrogate's where the rule answers
    it and the method's elsewhere, exactly as for a point outside the
    subsample.  |dlambda1| is taken where the chosen answer is a value and the
    judge's point is OK; the class (chaotic / regular / escaped) everywhere."""
    import numpy as np

    ans = ~flagged
    sur_status = np.where(pred[:, 4] >= 0.5, 2, 0)
    chosen_l1 = np.where(ans, pred[:, 0].astype(np.float64), lam_m[:, 0])
    chosen_ok = np.where(ans, sur_status == 0, st_m == 0)
    both = chosen_ok & (st_j == 0)
    d1 = np.abs(chosen_l1[both] - lam_j[both, 0])
    d1_sur = np.abs(chosen_l1[both & ans] - lam_j[both & ans, 0])
    comp = flagged & (st_m == 0) & (st_j == 0)
    d1_comp = np.abs(lam_m[comp, 0] - lam_j[comp, 0])
    cm = (st_m == 0) & (st_j == 0)
    dk = np.abs(lam_m[cm] - lam_j[cm])
    cls_final = np.where(ans, classify(pred[:, 0], sur_status), classify(lam_m[:, 0], st_m))
    cls_judge = classify(lam_j[:, 0], st_j)
    agree = cls_final == cls_judge
    out = dict(
        verify_points_valued=int(both.sum()),
        max_dlambda1_chosen_vs_fp64=float(d1.max()) if d1.size else 0.0,
        dlambda1_chosen_vs_fp64_p999=float(np.quantile(d1, 0.999)) if d1.size else 0.0,
        chosen_points_over_gate=int(np.sum(d1 > DLAMBDA1_GATE)),
        max_dlambda1_surrogate_answers=float(d1_sur.max()) if d1_sur.size else 0.0,
        surrogate_valued_answers_verify=int(d1_sur.size),
        max_dlambda1_computed_vs_fp64=float(d1_comp.max()) if d1_comp.size else 0.0,
        class_agreement=float(np.mean(agree)) if agree.size else 1.0,
        surrogate_answered_verify=int(ans.sum()),
        surrogate_class_agreement=float(np.mean(agree[ans])) if ans.any() else 1.0,
        # the method in the chosen precision against FP64, every exponent
        # (all verification points, answered or not)
        max_dlambda_method_vs_fp64_per_exponent=[float(x) for x in dk.max(axis=0)] if dk.size else [],
        dlambda_method_vs_fp64_p999_per_exponent=[float(x) for x in np.quantile(dk, 0.999, axis=0)] if dk.size else [],
    )
    out["gate_met"] = bool(out["max_dlambda1_chosen_vs_fp64"] <= DLAMBDA1_GATE and out["class_agreement"] >= 0.999)
    return out


def refine(ctx, fine: dict, train_n: int = 256, epochs: int = 80, batch: int = 2048,
           lr: float = 1e-3, seed: int = 7, verify_frac: float = 0.01, std_thr: float = 0.01,
           lam_thr: float = 0.02, bad_lo: float = 0.02, bad_hi: float = 0.98,
           infer_batch: int = 65536, judge: bool = True, hidden: int = 256,
           fourier_sigma: float = 6.0, calib_frac: float = 0.01,
           calib_target: float = 0.9995, dlam_cal: float = 0.5 * DLAMBDA1_GATE,
           longest_first: bool = False,
           rerun_gather: int = 4, overlap_points: int = 0, overlap_frac: float = 1 / 16,
           overlap_blocks_per_cu: int = 2, first_blocks_per_cu: int = 2,
           split_first: bool = True, split_blocks_per_cu: int = 1,
           judge_quad: bool = False) -> RefineResult:
    import torch

    dev = torch.device("cuda", ctx.device)
    tm = {}

    def mark_stream(name, t0, stream):
        stream.synchronize()
        tm[name] = time.perf_counter() - t0
        return time.perf_counter()

    fd = L.finalize(L.make_desc(fine))
    n = fd.n1 * fd.n2
    end1 = fd.start1 + fd.delta1 * (fd.n1 - 1)
    end2 = fd.start2 + fd.delta2 * (fd.n2 - 1)
    t_all = t = time.perf_counter()

    # 1. the training map, and the calibration and verification subsamples: none
    #    depends on the surrogate.  When the fine grid holds a train_n x train_n
    #    lattice (stride s = (n-1) // (train_n-1) >= 1 in both directions), the
    #    training map is that lattice, so the three sets are one chosen-form
    #    launch over the fine grid (one tail) and their values enter the final
    #    map
    gv = torch.Generator(device="cpu").manual_seed(seed + 1)
    n_ver = max(1, int(round(verify_frac * n)))
    n_cal = max(1, int(round(calib_frac * n)))
    perm = torch.randperm(n, generator=gv)
    verify = perm[:n_ver].sort().values.to(dev)
    calib = perm[n_ver:n_ver + n_cal].sort().values.to(dev)
    main = torch.cuda.current_stream(dev)
    n_cus = torch.cuda.get_device_properties(dev).multi_processor_count
    s1 = (fd.n1 - 1) // (train_n - 1) if train_n > 1 else 0
    s2 = (fd.n2 - 1) // (train_n - 1) if train_n > 1 else 0
    n_t = train_n * train_n
    lattice = s1 >= 1 and s2 >= 1
    cv_ctx, side_c = None, None
    if lattice:
        ar = torch.arange(train_n, device=dev, dtype=torch.int64)
        train_idx = ((ar * s1)[:, None] * fd.n2 + (ar * s2)[None, :]).reshape(-1)
        first = torch.cat([train_idx, calib, verify]).contiguous()
        if split_first:
            # training needs only the lattice: it runs on the main stream while
            # the calibration and verification points run beside it on a side
            # stream with a context of its own (needed only after training),
            # each persistent grid capped at split_blocks_per_cu 256-lane blocks
            # per CU (on the 2048^2 Rossler plane the lattice alone takes 594 ms
            # where the three sets in one launch take 819 ms)
            cv = first[n_t:]
            side_c = side_streams(dev)[0]
            side_c.wait_stream(main)
            cv_ctx = L.Context(ctx.device)
            cv_ctx.set_option("lane_max_blocks", split_blocks_per_cu * n_cus)
            ctx.set_option("lane_max_blocks", split_blocks_per_cu * n_cus)
            try:
                lam_t, st_t, na_t, nr_t = ctx.run(fd, idx_list=train_idx)
            finally:
                ctx.set_option("lane_max_blocks", 0)
            with torch.cuda.stream(side_c):
                lam_cv, st_cv, na_cv, nr_cv = cv_ctx.run(fd, idx_list=cv, stream=side_c)
            lam_f, st_f, na_f, nr_f = None, None, na_t, nr_t
        else:
            # about one point per lane: the launch lasts as long as its slowest
            # point, and lanes advance faster with fewer waves per SIMD (two
            # 256-lane blocks per CU: 820 ms against 1,083 ms at three for
            # 149,422 points of the 2048^2 Rossler plane)
            ctx.set_option("lane_max_blocks", first_blocks_per_cu * n_cus)
            try:
                lam_f, st_f, na_f, nr_f = ctx.run(fd, idx_list=first)
            finally:
                ctx.set_option("lane_max_blocks", 0)
            lam_t, st_t = lam_f[:n_t], st_f[:n_t]
        p_t = grid_params(fd, train_idx)
    else:
        tspec = dict(fine)
        tspec.update(W.grid_fields(fd.sweep_idx1, fd.start1, end1, train_n, fd.sweep_idx2, fd.start2,
                                   end2, train_n))
        td = L.finalize(L.make_desc(tspec))
        lam_t, st_t, _, _ = ctx.run(td)
        p_t = grid_params(td, torch.arange(n_t, device=dev))
        first = torch.cat([calib, verify]).contiguous()
        lam_f, st_f, na_f, nr_f = ctx.run(fd, idx_list=first)
        lam_f = torch.cat([lam_t, lam_f])
        st_f = torch.cat([st_t, st_f])
    if side_c is None:
        lam_c, st_c = lam_f[n_t:n_t + n_cal], st_f[n_t:n_t + n_cal]
        lam_v, st_v = lam_f[n_t + n_cal:], st_f[n_t + n_cal:]
    else:
        lam_c, st_c = lam_cv[:n_cal], st_cv[:n_cal]
        lam_v, st_v = lam_cv[n_cal:], st_cv[n_cal:]
    t = mark_stream("train_map", t, main)
    first_kernel_ms = ctx.kernel_ms()
    vmask = torch.zeros(n, dtype=torch.bool, device=dev)
    vmask[verify] = True
    cmask = torch.zeros(n, dtype=torch.bool, device=dev)
    cmask[calib] = True
    tmask = torch.zeros(n, dtype=torch.bool, device=dev)
    if lattice:
        tmask[train_idx] = True

    # the FP64 judge of the verification points, on a side stream with a
    # context of its own while the surrogate trains (training's small kernels
    # leave most of the GPU idle; beside the latency-bound first launch the
    # judge slowed it by ~40 %, beside the rerun it cost the rerun ~0.3 s)
    judge_ctx = None
    if judge:
        side_j = side_streams(dev)[1]
        side_j.wait_stream(main)
        judge_ctx = L.Context(ctx.device)
        # judge_quad: four lanes per point shorten the judge alone (2048^2
        # Rossler, 41,943 FP64 points: 713 ms against 836 ms) but its four
        # times as many waves slow the training kernels beside it (training
        # 0.49 -> 0.75 s), so it is off by default
        judge_ctx.set_option("lane_quad", 1 if judge_quad else 0)
        with torch.cuda.stream(side_j):
            pdesc = L.SweepDesc.from_buffer_copy(fd)
            pdesc.precision = L.PREC_FP64
            # the judge is the plain method: FP64, every deviation vector
            if pdesc.method == L.METHOD_VAR_TRACE:
                pdesc.method = L.METHOD_VAR
            lam_j, st_j, _, _ = judge_ctx.run(pdesc, idx_list=verify, stream=side_j)

    # 1b. while the surrogate trains, a seeded subset of the other points
    #     (overlap_frac of them, at most overlap_points) runs on a side stream
    #     with a context of its own, its persistent grid capped at
    #     overlap_blocks_per_cu 256-lane blocks per CU so the training kernels
    #     keep room on every CU.  These points are computed whatever the
    #     surrogate later says, so the final map does not depend on timing.
    omask = torch.zeros(n, dtype=torch.bool, device=dev)
    ov_ctx, ov_idx = None, None
    rest = torch.nonzero(~(vmask | cmask | tmask)).flatten()
    n_ov = min(int(overlap_points), int(overlap_frac * rest.numel()))
    if n_ov > 0 and lattice:
        go = torch.Generator(device="cpu").manual_seed(seed + 2)
        pick = torch.randperm(int(rest.numel()), generator=go)[:n_ov].to(dev)
        ov_idx = rest[pick].sort().values.contiguous()
        omask[ov_idx] = True
        ov_ctx = L.Context(ctx.device)
        ov_ctx.set_option("lane_max_blocks", overlap_blocks_per_cu * n_cus)
        ov_ctx.set_option("lane_gather", rerun_gather)
        side_o = side_streams(dev)[2]
        side_o.wait_stream(main)
        with torch.cuda.stream(side_o):
            lam_o, st_o, na_o, nr_o = ov_ctx.run(fd, idx_list=ov_idx, stream=side_o)

    # 2. training (one shuffle per epoch; batches are contiguous slices)
    ok = st_t == 0
    lam_ok = lam_t[ok]
    ym = lam_ok.mean(dim=0).tolist() if lam_ok.shape[0] else [0.0, 0.0, 0.0]
    ys = lam_ok.std(dim=0).clamp_min(1e-6).tolist() if lam_ok.shape[0] > 1 else [1.0, 1.0, 1.0]
    net = L.Mlp(device=ctx.device, max_batch=max(batch, infer_batch), hidden=hidden,
                fourier_sigma=fourier_sigma, lr=lr, seed=seed,
                ranges=(min(fd.start1, end1), max(fd.start1, end1), min(fd.start2, end2),
                        max(fd.start2, end2)), y_mean=ym, y_std=ys)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(n_t, generator=gen).to(dev)
        pp, lp, sp = p_t[perm], lam_t[perm], st_t[perm]
        for b0 in range(0, n_t - batch + 1, batch):
            # the epoch's loss is its last batch's (no copy on the other steps)
            last = b0 + 2 * batch > n_t
            net.train_step(pp[b0:b0 + batch], lp[b0:b0 + batch], sp[b0:b0 + batch], loss if last else None)
        losses.append(loss)
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
    t = mark_stream("train", t, main)

    # 3. inference
    pred = torch.empty((n, 5), dtype=torch.float32, device=dev)
    for c0 in range(0, n, infer_batch):
        c1 = min(n, c0 + infer_batch)
        net.forward(grid_params(fd, torch.arange(c0, c1, device=dev)), pred[c0:c1])
    t = mark_stream("infer", t, main)

    # 4. calibration: the std multiplier at which the surrogate's answers on the
    #    calibration subsample agree with the method (the verification
    #    subsample stays out of it), then the rerun set: every point the
    #    surrogate does not answer and that step 1 did not compute
    if side_c is not None:
        main.wait_stream(side_c)
    mu1, std1, pbad = pred[:, 0], pred[:, 3], pred[:, 4]

    def plane_valued(k, esc):
        a = answerable(mu1, std1, pbad, k, std_thr, lam_thr, bad_lo, bad_hi, esc)
        return int((a & torch.isfinite(mu1) & (pbad < 0.5)).sum().item())

    kappa, escapes, cal_agree, cal_answered, cal_table = calibrate_kappa(
        pred[calib].cpu().numpy(), lam_c.cpu().numpy(), st_c.cpu().numpy(), std_thr, lam_thr,
        bad_lo, bad_hi, calib_target, dlam_cal, plane_valued=plane_valued)
    tm["calibrate"] = time.perf_counter() - t
    if cal_answered == 0:
        flagged = torch.ones(n, dtype=torch.bool, device=dev)
    else:
        flagged = ~answerable(mu1, std1, pbad, kappa, std_thr, lam_thr, bad_lo, bad_hi, escapes)
        flagged |= ~torch.isfinite(mu1)
    rerun_idx = torch.nonzero(flagged & ~vmask & ~cmask & ~tmask & ~omask).flatten().to(torch.int64).contiguous()
    ordered = longest_first and lattice and rerun_idx.numel() > 0
    if ordered:
        # longest first: each point's expected cost is the step count of its
        # nearest lattice point, and the kernel takes the list in order, so the
        # launch ends on short points instead of waiting for a long one
        cost_t = (na_f[:n_t].to(torch.int64) + nr_f[:n_t].to(torch.int64))
        j1 = torch.clamp(torch.round(torch.div(rerun_idx, fd.n2, rounding_mode="floor") / s1), 0, train_n - 1)
        j2 = torch.clamp(torch.round(torch.remainder(rerun_idx, fd.n2) / s2), 0, train_n - 1)
        cost = cost_t[(j1 * train_n + j2).to(torch.int64)]
        rerun_idx = rerun_idx[torch.sort(cost, descending=True, stable=True).indices].contiguous()
    elif rerun_idx.numel() > 0:
        # the kernel's spread order applied to the list itself, taken in order
        rerun_idx = rerun_idx[spread_order(rerun_idx.numel(), dev)].contiguous()
    ctx.set_option("item_order", 1)
    # the rerun keeps every lane busy for many points: lanes at a
    # renormalisation wait until rerun_gather lanes of the wave do (same bits)
    ctx.set_option("lane_gather", rerun_gather)
    try:
        lam_r, st_r, na_r, nr_r = ctx.run(fd, idx_list=rerun_idx)
    finally:
        ctx.set_option("lane_gather", 0)
        ctx.set_option("item_order", 0)
    t = mark_stream("rerun", t, main)
    rerun_kernel_ms = ctx.kernel_ms() if rerun_idx.numel() else 0.0
    torch.cuda.synchronize(dev)
    if judge_ctx is not None:
        judge_ctx.close()
    split_kernel_ms = 0.0
    if cv_ctx is not None:
        split_kernel_ms = cv_ctx.kernel_ms()
        cv_ctx.close()
        na_f, nr_f = torch.cat([na_t, na_cv]), torch.cat([nr_t, nr_cv])
    overlap_kernel_ms = 0.0
    if ov_ctx is not None:
        overlap_kernel_ms = ov_ctx.kernel_ms()
        ov_ctx.close()

    # 5. assemble
    lam = pred[:, :3].to(torch.float64)
    status = torch.where(pbad >= 0.5, L.PT_ESCAPED, L.PT_OK).to(torch.int32)
    lam[pbad >= 0.5] = float("nan")
    lam[rerun_idx] = lam_r
    status[rerun_idx] = st_r
    lam[verify] = lam_v
    status[verify] = st_v
    lam[calib] = lam_c
    status[calib] = st_c
    if lattice:
        lam[train_idx] = lam_t
        status[train_idx] = st_t
    if ov_idx is not None:
        lam[ov_idx] = lam_o
        status[ov_idx] = st_o
    rerun_mask = flagged | vmask | cmask | tmask | omask
    tm["total_chosen"] = time.perf_counter() - t_all

    n_rerun = int(rerun_mask.sum().item())  # distinct points the method computed
    stats = dict(points=n, train_points=n_t, rerun=n_rerun,
                 flagged=int(flagged.sum().item()), verify=n_ver, calib=n_cal,
                 kappa=kappa, escape_answers=bool(escapes), calib_agreement=cal_agree,
                 calib_answered=cal_answered, calib_table=cal_table,
                 train_on_lattice=bool(lattice),
                 rerun_fraction=float(n_rerun) / n, losses=[float(x.item()) for x in losses],
                 rerun_kernel_ms=rerun_kernel_ms, rerun_points=int(rerun_idx.numel()),
                 rerun_longest_first=bool(ordered), rerun_gather=int(rerun_gather),
                 overlap_points=int(ov_idx.numel()) if ov_idx is not None else 0,
                 overlap_kernel_ms=overlap_kernel_ms,
                 overlap_acc_steps=float(na_o.double().sum().item()) if ov_idx is not None else 0.0,
                 overlap_rej_steps=float(nr_o.double().sum().item()) if ov_idx is not None else 0.0,
                 first_kernel_ms=first_kernel_ms, first_points=int(first.numel()),
                 split_first=cv_ctx is not None, split_kernel_ms=split_kernel_ms,
                 first_acc_steps=float(na_f.double().sum().item()),
                 first_rej_steps=float(nr_f.double().sum().item()),
                 rerun_acc_steps=float(na_r.double().sum().item()),
                 rerun_rej_steps=float(nr_r.double().sum().item()))
    if judge:
        v = verify.cpu().numpy()
        lj, sj = lam_j.cpu().numpy(), st_j.cpu().numpy()
        lc, sc = lam_v.cpu().numpy(), st_v.cpu().numpy()
        stats.update(verify_gates(flagged[verify].cpu().numpy(), pred[verify].cpu().numpy(),
                                  lam_v.cpu().numpy(), st_v.cpu().numpy(), lam_j.cpu().numpy(),
                                  st_j.cpu().numpy()))
    net.close()
    out = RefineResult(lam=lam.cpu().numpy(), status=status.cpu().numpy(),
                       rerun=rerun_mask.cpu().numpy(), timings=tm, stats=stats,
                       verify_idx=verify.cpu().numpy(), verify_lam=lam_v.cpu().numpy())
    if judge:
        out.verify_fp64, out.verify_fp64_status = lam_j.cpu().numpy(), st_j.cpu().numpy()
    return out

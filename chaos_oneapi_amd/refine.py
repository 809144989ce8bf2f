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
rerun_gather: the lane_gather option of the method's launches (2^20 points
of the 2048^2 Rossler plane: 2,537 ms at 4 against 2,564 at 0, 2,579 at 2,
2,561 at 6, 2,591-2,620 at 8 and 2,685 at 12; bit-identical).

Schedules.  With one_launch (the default, for the sweep's method with adaptive
steps on a 3-D flow) the method runs as one persistent launch over the
training lattice, the calibration points, the verification points and then
every other point — heavy-first by a probe run of the plane over order_probe
thousandths of the integration times (heavy_first), so that no heavy point
starts in the launch's last round — on a side stream (lyap_ctx_set_watch): it
signals the host
when the lattice and when the calibration points are done, while it goes on;
training, inference and the choice of the rule follow on the main stream, and
the points the rule answers are skipped from the moment their skip words are
written.  The FP64 judge runs on the main stream first, on a few blocks
(judge_blocks), beside the launch, and the launch's grid is the occupancy-sized
one less train_blocks blocks (lane_max_blocks = -train_blocks; by default
judge_blocks), so that the judge's slots pass to training and inference when
the judge ends rather than to launch blocks that would hold them to the end
(2048^2 Rossler plane: training done ~3.7 s into a step instead of after the
launch; 6.68-6.73 s a step at 32 against 6.74 s at 0, 6.78 s at 16,
6.83-6.84 s at 24, 6.74-6.76 s at 40, 6.79-6.80 s at 48, 6.84 s at 64).
So no wave idles on the slowest point of a short
launch: on the 2048^2 Rossler plane the lattice, calibration and verification
points as launches of their own (one point per lane) kept their waves mostly
masked for ~0.8 s, and a launch beside them could only use the issue slots
they left.  The split schedule (one_launch False, or a fine grid without the
lattice): the training map on the main stream with the calibration and
verification points beside it on a side stream, each capped at one block per
CU, the judge beside training, then a rerun of the flagged points.
The returned map takes the surrogate's answers where the rule answers a point
outside the three first sets (also where the one launch computed it before the
rule was known, so the map does not depend on timing) and the method's values
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
# step).  The same two streams every call keep one placement; a third one
# shared the main stream's queue (a launch on it started only when the main
# stream's launch ended), so refine() uses two.
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


def heavy_first(ctx, fd, idx, permille: int, radius: int = 2, stream=None):
    """Grid point indices idx reordered most expected steps first, by the rule
    of lyap_sweep_run's heavy-first order (options order_probe, order_radius):
    a probe run of the whole grid over permille thousandths of the integration
    times (at least one renormalisation interval of averaging) counts every
    point's steps, and a point's key is the largest count within `radius` rows
    and columns of it.  Ties keep idx's order."""
    import torch
    import torch.nn.functional as Fn

    pd = L.SweepDesc.from_buffer_copy(fd)
    f = permille / 1000.0
    pd.t_transient *= f
    pd.t_slave *= f
    pd.t_average = max(fd.t_average * f, fd.t_renorm)
    pd = L.finalize(pd)
    _, _, na, nr = ctx.run(pd, stream=stream)
    keys = (na.to(torch.float32) + nr.to(torch.float32)).reshape(1, 1, fd.n1, fd.n2)
    if radius > 0:
        keys = Fn.max_pool2d(keys, 2 * radius + 1, stride=1, padding=radius)
    order = torch.argsort(keys.reshape(-1)[idx], descending=True, stable=True)
    return idx[order]


def side_streams(dev):
    import torch

    key = dev.index if dev.index is not None else torch.cuda.current_device()
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = tuple(torch.cuda.Stream(dev) for _ in range(2))
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


def _judge_desc(fd):
    """The plain method judging the chosen form: FP64, every deviation vector."""
    pdesc = L.SweepDesc.from_buffer_copy(fd)
    pdesc.precision = L.PREC_FP64
    if pdesc.method == L.METHOD_VAR_TRACE:
        pdesc.method = L.METHOD_VAR
    return pdesc


def _train(dev, fd, p_t, lam_t, st_t, epochs, batch, lr, seed, hidden, fourier_sigma, max_batch):
    """The surrogate trained on the training map (one shuffle per epoch, batches
    contiguous slices of it) on the current stream; returns (net, losses)."""
    import torch

    n_t = int(p_t.shape[0])
    end1 = fd.start1 + fd.delta1 * (fd.n1 - 1)
    end2 = fd.start2 + fd.delta2 * (fd.n2 - 1)
    ok = st_t == 0
    lam_ok = lam_t[ok]
    ym = lam_ok.mean(dim=0).tolist() if lam_ok.shape[0] else [0.0, 0.0, 0.0]
    ys = lam_ok.std(dim=0).clamp_min(1e-6).tolist() if lam_ok.shape[0] > 1 else [1.0, 1.0, 1.0]
    net = L.Mlp(device=dev.index, max_batch=max_batch, hidden=hidden, fourier_sigma=fourier_sigma, lr=lr,
                seed=seed, ranges=(min(fd.start1, end1), max(fd.start1, end1), min(fd.start2, end2),
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
    return net, losses


def _infer(net, fd, n, infer_batch, dev):
    import torch

    pred = torch.empty((n, 5), dtype=torch.float32, device=dev)
    for c0 in range(0, n, infer_batch):
        c1 = min(n, c0 + infer_batch)
        net.forward(grid_params(fd, torch.arange(c0, c1, device=dev)), pred[c0:c1])
    return pred


def _rule(pred, calib, lam_c, st_c, std_thr, lam_thr, bad_lo, bad_hi, calib_target, dlam_cal):
    """The calibrated answer rule and the points it leaves to the method
    (flagged): (kappa, escapes, agreement, answered, table, flagged)."""
    import torch

    mu1, std1, pbad = pred[:, 0], pred[:, 3], pred[:, 4]

    def plane_valued(k, esc):
        a = answerable(mu1, std1, pbad, k, std_thr, lam_thr, bad_lo, bad_hi, esc)
        return int((a & torch.isfinite(mu1) & (pbad < 0.5)).sum().item())

    kappa, escapes, cal_agree, cal_answered, cal_table = calibrate_kappa(
        pred[calib].cpu().numpy(), lam_c.cpu().numpy(), st_c.cpu().numpy(), std_thr, lam_thr,
        bad_lo, bad_hi, calib_target, dlam_cal, plane_valued=plane_valued)
    if cal_answered == 0:
        flagged = torch.ones(pred.shape[0], dtype=torch.bool, device=pred.device)
    else:
        flagged = ~answerable(mu1, std1, pbad, kappa, std_thr, lam_thr, bad_lo, bad_hi, escapes)
        flagged |= ~torch.isfinite(mu1)
    return kappa, escapes, cal_agree, cal_answered, cal_table, flagged


def one_launch_supported(fd) -> bool:
    """Whether refine() can run the fine grid as one watched launch: the
    sweep's method with adaptive steps on a 3-D flow (lyap_ctx_set_watch)."""
    return (fd.method in (L.METHOD_VAR, L.METHOD_VAR_TRACE) and bool(fd.adaptive) and fd.dim == 3
            and fd.model in (L.MODEL_SHIMIZU, L.MODEL_LORENZ, L.MODEL_ROSSLER))


def refine(ctx, fine: dict, train_n: int = 256, epochs: int = 80, batch: int = 2048,
           lr: float = 1e-3, seed: int = 7, verify_frac: float = 0.01, std_thr: float = 0.01,
           lam_thr: float = 0.02, bad_lo: float = 0.02, bad_hi: float = 0.98,
           infer_batch: int = 65536, judge: bool = True, hidden: int = 256,
           fourier_sigma: float = 6.0, calib_frac: float = 0.01,
           calib_target: float = 0.9995, dlam_cal: float = 0.5 * DLAMBDA1_GATE,
           rerun_gather: int = 4, one_launch: bool = True, judge_blocks: int = 32, train_blocks: int | None = None,
           split_blocks_per_cu: int = 1, watch_timeout_s: float = 600.0,
           order_probe: int = 10, order_radius: int = 2) -> RefineResult:
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

    # the seeded calibration and verification subsamples (disjoint) and, when
    # the fine grid holds a train_n x train_n lattice (stride s = (n-1) //
    # (train_n-1) >= 1 in both directions), the training map as that lattice,
    # so its values enter the final map
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
    first_mask = torch.zeros(n, dtype=torch.bool, device=dev)
    first_mask[verify] = True
    first_mask[calib] = True
    train_idx = None
    if lattice:
        ar = torch.arange(train_n, device=dev, dtype=torch.int64)
        train_idx = ((ar * s1)[:, None] * fd.n2 + (ar * s2)[None, :]).reshape(-1)
        first_mask[train_idx] = True
    # every other point in the kernels' spread order
    rest = torch.nonzero(~first_mask).flatten()
    rest = rest[spread_order(rest.numel(), dev)].contiguous()
    launches = []  # per sweep launch: name, kernel ms, points computed, accepted / rejected steps

    def launch_stats(name, kernel_ms, st, na, nr):
        computed = st != L.PT_SKIPPED
        launches.append(dict(name=name, kernel_ms=kernel_ms, points=int(computed.sum().item()),
                             acc_steps=float(na.double().sum().item()),
                             rej_steps=float(nr.double().sum().item())))

    lam_j = st_j = None
    if one_launch and lattice and one_launch_supported(fd):
        # One launch of the method over the lattice, the calibration and the
        # verification points and then every other point, on a side stream.
        # The FP64 judge goes first, on the main stream, on judge_blocks
        # 256-lane blocks, so its blocks hold their slots from the start (on
        # the 2048^2 Rossler plane its 41,943 points on 32 blocks, about five
        # a lane, end well before the launch: 7.88 s a step against 7.99 s on
        # 256 blocks, one point a lane, whose mostly idle waves held a third
        # of every SIMD's slots); the launch reports when the
        # lattice (milestone 0) and the calibration points (milestone 1) are
        # done, training and inference follow the judge on the main stream, and
        # the answer rule's points are skipped from the moment their skip
        # words land: no wave waits on a short launch's slowest point.  Points
        # taken before the rule was known are computed all the same; the map
        # takes the rule's answers there too, so it does not depend on timing.
        n_first = n_t + n_cal + n_ver
        if order_probe > 0:
            # the points after the first milestones heavy-first (a probe run
            # of the plane), so that no heavy point starts in the launch's last
            # round
            t_o = time.perf_counter()
            rest = heavy_first(ctx, fd, torch.nonzero(~first_mask).flatten(), order_probe, order_radius,
                               stream=main)
            main.synchronize()
            tm["order_probe"] = time.perf_counter() - t_o
        items = torch.cat([train_idx, calib, verify, rest]).contiguous()
        skip = torch.zeros(items.numel(), dtype=torch.int32, device=dev)
        side = side_streams(dev)[0]
        side.wait_stream(main)
        judge_ctx = None
        if judge:
            judge_ctx = L.Context(ctx.device)
            judge_ctx.set_option("lane_max_blocks", judge_blocks)
            lam_j, st_j, _, _ = judge_ctx.run(_judge_desc(fd), idx_list=verify, stream=main)
        ctx.set_option("item_order", 1)
        ctx.set_option("lane_gather", rerun_gather)
        if train_blocks is None:
            train_blocks = judge_blocks
        ctx.set_option("lane_max_blocks", -train_blocks)
        ctx.set_watch(n_t, n_t + n_cal, skip)
        try:
            with torch.cuda.stream(side):
                lam_a, st_a, na_a, nr_a = ctx.run(fd, idx_list=items, stream=side)
        finally:
            ctx.set_option("lane_max_blocks", 0)
            ctx.set_option("lane_gather", 0)
            ctx.set_option("item_order", 0)
        if not ctx.watch_wait(0, watch_timeout_s):
            raise RuntimeError(f"the training lattice did not finish within {watch_timeout_s} s")
        tm["train_map"] = time.perf_counter() - t
        t = time.perf_counter()
        lam_t, st_t = lam_a[:n_t], st_a[:n_t]
        net, losses = _train(dev, fd, grid_params(fd, train_idx), lam_t, st_t, epochs, batch, lr, seed,
                             hidden, fourier_sigma, max(batch, infer_batch))
        t = mark_stream("train", t, main)
        pred = _infer(net, fd, n, infer_batch, dev)
        t = mark_stream("infer", t, main)
        if not ctx.watch_wait(1, watch_timeout_s):
            raise RuntimeError(f"the calibration points did not finish within {watch_timeout_s} s")
        lam_c, st_c = lam_a[n_t:n_t + n_cal], st_a[n_t:n_t + n_cal]
        kappa, escapes, cal_agree, cal_answered, cal_table, flagged = _rule(
            pred, calib, lam_c, st_c, std_thr, lam_thr, bad_lo, bad_hi, calib_target, dlam_cal)
        # the skip words of the points the rule answers (a kernel on the main
        # stream; the launch reads them past the caches)
        skip[n_first:] = (~flagged[rest]).to(torch.int32)
        tm["calibrate"] = time.perf_counter() - t
        t = time.perf_counter()
        side.synchronize()
        main.synchronize()
        tm["rest"] = time.perf_counter() - t
        launch_stats("one launch", ctx.kernel_ms(), st_a, na_a, nr_a)
        if judge_ctx is not None:
            judge_ctx.close()
        lam_v, st_v = lam_a[n_t + n_cal:n_first], st_a[n_t + n_cal:n_first]
        # the map: the method's values wherever it ran, then the rule's answers
        # on the points outside the first three sets
        lam = torch.full((n, fd.n_exps), float("nan"), dtype=torch.float64, device=dev)
        status = torch.full((n,), L.PT_SKIPPED, dtype=torch.int32, device=dev)
        lam[items] = lam_a
        status[items] = st_a
    else:
        # The split schedule: the lattice (or a coarse training grid of its
        # own) on the main stream with the calibration and verification points
        # beside it on a side stream, each capped at split_blocks_per_cu
        # 256-lane blocks per CU; the FP64 judge on the other side stream while
        # the surrogate trains; then the rerun of the flagged points.
        side_c = side_streams(dev)[0]
        side_c.wait_stream(main)
        cv = torch.cat([calib, verify]).contiguous()
        cv_ctx = L.Context(ctx.device)
        cv_ctx.set_option("lane_max_blocks", split_blocks_per_cu * n_cus)
        if lattice:
            ctx.set_option("lane_max_blocks", split_blocks_per_cu * n_cus)
            try:
                lam_t, st_t, na_t, nr_t = ctx.run(fd, idx_list=train_idx)
            finally:
                ctx.set_option("lane_max_blocks", 0)
            p_t = grid_params(fd, train_idx)
        else:
            tspec = dict(fine)
            tspec.update(W.grid_fields(fd.sweep_idx1, fd.start1, end1, train_n, fd.sweep_idx2, fd.start2,
                                       end2, train_n))
            td = L.finalize(L.make_desc(tspec))
            lam_t, st_t, na_t, nr_t = ctx.run(td)
            p_t = grid_params(td, torch.arange(n_t, device=dev))
        with torch.cuda.stream(side_c):
            lam_cv, st_cv, na_cv, nr_cv = cv_ctx.run(fd, idx_list=cv, stream=side_c)
        t = mark_stream("train_map", t, main)
        launch_stats("training map", ctx.kernel_ms(), st_t, na_t, nr_t)
        judge_ctx = None
        if judge:
            side_j = side_streams(dev)[1]
            side_j.wait_stream(main)
            judge_ctx = L.Context(ctx.device)
            with torch.cuda.stream(side_j):
                lam_j, st_j, _, _ = judge_ctx.run(_judge_desc(fd), idx_list=verify, stream=side_j)
        net, losses = _train(dev, fd, p_t, lam_t, st_t, epochs, batch, lr, seed, hidden, fourier_sigma,
                             max(batch, infer_batch))
        t = mark_stream("train", t, main)
        pred = _infer(net, fd, n, infer_batch, dev)
        t = mark_stream("infer", t, main)
        main.wait_stream(side_c)
        lam_c, st_c = lam_cv[:n_cal], st_cv[:n_cal]
        lam_v, st_v = lam_cv[n_cal:], st_cv[n_cal:]
        kappa, escapes, cal_agree, cal_answered, cal_table, flagged = _rule(
            pred, calib, lam_c, st_c, std_thr, lam_thr, bad_lo, bad_hi, calib_target, dlam_cal)
        tm["calibrate"] = time.perf_counter() - t
        rerun_idx = rest[flagged[rest]].contiguous()
        ctx.set_option("item_order", 1)
        ctx.set_option("lane_gather", rerun_gather)
        try:
            lam_r, st_r, na_r, nr_r = ctx.run(fd, idx_list=rerun_idx)
        finally:
            ctx.set_option("lane_gather", 0)
            ctx.set_option("item_order", 0)
        t = mark_stream("rest", t, main)
        torch.cuda.synchronize(dev)
        launch_stats("calibration and verification", cv_ctx.kernel_ms(), st_cv, na_cv, nr_cv)
        if rerun_idx.numel():
            launch_stats("rerun", ctx.kernel_ms(), st_r, na_r, nr_r)
        cv_ctx.close()
        if judge_ctx is not None:
            judge_ctx.close()
        lam = torch.full((n, fd.n_exps), float("nan"), dtype=torch.float64, device=dev)
        status = torch.full((n,), L.PT_SKIPPED, dtype=torch.int32, device=dev)
        lam[rerun_idx] = lam_r
        status[rerun_idx] = st_r
        lam[cv] = torch.cat([lam_c, lam_v])
        status[cv] = torch.cat([st_c, st_v])
        if lattice:
            lam[train_idx] = lam_t
            status[train_idx] = st_t

    # the rule's answers on the points outside the first three sets (the
    # surrogate's mu and its escape class); the method's values elsewhere
    answered = ~flagged & ~first_mask
    pbad = pred[:, 4]
    sur_lam = pred[:, :3].to(torch.float64)
    sur_lam[pbad >= 0.5] = float("nan")
    sur_status = torch.where(pbad >= 0.5, L.PT_ESCAPED, L.PT_OK).to(torch.int32)
    computed = status != L.PT_SKIPPED
    speculative = int((answered & computed).sum().item())
    lam[answered] = sur_lam[answered]
    status[answered] = sur_status[answered]
    rerun_mask = ~answered
    tm["total_chosen"] = time.perf_counter() - t_all

    n_rerun = int(rerun_mask.sum().item())  # points whose value is the method's
    stats = dict(points=n, train_points=n_t, rerun=n_rerun,
                 flagged=int(flagged.sum().item()), verify=n_ver, calib=n_cal,
                 kappa=kappa, escape_answers=bool(escapes), calib_agreement=cal_agree,
                 calib_answered=cal_answered, calib_table=cal_table,
                 train_on_lattice=bool(lattice), one_launch=launches[0]["name"] == "one launch",
                 rerun_fraction=float(n_rerun) / n, losses=[float(x.item()) for x in losses],
                 computed_points=int(computed.sum().item()), speculative_points=speculative,
                 rerun_gather=int(rerun_gather), train_blocks=int(train_blocks or 0), launches=launches)
    if judge:
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

"""Surrogate-guided refinement of a sweep (the learned part of the sweep).

Pipeline, all on the device through liblyap.so:
  1. a coarse training map (train_n x train_n over the same parameter ranges)
     computed by the method in the chosen precision;
  2. the surrogate network (lyap_mlp_*) trained on it with Adam;
  3. inference over the fine grid;
  4. the method is rerun where the prediction is uncertain (std of lambda1
     above std_thr), near the chaos boundary (|mu1| < lam_thr), or where the
     escape probability is not decided, plus on a seeded random verification
     subsample (verify_frac of the grid);
  5. the plain FP64 method judges the verification subsample.
The returned map takes the method's values where it ran and the surrogate's
elsewhere.  Gates reported on the verification subsample: max |dlambda1|
between the chosen form and FP64 on computed points, and agreement of the
chaotic/regular/escaped class between the final map (surrogate answers
included) and FP64.
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


def refine(ctx, fine: dict, train_n: int = 256, epochs: int = 40, batch: int = 4096,
           lr: float = 1e-3, seed: int = 7, verify_frac: float = 0.01, std_thr: float = 0.01,
           lam_thr: float = 0.02, bad_lo: float = 0.02, bad_hi: float = 0.98,
           infer_batch: int = 65536, judge: bool = True, hidden: int = 256,
           fourier_sigma: float = 6.0) -> RefineResult:
    import torch

    dev = torch.device("cuda", ctx.device)
    tm = {}

    def mark(name, t0):
        torch.cuda.synchronize()
        tm[name] = time.perf_counter() - t0
        return time.perf_counter()

    fd = L.finalize(L.make_desc(fine))
    n = fd.n1 * fd.n2
    end1 = fd.start1 + fd.delta1 * (fd.n1 - 1)
    end2 = fd.start2 + fd.delta2 * (fd.n2 - 1)
    t_all = t = time.perf_counter()

    # 1. training map
    tspec = dict(fine)
    tspec.update(W.grid_fields(fd.sweep_idx1, fd.start1, end1, train_n, fd.sweep_idx2, fd.start2,
                               end2, train_n))
    td = L.finalize(L.make_desc(tspec))
    lam_t, st_t, _, _ = ctx.run(td)
    p_t = grid_params(td, torch.arange(train_n * train_n, device=dev))
    t = mark("train_map", t)

    # 2. training
    ok = st_t == 0
    lam_ok = lam_t[ok]
    ym = lam_ok.mean(dim=0).tolist() if lam_ok.shape[0] else [0.0, 0.0, 0.0]
    ys = lam_ok.std(dim=0).clamp_min(1e-6).tolist() if lam_ok.shape[0] > 1 else [1.0, 1.0, 1.0]
    net = L.Mlp(device=ctx.device, max_batch=max(batch, infer_batch), hidden=hidden,
                fourier_sigma=fourier_sigma, lr=lr, seed=seed,
                ranges=(min(fd.start1, end1), max(fd.start1, end1), min(fd.start2, end2),
                        max(fd.start2, end2)), y_mean=ym, y_std=ys)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    n_t = train_n * train_n
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(n_t, generator=gen).to(dev)
        for b0 in range(0, n_t - batch + 1, batch):
            sel = perm[b0:b0 + batch]
            net.train_step(p_t[sel].contiguous(), lam_t[sel].contiguous(), st_t[sel].contiguous(), loss)
        losses.append(loss.item())
    t = mark("train", t)

    # 3. inference
    pred = torch.empty((n, 5), dtype=torch.float32, device=dev)
    for c0 in range(0, n, infer_batch):
        c1 = min(n, c0 + infer_batch)
        net.forward(grid_params(fd, torch.arange(c0, c1, device=dev)), pred[c0:c1])
    t = mark("infer", t)

    # 4. rerun set and verification subsample
    mu1, std1, pbad = pred[:, 0], pred[:, 3], pred[:, 4]
    undecided = (pbad > bad_lo) & (pbad < bad_hi)
    flagged = (std1 > std_thr) | (mu1.abs() < lam_thr) | undecided | ~torch.isfinite(mu1)
    flagged &= ~(pbad >= bad_hi)  # confidently escaping: answered by the surrogate
    gv = torch.Generator(device="cpu").manual_seed(seed + 1)
    n_ver = max(1, int(round(verify_frac * n)))
    verify = torch.randperm(n, generator=gv)[:n_ver].sort().values.to(dev)
    rerun_mask = flagged.clone()
    rerun_mask[verify] = True
    rerun_idx = torch.nonzero(rerun_mask).flatten().to(torch.int64).contiguous()
    lam_r, st_r, na_r, nr_r = ctx.run(fd, idx_list=rerun_idx)
    t = mark("rerun", t)
    rerun_kernel_ms = ctx.kernel_ms()

    # 5. assemble
    lam = pred[:, :3].to(torch.float64)
    status = torch.where(pbad >= 0.5, L.PT_ESCAPED, L.PT_OK).to(torch.int32)
    lam[pbad >= 0.5] = float("nan")
    lam[rerun_idx] = lam_r
    status[rerun_idx] = st_r
    tm["total_chosen"] = time.perf_counter() - t_all

    stats = dict(points=n, train_points=n_t, rerun=int(rerun_idx.numel()),
                 flagged=int(flagged.sum().item()), verify=n_ver,
                 rerun_fraction=float(rerun_idx.numel()) / n, losses=losses,
                 rerun_kernel_ms=rerun_kernel_ms,
                 rerun_acc_steps=float(na_r.double().sum().item()),
                 rerun_rej_steps=float(nr_r.double().sum().item()))
    if judge:
        t = time.perf_counter()
        pdesc = L.SweepDesc.from_buffer_copy(fd)
        pdesc.precision = L.PREC_FP64
        lam_j, st_j, _, _ = ctx.run(pdesc, idx_list=verify.contiguous())
        t = mark("judge", t)
        v = verify.cpu().numpy()
        lj, sj = lam_j.cpu().numpy(), st_j.cpu().numpy()
        # chosen-form computed values at the verification points
        pos = {int(k): i for i, k in enumerate(rerun_idx.cpu().numpy())}
        lr_np, sr_np = lam_r.cpu().numpy(), st_r.cpu().numpy()
        lc = np.array([lr_np[pos[k]] for k in v])
        sc = np.array([sr_np[pos[k]] for k in v])
        both = (sc == 0) & (sj == 0)
        stats["max_dlambda1_chosen_vs_fp64"] = float(np.max(np.abs(lc[both, 0] - lj[both, 0]))) if both.any() else 0.0
        stats["max_dlambda_chosen_vs_fp64"] = float(np.max(np.abs(lc[both] - lj[both]))) if both.any() else 0.0
        # the final map's class at the verification points, answering with the
        # surrogate wherever the point was not flagged
        fl = flagged.cpu().numpy()[v]
        pr = pred.cpu().numpy()[v]
        sur_status = np.where(pr[:, 4] >= 0.5, 2, 0)
        cls_final = np.where(fl, classify(lc[:, 0], sc), classify(pr[:, 0], sur_status))
        cls_judge = classify(lj[:, 0], sj)
        stats["class_agreement"] = float(np.mean(cls_final == cls_judge))
        stats["surrogate_answered_verify"] = int((~fl).sum())
        stats["surrogate_class_agreement"] = float(np.mean(cls_final[~fl] == cls_judge[~fl])) if (~fl).any() else 1.0
        sur_ok = (~fl) & (sj == 0) & (sur_status == 0)
        stats["surrogate_max_dlambda1"] = float(np.max(np.abs(pr[sur_ok, 0] - lj[sur_ok, 0]))) if sur_ok.any() else 0.0
    net.close()
    return RefineResult(lam=lam.cpu().numpy(), status=status.cpu().numpy(),
                        rerun=rerun_mask.cpu().numpy(), timings=tm, stats=stats)

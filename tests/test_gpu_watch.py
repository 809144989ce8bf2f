"""Progress signals and skips of one launch (lyap_ctx_set_watch), through the
C-ABI: the milestones are reported while the launch runs, skipped items carry
LYAP_PT_SKIPPED, NaN exponents and no steps, every other item is the plain
launch's bit for bit, and the watch refuses what it does not cover."""
import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import refine as R
from chaos_oneapi_amd import workloads as W

pytestmark = pytest.mark.gpu


def host(t):
    return tuple(x.cpu().numpy() for x in t)


@pytest.mark.parametrize("method,integ,prec", [(L.METHOD_VAR_TRACE, L.INT_DVERK, L.PREC_MIXED),
                                               (L.METHOD_VAR, L.INT_DOPRI5, L.PREC_FP64)])
def test_watch_milestones_and_preset_skips(gpu_ctx, method, integ, prec):
    spec = dict(W.rossler_dverk(64, precision=prec, t_transient=5.0, t_average=20.0), method=method,
                integrator=integ)
    desc = L.finalize(L.make_desc(spec))
    n = desc.n1 * desc.n2
    print(f"case watch rossler 64x64 method {method} integrator {integ} precision {prec}", flush=True)
    items = R.spread_order(n, torch.device("cuda", 0))
    plain = host(gpu_ctx.run(desc, idx_list=items))
    end0, end1 = 700, 1500
    skip = torch.zeros(n, dtype=torch.int32, device="cuda")
    pick = np.arange(n) % 3 == 0
    pick[:end1] = False
    skip[torch.as_tensor(np.flatnonzero(pick), device="cuda")] = 1
    gpu_ctx.set_option("item_order", 1)
    try:
        gpu_ctx.set_watch(end0, end1, skip)
        got = gpu_ctx.run(desc, idx_list=items)
        assert gpu_ctx.watch_wait(0, 60.0) and gpu_ctx.watch_wait(1, 60.0)
        torch.cuda.synchronize()
    finally:
        gpu_ctx.set_option("item_order", 0)
    lam, st, na, nr = host(got)
    assert np.all(st[pick] == L.PT_SKIPPED) and np.all(np.isnan(lam[pick]))
    assert np.all(na[pick] == 0) and np.all(nr[pick] == 0)
    keep = ~pick
    np.testing.assert_array_equal(st[keep], plain[1][keep])
    np.testing.assert_array_equal(lam[keep], plain[0][keep])
    np.testing.assert_array_equal(na[keep], plain[2][keep])
    # the next run on the context is not watched: every item computed
    lam2, st2, _, _ = host(gpu_ctx.run(desc, idx_list=items))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lam2, plain[0])


def test_watch_refuses_what_it_does_not_cover(gpu_ctx):
    fixed = L.finalize(L.make_desc(W.lorenz_fixed(8)))
    gpu_ctx.set_watch(4, 8)
    with pytest.raises(L.LyapError):
        gpu_ctx.run(fixed)
    ada = L.finalize(L.make_desc(W.rossler_dverk(8, t_transient=1.0, t_average=2.0)))
    gpu_ctx.set_watch(4, 100)
    with pytest.raises(ValueError, match="beyond"):
        gpu_ctx.run(ada)
    gpu_ctx.set_option("lane_quad", 1)
    try:
        gpu_ctx.set_watch(4, 8)
        with pytest.raises(L.LyapError):
            gpu_ctx.run(ada)
    finally:
        gpu_ctx.set_option("lane_quad", 0)
    torch.cuda.synchronize()


def test_one_launch_and_split_schedules_give_the_same_map(gpu_ctx):
    """The surrogate refinement's two schedules on a small plane: the same
    training data, network, rule and method values, so the same map bit for
    bit, although the one launch skips points while it runs."""
    spec = W.rossler_dverk(192, precision=L.PREC_MIXED, t_transient=20.0, t_average=100.0)
    spec = dict(spec, method=L.METHOD_VAR_TRACE)
    print("case refine schedules rossler 192x192", flush=True)
    kw = dict(train_n=64, epochs=60, batch=1024, infer_batch=65536)
    a = R.refine(gpu_ctx, spec, one_launch=True, **kw)
    b = R.refine(gpu_ctx, spec, one_launch=False, **kw)
    print({k: a.stats[k] for k in ("one_launch", "rerun", "computed_points", "speculative_points")},
          {k: b.stats[k] for k in ("one_launch", "rerun", "computed_points", "speculative_points")}, flush=True)
    assert a.stats["one_launch"] and not b.stats["one_launch"]
    assert a.stats["kappa"] == b.stats["kappa"] and a.stats["escape_answers"] == b.stats["escape_answers"]
    np.testing.assert_array_equal(a.status, b.status)
    np.testing.assert_array_equal(a.lam, b.lam)
    np.testing.assert_array_equal(a.rerun, b.rerun)
    for res in (a, b):
        st = res.stats
        assert st["computed_points"] == st["rerun"] + st["speculative_points"]
    np.testing.assert_array_equal(a.verify_fp64, b.verify_fp64)

"""Batched mixed-model launch (lyap_batch_*): one persistent kernel over
Lorenz, Rossler and Lorenz-96 N=8/16/32 classes, captured in a hipGraph.

Bar: bit-exact against the CPU oracle per class (same operations in the same
order as the single-class kernels), bit-exact against the single-class
kernels for every built layout, queue order and grid size, and identical
results on every graph replay."""
import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def device_classes(mix, precision=None, trace=False):
    dev = torch.device("cuda", 0)
    classes = []
    for c in mix:
        spec = dict(c["spec"])
        if precision is not None:
            spec["precision"] = precision
        if trace and spec["dim"] == 3:
            spec["method"] = L.METHOD_VAR_TRACE
        desc = L.finalize(L.make_desc(spec))
        classes.append(dict(desc=desc, count=c["count"],
                            params=torch.from_numpy(c["params"]).to(dev).contiguous()))
    return classes


def to_host(outs):
    return [tuple(t.cpu().numpy() for t in o) for o in outs]


def assert_same(a, b, what):
    lam_a, st_a, na_a, nr_a = a
    lam_b, st_b, na_b, nr_b = b
    np.testing.assert_array_equal(st_a.astype(np.uint32), st_b.astype(np.uint32), err_msg=what)
    np.testing.assert_array_equal(na_a.astype(np.uint32), na_b.astype(np.uint32), err_msg=what)
    np.testing.assert_array_equal(nr_a.astype(np.uint32), nr_b.astype(np.uint32), err_msg=what)
    np.testing.assert_array_equal(lam_a, lam_b, err_msg=what)  # NaN == NaN here


@pytest.mark.parametrize("prec,trace", [(L.PREC_MIXED, False), (L.PREC_FP64, False), (L.PREC_MIXED, True)])
def test_batch_matches_oracle(gpu_ctx, prec, trace):
    """trace: the 3-D classes run the trace closure (two vectors), the
    Lorenz-96 classes every vector, in the same launch."""
    mix = W.batch_mix(1500, seed=11, time_scale=0.01)
    classes = device_classes(mix, prec, trace)
    print("case batch-vs-oracle prec", prec, "trace", trace, [(c["name"], c["count"]) for c in mix],
          flush=True)
    b = L.Batch(gpu_ctx, classes)
    outs = to_host(b.launch())
    torch.cuda.synchronize()
    for c, cl, g in zip(mix, classes, outs):
        od = O.make_desc(cl["desc"].to_dict())
        o = O.sweep(od, count=c["count"], params_in=c["params"])
        assert_same(g, o, c["name"])
        ok = o[1] == 0
        print(f"{c['name']}: {c['count']} points, {ok.sum()} ok, exact", flush=True)
    b.close()


TUNINGS = [dict(layout=0, occupancy3=False, order=0), dict(layout=0, occupancy3=True, order=0),
           dict(layout=1, occupancy3=False, order=2), dict(layout=1, occupancy3=True, order=1),
           dict(layout=0, occupancy3=False, order=0, waves_per_cu=1),
           dict(layout=1, occupancy3=True, order=0, gather=24)]


@pytest.mark.parametrize("trace", [False, True])
def test_batch_matches_single_class_kernels(gpu_ctx, trace):
    mix = W.batch_mix(6000, seed=12, time_scale=0.01)
    classes = device_classes(mix, trace=trace)
    ref = []
    for cl in classes:
        lam, st, na, nr = gpu_ctx.run(cl["desc"], count=cl["count"], params=cl["params"])
        ref.append((lam.clone(), st.clone(), na.clone(), nr.clone()))
    torch.cuda.synchronize()
    ref = to_host(ref)
    for tu in TUNINGS:
        print("case batch tuning", tu, flush=True)
        b = L.Batch(gpu_ctx, classes, **tu)
        outs = to_host(b.launch())
        torch.cuda.synchronize()
        for c, r, g in zip(mix, ref, outs):
            assert_same(g, r, f"{c['name']} {tu}")
        print(f"  waves {b.waves}, {b.last_ms():.1f} ms", flush=True)
        b.close()


def test_batch_graph_replays(gpu_ctx):
    mix = W.batch_mix(3000, seed=13, time_scale=0.005)
    classes = device_classes(mix)
    b = L.Batch(gpu_ctx, classes)
    first = to_host([tuple(t.clone() for t in o) for o in b.launch()])
    torch.cuda.synchronize()
    for rep in range(3):
        print("case batch replay", rep, flush=True)
        for o in b.outputs:  # clobber the outputs with eager work between replays
            o[0].fill_(-1.0)
            o[1].zero_()
        torch.cuda.synchronize()
        outs = to_host(b.launch())
        torch.cuda.synchronize()
        for c, f, g in zip(mix, first, outs):
            assert_same(g, f, f"{c['name']} replay {rep}")
    b.close()


def test_batch_rejects_bad_classes(gpu_ctx):
    mix = W.batch_mix(200, seed=14, time_scale=0.005)
    classes = device_classes(mix)
    bad = [dict(c) for c in classes]
    d = L.SweepDesc.from_buffer_copy(bad[1]["desc"])
    d.integrator = L.INT_DVERK
    bad[1]["desc"] = d
    with pytest.raises(L.LyapError, match="must match class 0"):
        L.Batch(gpu_ctx, bad)
    dv = [dict(c) for c in classes]
    for c in dv:
        c["desc"] = L.SweepDesc.from_buffer_copy(c["desc"])
        c["desc"].integrator = L.INT_DVERK
    with pytest.raises(L.LyapError, match="adaptive Dopri5 only"):
        L.Batch(gpu_ctx, dv, layout=1)
    with pytest.raises(L.LyapError, match="1 to 16 classes"):
        L.Batch(gpu_ctx, classes * 4)
    nogrid = [dict(classes[0])]
    nogrid[0]["params"] = None
    with pytest.raises(L.LyapError, match="beyond the"):
        L.Batch(gpu_ctx, nogrid)


@pytest.mark.parametrize("trace", [False, True])
def test_dverk_batch_matches_oracle(gpu_ctx, trace):
    """The batch built for adaptive DVERK: its 3-D classes in the mixed form
    take the fifth-order binary32 tangent (as the single-class sweep does), the
    Lorenz-96 classes the sixth-order one; every class bit for bit against the
    oracle."""
    mix = W.batch_mix(600, seed=19, time_scale=0.005)
    classes = device_classes(mix, L.PREC_MIXED, trace)
    for c in classes:
        d = L.SweepDesc.from_buffer_copy(c["desc"])
        d.integrator = L.INT_DVERK
        c["desc"] = L.finalize(d)
    print("case dverk batch-vs-oracle trace", trace, [(c["name"], c["count"]) for c in mix], flush=True)
    b = L.Batch(gpu_ctx, classes)
    outs = to_host(b.launch())
    torch.cuda.synchronize()
    b.close()
    for c, cl, g in zip(mix, classes, outs):
        od = O.make_desc(cl["desc"].to_dict())
        o = O.sweep(od, count=c["count"], params_in=c["params"])
        assert_same(g, o, c["name"])

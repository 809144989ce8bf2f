"""lyap_batch_run_noalloc: the batched launch that allocates nothing (class
records and work counters in the library's fixed device array).  Its outputs
must equal the allocating path's (lyap_batch_create + lyap_batch_launch)
bit for bit, on two calls in a row, the second with a smaller batch (the
array keeps the first call's records and counters: the second call must
restage and reset them)."""
import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W

pytestmark = pytest.mark.gpu


def device_classes(mix):
    dev = torch.device("cuda", 0)
    out = []
    for c in mix:
        desc = L.finalize(L.make_desc(c["spec"]))
        n = c["count"]
        out.append(dict(desc=desc, count=n, params=torch.from_numpy(c["params"]).to(dev).contiguous(),
                        outputs=(torch.full((n, desc.n_exps), -3.0, dtype=torch.float64, device=dev),
                                 torch.full((n,), 77, dtype=torch.int32, device=dev),
                                 torch.full((n,), 77, dtype=torch.int32, device=dev),
                                 torch.full((n,), 77, dtype=torch.int32, device=dev))))
    return out


def host(outs):
    return [tuple(t.cpu().numpy() for t in o) for o in outs]


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = L.Context(0)
    yield c
    c.close()


CALLS = [  # (points, seed, tuning): the second call is smaller
    (1500, 31, dict()),
    (600, 32, dict(layout=1, occupancy3=True, gather=4)),
]


def test_noalloc_matches_allocating_path_twice(ctx):
    for n_points, seed, tuning in CALLS:
        mix = W.batch_mix(n_points, seed=seed, time_scale=0.01)
        print(f"case noalloc batch {n_points} points seed {seed} tuning {tuning}: "
              + ", ".join(f"{c['name']} {c['count']}" for c in mix), flush=True)
        classes = device_classes(mix)
        L.batch_run_noalloc(0, classes, **tuning)
        torch.cuda.synchronize()
        got = host([c["outputs"] for c in classes])
        ref_b = L.Batch(ctx, [dict(desc=c["desc"], count=c["count"], params=c["params"]) for c in classes],
                        **tuning)
        want = host(ref_b.launch())
        torch.cuda.synchronize()
        ref_b.close()
        for m, g, w in zip(mix, got, want):
            for a, b in zip(g, w):
                np.testing.assert_array_equal(a, b, err_msg=m["name"])
            assert np.all(g[1] != 77) and np.isfinite(g[0][g[1] == 0]).all()
            print(f"{m['name']}: {m['count']} points equal, {int((g[1] == 0).sum())} ok", flush=True)


def test_noalloc_rejects_more_classes_than_its_array(ctx):
    mix = W.batch_mix(64, seed=33, time_scale=0.01)
    classes = device_classes(mix)
    with pytest.raises(L.LyapError) as e:
        L.batch_run_noalloc(0, (classes * 4)[:17])
    assert e.value.code == L.ERR_INVALID and "scratch" in str(e.value)


def test_noalloc_under_graph_capture(ctx):
    """The allocation-free launch captured into a HIP graph (its staging and
    reset are kernels, no memset nodes) and replayed several times, with eager
    work between the replays: every replay equals the eager call."""
    mix = W.batch_mix(700, seed=34, time_scale=0.01)
    print("case noalloc under graph capture: " + ", ".join(f"{c['name']} {c['count']}" for c in mix),
          flush=True)
    classes = device_classes(mix)
    L.batch_run_noalloc(0, classes)  # eager (also the device's one-time set-up)
    torch.cuda.synchronize()
    want = host([c["outputs"] for c in classes])
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side, capture_error_mode="thread_local"):
        L.batch_run_noalloc(0, classes, stream=side)
    for rep in range(3):
        for c in classes:  # eager work between replays: poison the outputs
            for t in c["outputs"]:
                t.fill_(-5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = host([c["outputs"] for c in classes])
        for m, a_cls, b_cls in zip(mix, got, want):
            for a, b in zip(a_cls, b_cls):
                np.testing.assert_array_equal(a, b, err_msg=f"replay {rep} {m['name']}")
        print(f"replay {rep}: equal", flush=True)
    del g

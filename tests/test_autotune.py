"""Autotuner host logic (candidate selection, cache, sampling) with a stand-in
timer; the GPU timing path runs in tests/test_gpu_batch.py."""
import numpy as np
import torch

from chaos_oneapi_amd import autotune as A
from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W


def cpu_classes(n=5000, seed=3):
    out = []
    for c in W.batch_mix(n, seed=seed, time_scale=0.01):
        out.append(dict(desc=L.finalize(L.make_desc(c["spec"])), count=c["count"],
                        params=torch.from_numpy(c["params"]), name=c["name"]))
    return out


def test_pick_best_prefers_fastest_then_first():
    assert A.pick_best({"a": 3.0, "b": 1.0, "c": 2.0}) == "b"
    assert A.pick_best({"a": 1.0, "b": 1.0}) == "a"


def test_sample_indices_seeded_sorted_unique():
    a = A.sample_indices(10000, 1 / 16, 256, 9)
    b = A.sample_indices(10000, 1 / 16, 256, 9)
    assert len(a) == 625 and np.array_equal(a, b)
    assert np.all(np.diff(a) > 0)
    assert len(A.sample_indices(100, 1 / 16, 256, 9)) == 100  # small classes run whole


def test_sample_classes_keeps_rows():
    cls = cpu_classes()
    s = A.sample_classes(cls, frac=0.1, min_points=10, min_total=100)
    for c, sc in zip(cls, s):
        idx = A.sample_indices(c["count"], 0.1, 10, 5 + cls.index(c))
        assert sc["count"] == len(idx)
        assert torch.equal(sc["params"], c["params"][torch.as_tensor(idx)])


def test_tune_batch_picks_and_caches(tmp_path):
    cls = cpu_classes()
    cache = A.Cache(str(tmp_path / "tune.json"))
    cost = {(0, False, 0, 0, 0): 10.0, (0, True, 0, 0, 0): 7.0, (1, False, 0, 0, 0): 9.0,
            (1, True, 0, 0, 0): 8.0, (1, True, 0, 0, 8): 7.5, (1, True, 0, 0, 32): 7.2,
            (1, True, 0, 2, 0): 12.0}
    calls = []

    def timer(cand, sample):
        calls.append(cand)
        # 5000 points < the minimum sample: the whole batch is timed
        assert sum(c["count"] for c in sample) == sum(c["count"] for c in cls)
        return cost[cand]

    r = A.tune_batch(None, cls, cache=cache, timer=timer, frac=0.1, min_points=10)
    assert r.best == (0, True, 0, 0, 0) and not r.cached
    assert len(calls) == len(A.BATCH_CANDIDATES)
    r2 = A.tune_batch(None, cls, cache=A.Cache(str(tmp_path / "tune.json")), timer=timer)
    assert r2.cached and tuple(r2.best) == (0, True, 0, 0, 0)
    assert len(calls) == len(A.BATCH_CANDIDATES)  # served from the file
    assert A.batch_kwargs(r2.best) == dict(layout=0, occupancy3=True, waves_per_cu=0, order=0,
                                           gather=0)


def test_tune_batch_skips_unbuilt_variants(tmp_path):
    cls = cpu_classes()

    def timer(cand, sample):
        if cand[0] == 1:
            raise L.LyapError(L.ERR_UNSUPPORTED, "not built")
        return 5.0 - cand[1]

    r = A.tune_batch(None, cls, cache=A.Cache(None), timer=timer, frac=0.1, min_points=10)
    assert r.best == (0, True, 0, 0, 0)
    assert all(c[0] == 0 for c in r.timings)


def test_batch_key_buckets_counts():
    cls = cpu_classes()
    k1 = A.batch_key("dev", cls)
    more = [dict(c, count=c["count"] + 1) for c in cls]
    assert A.batch_key("dev", more) == k1 or any(
        A._bucket(c["count"]) != A._bucket(c["count"] + 1) for c in cls)
    assert A.batch_key("other", cls) != k1


def test_tune_l96_candidates_cover_built_layouts(tmp_path):
    spec = W.lorenz96_dopri(n=32, dim=16)
    desc = L.finalize(L.make_desc(spec))
    seen = []

    def timer(cand, idx):
        seen.append(cand)
        assert len(idx) == 1024  # the whole 32x32 grid: below the sample floor
        return {"waves": 3.0, "ppw": 2.0}[cand[0]] + 0.1 * cand[1]

    r = A.tune_l96(None, desc, cache=A.Cache(None), timer=timer)
    assert r.best == ("ppw", 1)
    assert seen == list(A.L96_CANDIDATES[16])


def test_tune_lanes_grid_caps():
    desc = L.finalize(L.make_desc(W.rossler_dverk(n=64)))
    r = A.tune_lanes(None, desc, cache=A.Cache(None),
                     timer=lambda c, idx: 2.0 - 0.01 * c[1] + (c[0] > 0) + 0.5 * c[2])
    assert r.best == (0, 16, 0)
    assert set(r.timings) == {(0, 0, 0), (0, 4, 0), (0, 8, 0), (0, 16, 0), (512, 0, 0), (512, 8, 0), (0, 0, 1)}
    # a small grid whose launch is its slowest point: four lanes per point wins
    r = A.tune_lanes(None, desc, cache=A.Cache(None), timer=lambda c, idx: 2.0 - 1.5 * c[2])
    assert r.best == (0, 0, 1)


def test_sample_keeps_a_minimum_total():
    cls = cpu_classes(n=5000)
    s = A.sample_classes(cls, frac=1 / 16, min_points=1, min_total=2500)
    tot = sum(c["count"] for c in s)
    assert 2400 <= tot <= 2600


def test_scaled_classes_shrink_times_and_keep_points():
    cls = []
    for c in W.batch_mix(2000, seed=4):
        cls.append(dict(desc=L.finalize(L.make_desc(c["spec"])), count=c["count"],
                        params=torch.from_numpy(c["params"]), name=c["name"]))
    sc = A.scaled_classes(cls, 0.1)
    for c, s in zip(cls, sc):
        assert s["count"] == c["count"] and s["params"] is c["params"]
        assert s["desc"].n_renorm_transient * 10 == c["desc"].n_renorm_transient
        assert s["desc"].n_renorm_average * 10 == c["desc"].n_renorm_average
        assert s["desc"].t_renorm == c["desc"].t_renorm and s["desc"].rtol == c["desc"].rtol
        assert c["desc"].t_average > 1.0  # the originals are untouched
    seen = []

    def timer(cand, sample):
        seen.append(sum(x["count"] for x in sample))
        assert sample[0]["desc"].n_renorm_average * 10 == cls[0]["desc"].n_renorm_average
        return 1.0 + cand[1]

    r = A.tune_batch(None, cls, cache=A.Cache(None), timer=timer, time_scale=0.1)
    assert r.best == (0, False, 0, 0, 0)
    assert seen == [sum(c["count"] for c in cls)] * len(A.BATCH_CANDIDATES)  # every point, every candidate


def test_tune_l96_time_scaled_runs_the_whole_grid():
    desc = L.finalize(L.make_desc(W.lorenz96_dopri(n=32, dim=16)))
    sd = A.scaled_desc(desc, 0.1)
    assert sd.n_renorm_average * 10 == desc.n_renorm_average
    assert sd.n_renorm_transient * 10 == desc.n_renorm_transient
    seen = []

    def timer(cand, idx):
        seen.append(idx)
        return 1.0 if cand == ("ppw", 4) else 2.0

    r = A.tune_l96(None, desc, cache=A.Cache(None), timer=timer, time_scale=0.1)
    assert r.best == ("ppw", 4)
    assert all(i is None for i in seen)  # no point subset: the scaled whole grid


def test_tune_lanes_time_scaled_runs_the_whole_grid():
    desc = L.finalize(L.make_desc(W.rossler_dverk(n=64)))
    seen = []

    def timer(cand, idx):
        seen.append(idx)
        return 1.0 if cand == (512, 8, 0) else 2.0

    r = A.tune_lanes(None, desc, cache=A.Cache(None), timer=timer, time_scale=0.1)
    assert r.best == (512, 8, 0)
    assert len(seen) == 7 and all(i is None for i in seen)


def test_order_candidates_need_an_adaptive_grid_of_few_rounds():
    c2 = L.finalize(L.make_desc(W.rossler_dverk(n=512)))
    assert A.resident_points(c2, 256) == 256 * 4 * 3 * 64
    assert A.order_candidates(c2, 512 * 512, A.resident_points(c2, 256)) == A.ORDER_CANDIDATES
    # 21 points a lane still pays; 85 do not (every heavy point starts early enough)
    assert A.order_candidates(c2, 2048 * 2048, A.resident_points(c2, 256)) == A.ORDER_CANDIDATES
    assert A.order_candidates(c2, 8192 * 2048, A.resident_points(c2, 256)) == (0,)
    # fixed steps: every point costs the same
    c1 = L.finalize(L.make_desc(W.lorenz_fixed(64)))
    assert A.order_candidates(c1, 4096, A.resident_points(c1, 256)) == (0,)
    c3 = L.finalize(L.make_desc(W.lorenz96_dopri(n=256, dim=16)))
    assert A.resident_points(c3, 256, l96_ppw=4) == 256 * 4 * 3 * 4
    assert A.order_candidates(c3, 65536, A.resident_points(c3, 256, l96_ppw=4)) == A.ORDER_CANDIDATES


def test_tune_order_times_every_candidate_and_keeps_the_fastest():
    c2 = L.finalize(L.make_desc(W.rossler_dverk(n=512)))
    seen = []

    def timer(c):
        seen.append(c)
        return {0: 950.0, 10: 800.0, 30: 770.0}[c]

    r = A.tune_order(None, c2, A.resident_points(c2, 256), cache=A.Cache(None), timer=timer)
    assert r.best == (30,) and seen == list(A.ORDER_CANDIDATES)
    big = L.finalize(L.make_desc(W.rossler_dverk(n=4096)))
    r = A.tune_order(None, big, A.resident_points(big, 256) // 4, cache=A.Cache(None), timer=timer)
    assert r.best == (0,) and r.timings == {} and len(seen) == 3

"""CPU model of the persistent work-stealing queues, stress-tested with random
interleavings and straggler-heavy costs (no GPU).

It restates, step for step, the scheduling rules of the kernels:
  - grab_items (csrc/lyap_lane.hpp): the lanes of a wave that want work are
    counted by a ballot, the wave's first wanting lane does one atomicAdd of
    that count, and the lanes take consecutive items in lane order;
  - item_position (csrc/lyap_device.hpp): item i is grid point (i * s) mod n
    with s from spread_stride (csrc/lyap_abi.cpp), coprime to n;
  - lane_drain: a lane that finishes takes the next item; the wave leaves when
    the counter is exhausted and no lane is active;
  - l96_wave_drain: the same per slot of 64/P lanes (the slot's first lane
    asks, the slot's lanes share the item);
  - k_batch: every wave walks the classes in queue order and drains each one
    through that class's own counter.
The interleaving of the waves' atomic operations and the per-item costs are
random; the checks are that every item is processed exactly once, that every
wave terminates, and that the positions are a permutation of the grid.
"""
import math
import random

import numpy as np
import pytest


def spread_stride(count: int) -> int:
    if count <= 2:
        return 1
    s = max(1, int(count * 0.6180339887498949))
    while math.gcd(s, count) != 1:
        s += 1
    return s


def item_position(item: int, stride: int, count: int) -> int:
    return (item * stride) % count


class Counter:
    def __init__(self):
        self.value = 0

    def atomic_add(self, k: int) -> int:
        old = self.value
        self.value += k
        return old


class Wave:
    """One wave64 draining one queue with `slots` independent slots (64 for
    the per-lane drain, P for the Lorenz-96 wave drain)."""

    def __init__(self, wid, slots, counter, count, stride, cost, done):
        self.wid, self.slots = wid, slots
        self.counter, self.count, self.stride = counter, count, stride
        self.cost, self.done = cost, done
        self.remaining = [0] * slots     # work left on the slot's current item
        self.item = [-1] * slots
        self.active = [False] * slots
        self.exhausted = [False] * slots
        self.finished = False

    def iteration(self):
        """One pass of the drain loop; returns False once the wave has left."""
        want = [not self.active[j] and not self.exhausted[j] for j in range(self.slots)]
        n_want = sum(want)
        if n_want:
            base = self.counter.atomic_add(n_want)  # one atomic for the wave
            k = 0
            for j in range(self.slots):
                if want[j]:
                    got = base + k
                    k += 1
                    if got < self.count:
                        pos = item_position(got, self.stride, self.count)
                        self.item[j] = pos
                        self.remaining[j] = self.cost(pos)
                        self.active[j] = True
                    else:
                        self.exhausted[j] = True
        if not any(self.active):
            self.finished = True
            return False
        for j in range(self.slots):
            if self.active[j]:
                self.remaining[j] -= 1
                if self.remaining[j] <= 0:
                    self.done.append(self.item[j])
                    self.active[j] = False
        return True


def run_queue(count, n_waves, slots, cost, rng):
    counter, done = Counter(), []
    stride = spread_stride(count)
    waves = [Wave(w, slots, counter, count, stride, cost, done) for w in range(n_waves)]
    live = list(waves)
    steps = 0
    while live:
        w = rng.choice(live)  # random interleaving of the waves
        if not w.iteration():
            live.remove(w)
        steps += 1
        assert steps < 50_000_000, "the queue does not terminate"
    return done, waves, counter


def straggler_cost(rng, count):
    # heavy tail: most points 1-20 iterations, a few 10-100x longer
    c = rng.integers(1, 21, count)
    heavy = rng.random(count) < 0.03
    c[heavy] *= rng.integers(10, 100, heavy.sum())
    return lambda pos: int(c[pos])


@pytest.mark.parametrize("count", [1, 2, 3, 63, 64, 65, 1000, 4097])
def test_positions_are_a_permutation(count):
    s = spread_stride(count)
    assert math.gcd(s, count) == 1
    pos = sorted(item_position(i, s, count) for i in range(count))
    assert pos == list(range(count))


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("slots", [64, 8, 4, 2, 1])
def test_drain_processes_every_item_once(seed, slots):
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    count = int(nrng.integers(1, 3000))
    n_waves = int(nrng.integers(1, 40))
    done, waves, counter = run_queue(count, n_waves, slots, straggler_cost(nrng, count), rng)
    assert sorted(done) == list(range(count))
    assert all(w.finished for w in waves)
    # every slot asked until it saw the end: the counter went past the count
    assert counter.value >= count


def test_wave_aggregated_grab_gives_consecutive_items_in_lane_order():
    counter = Counter()
    counter.value = 100
    w = Wave(0, 64, counter, 10**6, 1, lambda pos: 5, [])
    w.active = [j % 3 == 0 for j in range(64)]  # a third of the lanes busy
    w.iteration()
    got = [w.item[j] for j in range(64) if j % 3 != 0]
    assert got == list(range(100, 100 + len(got)))
    assert counter.value == 100 + len(got)


def test_batch_walks_classes_and_drains_each_once():
    """k_batch: waves walk the classes in queue order, each through its own
    counter; a wave moves on once a class is exhausted and its lanes are done."""
    nrng = np.random.default_rng(11)
    rng = random.Random(11)
    classes = [dict(count=int(nrng.integers(1, 2000)), slots=s) for s in (64, 64, 1, 4, 8)]
    for c in classes:
        c["counter"], c["done"] = Counter(), []
        c["stride"] = spread_stride(c["count"])
        c["cost"] = straggler_cost(nrng, c["count"])
    n_waves = 23
    phase = [0] * n_waves
    cur = [None] * n_waves
    live = list(range(n_waves))
    steps = 0
    while live:
        w = rng.choice(live)
        if cur[w] is None:
            if phase[w] == len(classes):
                live.remove(w)
                continue
            c = classes[phase[w]]
            cur[w] = Wave(w, c["slots"], c["counter"], c["count"], c["stride"], c["cost"], c["done"])
        if not cur[w].iteration():
            cur[w] = None
            phase[w] += 1
        steps += 1
        assert steps < 50_000_000
    for c in classes:
        assert sorted(c["done"]) == list(range(c["count"]))

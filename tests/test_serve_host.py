"""Host logic of the serving loop (CPU): the first-fit range allocator and request sizing."""

import numpy as np

import serve


def test_range_alloc_first_fit_and_merge():
    a = serve.RangeAlloc(100)
    x = a.alloc(30)
    y = a.alloc(30)
    z = a.alloc(30)
    assert (x, y, z) == (0, 30, 60) and a.alloc(20) is None and a.extent() == 90
    a.release(y)
    assert a.alloc(31) is None and a.alloc(10) == 30  # first fit into the hole
    a.release(x)
    a.release(30)
    assert a.free[0] == (0, 60)  # merged with the neighbouring hole
    assert a.extent() == 90 and a.in_use() == 30
    a.release(z)
    assert a.free == [(0, 100)] and a.extent() == 0


def test_range_alloc_random_sequences_never_overlap():
    rng = np.random.default_rng(3)
    a = serve.RangeAlloc(1000)
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            s = list(live)[rng.integers(len(live))]
            a.release(s)
            del live[s]
        else:
            n = int(rng.integers(1, 80))
            s = a.alloc(n)
            if s is not None:
                live[s] = n
        spans = sorted(live.items())
        for (s0, n0), (s1, _) in zip(spans, spans[1:]):
            assert s0 + n0 <= s1
        assert a.in_use() == sum(live.values()) and a.in_use() + sum(n for _, n in a.free) == 1000

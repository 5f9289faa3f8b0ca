"""Index logic of the segmented gathers, restated on the CPU.

The far / near / Dirichlet kernels split a CSR list (segment s owns pairs
start[s]..start[s+1]) over 64-lane waves: wave w owns the segments whose first pair lies in
[64 w, 64 w + 64), found by the wave map (wave_map in csrc/build.hip) as the smallest s with
start[s] >= 64 w, and lane l of the wave leads segment map[w] + l. Lists are sized by capacity and only their first n
segments are written (the rest are empty: start[s] == start[n] for s >= n). These tests
restate that arithmetic and check that every live segment has exactly one leader and no
leader falls on an unwritten tail segment.
"""
import numpy as np

WAVE = 64


def wave_map(start, nseg, waves, n_live=None):
    """The wave map: smallest s in [0, n] with start[s] >= 64 w (n = live count if given)."""
    n = nseg if n_live is None else min(n_live, nseg)
    out = np.empty(waves + 1, np.int64)
    for w in range(waves + 1):
        out[w] = np.searchsorted(start[:n], w * WAVE, side="left")
    return out


def leaders(wmap, waves):
    """(wave, lane) -> segment led; a wave owns [map[w], map[w+1]) and lane l leads map[w] + l."""
    led = []
    for w in range(waves):
        c0, c1 = wmap[w], wmap[w + 1]
        for lane in range(WAVE):
            cs = c0 + lane
            if cs < c1:
                led.append(cs)
    return np.array(led, np.int64)


def csr_with_tail(rng, n_live, cap, max_len=40):
    lens = np.zeros(cap, np.int64)
    lens[:n_live] = rng.integers(1, max_len, n_live)   # C3 and face segments are never empty
    start = np.zeros(cap + 1, np.int64)
    start[1:] = np.cumsum(lens)
    return start


def test_live_segments_have_exactly_one_leader():
    rng = np.random.default_rng(11)
    for n_live, cap in [(1, 1), (5, 64), (100, 1000), (9783, 158976 // 8), (64, 64), (65, 200)]:
        start = csr_with_tail(rng, n_live, cap)
        total = int(start[n_live])
        waves = (total + 2 * WAVE) // WAVE + 1          # any count >= total/64 + 1 works
        wmap = wave_map(start, cap, waves, n_live)
        led = leaders(wmap, waves)
        assert np.array_equal(np.sort(led), np.arange(n_live)), (n_live, cap)
        # a wave never owns more segments than it has lanes
        assert np.all(np.diff(wmap) <= WAVE)


def test_capacity_tail_is_never_led():
    """The mapping searched over the whole capacity gave the last populated wave the empty
    tail segments as well; their cell entries are unwritten, and the near-field leader wrote
    g[cell] for them. Bounded by the live count, the tail has no leader."""
    rng = np.random.default_rng(5)
    n_live, cap = 300, 5000
    start = csr_with_tail(rng, n_live, cap)
    waves = int(start[n_live]) // WAVE + 1
    unbounded = leaders(wave_map(start, cap, waves), waves)
    assert unbounded.max() >= n_live                    # the defect the live bound removes
    bounded = leaders(wave_map(start, cap, waves, n_live), waves)
    assert bounded.max() == n_live - 1


def test_pairs_are_covered_once():
    """Every pair index in [0, start[n]) is loaded by exactly one wave: wave w loads the pairs
    of its segments, start[map[w]] .. start[map[w+1]]."""
    rng = np.random.default_rng(3)
    n_live, cap = 777, 2000
    start = csr_with_tail(rng, n_live, cap)
    waves = int(start[n_live]) // WAVE + 1
    wmap = wave_map(start, cap, waves, n_live)
    seen = np.zeros(int(start[n_live]), np.int64)
    for w in range(waves):
        seen[start[wmap[w]]:start[wmap[w + 1]]] += 1
    assert np.all(seen == 1)


def wave_map_scatter(start, n, waves):
    """The wave map as the device writes it, without a search: thread i stores segment i for the
    waves with start[i-1] < 64 w <= start[i], and n for the waves with 64 w > start[n]; each
    entry also carries the wave's first pair start[map[w]]."""
    out = np.full((waves + 1, 2), -7, np.int64)
    written = np.zeros(waves + 1, np.int64)
    total = int(start[n])
    for i in range(max(n, waves) + 1):
        if i <= n:
            lo = 0 if i == 0 else int(start[i - 1]) // WAVE + 1
            hi = min(int(start[i]) // WAVE, waves)
            for w in range(lo, hi + 1):
                out[w] = (i, start[i])
                written[w] += 1
        if i <= waves and i * WAVE > total:
            out[i] = (n, total)
            written[i] += 1
    assert np.all(written == 1), "every wave entry is written exactly once"
    return out


def test_scatter_wave_map_equals_search():
    rng = np.random.default_rng(17)
    cases = [(1, 1, 0), (5, 64, 0), (100, 1000, 0), (2000, 2500, 0), (64, 64, 0), (65, 200, 0)]
    for n_live, cap, _ in cases:
        for max_len in (3, 40, 200):
            lens = np.zeros(cap, np.int64)
            lens[:n_live] = rng.integers(0, max_len, n_live)   # empty segments included
            start = np.zeros(cap + 1, np.int64)
            start[1:] = np.cumsum(lens)
            total = int(start[n_live])
            for waves in (total // WAVE + 1, total // WAVE + 37):
                ref = wave_map(start, cap, waves, n_live)
                got = wave_map_scatter(start, n_live, waves)
                assert np.array_equal(got[:, 0], ref), (n_live, cap, max_len, waves)
                assert np.array_equal(got[:, 1], start[ref])

"""GPU cell sort against the oracle: keys bit-exact (particleCellIndex), the
permutation equal to the stable sort, cell offsets equal to the histogram."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def particles(n, L=(1.0, 1.0), seed=0, spread=0.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-spread * L[0], (1 + spread) * L[0], n).astype(np.float32)
    y = rng.uniform(-spread * L[1], (1 + spread) * L[1], n).astype(np.float32)
    q = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return x, y, q


CASES = [
    (1.0, 1.0, 64, 64, 100_000, 0.0),
    (1.0, 1.0, 1024, 1024, 300_000, 0.0),
    (1.0, 1.0, 100, 50, 77_777, 0.0),      # dx not a power of two: FP64 keys
    (3.7, 1.3, 33, 17, 50_001, 0.5),       # positions outside [0, L): wrapped keys
    (128.0, 128.0, 128, 128, 4096 * 3, 0.0),
    (1.0, 1.0, 1, 1, 5000, 0.0),
    (1.0, 1.0, 4096, 2048, 200_000, 0.0),  # 23-bit keys: three radix passes
]


@pytest.mark.parametrize("case", CASES)
def test_keys_perm_offsets(hip, orc, case):
    Lx, Ly, nx, ny, n, spread = case
    say("sort", case)
    dom = hip.make_domain(Lx, Ly, nx, ny)
    od = orc.domain(Lx, Ly, nx, ny)
    x, y, q = particles(n, (Lx, Ly), seed=nx + n, spread=spread)
    with hip.Context(dom, n) as c:
        c.upload(x, y, q)
        keys_now = c.keys()
        keys_sorted, perm, off = c.sort_perm()
    ref_keys = orc.cell_keys(x.astype(np.float64), y.astype(np.float64), od)
    assert np.array_equal(keys_now, ref_keys)
    ref_perm = orc.sort_perm(ref_keys)
    assert np.array_equal(perm, ref_perm)
    assert np.array_equal(keys_sorted, ref_keys[ref_perm])
    assert np.array_equal(off, orc.cell_offsets(ref_keys, nx * ny))


def test_sort_moves_all_streams_together(hip):
    say("sort streams")
    dom = hip.make_domain(1.0, 1.0, 32, 32)
    n = 50_000
    rng = np.random.default_rng(1)
    cols = {k: rng.uniform(0, 1, n).astype(np.float32) for k in ("x", "y", "vx", "vy", "q", "m")}
    with hip.Context(dom, n) as c:
        c.upload(cols["x"], cols["y"], cols["q"], cols["vx"], cols["vy"], cols["m"])
        _, perm, _ = c.sort_perm()
        c.upload(cols["x"], cols["y"], cols["q"], cols["vx"], cols["vy"], cols["m"])
        c.sort()
        out = c.download()
        k2 = c.keys()
    for k, v in cols.items():
        assert np.array_equal(out[k], v[perm]), k
    assert np.all(np.diff(k2) >= 0)


@pytest.mark.parametrize("uniform", ["q", "m", "qm"])
def test_single_species_streams_across_sorts_and_runs(hip, uniform):
    """Charges or masses that are one value for every particle are not
    permuted by the sort (a permutation leaves them unchanged): the streams
    stay consistent through sorts, graph runs with an odd number of sorts and
    a later upload of per-particle values."""
    say("uniform streams", uniform)
    dom = hip.make_domain(1.0, 1.0, 32, 32)
    n = 40_000
    rng = np.random.default_rng(3)
    cols = {k: rng.uniform(0.5, 1.5, n).astype(np.float32) for k in ("x", "y", "vx", "vy", "q", "m")}
    cols["x"] -= 0.5
    cols["y"] -= 0.5
    for k in uniform:
        cols[k] = np.full(n, 0.75, np.float32)
    with hip.Context(dom, n) as c:
        c.upload(cols["x"], cols["y"], cols["q"], cols["vx"], cols["vy"], cols["m"])
        _, perm, _ = c.sort_perm()
        c.upload(cols["x"], cols["y"], cols["q"], cols["vx"], cols["vy"], cols["m"])
        c.sort()
        out = c.download()
        for k, v in cols.items():
            assert np.array_equal(out[k], v[perm]), k
        # 3 sorts in a graph run leave the particle buffers swapped; with dt = 0
        # nothing moves, so every particle's (x, y, vx, vy, q, m) must survive
        c.run(3, 1, 2, 0.0, 0.0, 0, sort_interval=1, use_graph=True)
        after = c.download()
        for k in uniform:
            assert np.all(after[k] == 0.75), k
        names = ("x", "y", "vx", "vy", "q", "m")
        rows = lambda d: np.stack([d[k] for k in names], axis=1)[np.lexsort([d[k] for k in names[::-1]])]
        assert np.array_equal(rows(after), rows(cols))
        # per-particle values after uniform ones: sorted as usual again
        cols2 = {k: rng.uniform(0.5, 1.5, n).astype(np.float32) for k in cols}
        cols2["x"] -= 0.5
        cols2["y"] -= 0.5
        c.upload(cols2["x"], cols2["y"], cols2["q"], cols2["vx"], cols2["vy"], cols2["m"])
        _, perm2, _ = c.sort_perm()
        c.upload(cols2["x"], cols2["y"], cols2["q"], cols2["vx"], cols2["vy"], cols2["m"])
        c.sort()
        out2 = c.download()
    for k, v in cols2.items():
        assert np.array_equal(out2[k], v[perm2]), k


@pytest.mark.parametrize("kind", ["equal", "descending", "empty", "one"])
def test_degenerate_inputs(hip, orc, kind):
    say("degenerate", kind)
    dom = hip.make_domain(1.0, 1.0, 16, 16)
    od = orc.domain(1.0, 1.0, 16, 16)
    if kind == "equal":
        x = np.full(10_000, 0.51, np.float32)
        y = np.full(10_000, 0.27, np.float32)
    elif kind == "descending":
        t = np.linspace(0.999, 0.0, 9000).astype(np.float32)
        x, y = t, t
    elif kind == "empty":
        x = y = np.zeros(0, np.float32)
    else:
        x, y = np.float32([0.7]), np.float32([0.2])
    q = np.ones(len(x), np.float32)
    with hip.Context(dom, max(len(x), 1)) as c:
        c.upload(x, y, q)
        ks, perm, off = c.sort_perm()
    rk = orc.cell_keys(x.astype(np.float64), y.astype(np.float64), od)
    assert np.array_equal(perm, orc.sort_perm(rk))
    assert np.array_equal(off, orc.cell_offsets(rk, 256))


def test_full_size_sort_properties(hip):
    """1024^2 grid, 64 ppc (67,108,864 particles): sortedness, permutation
    validity and offsets by size-independent properties."""
    say("full-size sort")
    dom = hip.make_domain(1.0, 1.0, 1024, 1024)
    n = 1024 * 1024 * 64
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=1, seed=42, amp=0.01, mode=4, q=1.0 / 64)
        keys0 = c.keys()
        ks, perm, off = c.sort_perm()
    assert np.all(np.diff(ks) >= 0)
    counts = np.bincount(perm, minlength=n)
    assert counts.max() == 1 and counts.min() == 1  # a permutation
    assert np.array_equal(keys0[perm], ks)
    # stability: within equal keys, source indices increase
    same = ks[1:] == ks[:-1]
    assert np.all(perm[1:][same] > perm[:-1][same])
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(ks, minlength=1024 * 1024))]))

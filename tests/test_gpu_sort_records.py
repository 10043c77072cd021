"""The staged sort (sets of at least 2^22 particles), whose radix passes carry
the particle streams with the keys: after a drift of a few cells, after a far
move, of an order that is still sorted (copy mode), on the node-keyed order,
and inside captured step graphs; against the stable sort of the same
particles, and its cell offsets through a bitwise deposit comparison."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX = NY = 256
PPC = 64
N = NX * NY * PPC  # 4,194,304: the smallest staged sort


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def node_keys(x, y, nx, ny, Lx=1.0, Ly=1.0):
    """Keys of the node-keyed order: floor cell + [FP32 fraction >= 1/2], wrapped."""
    def axis(p, n, L):
        f = p.astype(np.float64) / (L / n)
        fl = np.floor(f)
        up = (f - fl).astype(np.float32) >= np.float32(0.5)
        return np.mod(fl.astype(np.int64) + up, n)
    return axis(y, ny, Ly) * nx + axis(x, nx, Lx)


def drifting_set(seed, cells):
    """Uniform positions, velocities moving up to `cells` cells per axis in dt = 1."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, N).astype(np.float32)
    y = rng.uniform(0, 1, N).astype(np.float32)
    vx = rng.uniform(-cells / NX, cells / NX, N).astype(np.float32)
    vy = rng.uniform(-cells / NY, cells / NY, N).astype(np.float32)
    q = rng.uniform(0.5, 1.5, N).astype(np.float32)
    return x, y, q, vx, vy


def keys_of(orc, p, node):
    if node:
        return node_keys(p["x"], p["y"], NX, NY)
    return orc.cell_keys(p["x"].astype(np.float64), p["y"].astype(np.float64), orc.domain(1.0, 1.0, NX, NY))


@pytest.mark.parametrize("node", [False, True], ids=["cell", "node"])
@pytest.mark.parametrize("cells", [0.0, 1.5, 3.0, 40.0])
def test_staged_sort_equals_stable_sort(hip, orc, node, cells):
    say("staged sort", "node" if node else "cell", cells)
    dom = hip.make_domain(1.0, 1.0, NX, NY)
    x, y, q, vx, vy = drifting_set(7 + int(2 * cells), cells)
    with hip.Context(dom, N) as c:
        c.upload(x, y, q, vx, vy)
        (c.sort_nodes if node else c.sort)()
        assert c.sort_path() == 0  # uniform positions: not in order
        c.gather_push(1.0, gather=False, push=True)
        pre = c.download()
        (c.sort_nodes if node else c.sort)()
        path = c.sort_path()
        post = c.download()
        scheme = hip.DepositionScheme.TSC if node else hip.DepositionScheme.CIC
        c.deposit(scheme, hip.DepositionBackend.GPUPrivatized, 0.0)
        rho_sorted = c.rho()
    assert path == (2 if cells == 0.0 else 0), path
    ref = np.argsort(keys_of(orc, pre, node), kind="stable")
    for k in ("x", "y", "vx", "vy", "q", "m"):
        assert np.array_equal(post[k], pre[k][ref]), k
    # The offsets the sort built: a context that sorts the same ordered set
    # (copy mode) deposits bit for bit alike.
    with hip.Context(dom, N) as c2:
        c2.upload(post["x"], post["y"], post["q"], post["vx"], post["vy"])
        (c2.sort_nodes if node else c2.sort)()
        assert c2.sort_path() == 2
        c2.deposit(scheme, hip.DepositionBackend.GPUPrivatized, 0.0)
        rho_copy = c2.rho()
    assert np.array_equal(rho_sorted, rho_copy)


def test_staged_sort_in_graphs_equals_eager(hip):
    """A TSC run that sorts every 5 steps equals, bit for bit, the same run
    replayed from captured graphs (the device picks each sort's mode)."""
    say("staged sort in graphs")
    dom = hip.make_domain(1.0, 1.0, NX, NY)
    out = []
    for graph in (False, True):
        with hip.Context(dom, N) as c:
            c.init_particles(N, kind=0, seed=5, vth=0.001, q=-1.0 / N)
            c.run(20, scheme=hip.DepositionScheme.TSC, dt=0.5, background=1.0, sort_interval=5, use_graph=graph)
            out.append((c.download(), c.rho(), c.sort_path()))
    (p0, r0, s0), (p1, r1, s1) = out
    assert s0 == 0 and s1 == 0, (s0, s1)
    for k in ("x", "y", "vx", "vy"):
        assert np.array_equal(p0[k], p1[k]), k
    assert np.array_equal(r0, r1)

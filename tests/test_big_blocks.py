"""The blocked refactor and chunked solves of the big-lane kernel, run on
the CPU by the library (esl_debug_big_replay) and compared with the oracle's
serial replay (lane_kernels.zig:49-143): the blocks gather their slots,
apply their rows and scatter what they wrote, so L, U, the pivots, dx and
the re-pivot flag must equal the serial column replay's bit for bit, for
any block size down to a few slots (many blocks, ops of one target split
across blocks) and on the peel cases (void column, wide multiplier, growth).
"""
import numpy as np
import pytest

from test_oracle_lu import random_system


@pytest.fixture(scope="module")
def lanes():
    from espice_amd import lanes as L
    L.load()
    return L


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def grid_laplacian(rng, k):
    """k x k five-point grid, random positive conductances, a grounded
    conductance per node (diagonally dominant, structurally symmetric)."""
    n = k * k
    a = {}
    for i in range(k):
        for j in range(k):
            u = i * k + j
            a[(u, u)] = a.get((u, u), 0.0) + 10 ** rng.uniform(-4, -2)
            for v in ([u + 1] if j + 1 < k else []) + ([u + k] if i + 1 < k else []):
                g = 10 ** rng.uniform(-4, -2)
                a[(u, u)] = a.get((u, u), 0.0) + g
                a[(v, v)] = a.get((v, v), 0.0) + g
                a[(u, v)] = -g
                a[(v, u)] = -g
    cols = [[] for _ in range(n)]
    for (r, c), v in a.items():
        cols[c].append((r, v))
    cp, ri, vals = [0], [], []
    for c in range(n):
        for r, v in sorted(cols[c]):
            ri.append(r)
            vals.append(v)
        cp.append(len(ri))
    return np.array(cp, np.uint32), np.array(ri, np.uint32), np.array(vals)


def compare(lanes, oracle, cp, ri, lead, q, a, rhs, block_slots):
    got = lanes.debug_big_replay(cp, ri, lead, q, a, rhs, block_slots=block_slots)
    ref = oracle.lu_replay(cp, ri, lead, q, a, rhs)
    assert ref is not None
    assert np.array_equal(got["fail"], ref["fail"])
    ok = ref["fail"] == 0
    for key in ("lx", "ux", "udiag", "dx"):
        assert np.array_equal(bits(got[key][ok]), bits(ref[key][ok])), key
    return got, ok


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("block_slots", [3, 17, 256, 4032])
def test_random_systems_blocked_equal_serial(lanes, oracle, seed, block_slots):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 90))
    cp, ri, lead = random_system(rng, n, 0.08)
    if oracle.lu_factor(cp, ri, lead, np.arange(n)) is None:
        pytest.skip("singular lead")
    q = rng.permutation(n).astype(np.uint32)
    if oracle.lu_factor(cp, ri, lead, q) is None:
        q = np.arange(n, dtype=np.uint32)
    lanes_n = 6
    a = lead * (1 + 0.02 * rng.standard_normal((lanes_n, len(lead))))
    a[0] = lead
    rhs = rng.standard_normal((lanes_n, n))
    got, ok = compare(lanes, oracle, cp, ri, lead, q, a, rhs, block_slots)
    assert ok[0]
    if block_slots <= 17:
        assert got["blocks"] > 1


@pytest.mark.parametrize("block_slots", [64, 4032])
def test_grid_blocked_equal_serial(lanes, oracle, block_slots):
    rng = np.random.default_rng(7)
    cp, ri, lead = grid_laplacian(rng, 12)
    n = len(cp) - 1
    q = np.arange(n, dtype=np.uint32)   # banded fill: long columns, many blocks
    a = lead * np.exp(0.3 * rng.standard_normal((4, len(lead))))
    rhs = rng.standard_normal((4, n))
    got, ok = compare(lanes, oracle, cp, ri, lead, q, a, rhs, block_slots)
    assert ok.all()
    if block_slots == 64:
        assert got["blocks"] > 20


def test_peel_cases_fail_as_serial(lanes, oracle):
    # void column with a NaN where its slot must be zero (sparse_lu.zig:1400-1416)
    cp, ri = [0, 2, 4, 5, 6], [0, 3, 0, 1, 2, 3]
    vals = np.array([1, 0.5, 0, 0, 1, 1.0])
    compare(lanes, oracle, cp, ri, vals, [0, 1, 2, 3], np.array([vals, [1, np.nan, 0, 0, 0, 1.0]]),
            np.ones((2, 4)), 3)
    # wide multiplier and growth monitor (lane_kernels.zig:110-117, sparse_lu.zig:1003-1012)
    cp, ri = [0, 2, 4], [0, 1, 0, 1]
    for lead, bad in (([1.0, 0.5, 0.0, 1.0], [1e-4, 1.0, 0.0, 1.0]), ([1.0, 0.5, 0.5, 1.0], [1e-13, 1.0, 0.5, 1.0])):
        got, ok = compare(lanes, oracle, cp, ri, np.array(lead), [0, 1], np.array([lead, bad]), np.ones((2, 2)), 3)
        assert list(got["fail"]) == [0, 1]

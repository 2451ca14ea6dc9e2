"""The wavefront kernels' refactor and solve row programs, chain rows
included, run on the CPU by the library (esl_debug_row_replay) and compared
with the oracle's serial column replay (lane_kernels.zig:49-143).

A chain row hands one thread a run of multiply-subtracts on one target (a
column's terms into one factor slot, or one row's terms of the forward
solve), kept in a register and taken in their order; L, U, the pivots, dx and
the re-pivot flag must equal the serial replay's bit for bit with chains of
any length down to two terms, and without chains. The replay also checks
every row for independent ops. On the 64-OTA array's Jacobian (shared VDD
and bias rows, whose columns and forward-solve rows collect one term per
OTA) the chains must shorten both programs.
"""
import numpy as np
import pytest

from test_big_blocks import bits, grid_laplacian
from test_oracle_lu import random_system


@pytest.fixture(scope="module")
def lanes():
    from espice_amd import lanes as L
    L.load()
    return L


def min_degree(cp, ri):
    """Minimum degree on the pattern of A + A^T, ties to the lowest index
    (the engine's column order)."""
    n = len(cp) - 1
    adj = [set() for _ in range(n)]
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            i = int(ri[p])
            if i != j:
                adj[i].add(j)
                adj[j].add(i)
    gone = np.zeros(n, bool)
    q = []
    for _ in range(n):
        v = min((u for u in range(n) if not gone[u]), key=lambda u: (len(adj[u]), u))
        q.append(v)
        gone[v] = True
        nb = adj[v]
        for u in nb:
            adj[u] = (adj[u] | nb) - {u, v}
            adj[u] = {w for w in adj[u] if not gone[w]}
        adj[v] = set()
    return np.array(q, np.uint32)


def compare(lanes, oracle, cp, ri, lead, q, a, rhs, min_chain):
    got = lanes.debug_row_replay(cp, ri, lead, q, a, rhs, min_chain=min_chain)
    ref = oracle.lu_replay(cp, ri, lead, q, a, rhs)
    assert ref is not None
    assert np.array_equal(got["fail"], ref["fail"])
    ok = ref["fail"] == 0
    for key in ("lx", "ux", "udiag", "dx"):
        assert np.array_equal(bits(got[key][ok]), bits(ref[key][ok])), key
    return got, ok


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("min_chain", [0, 2, 3, 8])
def test_random_systems_rows_equal_serial(lanes, oracle, seed, min_chain):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 90))
    cp, ri, lead = random_system(rng, n, 0.1)
    if oracle.lu_factor(cp, ri, lead, np.arange(n)) is None:
        pytest.skip("singular lead")
    q = rng.permutation(n).astype(np.uint32)
    if oracle.lu_factor(cp, ri, lead, q) is None:
        q = np.arange(n, dtype=np.uint32)
    a = lead * (1 + 0.02 * rng.standard_normal((6, len(lead))))
    a[0] = lead
    rhs = rng.standard_normal((6, n))
    # a zero right-hand side entry: the forward solve's zero-multiplier terms
    rhs[1, : n // 3] = 0.0
    got, ok = compare(lanes, oracle, cp, ri, lead, q, a, rhs, min_chain)
    assert ok[0]
    if min_chain == 0:
        assert got["refactor_chains"] == 0 and got["solve_chains"] == 0


@pytest.mark.parametrize("min_chain", [0, 2, 8])
def test_grid_rows_equal_serial(lanes, oracle, min_chain):
    rng = np.random.default_rng(7)
    cp, ri, lead = grid_laplacian(rng, 9)
    n = len(cp) - 1
    q = min_degree(cp, ri)
    a = lead * (1 + 0.05 * rng.standard_normal((4, len(lead))))
    rhs = rng.standard_normal((4, n))
    got, ok = compare(lanes, oracle, cp, ri, lead, q, a, rhs, min_chain)
    assert ok.all()


def test_peel_cases_with_chains(lanes, oracle):
    """A lane whose pivot shrinks past the growth test, one with a zero
    pivot, and one with a non-finite entry fail exactly where the serial
    replay fails, with the chains on."""
    rng = np.random.default_rng(11)
    cp, ri, lead = grid_laplacian(rng, 6)
    n = len(cp) - 1
    q = min_degree(cp, ri)
    a = np.repeat(lead[None, :], 4, 0)
    diag = [p for j in range(n) for p in range(cp[j], cp[j + 1]) if ri[p] == j]
    a[1, diag[5]] *= 1e-14
    a[2, diag[n - 1]] = 0.0
    a[3, diag[3]] = np.inf
    rhs = rng.standard_normal((4, n))
    got, ok = compare(lanes, oracle, cp, ri, lead, q, a, rhs, 2)
    assert ok[0] and not ok[1:].all()


def test_ota_array_chains_shorten_the_programs(lanes, oracle):
    from espice_amd import circuits

    fam = circuits.ota_array(64)
    c = oracle.Circuit(fam.deck.arrays())
    cp, ri, _, _ = c.pattern()
    lead, _ = c.assemble_cold(fam.nominal())
    q = min_degree(cp, ri)
    rng = np.random.default_rng(3)
    vals = fam.sample(2026, 4)
    a = np.stack([c.assemble_cold(v)[0] for v in vals])
    rhs = rng.standard_normal((4, c.n))
    plain, ok0 = compare(lanes, oracle, cp, ri, lead, q, a, rhs, 0)
    chained, ok1 = compare(lanes, oracle, cp, ri, lead, q, a, rhs, 8)
    assert ok0.all() and ok1.all()
    assert chained["refactor_chains"] > 0 and chained["solve_chains"] > 0
    print("\nOTA-64 refactor rows %d -> %d (+%d chain blocks), solve rows %d -> %d (+%d chain blocks)" % (
        plain["refactor_rows"], chained["refactor_rows"], chained["refactor_chain_blocks"],
        plain["solve_rows"], chained["solve_rows"], chained["solve_chain_blocks"]))
    # a chain block costs a fraction of a row; each program must come out shorter
    assert chained["refactor_rows"] + chained["refactor_chain_blocks"] < 0.7 * plain["refactor_rows"]
    assert chained["solve_rows"] + chained["solve_chain_blocks"] < 0.7 * plain["solve_rows"]

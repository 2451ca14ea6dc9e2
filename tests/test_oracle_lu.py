"""Pins the oracle's sparse LU (a restatement of sparse_lu.zig) with the
reference's own closed forms and invariants, and against dense elimination."""
import numpy as np
import pytest


def test_zero_diagonal_tridiagonal_closed_form(oracle):
    # direct.zig:655-680: A = [0 1 0; 1 2 4; 0 1 3], A [1 2 3] = [2 17 11].
    cp, ri = [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2]
    vals = [0, 1, 1, 2, 1, 4, 3]
    x = oracle.lu_solve(cp, ri, vals, [0, 1, 2], [2, 17, 11])
    assert np.allclose(x, [1, 2, 3], atol=1e-12, rtol=0)


def test_ring_closed_form_and_solve_neg(oracle):
    # direct.zig:682-711: ring, A [1 2 3 4] = [10 12 18 20]; solveNeg = -solve.
    cp = [0, 3, 6, 9, 12]
    ri = [0, 1, 3, 0, 1, 2, 1, 2, 3, 0, 2, 3]
    vals = [4, 1, 1, 1, 4, 1, 1, 4, 1, 1, 1, 4]
    x = oracle.lu_solve(cp, ri, vals, [0, 1, 2, 3], [10, 12, 18, 20])
    assert np.allclose(x, [1, 2, 3, 4], atol=1e-13, rtol=0)
    r = oracle.lu_replay(cp, ri, vals, [0, 1, 2, 3], np.array([vals], float), np.array([[10, 12, 18, 20.0]]))
    assert np.allclose(r["dx"][0], [-1, -2, -3, -4], atol=1e-13, rtol=0)
    assert r["fail"][0] == 0


def test_singular_matrix_fails(oracle):
    # direct.zig:713-724: columns 0 and 1 both [1 1 0 0]^T, rank 3.
    cp = [0, 3, 6, 9, 12]
    ri = [0, 1, 3, 0, 1, 2, 1, 2, 3, 0, 2, 3]
    vals = [1, 1, 0, 1, 1, 0, 0, 2, 3, 1, 5, 7]
    assert oracle.lu_factor(cp, ri, vals, [0, 1, 2, 3]) is None


def test_void_column_factor_and_nan_replay(oracle):
    # sparse_lu.zig:1400-1416: column 1 is structurally present but void.
    cp, ri = [0, 2, 4, 5, 6], [0, 3, 0, 1, 2, 3]
    vals = [1, 0.5, 0, 0, 1, 1]
    t = oracle.lu_factor(cp, ri, vals, [0, 1, 2, 3])
    assert t is not None and t["flags"][1] & 1
    bad = np.array([[1, np.nan, 0, 0, 0, 1.0]])
    r = oracle.lu_replay(cp, ri, vals, [0, 1, 2, 3], bad, np.zeros((1, 4)))
    assert r["fail"][0] == 1


def test_overflow_off_pivot_fails(oracle):
    # sparse_lu.zig:1418-1448
    cp, ri, q = [0, 2, 3, 6], [0, 1, 1, 0, 1, 2], [0, 1, 2]
    assert oracle.lu_factor(cp, ri, [1, 1, 1, 1.5e308, -1.5e308, 1], q) is None
    r = oracle.lu_replay(cp, ri, [1, 1, 1, 1, 1, 1.0], q, np.array([[1, 1, 1, 1.5e308, -1.5e308, 1]]),
                         np.zeros((1, 3)))
    assert r["fail"][0] == 1


def test_one_by_one(oracle):
    # sparse_lu.zig:1450-1470
    assert oracle.lu_solve([0, 1], [0], [4.0], [0], [2.0])[0] == 0.5
    assert oracle.lu_factor([0, 1], [0], [np.nan], [0]) is None
    r = oracle.lu_replay([0, 1], [0], [4.0], [0], np.array([[np.inf]]), np.zeros((1, 1)))
    assert r["fail"][0] == 1


def random_system(rng, n, density):
    """sparse_lu.zig:1234-1311 TestSystem: full structural diagonal, values in
    [-1, 1], diagonal pushed past its column's off-diagonal sum, except a
    column with another entry gets a zero diagonal 15% of the time."""
    mask = rng.random((n, n)) < density
    np.fill_diagonal(mask, True)
    cp, ri, vals = [0], [], []
    for j in range(n):
        rows = np.nonzero(mask[:, j])[0]
        v = 2 * rng.random(len(rows)) - 1
        d = int(np.nonzero(rows == j)[0][0])
        s = np.abs(np.delete(v, d)).sum()
        v[d] = 0.0 if (s > 0 and rng.random() < 0.15) else v[d] + np.copysign(s, v[d])
        ri.extend(rows.tolist())
        vals.extend(v.tolist())
        cp.append(len(ri))
    return np.array(cp, np.uint32), np.array(ri, np.uint32), np.array(vals)


def dense(cp, ri, vals, n):
    a = np.zeros((n, n))
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            a[ri[p], j] = vals[p]
    return a


def test_random_systems_against_dense_elimination(oracle):
    # sparse_lu.zig:1313-1398: every order 1..64 plus random ones.
    rng = np.random.default_rng(0x5BA75E)
    solved = 0
    for t in range(200):
        n = t + 1 if t < 64 else int(rng.integers(1, 65))
        cp, ri, vals = random_system(rng, n, rng.random() * (0.6 if n > 32 else 0.3))
        q = rng.permutation(n).astype(np.uint32)
        f = oracle.lu_factor(cp, ri, vals, q)
        a = dense(cp, ri, vals, n)
        if f is None:
            continue
        solved += 1
        b = 2 * rng.random(n) - 1
        x = oracle.lu_solve(cp, ri, vals, q, b)
        # LAPACK's scaled residual ||Ax - b|| / (n ||A|| ||x|| n eps); a wrong
        # factor scores ~1e13 and up (sparse_lu.zig:1313-1316).
        res = np.abs(a @ x - b).max() / (n * np.abs(a).max() * np.abs(x).max() * n * 2.2e-16)
        assert res <= 1e6, (t, n, res)
        xd = np.linalg.solve(a, b)
        assert np.allclose(x, xd, rtol=1e-6, atol=1e-8 * np.abs(xd).max())
        # replay on the values that chose the pivots is bitwise the factor
        r = oracle.lu_replay(cp, ri, vals, q, vals[None, :], -b[None, :])
        assert r["fail"][0] == 0 or np.abs(f["lx"]).max(initial=0) > 1e3
        assert np.array_equal(r["lx"][0].view(np.uint64), f["lx"].view(np.uint64))
        assert np.array_equal(r["ux"][0].view(np.uint64), f["ux"].view(np.uint64))
        assert np.array_equal(r["udiag"][0].view(np.uint64), f["udiag"].view(np.uint64))
        # the lane solveNeg of -b is the solve of b, bitwise
        assert np.array_equal(r["dx"][0].view(np.uint64), x.view(np.uint64))
    assert solved >= 100


def test_power_of_two_scaling_is_exact(oracle):
    # sparse_lu.zig:1468-1520: A * 2^-20, b * 2^7 => x * 2^27, bitwise.
    rng = np.random.default_rng(0x2B17)
    checked = 0
    for t in range(64):
        n = 1 + t % 48
        cp, ri, vals = random_system(rng, n, 0.5 if n > 32 else 0.25)
        q = rng.permutation(n).astype(np.uint32)
        b = 2 * rng.random(n) - 1
        f1 = oracle.lu_factor(cp, ri, vals, q)
        f2 = oracle.lu_factor(cp, ri, vals * 2.0 ** -20, q)
        assert (f1 is None) == (f2 is None)
        if f1 is None:
            continue
        assert np.array_equal(f1["pinv"], f2["pinv"])
        x1 = oracle.lu_solve(cp, ri, vals, q, b)
        x2 = oracle.lu_solve(cp, ri, vals * 2.0 ** -20, q, b * 2.0 ** 7)
        assert np.array_equal((x1 * 2.0 ** 27).view(np.uint64), x2.view(np.uint64))
        checked += 1
    assert checked >= 32


def test_wide_multiplier_forces_repivot(oracle):
    # lane_kernels.zig:110-117: an unscaled multiplier past 1/pivot_tol fails
    # the replay even though the pivot itself is fine.
    cp, ri = [0, 2, 4], [0, 1, 0, 1]
    lead = [1.0, 0.5, 0.0, 1.0]
    r = oracle.lu_replay(cp, ri, lead, [0, 1], np.array([[1e-4, 1.0, 0.0, 1.0]]), np.zeros((1, 2)))
    assert r["fail"][0] == 1
    r = oracle.lu_replay(cp, ri, lead, [0, 1], np.array([[1.0, 0.9, 0.0, 1.0]]), np.zeros((1, 2)))
    assert r["fail"][0] == 0


def test_growth_monitor(oracle):
    # sparse_lu.zig:1003-1012: a pivot below 1e-12 of its column max fails.
    cp, ri = [0, 2, 4], [0, 1, 0, 1]
    lead = [1.0, 0.5, 0.5, 1.0]
    a = np.array([[1e-13, 1.0, 0.5, 1.0]])
    r = oracle.lu_replay(cp, ri, lead, [0, 1], a, np.zeros((1, 2)))
    assert r["fail"][0] == 1

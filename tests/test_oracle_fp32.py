"""Pins the oracle's FP32-factor mode: FP32 replay of an FP64 tape, the FP32
solve, and Newton with FP64 iterative refinement.

The FP32 factors have no counterpart in the reference; what fixes them is
arithmetic: a matrix whose LU factors are small integers rounds nowhere in
either precision, so its FP32 factors equal the FP64 ones exactly; FP32
solves of well-conditioned systems carry FP32 error; and a refinement that
is never allowed to succeed must hand every lane to FP64, reproducing the
FP64 Newton bit for bit.
"""
import numpy as np

from espice_amd import circuits
from test_oracle_lu import dense, random_system


def integer_lu_system(rng, n):
    """A = L U with unit-lower L and upper U of small integers, diagonal of U
    dominant so threshold pivoting keeps the diagonal."""
    lo = np.tril(rng.integers(-2, 3, (n, n)), -1) * (rng.random((n, n)) < 0.3)
    L = lo + np.eye(n)
    up = np.triu(rng.integers(-3, 4, (n, n)), 1) * (rng.random((n, n)) < 0.3)
    U = up + np.diag(rng.integers(40, 60, n) * rng.choice([-1, 1], n))
    A = L @ U
    cp, ri, vals = [0], [], []
    for j in range(n):
        rows = np.nonzero(A[:, j])[0]
        ri.extend(rows.tolist())
        vals.extend(A[rows, j].tolist())
        cp.append(len(ri))
    return np.array(cp, np.uint32), np.array(ri, np.uint32), np.array(vals, float), L, U


def test_integer_factors_are_exact_in_fp32(oracle):
    rng = np.random.default_rng(32)
    for t in range(20):
        n = int(rng.integers(2, 40))
        cp, ri, vals, L, U = integer_lu_system(rng, n)
        q = np.arange(n, dtype=np.uint32)
        r64 = oracle.lu_replay(cp, ri, vals, q, vals[None, :], np.zeros((1, n)))
        r32 = oracle.lu_replay32(cp, ri, vals, q, vals[None, :], np.ones((1, n)))
        assert r64 is not None and r32 is not None
        assert r32["fail"][0] == r64["fail"][0]
        # the diagonal pivots reproduce U's diagonal and integer multipliers
        assert np.array_equal(r32["udiag"][0].astype(np.float64), r64["udiag"][0])
        assert np.array_equal(r32["lx"][0].astype(np.float64), r64["lx"][0])
        assert np.array_equal(r32["ux"][0].astype(np.float64), r64["ux"][0])
        assert np.allclose(np.sort(np.abs(r64["udiag"][0])), np.sort(np.abs(np.diag(U))))


def test_fp32_solve_has_fp32_accuracy(oracle):
    rng = np.random.default_rng(0xF32)
    for t in range(40):
        n = int(rng.integers(2, 48))
        cp, ri, vals = random_system(rng, n, 0.2)
        # keep the diagonal nonzero so the systems stay well conditioned
        a = dense(cp, ri, vals, n)
        if np.any(np.diag(a) == 0):
            continue
        q = np.arange(n, dtype=np.uint32)
        b = 2 * rng.random(n) - 1
        r = oracle.lu_replay32(cp, ri, vals, q, vals[None, :], b[None, :])
        if r is None or r["fail"][0]:
            continue
        x = np.linalg.solve(a, b)
        err = np.abs(r["x"][0] - x).max() / np.abs(x).max()
        assert err < 1e-4, (t, err)
        assert err > 0 or n < 3   # it really ran in FP32


def test_refined_newton_matches_fp64(oracle):
    """On the OTA (well conditioned, limiting quiet after a few steps) the
    refined FP32 Newton takes the FP64 path: same iteration counts and
    node values within 1e-9 relative (1e-7 is the contract)."""
    fam = circuits.ota_array(8)
    c = oracle.Circuit(fam.deck.arrays())
    q = np.arange(c.n, dtype=np.uint32)
    vals = fam.sample(77, 48)
    a = c.op(vals, q, threads=8)
    b = c.op(vals, q, threads=8, precision=1)
    assert (a["reasons"] == 0).all() and (b["reasons"] == 0).all()
    assert np.array_equal(a["iterations"], b["iterations"])
    err = np.abs(a["x"] - b["x"]) / np.maximum(np.abs(a["x"]), 1e-6)
    assert err.max() < 1e-9, err.max()


def test_refined_newton_on_a_chaotic_circuit_converges_to_the_same_points(oracle):
    """The inverter chain's Newton paths are long (limiting, re-pivots) and
    split on last-bit differences, so FP32 lanes may take other paths; where
    both converge they meet within the Newton tolerance (reltol 1e-3 on the
    step, vntol 1e-6 V)."""
    fam = circuits.inverter_chain()
    c = oracle.Circuit(fam.deck.arrays())
    q = np.arange(c.n, dtype=np.uint32)
    vals = fam.sample(77, 48)
    a = c.op(vals, q, threads=8)
    b = c.op(vals, q, threads=8, precision=1)
    ok = (a["reasons"] == 0) & (b["reasons"] == 0)
    assert ok.mean() > 0.85
    nn = fam.deck.arrays()["n_nodes"]
    xa, xb = a["x"][ok][:, :nn], b["x"][ok][:, :nn]
    assert (np.abs(xa - xb) <= 2e-3 * np.abs(xa) + 1e-5).all()


def test_refinement_never_succeeding_is_the_fp64_newton(oracle):
    """refine_tol = 0 and one step: every lane's first FP32 solve counts as a
    stall, so every lane redoes its first iteration in FP64 and finishes in
    FP64: bitwise the FP64 Newton."""
    fam = circuits.ota_array(8)
    c = oracle.Circuit(fam.deck.arrays())
    q = np.arange(c.n, dtype=np.uint32)
    vals = fam.sample(3, 32)
    a = c.op(vals, q, threads=8)
    b = c.op(vals, q, threads=8, precision=1, refine_tol=0.0, refine_max=1)
    assert b["fp64_lanes"].all()
    assert np.array_equal(a["iterations"], b["iterations"])
    assert np.array_equal(a["x"].view(np.uint64), b["x"].view(np.uint64))

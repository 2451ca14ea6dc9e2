"""The oracle's .op gmin rung (orc_ckt.c gmin_ladder, following the control
of dc/op.zig:160-222), pinned by properties its formula does not state:

- points the plain rung solves are untouched, bit for bit;
- a point the ladder solves is a solution of the circuit without the
  shunts: the residual assembled at the published x passes the .op gates,
  and a plain Newton started there converges at once to the same x;
- the solution is the physical one: past the first stages every stage of a
  long inverter chain sits at a rail (within 3% of VDD), alternating stage
  to stage (which rail depends on the input against the first stage's
  switching threshold).
"""
import numpy as np
import pytest

from espice_amd import circuits

TOL_PLAIN = (1e-3, 1e-12, 1e-6, 1e-9, 100, 0)
TOL_LADDER = (1e-3, 1e-12, 1e-6, 1e-9, 100, 1)


@pytest.fixture(scope="module")
def chain(oracle):
    fam = circuits.inverter_chain(128)
    c = oracle.Circuit(fam.deck.arrays())
    vals = fam.sample(5, 24)
    q = oracle.natural_order(c.n)
    plain = c.op(vals, q, threads=8, tol=oracle.Tol(*TOL_PLAIN))
    ladder = c.op(vals, q, threads=8, tol=oracle.Tol(*TOL_LADDER))
    return fam, c, vals, q, plain, ladder


def test_ladder_solves_what_plain_newton_does_not(chain):
    _, _, _, _, plain, ladder = chain
    assert (ladder["reasons"] == 0).sum() > (plain["reasons"] == 0).sum() + 10
    both = plain["reasons"] == 0
    assert np.array_equal(ladder["x"][both].view(np.uint64), plain["x"][both].view(np.uint64))
    assert np.array_equal(ladder["iterations"][both], plain["iterations"][both])


def test_ladder_points_solve_the_unshunted_circuit(oracle, chain):
    fam, c, vals, q, plain, ladder = chain
    new = (ladder["reasons"] == 0) & (plain["reasons"] != 0)
    assert new.any()
    _, _, dg, cr = c.pattern()
    for l in np.nonzero(new)[0]:
        x = ladder["x"][l]
        g, r = c.assemble_at(vals[l], x)
        s = np.abs(g[dg])
        s[s == 0] = 1.0
        gate = np.maximum(1e-9, 10 * s * (1e-3 * np.abs(x) + 1e-6))
        assert (np.abs(r) <= gate).all()
    again = c.op(vals[new], q, start_x=ladder["x"][new], threads=8, tol=oracle.Tol(*TOL_PLAIN))
    assert (again["reasons"] == 0).all()
    assert (again["iterations"] <= 3).all()
    assert np.allclose(again["x"], ladder["x"][new], rtol=1e-3, atol=1e-6)


def test_ladder_points_sit_at_the_rails(chain):
    fam, c, vals, q, plain, ladder = chain
    d = fam.deck
    ok = ladder["reasons"] == 0
    for l in np.nonzero(ok)[0]:
        x = ladder["x"][l]
        vdd = x[d.node_index("vdd")]
        first_high = x[d.node_index("s11")] > 0.5 * vdd
        for k in range(10, 128):
            v = x[d.node_index(f"s{k + 1}")]
            want_high = first_high == (k % 2 == 0)
            target = vdd if want_high else 0.0
            assert abs(v - target) <= 0.03 * vdd, (l, k, v, vdd)

"""The oracle's .op ladder (orc_ckt.c op_ladder, following the control of
dc/op.zig:160-288: the gmin rung, then source stepping and a final run),
pinned by properties its formula does not state:

- points the plain rung solves are untouched, bit for bit;
- a point the ladder solves is a solution of the circuit without the
  shunts and at full sources: the residual assembled at the published x
  passes the .op gates, and a plain Newton started there converges at once
  to the same x;
- the solution is the physical one: past the first stages every stage of a
  long inverter chain sits at a rail (within 3% of VDD), alternating stage
  to stage (which rail depends on the input against the first stage's
  switching threshold);
- the source rung alone on a linear circuit follows the lambda schedule of
  dc/op.zig:243-276 (0, then steps of 0.25 growing 1.5x: 0.375, 0.9375,
  capped at 1) and the final run, with the iteration counts a linear
  Newton run has (a cold run 3 with its INITFIX pass, a warm run 2), and
  ends at the plain solution;
- the source scale multiplies every independent source (V and I) in the
  device evaluation, so on a linear circuit the solution scales with it.
"""
import numpy as np
import pytest

from espice_amd import circuits
from espice_amd.deck import Deck

LADDER, ALWAYS, NO_GMIN = 1, 2, 4
TOL = (1e-3, 1e-12, 1e-6, 1e-9, 100)


@pytest.fixture(scope="module", params=[LADDER, LADDER | NO_GMIN], ids=["ladder", "source_rung_only"])
def chain(oracle, request):
    fam = circuits.inverter_chain(128)
    c = oracle.Circuit(fam.deck.arrays())
    vals = fam.sample(5, 24)
    q = oracle.natural_order(c.n)
    plain = c.op(vals, q, threads=8, tol=oracle.Tol(*TOL, 0))
    ladder = c.op(vals, q, threads=8, tol=oracle.Tol(*TOL, request.param))
    return fam, c, vals, q, plain, ladder


def test_ladder_solves_what_plain_newton_does_not(chain):
    _, _, _, _, plain, ladder = chain
    assert (ladder["reasons"] == 0).sum() > (plain["reasons"] == 0).sum() + 2
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
    again = c.op(vals[new], q, start_x=ladder["x"][new], threads=8, tol=oracle.Tol(*TOL, 0))
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


def linear_deck():
    """A divider fed by a V source, loaded by an I source, and a ladder of
    resistors behind it: linear, one V and one I source, both live."""
    d = Deck("linear_sources")
    d.vsource("vs", "a", 0, 2.0)
    d.bind("vs", 0, "vs")
    d.resistor("r1", "a", "b", 1e3)
    d.resistor("r2", "b", 0, 2e3)
    d.isource("ib", 0, "b", 1e-4)
    d.bind("ib", 0, "ib")
    prev = "b"
    for k in range(6):
        d.resistor(f"rl{k}", prev, f"l{k}", 500.0)
        d.resistor(f"rg{k}", f"l{k}", 0, 4e3)
        prev = f"l{k}"
    return d


def test_source_rung_schedule_on_a_linear_circuit(oracle):
    d = linear_deck()
    c = oracle.Circuit(d.arrays())
    q = oracle.natural_order(c.n)
    rng = np.random.default_rng(3)
    vals = np.stack([rng.uniform(0.5, 3.0, 8), rng.uniform(-3e-4, 3e-4, 8)], axis=1)
    plain = c.op(vals, q, threads=4, tol=oracle.Tol(*TOL, 0))
    assert (plain["reasons"] == 0).all() and (plain["iterations"] == 3).all()
    # lambda 0 (cold: 3), 0.375, 0.9375, 1 (warm: 2 each), then the final run (2)
    src = c.op(vals, q, threads=4, tol=oracle.Tol(*TOL, LADDER | ALWAYS | NO_GMIN))
    assert (src["reasons"] == 0).all()
    assert np.array_equal(src["iterations"], plain["iterations"] + 3 + 2 + 2 + 2 + 2)
    assert np.allclose(src["x"], plain["x"], rtol=1e-12, atol=1e-15)
    # the gmin rung on a linear circuit: from 1e-3 the factor stays 10 (every
    # run easy); nine divisions by 10 leave 1.0000000000000002e-12, above
    # the target, so the clamp to 1e-12 takes an eleventh run; the first run
    # cold (3), the rest warm (2), then the clean run (2); the source rung
    # is not reached
    gm = c.op(vals, q, threads=4, tol=oracle.Tol(*TOL, LADDER | ALWAYS))
    assert (gm["reasons"] == 0).all()
    assert np.array_equal(gm["iterations"], plain["iterations"] + 3 + 10 * 2 + 2)
    assert np.allclose(gm["x"], plain["x"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("scale", [0.0, 0.25, 0.375, 1.0])
def test_source_scale_scales_every_source(oracle, scale):
    # at x = 0 the residual of the linear deck holds only its sources: -V on
    # the source's branch row, the I source's current on its two nodes
    d = linear_deck()
    c = oracle.Circuit(d.arrays())
    v, i = 1.7, 2.5e-4
    x0 = np.zeros(c.n)
    g1, r1 = c.assemble_at([v, i], x0)
    gs, rs = c.assemble_at([v, i], x0, src=scale)
    assert np.array_equal(gs, g1)
    assert np.allclose(rs, scale * r1, rtol=1e-15, atol=0)
    b = d.node_index("b")
    assert r1[b] == pytest.approx(-i) and np.count_nonzero(r1) == 2   # the V branch row and node b
    assert sorted(np.abs(r1[np.nonzero(r1)])) == pytest.approx(sorted([i, v]))

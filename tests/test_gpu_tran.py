"""Diode operating points and transient lanes on the GPU against the oracle.

Bitwise: the device runs the oracle's operations in the oracle's order
(device physics, companion, LU on shared tapes), so node voltages, probe
rows, iteration counts and reasons agree bit for bit; the hipGraph replay
and plain launches, and any number of steps per launch, give the same bits.
"""
import sys

import numpy as np
import pytest

from espice_amd import circuits
from test_oracle_tran import rc_deck, trapezoid_reference

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def lanes():
    from espice_amd import lanes as L
    L.load()
    return L


@pytest.mark.parametrize("precision", [0, 1])
def test_diode_stack_operating_points_bitwise(lanes, oracle, precision):
    fam = circuits.diode_stack(8)
    vals = fam.sample(21, 256)
    say(f"diode stack .op precision={precision}")
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    got = eng.solve_op(vals, precision=precision, tiling=lanes.TILING_AUTO if precision == 0 else lanes.TILING_WAVE)
    ref = oracle.Circuit(fam.deck.arrays()).op(vals, q, threads=8, precision=precision)
    assert (got["reasons"] == 0).all()
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(bits(got["x"]), bits(ref["x"]))
    eng.close()


def test_rc_ladder_transient_matches_the_recurrence(lanes):
    R, C, dt, steps, stages = 1e3, 1e-12, 1e-10, 40, 6
    d = rc_deck(R, C, stages)
    eng = lanes.Engine(d.arrays())
    probes = [d.node_index(f"n{k}") for k in range(stages)]
    r = eng.solve_tran(np.zeros((4, 0)), dt, steps, probes=probes, steps_per_launch=7)
    assert (r["reasons"] == 0).all()
    ref = trapezoid_reference(R, C, stages, dt, steps)
    assert np.allclose(r["probes"][0], ref, rtol=1e-9, atol=1e-12)
    assert (r["iterations"] == 2 * steps).all()
    eng.close()


@pytest.mark.parametrize("k,lanes_n,steps", [(6, 64, 40), (12, 128, 30)])
def test_rc_mesh_transient_bitwise_against_oracle(lanes, oracle, k, lanes_n, steps):
    fam = circuits.rc_mesh(k, diode_frac=0.1)
    vals = fam.sample(2031, lanes_n)
    d = fam.deck
    probes = [d.node_index("n0_0"), d.node_index(f"n{k - 1}_{k - 1}"), d.node_index(f"n{k // 2}_{k // 2}")]
    say(f"rc mesh {k}x{k}, {lanes_n} lanes, {steps} steps")
    eng = lanes.Engine(d.arrays())
    _, _, q = eng.pattern()
    got = eng.solve_tran(vals, 1e-12, steps, probes=probes, steps_per_launch=10)
    tm = eng.timing()
    ref = oracle.tran(oracle.Circuit(d.arrays()), vals, q, 1e-12, steps, probes=probes, threads=8)
    say(f"  reasons ok {(got['reasons'] == 0).sum()}, iterations/lane {got['iterations'].mean():.1f}, "
        f"repivots {tm.repivots}, launches {tm.launches}")
    assert np.array_equal(got["reasons"], ref["reasons"])
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(bits(got["probes"]), bits(ref["probes"]))
    assert np.array_equal(bits(got["x"]), bits(ref["x"]))
    eng.close()


def test_graph_and_launch_blocking_do_not_change_bits(lanes):
    fam = circuits.rc_mesh(8, diode_frac=0.1)
    vals = fam.sample(77, 64)
    probes = [fam.deck.node_index("n7_7")]
    eng = lanes.Engine(fam.deck.arrays())
    a = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=10, use_graph=True)
    b = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=10, use_graph=False)
    c = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=1, use_graph=True)
    d = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=25, use_graph=True)
    for r in (b, c, d):
        assert np.array_equal(bits(a["probes"]), bits(r["probes"]))
        assert np.array_equal(a["iterations"], r["iterations"])
    eng.close()


def test_switching_transient_repivots_on_the_device_bitwise(lanes, oracle):
    """Inverter chains switching under a step input change pivot order
    during the transient: lanes whose replay fails factor their own
    Jacobian on the device and must still match the oracle (replay, then a
    full factor when a peel test fails) bit for bit, probes included."""
    fam = circuits.inverter_chain_tran(16)
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    vals = fam.sample(8, 48)
    dt, steps = 2e-12, 60
    probes = [eng.n - 1, 3]
    r = eng.solve_tran(vals, dt, steps, probes=probes, steps_per_launch=13)
    tm = eng.timing()
    ref = oracle.tran(oracle.Circuit(fam.deck.arrays()), vals, q, dt, steps, probes=probes, threads=8)
    say(f"  lanes on their own LU {tm.device_repivot_lanes}, host re-pivots {tm.repivots}")
    assert (ref["reasons"] == 0).mean() > 0.9
    assert np.array_equal(r["reasons"], ref["reasons"])
    assert np.array_equal(r["iterations"], ref["iterations"])
    ok = r["reasons"] == 0
    assert np.array_equal(bits(r["x"][ok]), bits(ref["x"][ok]))
    assert np.array_equal(bits(r["probes"][ok]), bits(ref["probes"][ok]))
    eng.close()


def test_progress_groups_and_cancellation(lanes):
    """The progress callback sees each group of steps; grouping the launches
    (one graph per group) leaves the bits unchanged; a true return cancels."""
    fam = circuits.rc_mesh(8, diode_frac=0.1)
    vals = fam.sample(77, 32)
    probes = [fam.deck.node_index("n7_7")]
    eng = lanes.Engine(fam.deck.arrays())
    ref = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=5)
    seen = []
    got = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=5,
                         progress=lambda done, total: seen.append((done, total)), progress_steps=10)
    assert seen == [(10, 25), (20, 25), (25, 25)]
    assert np.array_equal(bits(got["probes"]), bits(ref["probes"]))
    assert np.array_equal(got["iterations"], ref["iterations"])
    with pytest.raises(lanes.EslError) as ex:
        eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=5,
                       progress=lambda done, total: done >= 10, progress_steps=10)
    assert "cancelled" in str(ex.value)
    again = eng.solve_tran(vals, 1e-12, 25, probes=probes, steps_per_launch=5)
    assert np.array_equal(bits(again["probes"]), bits(ref["probes"]))
    eng.close()

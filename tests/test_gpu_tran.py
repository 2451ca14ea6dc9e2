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


def adaptive_cases():
    """(family, points, t_stop, t_step, dt_max, probes of the deck)"""
    mesh = circuits.rc_mesh(8, diode_frac=0.15)
    chain = circuits.inverter_chain_tran(12)
    ladder = circuits.rc_ladder(40, tran=True, clamp_every=8)
    pairs = circuits.diffpair_array_tran(4)
    return [
        ("rc_mesh_8", mesh, 48, 200e-12, 2e-12, 50e-12, ["n0_0", "n7_7", "n4_4"]),
        ("inverter_chain_tran_12", chain, 40, 300e-12, 5e-12, 75e-12, None),
        ("rc_ladder_tran_40", ladder, 40, 2e-9, 20e-12, 0.5e-9, None),
        ("diffpair_array_tran_4", pairs, 24, 1e-9, 20e-12, 0.25e-9, None),
    ]


@pytest.mark.parametrize("case", [c[0] for c in adaptive_cases()])
def test_adaptive_transient_bitwise_against_oracle(lanes, oracle, case):
    """Per-lane step control on the device (esl_tran_adaptive_solve): every
    lane's accepted times, probe rows, final state, accepted and rejected
    steps, Newton iterations and reasons are bitwise the oracle's
    (orc_tran_adaptive) on the same points, with several attempts per
    launch so lanes stop and resume between launches."""
    name, fam, count, t_stop, t_step, dt_max, pnames = [c for c in adaptive_cases() if c[0] == case][0]
    d = fam.deck
    eng = lanes.Engine(d.arrays())
    _, _, q = eng.pattern()
    probes = [d.node_index(p) for p in pnames] if pnames else [eng.n - 1, 1]
    vals = fam.sample(77, count)
    say(f"adaptive {name}: {count} points, t_stop {t_stop:g}")
    got = eng.solve_tran_adaptive(vals, t_stop, t_step, probes=probes, dt_max=dt_max, max_steps=3000,
                                  attempts_per_launch=16)
    ref = oracle.tran_adaptive(oracle.Circuit(d.arrays()), vals, q, t_stop, t_step,
                               dt_max=(dt_max if dt_max > 0 else None), max_steps=3000, probes=probes, threads=8)
    say(f"  accepted {got['accepted'].mean():.1f} rejected {got['rejected'].mean():.2f} its "
        f"{got['iterations'].mean():.1f} ok {(got['reasons'] == 0).sum()}; launches {eng.timing().launches}")
    assert (ref["reasons"] == 0).mean() > 0.9
    for k in ("reasons", "accepted", "rejected", "iterations"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(bits(got["x"]), bits(ref["x"]))
    for l in range(count):
        a = int(ref["accepted"][l])
        assert np.array_equal(bits(got["times"][l, :a]), bits(ref["times"][l, :a])), l
        assert np.array_equal(bits(got["probes"][l, :a]), bits(ref["probes"][l, :a])), l
    eng.close()


def test_adaptive_breakpoint_landing_bitwise_against_oracle(lanes, oracle):
    """A step source at td > 0: the lanes land on the edge, drop to order 1
    and restart with a tenth of the step, bitwise as the oracle."""
    from test_oracle_adaptive import one_rc
    d = one_rc(td=2.5e-9)
    eng = lanes.Engine(d.arrays())
    _, _, q = eng.pattern()
    probes = [d.node_index("n0")]
    got = eng.solve_tran_adaptive(np.zeros((3, 0)), 8e-9, 1e-10, probes=probes, dt_max=1e-9, attempts_per_launch=5)
    ref = oracle.tran_adaptive(oracle.Circuit(d.arrays()), np.zeros((3, 0)), q, 8e-9, 1e-10, dt_max=1e-9,
                               max_steps=1000, probes=probes)
    for k in ("reasons", "accepted", "rejected", "iterations"):
        assert np.array_equal(got[k], ref[k]), k
    a = int(ref["accepted"][0])
    assert np.array_equal(bits(got["times"][:, :a]), bits(ref["times"][:, :a]))
    assert np.array_equal(bits(got["probes"][:, :a]), bits(ref["probes"][:, :a]))
    assert np.min(np.abs(got["times"][0, :a] - 2.5e-9)) <= 1e-19
    eng.close()

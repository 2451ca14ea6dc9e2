"""Transient .meas cards measured on the device as the lanes accept their
steps (esl_tran_options.meas, the subset of src/output/measure.zig that
reads one vector: WHEN, TRIG/TARG, FIND at AT= and at a WHEN, AVG, MIN,
MAX, PP, INTEG, RMS), against the oracle's passes (orc_meas.c) fed the
oracle transient's rows (its operating point at t = 0, then every accepted
step): bitwise, with and without early stop; early-stopped points take
exactly the steps up to the row that resolves their last card.
"""
import sys

import numpy as np
import pytest

from espice_amd import circuits

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


def cards_for(lib, out, mid, dt, steps):
    """A switching chain's cards: the output's crossings, delays, window
    statistics; lib: the lanes module or the oracle (same constructors)."""
    t_end = dt * steps
    c = lib.meas_clause
    return [
        lib.meas_card("when", c(out, val=0.5, rise=1)),
        lib.meas_card("when", c(mid, val=0.5, fall=1)),
        lib.meas_card("trig_targ", c(mid, val=0.5, fall=1), c(out, val=0.5, rise=1)),
        lib.meas_card("find", c(mid, at=0.4 * t_end)),
        lib.meas_card("find", c(mid), c(out, val=0.3, rise=1)),
        lib.meas_card("avg", c(out, frm=0.1 * t_end, to=0.6 * t_end)),
        lib.meas_card("integ", c(mid, frm=0.05 * t_end, to=0.5 * t_end)),
        lib.meas_card("max", c(out, to=0.7 * t_end)),
        lib.meas_card("pp", c(mid, to=0.8 * t_end)),
        lib.meas_card("rms", c(out, frm=0.2 * t_end, to=0.9 * t_end)),
        lib.meas_card("min", c(mid, frm=0.3 * t_end, to=0.75 * t_end)),
    ]


def oracle_measures(oracle, fam, vals, q, dt, steps, nodes, cards, early):
    """The oracle's values and rows over its own transient: the operating
    point's row, then the probe row of every step."""
    c = oracle.Circuit(fam.deck.arrays())
    op = c.op(vals, q, threads=8)
    tr = oracle.tran(c, vals, q, dt, steps, probes=nodes, threads=8)
    t = np.concatenate([[0.0], (np.arange(steps) + 1.0) * dt])
    col = {v: k for k, v in enumerate(nodes)}
    remap = []
    for cd in cards:
        r = type(cd).from_buffer_copy(cd)
        r.a.node, r.b.node = col[cd.a.node], col[cd.b.node]
        remap.append(r)
    vals_out, rows = [], []
    for l in range(len(vals)):
        x = np.vstack([op["x"][l][nodes][None, :], tr["probes"][l]])
        v, n_rows = oracle.meas_rows(remap, t, x, early=early)
        vals_out.append(v)
        rows.append(n_rows)
    return np.array(vals_out), np.array(rows), tr


@pytest.mark.parametrize("early", [False, True])
def test_streamed_measures_bitwise_against_the_oracle(lanes, oracle, early):
    fam = circuits.inverter_chain_tran(16)
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    vals = fam.sample(8, 48)
    dt, steps = 2e-12, 200
    # odd stages start high: s7 falls some 240 ps after the input steps up,
    # and s8 rises right after it
    out, mid = fam.deck.node_index("s8"), fam.deck.node_index("s7")
    nodes = [out, mid, 0]   # ground: the unused clauses name node 0
    got = eng.solve_tran(vals, dt, steps, probes=nodes, steps_per_launch=13,
                         meas=cards_for(lanes, out, mid, dt, steps), meas_stop=early)
    want, rows, tr = oracle_measures(oracle, fam, vals, q, dt, steps, nodes,
                                     cards_for(oracle, out, mid, dt, steps), early)
    ok = got["reasons"] == 0
    say(f"  early={early}: {int(ok.sum())}/48 converge, steps {got['steps_done'].min()}..{got['steps_done'].max()} "
        f"of {steps}, NaN cards {int(np.isnan(got['meas']).sum())}")
    assert ok.all() and (tr["reasons"] == 0).all()
    assert np.array_equal(bits(got["meas"]), bits(want))
    if early:
        assert np.array_equal(got["steps_done"], rows - 1)
        assert (got["steps_done"] < steps).any()
        # the steps it ran are the full run's steps, bit for bit
        for l in range(len(vals)):
            s = int(got["steps_done"][l])
            assert np.array_equal(bits(got["probes"][l, :s]), bits(tr["probes"][l, :s]))
    else:
        assert (got["steps_done"] == steps).all()
        assert np.array_equal(bits(got["probes"]), bits(tr["probes"]))
        assert not np.isnan(got["meas"]).any()
    eng.close()


def test_measure_cards_are_checked(lanes):
    fam = circuits.rc_ladder(8, tran=True, clamp_every=16)
    eng = lanes.Engine(fam.deck.arrays())
    vals = fam.sample(1, 2)
    with pytest.raises(lanes.EslError):
        eng.solve_tran(vals, 1e-12, 5, meas=[lanes.meas_card("when", lanes.meas_clause(eng.n, val=0.5))])
    bad = lanes.meas_card("when", lanes.meas_clause(1, val=0.5))
    bad.func = 99
    with pytest.raises(lanes.EslError):
        eng.solve_tran(vals, 1e-12, 5, meas=[bad])
    eng.close()

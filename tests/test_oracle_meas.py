"""The oracle's .meas subset (orc_meas.c: WHEN, TRIG/TARG, FIND at AT= or a
WHEN, AVG, MIN, MAX, PP, INTEG, RMS over a transient's rows, fed one row at
a time as src/output/measure.zig's Stream feeds a run).

Pinned by:
- the values ngspice 45 prints for the reference's own measure tests
  (tests/golden/ngspice_meas.json: a ramp, a triangle wave with every event
  count, a wave with a NaN sample, a single point);
- closed forms on finely sampled smooth waves (an exponential's crossing
  time, its integral by the composite Simpson panels, a ramp's mean, a
  FIND at a WHEN on lines, where interpolation is exact);
- early stop: the values over the rows up to the one that resolves every
  card equal the values over the whole run, and that row is the event's
  (or the first past TO=).
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(__file__)


def clause(o, spec):
    spec = dict(spec or {})
    for k in ("rise", "fall", "cross"):
        if spec.get(k) == "last":
            spec[k] = o.MEAS_LAST
    if "from" in spec:
        spec["frm"] = spec.pop("from")
    return o.meas_clause(**spec)


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "ngspice_meas.json")) as f:
        return json.load(f)


def test_meas_matches_ngspice_on_the_reference_waves(oracle, golden):
    for case in golden["cases"]:
        w = golden["waves"][case["wave"]]
        card = oracle.meas_card(case["func"], clause(oracle, case.get("a")), clause(oracle, case.get("b")))
        got, rows = oracle.meas_rows([card], np.array(w["t"], float), np.array(w["v"], float))
        assert rows == len(w["t"])
        want = case["value"]
        tag = (case["wave"], case["name"], got[0], want)
        if want is None:
            assert math.isnan(got[0]), tag
        else:
            # ngspice prints six significant digits (five for WHEN, RMS)
            assert abs(got[0] - want) <= 5e-6 * max(abs(want), 1e-300) + 1e-15, tag


def test_meas_closed_forms_on_smooth_waves(oracle):
    # a step of 2^-10: every sample time and width is exact, so the integral
    # runs on Simpson panels throughout (widths equal to 100 ulps)
    tau, h = 1.0, 2.0 ** -10
    t = np.arange(0, 5 * 1024 + 1) * h
    v = 1.0 - np.exp(-t / tau)
    e = np.exp(-t)
    ramp = 2.0 * t + 1.0
    x = np.stack([v, e, ramp], axis=1)
    cards = [
        oracle.meas_card("when", oracle.meas_clause(0, val=0.5)),
        oracle.meas_card("integ", oracle.meas_clause(1, frm=0.0, to=3.0)),
        oracle.meas_card("avg", oracle.meas_clause(2, frm=1.0, to=4.0)),
        oracle.meas_card("trig_targ", oracle.meas_clause(0, val=0.1, rise=1), oracle.meas_clause(0, val=0.9, rise=1)),
        oracle.meas_card("rms", oracle.meas_clause(2, frm=0.0, to=2.0)),
        oracle.meas_card("find", oracle.meas_clause(2), oracle.meas_clause(0, val=0.5)),
        oracle.meas_card("max", oracle.meas_clause(1, frm=0.5, to=2.0)),
    ]
    got, _ = oracle.meas_rows(cards, t, x)
    t50 = tau * math.log(2.0)
    assert abs(got[0] - t50) < 1e-6                              # crossing by linear interpolation: O(h^2)
    assert abs(got[1] - (1.0 - math.exp(-3.0))) < 1e-12          # Simpson panels on uniform steps
    assert abs(got[2] - 6.0) < 1e-12                             # mean of 2t + 1 over [1, 4]
    assert abs(got[3] - tau * math.log(9.0)) < 1e-6              # 10 % to 90 % rise time
    rms = math.sqrt((4 * 8 / 3 + 2 * 4 + 2) / 2.0)               # sqrt(mean of (2t + 1)^2 over [0, 2])
    assert abs(got[4] - rms) < 1e-9
    assert abs(got[5] - (2.0 * got[0] + 1.0)) < 1e-12            # FIND a line at the WHEN time
    assert abs(got[6] - math.exp(-0.5)) < 1e-12


def test_meas_early_stop_resolves_at_the_event_and_keeps_the_values(oracle):
    t = np.linspace(0, 10, 1001)
    x = np.stack([np.sin(t), np.cos(0.7 * t)], axis=1)
    cards = [
        oracle.meas_card("when", oracle.meas_clause(0, val=0.5, rise=2)),
        oracle.meas_card("avg", oracle.meas_clause(1, frm=1.0, to=6.0)),
        oracle.meas_card("trig_targ", oracle.meas_clause(0, val=0.0, fall=1), oracle.meas_clause(1, val=0.0, rise=1)),
    ]
    full, rows_full = oracle.meas_rows(cards, t, x)
    early, rows = oracle.meas_rows(cards, t, x, early=True)
    assert rows_full == len(t)
    assert np.array_equal(full.view(np.uint64), early.view(np.uint64))
    # the last card to resolve: the second rise of sin through 0.5 (t near
    # 2 pi + 0.52), later than the window's end at 6 and cos's first rise
    assert rows < len(t)
    assert t[rows - 1] >= full[0] and t[rows - 2] < full[0]
    # no card resolves before its event: the rows up to it give NaN for it
    part, _ = oracle.meas_rows(cards, t[: rows - 1], x[: rows - 1], early=True)
    assert math.isnan(part[0])

"""Builds decks and expected columns from the ngspice fixtures transcribed
under tests/golden (MOS1 and Gummel-Poon operating points and sweeps)."""
from __future__ import annotations

import json
import os

import numpy as np

from espice_amd.deck import Deck

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLDEN = [os.path.join(HERE, "ngspice_mos1.json"), os.path.join(HERE, "ngspice_bjt.json"),
          os.path.join(HERE, "ngspice_diode.json")]


def load():
    cases = {}
    for path in GOLDEN:
        cases.update(json.load(open(path))["cases"])
    return cases


def build(case: dict):
    """Returns (deck, live values [rows][P], reader) for one golden case."""
    d = Deck()
    for dev in case["devices"]:
        kind, name, nodes = dev[0], dev[1], dev[2]
        if kind == "V":
            d.vsource(name, nodes[0], nodes[1], dev[3])
        elif kind == "I":
            d.isource(name, nodes[0], nodes[1], dev[3])
        elif kind == "R":
            d.resistor(name, nodes[0], nodes[1], dev[3])
        elif kind == "C":
            d.capacitor(name, nodes[0], nodes[1], dev[3])
        elif kind == "M":
            d.mos1(name, *nodes, dev[3], **dev[4])
        elif kind == "Q":
            d.bjt(name, *nodes, dev[3])
        elif kind == "D":
            d.diode(name, *nodes, dev[3])
    for src, _col, _sign in case["live"]:
        d.bind(src, 0, src)
    cols = case["columns"]
    if case["live"]:
        live = np.stack([sign * np.asarray(cols[col]["values"], dtype=float)
                         for _, col, sign in case["live"]], axis=1)
    else:
        live = np.zeros((1, 0))
    vsrc = [d.names[i].lower() for i in range(len(d.kind)) if d.kind[i] == 2]

    def reader(x: np.ndarray, column: str) -> np.ndarray:
        """Column `v(node)` or `i(vsource)` from solutions x [rows][n]."""
        kind, arg = column[0], column[2:-1].lower()
        if kind == "v":
            return x[:, d.node_index(arg)]
        return x[:, d.n_nodes + vsrc.index(arg)]

    return d, live, reader


def check(case: dict, x: np.ndarray, reader) -> tuple[int, float]:
    """Counts values outside the fixture's rtol/atol; returns (misses, worst
    error in units of the tolerance)."""
    misses, worst = 0, 0.0
    for col, spec in case["columns"].items():
        got = reader(x, col)
        exp = np.asarray(spec["values"], dtype=float)
        tol = spec["rtol"] * np.abs(exp) + spec["atol"]
        err = np.abs(got - exp)
        misses += int((~(err <= tol)).sum())
        worst = max(worst, float(np.max(err / tol)))
    return misses, worst

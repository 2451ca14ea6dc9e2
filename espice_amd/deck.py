"""Deck: a circuit template as flat device arrays, the layout of `esl_deck`.

The reference builds circuits from netlist text (src/frontend/) and binds
device parameters to `.param`s named in `live_params` (include/espice.h:
86-109); every point then supplies one row of live values
(`espice_append_points`, include/espice.h:325). Here a circuit is built
programmatically and a binding ties one device parameter to one live
column. Node 0 is ground.
"""
from __future__ import annotations

import numpy as np

NPARAM = 32
NTERM = 4

R, C, V, I, D, MOS1, BJT = range(7)
KIND_NAMES = {R: "R", C: "C", V: "V", I: "I", D: "D", MOS1: "MOS1", BJT: "BJT"}

# MOS1 parameter slots (include/espice_lanes.h ESL_M_*), models/mos1.va:17-80.
M_SLOTS = {
    "type": 0, "vto": 1, "kp": 2, "gamma": 3, "phi": 4, "lambda": 5, "w": 6, "l": 7,
    "is": 8, "ld": 9, "tox": 10, "u0": 11, "m": 12, "cbd": 13, "cbs": 14, "pb": 15,
    "mj": 16, "fc": 17, "cgso": 18, "cgdo": 19, "cgbo": 20, "gmin": 21, "temp": 22,
    "tnom": 23, "given": 24,
}
M_DEFAULTS = {
    "type": 1.0, "vto": 0.0, "kp": 2.0e-5, "gamma": 0.0, "phi": 0.6, "lambda": 0.0,
    "w": 1.0e-4, "l": 1.0e-4, "is": 1.0e-14, "ld": 0.0, "tox": 1.0e-7, "u0": 600.0,
    "m": 1.0, "cbd": 0.0, "cbs": 0.0, "pb": 0.8, "mj": 0.5, "fc": 0.5, "cgso": 0.0,
    "cgdo": 0.0, "cgbo": 0.0, "gmin": 1.0e-12, "temp": 27.0, "tnom": 27.0,
}
# $param_given bits (ESL_MG_*)
MG_TOX, MG_KP, MG_PHI, MG_GAMMA = 1, 2, 4, 8
# Diode level 1 slots (ESL_D_*), models/diode.va (RS = 0 subset).
D_SLOTS = {"is": 0, "n": 1, "cjo": 2, "vj": 3, "m": 4, "tt": 5, "fc": 6, "area": 7, "gmin": 8,
           "temp": 9, "tnom": 10, "eg": 11, "xti": 12, "bv": 13, "ibv": 14, "nbv": 15, "ikf": 16,
           "isr": 17, "nr": 18}
D_DEFAULTS = {"is": 1e-14, "n": 1.0, "cjo": 0.0, "vj": 1.0, "m": 0.5, "tt": 0.0, "fc": 0.5, "area": 1.0,
              "gmin": 1e-12, "temp": 27.0, "tnom": 27.0, "eg": 1.11, "xti": 3.0, "bv": 0.0, "ibv": 1e-3,
              "nbv": 0.0, "ikf": 0.0, "isr": 0.0, "nr": 2.0}
# Gummel-Poon slots (ESL_Q_*), models/gummel_poon.va (RB = RC = RE = 0).
Q_SLOTS = {"type": 0, "is": 1, "bf": 2, "nf": 3, "vaf": 4, "ikf": 5, "ise": 6, "ne": 7, "br": 8, "nr": 9,
           "var": 10, "ikr": 11, "isc": 12, "nc": 13, "nkf": 14, "iss": 15, "ns": 16, "area": 17, "m": 18,
           "gmin": 19, "temp": 20, "tnom": 21, "eg": 22, "xti": 23, "xtb": 24, "areab": 25, "areac": 26,
           "subs": 27}
Q_DEFAULTS = {"type": 1.0, "is": 1e-16, "bf": 100.0, "nf": 1.0, "vaf": 0.0, "ikf": 0.0, "ise": 0.0, "ne": 1.5,
              "br": 1.0, "nr": 1.0, "var": 0.0, "ikr": 0.0, "isc": 0.0, "nc": 2.0, "nkf": 0.5, "iss": 0.0,
              "ns": 1.0, "area": 1.0, "m": 1.0, "gmin": 1e-12, "temp": 27.0, "tnom": 27.0, "eg": 1.11,
              "xti": 3.0, "xtb": 0.0, "areab": 1.0, "areac": 1.0, "subs": 1.0}
# V source waveform slots (ESL_V_*): value (t <= td), wave (1 = step), v2, td.
V_WAVE, V_V2, V_TD = 1, 2, 3
# Parameters a binding may move per point on the GPU lanes.
BINDABLE = {R: {0}, C: {0}, V: {0}, I: {0}, D: {D_SLOTS["is"]}, MOS1: {M_SLOTS["vto"], M_SLOTS["kp"]},
            BJT: {Q_SLOTS["is"], Q_SLOTS["bf"]}}


class Deck:
    """Builder for a circuit template with live parameters."""

    def __init__(self, title: str = ""):
        self.title = title
        self._nodes: dict[str, int] = {"0": 0, "gnd": 0}
        self.kind: list[int] = []
        self.terms: list[list[int]] = []
        self.params: list[np.ndarray] = []
        self.names: list[str] = []
        self._by_name: dict[str, int] = {}
        self.live: list[str] = []
        self.binds: list[tuple[int, int, int]] = []

    # -- nodes --------------------------------------------------------------
    def node(self, name) -> int:
        key = str(name).lower()
        if key not in self._nodes:
            self._nodes[key] = len(set(self._nodes.values()))
        return self._nodes[key]

    @property
    def n_nodes(self) -> int:
        return len(set(self._nodes.values()))

    def node_index(self, name) -> int:
        return self._nodes[str(name).lower()]

    # -- devices ------------------------------------------------------------
    def _add(self, name: str, kind: int, nodes, p: np.ndarray) -> int:
        if name.lower() in self._by_name:
            raise ValueError(f"duplicate device {name}")
        t = [self.node(x) for x in nodes] + [0] * (NTERM - len(nodes))
        self.kind.append(kind)
        self.terms.append(t)
        self.params.append(p)
        self.names.append(name)
        idx = len(self.kind) - 1
        self._by_name[name.lower()] = idx
        return idx

    def resistor(self, name, p, n, r):
        a = np.zeros(NPARAM)
        a[0] = r
        return self._add(name, R, (p, n), a)

    def capacitor(self, name, p, n, c):
        a = np.zeros(NPARAM)
        a[0] = c
        return self._add(name, C, (p, n), a)

    def vsource(self, name, p, n, dc, step_to=None, step_at=0.0):
        """DC source; with step_to, the value becomes step_to for t > step_at."""
        a = np.zeros(NPARAM)
        a[0] = dc
        if step_to is not None:
            a[V_WAVE], a[V_V2], a[V_TD] = 1.0, step_to, step_at
        return self._add(name, V, (p, n), a)

    def bjt(self, name, c, b, e, s, model: dict | None = None, **inst):
        """Gummel-Poon BJT (c, b, e, substrate); model in SPICE names, type
        +1 NPN / -1 PNP."""
        vals = dict(Q_DEFAULTS)
        for k, v in {**(model or {}), **inst}.items():
            k = k.lower()
            if k not in Q_SLOTS:
                raise ValueError(f"unsupported BJT parameter {k}")
            vals[k] = float(v)
        a = np.zeros(NPARAM)
        for k, slot in Q_SLOTS.items():
            a[slot] = vals[k]
        return self._add(name, BJT, (c, b, e, s), a)

    def diode(self, name, a_node, k_node, model: dict | None = None, **inst):
        """Junction diode (models/diode.va). A series RS > 0 becomes a
        resistor RS / AREA from the anode to an internal node `name#a`
        that the junction hangs from (diode.va's rsb branch; RS(T) = RS
        with TRS = 0)."""
        vals = dict(D_DEFAULTS)
        rs = 0.0
        for k, v in {**(model or {}), **inst}.items():
            k = k.lower()
            if k == "rs":
                rs = float(v)
                continue
            if k not in D_SLOTS:
                raise ValueError(f"unsupported diode parameter {k}")
            vals[k] = float(v)
        if rs > 0:
            inner = f"{name}#a"
            self.resistor(f"{name}#rs", a_node, inner, max(rs / vals["area"], 1e-9))
            a_node = inner
        a = np.zeros(NPARAM)
        for k, slot in D_SLOTS.items():
            a[slot] = vals[k]
        return self._add(name, D, (a_node, k_node), a)

    def isource(self, name, p, n, dc):
        a = np.zeros(NPARAM)
        a[0] = dc
        return self._add(name, I, (p, n), a)

    def mos1(self, name, d, g, s, b, model: dict, **inst):
        """MOS1 with a model dict (SPICE names, lower case) and instance
        overrides (w, l, m). `type` is +1 NMOS, -1 PMOS."""
        vals = dict(M_DEFAULTS)
        given = 0
        for k, v in {**model, **inst}.items():
            k = k.lower()
            if k == "uo":
                k = "u0"
            if k not in M_SLOTS or k == "given":
                raise ValueError(f"unsupported MOS1 parameter {k}")
            vals[k] = float(v)
            given |= {"tox": MG_TOX, "kp": MG_KP, "phi": MG_PHI, "gamma": MG_GAMMA}.get(k, 0)
        a = np.zeros(NPARAM)
        for k, slot in M_SLOTS.items():
            if k != "given":
                a[slot] = vals[k]
        a[M_SLOTS["given"]] = float(given)
        return self._add(name, MOS1, (d, g, s, b), a)

    # -- live parameters ----------------------------------------------------
    def live_param(self, name: str) -> int:
        if name in self.live:
            return self.live.index(name)
        self.live.append(name)
        return len(self.live) - 1

    def bind(self, device: str, param, live: str):
        """params[device][param] = live column `live` for every point."""
        dev = self._by_name[device.lower()]
        slot = param if isinstance(param, int) else self._slot(dev, param)
        self.binds.append((dev, slot, self.live_param(live)))

    def _slot(self, dev, name):
        if self.kind[dev] == MOS1:
            return M_SLOTS[name]
        if self.kind[dev] == D:
            return D_SLOTS[name]
        if self.kind[dev] == BJT:
            return Q_SLOTS[name]
        return 0

    def param(self, device: str, name) -> float:
        dev = self._by_name[device.lower()]
        slot = name if isinstance(name, int) else self._slot(dev, name)
        return float(self.params[dev][slot])

    # -- flat arrays --------------------------------------------------------
    def arrays(self) -> dict:
        nd = len(self.kind)
        b = np.array(self.binds, dtype=np.uint32).reshape(-1, 3)
        return {
            "n_nodes": self.n_nodes,
            "kind": np.array(self.kind, dtype=np.uint32),
            "nodes": np.array(self.terms, dtype=np.uint32).reshape(nd, NTERM),
            "params": np.array(self.params, dtype=np.float64).reshape(nd, NPARAM),
            "n_live": len(self.live),
            "bind_dev": np.ascontiguousarray(b[:, 0]),
            "bind_param": np.ascontiguousarray(b[:, 1]),
            "bind_live": np.ascontiguousarray(b[:, 2]),
        }

    def nominal(self) -> np.ndarray:
        """The live row that reproduces the template's own parameter values."""
        row = np.zeros(len(self.live))
        for dev, slot, li in self.binds:
            row[li] = self.params[dev][slot]
        return row

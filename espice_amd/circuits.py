"""Circuit families of the lane workloads, with seeded per-point samplers.

Draws are keyed by (seed, lane, variable) through a counter-based hash, so a
point's values depend only on its own key and never on how many points run
or in what order (the keying idea of src/analysis/sweep/mc.zig:40-44; the
reference's own stream is Zig's Wyhash + Xoshiro256 and is not reproduced).

Families:
  inverter_chain  16-stage MOS1 CMOS inverter chain, Monte Carlo .op
  ota_array       64 MOS1 5-transistor OTAs in unity-gain feedback, .op,
                  32 live parameters per point
  diode_stack     a resistor into a stack of junction diodes, .op
  rc_mesh         k x k RC mesh with diode clamps and a step source, the
                  transient workload (every R, C and IS live)
  diffpair_array  64 Gummel-Poon differential pairs, .op (the FP32 +
                  refinement workload)

The uneven batch draws its instances from eight families built from four
structures, each as an .op and as a 200-step transient variant:
  inverter chains of 8-256 stages, RC ladders of 16-2048 nodes (with diode
  clamps), BJT differential-pair arrays of 4-64 pairs, and MOS1 current-
  mirror trees of depth 2-8 (`mixed_batch`).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .deck import Deck

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix(z: np.ndarray) -> np.ndarray:
    """splitmix64 finalizer (Steele, Lea, Flood, "Fast splittable pseudorandom
    number generators", OOPSLA 2014)."""
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, lanes: np.ndarray, var: int, sub: int = 0) -> np.ndarray:
    """U(0,1), open at both ends, keyed by (seed, lane, var, sub)."""
    lanes = np.asarray(lanes, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = _mix(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(var * 4 + sub + 1))
        z = _mix(k ^ (lanes * np.uint64(0xD1B54A32D192ED03) + np.uint64(0x632BE59BD9B4E019)))
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, lanes: np.ndarray, var: int) -> np.ndarray:
    """N(0,1) by Box-Muller on two keyed uniforms."""
    u1 = uniform01(seed, lanes, var, 0)
    u2 = uniform01(seed, lanes, var, 1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Family:
    name: str
    deck: Deck
    # (live name, kind, a, b): kind "normal" (mean a, sigma b), "rel_normal"
    # (mean a, relative sigma b), "uniform_of" (U[0,1] times live column a)
    draws: list = field(default_factory=list)

    def sample(self, seed: int, lanes: int, first: int = 0) -> np.ndarray:
        return self.sample_at(seed, np.arange(first, first + lanes, dtype=np.uint64))

    def sample_at(self, seed: int, idx: np.ndarray) -> np.ndarray:
        """Rows for the points keyed by `idx` (any lane numbers)."""
        idx = np.asarray(idx, dtype=np.uint64)
        out = np.zeros((len(idx), len(self.deck.live)))
        for var, (name, kind, a, b) in enumerate(self.draws):
            col = self.deck.live.index(name)
            if kind == "normal":
                out[:, col] = a + b * normal(seed, idx, var)
            elif kind == "rel_normal":
                out[:, col] = a * (1.0 + b * normal(seed, idx, var))
            elif kind == "uniform_of":
                out[:, col] = uniform01(seed, idx, var) * out[:, self.deck.live.index(a)]
            elif kind == "log_normal":
                out[:, col] = a * np.exp(b * normal(seed, idx, var))
            elif kind == "log_uniform":
                la, lb = np.log(a), np.log(b)
                out[:, col] = np.exp(la + (lb - la) * uniform01(seed, idx, var))
            else:
                raise ValueError(kind)
        return out

    def nominal(self) -> np.ndarray:
        return self.deck.nominal()


def inverter_chain(stages: int = 16) -> Family:
    """16-stage CMOS inverter chain (MOS1). NMOS W/L = 1u/0.1u, KP = 200u,
    VTO ~ N(0.5, 10 mV); PMOS W/L = 2u/0.1u, KP = 100u, VTO ~ N(-0.5, 10 mV);
    VDD ~ N(1.0 V, 2%); input DC level ~ U[0, VDD]. LAMBDA = 0.05 on both
    (the workload statement leaves it open; zero would make the switching
    point an infinite-gain singular pair)."""
    d = Deck(f"inverter_chain_{stages}")
    nm = {"type": 1, "vto": 0.5, "kp": 200e-6, "lambda": 0.05}
    pm = {"type": -1, "vto": -0.5, "kp": 100e-6, "lambda": 0.05}
    d.vsource("vdd", "vdd", 0, 1.0)
    d.vsource("vin", "in", 0, 0.5)
    prev = "in"
    for k in range(stages):
        out = f"s{k + 1}"
        d.mos1(f"mn{k}", out, prev, 0, 0, nm, w=1e-6, l=0.1e-6)
        d.mos1(f"mp{k}", out, prev, "vdd", "vdd", pm, w=2e-6, l=0.1e-6)
        prev = out
    d.bind("vdd", 0, "vdd")
    d.bind("vin", 0, "vin")
    for k in range(stages):
        d.bind(f"mn{k}", "vto", f"vto_n{k}")
        d.bind(f"mp{k}", "vto", f"vto_p{k}")
    draws = [("vdd", "rel_normal", 1.0, 0.02), ("vin", "uniform_of", "vdd", None)]
    draws += [(f"vto_n{k}", "normal", 0.5, 0.01) for k in range(stages)]
    draws += [(f"vto_p{k}", "normal", -0.5, 0.01) for k in range(stages)]
    return Family(d.title, d, draws)


OTA_ROLES = ("m1", "m2", "m3", "m4", "m5")


def ota_array(count: int = 64, groups: int = 3) -> Family:
    """`count` MOS1 5-transistor OTAs in unity-gain feedback sharing VDD and a
    bias mirror (about 320 unknowns at 64). Per OTA k: M1 (diff pair, gate =
    in+_k), M2 (gate = out_k), PMOS mirror M3/M4 and tail M5 mirrored from a
    shared diode-connected reference fed by IBIAS. in+_k is a DC source at
    0.8 + 0.4 k/(count-1) V. 32 live parameters per point: VDD ~ N(1.8 V, 2%),
    IBIAS ~ N(20 uA, 5%), and for each of `groups` device groups (OTA k is in
    group k mod groups) and each of the 5 roles, VTO (sigma 10 mV) and KP
    (sigma 3%)."""
    d = Deck(f"ota_array_{count}")
    kpn, kpp = 200e-6, 100e-6
    lam = 0.05
    nm = {"type": 1, "vto": 0.5, "kp": kpn, "lambda": lam}
    pm = {"type": -1, "vto": -0.5, "kp": kpp, "lambda": lam}
    d.vsource("vdd", "vdd", 0, 1.8)
    d.isource("ibias", "vdd", "vbias", 20e-6)
    d.mos1("mref", "vbias", "vbias", 0, 0, nm, w=2e-6, l=0.2e-6)
    sizes = {"m1": (nm, 4e-6), "m2": (nm, 4e-6), "m3": (pm, 4e-6), "m4": (pm, 4e-6), "m5": (nm, 2e-6)}
    for k in range(count):
        inp, out, tail, mir = f"inp{k}", f"out{k}", f"tail{k}", f"mir{k}"
        d.vsource(f"vin{k}", inp, 0, 0.8 + 0.4 * k / max(count - 1, 1))
        d.mos1(f"m1_{k}", mir, inp, tail, 0, sizes["m1"][0], w=sizes["m1"][1], l=0.2e-6)
        d.mos1(f"m2_{k}", out, out, tail, 0, sizes["m2"][0], w=sizes["m2"][1], l=0.2e-6)
        d.mos1(f"m3_{k}", mir, mir, "vdd", "vdd", sizes["m3"][0], w=sizes["m3"][1], l=0.2e-6)
        d.mos1(f"m4_{k}", out, mir, "vdd", "vdd", sizes["m4"][0], w=sizes["m4"][1], l=0.2e-6)
        d.mos1(f"m5_{k}", tail, "vbias", 0, 0, sizes["m5"][0], w=sizes["m5"][1], l=0.2e-6)
    d.bind("vdd", 0, "vdd")
    d.bind("ibias", 0, "ibias")
    draws = [("vdd", "rel_normal", 1.8, 0.02), ("ibias", "rel_normal", 20e-6, 0.05)]
    for g in range(groups):
        for r in OTA_ROLES:
            m = sizes[r][0]
            draws.append((f"vto_{r}_g{g}", "normal", m["vto"], 0.01))
            draws.append((f"kp_{r}_g{g}", "rel_normal", m["kp"], 0.03))
    for name, *_ in draws:
        d.live_param(name)
    for k in range(count):
        g = k % groups
        for r in OTA_ROLES:
            d.bind(f"{r}_{k}", "vto", f"vto_{r}_g{g}")
            d.bind(f"{r}_{k}", "kp", f"kp_{r}_g{g}")
    return Family(d.title, d, draws)


def diode_stack(stages: int = 8) -> Family:
    """A DC supply through a resistor into a stack of `stages` junction
    diodes (the reference's convergence/diode_stack_* decks): VDD ~ N(5 V,
    2%), R ~ N(1 kOhm, 1%), each diode's IS ~ log-uniform [1e-16, 1e-14] A."""
    d = Deck(f"diode_stack_{stages}")
    d.vsource("vdd", "vdd", 0, 5.0)
    d.resistor("r1", "vdd", "a0", 1e3)
    for k in range(stages):
        nxt = 0 if k == stages - 1 else f"a{k + 1}"
        d.diode(f"d{k}", f"a{k}", nxt, {"is": 1e-15})
    d.bind("vdd", 0, "vdd")
    d.bind("r1", 0, "r1")
    draws = [("vdd", "rel_normal", 5.0, 0.02), ("r1", "rel_normal", 1e3, 0.01)]
    for k in range(stages):
        d.bind(f"d{k}", "is", f"is{k}")
        draws.append((f"is{k}", "log_uniform", 1e-16, 1e-14))
    return Family(d.title, d, draws)


def diffpair_array(count: int = 64) -> Family:
    """`count` Gummel-Poon NPN differential pairs (128 BJTs at 64) with
    resistive loads and tail current sources between VCC = 5 V and VEE = -5 V;
    pair k is driven differentially by +-vd_k/2 with vd_k spread over
    [-0.1, 0.1] V (the reference's bjt_diffpair decks). Per point: each BJT's
    IS ~ log-normal (median 1e-16 A, sigma_ln 0.2) and BF ~ N(100, 10), each
    load R ~ N(10 kOhm, 1%). Tail 0.2 mA: the loads drop about 1 V."""
    d = Deck(f"diffpair_array_{count}")
    d.vsource("vcc", "vcc", 0, 5.0)
    d.vsource("vee", "vee", 0, -5.0)
    qm = {"type": 1, "is": 1e-16, "bf": 100.0, "vaf": 100.0}
    draws = []
    for k in range(count):
        vd = -0.1 + 0.2 * k / max(count - 1, 1)
        d.vsource(f"vp{k}", f"bp{k}", 0, 0.5 * vd)
        d.vsource(f"vn{k}", f"bn{k}", 0, -0.5 * vd)
        d.resistor(f"rp{k}", "vcc", f"cp{k}", 10e3)
        d.resistor(f"rn{k}", "vcc", f"cn{k}", 10e3)
        d.isource(f"it{k}", f"e{k}", "vee", 0.2e-3)
        d.bjt(f"q1_{k}", f"cp{k}", f"bp{k}", f"e{k}", 0, qm)
        d.bjt(f"q2_{k}", f"cn{k}", f"bn{k}", f"e{k}", 0, qm)
        for q in (f"q1_{k}", f"q2_{k}"):
            d.bind(q, "is", f"is_{q}")
            d.bind(q, "bf", f"bf_{q}")
            draws.append((f"is_{q}", "log_normal", 1e-16, 0.2))
            draws.append((f"bf_{q}", "normal", 100.0, 10.0))
        for r in (f"rp{k}", f"rn{k}"):
            d.bind(r, 0, r)
            draws.append((r, "rel_normal", 10e3, 0.01))
    return Family(d.title, d, draws)


def rc_mesh(k: int = 64, diode_frac: float = 0.05, layout_seed: int = 1) -> Family:
    """k x k RC mesh (5-point stencil), every edge resistor R ~ log-uniform
    [100, 1e4] Ohm, every node a grounded C ~ log-uniform [1 fF, 100 fF],
    diode clamps to ground on a seeded `diode_frac` of the nodes with IS ~
    log-uniform [1e-16, 1e-14] A; a 1 V step (0 V before t = 0) into node
    (0, 0) through 50 Ohm. Every R, C and IS is a live parameter, so each
    point is its own mesh on one shared pattern."""
    d = Deck(f"rc_mesh_{k}")
    node = lambda i, j: f"n{i}_{j}"
    d.vsource("vs", "src", 0, 0.0, step_to=1.0, step_at=0.0)
    d.resistor("rs", "src", node(0, 0), 50.0)
    draws = []
    for i in range(k):
        for j in range(k):
            if j + 1 < k:
                nm = f"rh{i}_{j}"
                d.resistor(nm, node(i, j), node(i, j + 1), 1e3)
                d.bind(nm, 0, nm)
                draws.append((nm, "log_uniform", 100.0, 1e4))
            if i + 1 < k:
                nm = f"rv{i}_{j}"
                d.resistor(nm, node(i, j), node(i + 1, j), 1e3)
                d.bind(nm, 0, nm)
                draws.append((nm, "log_uniform", 100.0, 1e4))
    for i in range(k):
        for j in range(k):
            nm = f"c{i}_{j}"
            d.capacitor(nm, node(i, j), 0, 1e-14)
            d.bind(nm, 0, nm)
            draws.append((nm, "log_uniform", 1e-15, 1e-13))
    pick = uniform01(layout_seed, np.arange(k * k, dtype=np.uint64), 0) < diode_frac
    for e in np.nonzero(pick)[0]:
        i, j = divmod(int(e), k)
        nm = f"d{i}_{j}"
        d.diode(nm, node(i, j), 0, {"is": 1e-15})
        d.bind(nm, "is", nm)
        draws.append((nm, "log_uniform", 1e-16, 1e-14))
    return Family(d.title, d, draws)


def inverter_chain_tran(stages: int = 16) -> Family:
    """The inverter chain as a transient: a 10 fF load on every stage output
    and the input stepping from 0 V to 1 V at t = 0; VDD ~ N(1.0 V, 2%),
    per-device VTO as in `inverter_chain`."""
    d = Deck(f"inverter_chain_tran_{stages}")
    nm = {"type": 1, "vto": 0.5, "kp": 200e-6, "lambda": 0.05}
    pm = {"type": -1, "vto": -0.5, "kp": 100e-6, "lambda": 0.05}
    d.vsource("vdd", "vdd", 0, 1.0)
    d.vsource("vin", "in", 0, 0.0, step_to=1.0, step_at=0.0)
    prev = "in"
    for k in range(stages):
        out = f"s{k + 1}"
        d.mos1(f"mn{k}", out, prev, 0, 0, nm, w=1e-6, l=0.1e-6)
        d.mos1(f"mp{k}", out, prev, "vdd", "vdd", pm, w=2e-6, l=0.1e-6)
        d.capacitor(f"cl{k}", out, 0, 10e-15)
        prev = out
    d.bind("vdd", 0, "vdd")
    for k in range(stages):
        d.bind(f"mn{k}", "vto", f"vto_n{k}")
        d.bind(f"mp{k}", "vto", f"vto_p{k}")
    draws = [("vdd", "rel_normal", 1.0, 0.02)]
    draws += [(f"vto_n{k}", "normal", 0.5, 0.01) for k in range(stages)]
    draws += [(f"vto_p{k}", "normal", -0.5, 0.01) for k in range(stages)]
    return Family(d.title, d, draws)


def rc_ladder(nodes: int = 64, tran: bool = False, clamp_every: int = 16) -> Family:
    """RC ladder of `nodes` nodes: a source through R into node 1, R between
    consecutive nodes, C from every node to ground, and a junction diode
    clamp to ground on every `clamp_every`-th node. The source is 1 V DC
    (.op) or steps from 0 V to 1 V at t = 0 (transient). Per point: every R ~
    N(1 kOhm, 5%), every C ~ log-uniform [10 fF, 100 fF], the clamps' IS ~
    log-uniform [1e-16, 1e-14] A (shared), the source level ~ N(1 V, 2%)."""
    d = Deck(f"rc_ladder_{'tran_' if tran else ''}{nodes}")
    if tran:
        d.vsource("vs", "src", 0, 0.0, step_to=1.0, step_at=0.0)
    else:
        d.vsource("vs", "src", 0, 1.0)
        d.bind("vs", 0, "vs")
    draws = [] if tran else [("vs", "rel_normal", 1.0, 0.02)]
    prev = "src"
    for k in range(nodes):
        nd = f"n{k}"
        d.resistor(f"r{k}", prev, nd, 1e3)
        d.bind(f"r{k}", 0, f"r{k}")
        draws.append((f"r{k}", "rel_normal", 1e3, 0.05))
        d.capacitor(f"c{k}", nd, 0, 30e-15)
        d.bind(f"c{k}", 0, f"c{k}")
        draws.append((f"c{k}", "log_uniform", 10e-15, 100e-15))
        if (k + 1) % clamp_every == 0:
            d.diode(f"d{k}", nd, 0, {"is": 1e-15})
            d.bind(f"d{k}", "is", "is_clamp")
        prev = nd
    if nodes >= clamp_every:
        draws.append(("is_clamp", "log_uniform", 1e-16, 1e-14))
    return Family(d.title, d, draws)


def diffpair_array_tran(count: int = 8) -> Family:
    """`diffpair_array` as a transient: 100 fF on every collector, and each
    pair's differential input flipping sign at t = 0 (vd_k/2 -> -vd_k/2 on
    one side, the opposite on the other). Same per-point draws."""
    d = Deck(f"diffpair_array_tran_{count}")
    d.vsource("vcc", "vcc", 0, 5.0)
    d.vsource("vee", "vee", 0, -5.0)
    qm = {"type": 1, "is": 1e-16, "bf": 100.0, "vaf": 100.0}
    draws = []
    for k in range(count):
        vd = -0.1 + 0.2 * k / max(count - 1, 1)
        d.vsource(f"vp{k}", f"bp{k}", 0, 0.5 * vd, step_to=-0.5 * vd, step_at=0.0)
        d.vsource(f"vn{k}", f"bn{k}", 0, -0.5 * vd, step_to=0.5 * vd, step_at=0.0)
        d.resistor(f"rp{k}", "vcc", f"cp{k}", 10e3)
        d.resistor(f"rn{k}", "vcc", f"cn{k}", 10e3)
        d.capacitor(f"ccp{k}", f"cp{k}", 0, 100e-15)
        d.capacitor(f"ccn{k}", f"cn{k}", 0, 100e-15)
        d.isource(f"it{k}", f"e{k}", "vee", 0.2e-3)
        d.bjt(f"q1_{k}", f"cp{k}", f"bp{k}", f"e{k}", 0, qm)
        d.bjt(f"q2_{k}", f"cn{k}", f"bn{k}", f"e{k}", 0, qm)
        for q in (f"q1_{k}", f"q2_{k}"):
            d.bind(q, "is", f"is_{q}")
            d.bind(q, "bf", f"bf_{q}")
            draws.append((f"is_{q}", "log_normal", 1e-16, 0.2))
            draws.append((f"bf_{q}", "normal", 100.0, 10.0))
        for r in (f"rp{k}", f"rn{k}"):
            d.bind(r, 0, r)
            draws.append((r, "rel_normal", 10e3, 0.01))
    return Family(d.title, d, draws)


def mirror_tree(depth: int = 4, tran: bool = False) -> Family:
    """A binary tree of MOS1 current mirrors `depth` levels deep. The root
    current flows from a reference source through 20 kOhm into a diode-
    connected NMOS; each tree node is a diode-connected device whose gate
    drives two mirror outputs of the same type, each output feeding a
    diode-connected device of the other type (NMOS sinks into PMOS diodes
    at VDD, PMOS sources into NMOS diodes at ground); the last level's
    outputs drive 10 kOhm loads to the opposite rail. The reference source
    is 1.2 V DC (.op) or steps from 1.0 V to 1.2 V at t = 0 with 20 fF on
    every tree node (transient). Per point: VDD ~ N(1.8 V, 2%), and per
    level and type VTO (sigma 10 mV) and KP (sigma 3%)."""
    d = Deck(f"mirror_tree_{'tran_' if tran else ''}{depth}")
    nm = {"type": 1, "vto": 0.5, "kp": 200e-6, "lambda": 0.05}
    pm = {"type": -1, "vto": -0.5, "kp": 100e-6, "lambda": 0.05}
    d.vsource("vdd", "vdd", 0, 1.8)
    if tran:
        d.vsource("vref", "ref", 0, 1.0, step_to=1.2, step_at=0.0)
    else:
        d.vsource("vref", "ref", 0, 1.2)
    d.resistor("rref", "ref", "g0_0", 20e3)
    d.mos1("md0_0", "g0_0", "g0_0", 0, 0, nm, w=2e-6, l=0.2e-6)
    d.bind("md0_0", "vto", "vto_n0")
    d.bind("md0_0", "kp", "kp_n0")
    if tran:
        d.capacitor("c0_0", "g0_0", 0, 20e-15)
    draws = [("vdd", "rel_normal", 1.8, 0.02)]
    used = {"vto_n0", "kp_n0"}
    for lev in range(depth):
        ntype = lev % 2 == 0        # this level's diodes and outputs: NMOS on even levels
        model, rail = (nm, "0") if ntype else (pm, "vdd")
        other, orail = (pm, "vdd") if ntype else (nm, "0")
        tname = "n" if ntype else "p"
        oname = "p" if ntype else "n"
        for i in range(2 ** lev):
            gate = f"g{lev}_{i}"
            for c in range(2):
                j = 2 * i + c
                out = f"g{lev + 1}_{j}"
                mo = f"mo{lev}_{j}"
                d.mos1(mo, out, gate, rail, rail, model, w=2e-6, l=0.2e-6)
                d.bind(mo, "vto", f"vto_{tname}{lev}")
                d.bind(mo, "kp", f"kp_{tname}{lev}")
                if lev + 1 < depth:
                    md = f"md{lev + 1}_{j}"
                    d.mos1(md, out, out, orail, orail, other, w=2e-6, l=0.2e-6)
                    d.bind(md, "vto", f"vto_{oname}{lev + 1}")
                    d.bind(md, "kp", f"kp_{oname}{lev + 1}")
                    used |= {f"vto_{oname}{lev + 1}", f"kp_{oname}{lev + 1}"}
                else:
                    d.resistor(f"rl{j}", out, orail, 10e3)
                if tran:
                    d.capacitor(f"c{lev + 1}_{j}", out, 0, 20e-15)
        used |= {f"vto_{tname}{lev}", f"kp_{tname}{lev}"}
    d.bind("vdd", 0, "vdd")
    for name in sorted(used, key=lambda v: (int(v[v.index("_") + 2:]), v)):
        t = name[name.index("_") + 1]
        m = nm if t == "n" else pm
        if name.startswith("vto"):
            draws.append((name, "normal", m["vto"], 0.01))
        else:
            draws.append((name, "rel_normal", m["kp"], 0.03))
    return Family(d.title, d, draws)


# The uneven batch: eight families, each drawing its size per instance from
# a geometric grid over its range (sizes repeat, so instances share
# templates), .op or a 200-step transient with the step stated per family.
MIXED_FAMILIES = (
    # name, builder(size), size grid, analysis, dt
    ("inverter_chain", inverter_chain, (8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256), "op", None),
    ("inverter_chain_tran", inverter_chain_tran, (8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256), "tran", 2e-12),
    ("rc_ladder", lambda k: rc_ladder(k), (16, 24, 36, 54, 81, 122, 182, 273, 410, 615, 923, 1384, 2048), "op", None),
    ("rc_ladder_tran", lambda k: rc_ladder(k, tran=True),
     (16, 24, 36, 54, 81, 122, 182, 273, 410, 615, 923, 1384, 2048), "tran", 10e-12),
    ("diffpair_array", diffpair_array, (4, 6, 8, 12, 16, 24, 32, 48, 64), "op", None),
    ("diffpair_array_tran", diffpair_array_tran, (4, 6, 8, 12, 16, 24, 32, 48, 64), "tran", 20e-12),
    ("mirror_tree", lambda k: mirror_tree(k), (2, 3, 4, 5, 6, 7, 8), "op", None),
    ("mirror_tree_tran", lambda k: mirror_tree(k, tran=True), (2, 3, 4, 5, 6, 7, 8), "tran", 20e-12),
)
MIXED_STEPS = 200


def mixed_batch(instances: int = 10000, seed: int = 2024, families=None):
    """The uneven batch: instance i belongs to family i mod 8 and draws its
    size from the family's grid (keyed by (seed, i)); instances of one
    (family, size) share a template and take their parameters from that
    family's sampler keyed by (seed, instance) (Family.sample_at). Returns a list of groups
    {"family", "size", "analysis", "dt", "steps", "instances", "build"}
    ordered by family then size."""
    fams = MIXED_FAMILIES if families is None else [f for f in MIXED_FAMILIES if f[0] in families]
    idx = np.arange(instances, dtype=np.uint64)
    u = uniform01(seed, idx, 0)
    groups = {}
    for i in range(instances):
        f = fams[i % len(fams)]
        grid = f[2]
        size = grid[min(int(u[i] * len(grid)), len(grid) - 1)]
        groups.setdefault((f[0], size), []).append(i)
    out = []
    for f in fams:
        for size in f[2]:
            inst = groups.get((f[0], size))
            if not inst:
                continue
            out.append({"family": f[0], "size": size, "analysis": f[3], "dt": f[4],
                        "steps": MIXED_STEPS if f[3] == "tran" else 0, "instances": np.array(inst, dtype=np.uint64),
                        "build": (lambda b=f[1], s=size: b(s))})
    return out


FAMILIES = {"inverter_chain": inverter_chain, "ota_array": ota_array, "diode_stack": diode_stack,
            "rc_mesh": rc_mesh, "diffpair_array": diffpair_array, "inverter_chain_tran": inverter_chain_tran,
            "rc_ladder": rc_ladder, "diffpair_array_tran": diffpair_array_tran, "mirror_tree": mirror_tree}

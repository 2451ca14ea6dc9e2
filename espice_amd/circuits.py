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
        idx = np.arange(first, first + lanes, dtype=np.uint64)
        out = np.zeros((lanes, len(self.deck.live)))
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


FAMILIES = {"inverter_chain": inverter_chain, "ota_array": ota_array, "diode_stack": diode_stack,
            "rc_mesh": rc_mesh, "diffpair_array": diffpair_array}

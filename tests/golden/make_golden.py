"""Writes tests/golden/ngspice_mos1.json from the reference's own fixtures.

Run where the reference tree is present (the tests only read the JSON):
    python tests/golden/make_golden.py

Each case pairs a deck, transcribed by hand below from the fixture's `.sp`
netlist (device list, models, sources), with the expected columns copied
verbatim from its `.expected.json` (ngspice-44.2 oracle values and the
fixture's own rtol/atol). Sweep rows (`.dc`) become points: the swept source
values are read back from the fixture's own node-voltage columns.
"""
from __future__ import annotations

import hashlib
import json
import os

REF = "/root/reference/tests/fixtures"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ngspice_mos1.json")
OUT_BJT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ngspice_bjt.json")
OUT_DIODE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ngspice_diode.json")

NM_REF = {"type": 1, "vto": 0.7, "kp": 200e-6, "gamma": 0.4, "phi": 0.6, "lambda": 0.02,
          "cgso": 100e-12, "cgdo": 100e-12}

# name -> (fixture path, devices, live sources: [source, column, sign])
# devices: (kind, name, nodes, value-or-model, instance dict)
CASES = {
    "mos_saturation": ("reference/mos_saturation", [
        ("V", "vg", ["gate", "0"], 2.0), ("V", "vd", ["drain", "0"], 3.0), ("V", "vb", ["bulk", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"], NM_REF, {"w": 10e-6, "l": 1e-6})], []),
    "mos_cutoff": ("reference/mos_cutoff", [
        ("V", "vg", ["gate", "0"], 0.0), ("V", "vd", ["drain", "0"], 3.0), ("V", "vb", ["bulk", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"], NM_REF, {"w": 10e-6, "l": 1e-6})], []),
    "mos_linear": ("reference/mos_linear", [
        ("V", "vg", ["gate", "0"], 2.0), ("V", "vd", ["drain", "0"], 0.1), ("V", "vb", ["bulk", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"], NM_REF, {"w": 10e-6, "l": 1e-6})], []),
    "mos_threshold": ("reference/mos_threshold", [
        ("V", "vg", ["gate", "0"], 0.7), ("V", "vd", ["drain", "0"], 3.0), ("V", "vb", ["bulk", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"], NM_REF, {"w": 10e-6, "l": 1e-6})], []),
    "mos_body_effect": ("reference/mos_body_effect", [
        ("V", "vg", ["gate", "0"], 2.0), ("V", "vd", ["drain", "0"], 3.0), ("V", "vb", ["bulk", "0"], -1.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"], NM_REF, {"w": 10e-6, "l": 1e-6})], []),
    "bench_mosfet_nand2": ("op/bench_mosfet_nand2", [
        ("V", "vdd", ["vdd", "0"], 5.0), ("V", "va", ["a", "0"], 5.0), ("V", "vb", ["b", "0"], 5.0),
        ("M", "m1", ["out", "a", "n1", "0"], {"type": 1, "vto": 0.7, "kp": 120e-6}, {"l": 1e-6, "w": 10e-6}),
        ("M", "m2", ["n1", "b", "0", "0"], {"type": 1, "vto": 0.7, "kp": 120e-6}, {"l": 1e-6, "w": 10e-6}),
        ("M", "m3", ["out", "a", "vdd", "vdd"], {"type": -1, "vto": -0.7, "kp": 60e-6}, {"l": 1e-6, "w": 20e-6}),
        ("M", "m4", ["out", "b", "vdd", "vdd"], {"type": -1, "vto": -0.7, "kp": 60e-6}, {"l": 1e-6, "w": 20e-6})],
        []),
    "cmos_inverter_dc": ("reference/cmos_inverter_dc", [
        ("V", "vdd", ["vdd", "0"], 3.3), ("V", "vin", ["in", "0"], 0.0),
        ("M", "mp", ["out", "in", "vdd", "vdd"], {"type": -1, "vto": -0.7, "kp": 100e-6, "lambda": 0.02},
         {"w": 20e-6, "l": 1e-6}),
        ("M", "mn", ["out", "in", "0", "0"], {"type": 1, "vto": 0.7, "kp": 200e-6, "lambda": 0.02},
         {"w": 10e-6, "l": 1e-6}),
        ("C", "cl", ["out", "0"], 1e-12)], [["vin", "v(in)", 1.0]]),
    "bench_mosfet_cmos_inverter": ("dc/bench_mosfet_cmos_inverter", [
        ("V", "vdd", ["vdd", "0"], 5.0), ("V", "vin", ["in", "0"], 0.0),
        ("M", "m1", ["out", "in", "0", "0"], {"type": 1, "vto": 0.7, "kp": 120e-6, "lambda": 0.01},
         {"l": 1e-6, "w": 10e-6}),
        ("M", "m2", ["out", "in", "vdd", "vdd"], {"type": -1, "vto": -0.7, "kp": 60e-6, "lambda": 0.01},
         {"l": 1e-6, "w": 20e-6})], [["vin", "v(in)", 1.0]]),
    "device_mos1_transfer": ("dc/device_mos1_transfer", [
        ("V", "vds", ["drain", "0"], 3.0), ("V", "vgs", ["gate", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "0"],
         {"type": 1, "vto": 0.7, "kp": 120e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.04}, {"w": 10e-6, "l": 1e-6})],
        [["vds", "v(drain)", 1.0], ["vgs", "v(gate)", 1.0]]),
    "device_mos1_output": ("dc/device_mos1_output", [
        ("V", "vds", ["drain", "0"], 0.0), ("V", "vgs", ["gate", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "0"],
         {"type": 1, "vto": 0.7, "kp": 120e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.04, "cbd": 20e-15,
          "cbs": 20e-15, "pb": 0.8, "mj": 0.5, "cgso": 0.6e-9, "cgdo": 0.6e-9}, {"w": 10e-6, "l": 1e-6})],
        [["vds", "v(drain)", 1.0], ["vgs", "v(gate)", 1.0]]),
    "device_mos1_pmos": ("dc/device_mos1_pmos", [
        ("V", "vsd", ["0", "drain"], 0.0), ("V", "vsg", ["0", "gate"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "0"],
         {"type": -1, "vto": -0.7, "kp": 60e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.05}, {"w": 20e-6, "l": 1e-6})],
        [["vsd", "v(drain)", -1.0], ["vsg", "v(gate)", -1.0]]),
    "device_mos1_subthreshold": ("dc/device_mos1_subthreshold", [
        ("V", "vds", ["drain", "0"], 1.0), ("V", "vgs", ["gate", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "0"],
         {"type": 1, "vto": 0.7, "kp": 120e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.04}, {"w": 10e-6, "l": 2e-6})],
        [["vds", "v(drain)", 1.0], ["vgs", "v(gate)", 1.0]]),
    "device_mos1_body_effect": ("dc/device_mos1_body_effect", [
        ("V", "vds", ["drain", "0"], 3.0), ("V", "vgs", ["gate", "0"], 0.0), ("V", "vbs", ["bulk", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"],
         {"type": 1, "vto": 0.7, "kp": 120e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.04}, {"w": 10e-6, "l": 1e-6})],
        [["vds", "v(drain)", 1.0], ["vgs", "v(gate)", 1.0], ["vbs", "v(bulk)", 1.0]]),
}

# dc/device_mos1_temp: `.dc Vgs 0 5 0.01 TEMP -40 125 55`, the outer sweep
# over -40, 15, 70 and 125 degC. One case per temperature: the model at that
# temperature (TOX and UO are given, but so is KP, which wins), and the
# fixture's selected rows of that block (row index // 501).
TEMP_ROWS = 501
for k, t in enumerate((-40.0, 15.0, 70.0, 125.0)):
    CASES[f"device_mos1_temp_{int(t)}".replace("-", "m")] = ("dc/device_mos1_temp", [
        ("V", "vds", ["drain", "0"], 3.0), ("V", "vgs", ["gate", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "0"],
         {"type": 1, "vto": 0.7, "kp": 120e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.04, "tox": 40e-9,
          "u0": 600.0, "temp": t}, {"w": 10e-6, "l": 1e-6})],
        [["vds", "v(drain)", 1.0], ["vgs", "v(gate)", 1.0]], {"row_block": TEMP_ROWS, "block": k})

QM = {"type": 1, "is": 1e-15, "bf": 100, "br": 2, "vaf": 100}
QD = {"type": 1, "is": 1e-15, "bf": 200, "vaf": 100}
BJT_CASES = {
    "bjt_active": ("reference/bjt_active", [
        ("V", "vb", ["b", "0"], 0.7), ("V", "vc", ["supply", "0"], 5.0), ("R", "rc", ["supply", "c"], 1e3),
        ("Q", "q1", ["c", "b", "0", "0"], QM)], []),
    "bjt_cutoff": ("reference/bjt_cutoff", [
        ("V", "vb", ["b", "0"], 0.0), ("V", "vc", ["supply", "0"], 5.0), ("R", "rc", ["supply", "c"], 1e3),
        ("Q", "q1", ["c", "b", "0", "0"], QM)], []),
    "bjt_saturation": ("reference/bjt_saturation", [
        ("V", "vb", ["b", "0"], 0.9), ("V", "vc", ["supply", "0"], 0.2), ("R", "rc", ["supply", "c"], 1e3),
        ("Q", "q1", ["c", "b", "0", "0"], QM)], []),
    "bjt_reverse_active": ("reference/bjt_reverse_active", [
        ("V", "vb", ["b", "0"], 0.8), ("V", "vc", ["supply", "0"], 0.0), ("R", "rc", ["supply", "c"], 1e3),
        ("Q", "q1", ["c", "b", "0", "0"], QM)], []),
}
for tag, vp in (("0p1", 0.05), ("-0p1", -0.05), ("0", 0.0)):
    BJT_CASES[f"bjt_diffpair_{tag}"] = (f"reference/bjt_diffpair_{tag}", [
        ("V", "vcc", ["vcc", "0"], 5.0), ("V", "vp", ["bp", "0"], vp), ("V", "vn", ["bn", "0"], -vp),
        ("I", "itail", ["emit", "neg"], 1e-3), ("V", "vneg", ["neg", "0"], -5.0),
        ("R", "r1", ["vcc", "op"], 2e3), ("R", "r2", ["vcc", "on"], 2e3),
        ("Q", "q1", ["op", "bp", "emit", "0"], QD), ("Q", "q2", ["on", "bn", "emit", "0"], QD)], [])
BJT_CASES["bench_analog_diff_pair"] = ("op/bench_analog_diff_pair", [
    ("V", "vcc", ["vcc", "0"], 5.0), ("V", "vee", ["vee", "0"], -5.0), ("V", "vp", ["inp", "0"], 0.1),
    ("V", "vn", ["inn", "0"], -0.1), ("R", "rcp", ["vcc", "cp"], 4.7e3), ("R", "rcn", ["vcc", "cn"], 4.7e3),
    ("R", "rt", ["tail", "vee"], 4.7e3),
    ("Q", "q1", ["cp", "inp", "tail", "0"], {"type": 1, "is": 1e-16, "bf": 100}),
    ("Q", "q2", ["cn", "inn", "tail", "0"], {"type": 1, "is": 1e-16, "bf": 100})], [])
BJT_CASES["bench_analog_current_mirror"] = ("op/bench_analog_current_mirror", [
    ("V", "vcc", ["vcc", "0"], 5.0), ("I", "iref", ["vcc", "ref"], 1e-3), ("R", "rl", ["vcc", "out"], 1e3),
    ("Q", "q1", ["ref", "ref", "0", "0"], {"type": 1, "is": 1e-16, "bf": 200}),
    ("Q", "q2", ["out", "ref", "0", "0"], {"type": 1, "is": 1e-16, "bf": 200})], [])


DIODE_CASES = {
    "device_diode_iv_sweep": ("dc/device_diode_iv_sweep", [
        ("V", "v1", ["anode", "0"], 0.0),
        ("D", "d1", ["anode", "0"], {"is": 1e-14, "n": 1, "rs": 10, "bv": 100, "ibv": 1e-10})],
        [["v1", "v(anode)", 1.0]]),
    "device_diode_breakdown": ("dc/device_diode_breakdown", [
        ("V", "v1", ["anode", "0"], 0.0), ("R", "r1", ["anode", "cathode"], 100.0),
        ("D", "d1", ["cathode", "0"], {"is": 1e-14, "n": 1, "bv": 15, "ibv": 1e-3, "rs": 1})],
        [["v1", "v(anode)", 1.0]]),
    "device_diode_high_injection": ("dc/device_diode_high_injection", [
        ("V", "v1", ["anode", "0"], 0.0),
        ("D", "d1", ["anode", "0"], {"is": 1e-14, "n": 1, "rs": 50, "ikf": 10e-3})],
        [["v1", "v(anode)", 1.0]]),
    "device_diode_recombination": ("dc/device_diode_recombination", [
        ("V", "v1", ["anode", "0"], 0.0),
        ("D", "d1", ["anode", "0"], {"is": 1e-14, "n": 1, "isr": 1e-10, "nr": 2, "rs": 5})],
        [["v1", "v(anode)", 1.0]]),
}
# dc/device_diode_temp: `.dc V1 0 0.9 0.005 TEMP -40 125 55`, the outer sweep
# over -40, 15, 70 and 125 degC (181 rows each): one case per temperature.
for k, t in enumerate((-40.0, 15.0, 70.0, 125.0)):
    DIODE_CASES[f"device_diode_temp_{int(t)}".replace("-", "m")] = ("dc/device_diode_temp", [
        ("V", "v1", ["anode", "0"], 0.0),
        ("D", "d1", ["anode", "0"], {"is": 1e-14, "n": 1.05, "rs": 5, "xti": 3, "eg": 1.11, "temp": t})],
        [["v1", "v(anode)", 1.0]], {"row_block": 181, "block": k})


def write(cases, path):
    out = {"source": "OmarSiwy/ESPice 1.2.0 tests/fixtures (ngspice-44.2 oracle values)", "cases": {}}
    for name, (fixture, devices, live, *sel) in cases.items():
        exp = json.load(open(os.path.join(REF, fixture + ".expected.json")))
        net = open(os.path.join(REF, fixture + ".sp"), "rb").read()
        plot = exp["expect"]["plots"][0]
        keep = None
        if sel:   # the rows of one block of an outer sweep
            idx = plot["row_indices"]
            keep = [i for i, r in enumerate(idx) if r // sel[0]["row_block"] == sel[0]["block"]]
        cols = {k: {"values": v["values"] if keep is None else [v["values"][i] for i in keep],
                    "rtol": v.get("rtol", 1e-3), "atol": v.get("atol", 0.0)}
                for k, v in plot["columns"].items()}
        out["cases"][name] = {
            "fixture": fixture,
            "netlist_sha256": hashlib.sha256(net).hexdigest(),
            "oracle": exp["oracle"].get("version", exp["oracle"].get("kind", "")),
            "devices": devices,
            "live": live,
            "rows": plot.get("row_count", 1),
            "columns": cols,
        }
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {len(out['cases'])} cases")


def main():
    write(CASES, OUT)
    write(BJT_CASES, OUT_BJT)
    write(DIODE_CASES, OUT_DIODE)


if __name__ == "__main__":
    main()

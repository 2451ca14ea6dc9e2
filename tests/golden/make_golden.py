"""Writes tests/golden/ngspice_mos1.json from the reference's own fixtures.

Run in the build container (the reference tree is not on the GPU box):
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
    "device_mos1_body_effect": ("dc/device_mos1_body_effect", [
        ("V", "vds", ["drain", "0"], 3.0), ("V", "vgs", ["gate", "0"], 0.0), ("V", "vbs", ["bulk", "0"], 0.0),
        ("M", "m1", ["drain", "gate", "0", "bulk"],
         {"type": 1, "vto": 0.7, "kp": 120e-6, "gamma": 0.4, "phi": 0.65, "lambda": 0.04}, {"w": 10e-6, "l": 1e-6})],
        [["vds", "v(drain)", 1.0], ["vgs", "v(gate)", 1.0], ["vbs", "v(bulk)", 1.0]]),
}


def main():
    out = {"source": "OmarSiwy/ESPice 1.2.0 tests/fixtures (ngspice-44.2 oracle values)", "cases": {}}
    for name, (path, devices, live) in CASES.items():
        exp = json.load(open(os.path.join(REF, path + ".expected.json")))
        net = open(os.path.join(REF, path + ".sp"), "rb").read()
        plot = exp["expect"]["plots"][0]
        cols = {k: {"values": v["values"], "rtol": v.get("rtol", 1e-3), "atol": v.get("atol", 0.0)}
                for k, v in plot["columns"].items()}
        out["cases"][name] = {
            "fixture": path,
            "netlist_sha256": hashlib.sha256(net).hexdigest(),
            "oracle": exp["oracle"].get("version", exp["oracle"].get("kind", "")),
            "devices": devices,
            "live": live,
            "rows": plot.get("row_count", 1),
            "columns": cols,
        }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}: {len(out['cases'])} cases")


if __name__ == "__main__":
    main()

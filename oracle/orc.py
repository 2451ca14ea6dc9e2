"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to check results against and to time the CPU
path. The espice_amd package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class Deck(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("n_dev", C.c_uint32),
        ("kind", C.c_void_p), ("nodes", C.c_void_p), ("params", C.c_void_p),
        ("n_live", C.c_uint32), ("n_bind", C.c_uint32),
        ("bind_dev", C.c_void_p), ("bind_param", C.c_void_p), ("bind_live", C.c_void_p),
    ]


class Tol(C.Structure):
    _fields_ = [
        ("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double),
        ("residual_tol", C.c_double), ("max_iter", C.c_uint32), ("reserved", C.c_uint32),
    ]


class TranAdapt(C.Structure):
    _fields_ = [("t_stop", C.c_double), ("t_step", C.c_double), ("dt_max", C.c_double), ("dt_min", C.c_double),
                ("trtol", C.c_double), ("chgtol", C.c_double), ("max_steps", C.c_uint32), ("reserved", C.c_uint32)]


class Prec(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("refine_max", C.c_uint32), ("refine_tol", C.c_double),
                ("nominal_x", C.c_void_p)]


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_exp.restype = C.c_double
        L.orc_exp.argtypes = [C.c_double]
        L.orc_log.restype = C.c_double
        L.orc_log.argtypes = [C.c_double]
        L.orc_lu_factor_export.restype = C.c_int
        L.orc_lu_factor_export.argtypes = [C.c_uint32, u32p, u32p, f64p, u32p, C.c_double, u32p,
                                           u32p, u32p, u32p, u32p, u32p, u32p, u32p,
                                           f64p, f64p, f64p, C.c_size_t, C.c_size_t]
        L.orc_lu_replay_export.restype = C.c_int
        L.orc_lu_replay_export.argtypes = [C.c_uint32, u32p, u32p, f64p, u32p, C.c_double, f64p, f64p,
                                           C.c_size_t, C.c_double, f64p, f64p, f64p, f64p, u32p]
        L.orc_lu_solve_export.restype = C.c_int
        L.orc_lu_solve_export.argtypes = [C.c_uint32, u32p, u32p, f64p, u32p, C.c_double, f64p, f64p]
        L.orc_meas_rows.restype = C.c_size_t
        L.orc_meas_rows.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                    C.c_int, C.c_void_p]
        L.orc_ckt_new.restype = C.c_void_p
        L.orc_ckt_new.argtypes = [C.POINTER(Deck), C.c_char_p, C.c_size_t]
        L.orc_ckt_free.argtypes = [C.c_void_p]
        L.orc_ckt_n.restype = C.c_uint32
        L.orc_ckt_n.argtypes = [C.c_void_p]
        L.orc_ckt_nnz.restype = C.c_uint32
        L.orc_ckt_nnz.argtypes = [C.c_void_p]
        L.orc_ckt_pattern.argtypes = [C.c_void_p, u32p, u32p, u32p, u8p]
        L.orc_ckt_assemble_cold.argtypes = [C.c_void_p, f64p, f64p, f64p]
        L.orc_ckt_assemble_at.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p]
        L.orc_ckt_assemble_scaled.argtypes = [C.c_void_p, f64p, f64p, C.c_double, f64p, f64p]
        L.orc_op_solve2.restype = C.c_int
        L.orc_op_solve2.argtypes = [C.c_void_p, f64p, C.c_size_t, C.POINTER(Tol), C.POINTER(Prec), u32p,
                                    C.c_void_p, C.c_int, f64p, u32p, u32p, u32p, u32p, u32p]
        L.orc_lu_replay32_export.restype = C.c_int
        L.orc_lu_replay32_export.argtypes = [C.c_uint32, u32p, u32p, f64p, u32p, C.c_double, f64p, f64p,
                                             C.c_size_t, C.c_double, f32p, f32p, f32p, f32p, u32p]
        L.orc_tran_solve.restype = C.c_int
        L.orc_tran_solve.argtypes = [C.c_void_p, f64p, C.c_size_t, C.POINTER(Tol), C.POINTER(Tol), C.c_double,
                                     C.c_uint32, C.c_void_p, C.c_uint32, u32p, C.c_int, f64p, C.c_void_p, u32p, u32p]
        L.orc_ckt_assemble_charges.argtypes = [C.c_void_p, f64p, f64p, C.c_double, f64p, f64p, f64p, f64p]
        L.orc_tran_solve_traced.restype = C.c_int
        L.orc_tran_solve_traced.argtypes = [C.c_void_p, f64p, C.c_size_t, C.POINTER(Tol), C.POINTER(Tol),
                                            C.c_double, C.c_uint32, C.c_void_p, C.c_uint32, u32p, C.c_int, f64p,
                                            C.c_void_p, u32p, u32p, C.c_void_p, C.c_void_p]
        L.orc_tran_adaptive.restype = C.c_int
        L.orc_tran_adaptive.argtypes = [C.c_void_p, f64p, C.c_size_t, C.POINTER(Tol), C.POINTER(Tol),
                                        C.POINTER(TranAdapt), C.c_void_p, C.c_uint32, u32p, C.c_int, f64p,
                                        C.c_void_p, C.c_void_p, u32p, u32p, u32p, u32p]
        L.orc_op_solve.restype = C.c_int
        L.orc_op_solve.argtypes = [C.c_void_p, f64p, C.c_size_t, C.POINTER(Tol), u32p, C.c_void_p,
                                   C.c_int, f64p, u32p, u32p, u32p]
        _LIB = L
    return _LIB


def exp(x: float) -> float:
    return lib().orc_exp(float(x))


def log(x: float) -> float:
    return lib().orc_log(float(x))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lu_factor(col_ptr, row_idx, vals, q, pivot_tol=1e-3):
    """Full factor. Returns dict of the tape and factors, or None if singular."""
    col_ptr, row_idx, q = _u32(col_ptr), _u32(row_idx), _u32(q)
    vals = _f64(vals)
    n = len(col_ptr) - 1
    nnz = int(col_ptr[-1])
    cap = max(4 * nnz, 64) + n * n
    cap = min(cap, 50_000_000)
    sizes = np.zeros(3, np.uint32)
    out = {k: np.zeros(s, np.uint32) for k, s in
           [("pinv", n), ("lp", n + 1), ("up", n + 1), ("flags", n), ("prow", max(nnz, 1))]}
    li = np.zeros(cap, np.uint32)
    ui = np.zeros(cap, np.uint32)
    lx = np.zeros(cap)
    ux = np.zeros(cap)
    ud = np.zeros(n)
    rc = lib().orc_lu_factor_export(n, col_ptr, row_idx, vals, q, pivot_tol, sizes, out["pinv"], out["lp"],
                                    li, out["up"], ui, out["flags"], out["prow"], lx, ux, ud, cap, cap)
    if rc < 0:
        return None
    if rc > 0:
        raise RuntimeError("factor capacity")
    nl, nu = int(sizes[0]), int(sizes[1])
    out.update(li=li[:nl].copy(), ui=ui[:nu].copy(), lx=lx[:nl].copy(), ux=ux[:nu].copy(), udiag=ud,
               nnz_l=nl, nnz_u=nu, n_void_slots=int(sizes[2]), prow=out["prow"][:nnz])
    return out


def lu_replay(col_ptr, row_idx, lead_vals, q, a, rhs, growth=1e-12, pivot_tol=1e-3):
    """Lane replay of the tape built from lead_vals, over a [lanes][nnz]."""
    col_ptr, row_idx, q = _u32(col_ptr), _u32(row_idx), _u32(q)
    lead_vals, a, rhs = _f64(lead_vals), _f64(a), _f64(rhs)
    t = lu_factor(col_ptr, row_idx, lead_vals, q, pivot_tol)
    if t is None:
        return None
    n = len(col_ptr) - 1
    lanes = a.shape[0]
    lx = np.zeros((lanes, max(t["nnz_l"], 1)))
    ux = np.zeros((lanes, max(t["nnz_u"], 1)))
    ud = np.zeros((lanes, n))
    dx = np.zeros((lanes, n))
    fail = np.zeros(lanes, np.uint32)
    rc = lib().orc_lu_replay_export(n, col_ptr, row_idx, lead_vals, q, pivot_tol, a, rhs, lanes, growth,
                                    lx, ux, ud, dx, fail)
    if rc != 0:
        return None
    return dict(lx=lx[:, :t["nnz_l"]], ux=ux[:, :t["nnz_u"]], udiag=ud, dx=dx, fail=fail, tape=t)


def lu_replay32(col_ptr, row_idx, lead_vals, q, a, b, growth=1e-12, pivot_tol=1e-3):
    """FP32 lane replay of the (FP64) tape built from lead_vals over a
    [lanes][nnz]; x = A32^-1 b in FP32."""
    col_ptr, row_idx, q = _u32(col_ptr), _u32(row_idx), _u32(q)
    lead_vals, a, b = _f64(lead_vals), _f64(a), _f64(b)
    t = lu_factor(col_ptr, row_idx, lead_vals, q, pivot_tol)
    if t is None:
        return None
    n = len(col_ptr) - 1
    lanes = a.shape[0]
    lx = np.zeros((lanes, max(t["nnz_l"], 1)), np.float32)
    ux = np.zeros((lanes, max(t["nnz_u"], 1)), np.float32)
    ud = np.zeros((lanes, n), np.float32)
    x = np.zeros((lanes, n), np.float32)
    fail = np.zeros(lanes, np.uint32)
    rc = lib().orc_lu_replay32_export(n, col_ptr, row_idx, lead_vals, q, pivot_tol, a, b, lanes, growth,
                                      lx, ux, ud, x, fail)
    if rc != 0:
        return None
    return dict(lx=lx[:, :t["nnz_l"]], ux=ux[:, :t["nnz_u"]], udiag=ud, x=x, fail=fail, tape=t)


def lu_solve(col_ptr, row_idx, vals, q, b, pivot_tol=1e-3):
    col_ptr, row_idx, q = _u32(col_ptr), _u32(row_idx), _u32(q)
    n = len(col_ptr) - 1
    x = np.zeros(n)
    rc = lib().orc_lu_solve_export(n, col_ptr, row_idx, _f64(vals), q, pivot_tol, _f64(b), x)
    return None if rc != 0 else x


class Circuit:
    """The oracle's MNA circuit for a deck (espice_amd.deck.Deck.arrays())."""

    def __init__(self, arrays: dict):
        self._keep = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        a = self._keep
        d = Deck()
        d.n_nodes = a["n_nodes"]
        d.n_dev = len(a["kind"])
        d.kind = a["kind"].ctypes.data
        d.nodes = a["nodes"].ctypes.data
        d.params = a["params"].ctypes.data
        d.n_live = a["n_live"]
        d.n_bind = len(a["bind_dev"])
        d.bind_dev = a["bind_dev"].ctypes.data
        d.bind_param = a["bind_param"].ctypes.data
        d.bind_live = a["bind_live"].ctypes.data
        err = C.create_string_buffer(256)
        self.h = lib().orc_ckt_new(C.byref(d), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())
        self.n = lib().orc_ckt_n(self.h)
        self.nnz = lib().orc_ckt_nnz(self.h)
        self.n_live = a["n_live"]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ckt_free(self.h)
            self.h = None

    def pattern(self):
        cp = np.zeros(self.n + 1, np.uint32)
        ri = np.zeros(max(self.nnz, 1), np.uint32)
        dg = np.zeros(self.n, np.uint32)
        cr = np.zeros(self.n, np.uint8)
        lib().orc_ckt_pattern(self.h, cp, ri, dg, cr)
        return cp, ri[: self.nnz], dg, cr

    def assemble_cold(self, live):
        g = np.zeros(self.nnz)
        r = np.zeros(self.n)
        lib().orc_ckt_assemble_cold(self.h, _f64(live), g, r)
        return g, r

    def assemble_at(self, live, x, src=None):
        """Jacobian (pattern slots) and residual at x; src: every
        independent source scaled by it."""
        g = np.zeros(self.nnz)
        r = np.zeros(self.n)
        if src is None:
            lib().orc_ckt_assemble_at(self.h, _f64(live), _f64(x), g, r)
        else:
            lib().orc_ckt_assemble_scaled(self.h, _f64(live), _f64(x), float(src), g, r)
        return g, r

    def assemble_charges(self, live, x, t=0.0):
        """Jacobian g, residual f, charge Jacobian c (the c plane, pattern
        slots) and row charges q at x (no limiting), sources at time t."""
        g = np.zeros(self.nnz)
        r = np.zeros(self.n)
        cp = np.zeros(self.nnz)
        q = np.zeros(self.n)
        lib().orc_ckt_assemble_charges(self.h, _f64(live), _f64(x), float(t), g, r, cp, q)
        return g, r, cp, q

    def op(self, values, q, start_x=None, threads=1, tol=None, precision=0, refine_tol=1e-12, refine_max=8,
           nominal_x=None):
        """Operating points; precision 1 = FP32 factors + FP64 refinement;
        nominal_x with start_x = a learned start with the nominal fallback."""
        values = _f64(np.atleast_2d(values))
        count = values.shape[0]
        t = Tol(1e-3, 1e-12, 1e-6, 1e-9, 100, 0) if tol is None else tol
        nom = None if nominal_x is None else _f64(nominal_x)
        p = Prec(int(precision), int(refine_max), float(refine_tol), None if nom is None else nom.ctypes.data)
        x = np.zeros((count, self.n))
        it = np.zeros(count, np.uint32)
        rs = np.zeros(count, np.uint32)
        rp = np.zeros(count, np.uint32)
        fb = np.zeros(count, np.uint32)
        wb = np.zeros(count, np.uint32)
        sx = None if start_x is None else _f64(start_x)
        lib().orc_op_solve2(self.h, values, count, C.byref(t), C.byref(p), _u32(q),
                            None if sx is None else sx.ctypes.data, int(threads), x, it, rs, rp, fb, wb)
        return dict(x=x, iterations=it, reasons=rs, repivots=rp, fp64_lanes=fb, fell_back=wb)


NO_PREDICT = 16   # ORC_TOL_NO_PREDICT: steps start at the accepted point


def tran(circuit, values, q, dt, steps, probes=(), threads=1, tol=None, tol_op=None, trace=False):
    """Transient lanes on the oracle (see orc_tran_solve). trace: also
    returns each step's first iterate (first_x [count][steps][n]) and the
    limiting state after the predictor's limit pass (first_lim
    [count][steps][n_inst][4]); step 0's rows are zero."""
    values = _f64(np.atleast_2d(values))
    count = values.shape[0]
    t = Tol(1e-3, 1e-12, 1e-6, 1e-9, 50, 0) if tol is None else tol
    to = Tol(1e-3, 1e-12, 1e-6, 1e-9, 100, 0) if tol_op is None else tol_op
    pr = np.ascontiguousarray(probes, dtype=np.uint32)
    x = np.zeros((count, circuit.n))
    po = np.zeros((count, steps, max(len(pr), 1)))
    it = np.zeros(count, np.uint32)
    rs = np.zeros(count, np.uint32)
    fx = fl = None
    if trace:
        n_inst = int(len(circuit._keep["kind"]))
        fx = np.zeros((count, steps, circuit.n))
        fl = np.zeros((count, steps, n_inst, 4))
    lib().orc_tran_solve_traced(circuit.h, values, count, C.byref(to), C.byref(t), float(dt), int(steps),
                                pr.ctypes.data if len(pr) else None, len(pr), _u32(q), int(threads), x,
                                po.ctypes.data, it, rs, None if fx is None else fx.ctypes.data,
                                None if fl is None else fl.ctypes.data)
    out = dict(x=x, probes=po[:, :, :len(pr)], iterations=it, reasons=rs)
    if trace:
        out.update(first_x=fx, first_lim=fl)
    return out


# ---- .meas cards (orc_meas.c) ----
MEAS_FUNCS = {"when": 1, "trig_targ": 2, "find": 3, "avg": 4, "min": 5, "max": 6, "pp": 7, "integ": 8, "rms": 9}
MEAS_UNSET, MEAS_LAST, MEAS_NO_AT = -1, -2, 1e99


class MeasClause(C.Structure):
    _fields_ = [("node", C.c_uint32), ("level_node", C.c_uint32), ("has_level_node", C.c_uint32),
                ("pad", C.c_uint32), ("val", C.c_double), ("td", C.c_double), ("frm", C.c_double),
                ("to", C.c_double), ("at", C.c_double), ("rise", C.c_int32), ("fall", C.c_int32),
                ("cross", C.c_int32), ("pad2", C.c_int32)]


class MeasCard(C.Structure):
    _fields_ = [("func", C.c_uint32), ("pad", C.c_uint32 * 3), ("a", MeasClause), ("b", MeasClause)]


def meas_clause(node=0, val=1e99, td=0.0, frm=0.0, to=0.0, at=MEAS_NO_AT, rise=MEAS_UNSET, fall=MEAS_UNSET,
                cross=MEAS_UNSET, level_node=None):
    """One clause with the reference's defaults (core/deck.zig:311-340)."""
    return MeasClause(node, 0 if level_node is None else level_node, 0 if level_node is None else 1, 0, val, td, frm,
                      to, at, rise, fall, cross, 0)


def meas_card(func, a, b=None):
    c = MeasCard()
    c.func = MEAS_FUNCS[func]
    c.a = a
    c.b = b if b is not None else meas_clause()
    return c


def meas_rows(cards, t, x, early=False):
    """Each card over the rows t[r], x[r][node], fed one row at a time;
    returns (values, rows read: all, or up to the row that resolved every
    card when early)."""
    L = lib()
    arr = (MeasCard * len(cards))(*cards)
    t = _f64(t)
    x = _f64(np.atleast_2d(x))
    out = np.zeros(len(cards))
    rows = L.orc_meas_rows(arr, len(cards), t.ctypes.data, x.ctypes.data, len(t), x.shape[1], int(early),
                           out.ctypes.data)
    return out, int(rows)


def natural_order(n):
    return np.arange(n, dtype=np.uint32)


def tran_adaptive(circuit, values, q, t_stop, t_step, dt_max=None, dt_min=1e-18, trtol=7.0, chgtol=1e-14,
                  max_steps=4000, probes=(), threads=1, tol=None, tol_op=None):
    """Adaptive transient lanes on the oracle (orc_tran_adaptive). dt_max
    defaults to t_stop / 50 (ngspice's tmax). Returns x (last accepted
    state), times [count][max_steps], probes [count][max_steps][n_probes]
    (rows beyond `accepted` are zero), accepted, rejected, iterations,
    reasons."""
    values = _f64(np.atleast_2d(values))
    count = values.shape[0]
    t = Tol(1e-3, 1e-12, 1e-6, 1e-9, 100, 0) if tol is None else tol
    to = Tol(1e-3, 1e-12, 1e-6, 1e-9, 100, 0) if tol_op is None else tol_op
    o = TranAdapt(float(t_stop), float(t_step), float(t_stop / 50.0 if dt_max is None else dt_max), float(dt_min),
                  float(trtol), float(chgtol), int(max_steps), 0)
    pr = np.ascontiguousarray(probes, dtype=np.uint32)
    x = np.zeros((count, circuit.n))
    tm = np.zeros((count, max_steps))
    po = np.zeros((count, max_steps, max(len(pr), 1)))
    acc = np.zeros(count, np.uint32)
    rej = np.zeros(count, np.uint32)
    it = np.zeros(count, np.uint32)
    rs = np.zeros(count, np.uint32)
    lib().orc_tran_adaptive(circuit.h, values, count, C.byref(to), C.byref(t), C.byref(o),
                            pr.ctypes.data if len(pr) else None, len(pr), _u32(q), int(threads), x,
                            tm.ctypes.data, po.ctypes.data, acc, rej, it, rs)
    return dict(x=x, times=tm, probes=po[:, :, :len(pr)], accepted=acc, rejected=rej, iterations=it, reasons=rs)

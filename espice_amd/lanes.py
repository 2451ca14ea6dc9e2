"""Python host mirror of the lane engine's C ABI (include/espice_lanes.h).

`Engine` plays the role of the reference's points API for one template
(include/espice.h:244-360: a deck with live parameters, rows of values in,
one result per point out) and of `GpuLanes` underneath it
(src/analysis/gpu_lanes.zig:251-1027). `Tape` is the LU boundary alone
(`arp_lane_refactor` + `arp_lane_solve`, src/solver/lane_device.zig:26-38).

There is no fallback: without the built library or without a HIP device the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libespice_lanes.so")

OK, INVALID_ARGUMENT, BUFFER_TOO_SMALL, OUT_OF_MEMORY = 0, 1, 2, 3
FAILED, ABI_MISMATCH, INVALID_HANDLE, DEVICE_UNAVAILABLE = 6, 7, 8, 9
REASON_OK, REASON_NO_CONVERGENCE, REASON_SINGULAR = 0, 1, 3
FP64, FP32_REFINED = 0, 1
START_COLD, START_GIVEN, START_LEARNED = 0, 1, 2
TILING_AUTO, TILING_WAVE, TILING_THREAD, TILING_BLOCK = 0, 1, 2, 3
OP_HOST_REPIVOT, OP_LADDER = 1, 2
BATCH_OP, BATCH_TRAN = 0, 1

STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "BUFFER_TOO_SMALL", 3: "OUT_OF_MEMORY",
                4: "INVALID_QUERY", 5: "RESULT_UNAVAILABLE", 6: "FAILED", 7: "ABI_MISMATCH",
                8: "INVALID_HANDLE", 9: "DEVICE_UNAVAILABLE"}

EXPORTS = [
    "esl_abi_version", "esl_default_tolerances", "esl_default_op_options", "esl_engine_create",
    "esl_engine_destroy", "esl_engine_info_get", "esl_engine_pattern", "esl_op_solve", "esl_last_timing",
    "esl_tape_create", "esl_tape_destroy", "esl_tape_sizes", "esl_tape_export", "esl_tape_replay",
    "esl_tape_replay_noalloc",
    "esl_mlp_default_train_options", "esl_mlp_create", "esl_mlp_destroy", "esl_mlp_param_count",
    "esl_mlp_get_params", "esl_mlp_set_params", "esl_mlp_set_scaling", "esl_mlp_gradient", "esl_mlp_train",
    "esl_mlp_predict", "esl_engine_set_warm_start", "esl_default_tran_options", "esl_tran_solve",
    "esl_default_batch_options", "esl_batch_solve", "esl_error_message", "esl_default_tran_adaptive_options",
    "esl_tran_adaptive_solve",
]


class EslError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")
        self.status = status


class _Deck(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("n_dev", C.c_uint32),
        ("kind", C.c_void_p), ("nodes", C.c_void_p), ("params", C.c_void_p),
        ("n_live", C.c_uint32), ("n_bind", C.c_uint32),
        ("bind_dev", C.c_void_p), ("bind_param", C.c_void_p), ("bind_live", C.c_void_p),
    ]


class Tolerances(C.Structure):
    _fields_ = [("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double),
                ("residual_tol", C.c_double), ("max_iter", C.c_uint32), ("reserved", C.c_uint32)]


class OpOptions(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("tol", Tolerances), ("precision", C.c_uint32),
                ("start", C.c_uint32), ("x0", C.c_void_p), ("refine_tol", C.c_double),
                ("refine_max", C.c_uint32), ("tiling", C.c_uint32), ("probes", C.c_void_p),
                ("n_probes", C.c_uint32), ("flags", C.c_uint32)]


class EngineInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("nnz", C.c_uint32), ("n_nodes", C.c_uint32), ("n_live", C.c_uint32),
                ("nnz_l", C.c_uint32), ("nnz_u", C.c_uint32), ("flops", C.c_uint64), ("levels", C.c_uint32),
                ("lds_bytes", C.c_uint32), ("tapes", C.c_uint32), ("reserved", C.c_uint32),
                ("refactor_steps", C.c_uint32), ("solve_steps", C.c_uint32), ("tiling", C.c_uint32),
                ("reserved2", C.c_uint32)]


class Timing(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("launches", C.c_uint32),
                ("repivots", C.c_uint32), ("lane_iterations", C.c_uint64), ("tiling", C.c_uint32),
                ("fp64_lanes", C.c_uint32), ("warm_fallbacks", C.c_uint32), ("device_repivot_lanes", C.c_uint32),
                ("warm_ms", C.c_double), ("lead_ms", C.c_double), ("download_ms", C.c_double)]


# progress(ctx, steps_done, steps) -> nonzero cancels (esl_tran_options.progress)
PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32)


class TranOptions(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("tol", Tolerances), ("dt", C.c_double), ("steps", C.c_uint32),
                ("steps_per_launch", C.c_uint32), ("probes", C.c_void_p), ("n_probes", C.c_uint32),
                ("use_graph", C.c_uint32), ("progress", PROGRESS_FN), ("progress_ctx", C.c_void_p),
                ("progress_steps", C.c_uint32), ("reserved", C.c_uint32), ("meas", C.c_void_p),
                ("n_meas", C.c_uint32), ("meas_stop", C.c_uint32), ("meas_out", C.c_void_p),
                ("steps_done", C.c_void_p)]


# .meas cards (esl_meas_card): the subset of src/output/measure.zig that the
# transient lanes measure as they step
MEAS = {"when": 1, "trig_targ": 2, "find": 3, "avg": 4, "min": 5, "max": 6, "pp": 7, "integ": 8, "rms": 9}
MEAS_UNSET, MEAS_LAST, MEAS_NO_AT = -1, -2, 1e99


class MeasClause(C.Structure):
    _fields_ = [("node", C.c_uint32), ("level_node", C.c_uint32), ("has_level_node", C.c_uint32),
                ("reserved", C.c_uint32), ("val", C.c_double), ("td", C.c_double), ("frm", C.c_double),
                ("to", C.c_double), ("at", C.c_double), ("rise", C.c_int32), ("fall", C.c_int32),
                ("cross", C.c_int32), ("reserved2", C.c_int32)]


class MeasCard(C.Structure):
    _fields_ = [("func", C.c_uint32), ("reserved", C.c_uint32 * 3), ("a", MeasClause), ("b", MeasClause)]


def meas_clause(node=0, val=1e99, td=0.0, frm=0.0, to=0.0, at=MEAS_NO_AT, rise=MEAS_UNSET, fall=MEAS_UNSET,
                cross=MEAS_UNSET, level_node=None) -> MeasClause:
    """A clause with the reference's defaults (core/deck.zig:311-340)."""
    return MeasClause(node, 0 if level_node is None else level_node, 0 if level_node is None else 1, 0, val, td,
                      frm, to, at, rise, fall, cross, 0)


def meas_card(func: str, a: MeasClause, b: MeasClause | None = None) -> MeasCard:
    c = MeasCard()
    c.func = MEAS[func]
    c.a = a
    c.b = b if b is not None else meas_clause()
    return c


class TranAdaptiveOptions(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("tol", Tolerances), ("t_stop", C.c_double), ("t_step", C.c_double),
                ("dt_max", C.c_double), ("dt_min", C.c_double), ("trtol", C.c_double), ("chgtol", C.c_double),
                ("max_steps", C.c_uint32), ("attempts_per_launch", C.c_uint32), ("probes", C.c_void_p),
                ("n_probes", C.c_uint32), ("reserved", C.c_uint32)]


class BatchItem(C.Structure):
    _fields_ = [("engine", C.c_void_p), ("values", C.c_void_p), ("count", C.c_size_t), ("analysis", C.c_uint32),
                ("flags", C.c_uint32), ("tol", Tolerances), ("tran", C.POINTER(TranOptions)),
                ("x", C.c_void_p), ("probe_out", C.c_void_p), ("iterations", C.c_void_p), ("reasons", C.c_void_p)]


class BatchOptions(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("tiling", C.c_uint32), ("tune", C.c_uint32),
                ("reserved", C.c_uint32 * 6)]


class BatchReport(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("kernel_ms", C.c_double), ("repivot_ms", C.c_double),
                ("setup_ms", C.c_double), ("total_ms", C.c_double), ("predicted_ms", C.c_double),
                ("lead_ms", C.c_double),
                ("grid", C.c_uint32), ("block_jobs", C.c_uint32), ("wave_jobs", C.c_uint32),
                ("repivots", C.c_uint32), ("lds_bytes", C.c_uint32), ("repivot_launches", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class MlpTrainOptions(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("epochs", C.c_uint32), ("batch", C.c_uint32),
                ("lr", C.c_float), ("lr_final", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("weight_decay", C.c_float), ("seed", C.c_uint64)]


_lib = None


def load() -> C.CDLL:
    """Loads the built library; raises when it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, u32p = C.c_void_p, C.POINTER(C.c_uint32)
    L.esl_abi_version.restype = C.c_uint32
    L.esl_default_op_options.argtypes = [C.POINTER(OpOptions)]
    L.esl_engine_create.restype = C.c_uint32
    L.esl_engine_create.argtypes = [C.POINTER(_Deck), C.POINTER(vp), C.c_char_p, C.c_size_t]
    L.esl_engine_destroy.restype = C.c_uint32
    L.esl_engine_destroy.argtypes = [vp]
    L.esl_engine_info_get.restype = C.c_uint32
    L.esl_engine_info_get.argtypes = [vp, C.POINTER(EngineInfo)]
    L.esl_engine_pattern.restype = C.c_uint32
    L.esl_engine_pattern.argtypes = [vp, vp, vp, vp]
    L.esl_op_solve.restype = C.c_uint32
    L.esl_op_solve.argtypes = [vp, vp, C.c_size_t, C.POINTER(OpOptions), vp, vp, vp]
    L.esl_last_timing.restype = C.c_uint32
    L.esl_last_timing.argtypes = [vp, C.POINTER(Timing)]
    L.esl_tape_create.restype = C.c_uint32
    L.esl_debug_big_replay.argtypes = [C.c_uint32, vp, vp, vp, vp, C.c_double, C.c_uint32, vp, vp, C.c_uint32,
                                       vp, vp, vp, vp, vp, vp]
    L.esl_debug_big_replay.restype = C.c_int
    L.esl_debug_row_replay.argtypes = [C.c_uint32, vp, vp, vp, vp, C.c_double, C.c_uint32, vp, vp, C.c_uint32,
                                       vp, vp, vp, vp, vp, vp]
    L.esl_debug_row_replay.restype = C.c_int
    L.esl_tape_create.argtypes = [C.c_uint32, vp, vp, vp, vp, C.c_double, C.POINTER(vp)]
    L.esl_tape_destroy.restype = C.c_uint32
    L.esl_tape_destroy.argtypes = [vp]
    L.esl_tape_sizes.restype = C.c_uint32
    L.esl_tape_sizes.argtypes = [vp, u32p, u32p, C.POINTER(C.c_uint64), u32p]
    L.esl_tape_export.restype = C.c_uint32
    L.esl_tape_export.argtypes = [vp] * 8
    L.esl_tape_replay.restype = C.c_uint32
    L.esl_tape_replay.argtypes = [vp, vp, vp, C.c_size_t, C.c_double, vp, vp, vp, vp, vp]
    L.esl_tape_replay_noalloc.restype = C.c_uint32
    L.esl_tape_replay_noalloc.argtypes = [vp, vp, vp, C.c_size_t, C.c_double, vp, vp, vp, vp, vp]
    L.esl_mlp_default_train_options.argtypes = [C.POINTER(MlpTrainOptions)]
    L.esl_mlp_create.restype = C.c_uint32
    L.esl_mlp_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(vp)]
    L.esl_mlp_destroy.restype = C.c_uint32
    L.esl_mlp_destroy.argtypes = [vp]
    L.esl_mlp_param_count.restype = C.c_uint32
    L.esl_mlp_param_count.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.esl_mlp_get_params.restype = C.c_uint32
    L.esl_mlp_get_params.argtypes = [vp, vp, C.c_size_t]
    L.esl_mlp_set_params.restype = C.c_uint32
    L.esl_mlp_set_params.argtypes = [vp, vp, C.c_size_t]
    L.esl_mlp_set_scaling.restype = C.c_uint32
    L.esl_mlp_set_scaling.argtypes = [vp, vp, vp, vp, vp]
    L.esl_mlp_gradient.restype = C.c_uint32
    L.esl_mlp_gradient.argtypes = [vp, vp, vp, C.c_size_t, C.POINTER(C.c_float), vp]
    L.esl_mlp_train.restype = C.c_uint32
    L.esl_mlp_train.argtypes = [vp, vp, vp, C.c_size_t, C.POINTER(MlpTrainOptions), vp]
    L.esl_mlp_predict.restype = C.c_uint32
    L.esl_mlp_predict.argtypes = [vp, vp, C.c_size_t, vp]
    L.esl_default_tran_options.argtypes = [C.POINTER(TranOptions)]
    L.esl_tran_solve.restype = C.c_uint32
    L.esl_tran_solve.argtypes = [vp, vp, C.c_size_t, C.POINTER(TranOptions), vp, vp, vp, vp]
    L.esl_default_tran_adaptive_options.argtypes = [C.POINTER(TranAdaptiveOptions)]
    L.esl_tran_adaptive_solve.restype = C.c_uint32
    L.esl_tran_adaptive_solve.argtypes = [vp, vp, C.c_size_t, C.POINTER(TranAdaptiveOptions), vp, vp, vp, vp, vp,
                                          vp, vp]
    L.esl_engine_set_warm_start.restype = C.c_uint32
    L.esl_engine_set_warm_start.argtypes = [vp, vp, vp]
    L.esl_error_message.restype = C.c_uint32
    L.esl_error_message.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p]
    L.esl_default_batch_options.argtypes = [C.POINTER(BatchOptions)]
    L.esl_batch_solve.restype = C.c_uint32
    L.esl_batch_solve.argtypes = [C.POINTER(BatchItem), C.c_size_t, C.POINTER(BatchOptions), C.POINTER(BatchReport)]
    _lib = L
    return L


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def last_error() -> str:
    """The calling thread's last failure text from the library."""
    buf = C.create_string_buffer(512)
    load().esl_error_message(buf, 512, None)
    return buf.value.decode(errors="replace")


def _check(st: int, what: str):
    if st != OK:
        msg = last_error()
        raise EslError(st, f"{what} ({msg})" if msg else what)


def default_options(precision=FP64, start=START_COLD) -> OpOptions:
    o = OpOptions()
    load().esl_default_op_options(C.byref(o))
    o.precision = precision
    o.start = start
    return o


class Engine:
    """A compiled deck on the GPU: solves operating points of many points."""

    def __init__(self, arrays: dict):
        L = load()
        self._a = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        a = self._a
        d = _Deck()
        d.n_nodes = a["n_nodes"]
        d.n_dev = len(a["kind"])
        d.kind, d.nodes, d.params = _ptr(a["kind"]), _ptr(a["nodes"]), _ptr(a["params"])
        d.n_live = a["n_live"]
        d.n_bind = len(a["bind_dev"])
        d.bind_dev, d.bind_param, d.bind_live = _ptr(a["bind_dev"]), _ptr(a["bind_param"]), _ptr(a["bind_live"])
        h = C.c_void_p()
        diag = C.create_string_buffer(256)
        st = L.esl_engine_create(C.byref(d), C.byref(h), diag, 256)
        if st != OK:
            raise EslError(st, f"esl_engine_create ({diag.value.decode()})")
        self.h = h
        info = self.info()
        self.n, self.nnz, self.n_live = info.n, info.nnz, info.n_live

    def close(self):
        if getattr(self, "h", None):
            load().esl_engine_destroy(self.h)
            self.h = None

    __del__ = close

    def info(self) -> EngineInfo:
        i = EngineInfo()
        _check(load().esl_engine_info_get(self.h, C.byref(i)), "esl_engine_info_get")
        return i

    def pattern(self):
        info = self.info()
        cp = np.zeros(info.n + 1, np.uint32)
        ri = np.zeros(max(info.nnz, 1), np.uint32)
        q = np.zeros(info.n, np.uint32)
        _check(load().esl_engine_pattern(self.h, _ptr(cp), _ptr(ri), _ptr(q)), "esl_engine_pattern")
        return cp, ri[: info.nnz], q

    def solve_op(self, values, x0=None, precision=FP64, tol: Tolerances | None = None, tiling=TILING_AUTO,
                 refine_tol=None, refine_max=None, learned=False, probes=None, host_repivot=False,
                 ladder=False):
        """Operating points of rows `values` [count][n_live]. Returns dict
        with x [count][n], iterations, reasons. x0 [count][n]: given starts;
        learned=True: starts from the attached network (set_warm_start);
        probes: unknown indices to return instead of all of x; ladder=True:
        points the plain rung leaves unconverged climb the .op ladder's gmin
        and source-stepping rungs (FP64)."""
        values = np.ascontiguousarray(np.atleast_2d(np.asarray(values, dtype=np.float64)))
        count = values.shape[0]
        start = START_LEARNED if learned else (START_COLD if x0 is None else START_GIVEN)
        o = default_options(precision, start)
        o.tiling = tiling
        o.flags = (OP_HOST_REPIVOT if host_repivot else 0) | (OP_LADDER if ladder else 0)
        if refine_tol is not None:
            o.refine_tol = refine_tol
        if refine_max is not None:
            o.refine_max = refine_max
        if tol is not None:
            o.tol = tol
        xs = None
        if x0 is not None and not learned:
            xs = np.ascontiguousarray(x0, dtype=np.float64).reshape(count, self.n)
            o.x0 = xs.ctypes.data
        pr = None
        if probes is not None:
            pr = np.ascontiguousarray(probes, dtype=np.uint32)
            o.probes, o.n_probes = pr.ctypes.data, len(pr)
        x = np.zeros((count, self.n if pr is None else len(pr)))
        it = np.zeros(count, np.uint32)
        rs = np.zeros(count, np.uint32)
        _check(load().esl_op_solve(self.h, _ptr(values), count, C.byref(o), _ptr(x), _ptr(it), _ptr(rs)),
               "esl_op_solve")
        return {"x": x, "iterations": it, "reasons": rs}

    def solve_tran(self, values, dt, steps, probes=(), steps_per_launch=10, use_graph=True, max_iter=None,
                   progress=None, progress_steps=0, meas=None, meas_stop=False):
        """Transient lanes: operating point, then `steps` fixed steps of dt.
        Returns x [count][n] (last state), probes [count][steps][n_probes],
        iterations, reasons. progress(steps_done, steps), when given, is
        called after each group of about progress_steps steps; a true return
        cancels the solve (EslError). meas: MeasCard list measured on the
        device as the steps are accepted ("meas" [count][n_meas],
        "steps_done" [count]); meas_stop: a point stops once its cards
        resolve."""
        values = np.ascontiguousarray(np.atleast_2d(np.asarray(values, dtype=np.float64)))
        count = values.shape[0]
        o = TranOptions()
        load().esl_default_tran_options(C.byref(o))
        o.dt, o.steps, o.steps_per_launch, o.use_graph = dt, steps, steps_per_launch, 1 if use_graph else 0
        if max_iter is not None:
            o.tol.max_iter = max_iter
        if progress is not None:
            cb = PROGRESS_FN(lambda ctx, done, total: 1 if progress(int(done), int(total)) else 0)
            o.progress, o.progress_steps = cb, progress_steps
        pr = np.ascontiguousarray(probes, dtype=np.uint32)
        if len(pr):
            o.probes, o.n_probes = pr.ctypes.data, len(pr)
        cards, mo, sd = None, None, np.zeros(count, np.uint32)
        if meas:
            cards = (MeasCard * len(meas))(*meas)
            mo = np.zeros((count, len(meas)))
            o.meas, o.n_meas, o.meas_stop = C.cast(cards, C.c_void_p), len(meas), 1 if meas_stop else 0
            o.meas_out, o.steps_done = mo.ctypes.data, sd.ctypes.data
        x = np.zeros((count, self.n))
        po = np.zeros((count, steps, max(len(pr), 1)))
        it = np.zeros(count, np.uint32)
        rs = np.zeros(count, np.uint32)
        _check(load().esl_tran_solve(self.h, _ptr(values), count, C.byref(o), _ptr(x),
                                     _ptr(po) if len(pr) else None, _ptr(it), _ptr(rs)), "esl_tran_solve")
        out = {"x": x, "probes": po[:, :, :len(pr)], "iterations": it, "reasons": rs}
        if meas:
            out.update(meas=mo, steps_done=sd)
        return out

    def solve_tran_adaptive(self, values, t_stop, t_step, probes=(), dt_max=0.0, dt_min=1e-18, trtol=7.0,
                            chgtol=1e-14, max_steps=1000, attempts_per_launch=64, max_iter=None):
        """Adaptive transient lanes (esl_tran_adaptive_solve): each point's
        operating point, then steps the lane chooses up to t_stop. Returns
        x [count][n] (last accepted state), times [count][max_steps],
        probes [count][max_steps][n_probes] (rows past `accepted` unused),
        accepted, rejected, iterations, reasons."""
        values = np.ascontiguousarray(np.atleast_2d(np.asarray(values, dtype=np.float64)))
        count = values.shape[0]
        o = TranAdaptiveOptions()
        load().esl_default_tran_adaptive_options(C.byref(o))
        o.t_stop, o.t_step, o.dt_max, o.dt_min = t_stop, t_step, dt_max, dt_min
        o.trtol, o.chgtol, o.max_steps, o.attempts_per_launch = trtol, chgtol, max_steps, attempts_per_launch
        if max_iter is not None:
            o.tol.max_iter = max_iter
        pr = np.ascontiguousarray(probes, dtype=np.uint32)
        if len(pr):
            o.probes, o.n_probes = pr.ctypes.data, len(pr)
        x = np.zeros((count, self.n))
        tm = np.zeros((count, max_steps))
        po = np.zeros((count, max_steps, max(len(pr), 1)))
        acc = np.zeros(count, np.uint32)
        rej = np.zeros(count, np.uint32)
        it = np.zeros(count, np.uint32)
        rs = np.zeros(count, np.uint32)
        _check(load().esl_tran_adaptive_solve(self.h, _ptr(values), count, C.byref(o), _ptr(x), _ptr(tm),
                                              _ptr(po) if len(pr) else None, _ptr(acc), _ptr(rej), _ptr(it),
                                              _ptr(rs)), "esl_tran_adaptive_solve")
        return {"x": x, "times": tm, "probes": po[:, :, :len(pr)], "accepted": acc, "rejected": rej,
                "iterations": it, "reasons": rs}

    def set_warm_start(self, net: "Mlp | None", x_nominal=None):
        """Attaches the warm-start network and the nominal operating point."""
        if net is None:
            _check(load().esl_engine_set_warm_start(self.h, None, None), "esl_engine_set_warm_start")
            self._warm = None
            return
        xn = np.ascontiguousarray(x_nominal, dtype=np.float64).reshape(self.n)
        _check(load().esl_engine_set_warm_start(self.h, net.h, _ptr(xn)), "esl_engine_set_warm_start")
        self._warm = net   # keep the network alive while attached

    def timing(self) -> Timing:
        t = Timing()
        _check(load().esl_last_timing(self.h, C.byref(t)), "esl_last_timing")
        return t


def solve_batch(jobs, tiling=TILING_AUTO, tune=False):
    """One persistent launch over several templates (esl_batch_solve).

    jobs: list of dicts {"engine": Engine, "values": [count][n_live],
    "analysis": BATCH_OP | BATCH_TRAN, and for BATCH_TRAN "dt", "steps",
    optional "probes", "max_iter"; for BATCH_OP optional "tol" and
    "ladder" (True: unconverged points climb the .op ladder)}. Returns
    (results, report): per job a dict like Engine.solve_op / solve_tran."""
    L = load()
    n = len(jobs)
    items = (BatchItem * n)()
    keep, out = [], []
    for k, j in enumerate(jobs):
        eng = j["engine"]
        values = np.ascontiguousarray(np.atleast_2d(np.asarray(j["values"], dtype=np.float64)))
        count = values.shape[0]
        it = items[k]
        it.engine, it.values, it.count = eng.h, _ptr(values), count
        it.analysis = j.get("analysis", BATCH_OP)
        it.flags = OP_LADDER if j.get("ladder") else 0
        t = Tolerances()
        o = OpOptions()
        L.esl_default_op_options(C.byref(o))
        t = j.get("tol", o.tol)
        it.tol = t
        res = {"x": np.zeros((count, eng.n)), "iterations": np.zeros(count, np.uint32),
               "reasons": np.zeros(count, np.uint32)}
        keep.append(values)
        if it.analysis == BATCH_TRAN:
            to = TranOptions()
            L.esl_default_tran_options(C.byref(to))
            to.dt, to.steps = j["dt"], j["steps"]
            if j.get("max_iter") is not None:
                to.tol.max_iter = j["max_iter"]
            pr = np.ascontiguousarray(j.get("probes", ()), dtype=np.uint32)
            if len(pr):
                to.probes, to.n_probes = pr.ctypes.data, len(pr)
            po = np.zeros((count, j["steps"], max(len(pr), 1)))
            res["probes"] = po
            res["_np"] = len(pr)
            keep += [to, pr, po]
            it.tran = C.pointer(to)
            it.probe_out = _ptr(po) if len(pr) else None
        it.x, it.iterations, it.reasons = _ptr(res["x"]), _ptr(res["iterations"]), _ptr(res["reasons"])
        out.append(res)
    bo = BatchOptions()
    L.esl_default_batch_options(C.byref(bo))
    bo.tiling, bo.tune = tiling, 1 if tune else 0
    rep = BatchReport()
    rep.struct_size = C.sizeof(BatchReport)
    _check(L.esl_batch_solve(items, n, C.byref(bo), C.byref(rep)), "esl_batch_solve")
    for res in out:
        if "probes" in res:
            res["probes"] = res["probes"][:, :, :res.pop("_np")]
    return out, rep


def debug_big_replay(col_ptr, row_idx, lead_vals, q, a, rhs, block_slots=4032, pivot_tol=1e-3):
    """The big-lane kernel's blocked refactor and chunked solves run on the
    CPU (esl_debug_big_replay) for a [lanes][nnz], rhs [lanes][n]: a
    development check of the block programs, no device needed."""
    L = load()
    cp, ri, qq = (np.ascontiguousarray(v, dtype=np.uint32) for v in (col_ptr, row_idx, q))
    lead = np.ascontiguousarray(lead_vals, dtype=np.float64)
    a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64)
    rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.float64)
    n, lanes = len(cp) - 1, a.shape[0]
    sizes = np.zeros(4, np.uint32)
    args = (n, _ptr(cp), _ptr(ri), _ptr(lead), _ptr(qq), pivot_tol, block_slots)
    _check(L.esl_debug_big_replay(*args, None, None, 0, None, None, None, None, None, _ptr(sizes)), "big replay")
    nl, nu = int(sizes[1]), int(sizes[2])
    out = dict(lx=np.zeros((lanes, nl)), ux=np.zeros((lanes, nu)), udiag=np.zeros((lanes, n)),
               dx=np.zeros((lanes, n)), fail=np.zeros(lanes, np.uint32))
    _check(L.esl_debug_big_replay(*args, _ptr(a), _ptr(rhs), lanes, _ptr(out["lx"]), _ptr(out["ux"]),
                                  _ptr(out["udiag"]), _ptr(out["dx"]), _ptr(out["fail"]), _ptr(sizes)),
           "big replay")
    out["blocks"], out["rows"] = int(sizes[0]), int(sizes[3])
    return out


def debug_row_replay(col_ptr, row_idx, lead_vals, q, a, rhs, min_chain=8, pivot_tol=1e-3):
    """The wavefront kernels' refactor and solve row programs (chain rows
    for runs of at least min_chain terms) run on the CPU
    (esl_debug_row_replay) for a [lanes][nnz], rhs [lanes][n], each row
    checked for independent ops: a development check of the row programs,
    no device needed."""
    L = load()
    cp, ri, qq = (np.ascontiguousarray(v, dtype=np.uint32) for v in (col_ptr, row_idx, q))
    lead = np.ascontiguousarray(lead_vals, dtype=np.float64)
    a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64)
    rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.float64)
    n, lanes = len(cp) - 1, a.shape[0]
    sizes = np.zeros(10, np.uint32)
    args = (n, _ptr(cp), _ptr(ri), _ptr(lead), _ptr(qq), pivot_tol, min_chain)
    _check(L.esl_debug_row_replay(*args, None, None, 0, None, None, None, None, None, _ptr(sizes)), "row replay")
    nl, nu = int(sizes[1]), int(sizes[2])
    out = dict(lx=np.zeros((lanes, nl)), ux=np.zeros((lanes, nu)), udiag=np.zeros((lanes, n)),
               dx=np.zeros((lanes, n)), fail=np.zeros(lanes, np.uint32))
    _check(L.esl_debug_row_replay(*args, _ptr(a), _ptr(rhs), lanes, _ptr(out["lx"]), _ptr(out["ux"]),
                                  _ptr(out["udiag"]), _ptr(out["dx"]), _ptr(out["fail"]), _ptr(sizes)),
           "row replay")
    keys = ("refactor_rows", None, None, "solve_rows", "refactor_chains", "solve_chains", "refactor_chain_blocks",
            "solve_chain_blocks", "refactor_critical", "solve_critical")
    out.update({k: int(v) for k, v in zip(keys, sizes) if k})
    return out


class Tape:
    """A frozen pivot tape replayed by many lanes on the GPU."""

    def __init__(self, col_ptr, row_idx, lead_vals, q, pivot_tol=1e-3):
        L = load()
        self._cp = np.ascontiguousarray(col_ptr, dtype=np.uint32)
        self._ri = np.ascontiguousarray(row_idx, dtype=np.uint32)
        lv = np.ascontiguousarray(lead_vals, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.uint32)
        self.n = len(self._cp) - 1
        self.nnz = int(self._cp[-1])
        h = C.c_void_p()
        _check(L.esl_tape_create(self.n, _ptr(self._cp), _ptr(self._ri), _ptr(lv), _ptr(q), pivot_tol,
                                 C.byref(h)), "esl_tape_create")
        self.h = h
        nl, nu, lev = C.c_uint32(), C.c_uint32(), C.c_uint32()
        fl = C.c_uint64()
        _check(L.esl_tape_sizes(self.h, C.byref(nl), C.byref(nu), C.byref(fl), C.byref(lev)), "esl_tape_sizes")
        self.nnz_l, self.nnz_u, self.flops, self.levels = nl.value, nu.value, fl.value, lev.value

    def close(self):
        if getattr(self, "h", None):
            load().esl_tape_destroy(self.h)
            self.h = None

    __del__ = close

    def export(self) -> dict:
        n, nnz = self.n, self.nnz
        out = {k: np.zeros(s, np.uint32) for k, s in [("pinv", n), ("lp", n + 1), ("li", max(self.nnz_l, 1)),
                                                      ("up", n + 1), ("ui", max(self.nnz_u, 1)),
                                                      ("flags", n), ("prow", max(nnz, 1))]}
        _check(load().esl_tape_export(self.h, *[_ptr(out[k]) for k in
                                               ["pinv", "lp", "li", "up", "ui", "flags", "prow"]]),
               "esl_tape_export")
        out["li"] = out["li"][: self.nnz_l]
        out["ui"] = out["ui"][: self.nnz_u]
        out["prow"] = out["prow"][:nnz]
        return out

    def replay(self, a, rhs, growth=1e-12, noalloc=False) -> dict:
        """Replays the tape over a [lanes][nnz] and solves -rhs [lanes][n].
        noalloc: through esl_tape_replay_noalloc (the library's fixed
        scratch; no device allocation in the call)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        lanes = a.shape[0]
        lx = np.zeros((lanes, max(self.nnz_l, 1)))
        ux = np.zeros((lanes, max(self.nnz_u, 1)))
        ud = np.zeros((lanes, self.n))
        dx = np.zeros((lanes, self.n))
        fail = np.zeros(lanes, np.uint32)
        fn = load().esl_tape_replay_noalloc if noalloc else load().esl_tape_replay
        _check(fn(self.h, _ptr(a), _ptr(rhs), lanes, growth, _ptr(lx), _ptr(ux), _ptr(ud), _ptr(dx), _ptr(fail)),
               "esl_tape_replay_noalloc" if noalloc else "esl_tape_replay")
        return {"lx": lx[:, : self.nnz_l], "ux": ux[:, : self.nnz_u], "udiag": ud, "dx": dx, "fail": fail}


def default_train_options(**kw) -> MlpTrainOptions:
    o = MlpTrainOptions()
    load().esl_mlp_default_train_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Mlp:
    """The warm-start network on the GPU (BF16 MFMA, FP32 master weights)."""

    def __init__(self, n_in, n_out, n_hidden=256, n_hidden_layers=3, seed=1):
        h = C.c_void_p()
        _check(load().esl_mlp_create(n_in, n_hidden, n_hidden_layers, n_out, seed, C.byref(h)), "esl_mlp_create")
        self.h = h
        self.n_in, self.n_out = n_in, n_out
        self.dims = [n_in] + [n_hidden] * n_hidden_layers + [n_out]
        c = C.c_size_t()
        _check(load().esl_mlp_param_count(self.h, C.byref(c)), "esl_mlp_param_count")
        self.n_params = c.value

    def close(self):
        if getattr(self, "h", None):
            load().esl_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(load().esl_mlp_get_params(self.h, _ptr(p), p.size), "esl_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(load().esl_mlp_set_params(self.h, _ptr(p), p.size), "esl_mlp_set_params")

    def set_scaling(self, in_mean, in_scale, out_mean, out_scale):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (in_mean, in_scale, out_mean, out_scale)]
        _check(load().esl_mlp_set_scaling(self.h, *[_ptr(v) for v in a]), "esl_mlp_set_scaling")

    def gradient(self, inputs, targets):
        inputs = np.ascontiguousarray(inputs, dtype=np.float64)
        targets = np.ascontiguousarray(targets, dtype=np.float64)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        _check(load().esl_mlp_gradient(self.h, _ptr(inputs), _ptr(targets), inputs.shape[0], C.byref(loss),
                                       _ptr(g)), "esl_mlp_gradient")
        return float(loss.value), g

    def train(self, inputs, targets, **kw):
        inputs = np.ascontiguousarray(inputs, dtype=np.float64)
        targets = np.ascontiguousarray(targets, dtype=np.float64)
        o = default_train_options(**kw)
        losses = np.zeros(o.epochs, np.float32)
        _check(load().esl_mlp_train(self.h, _ptr(inputs), _ptr(targets), inputs.shape[0], C.byref(o),
                                    _ptr(losses)), "esl_mlp_train")
        return losses

    def predict(self, inputs):
        inputs = np.ascontiguousarray(np.atleast_2d(inputs), dtype=np.float64)
        out = np.zeros((inputs.shape[0], self.n_out + 1))
        _check(load().esl_mlp_predict(self.h, _ptr(inputs), inputs.shape[0], _ptr(out)), "esl_mlp_predict")
        return out

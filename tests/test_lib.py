"""The C-ABI library loads and exports every entry point include/espice_lanes.h
declares (no compute calls here: they need a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "espice_lanes.h")).read()
    return sorted(set(re.findall(r"\b(esl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ["esl_engine_create", "esl_op_solve", "esl_tape_create", "esl_tape_replay", "esl_engine_destroy"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from espice_amd import lanes
    lib = lanes.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(lanes.EXPORTS) <= set(declared())


def test_abi_version_and_defaults():
    from espice_amd import lanes
    lib = lanes.load()
    assert lib.esl_abi_version() == 5
    o = lanes.default_options()
    assert o.struct_size == ctypes.sizeof(lanes.OpOptions)
    assert o.tol.reltol == 1e-3 and o.tol.abstol == 1e-12 and o.tol.vntol == 1e-6
    assert o.tol.residual_tol == 1e-9 and o.tol.max_iter == 100
    t = lanes.TranOptions()
    lib.esl_default_tran_options(ctypes.byref(t))
    assert t.struct_size == ctypes.sizeof(lanes.TranOptions)
    assert t.steps_per_launch == 10 and t.use_graph == 1 and t.tol.max_iter == 50
    assert not t.progress and t.progress_ctx is None and t.progress_steps == 0
    assert t.n_meas == 0 and not t.meas and t.meas_stop == 0 and not t.meas_out and not t.steps_done
    # the .meas card has the size the library asserts for esl_meas_card
    assert ctypes.sizeof(lanes.MeasCard) == 160 and ctypes.sizeof(lanes.MeasClause) == 72


def test_invalid_handles_are_rejected():
    from espice_amd import lanes
    lib = lanes.load()
    info = lanes.EngineInfo()
    assert lib.esl_engine_info_get(ctypes.c_void_p(12345), ctypes.byref(info)) == lanes.INVALID_HANDLE
    assert lib.esl_engine_destroy(ctypes.c_void_p(12345)) == lanes.INVALID_HANDLE
    assert lib.esl_engine_destroy(None) == lanes.OK


def test_invalid_network_handles_are_rejected():
    from espice_amd import lanes
    lib = lanes.load()
    c = ctypes.c_size_t()
    assert lib.esl_mlp_param_count(ctypes.c_void_p(777), ctypes.byref(c)) == lanes.INVALID_HANDLE
    assert lib.esl_mlp_destroy(ctypes.c_void_p(777)) == lanes.INVALID_HANDLE
    o = lanes.default_train_options()
    assert o.struct_size == ctypes.sizeof(lanes.MlpTrainOptions) and o.epochs > 0 and o.batch > 0


def test_batch_defaults_and_argument_checks():
    """Defaults of the batch options, and the checks that run before any
    device work (no GPU needed)."""
    from espice_amd import lanes
    lib = lanes.load()
    o = lanes.BatchOptions()
    lib.esl_default_batch_options(ctypes.byref(o))
    assert o.struct_size == ctypes.sizeof(lanes.BatchOptions) and o.tiling == lanes.TILING_AUTO and o.tune == 0
    rep = lanes.BatchReport()
    rep.struct_size = ctypes.sizeof(lanes.BatchReport)
    assert lib.esl_batch_solve(None, 1, ctypes.byref(o), ctypes.byref(rep)) == lanes.INVALID_ARGUMENT
    items = (lanes.BatchItem * 1)()
    assert lib.esl_batch_solve(items, 0, ctypes.byref(o), ctypes.byref(rep)) == lanes.INVALID_ARGUMENT
    bad = lanes.BatchOptions()
    bad.struct_size = 8
    assert lib.esl_batch_solve(items, 1, ctypes.byref(bad), ctypes.byref(rep)) == lanes.ABI_MISMATCH
    # a null engine handle
    assert lib.esl_batch_solve(items, 1, ctypes.byref(o), ctypes.byref(rep)) == lanes.INVALID_HANDLE


def test_error_message_reports_the_last_failure():
    from espice_amd import lanes
    lib = lanes.load()
    need = ctypes.c_size_t(0)
    assert lib.esl_error_message(None, 0, ctypes.byref(need)) == lanes.OK and need.value >= 1
    assert isinstance(lanes.last_error(), str)

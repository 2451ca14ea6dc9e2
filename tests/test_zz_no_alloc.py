"""esl_tape_replay_noalloc (the replay without device allocation) against
esl_tape_replay, through the C ABI: bitwise equal factors, solutions and peel
flags on two calls in a row, the second with fewer lanes than the first; a
call beyond the fixed scratch is refused before anything runs."""
import ctypes
import sys

import numpy as np
import pytest

from test_oracle_lu import random_system

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def lane_matrices(rng, vals, n, lanes_n):
    a = vals[None, :] * (0.8 + 0.4 * rng.random((lanes_n, len(vals))))
    a[3] = vals
    a[5, 0] = np.nan   # a lane that must re-pivot
    rhs = 2 * rng.random((lanes_n, n)) - 1
    return a, rhs


def test_noalloc_replay_equals_the_allocating_replay():
    from espice_amd import lanes
    rng = np.random.default_rng(29)
    n = 180
    while True:
        cp, ri, vals = random_system(rng, n, 0.03)
        q = rng.permutation(n).astype(np.uint32)
        try:
            tape = lanes.Tape(cp, ri, vals, q)
            break
        except lanes.EslError:
            continue
    per_lane = 8 * (len(vals) + 3 * n + tape.nnz_l + tape.nnz_u) + 4
    say(f"noalloc replay n={n} nnz={len(vals)} nnz_l={tape.nnz_l} nnz_u={tape.nnz_u} bytes/lane={per_lane}")
    first = min(4096, (24 << 20) // per_lane)
    for lanes_n in (first, first // 3 + 7):
        a, rhs = lane_matrices(rng, vals, n, lanes_n)
        say(f"  call with {lanes_n} lanes")
        ref = tape.replay(a, rhs)
        got = tape.replay(a, rhs, noalloc=True)
        assert np.array_equal(got["fail"], ref["fail"])
        assert ref["fail"][5] == 1
        for k in ("lx", "ux", "udiag", "dx"):
            assert np.array_equal(bits(got[k]), bits(ref[k])), (lanes_n, k)
    tape.close()


def test_noalloc_replay_refuses_more_than_its_scratch():
    from espice_amd import lanes
    L = lanes.load()
    rng = np.random.default_rng(31)
    n = 40
    cp, ri, vals = random_system(rng, n, 0.1)
    q = np.arange(n, dtype=np.uint32)
    tape = lanes.Tape(cp, ri, vals, q)
    per_lane = 8 * (len(vals) + 3 * n + tape.nnz_l + tape.nnz_u) + 4
    lanes_n = (40 << 20) // per_lane          # past the 32 MiB scratch
    a = np.repeat(vals[None, :], lanes_n, 0)
    rhs = np.ones((lanes_n, n))
    out = np.zeros((lanes_n, n))
    st = L.esl_tape_replay_noalloc(tape.h, a.ctypes.data, rhs.ctypes.data, lanes_n, 1e-12, None, None, None,
                                   out.ctypes.data, None)
    assert st == lanes.BUFFER_TOO_SMALL
    assert "scratch" in lanes.last_error()
    # the scratch is untouched by the refusal: a call that fits still works
    got = tape.replay(a[:100], rhs[:100], noalloc=True)
    ref = tape.replay(a[:100], rhs[:100])
    assert np.array_equal(bits(got["dx"]), bits(ref["dx"]))
    tape.close()

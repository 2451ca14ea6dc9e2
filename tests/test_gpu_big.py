"""Lanes beyond a workgroup's LDS: the 64 x 64 RC mesh with diode clamps
(4,099 unknowns; its factors hold about 140,000 slots per lane), run one
thread per lane with every working set in HBM. The arithmetic is the
wavefront kernels' (same stamps, same column replay, same gates), so the
lanes must equal the oracle bit for bit, operating points and transient
probes alike.
"""
import sys

import numpy as np
import pytest

from espice_amd import circuits

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def lanes():
    from espice_amd import lanes as L
    L.load()
    return L


@pytest.fixture(scope="module")
def mesh(lanes):
    fam = circuits.rc_mesh(64)
    eng = lanes.Engine(fam.deck.arrays())
    yield fam, eng
    eng.close()


def test_mesh_operating_points_bitwise(lanes, oracle, mesh):
    fam, eng = mesh
    vals = fam.sample(41, 8)
    _, _, q = eng.pattern()
    say(f"rc mesh 64x64 .op, n {eng.n}, 8 lanes")
    got = eng.solve_op(vals)
    tm = eng.timing()
    assert tm.tiling == lanes.TILING_WAVE   # one wavefront per lane (no LDS kernel fits the mesh)
    ref = oracle.Circuit(fam.deck.arrays()).op(vals, q, threads=8)
    say(f"  iterations {got['iterations']}, kernel {tm.kernel_ms:.1f} ms")
    assert (got["reasons"] == 0).all()
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(bits(got["x"]), bits(ref["x"]))


@pytest.mark.parametrize("graph", [True, False])
def test_mesh_transient_bitwise(lanes, oracle, mesh, graph):
    fam, eng = mesh
    d = fam.deck
    vals = fam.sample(2031, 6)
    _, _, q = eng.pattern()
    probes = [d.node_index("n0_0"), d.node_index("n63_63"), d.node_index("n32_32")]
    steps = 16
    say(f"rc mesh 64x64 transient, 6 lanes, {steps} steps, graph {graph}")
    got = eng.solve_tran(vals, 1e-12, steps, probes=probes, steps_per_launch=5, use_graph=graph)
    tm = eng.timing()
    ref = oracle.tran(oracle.Circuit(d.arrays()), vals, q, 1e-12, steps, probes=probes, threads=8)
    say(f"  iterations/lane {got['iterations'].mean():.1f}, repivots {tm.repivots}, kernel {tm.kernel_ms:.1f} ms")
    assert (ref["reasons"] == 0).all()
    assert np.array_equal(got["reasons"], ref["reasons"])
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(bits(got["probes"]), bits(ref["probes"]))
    assert np.array_equal(bits(got["x"]), bits(ref["x"]))


def test_mesh_rejects_fp32_and_batches(lanes, mesh):
    fam, eng = mesh
    vals = fam.sample(3, 2)
    with pytest.raises(lanes.EslError):
        eng.solve_op(vals, precision=lanes.FP32_REFINED)
    with pytest.raises(lanes.EslError):
        lanes.solve_batch([{"engine": eng, "values": vals}])

"""The persistent uneven batch (esl_batch_solve) on the GPU.

A batch runs every point of several templates, .op and transient mixed, in
one launch, dealt ahead of time to workgroups (lane per block: one lane on
four wavefronts) and to wavefronts (lane per wavefront). The arithmetic of a
lane is the per-engine kernels' own, so each item must equal esl_op_solve /
esl_tran_solve on the same points bit for bit, whatever tiling the job ran
in, and the oracle on the same points.
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


def small_mix():
    """One small template per family, a few dozen points each."""
    return [
        (circuits.inverter_chain(8), "op", None, 40),
        (circuits.inverter_chain_tran(8), "tran", 2e-12, 24),
        (circuits.rc_ladder(24), "op", None, 33),
        (circuits.rc_ladder(36, tran=True), "tran", 10e-12, 20),
        (circuits.diffpair_array(4), "op", None, 37),
        (circuits.diffpair_array_tran(4), "tran", 20e-12, 18),
        (circuits.mirror_tree(3), "op", None, 29),
        (circuits.mirror_tree(3, tran=True), "tran", 20e-12, 21),
    ]


def make_jobs(lanes, mix, steps, seed=5):
    jobs, engines = [], []
    for k, (fam, an, dt, count) in enumerate(mix):
        eng = lanes.Engine(fam.deck.arrays())
        engines.append(eng)
        vals = fam.sample(seed + k, count)
        j = {"engine": eng, "values": vals, "analysis": lanes.BATCH_OP if an == "op" else lanes.BATCH_TRAN}
        if an == "tran":
            j.update(dt=dt, steps=steps, probes=[1, eng.n // 2])
            # the operating points as esl_tran_solve runs them: the transient's
            # tolerances with at least 100 iterations
            t = lanes.default_options().tol
            t.max_iter = 100
            j["tol"] = t
        jobs.append(j)
    return jobs, engines


def per_engine(lanes, jobs):
    """The same points through the per-engine entry points."""
    out = []
    for j in jobs:
        eng = j["engine"]
        if j["analysis"] == lanes.BATCH_OP:
            out.append(eng.solve_op(j["values"], tiling=lanes.TILING_WAVE))
        else:
            out.append(eng.solve_tran(j["values"], j["dt"], j["steps"], probes=j["probes"]))
    return out


def assert_same(got, ref, what):
    assert np.array_equal(got["reasons"], ref["reasons"]), what
    assert np.array_equal(got["iterations"], ref["iterations"]), what
    ok = got["reasons"] == 0
    assert np.array_equal(bits(got["x"][ok]), bits(ref["x"][ok])), what
    if "probes" in ref:
        assert np.array_equal(bits(got["probes"][ok]), bits(ref["probes"][ok])), what


@pytest.mark.parametrize("tiling", ["wave", "block", "auto"])
def test_mixed_batch_equals_per_engine_solves(lanes, tiling):
    mix = small_mix()
    jobs, engines = make_jobs(lanes, mix, steps=30)
    t = {"wave": lanes.TILING_WAVE, "block": lanes.TILING_BLOCK, "auto": lanes.TILING_AUTO}[tiling]
    say(f"batch of {sum(m[3] for m in mix)} points over {len(mix)} templates, tiling {tiling}")
    got, rep = lanes.solve_batch(jobs, tiling=t)
    say(f"  grid {rep.grid} lds {rep.lds_bytes} block jobs {rep.block_jobs} wave jobs {rep.wave_jobs} "
        f"kernel {rep.kernel_ms:.2f} ms repivots {rep.repivots}")
    assert rep.block_jobs + rep.wave_jobs == sum(m[3] for m in mix)
    if tiling == "wave":
        assert rep.block_jobs == 0
    if tiling == "block":
        assert rep.wave_jobs == 0
    ref = per_engine(lanes, jobs)
    for (fam, an, _, _), g, r in zip(mix, got, ref):
        assert (r["reasons"] == 0).mean() > 0.5, fam.name
        assert_same(g, r, f"{fam.name} {an}")
    for e in engines:
        e.close()


def test_batch_operating_points_match_oracle(lanes, oracle):
    mix = [(circuits.diffpair_array(6), "op", None, 40), (circuits.mirror_tree(4), "op", None, 40),
           (circuits.rc_ladder(54), "op", None, 40)]
    jobs, engines = make_jobs(lanes, mix, steps=0, seed=11)
    got, _ = lanes.solve_batch(jobs, tiling=lanes.TILING_BLOCK)
    for (fam, _, _, _), j, g in zip(mix, jobs, got):
        _, _, q = j["engine"].pattern()
        ref = oracle.Circuit(fam.deck.arrays()).op(j["values"], q, threads=8)
        say(f"  {fam.name}: iterations {g['iterations'].min()}..{g['iterations'].max()}")
        assert (g["reasons"] == 0).all()
        assert np.array_equal(g["iterations"], ref["iterations"])
        assert np.array_equal(bits(g["x"]), bits(ref["x"]))
    for e in engines:
        e.close()


def test_batch_transient_matches_oracle(lanes, oracle):
    fam = circuits.rc_ladder(36, tran=True)
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    vals = fam.sample(3, 24)
    dt, steps = 10e-12, 40
    probes = [eng.n - 1, 5]
    got, _ = lanes.solve_batch([{"engine": eng, "values": vals, "analysis": lanes.BATCH_TRAN, "dt": dt,
                                 "steps": steps, "probes": probes}], tiling=lanes.TILING_AUTO, tune=True)
    ref = oracle.tran(oracle.Circuit(fam.deck.arrays()), vals, q, dt, steps, probes=probes, threads=8)
    g = got[0]
    assert (g["reasons"] == 0).all()
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert np.array_equal(bits(g["x"]), bits(ref["x"]))
    assert np.array_equal(bits(g["probes"]), bits(ref["probes"]))
    eng.close()


def test_autotuner_keeps_its_choice_and_results(lanes):
    """The first AUTO batch measures both tilings and settles one per
    template; a second batch runs the settled tilings with the same bits."""
    mix = small_mix()[:4]
    jobs, engines = make_jobs(lanes, mix, steps=10)
    first, r1 = lanes.solve_batch(jobs, tiling=lanes.TILING_AUTO, tune=True)
    second, r2 = lanes.solve_batch(jobs, tiling=lanes.TILING_AUTO)
    say(f"  tuning: block {r1.block_jobs} wave {r1.wave_jobs}; settled: block {r2.block_jobs} wave {r2.wave_jobs}, "
        f"predicted {r2.predicted_ms:.2f} ms, ran {r2.kernel_ms:.2f} ms")
    assert r1.block_jobs > 0 and r1.wave_jobs > 0
    assert r2.predicted_ms > 0
    for a, b in zip(first, second):
        assert_same(a, b, "tuned rerun")
    for e in engines:
        e.close()


def test_batch_rejects_bad_items(lanes):
    fam = circuits.rc_ladder(16)
    eng = lanes.Engine(fam.deck.arrays())
    vals = fam.sample(1, 4)
    with pytest.raises(lanes.EslError) as ex:
        lanes.solve_batch([{"engine": eng, "values": vals}, {"engine": eng, "values": vals}])
    assert ex.value.status == lanes.INVALID_ARGUMENT
    bad = vals.copy()
    bad[1, 0] = np.nan
    with pytest.raises(lanes.EslError):
        lanes.solve_batch([{"engine": eng, "values": bad}])
    with pytest.raises(lanes.EslError):
        lanes.solve_batch([{"engine": eng, "values": vals, "analysis": lanes.BATCH_TRAN, "dt": -1.0, "steps": 3}])
    eng.close()

"""The warm-start network's HIP kernels against the CPU reference
(oracle/mlp_ref.py), and what the trained network gains in Newton.

The device multiplies in BF16 with FP32 accumulation; the reference applies
the same BF16 roundings and accumulates exactly, so the checks compare norms
with stated tolerances. The warm-start test is bitwise: the engine's lanes
started from the network must equal the oracle's Newton from the same
guesses with the same nominal fallback rule.
"""
import sys

import numpy as np
import pytest

from espice_amd import circuits
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lanes():
    from espice_amd import lanes as L
    L.load()
    return L


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def random_net(lanes, rng, n_in, n_out, hidden, layers):
    net = lanes.Mlp(n_in, n_out, n_hidden=hidden, n_hidden_layers=layers, seed=3)
    p = net.params()
    p = (p + 0.05 * rng.standard_normal(p.size)).astype(np.float32)   # nonzero biases too
    net.set_params(p)
    in_mean, in_scale = rng.standard_normal(n_in), 0.5 + rng.random(n_in)
    out_mean, out_scale = rng.standard_normal(n_out), 0.5 + rng.random(n_out)
    net.set_scaling(in_mean, in_scale, out_mean, out_scale)
    return net, p, (in_mean, in_scale, out_mean, out_scale)


@pytest.mark.parametrize("n_in,n_out,hidden,layers,rows", [(32, 20, 64, 2, 37), (32, 323, 256, 3, 600)])
def test_gradient_matches_cpu_reference(lanes, n_in, n_out, hidden, layers, rows):
    rng = np.random.default_rng(rows)
    say(f"gradient check dims {n_in}->{hidden}x{layers}->{n_out}, {rows} rows")
    net, p, (im, isc, om, osc) = random_net(lanes, rng, n_in, n_out, hidden, layers)
    x = im + isc * rng.standard_normal((rows, n_in))
    t = np.concatenate([np.zeros((rows, 1)), om + osc * rng.standard_normal((rows, n_out))], axis=1)
    loss, g = net.gradient(x, t)
    rloss, rg = mlp_ref.loss_and_grads(p, net.dims, x, t, im, isc, om, osc)
    assert abs(loss - rloss) <= 1e-3 * rloss, (loss, rloss)
    # per-layer weight and bias gradients, relative 2-norm error
    off = 0
    for i, o in zip(net.dims[:-1], net.dims[1:]):
        gw, rw = g[off:off + i * o], rg[off:off + i * o]
        off += i * o
        gb, rb = g[off:off + o], rg[off:off + o]
        off += o
        assert rel(gw, rw) < 2e-2, ("W", i, o, rel(gw, rw))
        assert rel(gb, rb) < 2e-2, ("b", i, o, rel(gb, rb))
    # a wrong operand layout or a transposed product fails by O(1)
    assert rel(g, rg) < 1e-2
    net.close()


def test_predict_matches_cpu_reference(lanes):
    rng = np.random.default_rng(5)
    net, p, (im, isc, om, osc) = random_net(lanes, rng, 32, 323, 256, 3)
    x = im + isc * rng.standard_normal((300, 32))
    y = net.predict(x)
    ry = mlp_ref.predict(p, net.dims, x, im, isc, om, osc)
    assert (y[:, 0] == 0).all()
    assert rel(y[:, 1:] - om, ry[:, 1:] - om) < 1e-2
    net.close()


def test_training_reduces_loss(lanes):
    rng = np.random.default_rng(9)
    n_in, n_out, rows = 8, 12, 4096
    A = rng.standard_normal((n_in, n_out))
    x = rng.standard_normal((rows, n_in))
    y = np.tanh(x @ A) + 0.1 * (x[:, :1] ** 2)
    t = np.concatenate([np.zeros((rows, 1)), y], axis=1)
    net = lanes.Mlp(n_in, n_out, n_hidden=64, n_hidden_layers=2, seed=2)
    net.set_scaling(np.zeros(n_in), np.ones(n_in), y.mean(0), y.std(0))
    losses = net.train(x, t, epochs=100, batch=256)
    say(f"losses {losses[0]:.4f} -> {losses[-1]:.5f}")
    assert losses[-1] < 0.1 * losses[0]
    # inference agrees with what training measured: the standardized error
    pred = net.predict(x[:512])[:, 1:]
    mse = (((pred - y[:512]) / y.std(0)) ** 2).mean()
    assert mse < 2.0 * losses[-1] + 1e-3, (mse, losses[-1])
    net.close()


def train_warm_start(lanes, fam, eng, rows, seed=11, epochs=30):
    """Operating points solved by the engine itself train the network."""
    vals = fam.sample(seed, rows)
    sol = eng.solve_op(vals, tiling=lanes.TILING_WAVE)
    ok = sol["reasons"] == 0
    vals, xs = vals[ok], sol["x"][ok]
    net = lanes.Mlp(vals.shape[1], eng.n - 1, seed=7)
    net.set_scaling(vals.mean(0), vals.std(0) + 1e-30, xs[:, 1:].mean(0), xs[:, 1:].std(0) + 1e-12)
    losses = net.train(vals, xs, epochs=epochs, batch=512)
    return net, losses


def test_learned_warm_start_saves_iterations_and_matches_oracle(lanes, oracle):
    fam = circuits.ota_array()
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    say("training the warm start on 8192 operating points")
    net, losses = train_warm_start(lanes, fam, eng, 8192)
    say(f"  loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    c = oracle.Circuit(fam.deck.arrays())
    x_nom = c.op(fam.nominal()[None, :], q)["x"][0]
    eng.set_warm_start(net, x_nom)
    vals = fam.sample(2029, 2048)   # points the network never saw
    cold = eng.solve_op(vals, tiling=lanes.TILING_WAVE)
    nom = eng.solve_op(vals, x0=np.repeat(x_nom[None, :], len(vals), 0), tiling=lanes.TILING_WAVE)
    got = eng.solve_op(vals, learned=True)
    tm = eng.timing()
    say(f"  iterations cold {cold['iterations'].mean():.2f} nominal {nom['iterations'].mean():.2f} "
        f"learned {got['iterations'].mean():.2f}; fallbacks {tm.warm_fallbacks}; inference {tm.warm_ms:.3f} ms")
    assert (got["reasons"] == 0).all()
    assert got["iterations"].mean() < nom["iterations"].mean() < cold["iterations"].mean()
    # the learned guesses are the network's predictions, bitwise
    x0 = net.predict(vals)
    pick = np.arange(0, len(vals), 37)
    ref = c.op(vals[pick], q, start_x=x0[pick], threads=8, nominal_x=x_nom)
    assert np.array_equal(got["iterations"][pick], ref["iterations"])
    assert np.array_equal(got["x"][pick].view(np.uint64), ref["x"].view(np.uint64))
    # same operating points as the cold start, within the Newton tolerance
    nn = fam.deck.arrays()["n_nodes"]
    assert (np.abs(got["x"][:, :nn] - cold["x"][:, :nn]) <= 1e-3 * np.abs(cold["x"][:, :nn]) + 2e-6).all()
    eng.close()
    net.close()


def test_nominal_fallback_when_the_guess_is_bad(lanes, oracle):
    """An untrained network's guesses are far off: lanes whose first
    residual does not fall restart from the nominal point, bitwise as the
    oracle's rule, and still converge."""
    fam = circuits.ota_array(8)
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    c = oracle.Circuit(fam.deck.arrays())
    x_nom = c.op(fam.nominal()[None, :], q)["x"][0]
    net = lanes.Mlp(fam.deck.arrays()["n_live"], eng.n - 1, n_hidden=64, n_hidden_layers=2, seed=5)
    nl = fam.deck.arrays()["n_live"]
    net.set_scaling(np.zeros(nl), np.ones(nl), np.zeros(eng.n - 1), 3.0 * np.ones(eng.n - 1))
    eng.set_warm_start(net, x_nom)
    vals = fam.sample(31, 256)
    got = eng.solve_op(vals, learned=True)
    tm = eng.timing()
    ref = c.op(vals, q, start_x=net.predict(vals), threads=8, nominal_x=x_nom)
    say(f"  fallbacks {tm.warm_fallbacks} (oracle {int(ref['fell_back'].sum())})")
    assert tm.warm_fallbacks == int(ref["fell_back"].sum()) > 0
    assert np.array_equal(got["reasons"], ref["reasons"])
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(got["x"].view(np.uint64), ref["x"].view(np.uint64))
    eng.close()
    net.close()


def test_headline_learned_fp32_full_size_bitwise_against_oracle(lanes, oracle):
    """The headline mode at its workload: 16,384 OTA points on the
    benchmark's seed, learned starts from a network trained on 50,000
    operating points the engine solved, FP32 factors + FP64 refinement.
    Every point is bitwise the oracle's FP32 mode from the same guesses
    (iterations, reasons, FP64 fallbacks and nominal restarts included), and
    within 1e-7 relative of the oracle's plain FP64 lanes from those
    guesses (the contract)."""
    fam = circuits.ota_array()
    eng = lanes.Engine(fam.deck.arrays())
    _, _, q = eng.pattern()
    c = oracle.Circuit(fam.deck.arrays())
    x_nom = eng.solve_op(fam.nominal()[None, :])["x"][0]
    tv = fam.sample(7, 50000)
    sol = eng.solve_op(tv, x0=np.repeat(x_nom[None, :], len(tv), 0), tiling=lanes.TILING_WAVE)
    ok = sol["reasons"] == 0
    say(f"training the warm start on {ok.sum()} operating points")
    net = lanes.Mlp(tv.shape[1], eng.n - 1, n_hidden=256, n_hidden_layers=3, seed=7)
    net.set_scaling(tv[ok].mean(0), tv[ok].std(0) + 1e-30, sol["x"][ok][:, 1:].mean(0),
                    sol["x"][ok][:, 1:].std(0) + 1e-12)
    losses = net.train(tv[ok], sol["x"][ok], epochs=30, batch=512)
    say(f"  loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    eng.set_warm_start(net, x_nom)
    vals = fam.sample(2026, 16384)
    got = eng.solve_op(vals, learned=True, precision=lanes.FP32_REFINED)
    tm = eng.timing()
    x0 = net.predict(vals)
    ref = c.op(vals, q, start_x=x0, threads=16, precision=1, nominal_x=x_nom)
    say(f"  its/pt {got['iterations'].mean():.3f}; fp64 fallbacks {tm.fp64_lanes} (oracle "
        f"{int(ref['fp64_lanes'].sum())}); nominal restarts {tm.warm_fallbacks} (oracle {int(ref['fell_back'].sum())})")
    assert np.array_equal(got["reasons"], ref["reasons"])
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(got["x"].view(np.uint64), ref["x"].view(np.uint64))
    assert tm.fp64_lanes == int(ref["fp64_lanes"].sum())
    assert tm.warm_fallbacks == int(ref["fell_back"].sum())
    r64 = c.op(vals, q, start_x=x0, threads=16, nominal_x=x_nom)
    ok = (ref["reasons"] == 0) & (r64["reasons"] == 0)
    err = np.abs(got["x"][ok] - r64["x"][ok]) / np.maximum(np.abs(r64["x"][ok]), 1e-6)
    say(f"  max relative error against the FP64 lanes from the same guesses {err.max():.3e}")
    assert ok.mean() > 0.999 and err.max() <= 1e-7
    assert got["iterations"].mean() < 3.0
    eng.close()
    net.close()

"""Pins the oracle's charges, its transient predictor and the Monte Carlo
sampler against mathematics that does not go through the oracle's code.

- Diode charge: q(v) equals the integral of the textbook SPICE junction
  capacitance (CJO (1 - v/VJ)^-M below FC VJ, its linear extension above),
  computed by quadrature, plus TT times the diode current (models/diode.va:
  329-359: depletion below fc*vj, the f1/f2/f3 continuation above, TT*cd
  with cd including the gmin shunt).
- The c plane: the charge Jacobian of every model with charges (capacitor,
  diode) equals central finite differences of the row charges.
- The predictor (tran.zig:37-41,1131-1138): each step's first iterate is
  cur + (dt/dt_prev)(cur - prev) of the last two accepted points, then one
  limit pass against the accepted point (DEVpnjlim for a junction); turning
  the predictor off moves the accepted points by less than the Newton
  tolerance and costs iterations.
- The sampler: the reference's fixture mc/gaussian_source.sp (4,096 trials,
  5 % relative sigma: mean within 0.5 %, sigma within 10 %), repeatability,
  Kolmogorov-Smirnov tests of the normal and uniform draws, and independence
  across keys.
"""
import numpy as np
import pytest
from scipy import integrate, stats

from espice_amd import circuits
from espice_amd.deck import Deck

KB, QE = 1.38064852e-23, 1.6021766208e-19
VT27 = 300.15 * KB / QE


def cdep(v, cjo, vj, m, fc):
    """SPICE junction depletion capacitance (Nagel, SPICE2, 1975): the
    power law below fc*vj, its tangent-matched linear extension above."""
    if v < fc * vj:
        return cjo * (1.0 - v / vj) ** (-m)
    return cjo / (1.0 - fc) ** (1.0 + m) * (1.0 - fc * (1.0 + m) + m * v / vj)


def one_diode(model):
    d = Deck("d1")
    d.vsource("va", "a", 0, 0.0)
    d.diode("d1", "a", 0, model)
    return d


@pytest.mark.parametrize("model", [
    {"is": 1e-14, "cjo": 2e-12, "vj": 0.8, "m": 0.5, "fc": 0.5, "tt": 0.0},
    {"is": 3e-15, "cjo": 1e-12, "vj": 0.7, "m": 0.33, "fc": 0.6, "tt": 5e-9},
    {"is": 1e-16, "cjo": 0.0, "tt": 1e-8},
])
def test_diode_charge_is_the_integral_of_its_capacitance(oracle, model):
    d = one_diode(model)
    c = oracle.Circuit(d.arrays())
    row = d.node_index("a")
    full = {"vj": 1.0, "m": 0.5, "fc": 0.5, "gmin": 1e-12, **model}
    for v in (-5.0, -1.0, -0.2, 0.0, 0.1, 0.3, 0.45, 0.55, 0.62, 0.7, 0.75):
        x = np.zeros(c.n)
        x[row] = v
        _, _, cp, q = c.assemble_charges(d.nominal(), x)
        qdep = 0.0
        if full["cjo"] > 0:
            args = (full["cjo"], full["vj"], full["m"], full["fc"])
            brk = full["fc"] * full["vj"]
            if v > brk:
                qdep = integrate.quad(cdep, 0.0, brk, args=args, epsabs=0, epsrel=1e-13)[0] + \
                    integrate.quad(cdep, brk, v, args=args, epsabs=0, epsrel=1e-13)[0]
            else:
                qdep = integrate.quad(cdep, 0.0, v, args=args, epsabs=0, epsrel=1e-13)[0]
        cd = full["is"] * np.expm1(v / VT27) + full["gmin"] * v if v >= -3 * VT27 else None
        if cd is None:   # the reverse cubic tail of dioload.c
            w = 3.0 * VT27 / (v * np.e)
            cd = -full["is"] * (1.0 + w ** 3) + full["gmin"] * v
        expect = qdep + full["tt"] * cd
        scale = max(abs(expect), 1e-30)
        assert abs(q[row] - expect) <= 1e-9 * scale + 1e-27, (v, q[row], expect)
        # the diagonal of the c plane is the capacitance at v (plus TT gd)
        slots = c.pattern()
        cptr, ridx = slots[0], slots[1]
        diag = [p for p in range(cptr[row], cptr[row + 1]) if ridx[p] == row][0]
        cap = cdep(v, full["cjo"], full["vj"], full["m"], full["fc"]) if full["cjo"] > 0 else 0.0
        gd = full["is"] * np.exp(min(v / VT27, 80.0)) / VT27 + full["gmin"] if v >= -3 * VT27 else None
        if gd is not None:
            want = cap + full["tt"] * gd
            assert abs(cp[diag] - want) <= 1e-9 * abs(want) + 1e-27, (v, cp[diag], want)


def charge_fd_check(oracle, deck, live, x, rel=1e-6):
    c = oracle.Circuit(deck.arrays())
    cptr, ridx, _, _ = c.pattern()
    _, _, cp, _ = c.assemble_charges(live, x)
    Cm = np.zeros((c.n, c.n))
    for j in range(c.n):
        for p in range(cptr[j], cptr[j + 1]):
            Cm[ridx[p], j] += cp[p]
    for j in range(1, c.n):
        h = 1e-6 * max(1.0, abs(x[j]))
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        fd = (c.assemble_charges(live, xp)[3] - c.assemble_charges(live, xm)[3]) / (2 * h)
        scale = np.abs(Cm[:, j]).max() + 1e-30
        assert np.abs(fd - Cm[:, j]).max() <= rel * scale, (j, np.abs(fd - Cm[:, j]).max(), scale)
    return Cm


def test_charge_jacobian_matches_finite_differences_mesh(oracle):
    fam = circuits.rc_mesh(5, diode_frac=0.4)
    c = oracle.Circuit(fam.deck.arrays())
    rng = np.random.default_rng(11)
    vals = fam.sample(2, 3)
    for l in range(3):
        x = np.zeros(c.n)
        x[1:] = rng.uniform(-0.3, 0.75, c.n - 1)
        Cm = charge_fd_check(oracle, fam.deck, vals[l], x)
        assert np.abs(Cm).max() > 0


def test_charge_jacobian_matches_finite_differences_diode_with_charges(oracle):
    d = Deck("dq")
    d.vsource("vs", "in", 0, 1.0)
    d.resistor("r1", "in", "a", 1e3)
    d.diode("d1", "a", "b", {"is": 1e-14, "cjo": 2e-12, "vj": 0.8, "m": 0.4, "fc": 0.5, "tt": 3e-9})
    d.capacitor("c1", "b", 0, 5e-13)
    d.diode("d2", "b", 0, {"is": 5e-15, "cjo": 1e-12, "vj": 0.6, "m": 0.5, "fc": 0.4, "tt": 1e-9})
    a, b = d.node_index("a"), d.node_index("b")
    for va, vb in ((0.9, 0.45), (0.2, 0.1), (-0.5, 0.35), (1.4, 0.6)):
        x = np.zeros(oracle.Circuit(d.arrays()).n)
        x[d.node_index("in")] = 1.0
        x[a], x[b] = va, vb
        charge_fd_check(oracle, d, d.nominal(), x)


def pnjlim(vnew, vold, vt, vcrit):
    """DEVpnjlim of ngspice devsup.c (the form the reference quotes in
    dev/solvers/newton-raphson-convergence.md:42-56, with the reverse-bias
    clamp of :53-55)."""
    if vnew > vcrit and abs(vnew - vold) > vt + vt:
        if vold > 0:
            arg = (vnew - vold) / vt
            if arg > 0:
                return vold + vt * (2.0 + np.log(arg - 2.0))
            return vold - vt * (2.0 + np.log(2.0 - arg))
        return vt * np.log(vnew / vt)
    if vnew < 0:
        floor = -vold - 1.0 if vold > 0 else 2.0 * vold - 1.0
        return max(vnew, floor)
    return vnew


def diode_transient():
    """A current step into a capacitor with a junction clamp: the clamp's
    voltage rises fast enough for the predicted jump to exceed 2 vt."""
    d = Deck("clamp")
    d.vsource("vs", "in", 0, 0.0, step_to=5.0, step_at=0.0)
    d.resistor("r1", "in", "a", 200.0)
    d.capacitor("c1", "a", 0, 2e-12)
    d.diode("d1", "a", 0, {"is": 1e-14, "cjo": 5e-13, "vj": 0.8, "m": 0.5, "tt": 1e-10})
    d.resistor("r2", "a", "b", 1e3)
    d.capacitor("c2", "b", 0, 1e-12)
    return d


def test_predictor_first_iterate_is_the_extrapolation(oracle):
    d = diode_transient()
    c = oracle.Circuit(d.arrays())
    q = np.arange(c.n, dtype=np.uint32)
    steps = 60
    r = oracle.tran(c, np.zeros((2, 0)), q, 20e-12, steps, probes=list(range(c.n)), trace=True)
    assert (r["reasons"] == 0).all()
    acc = r["probes"][0]                               # accepted points, every unknown
    fx = r["first_x"][0]
    op = oracle.Circuit(d.arrays()).op(np.zeros((1, 0)), q)["x"][0]
    prev = np.vstack([op[None, :], acc[:-1]])          # the point before each accepted one
    for s in range(1, steps):
        expect = acc[s - 1] + 1.0 * (acc[s - 1] - prev[s - 1])   # dt/dt_prev = 1
        assert np.array_equal(fx[s], expect), s


def test_predictor_limit_pass_is_pnjlim_against_the_accepted_point(oracle):
    d = diode_transient()
    c = oracle.Circuit(d.arrays())
    q = np.arange(c.n, dtype=np.uint32)
    steps = 60
    r = oracle.tran(c, np.zeros((1, 0)), q, 20e-12, steps, probes=list(range(c.n)), trace=True)
    acc = r["probes"][0]
    a = d.node_index("a")
    kinds = d.arrays()["kind"]
    inst = sorted(range(len(kinds)), key=lambda i: (int(kinds[i]), i))   # instances in stamp order
    k = inst.index(d.names.index("d1"))
    isat = 1e-14
    vcrit = VT27 * np.log(VT27 / (np.sqrt(2.0) * isat))
    moved = 0
    for s in range(1, steps):
        v_trial = r["first_x"][0, s, a]
        v_acc = acc[s - 1, a]
        want = pnjlim(v_trial, v_acc, VT27, vcrit)
        got = r["first_lim"][0, s, k, 0]
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (s, got, want, v_trial, v_acc)
        moved += want != v_trial
    assert moved > 0   # the workload does exercise the limiter on predicted jumps


def test_predictor_off_stays_within_newton_tolerance(oracle):
    for fam, dt, steps in ((circuits.rc_mesh(5, diode_frac=0.4), 1e-12, 40), (None, 20e-12, 60)):
        if fam is None:
            d = diode_transient()
            vals = np.zeros((1, 0))
        else:
            d = fam.deck
            vals = fam.sample(4, 4)
        c = oracle.Circuit(d.arrays())
        q = np.arange(c.n, dtype=np.uint32)
        tol = oracle.Tol(1e-3, 1e-12, 1e-6, 1e-9, 50, 0)
        off = oracle.Tol(1e-3, 1e-12, 1e-6, 1e-9, 50, oracle.NO_PREDICT)
        on = oracle.tran(c, vals, q, dt, steps, probes=list(range(c.n)), tol=tol)
        nop = oracle.tran(c, vals, q, dt, steps, probes=list(range(c.n)), tol=off)
        assert (on["reasons"] == 0).all() and (nop["reasons"] == 0).all()
        # both accept points inside the Newton tolerance of the step solution:
        # reltol |x| + vntol per unknown, a few times over for the two paths
        tl = 1e-3 * np.maximum(np.abs(on["probes"]), np.abs(nop["probes"])) + 1e-6
        assert (np.abs(on["probes"] - nop["probes"]) <= 4 * tl).all()
        if fam is None:
            # on a smooth waveform the predictor saves Newton iterations (the
            # mesh's stiff 50-ohm corner rings under the trapezoid, where
            # extrapolation does not pay)
            assert on["iterations"].sum() < nop["iterations"].sum()


# ---- the Monte Carlo sampler ----

def test_gaussian_source_fixture_moments(oracle):
    # mc/gaussian_source.sp: Vin = 1 V with 5 % relative sigma into 1 kOhm,
    # 4,096 trials; mean in [0.995, 1.005], sigma in [0.045, 0.055]; a
    # repeated run is bitwise the same.
    d = Deck("gauss")
    d.vsource("vin", "out", 0, 1.0)
    d.resistor("r1", "out", 0, 1e3)
    d.bind("vin", 0, "vin")
    d.bind("r1", 0, "r1")
    fam = circuits.Family("gauss", d, [("vin", "rel_normal", 1.0, 0.05), ("r1", "rel_normal", 1e3, 0.05)])
    vals = fam.sample(42, 4096)
    c = oracle.Circuit(d.arrays())
    r = c.op(vals, np.arange(c.n, dtype=np.uint32), threads=8)
    assert (r["reasons"] == 0).all()
    v = r["x"][:, d.node_index("out")]
    assert 0.995 <= v.mean() <= 1.005 and 0.045 <= v.std(ddof=1) <= 0.055
    assert np.isfinite(v).all()
    again = c.op(fam.sample(42, 4096), np.arange(c.n, dtype=np.uint32), threads=3)
    assert np.array_equal(r["x"].view(np.uint64), again["x"].view(np.uint64))


def test_normal_and_uniform_draws_have_their_distributions():
    lanes = np.arange(200_000, dtype=np.uint64)
    z = circuits.normal(2026, lanes, 3)
    assert abs(z.mean()) < 6 / np.sqrt(len(z))
    assert abs(z.var() - 1.0) < 6 * np.sqrt(2.0 / len(z))
    assert abs(stats.skew(z)) < 6 * np.sqrt(6.0 / len(z))
    assert abs(stats.kurtosis(z)) < 6 * np.sqrt(24.0 / len(z))
    assert stats.kstest(z, "norm").pvalue > 1e-4
    u = circuits.uniform01(2026, lanes, 5)
    assert (u > 0).all() and (u < 1).all()
    assert stats.kstest(u, "uniform").pvalue > 1e-4


def test_draws_are_independent_across_keys():
    lanes = np.arange(100_000, dtype=np.uint64)
    a = circuits.normal(7, lanes, 0)
    b = circuits.normal(7, lanes, 1)          # another variable
    c = circuits.normal(8, lanes, 0)          # another seed
    d = circuits.normal(7, lanes + 1, 0)      # the neighbouring lane
    bound = 6 / np.sqrt(len(lanes))
    for other in (b, c, d):
        assert abs(np.corrcoef(a, other)[0, 1]) < bound
    # a point's draw depends only on its key, not on the batch around it
    idx = np.array([5, 99_999, 17, 40_000], dtype=np.uint64)
    assert np.array_equal(circuits.normal(7, idx, 0), a[idx.astype(np.int64)])


@pytest.mark.parametrize("model,volts", [
    ({"is": 1e-14, "bv": 15.0, "ibv": 1e-3}, (-20.0, -16.0, -15.2, -14.0, -1.0, 0.3, 0.7)),
    ({"is": 1e-14, "ikf": 10e-3}, (0.2, 0.6, 0.75, 0.9, 1.1)),
    ({"is": 1e-14, "isr": 1e-10, "nr": 2.0, "m": 0.4, "vj": 0.8}, (-0.05, 0.1, 0.3, 0.5, 0.7)),
    ({"is": 1e-14, "bv": 100.0, "ibv": 1e-10, "ikf": 1e-3, "isr": 1e-11}, (-101.0, -99.0, -1.0, 0.4, 0.8)),
])
def test_diode_breakdown_knee_recombination_jacobian(oracle, model, volts):
    """The breakdown branch, the IKF knee and the ISR recombination current
    (diode.va:264-314): the Jacobian equals central finite differences of
    the residual on either side of each branch's threshold."""
    d = one_diode(model)
    c = oracle.Circuit(d.arrays())
    cp, ri, _, _ = c.pattern()
    a = d.node_index("a")
    diag = [p for p in range(cp[a], cp[a + 1]) if ri[p] == a][0]
    for v in volts:
        x = np.zeros(c.n)
        x[a] = v
        g, _ = c.assemble_at(d.nominal(), x)
        h = 1e-7 * max(1.0, abs(v))
        xp, xm = x.copy(), x.copy()
        xp[a] += h
        xm[a] -= h
        fd = (c.assemble_at(d.nominal(), xp)[1][a] - c.assemble_at(d.nominal(), xm)[1][a]) / (2 * h)
        assert abs(fd - g[diag]) <= 1e-5 * abs(g[diag]) + 1e-12, (v, fd, g[diag])

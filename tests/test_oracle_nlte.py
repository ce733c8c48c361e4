"""The CPU oracle's sparse level-population update (oracle/mgx_oracle_nlte.c),
pinned by what the mathematics and the reference's own benchmark fix:

* the long-double column-pivoted QR against numpy's LU solve;
* a two-level atom, one statistical-equilibrium step, against the closed
  form n_u = n_tot a / (a + b) with the effective mean intensity, the
  approximated-Lambda term, interpolated collision rates and the ortho/para
  H2 fraction written out by hand;
* detailed balance: a 10-level rotor with J = B_nu(T) in every line and
  collision rates that obey detailed balance at T must sit at LTE;
* the negative-population correction against numpy's solve of the same
  system followed by the documented clamp-and-renormalise;
* Ng acceleration against the constrained least-squares problem it solves
  (min |sum c_i r_i| subject to sum c_i = 1), set up and solved by numpy
  through its KKT system;
* J_lin of the sparse path against the line integral of the dense J;
* the reference's van Zadelhoff 1a benchmark: the whole non-LTE loop (with
  adaptive Ng) against RATRAN's populations within the reference's bounds
  (max relative error < 0.12 / 0.13, vanZadelhoff_1_1D.py:162).
"""
import os

import numpy as np
import pytest

from magritte_amd import _abi, models
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAMDA = os.path.join(ROOT, "tests", "golden", "lamda_test.txt")
HH4PI = models.HH / (4.0 * np.pi)


def _rotor_1d(nlev=10, T=50.0, P=12, dens=1.0e10, rate=1.0e-16, op="n", tmp=None):
    """Small 1D model: a linear rotor, H2 density in column 0, the species in column 1."""
    pos = np.zeros((P, 3))
    pos[:, 0] = np.arange(P) * 1.0e12
    nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
    m = models.ModelData()
    m.dimension = 1
    m.set_geometry(pos, np.zeros((P, 3)), nbs, [0, P - 1], [[-1.0, 0, 0], [1.0, 0, 0]],
                   [0.5, 0.5])
    ab = np.stack([np.full(P, dens), np.full(P, 1.0e-4 * dens)], axis=1)
    m.set_thermo(np.full(P, T), np.full(P, 150.0), ab)
    ld = models.linedata_linear_rotor(57.6e9, 0.11, nlev)
    tmp = np.array([10.0, 20.0, 50.0, 100.0, 200.0]) if tmp is None else tmp
    ld["colpar"] = [models.colpar_constant(ld, 0, rate, tmp=tmp, orth_or_para=op)]
    m.set_lines([ld], 5)
    return m.finalize()


def _planck(nu, T):
    return 2.0 * models.HH * nu**3 / models.CC**2 / np.expm1(models.HH * nu / (models.KB * T))


def test_cpqr_solve_matches_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 17, 41):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        x = oracle.cpqr_solve(A, b)
        ref = np.linalg.solve(A, b)
        assert np.max(np.abs(x - ref)) <= 1e-13 * np.max(np.abs(ref)), n
    # a rank-one consistent system: the rank-revealing solve still satisfies A x = b
    u = rng.standard_normal(4)
    A = np.outer(u, u)
    b = A @ rng.standard_normal(4)
    x = oracle.cpqr_solve(A, b)
    assert np.max(np.abs(A @ x - b)) < 1e-13 * np.max(np.abs(b))


@pytest.mark.parametrize("op,T", [("n", 35.0), ("o", 35.0), ("p", 35.0), ("n", 65.0),
                                  ("n", 10.0), ("n", 200.0)])
def test_two_level_closed_form(op, T):
    """One statistical-equilibrium step of a two-level species: lower 0, upper 1.
    Collision rates at T by the reference's search + interpolation
    (collisionPartner.tpp:5-23): linear between the tabulated temperatures
    below the last interval, the end values outside the table, and the last
    tabulated value throughout the last interval (search returns ntmp - 1
    there, which the reference treats as the upper end)."""
    tmp = np.array([20.0, 50.0, 80.0])
    P = 6
    pos = np.zeros((P, 3))
    pos[:, 0] = np.arange(P) * 1.0e12
    nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
    m = models.ModelData()
    m.dimension = 1
    m.set_geometry(pos, np.zeros((P, 3)), nbs, [0, P - 1], [[-1.0, 0, 0], [1.0, 0, 0]],
                   [0.5, 0.5])
    dens = np.linspace(1.0e8, 1.0e11, P)
    m.set_thermo(np.full(P, T), np.full(P, 150.0), np.stack([dens, 1e-4 * dens], axis=1))
    ld = models.linedata_two_level(115.27e9, 7.2e-8)
    Cd = np.array([[1.0e-16], [3.0e-16], [4.0e-16]])
    ld["colpar"] = [models._colpar(ld, 0, op, [1], [0], tmp, Cd)]
    m.set_lines([ld], 5)
    m.finalize()
    prm = _abi.default_nlte_params()
    nl = oracle.NLTE(m, prm)
    nl.set_lte()
    pop0 = nl.populations()["population"].reshape(P, 2)
    rng = np.random.default_rng(5)
    Jlin = rng.uniform(1e-17, 1e-15, (P, 1))
    lam = rng.uniform(0.0, 0.5, (P, 1))
    nl.set_Jeff(Jlin, lam)
    nl.stateq(lam)
    got = nl.populations()["population"].reshape(P, 2)

    A, Bs, Ba, nu = ld["A"][0], ld["Bs"][0], ld["Ba"][0], ld["frequency"][0]
    Ce = m.linedata[0]["colpar"][0]["Ce"][:, 0]
    if T <= tmp[0]:
        cd, ce = Cd[0, 0], Ce[0]
    elif T > tmp[1]:
        cd, ce = Cd[2, 0], Ce[2]
    else:
        step = (T - tmp[0]) / (tmp[1] - tmp[0])
        cd = Cd[0, 0] + (Cd[1, 0] - Cd[0, 0]) * step
        ce = Ce[0] + (Ce[1] - Ce[0]) * step
    # expf of the float argument, correctly rounded to float, then double arithmetic
    e32 = float(np.float32(np.exp(np.float64(np.float32(-170.5 / T)))))
    fpara = 1.0 / (1.0 + 9.0 * e32)
    scale = {"n": 1.0, "o": 1.0 - fpara, "p": fpara}[op]
    for p in range(P):
        n0, n1 = pop0[p]
        tot = n0 + n1
        abn = dens[p] * scale
        Jeff = Jlin[p, 0] - HH4PI * lam[p, 0] * n1
        kap = max(nu * HH4PI * (n0 * Ba - n1 * Bs), prm.min_line_opacity)
        up = Ba * Jeff + ce * abn                               # 0 -> 1
        down = A + Bs * Jeff + cd * abn - lam[p, 0] * kap / nu  # 1 -> 0
        n1_ref = tot * up / (up + down)
        assert abs(got[p, 1] - n1_ref) <= 1e-12 * n1_ref, (p, got[p, 1], n1_ref)
        assert abs(got[p, 0] + got[p, 1] - tot) <= 1e-14 * tot


def test_detailed_balance_gives_lte():
    T = 50.0
    m = _rotor_1d(nlev=10, T=T)
    prm = _abi.default_nlte_params()
    nl = oracle.NLTE(m, prm)
    nl.set_lte()
    lte = nl.populations()["population"].copy()
    # perturb the start so that the step has something to do
    rng = np.random.default_rng(8)
    nl.set_populations(lte * rng.uniform(0.5, 1.5, lte.shape))
    Jlin = np.tile(_planck(m.linedata[0]["frequency"], T), (m.npoints, 1))
    lam = np.zeros((m.npoints, m.nlines))
    nl.set_Jeff(Jlin, lam)
    nl.stateq(lam)
    got = nl.populations()["population"]
    E, g = m.linedata[0]["energy"], m.linedata[0]["weight"]
    b = g * np.exp(-E / (models.KB * T))
    ref = np.tile(b / b.sum(), m.npoints) * 1.0e-4 * 1.0e10
    assert np.max(np.abs(got - ref) / ref) < 1e-12


def test_lte_populations():
    m = _rotor_1d(nlev=8, T=20.0)
    nl = oracle.NLTE(m, _abi.default_nlte_params())
    nl.set_lte()
    got = nl.populations()["population"].reshape(m.npoints, 8)
    E, g = m.linedata[0]["energy"], m.linedata[0]["weight"]
    b = g * np.exp(-E / (models.KB * 20.0))
    ref = b / b.sum() * 1.0e6
    assert np.max(np.abs(got - ref) / ref) < 1e-14
    # and the line data written into the model from them
    d = m.linedata[0]
    k = np.arange(d["nrad"])
    ni, nj = ref[d["irad"]], ref[d["jrad"]]
    eps = d["frequency"] * HH4PI * d["A"] * ni
    kap = d["frequency"] * HH4PI * (nj * d["Ba"] - ni * d["Bs"])
    assert np.allclose(m.emissivity[0, k], eps, rtol=1e-13, atol=0)
    assert np.allclose(m.opacity[0, k], kap, rtol=1e-12, atol=0)
    assert np.allclose(m.interp_crit[0, k], kap / 1.0e6, rtol=1e-12, atol=0)


@pytest.mark.parametrize("inv_frac", [-1.0, 1.01])
def test_negative_population_correction(inv_frac):
    """Unphysical negative J drives the solution negative for some levels."""
    m = _rotor_1d(nlev=6, T=20.0, dens=1.0e6)
    prm = _abi.default_nlte_params(population_inversion_fraction=inv_frac)
    nl = oracle.NLTE(m, prm)
    nl.set_lte()
    rng = np.random.default_rng(9)
    P, L = m.npoints, m.nlines
    Jlin = rng.uniform(-1e-16, 1e-16, (P, L))
    lam = np.zeros((P, L))
    nl.set_Jeff(Jlin, lam)
    raw = []
    for p in range(P):
        M, y = nl.system(lam, 0, p)
        raw.append(np.linalg.solve(M, y))
    nl.stateq(lam)
    got = nl.populations()["population"].reshape(P, 6)
    tot = 1.0e2
    n_neg = 0
    for p in range(P):
        x = raw[p].copy()
        n_neg += int((x < 0).sum())
        x[x < 0] = 0.0
        x = x * tot / x.sum()
        assert np.all(got[p] >= 0.0)
        assert abs(got[p].sum() - tot) < 1e-12 * tot
        if inv_frac < 0:
            assert np.allclose(got[p], x, rtol=1e-10, atol=1e-12 * tot), p
    assert n_neg > 0, "the case must exercise the correction"


def test_ng_solves_the_constrained_least_squares():
    m = _rotor_1d(nlev=6, T=20.0)
    prm = _abi.default_nlte_params()
    nl = oracle.NLTE(m, prm)
    nl.set_lte()
    rng = np.random.default_rng(12)
    P, L = m.npoints, m.nlines
    lam = np.zeros((P, L))
    hist = [nl.populations()["population"].copy()]
    for it in range(4):
        Jlin = np.tile(_planck(m.linedata[0]["frequency"], 20.0 + 5.0 * it), (P, 1))
        Jlin *= rng.uniform(0.9, 1.1, Jlin.shape)
        nl.set_Jeff(Jlin, lam)
        nl.stateq(lam)
        hist.append(nl.populations()["population"].copy())
    # populations list of the reference: [lte, lte, x1, x2, x3], current x4
    xs = [hist[0]] + hist
    for order in (2, 3):
        sub = xs[-(order + 1):]
        R = np.array([sub[i + 1] - sub[i] for i in range(order)])
        G = R @ R.T
        K = np.zeros((order + 1, order + 1))
        K[:order, :order] = G
        K[:order, order] = 1.0
        K[order, :order] = 1.0
        rhs = np.zeros(order + 1)
        rhs[order] = 1.0
        c = np.linalg.solve(K, rhs)[:order]
        ref = sum(ci * xi for ci, xi in zip(c, sub[1:]))
        # trial: does not change the populations
        nl.trial_ng(order)
        tr = nl.populations()["trial"]
        assert np.max(np.abs(tr - ref) / np.abs(ref)) < 1e-9, order
        assert np.array_equal(nl.populations()["population"], hist[-1])
    nl.ng(3)
    got = nl.populations()["population"]
    sub = xs[-4:]
    R = np.array([sub[i + 1] - sub[i] for i in range(3)])
    K = np.block([[R @ R.T, np.ones((3, 1))], [np.ones((1, 3)), np.zeros((1, 1))]])
    c = np.linalg.solve(K, np.r_[0.0, 0.0, 0.0, 1.0])[:3]
    ref = sum(ci * xi for ci, xi in zip(c, sub[1:]))
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-9
    assert nl.stats()[1][0] == 0  # residuals cleared


def test_Jlin_is_the_line_integral_of_J(line_model):
    m = line_model
    u, J, lam = oracle.solve(m, m.options())
    nl = oracle.NLTE(m, _abi.default_nlte_params())
    e0, k0, i0 = m.emissivity.copy(), m.opacity.copy(), m.interp_crit.copy()
    try:
        nl.set_populations(m.population)
        nl.compute_Jeff(u, lam)
        got = nl.populations()
    finally:
        m.emissivity[...], m.opacity[...], m.interp_crit[...] = e0, k0, i0
    w = m.quad_weights[0]
    Q = m.nquads
    ref = np.stack([(J[:, l * Q:(l + 1) * Q] * w).sum(axis=1) for l in range(m.nlines)], axis=1)
    assert np.max(np.abs(got["J_lin"] - ref) / np.abs(ref)) < 1e-13
    n_up = m.population.reshape(m.npoints, 2)[:, 1]
    assert np.allclose(got["J_eff"][:, 0], got["J_lin"][:, 0] - HH4PI * lam[:, 0] * n_up,
                       rtol=1e-14, atol=0)


def _ratran_errors(m, pops):
    R = np.loadtxt(os.path.join(ROOT, "tests", "golden", "ratran_vanZadelhoff_1a.out"), skiprows=14)
    ra, rb, lp0, lp1 = R[:, 1], R[:, 2], R[:, 9], R[:, 10]
    rc = 0.5 * (ra + rb)
    rs = m.position[:, 0]
    abun = m.abundance[:, 0]

    def interp(y):  # scipy interp1d(fill_value='extrapolate') as linear extrapolation
        from scipy.interpolate import interp1d

        return interp1d(rc, y, fill_value="extrapolate")(rs)

    e0 = np.abs(2.0 * (pops[:, 0] / abun - interp(lp0)) / (pops[:, 0] / abun + interp(lp0)))
    e1 = np.abs(2.0 * (pops[:, 1] / abun - interp(lp1)) / (pops[:, 1] / abun + interp(lp1)))
    return float(np.max(e0[1:])), float(np.max(e1[1:]))


def test_van_zadelhoff_1a_against_ratran():
    m = models.model_vanzadelhoff_1d(LAMDA, "a")
    prm = _abi.default_nlte_params()
    nl, it, em, ex = oracle.compute_level_populations_sparse(m, m.options(), prm, True, 100,
                                                             nthreads=8)
    pops = nl.populations()["population"].reshape(m.npoints, 2)
    e0, e1 = _ratran_errors(m, pops)
    print(f"van Zadelhoff 1a: {it} iterations, max rel err vs RATRAN {e0:.4f} / {e1:.4f}")
    assert e0 < 0.12 and e1 < 0.13
    assert ex[-1] < prm.pop_prec * 10

"""The reference's c2p round-trip sweeps (test/test_c2p.cpp:130-455) run on the
oracle, plus an independent check of the Kastaun root against bisection.

EOS: hybrid piecewise polytrope, one piece, Gamma=2, K=168.5819, Gamma_th=1.8
(test/configs/c2p_test_gamma2.yaml); atmosphere and c2p parameters at their
schema defaults (parameters/grmhd.yaml).
"""
import ctypes as C
import math

import numpy as np

import grace_amd as ga
from _oracle import olib, p, c2p_batch

DIRS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
WS = [1.05, 1.5, 2.0, 3.0]
LOG10_RHO = [-12.0, -8.0, -5.0, -3.0, -1.5, -0.5]
TEMPS = [1e-3, 1e-1, 1.0, 10.0]


def hybrid_params():
    return ga.default_params(eos_kind=ga.EOS_HYBRID_PWPOLY, pwp_n=1, pwp_kappa0=1.685819e+02, pwp_gamma=[2.0],
                             gamma_th=1.8, eps_max=1e5, entropy_min=1e-13)


def mink():
    return np.array([1, 0, 0, 1, 0, 1, 1.0, 0, 0, 0], dtype=np.float64)


def schw_cks(R=4.0, M=1.0):
    H = M / R
    fac = 1 + 2 * H
    return np.array([fac, 0, 0, 1, 0, 1, 1 / math.sqrt(fac), 2 * H / fac, 0, 0], dtype=np.float64)


def prim_residual(a, b):
    """State-normalised infinity norm (test_c2p.cpp:87-103)."""
    zmag = np.sqrt(b[2] ** 2 + b[3] ** 2 + b[4] ** 2)
    scale = np.maximum(np.maximum(np.abs(b[0]), np.abs(b[7])), np.maximum(zmag, 1e-30))
    err = np.zeros_like(scale)
    for i in (0, 7, 2, 3, 4):  # RHO EPS ZX ZY ZZ
        err = np.maximum(err, np.abs(a[i] - b[i]))
    return err / scale


def build_states(params, metric, samples):
    """samples: list of (rho, temp, z[3], B[3]) -> prims (13,n), met (10,n), cons (10,n)."""
    L = olib()
    n = len(samples)
    prims = np.zeros((13, n))
    met = np.tile(metric[:, None], (1, n)).copy()
    for s, (rho, temp, z, B) in enumerate(samples):
        eps, cs2, ent = C.c_double(), C.c_double(), C.c_double()
        pr = L.orc_eos_from_temp(C.byref(params), temp, rho, C.byref(eps), C.byref(cs2), C.byref(ent), None)
        prims[:, s] = [rho, pr, z[0], z[1], z[2], 0.0, temp, eps.value, ent.value, B[0], B[1], B[2], 0.0]
    cons = np.zeros((10, n))
    L.orc_p2c_batch(C.byref(params), n, p(prims), p(met), p(cons))
    g = metric
    det = g[0] * g[3] * g[5]  # diagonal metrics only in these sweeps
    for i in range(3):
        cons[7 + i] = math.sqrt(det) * prims[9 + i]
    return prims, met, cons


def hydro_samples(metric):
    g = metric
    out = []
    for i_dir in range(6):
        for i_w in range(4):
            for i_t in range(4):
                for i_rho in range(6):
                    W = WS[i_w]
                    zm = math.sqrt(W * W - 1)
                    d = DIRS[i_dir]
                    n2 = g[0] * d[0] ** 2 + 2 * g[1] * d[0] * d[1] + 2 * g[2] * d[0] * d[2] + g[3] * d[1] ** 2 \
                        + 2 * g[4] * d[1] * d[2] + g[5] * d[2] ** 2
                    out.append((10 ** LOG10_RHO[i_rho], TEMPS[i_t], zm * d / math.sqrt(n2), np.zeros(3)))
    return out


def run_sweep(metric, samples, tol):
    params = hybrid_params()
    prims, met, cons = build_states(params, metric, samples)
    p1, _, err, fl, ev, fatal = c2p_batch(params, cons, met)
    res = prim_residual(p1, prims)
    assert fatal == 0
    assert int(fl.sum()) == 0, f"floored samples: {np.nonzero(fl)[0][:10]}"
    assert res.max() < tol, f"max residual {res.max():.3e} at {res.argmax()}"
    return res


def test_c2p_minkowski_hydro():
    res = run_sweep(mink(), hydro_samples(mink()), 1e-10)
    assert len(res) == 576


def test_c2p_minkowski_bsweep():
    rho0, t0, W0 = 1e-3, 1e-1, 1.5
    z0 = 1.118033988749895
    samples = []
    for bsq in [0.0, 1e-4, 1e-2]:
        for d in DIRS:
            Bm = math.sqrt(bsq * rho0)
            samples.append((rho0, t0, np.array([z0, 0, 0]), Bm * d))
    res = run_sweep(mink(), samples, 1e-8)
    assert len(res) == 18


def test_c2p_schwarzschild_hydro():
    g = schw_cks()
    res = run_sweep(g, hydro_samples(g), 1e-10)
    assert len(res) == 576


def test_c2p_sweep_detects_planted_error():
    """Planting an error in the oracle's inputs (a 1e-6 relative kick of tau)
    breaks the 1e-10 bound of the hydro sweep."""
    params = hybrid_params()
    prims, met, cons = build_states(params, mink(), hydro_samples(mink()))
    cons[4] *= 1 + 1e-6
    p1, _, _, _, _, _ = c2p_batch(params, cons, met)
    assert prim_residual(p1, prims).max() > 1e-10


def test_kastaun_root_matches_bisection():
    """The Kastaun master-function root found by Brent agrees with an
    independent bisection (ideal gas, Kerr-Schild c2p states)."""
    from grace_amd.workloads import c2p_states
    L = olib()
    st = c2p_states(200, seed=7)
    params = st["params"]
    nchk = 0
    for s in range(200):
        cons = st["cons"][:, s].copy()
        met = st["met"][:, s].copy()
        mu0 = C.c_double()
        f = lambda mu: L.orc_kastaun_f(C.byref(params), p(cons), p(met), mu, C.byref(mu0))
        f(0.5)
        a, b = 0.0, mu0.value
        fa, fb = f(a), f(b)
        if not (fa * fb < 0):
            continue
        for _ in range(200):
            m = 0.5 * (a + b)
            fm = f(m)
            if fa * fm <= 0:
                b, fb = m, fm
            else:
                a, fa = m, fm
        mu_bis = 0.5 * (a + b)
        # residual at the bisection root is ~0; the Brent root has |f| below the c2p tolerance
        assert abs(f(mu_bis)) < 1e-13
        nchk += 1
    assert nchk > 150


def _c4_states(n, seed):
    from grace_amd.workloads import c2p_states
    return c2p_states(n, seed=seed)


def test_master_functions_have_one_root_below_inv_h0():
    """The energy and entropy master functions change sign exactly once on
    (0, 1/h0] (dense grid, log- and linear-spaced, refined near 1/h0), so a
    sign-changing sub-bracket holds the same root as the plain solve."""
    L = olib()
    L.orc_master_grid.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int64, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
    st = _c4_states(150, seed=11)
    params = st["params"]
    roots = np.zeros((4, 150))
    c2p_batch(params, st["cons"], st["met"], r=st["rad"], roots=roots)
    grid = np.unique(np.concatenate([np.logspace(-10, 0, 700), np.linspace(1e-4, 1, 700),
                                     1 - np.logspace(-13, -1, 300)]))
    out = np.zeros_like(grid)
    h0 = C.c_double()
    checked = 0
    for s in range(150):
        cons = np.ascontiguousarray(st["cons"][:, s])
        met = np.ascontiguousarray(st["met"][:, s])
        for which in (0, 1):
            L.orc_master_grid(C.byref(params), p(cons), p(met), which, grid.size, p(grid), p(out), C.byref(h0))
            mus = grid / h0.value
            sg = np.sign(out)
            ch = np.nonzero(sg[1:] != sg[:-1])[0]
            assert ch.size == 1, (s, which, ch.size)
            x = roots[which, s]
            if x > 0:   # the plain root lies in the one sign-change cell
                assert mus[ch[0]] * h0.value <= x * (1 + 1e-12) and x <= mus[ch[0] + 1] * h0.value * (1 + 1e-12)
                checked += 1
    assert checked > 150


def test_guess_bracket_semantics():
    """A sign-changing guessed bracket (units of 1/h0) gives the plain root to
    the c2p tolerance and skips the mu+ Newton; a bracket that does not change
    sign, or an invalid one, leaves the plain path bit-identical."""
    n = 4000
    st = _c4_states(n, seed=5)
    params = st["params"]
    roots = np.zeros((4, n))
    p0, c0, e0, f0, ev0, _ = c2p_batch(params, st["cons"], st["met"], r=st["rad"], roots=roots)
    xe, xs = roots[0], roots[1]
    assert np.all(xe > 0) and np.all(roots[2] > 0)
    # good brackets around the plain roots
    g = np.zeros((4, n))
    g[0], g[1] = xe * (1 - 1e-3), np.minimum(xe * (1 + 1e-3), 1.0)
    ran_s = xs > 0
    g[2] = np.where(ran_s, xs * (1 - 1e-3), 0.0)
    g[3] = np.where(ran_s, np.minimum(xs * (1 + 1e-3), 1.0), 0.0)
    r1 = np.zeros((4, n))
    p1, c1, e1, f1, ev1, _ = c2p_batch(params, st["cons"], st["met"], r=st["rad"], guess=g, roots=r1)
    used_e = r1[2] == -1
    assert used_e.mean() > 0.99
    assert np.all(ev1[used_e, 1] <= ev0[used_e, 1])
    # both roots sit inside the Brent tolerance of a flat master function
    assert np.all(np.abs(r1[0] - xe) <= 1e-11 * xe)
    assert np.array_equal(e1, e0)
    assert prim_residual(p1, p0).max() < 1e-11
    # a bracket beside the root (no sign change) and an inverted one
    for bad in (np.stack([np.minimum(xe * 1.5, 0.999), np.full(n, 1.0), np.zeros(n), np.zeros(n)]),
                np.stack([xe * 1.001, xe * 0.999, xs * 1.001, xs * 0.999])):
        r2 = np.zeros((4, n))
        p2, c2, e2, f2, ev2, _ = c2p_batch(params, st["cons"], st["met"], r=st["rad"], guess=bad, roots=r2)
        same = (bad[0] >= bad[1]) | (bad[0] > xe)
        assert np.array_equal(p2[:, same], p0[:, same]) and np.array_equal(e2[same], e0[same])
        assert np.all(r2[2][same] == roots[2][same])

"""Pins of the oracle's point functions against the reference's own tests and
against closed forms / invariants (CPU only).

Reference tests restated: test/test_grmhd_kernels.cpp:80-469 (closed forms +
the two geometric-source golden vectors), test/test_metric.cpp:68-186.
"""
import ctypes as C
import math

import numpy as np
import pytest

from _oracle import olib, p, arr

L = olib()
FLAT = arr(1, 0, 0, 1, 0, 1)
ZERO3 = arr(0, 0, 0)
RHO, EPS, GAMMA = 1e-3, 1e-2, 2.0
PRESS = (GAMMA - 1) * RHO * EPS
H = 1 + EPS + PRESS / RHO
CS2 = GAMMA * PRESS / (RHO * H)


def W_of(g, z):
    return L.orc_W(p(g), p(z))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ---- test_grmhd_kernels.cpp:80-116 ----
def test_lorentz_factor():
    assert rel(W_of(FLAT, arr(0, 0, 0)), 1.0) < 1e-15
    assert rel(W_of(FLAT, arr(0.5, 0, 0)), math.sqrt(1.25)) < 1e-15
    g = arr(1 + 2 / 4.0, 0, 0, 1, 0, 1)
    z = arr(0.3, -0.2, 0.1)
    ex = math.sqrt(1 + g[0] * z[0] ** 2 + g[3] * z[1] ** 2 + g[5] * z[2] ** 2)
    assert rel(W_of(g, z), ex) < 1e-15


# ---- :119-160 ----
def test_vtilde():
    vt = np.zeros(3)
    L.orc_vtilde(1.0, p(ZERO3), p(ZERO3), 1.0, p(vt))
    assert np.all(np.abs(vt) < 1e-15)
    beta = arr(0.10, -0.05, 0.02)
    L.orc_vtilde(0.7, p(beta), p(ZERO3), 1.0, p(vt))
    assert np.allclose(vt, -beta, atol=1e-15, rtol=0)
    z = arr(0.4, -0.3, 0.2)
    W = math.sqrt(1 + (z ** 2).sum())
    L.orc_vtilde(0.6, p(beta), p(z), W, p(vt))
    for j in range(3):
        assert rel(vt[j], 0.6 * z[j] / W - beta[j]) < 1e-15


def fluxes(alp, beta, g, rho, press, eps, z, s, W, b2, b, vt, d):
    cons = np.zeros(6)
    fl = np.zeros(6)
    L.orc_fluxes(alp, p(beta), p(g), rho, press, eps, p(z), s, W, b2, p(b), p(vt), d, p(cons), p(fl))
    # cons: D, S0, S1, S2, tau, S*
    return cons, fl


# ---- :163-215 conserved ----
def test_conserved_static_and_boost():
    b = np.zeros(4)
    cons, _ = fluxes(1.0, ZERO3, FLAT, RHO, PRESS, EPS, arr(0, 0, 0), 0.0, 1.0, 0.0, b, ZERO3, 0)
    assert rel(cons[0], RHO) < 1e-15
    assert np.all(np.abs(cons[1:4]) < 1e-15)
    assert rel(cons[4], RHO * EPS) < 1e-13
    z = arr(0.5, 0, 0)
    W = math.sqrt(1.25)
    vt = np.zeros(3)
    L.orc_vtilde(1.0, p(ZERO3), p(z), W, p(vt))
    cons, fl = fluxes(1.0, ZERO3, FLAT, RHO, PRESS, EPS, z, 0.0, W, 0.0, b, vt, 0)
    assert rel(cons[0], RHO * W) < 1e-14
    assert rel(cons[1], RHO * H * W * z[0]) < 1e-13
    assert abs(cons[2]) < 1e-15 and abs(cons[3]) < 1e-15
    assert rel(cons[4], RHO * H * W * W - PRESS - RHO * W) < 1e-13
    # ---- :218-270 hydro fluxes ----
    assert rel(fl[0], RHO * W * vt[0]) < 1e-13


def test_fluxes_static():
    b = np.zeros(4)
    for d in range(3):
        _, fl = fluxes(1.0, ZERO3, FLAT, RHO, PRESS, EPS, arr(0, 0, 0), 0.0, 1.0, 0.0, b, ZERO3, d)
        assert abs(fl[0]) < 1e-15 and abs(fl[4]) < 1e-15
        for i in range(3):
            assert abs(fl[1 + i] - (PRESS if i == d else 0.0)) < 1e-13


# ---- :273-318 wavespeeds ----
def test_wavespeeds_hydro():
    cm, cp = C.c_double(), C.c_double()
    L.orc_cm_cp(1.0, p(ZERO3), RHO, PRESS, EPS, CS2, 1.0, 0.0, p(ZERO3), 1.0, 0, C.byref(cm), C.byref(cp))
    assert rel(cp.value, math.sqrt(CS2)) < 1e-12
    assert rel(cm.value, -math.sqrt(CS2)) < 1e-12
    z = arr(0.5, 0, 0)
    W = math.sqrt(1.25)
    vt = np.zeros(3)
    L.orc_vtilde(1.0, p(ZERO3), p(z), W, p(vt))
    L.orc_cm_cp(1.0, p(ZERO3), RHO, PRESS, EPS, CS2, W, 0.0, p(vt), 1.0, 0, C.byref(cm), C.byref(cp))
    assert cp.value > 0 and cm.value < cp.value and cm.value < vt[0] < cp.value
    # relativistic velocity addition along the flow: (v +- cs)/(1 +- v cs)
    v, cs = vt[0], math.sqrt(CS2)
    assert rel(cp.value, (v + cs) / (1 + v * cs)) < 1e-13
    assert rel(cm.value, (v - cs) / (1 - v * cs)) < 1e-13


def test_wavespeeds_alfven_limit():
    """Cold, strongly magnetised fluid at rest: fast speed -> Alfven speed
    vA^2 = b^2/(rho h + b^2) along any direction."""
    cm, cp = C.c_double(), C.c_double()
    b2 = 3.0 * RHO
    vA = math.sqrt(b2 / (RHO * (1 + EPS) + PRESS + b2))
    w2 = vA ** 2 + CS2 * (1 - vA ** 2)
    L.orc_cm_cp(1.0, p(ZERO3), RHO, PRESS, EPS, CS2, 1.0, b2, p(ZERO3), 1.0, 1, C.byref(cm), C.byref(cp))
    assert rel(cp.value, math.sqrt(w2)) < 1e-13 and rel(cm.value, -math.sqrt(w2)) < 1e-13


# ---- :321-390 geometric sources vanish on Minkowski ----
def geom(alp, beta, g, gu, K, da, dg, db, rho, press, eps, B, z, W):
    out = np.zeros(4)
    L.orc_geom_sources(alp, p(beta), p(g), p(gu), p(K), p(da), p(dg), p(db), rho, press, eps, p(B), p(z), W, p(out))
    return out


def test_geom_sources_minkowski():
    z0 = np.zeros(6)
    out = geom(1.0, ZERO3, FLAT, FLAT, z0, ZERO3, np.zeros(18), np.zeros(9), RHO, PRESS, EPS, ZERO3, ZERO3, 1.0)
    assert np.all(np.abs(out) < 1e-14)
    z = arr(0.4, -0.3, 0.2)
    B = arr(0.01, -0.02, 0.03)
    out = geom(1.0, ZERO3, FLAT, FLAT, z0, ZERO3, np.zeros(18), np.zeros(9), RHO, PRESS, EPS, B, z, W_of(FLAT, z))
    assert np.all(np.abs(out) < 1e-13)


# ---- :393-469 golden vectors (rel 1e-12) ----
GOLD_IN = dict(
    gdd=arr(+1.19999999999999996e+00, +1.00000000000000006e-01, -5.00000000000000028e-02, +1.14999999999999991e+00,
            +8.00000000000000017e-02, +1.17999999999999994e+00),
    guu=arr(+8.41366894150114231e-01, -7.60008596818554372e-02, +4.08037402390289405e-02, +8.80550943936906916e-01,
            -6.29187444907163829e-02, +8.53452276755261696e-01),
    Kdd=arr(+7.32688520015280093e-02, +2.03453496023656782e-01, -1.75658814405620894e-01, -2.72601936937604428e-01,
            -5.97551463336409405e-01, +2.91486894595860413e-01),
    betau=arr(+1.00000000000000006e-01, -5.00000000000000028e-02, +2.50000000000000014e-02),
    dgdd=arr(+8.32865028802352858e-03, +1.02865672473414063e-01, -3.91797438252806718e-01, +6.13249014642980317e-01,
             +4.71600288839008341e-01, -6.09119407438451152e-02, -2.77931728951141155e-01, -1.78345692029290043e-01,
             -3.99312771212324613e-01, +1.32820746202705325e-02, -3.08083859850273589e-01, +2.92771260271021072e-01,
             -1.79734287329371828e-01, -1.76978531969679137e-01, +7.68626502100995701e-01, -3.62859761276027648e-01,
             -7.83957943972732330e-01, +1.83107596203723882e-01),
    dbetau=arr(-1.03893425570468656e+00, +1.32189317912860316e-01, +4.90860070946320948e-01, +3.50575153770565986e-02,
               +4.63753714810319351e-03, -1.02530467612477705e+00, +2.77651842779394309e-01, +1.98146283623305774e-01,
               -5.85998879187250465e-01),
    dalp=arr(+2.86975612086930454e+00, -1.83602430111033832e+00, -2.51893554257572383e+00),
    z=arr(+5.00000000000000000e-01, -1.00000000000000006e-01, +5.00000000000000028e-02),
)
GOLD = [
    (arr(0, 0, 0), -2.46366540636220083e-05, [-7.82000088795445234e-05, +4.33711030551046470e-05, +6.21362117458239347e-05]),
    (arr(1.0e-3, -8.0e-4, 3.0e-4), -2.48104573815721489e-05,
     [-8.21445746274834321e-05, +4.58890193752384476e-05, +6.54992152417176692e-05]),
]


@pytest.mark.parametrize("case", [0, 1])
def test_geom_sources_golden(case):
    B, tau_ref, s_ref = GOLD[case]
    g = GOLD_IN
    out = geom(0.5, g["betau"], g["gdd"], g["guu"], g["Kdd"], g["dalp"], g["dgdd"], g["dbetau"], 1e-5, 1e-6, 1e-3,
               B, g["z"], 1.5)
    assert rel(out[0], tau_ref) < 1e-12
    for i in range(3):
        assert rel(out[1 + i], s_ref[i]) < 1e-12


def test_geom_sources_golden_detects_sign_flip():
    """A plausible mistake (dropped lapse factor on the d_i alp term) fails the pin."""
    g = GOLD_IN
    B, tau_ref, _ = GOLD[0]
    out = geom(0.5, g["betau"], g["gdd"], g["guu"], g["Kdd"], 2 * g["dalp"], g["dgdd"], g["dbetau"], 1e-5, 1e-6, 1e-3,
               B, g["z"], 1.5)
    assert rel(out[0], tau_ref) > 1e-3


# ---- metric_utils / test_metric.cpp ----
def test_metric_inverse():
    g = GOLD_IN["gdd"]
    gi = np.zeros(6)
    sq = C.c_double()
    L.orc_metric(p(g), p(gi), C.byref(sq))
    G = np.array([[g[0], g[1], g[2]], [g[1], g[3], g[4]], [g[2], g[4], g[5]]])
    Gi = np.array([[gi[0], gi[1], gi[2]], [gi[1], gi[3], gi[4]], [gi[2], gi[4], gi[5]]])
    assert np.allclose(G @ Gi, np.eye(3), atol=1e-15)
    assert abs(sq.value - math.sqrt(np.linalg.det(G))) < 1e-15
    assert np.allclose(gi, GOLD_IN["guu"], rtol=1e-14, atol=1e-16)
    L.orc_metric(p(FLAT), p(gi), C.byref(sq))
    assert np.array_equal(gi, FLAT) and sq.value == 1.0


# ---- reconstruction (weno_reconstruction.hh, reconstruction.hh, limiters.hh) ----
def rec(kind, u):
    out = np.zeros(2)
    L.orc_recon(kind, p(np.asarray(u, dtype=np.float64)), p(out))
    return out


def test_wenoz_constant_and_linear_exact():
    assert np.array_equal(rec(0, [2.5] * 6), [2.5, 2.5])
    x = np.arange(-3, 3) + 0.5  # cell centres relative to face at 0
    for a, b in [(1.0, 0.3), (-2.0, 1.7)]:
        u = a + b * x
        out = rec(0, u)
        assert abs(out[0] - a) < 1e-14 and abs(out[1] - a) < 1e-14


def test_wenoz_fifth_order_on_smooth_data():
    """Cell averages of sin(x): face error of WENO-Z decays at 5th order."""
    errs = []
    for h in [0.1, 0.05, 0.025]:
        edges = (np.arange(-3, 4)) * h + 0.3
        avg = (np.cos(edges[:-1]) - np.cos(edges[1:])) / h  # exact cell averages of sin
        out = rec(0, avg)
        errs.append(max(abs(out[0] - math.sin(0.3)), abs(out[1] - math.sin(0.3))))
    r1 = math.log2(errs[0] / errs[1])
    r2 = math.log2(errs[1] / errs[2])
    assert r1 > 4.5 and r2 > 4.5


def test_wenoz_mirror_symmetry():
    rng = np.random.default_rng(3)
    for _ in range(50):
        u = rng.normal(size=6)
        a = rec(0, u)
        # mirror about the face i-1/2: the cell centres at -2.5..2.5 map onto themselves reversed
        m = u[::-1].copy()
        b = rec(0, m)
        assert abs(a[0] - b[1]) < 1e-14 * max(1, abs(a[0])) and abs(a[1] - b[0]) < 1e-14 * max(1, abs(a[1]))


def test_wenoz_hand_case():
    """Hand-computed WENO-Z value on a discontinuity u = (0,0,0,1,1,1)."""
    out = rec(0, [0, 0, 0, 1, 1, 1])
    # left side: the smooth stencil {0,0,0} takes all but ~(1e-42)^2 of the weight
    assert abs(out[0]) < 1e-60 and abs(out[1] - 1.0) < 1e-60


def test_mc2_minmod_donor_closed_forms():
    u = [0, 1.0, 2.0, 4.0, 5.0, 0]
    # donor
    assert np.array_equal(rec(3, u), [2.0, 4.0])
    # MC2 beta=2 at i (slopes 2 and 1): min(2*|1|, 2*2, 1.5) = 1.5 -> uR = 4 - 0.75
    # at im (slopes 1 and 2): min(2*2, 2*1, 1.5) = 1.5 -> uL = 2 + 0.75
    assert np.allclose(rec(1, u), [2.75, 3.25], atol=0, rtol=1e-15)
    # minmod: at i min(2,1)=1 -> 3.5 ; at im min(1,2)=1 -> 2.5
    assert np.allclose(rec(2, u), [2.5, 3.5], atol=0, rtol=1e-15)
    # extremum -> zero slope
    assert np.allclose(rec(1, [0, 1.0, 3.0, 2.0, 5.0, 0]), [3.0, 2.0])


# ---- Brent (rootfinding.hh:73-231) ----
def test_brent_roots():
    n = C.c_int64()
    for r0 in [0.1, 0.37, 0.9]:
        x = L.orc_brent_poly(r0, 0.5, 0.0, 1.0, 1e-15, C.byref(n))
        assert abs(x - r0) < 4e-15
        assert n.value < 20

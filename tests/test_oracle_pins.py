"""Mathematics-based pins of the oracle functions the reference's own tests
leave open (CPU only).

The reference pins its GRMHD fluxes and wave speeds only for B = 0
(test/test_grmhd_kernels.cpp:211-313) and has no test of the edge EMFs,
the face interpolation or the substep assembly with B != 0.  Each check
below computes the expected value by a route independent of the oracle's
closed forms, so that a plausible slip in the oracle (a sign of a b^mu term,
a transposed metric index, swapped HLL speeds, a wrong stencil weight, a
dropped source term) fails it:

* fluxes / conserved (grmhd_subexpressions.hh:97-156, :318-350): built here
  from the 4-metric, u^mu, the projected b^mu and
  T^{mu nu} = (rho h + b^2) u u + (p + b^2/2) g^{-1} - b b with numpy;
* wave speeds (grmhd_subexpressions.hh:60-95): roots of the covariant
  dispersion relation (1 - w^2)(k.u)^2 = w^2 k.k for k = (-lambda, e_d), and
  relativistic velocity addition in flat space;
* UCT and GS edge EMFs (evolve.cpp:1069-1340, grmhd_helpers.hh:492-627):
  E = -vtilde x B for a uniform state; on data varying along one axis both
  reduce to the 1-D HLL flux of the induction equation; on a flow
  supersonic along two axes both take the E field of the upwind corner cell;
* fd_der (fd_subexpressions.hh:21-57) and face_interp_4
  (grmhd_helpers.hh:360-380): polynomial exactness and convergence order;
* the substep assembly with B != 0: the circularly polarised Alfven wave
  (an exact nonlinear solution) converges at the scheme's order; Balsara-1
  conserves D, tau and B^y and gains S_x at the rate the boundary stresses
  fix; in the Kerr-Schild metric the Fishbone-Moncrief torus (B = 0, an
  exact stationary solution) has a residual right-hand side that falls with
  resolution.
"""
import ctypes as C
import math

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O
from _oracle import olib, p, arr

L = olib()
SIDX = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


# ---------------------------------------------------------------------------
# independent 4-tensor algebra
# ---------------------------------------------------------------------------
def sym3(g6):
    g = np.zeros((3, 3))
    for a, (i, j) in enumerate(SIDX):
        g[i, j] = g[j, i] = g6[a]
    return g


def four_metric(alp, beta, gam):
    bl = gam @ beta
    g4 = np.zeros((4, 4))
    g4[0, 0] = -alp * alp + beta @ bl
    g4[0, 1:] = g4[1:, 0] = bl
    g4[1:, 1:] = gam
    return g4, np.linalg.inv(g4)


def fluid(alp, beta, gam, rho, press, eps, z, B):
    """u^mu, b^mu, b^2 and T^{mu nu} from 4-tensors (B: Eulerian field, undensitised)."""
    g4, g4i = four_metric(alp, beta, gam)
    n_up = g4i @ np.array([-alp, 0.0, 0.0, 0.0])
    Wl = math.sqrt(1.0 + z @ gam @ z)
    u = Wl * n_up + np.concatenate([[0.0], z])
    Bv = np.concatenate([[0.0], B])
    u_low = g4 @ u
    b = (Bv + u * (u_low @ Bv)) / Wl
    b2 = b @ g4 @ b
    rhoh = rho * (1.0 + eps) + press
    T = (rhoh + b2) * np.outer(u, u) + (press + 0.5 * b2) * g4i - np.outer(b, b)
    return dict(g4=g4, g4i=g4i, u=u, b=b, b2=b2, T=T, W=Wl, rhoh=rhoh)


def conserved_and_flux(alp, beta, gam, rho, press, eps, z, B, s, d):
    f = fluid(alp, beta, gam, rho, press, eps, z, B)
    T, g4, u = f["T"], f["g4"], f["u"]
    D = rho * f["W"]
    S = np.array([alp * (T[0] @ g4[:, 1 + j]) for j in range(3)])
    tau = alp * alp * T[0, 0] - D
    vt = u[1 + d] / u[0]
    FD = D * vt
    FS = np.array([alp * (T[1 + d] @ g4[:, 1 + j]) for j in range(3)])
    Ftau = alp * alp * T[0, 1 + d] - FD
    return np.array([D, S[0], S[1], S[2], tau, s * D]), np.array([FD, FS[0], FS[1], FS[2], Ftau, s * FD]), f


def char_speeds(alp, beta, gam, rho, press, eps, z, B, gamma_eos, d):
    """Fast speeds along d from (1 - w^2)(k.u)^2 = w^2 k_mu k_nu g^{mu nu},
    k = (-lambda, e_d), w^2 = vA^2 + cs^2 (1 - vA^2)."""
    f = fluid(alp, beta, gam, rho, press, eps, z, B)
    u, g4i = f["u"], f["g4i"]
    cs2 = gamma_eos * press / f["rhoh"]
    vA2 = f["b2"] / (f["rhoh"] + f["b2"])
    w2 = vA2 + cs2 * (1.0 - vA2)
    a2 = (1 - w2) * u[0] ** 2 - w2 * g4i[0, 0]
    a1 = -2 * (1 - w2) * u[0] * u[1 + d] + 2 * w2 * g4i[0, 1 + d]
    a0 = (1 - w2) * u[1 + d] ** 2 - w2 * g4i[1 + d, 1 + d]
    r = np.sort(np.roots([a2, a1, a0]).real)
    return r[0], r[1], cs2


def random_state(rng, curved=True):
    if curved:
        A = rng.normal(size=(3, 3)) * 0.25
        gam = np.eye(3) + 0.5 * (A + A.T) + 0.3 * A @ A.T
        assert np.all(np.linalg.eigvalsh(gam) > 0.2)
        alp = rng.uniform(0.4, 1.5)
        beta = rng.normal(size=3) * 0.15
    else:
        gam, alp, beta = np.eye(3), 1.0, np.zeros(3)
    rho = 10 ** rng.uniform(-3, 0)
    eps = 10 ** rng.uniform(-3, 0.5)
    gm = 4.0 / 3.0
    press = (gm - 1) * rho * eps
    z = rng.normal(size=3) * rng.uniform(0.1, 2.0)
    B = rng.normal(size=3) * math.sqrt(rho) * rng.uniform(0.1, 3.0)
    return gam, alp, beta, rho, press, eps, z, B, gm


def oracle_point(gam, alp, beta, rho, press, eps, z, B, s, d):
    g6 = np.array([gam[i, j] for i, j in SIDX])
    Wl = L.orc_W(p(g6), p(z))
    b = np.zeros(4)
    b2 = C.c_double()
    L.orc_smallb(alp, p(beta), p(g6), p(B), p(z), Wl, p(b), C.byref(b2))
    vt = np.zeros(3)
    L.orc_vtilde(alp, p(beta), p(z), Wl, p(vt))
    cons = np.zeros(6)
    fl = np.zeros(6)
    L.orc_fluxes(alp, p(beta), p(g6), rho, press, eps, p(z), s, Wl, b2.value, p(b), p(vt), d, p(cons), p(fl))
    # reorder oracle (D, S0, S1, S2, tau, S*) -> same order as conserved_and_flux
    return cons, fl, Wl, b, b2.value, vt, g6


# ---------------------------------------------------------------------------
# point algebra with B != 0 and v != 0 in a generic curved metric
# ---------------------------------------------------------------------------
def test_fluxes_and_conserved_from_stress_energy_tensor():
    rng = np.random.default_rng(7)
    worst = 0.0
    for trial in range(300):
        gam, alp, beta, rho, press, eps, z, B, gm = random_state(rng, curved=trial % 5 != 0)
        s = rng.uniform(-2, 2)
        for d in range(3):
            ref_c, ref_f, f = conserved_and_flux(alp, beta, gam, rho, press, eps, z, B, s, d)
            cons, fl, Wl, b, b2, vt, _ = oracle_point(gam, alp, beta, rho, press, eps, z, B, s, d)
            # u.u = -1 and b.u = 0 hold for the independent construction
            assert abs(f["u"] @ f["g4"] @ f["u"] + 1) < 1e-12 * f["W"] ** 2
            assert abs(f["b"] @ f["g4"] @ f["u"]) < 1e-12 * (1 + np.abs(f["b"]).max() * f["W"])
            scale = f["rhoh"] * f["W"] ** 2 + press + f["b2"] * f["W"] ** 2 + rho * f["W"]
            assert abs(Wl - f["W"]) < 1e-14 * f["W"]
            assert np.allclose(b, f["b"], rtol=0, atol=1e-13 * (1 + np.abs(f["b"]).max()))
            assert abs(b2 - f["b2"]) < 1e-13 * (1 + f["b2"])
            assert np.allclose(vt, [f["u"][1 + k] / f["u"][0] for k in range(3)], rtol=0, atol=1e-14 * (1 + alp))
            err = max(np.abs(cons - ref_c).max(), np.abs(fl - ref_f).max()) / scale
            worst = max(worst, err)
    assert worst < 1e-12, worst


def test_wave_speeds_dispersion_relation():
    rng = np.random.default_rng(11)
    worst = 0.0
    for trial in range(300):
        gam, alp, beta, rho, press, eps, z, B, gm = random_state(rng, curved=trial % 4 != 0)
        for d in range(3):
            lm, lp, cs2 = char_speeds(alp, beta, gam, rho, press, eps, z, B, gm, d)
            _, _, Wl, b, b2, vt, g6 = oracle_point(gam, alp, beta, rho, press, eps, z, B, 0.0, d)
            gi = np.linalg.inv(gam)
            cm, cp = C.c_double(), C.c_double()
            L.orc_cm_cp(alp, p(beta), rho, press, eps, cs2, Wl, b2, p(vt), gi[d, d], d, C.byref(cm), C.byref(cp))
            sc = 1.0 + abs(lm) + abs(lp)
            worst = max(worst, abs(cm.value - lm) / sc, abs(cp.value - lp) / sc)
    assert worst < 1e-12, worst


def test_wave_speeds_velocity_addition_flat():
    """Flat space: along the flow lambda = (v -+ w)/(1 -+ v w); across it
    lambda = +-w sqrt((1 - v^2)/(1 - v^2 w^2))."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        rho, gm = 10 ** rng.uniform(-2, 0), 4.0 / 3.0
        press = (gm - 1) * rho * 10 ** rng.uniform(-2, 0)
        eps = press / ((gm - 1) * rho)
        d = int(rng.integers(3))
        v = rng.uniform(-0.95, 0.95)
        e = np.zeros(3); e[d] = 1.0
        # B parallel to v: vA enters w through b^2 only
        B = e * math.sqrt(rho) * rng.uniform(0, 2)
        for along in (True, False):
            vel = e * v if along else np.roll(e, 1) * v
            Bf = B if along else np.roll(e, 1) * B[d]
            Wl = 1 / math.sqrt(1 - v * v)
            z = Wl * vel
            b2 = (Bf @ Bf) / Wl ** 2 + (Bf @ z) ** 2 / Wl ** 2
            h = 1 + eps + press / rho
            cs2 = gm * press / (rho * h)
            vA2 = b2 / (rho * h + b2)
            w = math.sqrt(vA2 + cs2 * (1 - vA2))
            if along:
                ex = sorted([(v - w) / (1 - v * w), (v + w) / (1 + v * w)])
            else:
                a = w * math.sqrt((1 - v * v) / (1 - v * v * w * w))
                ex = [-a, a]
            vt = np.zeros(3)
            L.orc_vtilde(1.0, p(np.zeros(3)), p(z), Wl, p(vt))
            cm, cp = C.c_double(), C.c_double()
            L.orc_cm_cp(1.0, p(np.zeros(3)), rho, press, eps, cs2, Wl, b2, p(vt), 1.0, d, C.byref(cm), C.byref(cp))
            assert abs(cm.value - ex[0]) < 1e-13 and abs(cp.value - ex[1]) < 1e-13, (along, cm.value, cp.value, ex)


# ---------------------------------------------------------------------------
# FD derivative and 4-point face interpolation
# ---------------------------------------------------------------------------
def fd_oracle(order, u, ih):
    return L.orc_fd_der(order, p(np.ascontiguousarray(u, dtype=np.float64)), ih)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_fd_der_exact_on_polynomials_and_converges(order):
    rng = np.random.default_rng(order)
    h = 0.37
    xs = np.arange(-3, 4) * h
    for deg in range(order + 2):
        c = rng.normal(size=deg + 1)
        u = np.polyval(c, xs + 0.3) if deg else np.full(7, c[0])
        exact = np.polyval(np.polyder(c), 0.3) if deg else 0.0
        err = abs(fd_oracle(order, u, 1 / h) - exact)
        if deg <= order:
            assert err < 1e-12 * (1 + np.abs(c).sum()), (order, deg, err)
        elif deg == order + 1:
            assert err > 1e-8, (order, deg, err)   # exactly the claimed order, not higher
    # convergence on sin: error ratio 2^order when h halves
    errs = []
    for hh in (0.2, 0.1, 0.05):
        u = np.sin(1.3 + np.arange(-3, 4) * hh)
        errs.append(abs(fd_oracle(order, u, 1 / hh) - math.cos(1.3)))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(abs(r - order) < 0.15 for r in rates), rates


def test_fd_der_check_fails_on_a_planted_wrong_weight():
    """The exactness check above rejects a stencil with one wrong weight
    (the 4th-order weight 8 replaced by 7), computed here in numpy."""
    h = 0.37
    xs = np.arange(-3, 4) * h
    u = xs ** 3 + xs
    bad = (1.0 / 12.0) / h * (7.0 * (u[4] - u[2]) - (u[5] - u[1]))
    assert abs(bad - 1.0) > 1e-3
    assert abs(fd_oracle(4, u, 1 / h) - 1.0) < 1e-13


def test_face_interp_exact_on_cubics():
    rng = np.random.default_rng(5)
    for _ in range(50):
        c = rng.normal(size=4)
        x0 = rng.uniform(-2, 2)
        h = rng.uniform(0.05, 1.0)
        xs = x0 + np.array([-1.5, -0.5, 0.5, 1.5]) * h
        u = np.polyval(c, xs)
        out = L.orc_face4(p(u))
        assert abs(out - np.polyval(c, x0)) < 1e-13 * (1 + np.abs(u).max())
    # a quartic is not reproduced (fourth order, not higher)
    u = (np.array([-1.5, -0.5, 0.5, 1.5])) ** 4
    assert abs(L.orc_face4(p(u))) > 0.5
    # mirrored data give the bit-identical value (pair-symmetric sums)
    u = rng.normal(size=4)
    assert L.orc_face4(p(u)) == L.orc_face4(p(u[::-1].copy()))


# ---------------------------------------------------------------------------
# single-block workloads from primitive fields
# ---------------------------------------------------------------------------
def block_from_prims(n, ngz, idx, bc, params, prim_fn, bface_fn, metric=None):
    """One block of n = (nx, ny, nz) interior cells.  prim_fn(x, y, z) ->
    (rho, press, z^i (3,)) on cell centres (arrays); bface_fn(d, x, y, z) ->
    undensitised B^d on the d-faces.  metric: (gam6, alp, beta) constant in
    space (default flat).  Conserved from the oracle's prims_to_conservs."""
    grid = ga.Grid(n[0], n[1], n[2], ngz, 1)
    spec = [dict(nb=(1, 1, 1), q0=0, bc=bc, idx=idx)]
    state, B, aux = W.alloc(grid)
    N = [n[d] + 2 * ngz for d in range(3)]
    gam6, alp, beta = metric if metric is not None else (arr(1, 0, 0, 1, 0, 1), 1.0, np.zeros(3))
    sqrtg = math.sqrt(np.linalg.det(sym3(gam6)))
    for a in range(6):
        state[:, ga.GXX + a] = gam6[a]
    state[:, ga.ALP] = alp
    for a in range(3):
        state[:, ga.BETAX + a] = beta[a]
    cc = [(np.arange(N[d]) - ngz + 0.5) / idx[d] for d in range(3)]
    fc = [(np.arange(N[d] + 1) - ngz) / idx[d] for d in range(3)]
    zc, yc, xc = np.meshgrid(cc[2], cc[1], cc[0], indexing="ij")
    rho, press, zv = prim_fn(xc, yc, zc)
    for d in range(3):
        X = [cc[0], cc[1], cc[2]]
        X[d] = fc[d]
        zz, yy, xx = np.meshgrid(X[2], X[1], X[0], indexing="ij")
        B[d][0, 0] = sqrtg * bface_fn(d, xx, yy, zz)
    gm = params.gamma
    eps = press / ((gm - 1) * rho)
    Bc = [0.5 * (B[0][0, 0][:, :, :-1] + B[0][0, 0][:, :, 1:]) / sqrtg,
          0.5 * (B[1][0, 0][:, :-1, :] + B[1][0, 0][:, 1:, :]) / sqrtg,
          0.5 * (B[2][0, 0][:-1] + B[2][0, 0][1:]) / sqrtg]
    aux[0, ga.RHO] = rho; aux[0, ga.PRESS] = press; aux[0, ga.EPS] = eps; aux[0, ga.TEMP] = eps * (gm - 1)
    aux[0, ga.ENTROPY] = np.log(eps) / (gm - 1) - np.log(rho)
    for a in range(3):
        aux[0, ga.ZVECX + a] = zv[a]
        aux[0, ga.BX + a] = Bc[a]
    nc = state[0, 0].size
    prims = np.zeros((13, nc))
    prims[0] = rho.ravel(); prims[1] = press.ravel()
    for a in range(3):
        prims[2 + a] = np.broadcast_to(zv[a], rho.shape).ravel()
        prims[9 + a] = Bc[a].ravel()
    prims[6] = (eps * (gm - 1)).ravel(); prims[7] = eps.ravel()
    prims[8] = (np.log(eps) / (gm - 1) - np.log(rho)).ravel()
    met = np.zeros((10, nc))
    for a in range(6):
        met[a] = gam6[a]
    met[6] = alp
    for a in range(3):
        met[7 + a] = beta[a]
    cons = np.zeros((10, nc))
    L.orc_p2c_batch(C.byref(params), nc, p(prims), p(met), p(cons))
    for v in range(7):
        state[0, v] = cons[v].reshape(state[0, v].shape)
    return W.Workload("block", params, grid, spec, state, B, aux, zones=n[0] * n[1] * n[2],
                      extra=dict(sqrtg=sqrtg, gam6=gam6, alp=alp, beta=beta))


def run_substep(wl, dt=1e-4):
    S = O.Stepper(wl)
    st, B, aux = wl.copy_arrays()
    S.ghosts(st, B)
    S.aux(st, B, aux)
    new, nB = st.copy(), [b.copy() for b in B]
    S.substep(dt, 1.0, new, st, nB, B, aux)
    g = wl.grid
    Ni, Nj, Nk = g.nx + 2 * g.ngz, g.ny + 2 * g.ngz, g.nz + 2 * g.ngz
    emf = S.scratch(2).reshape(1, 3, Nk + 1, Nj + 1, Ni + 1)
    return S, st, B, aux, emf


def cell_E(wl, aux):
    """-vtilde x Btilde per cell from the primitives (Btilde = sqrt(gamma) B)."""
    ex = wl.extra
    gam = sym3(ex["gam6"])
    z = np.stack([aux[0, ga.ZVECX + a] for a in range(3)])
    Wl = np.sqrt(1 + np.einsum("i...,ij,j...->...", z, gam, z))
    vt = ex["alp"] * z / Wl - ex["beta"].reshape(3, 1, 1, 1)
    Bt = ex["sqrtg"] * np.stack([aux[0, ga.BX + a] for a in range(3)])
    return -np.cross(vt, Bt, axis=0), vt


def hll_1d(wl, aux, D, sl_L, sl_R):
    """Textbook 1-D HLL flux of the induction equation along D,
    F^i = vt^D Bt^i - vt^i Bt^D, between donor states sl_L and sl_R (cell
    index slices), with the wave speeds from the dispersion relation."""
    ex = wl.extra
    gam = sym3(ex["gam6"])
    gm = wl.params.gamma

    def state(sl):
        rho, press, eps = aux[0, ga.RHO][sl], aux[0, ga.PRESS][sl], aux[0, ga.EPS][sl]
        z = np.array([aux[0, ga.ZVECX + a][sl] for a in range(3)])
        Bu = np.array([aux[0, ga.BX + a][sl] for a in range(3)])
        lm, lp, _ = char_speeds(ex["alp"], ex["beta"], gam, rho, press, eps, z, Bu, gm, D)
        Wl = math.sqrt(1 + z @ gam @ z)
        vt = ex["alp"] * z / Wl - ex["beta"]
        Bt = ex["sqrtg"] * Bu
        F = vt[D] * Bt - vt * Bt[D]
        return lm, lp, F, Bt

    lmL, lpL, FL, BL = state(sl_L)
    lmR, lpR, FR, BR = state(sl_R)
    cmin, cmax = -min(0.0, lmL, lmR), max(0.0, lpL, lpR)
    return (cmax * FL + cmin * FR - cmax * cmin * (BR - BL)) / (cmax + cmin)


CURVED = (arr(1.3, 0.1, -0.05, 1.1, 0.08, 0.9), 0.8, arr(0.05, -0.1, 0.07))


@pytest.mark.parametrize("emf", [ga.EMF_UCT, ga.EMF_GS])
def test_edge_emf_uniform_state(emf):
    """A uniform moving state in a constant curved metric: every edge EMF is
    -vtilde x Btilde."""
    params = ga.default_params(gamma=4.0 / 3.0, emf=emf, flat_metric=0)
    z0, B0 = arr(0.3, -0.5, 0.2), arr(0.4, -0.2, 0.7)
    wl = block_from_prims((6, 6, 6), 3, (6.0,) * 3, (ga.BC_PERIODIC,) * 6, params,
                          lambda x, y, z: (np.ones_like(x), 0.5 * np.ones_like(x), z0.reshape(3, 1, 1, 1)),
                          lambda d, x, y, z: B0[d] * np.ones_like(x), metric=CURVED)
    S, st, B, aux, emf_a = run_substep(wl)
    Ec, _ = cell_E(wl, aux)
    ng, n = 3, 6
    sl = slice(ng, ng + n)
    for c in range(3):
        got = emf_a[0, c, sl, sl, sl]
        assert np.allclose(got, Ec[c, 3, 3, 3], rtol=0, atol=1e-13 * np.abs(Ec[:, 3, 3, 3]).max()), c


@pytest.mark.parametrize("emf", [ga.EMF_UCT, ga.EMF_GS])
@pytest.mark.parametrize("D", [0, 1, 2])
def test_edge_emf_reduces_to_1d_hll(emf, D):
    """Data varying along D only, donor-cell reconstruction: the two edge
    components transverse to D, at every D-face, equal the 1-D HLL flux of
    the induction equation (E_{D+2} = -F^{D+1}, E_{D+1} = F^{D+2}); the
    component along D equals the cell's -vtilde x Btilde."""
    params = ga.default_params(gamma=4.0 / 3.0, emf=emf, flat_metric=0, recon=ga.RECON_DONOR)
    n, ng = 8, 3
    k = 2 * math.pi / n * 8.0 / 8.0

    def coord(x, y, z):
        return (x, y, z)[D]

    def prims(x, y, z):
        s = coord(x, y, z) * 2 * math.pi
        rho = 1.0 + 0.5 * np.sin(s)
        press = 0.3 + 0.2 * np.cos(2 * s)
        zv = np.stack([0.4 * np.sin(s + a) + 0.2 * (a - 1) for a in range(3)])
        return rho, press, zv

    def bface(d, x, y, z):
        if d == D:
            return 0.6 * np.ones_like(x)   # div B = 0: the normal component is constant
        s = coord(x, y, z) * 2 * math.pi
        return 0.5 * np.cos(s + d) + 0.3 * (d - 1)

    wl = block_from_prims((n, n, n), ng, (float(n),) * 3, (ga.BC_PERIODIC,) * 6, params, prims, bface,
                          metric=CURVED)
    S, st, B, aux, emf_a = run_substep(wl)
    Ec, _ = cell_E(wl, aux)
    c1, c2 = (D + 1) % 3, (D + 2) % 3
    worst = 0.0
    for f in range(ng, ng + n + 1):          # D-faces between cells f-1 and f
        def csl(i):
            s = [ng, ng, ng]
            s[D] = i
            return (s[2], s[1], s[0])
        F = hll_1d(wl, aux, D, csl(f - 1), csl(f))
        for comp, want in ((c2, -F[c1]), (c1, F[c2])):
            # edges of component comp at D-face f: index f along D, the other
            # transverse index over the interior edge range
            idx = [slice(ng, ng + n + 1)] * 3
            idx[D] = f
            idx[comp] = slice(ng, ng + n)
            got = emf_a[0, comp][idx[2], idx[1], idx[0]]
            worst = max(worst, float(np.abs(got - want).max()) / (1 + abs(want)))
    assert worst < 1e-13, worst
    # the component along D at every cell position along D (uniform across):
    # the cell value
    for i in range(n):
        e = [ng + 1, ng + 2, ng + 3]
        e[D] = ng + i
        cc = [ng, ng, ng]
        cc[D] = ng + i
        e_val = emf_a[0, D][e[2], e[1], e[0]]
        c_val = Ec[D][cc[2], cc[1], cc[0]]
        assert abs(e_val - c_val) < 1e-13 * (1 + abs(c_val)), (i, e_val, c_val)


@pytest.mark.parametrize("emf", [ga.EMF_UCT, ga.EMF_GS])
@pytest.mark.parametrize("plane", [(0, 1), (1, 2), (2, 0)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_edge_emf_supersonic_upwind_corner(emf, plane, sign):
    """A cold, weakly magnetised flow supersonic along both axes of a plane,
    donor-cell reconstruction: the edge EMF normal to the plane equals the E
    field of the upwind corner cell, (i-1, j-1) for flow towards +, (i, j)
    for flow towards - (both schemes upwind this way; swapped HLL speeds,
    L/R states or upwind signs give a neighbouring cell instead)."""
    a, b = plane
    c = 3 - a - b
    params = ga.default_params(gamma=4.0 / 3.0, emf=emf, flat_metric=1, recon=ga.RECON_DONOR)
    n, ng = 6, 3

    def prims(x, y, z):
        X = (x, y, z)
        rho = 1.0 + 0.3 * np.sin(2 * math.pi * X[a]) * np.cos(2 * math.pi * X[b])
        press = 1e-4 * (1 + 0.5 * np.cos(2 * math.pi * (X[a] + X[b])))
        # velocity varying across the plane, so that E is not a sum of a
        # function of X_a and one of X_b (for such an E every upwind choice
        # of the transverse derivatives gives the same edge value)
        s = np.sin(2 * math.pi * X[a]) * np.cos(2 * math.pi * X[b])
        vel = np.zeros((3,) + x.shape)
        vel[a] = sign * (0.55 + 0.05 * s)
        vel[b] = sign * (0.5 - 0.04 * s)
        vel[c] = 0.1 + 0.03 * s
        Wl = 1 / np.sqrt(1 - (vel ** 2).sum(axis=0))
        return rho, press, Wl * vel

    def bface(d, x, y, z):
        X = (x, y, z)
        # B^a depends on X_b only and B^b on X_a only: face values equal the
        # neighbouring cell averages along the plane
        if d == a:
            return 1e-3 * (1 + 0.5 * np.sin(2 * math.pi * X[b]))
        if d == b:
            return 1e-3 * (0.5 + 0.5 * np.cos(2 * math.pi * X[a]))
        return 1e-3 * (0.3 + 0.4 * np.sin(2 * math.pi * (X[a] - X[b])))

    wl = block_from_prims((n, n, n), ng, (float(n),) * 3, (ga.BC_PERIODIC,) * 6, params, prims, bface)
    S, st, B, aux, emf_a = run_substep(wl)
    Ec, _ = cell_E(wl, aux)
    off = 1 if sign > 0 else 0
    worst = 0.0
    for ia in range(ng, ng + n + 1):
        for ib in range(ng, ng + n + 1):
            for ic in range(ng, ng + n):
                e = [0, 0, 0]
                e[a], e[b], e[c] = ia, ib, ic
                cc = [0, 0, 0]
                cc[a], cc[b], cc[c] = ia - off, ib - off, ic
                got = emf_a[0, c, e[2], e[1], e[0]]
                want = Ec[c, cc[2], cc[1], cc[0]]
                worst = max(worst, abs(got - want) / 1e-3)
    assert worst < 1e-12, worst


# ---------------------------------------------------------------------------
# substep / RK3 assembly with B != 0
# ---------------------------------------------------------------------------
def alfven_speed(rhoh, B0, eta):
    """Circularly polarised Alfven wave (Del Zanna, Zanotti, Roedig & Bucciantini
    2007, Sect. 4.2): exact speed of the nonlinear wave."""
    K = rhoh + B0 * B0 * (1 + eta * eta)
    return math.sqrt(2 * B0 * B0 / K / (1 + math.sqrt(1 - (2 * eta * B0 * B0 / K) ** 2)))


def alfven_run(N, emf, t_end, cfl=0.25, recon=ga.RECON_WENOZ):
    gm, rho0, p0, B0, eta = 4.0 / 3.0, 1.0, 1.0, 1.0, 1.0
    vA = alfven_speed(rho0 + gm / (gm - 1) * p0, B0, eta)
    kx = 2 * math.pi
    params = ga.default_params(gamma=gm, emf=emf, flat_metric=1, recon=recon)

    def prims(x, y, z):
        ph = kx * x
        zv = np.stack([np.zeros_like(x), -vA * eta * np.cos(ph), -vA * eta * np.sin(ph)])
        v2 = zv[1] ** 2 + zv[2] ** 2
        Wl = 1 / np.sqrt(1 - v2)
        return rho0 * np.ones_like(x), p0 * np.ones_like(x), Wl * zv

    def bface(d, x, y, z):
        if d == 0:
            return B0 * np.ones_like(x)
        return eta * B0 * (np.cos(kx * x) if d == 1 else np.sin(kx * x))

    wl = block_from_prims((N, 1, 1), 3, (float(N),) * 3, (ga.BC_PERIODIC,) * 6, params, prims, bface)
    S = O.Stepper(wl)
    st, B, aux = wl.copy_arrays()
    S.ghosts(st, B)
    S.aux(st, B, aux)
    dt = cfl / N
    nsteps = int(math.ceil(t_end / dt))
    dt = t_end / nsteps
    for _ in range(nsteps):
        S.rk3(dt, st, B, aux)
    x = (np.arange(N) + 0.5) / N
    by = B[1][0, 0, 3, 3, 3:3 + N]
    ex = eta * B0 * np.cos(kx * (x - vA * t_end))
    vy = aux[0, ga.ZVECY, 3, 3, 3:3 + N]
    Wn = np.sqrt(1 + aux[0, ga.ZVECX, 3, 3, 3:3 + N] ** 2 + vy ** 2 + aux[0, ga.ZVECZ, 3, 3, 3:3 + N] ** 2)
    ex_vy = -vA * eta * np.cos(kx * (x - vA * t_end))
    return np.abs(by - ex).mean(), np.abs(vy / Wn - ex_vy).mean(), S.fatal()


@pytest.mark.parametrize("emf", [ga.EMF_UCT, ga.EMF_GS])
def test_alfven_wave_converges(emf):
    """The nonlinear circularly polarised Alfven wave moves at its exact
    speed: the error of B^y and v^y after a quarter period falls at second
    order or better as the grid is refined (a sign slip in a b^mu term, in
    the EMFs or in the CT update moves the wave at another speed or not at
    all, and the error stays O(1))."""
    t_end = 0.25 / 0.3818
    e32 = alfven_run(32, emf, t_end)
    e64 = alfven_run(64, emf, t_end)
    assert e32[2] == 0 and e64[2] == 0
    rate_b = math.log2(e32[0] / e64[0])
    rate_v = math.log2(e32[1] / e64[1])
    assert e64[0] < 2e-3 and e64[1] < 1e-3, (e32, e64)
    assert rate_b > 1.8 and rate_v > 1.8, (rate_b, rate_v, e32, e64)


def test_balsara1_integral_invariants():
    """Balsara-1 (GS, outflow) to t = 0.4, before any wave reaches the
    boundaries: sum D, sum tau and sum B^y do not change, and
    sum S_x dx = t (T^xx_L - T^xx_R) = 0.9 t exactly, with
    T^xx = p + B^2/2 - B_x^2 of the two static end states."""
    wl = W.balsara1(N=200)
    S = O.Stepper(wl)
    st, B, aux = wl.copy_arrays()
    g = 3
    dx = wl.extra["dx"]
    sums0 = [st[0, v, g, g, g:-g].sum() * dx for v in (ga.DENS, ga.TAU)]
    by0 = B[1][0, 0, g, g, g:-g].sum() * dx
    t, steps = 0.0, 0
    while t < 0.4 - 1e-14:
        dt = min(S.dt(st, aux, 0.25), 0.4 - t)
        S.rk3(dt, st, B, aux)
        t += dt
        steps += 1
    sums1 = [st[0, v, g, g, g:-g].sum() * dx for v in (ga.DENS, ga.TAU)]
    for a, b_ in zip(sums0, sums1):
        assert abs(a - b_) < 1e-12 * abs(a), (a, b_)
    sx = st[0, ga.SX, g, g, g:-g].sum() * dx
    assert abs(sx - 0.9 * t) < 1e-11, (sx, 0.9 * t)
    assert abs(B[1][0, 0, g, g, g:-g].sum() * dx - by0) < 1e-12
    # the solution is not trivially unchanged: the density plateau exists
    rho = aux[0, ga.RHO, g, g, g:-g]
    assert rho.min() < 0.13 and rho.max() > 0.99 and steps > 100


def torus_rhs(n_blocks, n=16):
    """Interior residual |dU/dt| / |flux divergence| of the B = 0
    Fishbone-Moncrief torus (an exact stationary solution) after one tiny
    substep, in the torus body (rho > 0.4 rho_max), Kerr-Schild a = 0.9375."""
    wl = W.fm_torus(nb=(n_blocks,) * 3, n=n, extent=24.0, beta=0.0, noise=0.0)
    S = O.Stepper(wl)
    st, B, aux = wl.copy_arrays()
    S.metric_ghosts(st)
    S.ghosts(st, B)
    S.aux(st, B, aux)
    dt = 1e-6
    new, nB = st.copy(), [b.copy() for b in B]
    S.substep(dt, 1.0, new, st, nB, B, aux)
    g = wl.grid.ngz
    sl = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    rho = aux[:, ga.RHO][sl]
    body = rho > 0.4 * rho.max()
    out = {}
    for v, name in ((ga.DENS, "D"), (ga.SX, "Sx"), (ga.TAU, "tau")):
        rhs = (new[:, v][sl] - st[:, v][sl]) / dt
        out[name] = float(np.abs(rhs[body]).mean())
    # scale: the flux divergence alone (sources excluded) from the oracle's fluxes
    return out, int(body.sum())


def test_torus_stationary_residual_converges():
    """The curved-metric assembly (face metric, lapse/shift in the fluxes,
    geometric sources from the metric derivatives) keeps the stationary
    torus stationary up to truncation error: the body-averaged |dS_x/dt| and
    |dD/dt| fall by at least 2^1.5 when the spacing halves (a dropped or
    sign-flipped source term leaves an O(1) residual)."""
    r1, n1 = torus_rhs(2)
    r2, n2 = torus_rhs(4)
    assert n1 > 50 and n2 > 8 * 50
    for k in ("D", "Sx", "tau"):
        rate = math.log2(r1[k] / r2[k])
        assert rate > 1.5, (k, r1, r2, rate)


# ---------------------------------------------------------------------------
# radius-based excision (c2p.cpp:345-348 with rtp from cart_to_sph)
# ---------------------------------------------------------------------------
def ks_radius(x, y, z, a):
    """Kerr-Schild radius: the root r > 0 of (x^2 + y^2)/(r^2 + a^2) + z^2/r^2 = 1."""
    R2 = x * x + y * y + z * z
    r2 = 0.5 * (R2 - a * a + np.sqrt((R2 - a * a) ** 2 + 4 * a * a * z * z))
    return np.sqrt(r2)


def test_ks_radius_definition():
    rng = np.random.default_rng(9)
    x, y, z = rng.uniform(-5, 5, size=(3, 1000))
    a = 0.9375
    r = ks_radius(x, y, z, a)
    assert np.allclose((x * x + y * y) / (r * r + a * a) + z * z / (r * r), 1.0, rtol=0, atol=1e-12)
    assert np.allclose(ks_radius(3.0, 0.0, 0.0, a), math.sqrt(9 - a * a))
    assert np.allclose(ks_radius(0.0, 0.0, 2.5, a), 2.5)


@pytest.mark.parametrize("cks", [0, 1])
def test_radius_excision_resets_exactly_inside(cks):
    """excise_by_radius: after the c2p pass every cell whose radius (Kerr-Schild
    with cks, else |x|) is <= r_ex carries the atmosphere reset, and the
    lapse criterion no longer applies (a cell outside r_ex with lapse below
    excision_lapse is not excised)."""
    a, r_ex = 0.9375, 3.0
    wl = W.fm_torus(nb=(2, 2, 2), n=8, extent=6.0, beta=0.0, noise=0.0)
    wl.params.excise_by_radius = 1
    wl.params.excision_radius = r_ex
    wl.params.excision_cks = cks
    wl.params.excision_spin = a
    wl.params.excision_lapse = 0.9        # would excise much more if it applied
    S = O.Stepper(wl)
    st, B, aux = wl.copy_arrays()
    S.aux(st, B, aux)
    g = wl.grid
    N = g.nx + 2 * g.ngz
    dx = wl.extra["dx"]
    cc = (np.arange(N) - g.ngz + 0.5) * dx
    atmo_bit = 1 << ga.C2P_BITS.index("ATMO_RESET")
    inside_all, flagged_outside = 0, 0
    for q in range(g.nq):
        b = (q % 2, (q // 2) % 2, q // 4)
        zc, yc, xc = np.meshgrid(-6.0 + b[2] * 8 * dx + cc, -6.0 + b[1] * 8 * dx + cc, -6.0 + b[0] * 8 * dx + cc,
                                 indexing="ij")
        r = ks_radius(xc, yc, zc, a) if cks else np.sqrt(xc ** 2 + yc ** 2 + zc ** 2)
        bits = aux[q, ga.C2P_ERR].astype(np.uint64)
        reset = (bits & np.uint64(atmo_bit)) != 0
        inside = r <= r_ex
        assert reset[inside].all()
        inside_all += int(inside.sum())
        # outside r_ex the torus body (rho well above the floor) is never reset
        body = ~inside & (aux[q, ga.RHO] > 1e3 * wl.params.rho_fl)
        flagged_outside += int((reset & body).sum())
        alp = st[q, ga.ALP]
        assert (alp[~inside] < 0.9).any()  # the lapse criterion would have applied here
    assert inside_all > 100 and flagged_outside == 0

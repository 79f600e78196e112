"""Host-side checks of the Fishbone-Moncrief torus initial data
(grmhd_setup_fm_torus / grmhd_fm_torus_point, include/grmhd_setup.h).

Pinned by what the physics fixes, not by retyping the formulas: the torus
density is 1 at the pressure maximum r_max on the equator; the flow is a unit
timelike 4-velocity of the Kerr-Schild metric (W = alpha u^t); u^t u_phi is
the same constant everywhere in the torus (Fishbone & Moncrief's l); the
enthalpy obeys the relativistic Euler equation of a stationary azimuthal flow,
d ln h = -d ln(-u_t) + Omega / (1 - Omega l_s) d l_s with l_s = -u_phi/u_t,
checked by central differences along r and theta; the face fluxes from the
vector potential are divergence free to round-off; the field is scaled to
max p / max(b^2/2) = beta.  A dropped term of ln h, a wrong sign of the
frame-dragging velocity or a wrong l breaks one of them."""
import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W

SPIN, GAMMA, RIN, RMAX = 0.9375, 4.0 / 3.0, 6.0, 12.0


def ks4(x, y, z):
    """Kerr-Schild 4-metric (Cartesian) from the library's 3+1 data."""
    g = np.zeros(6)
    alp = np.zeros(1)
    beta = np.zeros(3)
    ga.lib().grmhd_kerr_schild(SPIN, x, y, z, g.ctypes.data, alp.ctypes.data, beta.ctypes.data)
    gij = np.array([[g[0], g[1], g[2]], [g[1], g[3], g[4]], [g[2], g[4], g[5]]])
    bl = gij @ beta
    G = np.zeros((4, 4))
    G[0, 0] = -alp[0] ** 2 + beta @ bl
    G[0, 1:] = G[1:, 0] = bl
    G[1:, 1:] = gij
    return G, alp[0], beta, gij


def cart(r, th, ph=0.0):
    """Cartesian Kerr-Schild point of Boyer-Lindquist-like (r, theta, phi)."""
    st = np.sin(th)
    return st * (r * np.cos(ph) - SPIN * np.sin(ph)), st * (r * np.sin(ph) + SPIN * np.cos(ph)), r * np.cos(th)


def flow(r, th):
    """ln h, u_t, u_phi, Omega at (r, theta) from the point function."""
    x, y, z = cart(r, th)
    rho, p, zx, zy, zz, ut = W.fm_torus_point(x, y, z, SPIN, GAMMA, RIN, RMAX)
    G, alp, beta, gij = ks4(x, y, z)
    zv = np.array([zx, zy, zz])
    u = np.concatenate([[ut], zv - ut * beta])          # u^i = z^i - u^t beta^i
    ul = G @ u
    dx_dphi = np.array([0.0, -y, x, 0.0])              # d(t,x,y,z)/dphi at fixed r, theta
    u_phi = ul @ dx_dphi
    uphi = u[1] * (-y) / (x * x + y * y) + u[2] * x / (x * x + y * y)   # dphi/dx^i u^i
    h = 1.0 + GAMMA / (GAMMA - 1.0) * p / rho
    return np.log(h), ul[0], u_phi, uphi / ut, rho, u, G, alp, zv, gij


def test_density_peak_and_normalisation():
    lnh, u_t, u_phi, Om, rho, u, G, alp, zv, gij = flow(RMAX, 0.5 * np.pi)
    assert abs(rho - 1.0) < 1e-12
    for r, th in [(RMAX, 0.5 * np.pi), (9.0, 1.3), (16.0, 1.1), (20.0, 1.45)]:
        lnh, u_t, u_phi, Om, rho, u, G, alp, zv, gij = flow(r, th)
        assert rho > 0
        assert abs(u @ G @ u + 1.0) < 1e-12                      # unit timelike
        W_ = np.sqrt(1.0 + zv @ gij @ zv)
        assert abs(W_ - alp * u[0]) < 1e-12 * W_                 # W = alpha u^t


def test_constant_l_and_hydrostatic_equilibrium():
    ls = []
    for r, th in [(RMAX, 0.5 * np.pi), (9.0, 1.3), (16.0, 1.1), (20.0, 1.45), (14.0, 1.3)]:
        lnh, u_t, u_phi, Om, rho, *_ = flow(r, th)
        ls.append(u_phi * flow(r, th)[5][0])                      # u_phi u^t
        # Euler: d ln h = -d ln(-u_t) + Omega/(1 - Omega l_s) d l_s
        for d in range(2):
            e = 1e-5
            a = (r + e, th) if d == 0 else (r, th + e)
            b = (r - e, th) if d == 0 else (r, th - e)
            fa, fb = flow(*a), flow(*b)
            dlnh = (fa[0] - fb[0]) / (2 * e)
            dln_ut = (np.log(-fa[1]) - np.log(-fb[1])) / (2 * e)
            lsa, lsb = -fa[2] / fa[1], -fb[2] / fb[1]
            l_s = -u_phi / u_t
            rhs = -dln_ut + Om / (1.0 - Om * l_s) * (lsa - lsb) / (2 * e)
            assert abs(dlnh - rhs) < 1e-6 * max(1.0, abs(dlnh)), (r, th, d, dlnh, rhs)
    ls = np.array(ls)
    assert np.all(np.abs(ls - ls[0]) < 1e-10 * abs(ls[0]))
    assert 4.2 < ls[0] < 4.4                                     # l = 4.28 for a=0.9375, r_in=6, r_max=12


def test_grid_data_divergence_free_and_beta():
    wl = W.fm_torus(nb=(2, 2, 2), n=8, extent=24.0)
    g = wl.grid
    sl = (slice(None), slice(g.ngz, g.ngz + g.nz), slice(g.ngz, g.ngz + g.ny), slice(g.ngz, g.ngz + g.nx))
    bdiv = wl.aux[:, ga.BDIV][sl]
    bmax = max(np.abs(b).max() for b in wl.B)
    assert bmax > 0
    assert np.abs(bdiv).max() * wl.extra["dx"] < 1e-13 * bmax
    pmax = wl.aux[:, ga.PRESS][sl].max()
    assert abs(pmax / wl.extra["pmag_max"] - 100.0) < 1e-9
    assert np.isfinite(wl.state).all() and np.isfinite(wl.aux).all()
    # atmosphere outside, torus inside
    rho = wl.aux[:, ga.RHO][sl]
    assert rho.min() == 1e-10 and 0.5 < rho.max() <= 1.0


def test_torus_forest_fully_refined_equals_uniform():
    """A forest with every coarse block refined holds, block for block, the
    uniform torus of the fine spacing (noise off: its key includes the level)."""
    wf = W.fm_torus_forest(nb=(2, 2, 2), n=8, ngz=4, extent=16.0, refined=list(range(8)), noise=0.0)
    wu = W.fm_torus(nb=(4, 4, 4), n=8, ngz=4, extent=16.0, noise=0.0)
    # forest block q = 8 c + child: fine coords 2 cb + child bits; uniform q = bi + 4 (bj + 4 bk)
    for q in range(wf.grid.nq):
        c, ch = divmod(q, 8)
        cb = (c % 2, (c // 2) % 2, c // 4)
        fc = [2 * cb[d] + ((ch >> d) & 1) for d in range(3)]
        qu = fc[0] + 4 * (fc[1] + 4 * fc[2])
        np.testing.assert_allclose(wf.state[q], wu.state[qu], rtol=1e-14, atol=1e-300)
        np.testing.assert_allclose(wf.aux[q], wu.aux[qu], rtol=1e-14, atol=1e-300)
        for d in range(3):
            np.testing.assert_allclose(wf.B[d][q], wu.B[d][qu], rtol=1e-13, atol=1e-18)
    assert wf.extra["bscale"] == pytest.approx(wu.extra["bscale"], rel=1e-14)


def test_torus_forest_levels_agree_on_the_metric():
    """Coarse and fine blocks sample the same Kerr-Schild metric: a coarse
    cell's value equals the fine value at the same centre only up to the
    half-cell offset, but the analytic metric at a shared point is identical."""
    wf = W.fm_torus_forest(nb=(2, 2, 2), n=8, ngz=4, extent=16.0, refined=[0], noise=0.0)
    assert wf.grid.nq == 15
    assert np.all(np.isfinite(wf.state[:, :7])) and np.all(np.isfinite(wf.aux[:, 0]))
    # divergence of the face field in every interior cell of every block
    n, g = 8, 4
    for q in range(wf.grid.nq):
        lev = 0 if q < 7 else 1
        idx = (1.0 / wf.extra["dx"]) * (2 if lev else 1)
        Bx, By, Bz = (b[q, 0] for b in wf.B)
        div = (Bx[g:g + n, g:g + n, g + 1:g + n + 1] - Bx[g:g + n, g:g + n, g:g + n]) * idx + \
              (By[g:g + n, g + 1:g + n + 1, g:g + n] - By[g:g + n, g:g + n, g:g + n]) * idx + \
              (Bz[g + 1:g + n + 1, g:g + n, g:g + n] - Bz[g:g + n, g:g + n, g:g + n]) * idx
        scale = max(np.abs(b[q]).max() for b in wf.B) * idx + 1e-300
        assert np.abs(div).max() <= 1e-13 * scale, (q, np.abs(div).max(), scale)

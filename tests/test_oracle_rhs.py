"""Invariants of the oracle's block-level substep / RK3 step (CPU only).

Pins what the mathematics fixes for the block pipeline (evolve.cpp:813-1479,
auxiliaries.cpp:81-180): a uniform state does not change, CT keeps div B at
round-off, mirror-symmetric data stays mirror symmetric, the ghost fill
copies the clamped global cell for outflow and the wrapped one for periodic.
"""
import ctypes

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O


def uniform_blast(nb=(2, 2, 2), n=8, B=(0.1, 0.05, -0.02), v=(0.1, -0.2, 0.05)):
    wl = W.blast(nb=nb, n=n, noise=0.0)
    # overwrite with a uniform moving magnetised state
    st, Bf, aux = wl.state, wl.B, wl.aux
    gam = 4.0 / 3.0
    rho, p = 1e-2, 1e-3
    eps = p / ((gam - 1) * rho)
    v = np.array(v)
    Wl = 1 / np.sqrt(1 - (v ** 2).sum())
    z = Wl * v
    for d in range(3):
        Bf[d][...] = B[d]
    aux[:, ga.RHO] = rho
    aux[:, ga.PRESS] = p
    aux[:, ga.EPS] = eps
    aux[:, ga.TEMP] = eps * (gam - 1)
    aux[:, ga.ENTROPY] = np.log(eps) / (gam - 1) - np.log(rho)
    for d in range(3):
        aux[:, ga.ZVECX + d] = z[d]
        aux[:, ga.BX + d] = B[d]
    # conserved from the library's own prims_to_conservs
    n_cells = st[:, 0].size
    prims = np.zeros((13, n_cells))
    prims[0] = rho; prims[1] = p; prims[2:5] = z[:, None]; prims[6] = eps * (gam - 1); prims[7] = eps
    prims[8] = np.log(eps) / (gam - 1) - np.log(rho)
    prims[9:12] = np.array(B)[:, None]
    met = np.zeros((10, n_cells)); met[0] = met[3] = met[5] = met[6] = 1.0
    cons = np.zeros((10, n_cells))
    O.olib().orc_p2c_batch(ctypes.byref(wl.params), n_cells, O.p(prims), O.p(met), O.p(cons))
    for v_ in range(7):
        st[:, v_] = cons[v_].reshape(st[:, v_].shape)
    return wl


def test_uniform_state_is_stationary():
    wl = uniform_blast()
    st, B, aux = wl.copy_arrays()
    s0 = st.copy()
    S = O.Stepper(wl)
    S.rk3(0.25 / 16, st, B, aux)
    g = wl.grid.ngz
    for v in range(7):
        a = st[:, v, g:-g, g:-g, g:-g]
        b = s0[:, v, g:-g, g:-g, g:-g]
        scale = max(np.max(np.abs(b)), 1e-300)
        assert np.max(np.abs(a - b)) / scale < 1e-13, v
    for d in range(3):
        assert np.max(np.abs(B[d] - wl.B[d])) < 1e-15


def test_divB_roundoff_blast():
    wl = W.blast(nb=(2, 2, 2), n=8)
    st, B, aux = wl.copy_arrays()
    S = O.Stepper(wl)
    dt = 0.25 * wl.extra["dx"]
    for _ in range(3):
        S.rk3(dt, st, B, aux)
    g = wl.grid.ngz
    assert np.max(np.abs(aux[:, ga.BDIV, g:-g, g:-g, g:-g])) < 1e-12
    assert S.fatal() == 0
    # the blast wave moved: the pressure changed inside
    assert np.max(np.abs(aux[:, ga.PRESS] - wl.aux[:, ga.PRESS])) > 1e-3


def test_mirror_symmetry_blast():
    """A blast without noise is symmetric under x -> -x; the evolved density
    is symmetric to round-off (block layout mirrors too)."""
    wl = W.blast(nb=(2, 2, 2), n=8, noise=0.0)
    st, B, aux = wl.copy_arrays()
    S = O.Stepper(wl)
    S.rk3(0.25 * wl.extra["dx"], st, B, aux)
    g, n = wl.grid.ngz, 8
    rho = aux[:, ga.RHO]
    # assemble the global interior density (blocks ordered bi + 2*(bj + 2*bk))
    G = np.zeros((16, 16, 16))
    for q in range(8):
        bi, bj, bk = q % 2, (q // 2) % 2, q // 4
        G[bk * n:(bk + 1) * n, bj * n:(bj + 1) * n, bi * n:(bi + 1) * n] = rho[q, g:-g, g:-g, g:-g]
    assert np.max(np.abs(G - G[:, :, ::-1])) / np.max(np.abs(G)) < 1e-13


def test_ghost_fill_outflow_and_periodic():
    wl = W.balsara1(N=16)
    st, B, aux = wl.copy_arrays()
    rng = np.random.default_rng(0)
    st[:, :7] = rng.normal(size=st[:, :7].shape)
    for b in B:
        b[...] = rng.normal(size=b.shape)
    S = O.Stepper(wl)
    S.ghosts(st, B)
    g = wl.grid.ngz
    # outflow along x: ghosts copy the first/last interior cell
    for v in range(7):
        for i in range(g):
            assert np.array_equal(st[0, v, :, :, i], st[0, v, :, :, g])
        assert np.array_equal(st[0, v, g, g, 0], st[0, v, g, g, g])
        assert np.array_equal(st[0, v, g, g, -1], st[0, v, g, g, -1 - g])
        # periodic along y/z with one cell: every ghost row equals the interior row
        for j in range(st.shape[3]):
            assert np.array_equal(st[0, v, g, j, g:-g], st[0, v, g, g, g:-g])
    # x faces: outflow clamps to the boundary faces g and N+g
    bx = B[0]
    assert np.array_equal(bx[0, 0, g, g, 0], bx[0, 0, g, g, g])
    assert np.array_equal(bx[0, 0, g, g, -1], bx[0, 0, g, g, 16 + g])


def test_balsara1_gs_runs():
    wl = W.balsara1(N=200)
    st, B, aux = wl.copy_arrays()
    S = O.Stepper(wl)
    dt = 0.25 * wl.extra["dx"]
    for _ in range(20):
        dtc = S.dt(st, aux, 0.25)
        assert dtc > 0
        S.rk3(min(dt, dtc), st, B, aux)
    g = wl.grid.ngz
    rho = aux[0, ga.RHO, g, g, g:-g]
    assert np.all(np.isfinite(rho)) and rho.min() > 0.05 and rho.max() < 1.2
    # B^x is constant (=0.5) in 1D: d Bx/dt = -dEz/dy + dEy/dz = 0
    assert np.max(np.abs(B[0][0, 0, g, g, g:-g] - 0.5)) < 1e-13

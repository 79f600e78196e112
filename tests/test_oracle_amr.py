"""CPU oracle of the two-level forest: coarse-fine ghost exchange and reflux.

Pins the oracle's restatement (oracle/grmhd_oracle_amr.hh) with properties the
operators must have, each chosen so that a plausible slip fails:
  * constant data stays constant through the exchange (every branch);
  * restriction into coarse ghosts is the mean of the 8 fine cells covering
    the coarse cell (wrong child / offset fails);
  * a field linear in x, y, z is reproduced exactly by prolongation and
    restriction (a flipped slope sign or a wrong coarse centre fails);
  * prolonged cells average back to their coarse cell (conservation);
  * face B after the exchange is divergence free in every cell of every
    block, ghosts included (Toth-Roe signs, area-average restriction);
  * on a periodic forest an RK3 step conserves the volume integrals of the
    hydro conserved variables (flux reflux), keeps div B at round-off and
    keeps every coarse face on a coarse-fine interface equal to the mean of
    its four fine faces (EMF reflux) -- the invariants of the reference's
    test_flux_reflux.cpp:215-383 and test_ct_flux_conservation.cpp:342-362.
"""
import numpy as np
import pytest

from grace_amd import workloads as W
from _oracle import Stepper

NH = 7


def geometry(o, wl):
    g = wl.grid
    nq = g.nq
    blocks = [o.forest_block(q) for q in range(nq)]
    index = {b: q for q, b in enumerate(blocks)}
    return blocks, index


def ghost_mask(n, ng):
    N = n + 2 * ng
    m = np.ones((N, N, N), dtype=bool)
    m[ng:n + ng, ng:n + ng, ng:n + ng] = False
    return m   # [k, j, i]


def covering(index, nb, n, lev, gidx):
    """Block (level, q) covering global cell gidx (i, j, k) at level lev, periodic."""
    Ntot = [nb[d] * n << lev for d in range(3)]
    g = [gidx[d] % Ntot[d] for d in range(3)]
    if lev == 1:
        crd = tuple(g[d] // n for d in range(3))
        if (1, crd) in index:
            return 1, index[(1, crd)], tuple(g[d] % n for d in range(3))
        cg = [g[d] // 2 for d in range(3)]
        crd = tuple(cg[d] // n for d in range(3))
        return 0, index[(0, crd)], tuple(cg[d] % n for d in range(3))
    crd = tuple(g[d] // n for d in range(3))
    if (0, crd) in index:
        return 0, index[(0, crd)], tuple(g[d] % n for d in range(3))
    return 1, None, None


def test_forest_layout():
    wl = W.forest_wave(nb=(2, 2, 2), n=8, refined=[0, 7])
    o = Stepper(wl)
    blocks, index = geometry(o, wl)
    assert wl.grid.nq == 8 + 14
    assert [b[0] for b in blocks] == [0] * 6 + [1] * 16
    assert blocks[:6] == [(0, (1, 0, 0)), (0, (0, 1, 0)), (0, (1, 1, 0)), (0, (0, 0, 1)), (0, (1, 0, 1)), (0, (0, 1, 1))]
    assert blocks[6] == (1, (0, 0, 0)) and blocks[7] == (1, (1, 0, 0)) and blocks[13] == (1, (1, 1, 1))
    assert blocks[14] == (1, (2, 2, 2)) and blocks[21] == (1, (3, 3, 3))


def test_constant_state_is_kept():
    wl = W.forest_wave(nb=(2, 2, 2), n=8, refined=[0], amp=0.0, b_loop=0.0)
    o = Stepper(wl)
    st, B, aux = wl.copy_arrays()
    ng, n = wl.grid.ngz, wl.grid.nx
    gm = ghost_mask(n, ng)
    ref = st[0, :NH, ng, ng, ng].copy()
    st[:, :NH][:, :, gm] = -7.0
    for d in range(3):
        keep = B[d][:, :, ng + 1:n + ng - 1, ng + 1:n + ng - 1, ng + 1:n + ng - 1].copy()
        B[d][:] = -7.0
        B[d][:, :, ng + 1:n + ng - 1, ng + 1:n + ng - 1, ng + 1:n + ng - 1] = keep
    # faces of the interior (normal range [ng, n+ng]) back from the input
    for d in range(3):
        sl = [slice(None), slice(None), slice(ng, n + ng), slice(ng, n + ng), slice(ng, n + ng)]
        sl[4 - d] = slice(ng, n + ng + 1)
        B[d][tuple(sl)] = wl.B[d][tuple(sl)]
    o.ghosts(st, B)
    for v in range(NH):
        assert np.all(st[:, v] == ref[v]), v
    for d in range(3):
        assert np.all(B[d] == wl.B[d][0, 0, ng, ng, ng]), d


def _linear_state(wl, o, coef):
    """DENS = c0 + c . x at every cell centre of every block (interior and ghost)."""
    blocks, index = geometry(o, wl)
    st, B, aux = wl.copy_arrays()
    n, ng = wl.grid.nx, wl.grid.ngz
    dx0 = wl.extra["dx"]
    ax = np.arange(-ng, n + ng) + 0.5
    for q, (lev, crd) in enumerate(blocks):
        dx = dx0 / (1 << lev)
        x = (crd[0] * n + ax) * dx
        y = (crd[1] * n + ax) * dx
        z = (crd[2] * n + ax) * dx
        st[q, 0] = coef[0] + coef[1] * x[None, None, :] + coef[2] * y[None, :, None] + coef[3] * z[:, None, None]
    return st, B, aux, blocks, index


def test_linear_field_exact_through_exchange():
    # refined 2x2x2 centre of a 4^3 forest: every coarse-fine stencil stays away from the periodic wrap
    wl = W.forest_wave(nb=(4, 4, 4), n=8)
    o = Stepper(wl)
    coef = (0.7, 1.3, -0.6, 0.45)
    st, B, aux, blocks, index = _linear_state(wl, o, coef)
    exact = st.copy()
    n, ng = wl.grid.nx, wl.grid.ngz
    gm = ghost_mask(n, ng)
    st[:, 0][:, gm] = 0.0
    o.ghosts(st, B)
    # compare only ghosts whose global position is inside the domain (no wrap)
    worst = 0.0
    for q, (lev, crd) in enumerate(blocks):
        N = 4 * n << lev
        ins = [((crd[d] * n + np.arange(-ng, n + ng)) >= 0) & ((crd[d] * n + np.arange(-ng, n + ng)) < N)
               for d in range(3)]
        m = gm & ins[0][None, None, :] & ins[1][None, :, None] & ins[2][:, None, None]
        worst = max(worst, np.abs(st[q, 0][m] - exact[q, 0][m]).max())
    assert worst < 1e-13, worst


def test_restriction_is_fine_mean_and_prolongation_conserves():
    wl = W.forest_wave(nb=(2, 2, 2), n=8, refined=[0, 3])
    o = Stepper(wl)
    st, B, aux = wl.copy_arrays()
    n, ng = wl.grid.nx, wl.grid.ngz
    o.ghosts(st, B)
    blocks, index = geometry(o, wl)
    nb = (2, 2, 2)
    n_restr = n_prol = 0
    for q, (lev, crd) in enumerate(blocks):
        for k in range(n + 2 * ng):
            for j in range(n + 2 * ng):
                for i in range(n + 2 * ng):
                    loc = (i, j, k)
                    if all(ng <= loc[d] < n + ng for d in range(3)):
                        continue
                    g = [crd[d] * n + loc[d] - ng for d in range(3)]
                    cov, qs, sl = covering(index, nb, n, lev, g)
                    if lev == 0 and cov == 1:
                        vals = []
                        for c in range(8):
                            gf = [2 * g[d] + ((c >> d) & 1) for d in range(3)]
                            _, qf, lf = covering(index, nb, n, 1, gf)
                            vals.append(st[qf, :NH, lf[2] + ng, lf[1] + ng, lf[0] + ng])
                        np.testing.assert_allclose(st[q, :NH, k, j, i], np.mean(vals, axis=0), rtol=1e-15, atol=1e-15)
                        n_restr += 1
                    elif lev == 1 and cov == 0 and all(x % 2 == 0 for x in loc):
                        # a whole 2x2x2 group of prolonged ghosts: its mean is the coarse cell
                        grp = st[q, :NH, k:k + 2, j:j + 2, i:i + 2]
                        if k + 1 >= n + 2 * ng or j + 1 >= n + 2 * ng or i + 1 >= n + 2 * ng:
                            continue
                        if any(ng <= loc[d] + 1 < n + ng and ng <= loc[d] < n + ng for d in range(3)) and \
                                all(ng <= loc[d] < n + ng or ng <= loc[d] + 1 < n + ng for d in range(3)):
                            continue
                        coarse = st[qs, :NH, sl[2] + ng, sl[1] + ng, sl[0] + ng]
                        np.testing.assert_allclose(grp.reshape(NH, -1).mean(axis=1), coarse, rtol=1e-14, atol=1e-15)
                        n_prol += 1
    assert n_restr > 100 and n_prol > 100, (n_restr, n_prol)


def div_b(B, idx, n, ng):
    """div B * dx_min (face differences weighted by the inverse spacings of the
    coarse level, which differ from every level's by the same factor) in every
    cell of every block."""
    w = np.asarray(idx, dtype=np.float64) / max(idx)
    return (B[0][:, 0, :, :, 1:] - B[0][:, 0, :, :, :-1]) * w[0] + (B[1][:, 0, :, 1:, :] - B[1][:, 0, :, :-1, :]) * w[1] + \
           (B[2][:, 0, 1:, :, :] - B[2][:, 0, :-1, :, :]) * w[2]


@pytest.mark.parametrize("nb,refined,bc", [((2, 2, 2), [0], 1), ((2, 2, 2), [0, 3, 5], 1),
                                           ((2, 2, 2), [0, 1, 2, 3, 4, 5, 6], 1), ((3, 2, 2), [0, 4], 0),
                                           ((3, 3, 2), [1, 3, 4, 10], (1, 1, 0, 0, 1, 1))])
def test_exchange_keeps_div_b_zero(nb, refined, bc):
    wl = W.forest_wave(nb=nb, n=8, refined=refined, b_loop=0.3, bc=bc)
    o = Stepper(wl)
    st, B, aux = wl.copy_arrays()
    n, ng = wl.grid.nx, wl.grid.ngz
    for d in range(3):   # wipe every ghost face, keep the interior ones
        keep = wl.B[d].copy()
        B[d][:] = 1e3
        sl = [slice(None), slice(None), slice(ng, n + ng), slice(ng, n + ng), slice(ng, n + ng)]
        sl[4 - d] = slice(ng, n + ng + 1)
        B[d][tuple(sl)] = keep[tuple(sl)]
    o.ghosts(st, B)
    scale = max(np.abs(b).max() for b in B)
    assert scale < 10
    db = div_b(B, wl.forest.idx, n, ng)
    # cells beyond an outflow face hold clamped copies (not divergence free,
    # as on uniform grids); every other cell, ghosts included, must be
    blocks, _ = geometry(o, wl)
    bcs = list(wl.forest.bc)
    for q, (lev, crd) in enumerate(blocks):
        keep = np.ones((n + 2 * ng,) * 3, dtype=bool)
        for d in range(3):
            gi = crd[d] * n + np.arange(-ng, n + ng)
            N = wl.forest.nb[d] * n << lev
            ok = np.ones(n + 2 * ng, dtype=bool)
            if bcs[2 * d] == 0:
                ok &= gi >= 0
            if bcs[2 * d + 1] == 0:
                ok &= gi < N
            shape = [1, 1, 1]
            shape[2 - d] = n + 2 * ng
            keep &= ok.reshape(shape)
        assert np.abs(db[q][keep]).max() < 1e-14 * scale * 8, (q, np.abs(db[q][keep]).max())


def interface_mismatch(o, wl, B):
    """max |coarse face - mean of the 4 fine faces on it| over coarse-fine interfaces."""
    blocks, index = geometry(o, wl)
    n, ng = wl.grid.nx, wl.grid.ngz
    nb = tuple(wl.forest.nb)
    worst = 0.0
    for q, (lev, crd) in enumerate(blocks):
        if lev != 0:
            continue
        for d in range(3):
            for side in (0, 1):
                nbc = list(crd)
                nbc[d] = (nbc[d] + (1 if side else -1)) % nb[d]
                if (0, tuple(nbc)) in index:
                    continue
                t0, t1 = [e for e in range(3) if e != d]
                for b in range(n):
                    for a in range(n):
                        lc = [0, 0, 0]
                        lc[d] = n + ng if side else ng
                        lc[t0] = ng + a; lc[t1] = ng + b
                        bc_ = B[d][q, 0, lc[2], lc[1], lc[0]]
                        fv = []
                        for u in range(2):
                            for w in range(2):
                                gf = [0, 0, 0]
                                gf[d] = 2 * (crd[d] * n + (n if side else 0))
                                gf[t0] = 2 * (crd[t0] * n + a) + u
                                gf[t1] = 2 * (crd[t1] * n + b) + w
                                Nf = [2 * nb[e] * n for e in range(3)]
                                # the fine cell on the far side of the face: above it (side 1) or below (side 0)
                                cell = list(gf)
                                cell[d] = gf[d] if side else gf[d] - 1
                                cell = [cell[e] % Nf[e] for e in range(3)]
                                qf = index[(1, tuple(cell[e] // n for e in range(3)))]
                                lf = [cell[e] % n + ng for e in range(3)]
                                lf[d] += 0 if side else 1
                                fv.append(B[d][qf, 0, lf[2], lf[1], lf[0]])
                        worst = max(worst, abs(bc_ - np.mean(fv)))
    return worst


def volume_integrals(o, wl, st):
    blocks, _ = geometry(o, wl)
    n, ng = wl.grid.nx, wl.grid.ngz
    dx0 = wl.extra["dx"]
    tot = np.zeros(5)
    for q, (lev, crd) in enumerate(blocks):
        dv = (dx0 / (1 << lev)) ** 3
        tot += dv * st[q, :5, ng:n + ng, ng:n + ng, ng:n + ng].sum(axis=(1, 2, 3))
    return tot


@pytest.mark.parametrize("emf", [0, 1])
def test_rk3_conserves_and_keeps_interfaces_consistent(emf):
    wl = W.forest_wave(nb=(2, 2, 2), n=8, refined=[0, 3], b_loop=0.2, p_in=3.0, r_in=0.3, emf=emf)
    o = Stepper(wl)
    st, B, aux = wl.copy_arrays()
    n, ng = wl.grid.nx, wl.grid.ngz
    o.ghosts(st, B)
    o.aux(st, B, aux)
    m0 = interface_mismatch(o, wl, B)
    assert m0 < 1e-15, m0
    tot0 = volume_integrals(o, wl, st)
    for _ in range(2):
        dt = o.dt(st, aux, 0.25)
        o.rk3(dt, st, B, aux)
    tot1 = volume_integrals(o, wl, st)
    assert np.all(np.isfinite(st))
    rel = np.abs(tot1 - tot0) / np.maximum(np.abs(tot0), 1e-3)
    assert rel.max() < 5e-14, (tot0, tot1, rel)
    gm = ~ghost_mask(n, ng)
    db = div_b(B, wl.forest.idx, n, ng)[:, gm]
    assert np.abs(db).max() < 1e-14, np.abs(db).max()
    m1 = interface_mismatch(o, wl, B)
    assert m1 < 1e-14, m1
    assert o.fatal() == 0


FOFC_SMOOTH = dict(rho0=1.0, p0=1e-3, p_in=10.0, r_in=0.25, amp=0.9, zvel=(0.9, 0.5, 0.2), b_loop=0.2)
FOFC_HARSH = dict(rho0=1.0, p0=1e-6, p_in=1e3, r_in=0.2, amp=0.99, zvel=(3.0, 2.0, 1.0), b_loop=1.0)


def _fofc_run(refined, kw, ghost_emf, steps=3):
    wl = W.forest_wave(nb=(2, 2, 2), n=8, refined=refined, emf=1, **kw)
    wl.params.fofc = 1
    wl.params.fofc_ghost_emf = ghost_emf
    o = Stepper(wl)
    st, B, aux = wl.copy_arrays()
    o.ghosts(st, B)
    o.aux(st, B, aux)
    tot0 = volume_integrals(o, wl, st)
    for _ in range(steps):
        o.rk3(o.dt(st, aux, 0.25), st, B, aux)
    tot1 = volume_integrals(o, wl, st)
    rel = float((np.abs(tot1 - tot0) / np.maximum(np.abs(tot0), 1e-3)).max())
    return o, wl, B, rel


def test_fofc_reference_edge_range_flags_spuriously():
    """The reference computes GS edge EMFs on [ngz, n+ngz] (evolve.cpp:1090-1104)
    but its FOFC flag pass also advances face B for the first ghost layer
    (evolve.cpp:492-525), reading edges nothing wrote: on smooth periodic data
    cells are flagged and the two sides of a block face disagree, so the run
    stops conserving.  With the EMFs one ghost layer wider
    (fofc_ghost_emf = 1) nothing is flagged and conservation holds."""
    o_ref, _, _, rel_ref = _fofc_run([], FOFC_SMOOTH, 0)
    o_fix, _, _, rel_fix = _fofc_run([], FOFC_SMOOTH, 1)
    assert o_ref.fofc_flagged() > 100 and rel_ref > 1e-8, (o_ref.fofc_flagged(), rel_ref)
    assert o_fix.fofc_flagged() == 0 and rel_fix < 5e-14, (o_fix.fofc_flagged(), rel_fix)


@pytest.mark.parametrize("refined", [[], [0, 3]])
def test_fofc_with_ghost_emfs_conserves_on_forests(refined):
    """Strong blast with FOFC active in tens of thousands of cells: with the
    ghost-layer EMFs the first-order fallback keeps the volume integrals
    (and, on the forest, the coarse-fine face consistency) at round-off."""
    o, wl, B, rel = _fofc_run(refined, FOFC_HARSH, 1)
    assert o.fofc_flagged() > 10000, o.fofc_flagged()
    assert rel < 5e-14, rel
    assert o.fatal() == 0
    if refined:
        assert interface_mismatch(o, wl, B) < 1e-14


def _uniform_twin(wf):
    """The uniform periodic grid of the forest's fine spacing, block for block:
    (instance spec, map forest q -> uniform q)."""
    nbc = list(wf.forest.nb)
    nbf = [2 * x for x in nbc]
    idx = [2 * x for x in wf.forest.idx]
    spec = [dict(nb=tuple(nbf), q0=0, bc=tuple(wf.forest.bc), idx=tuple(idx))]
    qmap = []
    for q in range(wf.grid.nq):
        c, ch = divmod(q, 8)
        cb = (c % nbc[0], (c // nbc[0]) % nbc[1], c // (nbc[0] * nbc[1]))
        fc = [2 * cb[d] + ((ch >> d) & 1) for d in range(3)]
        qmap.append(fc[0] + nbf[0] * (fc[1] + nbf[1] * fc[2]))
    return spec, qmap


def test_fully_refined_forest_equals_uniform_grid():
    """Every coarse block refined: no coarse-fine interface, so the forest's
    exchange and RK3 must reproduce the uniform grid of the fine spacing
    (block order permuted)."""
    import copy
    from grace_amd import make_instances
    wf = W.forest_wave(nb=(2, 2, 2), n=8, refined=list(range(8)), b_loop=0.2, p_in=3.0, r_in=0.3)
    spec, qmap = _uniform_twin(wf)
    wu = copy.copy(wf)
    wu.forest = None
    wu.inst_specs = spec
    wu.instances = make_instances(spec)
    inv = np.argsort(qmap)
    wu.state = wf.state[inv].copy()
    wu.B = [b[inv].copy() for b in wf.B]
    wu.aux = wf.aux[inv].copy()
    res = []
    for w in (wf, wu):
        o = Stepper(w)
        st, B, aux = w.copy_arrays()
        o.ghosts(st, B)
        o.aux(st, B, aux)
        dt = o.dt(st, aux, 0.25)
        o.rk3(dt, st, B, aux)
        res.append((st, B, aux, dt))
    (sf, Bf, af, dtf), (su, Bu, au, dtu) = res
    # The initial data holds both copies of a face on a periodic block
    # boundary (A_z sampled at x = 0 and at x = L differ in the last bit); the
    # uniform exchange takes the lower block's copy at the wrap, the forest's
    # the upper one's, so agreement is to round-off, not bit for bit.
    assert abs(dtf - dtu) <= 1e-15 * dtu
    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    assert rel(sf[:, :7], su[qmap][:, :7]) < 1e-13
    for v in (0, 1, 2, 3, 4, 5, 6, 11):
        assert rel(af[:, v], au[qmap][:, v]) < 1e-13, v
    for d in range(3):
        assert rel(Bf[d], Bu[d][qmap]) < 1e-13, d


def _metric_field(wl, o, fun):
    """Fill state var GXX with fun(u, v, w) of cell-centre coordinates
    (centred on the domain, in coarse cells) in every cell of every block."""
    blocks, index = geometry(o, wl)
    st, B, aux = wl.copy_arrays()
    n, ng = wl.grid.nx, wl.grid.ngz
    for q, (lev, crd) in enumerate(blocks):
        h = 1.0 / (1 << lev)
        c = [((crd[d] * n + np.arange(-ng, n + ng) + 0.5) - 0.5 * (wl.forest.nb[d] * n << lev)) * h for d in range(3)]
        st[q, 7] = fun(c[0][None, None, :], c[1][None, :, None], c[2][:, None, None])
    return st, B, aux, blocks, index


def test_metric_exchange_reproduces_quartics():
    """Cowling-metric ghost exchange (copies + 5-point Lagrange prolongation /
    restriction): a quartic polynomial of position is reproduced in every
    in-domain ghost cell of every block."""
    wl = W.forest_wave(nb=(4, 4, 4), n=8)
    o = Stepper(wl)
    f = lambda x, y, z: 1.0 + 0.3 * x - 0.2 * y + 0.1 * z + 0.01 * x * x * y - 0.007 * y * z * z + 4e-4 * x ** 4 \
        - 3e-4 * (x * y) ** 2 + 2.5e-4 * z ** 3 * y
    st, B, aux, blocks, index = _metric_field(wl, o, f)
    exact = st.copy()
    n, ng = wl.grid.nx, wl.grid.ngz
    gm = ghost_mask(n, ng)
    st[:, 7][:, gm] = 0.0
    o.metric_ghosts(st)
    worst = 0.0
    for q, (lev, crd) in enumerate(blocks):
        N = 4 * n << lev
        ins = [((crd[d] * n + np.arange(-ng, n + ng)) >= 0) & ((crd[d] * n + np.arange(-ng, n + ng)) < N)
               for d in range(3)]
        m = gm & ins[0][None, None, :] & ins[1][None, :, None] & ins[2][:, None, None]
        worst = max(worst, float(np.abs(st[q, 7][m] - exact[q, 7][m]).max()))
    assert worst < 1e-13 * np.abs(exact[:, 7]).max(), worst


def test_metric_exchange_is_mirror_symmetric():
    """Data even in x on a forest symmetric under x -> -x: the exchanged ghosts
    are bit-for-bit mirror images (the half-dependent restriction stencils and
    the pair-symmetric contractions of pr_helpers.hh:52-58, 222-330; a fixed
    stencil choice leaves O(h^5) asymmetries)."""
    wl = W.forest_wave(nb=(4, 4, 4), n=8)
    o = Stepper(wl)
    f = lambda x, y, z: np.cos(0.37 * x) * (1.0 + 0.1 * y) + np.exp(-0.01 * (x * x + z * z)) * np.sin(0.2 * z + 0.3)
    st, B, aux, blocks, index = _metric_field(wl, o, f)
    n, ng = wl.grid.nx, wl.grid.ngz
    gm = ghost_mask(n, ng)
    st[:, 7][:, gm] = 0.0
    o.metric_ghosts(st)
    for q, (lev, crd) in enumerate(blocks):
        mc = ((wl.forest.nb[0] << lev) - 1 - crd[0], crd[1], crd[2])
        qm = index[(lev, mc)]
        assert np.array_equal(st[q, 7], st[qm, 7][:, :, ::-1]), (q, qm)

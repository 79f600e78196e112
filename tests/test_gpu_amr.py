"""GPU parity on two-level forests: the HIP coarse-fine ghost exchange,
refluxing and RK3 (through the C ABI) against the CPU oracle on identical
inputs, plus the invariants on the GPU's own results.

Tolerances: ghost fill bit-exact (same operations in the same order);
conserved variables and face B after a substep / RK3 step L-infinity relative
<= 1e-12 per variable over interior cells; volume integrals of the conserved
variables conserved to 5e-14 relative on periodic forests; div B and the
coarse-fine face consistency at round-off (1e-14).
"""
import sys

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O
from test_oracle_amr import volume_integrals, interface_mismatch, div_b, ghost_mask
from test_gpu_parity import check_cons, to_dev, to_host, interior, face_interior, linf_rel

pytestmark = pytest.mark.gpu

FORESTS = {
    "one": dict(nb=(2, 2, 2), refined=[0]),
    "three": dict(nb=(2, 2, 2), refined=[0, 3, 5]),
    "centre": dict(nb=(4, 4, 4), refined=None),
    "outflow": dict(nb=(3, 2, 2), refined=[0, 4], bc=0),
    "mixed_bc": dict(nb=(3, 3, 2), refined=[1, 3, 4, 10], bc=(1, 1, 0, 0, 1, 1)),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def ctx_of(wl):
    return ga.Context(wl.params, wl.grid, wl.instances, forest=wl.forest)


@pytest.mark.parametrize("name", list(FORESTS))
def test_gpu_forest_ghost_fill_bitexact(dev, name):
    import torch
    wl = W.forest_wave(n=8, ngz=4, b_loop=0.3, **FORESTS[name])
    say(f"[forest ghosts] {name} nq={wl.grid.nq}")
    rng = np.random.default_rng(3)
    st, B, aux = wl.copy_arrays()
    st[:, :7] *= 1.0 + 0.01 * rng.normal(size=st[:, :7].shape)
    for b in B:
        b *= 1.0 + 0.01 * rng.normal(size=b.shape)
    sg, Bg = to_dev(st, dev), [to_dev(b, dev) for b in B]
    o = O.Stepper(wl)
    o.ghosts(st, B)
    ctx = ctx_of(wl)
    ctx.fill_ghosts(sg, Bg)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(sg), st)
    for d in range(3):
        assert np.array_equal(to_host(Bg[d]), B[d]), d


def test_gpu_forest_block_layout(dev):
    wl = W.forest_wave(n=8, ngz=4, **FORESTS["three"])
    ctx = ctx_of(wl)
    o = O.Stepper(wl)
    for q in range(wl.grid.nq):
        assert ctx.forest_block(q) == o.forest_block(q)


@pytest.mark.parametrize("emf", [0, 1])
def test_gpu_forest_substep(dev, emf):
    import torch
    wl = W.forest_wave(n=8, ngz=4, b_loop=0.2, p_in=3.0, r_in=0.3, emf=emf, **FORESTS["three"])
    say(f"[forest substep] emf={emf}")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    o.ghosts(st, B)
    ctx = ctx_of(wl)
    # both sides run their own c2p: it also sets the cell-centred E the GS EMFs read
    sg, Bg, ag = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
    o.aux(st, B, aux)
    ctx.compute_auxiliaries(sg, Bg, ag)
    dt = 0.25 * wl.extra["dx"] / 2
    new_o, nB_o = st.copy(), [b.copy() for b in B]
    o.substep(dt, 1.0, new_o, st, nB_o, B, aux)
    ng_, nBg = sg.clone(), [b.clone() for b in Bg]
    dtd = torch.tensor([dt], dtype=torch.float64, device=dev)
    ctx.advance_substep(dtd, 1.0, ng_, sg, nBg, Bg, ag)
    torch.cuda.synchronize()
    gn = to_host(ng_)
    check_cons(wl, gn, new_o, aux, 1e-12)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, to_host(nBg[d]), d), face_interior(wl.grid, nB_o[d], d))
        assert e < 1e-12, (d, e)


@pytest.mark.parametrize("name,emf", [("three", 0), ("three", 1), ("outflow", 0), ("mixed_bc", 0)])
def test_gpu_forest_rk3(dev, name, emf):
    import torch
    wl = W.forest_wave(n=8, ngz=4, b_loop=0.2, p_in=3.0, r_in=0.3, emf=emf, **FORESTS[name])
    say(f"[forest rk3] {name} emf={emf}")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    o.ghosts(st, B)
    ctx = ctx_of(wl)
    sg, Bg, ag = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
    o.aux(st, B, aux)
    ctx.compute_auxiliaries(sg, Bg, ag)
    periodic = all(b == 1 for b in wl.forest.bc)
    tot0 = volume_integrals(o, wl, st)
    dtd = torch.zeros(1, dtype=torch.float64, device=dev)
    for _ in range(2):
        dt = o.dt(st, aux, 0.25)
        ctx.compute_dt(sg, ag, 0.25, dtd)
        torch.cuda.synchronize()
        assert abs(float(dtd[0]) - dt) <= 1e-14 * dt
        o.rk3(dt, st, B, aux)
        dtd.fill_(dt)
        ctx.rk3_step(dtd, sg, Bg, ag)
    torch.cuda.synchronize()
    gs, gB, gaux = to_host(sg), [to_host(b) for b in Bg], to_host(ag)
    check_cons(wl, gs, st, aux, 1e-12)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, gB[d], d), face_interior(wl.grid, B[d], d))
        assert e < 1e-12, (d, e)
    for v in (ga.RHO, ga.PRESS, ga.ZVECX, ga.BX):
        assert linf_rel(gaux[:, v], aux[:, v]) < 1e-12, v
    n, ngz = wl.grid.nx, wl.grid.ngz
    db = div_b(gB, wl.forest.idx, n, ngz)[:, ~ghost_mask(n, ngz)]
    assert np.abs(db).max() < 1e-14, np.abs(db).max()
    if periodic:
        tot1 = volume_integrals(o, wl, gs)
        rel = np.abs(tot1 - tot0) / np.maximum(np.abs(tot0), 1e-3)
        assert rel.max() < 5e-14, rel
        assert interface_mismatch(o, wl, gB) < 1e-14


def test_gpu_forest_graph_matches_eager(dev):
    import torch
    wl = W.forest_wave(n=8, ngz=4, b_loop=0.2, p_in=3.0, r_in=0.3, **FORESTS["three"])
    say("[forest graph] 3 replays vs eager")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    o.ghosts(st, B)
    ctx = ctx_of(wl)
    sa, Ba, aa = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
    ctx.compute_auxiliaries(sa, Ba, aa)
    sb, Bb, ab = sa.clone(), [b.clone() for b in Ba], aa.clone()
    dtd = torch.tensor([0.25 * wl.extra["dx"] / 2], dtype=torch.float64, device=dev)
    s = torch.cuda.Stream()
    for it in range(3):
        ctx.rk3_step(dtd, sa, Ba, aa)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            ctx.rk3_step(dtd, sb, Bb, ab, stream=s.cuda_stream, graph=True)
        s.synchronize()
        assert torch.equal(sa, sb), it
        for d in range(3):
            assert torch.equal(Ba[d], Bb[d]), it
        assert torch.equal(aa, ab), it


@pytest.mark.parametrize("name,emf", [("three", W.EMF_UCT), ("outflow", W.EMF_UCT), ("three", W.EMF_GS)])
def test_gpu_forest_concurrent_stages_match_serial(dev, name, emf):
    """Forest RK3 with the stage launches side by side (the default) against
    the one-stream order (stage_streams = 0): same bits, eager and graph."""
    import torch
    wl = W.forest_wave(n=8, ngz=4, b_loop=0.2, p_in=3.0, r_in=0.3, emf=emf, **FORESTS[name])
    say(f"[forest concurrent] {name} emf {emf}")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    o.ghosts(st, B)
    p1 = type(wl.params).from_buffer_copy(wl.params)
    p1.stage_streams = 0
    # one context per array set: a context keeps the cell-centred E of its
    # last c2p, which the next GS step's edge EMFs read
    ser = ga.Context(p1, wl.grid, wl.instances, forest=wl.forest)
    con, cgr = ctx_of(wl), ctx_of(wl)
    sa, Ba, aa = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
    sb, Bb, ab = sa.clone(), [b.clone() for b in Ba], aa.clone()
    sc, Bc, ac = sa.clone(), [b.clone() for b in Ba], aa.clone()
    ser.compute_auxiliaries(sa, Ba, aa)
    con.compute_auxiliaries(sb, Bb, ab)
    cgr.compute_auxiliaries(sc, Bc, ac)
    dtd = torch.tensor([0.25 * wl.extra["dx"] / 2], dtype=torch.float64, device=dev)
    s = torch.cuda.Stream()
    for it in range(3):
        ser.rk3_step(dtd, sa, Ba, aa)
        con.rk3_step(dtd, sb, Bb, ab)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            cgr.rk3_step(dtd, sc, Bc, ac, stream=s.cuda_stream, graph=True)
        s.synchronize()
        for x, y in ((sa, sb), (sa, sc), (aa, ab), (aa, ac)):
            assert torch.equal(x, y), it
        for d in range(3):
            assert torch.equal(Ba[d], Bb[d]) and torch.equal(Ba[d], Bc[d]), it


def test_forest_rejects_bad_input(dev):
    wl = W.forest_wave(n=8, ngz=4, **FORESTS["one"])
    p = wl.params
    f = ga.make_forest((2, 2, 2), (1,) * 6, (16.0, 16.0, 8.0), [0])   # non-cubic cells
    with pytest.raises(RuntimeError):
        ga.Context(p, wl.grid, wl.instances, forest=f)
    g = ga.Grid(8, 8, 8, 4, wl.grid.nq + 1)
    with pytest.raises(RuntimeError):
        ga.Context(p, g, wl.instances, forest=wl.forest)
    g = ga.Grid(8, 8, 8, 3, wl.grid.nq)
    with pytest.raises(RuntimeError):
        ga.Context(p, g, wl.instances, forest=wl.forest)
    f = ga.make_forest((2, 2, 2), (1, 0, 1, 1, 1, 1), (16.0, 16.0, 16.0), [0])   # one-sided periodicity
    with pytest.raises(RuntimeError, match="both be periodic or both not"):
        ga.Context(p, wl.grid, wl.instances, forest=f)


def _torus_forest_small():
    # 2^3 coarse blocks of 8^3 on [-16, 16]^3 (hole, excised ring, torus
    # inner edge and atmosphere), two of them refined
    return W.fm_torus_forest(nb=(2, 2, 2), n=8, ngz=4, extent=16.0, refined=[0, 7])


def test_gpu_torus_forest_rk3(dev):
    """Curved Cowling metric on a forest: one RK3 step of the FM torus against
    the oracle (metric of every ghost cell from the initial data), and the
    same step with the static-metric derivative cache bit-identical."""
    import torch
    wl = _torus_forest_small()
    say(f"[torus forest] nq={wl.grid.nq}")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    st0 = st.copy()
    o.metric_ghosts(st)          # the reference's ghost metric (copies + 5-point Lagrange)
    assert not np.array_equal(st[:, 7:23], st0[:, 7:23])
    init = st.copy()
    dt = 0.25 * wl.extra["dx"] / 2
    o.rk3(dt, st, B, aux)
    res = []
    for cache in (False, True):
        ctx = ctx_of(wl)
        sg, Bg, ag = to_dev(st0, dev), [to_dev(b, dev) for b in wl.B], to_dev(wl.aux, dev)
        ctx.fill_metric_ghosts(sg)
        torch.cuda.synchronize()
        assert linf_rel(to_host(sg)[:, 7:23], init[:, 7:23]) < 1e-14
        if cache:
            ctx.set_static_metric(sg)
        dtd = torch.tensor([dt], dtype=torch.float64, device=dev)
        ctx.rk3_step(dtd, sg, Bg, ag)
        torch.cuda.synchronize()
        res.append((to_host(sg), [to_host(b) for b in Bg], to_host(ag)))
    gs, gB, gaux = res[0]
    for arr in [gs, gaux] + gB:
        assert np.isfinite(arr).all()
    check_cons(wl, gs, st, aux, 1e-12)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, gB[d], d), face_interior(wl.grid, B[d], d))
        assert e < 1e-12, (d, e)
    for v in (ga.RHO, ga.PRESS, ga.ZVECX):
        assert linf_rel(gaux[:, v], aux[:, v]) < 1e-11, v
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][2], res[1][2])
    for d in range(3):
        assert np.array_equal(res[0][1][d], res[1][1][d])


def test_gpu_torus_forest_learned_c2p(dev):
    """Learned c2p guess in the forest's RK3 cell passes (network trained on
    the forest's cells): agrees with the plain step to 1e-9, guess used in
    most cells."""
    import torch
    wl = W.fm_torus_forest(nb=(2, 2, 2), n=16, ngz=4, extent=24.0, refined=[0, 7])
    say(f"[torus forest learned] nq={wl.grid.nq}")
    dt = torch.tensor([0.25 * wl.extra["dx"] / 2], dtype=torch.float64, device=dev)
    out = {}
    mlp = ga.Mlp()
    for mode in ("plain", "learned"):
        ctx = ctx_of(wl)
        st, B, aux = to_dev(wl.state, dev), [to_dev(b, dev) for b in wl.B], to_dev(wl.aux, dev)
        if mode == "learned":
            rep = mlp.fit_cells(ctx, st, B, max_states=1 << 15, opts=ga.Mlp.default_fit_opts(epochs=40))
            say(f"  fit report {rep}")
            ctx.set_learned_c2p(mlp)
        ctx.rk3_step(dt, st, B, aux)
        torch.cuda.synchronize()
        res = (to_host(st), [to_host(b) for b in B], to_host(aux))
        ctx.compute_auxiliaries(st, B, aux)
        torch.cuda.synchronize()
        out[mode] = res + (ctx.c2p_stats(),)
    (sp, Bp, ap, _), (sl, Bl, al, stl) = out["plain"], out["learned"]
    for v in range(7):
        assert linf_rel(interior(wl.grid, sl[:, v]), interior(wl.grid, sp[:, v])) < 1e-9, v
    for d in range(3):
        assert linf_rel(face_interior(wl.grid, Bl[d], d), face_interior(wl.grid, Bp[d], d)) < 1e-9, d
    assert np.array_equal(al[:, ga.C2P_ERR], ap[:, ga.C2P_ERR])
    assert stl["n_guess_used"] > 0.5 * stl["n_cells"]


@pytest.mark.parametrize("ghost_emf,case", [(0, "smooth"), (0, "harsh"), (1, "smooth"), (1, "harsh")])
def test_gpu_forest_fofc_substep(dev, ghost_emf, case):
    """FOFC (GS) on a forest, one substep, both edge ranges, against the oracle
    at 1e-12.  (With the reference's edge range the flag pass decides on
    tentative ghost-layer states built from unwritten edges; over RK3 stages
    the last-bit differences of the stage combination -- contracted on the
    GPU, not in the oracle -- flip some of those decisions, so that mode is
    compared per substep.)"""
    import torch
    from test_oracle_amr import FOFC_HARSH, FOFC_SMOOTH
    kw = FOFC_HARSH if case == "harsh" else FOFC_SMOOTH
    wl = W.forest_wave(nb=(2, 2, 2), n=8, ngz=4, refined=[0, 3], emf=1, **kw)
    wl.params.fofc = 1
    wl.params.fofc_ghost_emf = ghost_emf
    say(f"[forest fofc substep] ghost_emf={ghost_emf} {case}")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    o.ghosts(st, B)
    ctx = ctx_of(wl)
    sg, Bg, ag = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
    o.aux(st, B, aux)
    ctx.compute_auxiliaries(sg, Bg, ag)
    dt = o.dt(st, aux, 0.25)
    new_o, nB_o = st.copy(), [b.copy() for b in B]
    o.substep(dt, 1.0, new_o, st, nB_o, B, aux)
    assert o.fofc_flagged() > 0
    ng_, nBg = sg.clone(), [b.clone() for b in Bg]
    ctx.advance_substep(torch.tensor([dt], dtype=torch.float64, device=dev), 1.0, ng_, sg, nBg, Bg, ag)
    torch.cuda.synchronize()
    gn = to_host(ng_)
    check_cons(wl, gn, new_o, aux, 1e-12)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, to_host(nBg[d]), d), face_interior(wl.grid, nB_o[d], d))
        assert e < 1e-12, (d, e)


@pytest.mark.parametrize("case", ["smooth", "harsh"])
def test_gpu_forest_fofc_rk3(dev, case):
    """FOFC with the ghost-layer edge EMFs (fofc_ghost_emf = 1) on a forest:
    two RK3 steps against the oracle (1e-12; the entropy tracer 1e-11 on the
    harsh blast, pressure ratio 1e9, where FOFC is active in tens of
    thousands of cells), and the GPU's own result conserves the volume
    integrals and keeps the coarse-fine faces consistent."""
    import torch
    from test_oracle_amr import FOFC_HARSH, FOFC_SMOOTH
    kw = FOFC_HARSH if case == "harsh" else FOFC_SMOOTH
    wl = W.forest_wave(nb=(2, 2, 2), n=8, ngz=4, refined=[0, 3], emf=1, **kw)
    wl.params.fofc = 1
    wl.params.fofc_ghost_emf = 1
    say(f"[forest fofc rk3] {case}")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    o.ghosts(st, B)
    ctx = ctx_of(wl)
    sg, Bg, ag = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
    o.aux(st, B, aux)
    ctx.compute_auxiliaries(sg, Bg, ag)
    tot0 = volume_integrals(o, wl, st)
    dtd = torch.zeros(1, dtype=torch.float64, device=dev)
    for _ in range(2):
        dt = o.dt(st, aux, 0.25)
        o.rk3(dt, st, B, aux)
        dtd.fill_(dt)
        ctx.rk3_step(dtd, sg, Bg, ag)
    torch.cuda.synchronize()
    assert o.fofc_flagged() > 0
    gs, gB = to_host(sg), [to_host(b) for b in Bg]
    for v in range(7):
        e = linf_rel(interior(wl.grid, gs[:, v]), interior(wl.grid, st[:, v]))
        assert e < (1e-11 if v == 6 else 1e-12), (v, e)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, gB[d], d), face_interior(wl.grid, B[d], d))
        assert e < 1e-12, (d, e)
    tot1 = volume_integrals(o, wl, gs)
    assert (np.abs(tot1 - tot0) / np.maximum(np.abs(tot0), 1e-3)).max() < 5e-14
    assert interface_mismatch(o, wl, gB) < 1e-14


def test_gpu_forest_workgroup_shapes_bit_identical(dev):
    """Workgroup sizes (grmhd_autotune / grmhd_set_tuning) change only the
    launch shapes: an RK3 step on a forest is bit-identical under two
    different tunings, and grmhd_autotune leaves the caller's arrays alone."""
    import torch
    wl = W.forest_wave(n=8, ngz=4, b_loop=0.2, p_in=3.0, r_in=0.3, **FORESTS["three"])
    dtd = torch.tensor([0.25 * wl.extra["dx"] / 2], dtype=torch.float64, device=dev)
    out = []
    for mode in ("default", "other", "autotuned"):
        ctx = ctx_of(wl)
        sg, Bg, ag = to_dev(wl.state, dev), [to_dev(b, dev) for b in wl.B], to_dev(wl.aux, dev)
        ctx.compute_auxiliaries(sg, Bg, ag)
        if mode == "other":
            ctx.set_tuning(flux=(64, 128, 64), emf=128, update=128, gather_faces=64, gather_cells=128, c2p=128)
        elif mode == "autotuned":
            before = (sg.clone(), [b.clone() for b in Bg], ag.clone())
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                ctx.autotune(sg, Bg, ag, reps=1, stream=s.cuda_stream)
            s.synchronize()
            assert torch.equal(sg, before[0]) and torch.equal(ag, before[2])
            for d in range(3):
                assert torch.equal(Bg[d], before[1][d])
        ctx.rk3_step(dtd, sg, Bg, ag)
        torch.cuda.synchronize()
        out.append((sg, Bg, ag))
    for o in out[1:]:
        assert torch.equal(o[0], out[0][0]) and torch.equal(o[2], out[0][2])
        for d in range(3):
            assert torch.equal(o[1][d], out[0][1][d])


def test_gpu_fully_refined_forest_equals_uniform_grid(dev):
    """Every coarse block refined: the library's forest path (no coarse-fine
    interface left) reproduces its uniform-grid path on the fine spacing to
    round-off after two RK3 steps (block order permuted; the periodic faces'
    two initial copies differ in the last bit, see the oracle test)."""
    import copy
    import torch
    from test_oracle_amr import _uniform_twin
    wf = W.forest_wave(nb=(2, 2, 2), n=8, ngz=4, refined=list(range(8)), b_loop=0.2, p_in=3.0, r_in=0.3)
    spec, qmap = _uniform_twin(wf)
    inv = np.argsort(qmap)
    dtd = torch.tensor([0.25 * wf.extra["dx"] / 2], dtype=torch.float64, device=dev)
    res = []
    for forest in (True, False):
        if forest:
            ctx = ctx_of(wf)
            st, B, aux = wf.state, wf.B, wf.aux
        else:
            ctx = ga.Context(wf.params, wf.grid, ga.make_instances(spec))
            st, B, aux = wf.state[inv], [b[inv] for b in wf.B], wf.aux[inv]
        sg, Bg, ag = to_dev(st, dev), [to_dev(b, dev) for b in B], to_dev(aux, dev)
        ctx.fill_ghosts(sg, Bg)
        ctx.compute_auxiliaries(sg, Bg, ag)
        for _ in range(2):
            ctx.rk3_step(dtd, sg, Bg, ag)
        torch.cuda.synchronize()
        res.append((to_host(sg), [to_host(b) for b in Bg], to_host(ag)))
    (sf, Bf, af), (su, Bu, au) = res
    for v in range(7):
        assert linf_rel(sf[:, v], su[qmap][:, v]) < 1e-13, v
    for d in range(3):
        assert linf_rel(Bf[d], Bu[d][qmap]) < 1e-13, d
    for v in (ga.RHO, ga.PRESS, ga.ZVECX, ga.BX):
        assert linf_rel(af[:, v], au[qmap][:, v]) < 1e-13, v


@pytest.mark.parametrize("name", ["centre", "outflow"])
def test_gpu_forest_metric_exchange(dev, name):
    """grmhd_fill_metric_ghosts against the oracle: copies bit-exact, 5-point
    Lagrange restriction / prolongation to 1e-14 (the GPU contracts the pair
    sums into FMAs), ghosts beyond outflow faces untouched."""
    import torch
    from test_oracle_amr import _metric_field
    wl = W.forest_wave(n=8, ngz=4, **FORESTS[name])
    o = O.Stepper(wl)
    f = lambda x, y, z: np.cos(0.37 * x) * (1.0 + 0.1 * y) + np.exp(-0.01 * (x * x + z * z)) * np.sin(0.2 * z + 0.3)
    st, B, aux, blocks, index = _metric_field(wl, o, f)
    st[:, 8] = 0.5 * st[:, 7] + 0.25     # a second metric variable
    n, ngz = wl.grid.nx, wl.grid.ngz
    gm = ghost_mask(n, ngz)
    st[:, 7:9][:, :, gm] = -3.0           # ghosts to be filled (or kept, beyond outflow faces)
    sg = to_dev(st, dev)
    o.metric_ghosts(st)
    ctx = ctx_of(wl)
    ctx.fill_metric_ghosts(sg)
    torch.cuda.synchronize()
    gs = to_host(sg)
    for v in (7, 8):
        assert linf_rel(gs[:, v], st[:, v]) < 1e-14, v
    assert np.array_equal(gs[:, 9:], st[:, 9:])
    assert np.array_equal(gs[:, :7], st[:, :7])

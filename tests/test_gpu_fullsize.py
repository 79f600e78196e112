"""GPU parity at the benchmark sizes, against the CPU oracle on the same inputs.

- 3-D blast, 128^3 cells in 512 blocks of 16^3 (ngz 3, WENO-Z + HLLE + UCT):
  one RK3 step replayed from the graph, every cell against the oracle's step.
- Fishbone-Moncrief torus on the 440-block two-level forest (6^3 coarse
  blocks of 16^3, central 4x4x2 refined, Kerr-Schild a = 0.9375): one RK3
  step with the static-metric cache, against the oracle.
- c2p alone on 4,194,304 Kerr-Schild states (a = 0.9, r in [1.5, 50] M):
  plain and learned (network trained on other states) against the oracle.
- 64 uneven blast instances (sides from {32, 48, 64, 80, 96}, 5,361 blocks):
  one RK3 graph step of the whole list; instances never exchange data, so the
  oracle steps five of them (one of each side) alone and those must agree;
  every instance keeps div B at round-off and finite values.

Tolerances as in test_gpu_parity: conserved and face B L-infinity relative
<= 1e-12 per variable (atmosphere and body separately; the torus's
atmosphere 1e-10, see that test) and the per-cell floored error; c2p primitives <= 1e-10 of the state scale where the error
bits agree, a bounded count of states whose bits differ (a rounding
difference straddling a branch threshold) and those within 1e-8.
"""
import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O
from test_gpu_parity import (say, to_dev, to_host, interior, face_interior, linf_rel, cell_rel, check_cons,
                             run_rk3_pair, check_close, CELL_TOL)

pytestmark = pytest.mark.gpu


def test_gpu_rk3_blast128_full(dev):
    wl = W.blast()
    assert wl.grid.nq == 512 and wl.zones == 128 ** 3
    say("[full] blast 128^3, 512 blocks of 16^3, one RK3 graph step vs oracle")
    ref, got, _ = run_rk3_pair(wl, dev, steps=1, graph=True)
    check_close(wl, ref, got, 1e-12)
    for v in (ga.RHO, ga.PRESS, ga.EPS, ga.ZVECX, ga.ZVECY, ga.ZVECZ, ga.BX, ga.BY, ga.BZ):
        e = linf_rel(got[2][:, v], ref[2][:, v])
        assert e < 1e-12, (v, e)
    assert np.array_equal(got[2][:, ga.C2P_ERR], ref[2][:, ga.C2P_ERR])
    bdiv = interior(wl.grid, got[2][:, ga.BDIV])
    say(f"  max |div B| {np.max(np.abs(bdiv)):.2e}")
    assert np.max(np.abs(bdiv)) < 1e-12


def test_gpu_rk3_torus_forest440_full(dev):
    import torch
    wl = W.fm_torus_forest()
    assert wl.grid.nq == 440
    say(f"[full] FM torus on the 440-block forest ({wl.zones} zones), one RK3 step vs oracle")
    st, B, aux = wl.copy_arrays()
    o = O.Stepper(wl)
    st0 = st.copy()
    o.metric_ghosts(st)
    dt = 0.25 * wl.extra["dx"] / 2
    o.rk3(dt, st, B, aux)
    ctx = ga.Context(wl.params, wl.grid, wl.instances, forest=wl.forest)
    sg, Bg, ag = to_dev(st0, dev), [to_dev(b, dev) for b in wl.B], to_dev(wl.aux, dev)
    ctx.fill_metric_ghosts(sg)
    ctx.set_static_metric(sg)
    dtd = torch.tensor([dt], dtype=torch.float64, device=dev)
    ctx.rk3_step(dtd, sg, Bg, ag)
    torch.cuda.synchronize()
    gs, gB, gaux = to_host(sg), [to_host(b) for b in Bg], to_host(ag)
    for a in [gs, gaux] + gB:
        assert np.isfinite(a).all()
    # atmosphere cells beside the torus take fluxes built from body values
    # 1e5-1e8 times their own density: the round-off of those (FMA
    # contraction on the GPU, none in the oracle) is ~1e-11 of the
    # atmosphere's scale at this resolution; every cell's floored error is
    # still bounded by CELL_TOL
    check_cons(wl, gs, st, aux, 1e-12, atm_tol=1e-10, what="torus forest 440")
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, gB[d], d), face_interior(wl.grid, B[d], d))
        assert e < 1e-12, (d, e)
    # primitives: every cell of the torus body to 1e-11 of the variable's
    # scale; cells within three decades of the atmosphere density invert
    # conserved values whose round-off is set by the neighbouring body (see
    # above), so there each cell's floored error is bounded instead
    body = aux[:, ga.RHO] >= 1e3 * wl.params.rho_fl
    for v in (ga.RHO, ga.PRESS, ga.ZVECX, ga.ZVECY, ga.ZVECZ):
        e = linf_rel(gaux[:, v], aux[:, v], body)
        ec = cell_rel(gaux[:, v], aux[:, v])
        assert e < 1e-11 and ec < CELL_TOL, (v, e, ec)
    same = gaux[:, ga.C2P_ERR] == aux[:, ga.C2P_ERR]
    say(f"  C2P_ERR identical in {same.mean():.7f} of cells")
    assert same.mean() > 0.9999


def _c2p_compare(label, gp, ge, op, oe, n):
    from test_oracle_c2p import prim_residual
    same = ge == oe
    nd = int((~same).sum())
    res = prim_residual(gp, op)
    say(f"  {label}: error bits differ in {nd} of {n} states; max state-normalised prim diff "
        f"{res[same].max():.3e} (same bits), {res[~same].max() if nd else 0.0:.3e} (differing)")
    assert nd <= 1e-4 * n
    assert res[same].max() < 1e-10
    if nd:
        assert res[~same].max() < 1e-8


def test_gpu_c2p_4m_states_plain_and_learned(dev):
    import torch
    n = 1 << 22
    say(f"[full] c2p on {n} Kerr-Schild states, plain and learned, vs oracle")
    st = W.c2p_states(n, seed=W.SEED)
    params = st["params"]
    grid = ga.Grid(1, 1, 1, 3, 1)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))])
    ctx = ga.Context(params, grid, inst)
    op, oc, oe, of, ev, fatal = O.c2p_batch(params, st["cons"], st["met"], r=st["rad"])
    # network trained on other states of the same distribution
    ntr = 1 << 20
    tr = W.c2p_states(ntr, seed=W.SEED + 11)
    mlp = ga.Mlp()
    rep = mlp.fit(ctx, ntr, to_dev(tr["cons"], dev), to_dev(tr["met"], dev), to_dev(tr["rad"], dev),
                  opts=ga.Mlp.default_fit_opts(epochs=40))
    say(f"  fit report {rep}")
    del tr
    for mode in ("plain", "learned"):
        cons = to_dev(st["cons"], dev)
        met = to_dev(st["met"], dev)
        rad = to_dev(st["rad"], dev)
        prims = torch.zeros((13, n), dtype=torch.float64, device=dev)
        err = torch.zeros(n, dtype=torch.int64, device=dev)
        fl = torch.zeros(n, dtype=torch.uint8, device=dev)
        if mode == "plain":
            ctx.c2p_batch(n, cons, met, prims, err, fl, r=rad)
        else:
            ga.c2p_batch_learned(ctx, mlp, n, cons, met, prims, err, fl, r=rad)
        torch.cuda.synchronize()
        stats = ctx.c2p_stats()
        say(f"  {mode} stats {stats}")
        _c2p_compare(mode, to_host(prims), to_host(err).astype(np.uint64), op, oe, n)
        assert stats["n_fatal"] == fatal
        if mode == "learned":
            assert stats["n_guess_used"] > 0.5 * n
        del cons, met, rad, prims, err, fl


def _instance_subset(wl, picks):
    """Workload of the instances `picks` of wl alone (their blocks copied, q0
    renumbered from 0)."""
    specs, blocks = [], []
    q0 = 0
    for i in picks:
        s = dict(wl.inst_specs[i])
        nbk = s["nb"][0] * s["nb"][1] * s["nb"][2]
        blocks.append((s["q0"], nbk))
        s["q0"] = q0
        specs.append(s)
        q0 += nbk
    sel = np.concatenate([np.arange(a, a + b) for a, b in blocks])
    g = wl.grid
    grid = ga.Grid(g.nx, g.ny, g.nz, g.ngz, q0)
    sub = W.Workload("subset", wl.params, grid, specs, wl.state[sel].copy(), [b[sel].copy() for b in wl.B],
                     wl.aux[sel].copy(), zones=q0 * g.nx * g.ny * g.nz)
    return sub, sel


def test_gpu_rk3_batched_64_instances_full(dev):
    import torch
    wl = W.multi_blast()
    sides = [int(s) for s in wl.extra["sides"]]
    say(f"[full] 64 uneven blast instances, {wl.grid.nq} blocks, {wl.zones} zones")
    assert len(sides) == 64
    picks = [sides.index(s) for s in (32, 48, 64, 80, 96)]
    sub, sel = _instance_subset(wl, picks)
    # one dt for all instances, as the batched step takes one device dt
    dt = 0.25 / 96.0
    st, B, aux = sub.copy_arrays()
    O.Stepper(sub).rk3(dt, st, B, aux)
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = to_dev(wl.state, dev), [to_dev(b, dev) for b in wl.B], to_dev(wl.aux, dev)
    dtd = torch.tensor([dt], dtype=torch.float64, device=dev)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx.rk3_step(dtd, sg, Bg, ag, stream=s.cuda_stream, graph=True)
    s.synchronize()
    gs = to_host(sg)
    gaux = to_host(ag)
    gB = [to_host(b) for b in Bg]
    del sg, Bg, ag
    for a in [gs, gaux] + gB:
        assert np.isfinite(a).all()
    bdiv = interior(wl.grid, gaux[:, ga.BDIV])
    say(f"  all instances: max |div B| {np.max(np.abs(bdiv)):.2e}")
    assert np.max(np.abs(bdiv)) < 1e-12
    got = (gs[sel], [b[sel] for b in gB], gaux[sel])
    check_close(sub, (st, B, aux), got, 1e-12)
    assert np.array_equal(got[2][:, ga.C2P_ERR], aux[:, ga.C2P_ERR])

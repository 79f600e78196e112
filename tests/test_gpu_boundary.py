"""GPU tests of the boundary's less-travelled entry points and options,
against the oracle: prims_to_conservs over a batch (c2p.cpp:416-455), LLF
with the light-cone floor (grmhd.hh:874-883, rusanov_c_limit > 0),
radius-based excision (c2p.cpp:345-348) in the RK3 cell passes, the
option combinations the library rejects instead of computing a wrong
answer, the host-dt entry points (advance_substep(t, dt, dtfact, ...),
find_stable_timestep) and the fatal-c2p report at sync points.  Tolerances
as in test_gpu_parity."""
import ctypes as C

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O
from test_gpu_parity import (check_cons, dev_arrays, to_dev, to_host, interior, face_interior, linf_rel, say,
                             run_rk3_pair, check_close)

pytestmark = pytest.mark.gpu


def test_gpu_prims_to_conservs_batch(dev):
    """p2c of the Kerr-Schild c2p states against the oracle, and the round
    trip p2c -> c2p back to the primitives."""
    import torch
    n = 20000
    say(f"[p2c batch] n={n}")
    st = W.c2p_states(n, seed=W.SEED + 7)
    grid = ga.Grid(1, 1, 1, 3, 1)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))])
    ctx = ga.Context(st["params"], grid, inst)
    prims, met = to_dev(st["prims"], dev), to_dev(st["met"], dev)
    cons = torch.zeros((10, n), dtype=torch.float64, device=dev)
    cons[7:10] = to_dev(st["cons"][7:10], dev)   # densitised B passes through
    ga.check(ctx.L.grmhd_prims_to_conservs_batch(ctx.h, n, ga._ptr(prims), ga._ptr(met), ga._ptr(cons), None),
             "prims_to_conservs_batch")
    torch.cuda.synchronize()
    gc = to_host(cons)
    oc = np.zeros((10, n))
    oc[7:10] = st["cons"][7:10]
    O.olib().orc_p2c_batch(C.byref(st["params"]), n, O.p(st["prims"]), O.p(st["met"]), O.p(oc))
    scale = np.abs(oc[0:7]).max(axis=0) + np.abs(oc[4]) + 1e-300
    err = np.abs(gc[0:7] - oc[0:7]).max(axis=0) / scale
    say(f"  max per-state relative difference {err.max():.3e}")
    assert err.max() < 1e-13
    # round trip through the library's c2p
    pr = torch.zeros((13, n), dtype=torch.float64, device=dev)
    eb = torch.zeros(n, dtype=torch.int64, device=dev)
    fl = torch.zeros(n, dtype=torch.uint8, device=dev)
    ctx.c2p_batch(n, cons, met, pr, eb, fl)
    torch.cuda.synchronize()
    from test_oracle_c2p import prim_residual
    res = prim_residual(to_host(pr), st["prims"])
    ok = to_host(fl) == 0
    say(f"  round trip: {ok.mean():.4f} not floored, max state residual {res[ok].max():.3e}")
    assert ok.mean() > 0.99 and np.quantile(res[ok], 0.99) < 1e-8


def test_gpu_llf_light_cone_floor(dev):
    """LLF with rusanov_c_limit above the ambient density: every face takes
    the light-cone speed; substep against the oracle."""
    import torch
    wl = W.blast(nb=(2, 2, 2), n=16, riemann=ga.RIEMANN_LLF)
    wl.params.rusanov_c_limit = 0.05      # ambient rho 1e-2 < 0.05: the floor applies
    say("[llf c_limit] blast substep")
    dt = 0.25 * wl.extra["dx"]
    st, B, aux = wl.copy_arrays()
    new_o, nB_o = st.copy(), [b.copy() for b in B]
    O.Stepper(wl).substep(dt, 1.0, new_o, st, nB_o, B, aux)
    # the option changes the result (the floor is active)
    wl2 = W.blast(nb=(2, 2, 2), n=16, riemann=ga.RIEMANN_LLF)
    s2, B2, a2 = wl2.copy_arrays()
    n2, nB2 = s2.copy(), [b.copy() for b in B2]
    O.Stepper(wl2).substep(dt, 1.0, n2, s2, nB2, B2, a2)
    assert np.abs(interior(wl.grid, n2[:, ga.DENS] - new_o[:, ga.DENS])).max() > 1e-8
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = dev_arrays(wl, dev)
    ng, nBg = sg.clone(), [b.clone() for b in Bg]
    ctx.advance_substep(torch.tensor([dt], dtype=torch.float64, device=dev), 1.0, ng, sg, nBg, Bg, ag)
    torch.cuda.synchronize()
    check_cons(wl, to_host(ng), new_o, aux, 1e-12, what="llf c_limit")
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, to_host(nBg[d]), d), face_interior(wl.grid, nB_o[d], d))
        assert e < 1e-12, (d, e)


def test_gpu_radius_excision_torus(dev):
    """excise_by_radius with the Kerr-Schild radius (r_ex = 2.5 M, outside
    the horizon of a = 0.9375): one RK3 step of the 32^3 torus against the
    oracle, the excised cells at atmosphere on both sides."""
    wl = W.fm_torus(nb=(2, 2, 2), n=16, extent=24.0)
    wl.params.excise_by_radius = 1
    wl.params.excision_radius = 2.5
    wl.params.excision_cks = 1
    wl.params.excision_spin = 0.9375
    say("[radius excision] FM torus 32^3, one step")
    ref, got, _ = run_rk3_pair(wl, dev, steps=1)
    check_close(wl, ref, got, 1e-12)
    atmo_bit = np.uint64(1 << ga.C2P_BITS.index("ATMO_RESET"))
    gbits, obits = got[2][:, ga.C2P_ERR].astype(np.uint64), ref[2][:, ga.C2P_ERR].astype(np.uint64)
    assert np.array_equal(gbits, obits)
    assert ((gbits & atmo_bit) != 0).sum() > 0


def _create_status(params, grid, inst, n_inst):
    h = C.c_void_p()
    return ga.lib().grmhd_ctx_create(C.byref(params), C.byref(grid), inst, n_inst, C.byref(h)), h


def test_gpu_rejected_option_combinations(dev):
    """Radius excision with periodic faces, and a curved metric with unequal
    spacings, are refused with GRMHD_ERR_UNSUPPORTED (4) and a message."""
    grid = ga.Grid(8, 8, 8, 3, 1)
    p = ga.default_params(excise_by_radius=1, flat_metric=1)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(1,) * 6, idx=(8.0,) * 3)])
    st, h = _create_status(p, grid, inst, 1)
    assert st == 4 and b"periodic" in ga.lib().grmhd_last_error()
    p = ga.default_params(flat_metric=0)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(8.0, 8.0, 4.0))])
    st, h = _create_status(p, grid, inst, 1)
    assert st == 4 and b"equal spacing" in ga.lib().grmhd_last_error()
    # the same spacing with the flat metric is accepted
    p = ga.default_params(flat_metric=1)
    st, h = _create_status(p, grid, inst, 1)
    assert st == 0
    ga.lib().grmhd_ctx_destroy(h)


def test_gpu_host_dt_entry_points(dev):
    """advance_substep(t, dt, dtfact, ...) with a host dt (evolve.hh:252-256)
    gives the bits of the device-dt substep; find_stable_timestep returns the
    device dt on the host (find_stable_timestep.cpp:72-126), equal to the
    oracle's to 1e-14."""
    import torch
    wl = W.blast(nb=(2, 2, 2), n=16)
    say("[host dt] substep with a host dt, find_stable_timestep")
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = dev_arrays(wl, dev)
    dt = ctx.find_stable_timestep(sg, ag, 0.25)
    dtd = torch.zeros(1, dtype=torch.float64, device=dev)
    ctx.compute_dt(sg, ag, 0.25, dtd)
    torch.cuda.synchronize()
    assert dt == float(dtd.item())
    dto = O.Stepper(wl).dt(wl.state, wl.aux, 0.25)
    assert abs(dt - dto) <= 1e-14 * dto
    outs = []
    for host in (False, True):
        ng, nBg = sg.clone(), [b.clone() for b in Bg]
        if host:
            ctx.advance_substep_t(0.125, dt, 0.5, ng, sg, nBg, Bg, ag)
        else:
            ctx.advance_substep(dtd, 0.5, ng, sg, nBg, Bg, ag)
        assert ctx.sync() == 0
        outs.append((ng, nBg))
    assert torch.equal(outs[0][0], outs[1][0])
    for d in range(3):
        assert torch.equal(outs[0][1][d], outs[1][1][d])
    assert not torch.equal(outs[0][0], sg)


def test_gpu_c2p_fatal_reported_at_sync(dev):
    """A cell whose density exceeds the EOS range (where the reference's c2p
    calls Kokkos::abort, c2p.hh:160-322) is counted by the c2p passes and
    reported as GRMHD_ERR_C2P_FATAL (5) at the next sync point, once; clean
    data report 0.  The oracle counts the same cell as fatal."""
    import torch
    wl = W.blast(nb=(2, 2, 2), n=16)
    say("[fatal] one cell above the EOS density range")
    st, B, aux = wl.copy_arrays()
    g = wl.grid.ngz
    st[3, ga.DENS, g + 5, g + 6, g + 7] = 1e101
    st[3, ga.TAU, g + 5, g + 6, g + 7] = 1e100
    for v in (ga.SX, ga.SY, ga.SZ):
        st[3, v, g + 5, g + 6, g + 7] = 0.0
    o = O.Stepper(wl)
    s_o, B_o, a_o = st.copy(), [b.copy() for b in B], aux.copy()
    o.aux(s_o, B_o, a_o)
    assert o.fatal() == 1
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = to_dev(wl.state, dev), [to_dev(b, dev) for b in wl.B], to_dev(wl.aux, dev)
    ctx.compute_auxiliaries(sg, Bg, ag)
    assert ctx.sync() == 0
    sb = to_dev(st, dev)
    ctx.compute_auxiliaries(sb, Bg, ag)
    assert ctx.sync() == 5
    assert b"reference aborts" in ga.lib().grmhd_last_error()
    assert ctx.sync() == 0          # cleared by the sync point that reported it
    # an RK3 step over the same data counts it in its c2p passes too
    dtd = torch.tensor([0.25 * wl.extra["dx"]], dtype=torch.float64, device=dev)
    ctx.rk3_step(dtd, sb, Bg, ag)
    assert ctx.sync() == 5
    # find_stable_timestep is a sync point too (on a fresh copy of the bad
    # state: the step's c2p has limited that cell's density)
    sc = to_dev(st, dev)
    ctx.compute_auxiliaries(sc, Bg, ag)
    with pytest.raises(RuntimeError, match="status 5"):
        ctx.find_stable_timestep(sc, ag, 0.25)

"""GPU parity in a curved Cowling metric: the blast wave of the flat tests
placed at r ~ 8 M in the Kerr-Schild metric of a spin-0.9 hole, so the face
metric interpolation, the geometric source terms with 6th-order metric
derivatives and the extrinsic curvature, and the curved c2p all run
(kernels with FLAT = false).  Tolerances as for the flat substep and step:
L-infinity relative <= 1e-12 per conserved variable / face-B component."""
import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O
from test_gpu_parity import check_cons, dev_arrays, to_host, interior, face_interior, linf_rel, say, run_rk3_pair, check_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("emf", [ga.EMF_UCT, ga.EMF_GS])
def test_gpu_curved_substep(dev, emf):
    import torch
    wl = W.blast_ks(emf=emf)
    say(f"[curved substep] Kerr-Schild blast, emf={emf}")
    dt = 0.25 * wl.extra["dx"]
    st, B, aux = wl.copy_arrays()
    new_o, nB_o = st.copy(), [b.copy() for b in B]
    O.Stepper(wl).substep(dt, 1.0, new_o, st, nB_o, B, aux)
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = dev_arrays(wl, dev)
    ng, nBg = sg.clone(), [b.clone() for b in Bg]
    ctx.advance_substep(torch.tensor([dt], dtype=torch.float64, device=dev), 1.0, ng, sg, nBg, Bg, ag)
    torch.cuda.synchronize()
    gn = to_host(ng)
    check_cons(wl, gn, new_o, aux, 1e-12)
    for v in range(7):
        du_o = interior(wl.grid, new_o[:, v] - st[:, v])
        if np.max(np.abs(du_o)) > 0:
            assert linf_rel(interior(wl.grid, gn[:, v] - st[:, v]), du_o) < 1e-9, v
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, to_host(nBg[d]), d), face_interior(wl.grid, nB_o[d], d))
        assert e < 1e-12, (d, e)


def test_gpu_curved_rk3(dev):
    wl = W.blast_ks()
    say("[curved rk3] Kerr-Schild blast, 2 steps")
    ref, got, _ = run_rk3_pair(wl, dev, steps=2)
    check_close(wl, ref, got, 1e-12)
    for v in (ga.RHO, ga.PRESS, ga.ZVECX, ga.ZVECY, ga.ZVECZ):
        assert linf_rel(got[2][:, v], ref[2][:, v]) < 1e-11, v
    assert np.max(np.abs(interior(wl.grid, got[2][:, ga.BDIV]))) < 1e-12


def test_gpu_static_metric_cache_bit_identical(dev):
    """grmhd_set_static_metric: an RK3 step reading the cached metric
    derivatives is bit-identical to the step that takes them every substep
    (Kerr-Schild blast, two steps, eager and graph)."""
    import torch
    wl = W.blast_ks()
    say("[static metric] cached derivatives vs per-substep")
    dt = torch.tensor([0.25 * wl.extra["dx"]], dtype=torch.float64, device=dev)
    res = []
    for cache in (False, True):
        ctx = ga.Context(wl.params, wl.grid, wl.instances)
        sg, Bg, ag = dev_arrays(wl, dev)
        if cache:
            ctx.set_static_metric(sg)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            ctx.rk3_step(dt, sg, Bg, ag, stream=s.cuda_stream)
            ctx.rk3_step(dt, sg, Bg, ag, stream=s.cuda_stream, graph=True)
        s.synchronize()
        res.append((to_host(sg), [to_host(b) for b in Bg], to_host(ag)))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][2], res[1][2])
    for d in range(3):
        assert np.array_equal(res[0][1][d], res[1][1][d])


def test_gpu_torus_rk3(dev):
    """Fishbone-Moncrief torus around a spin-0.9375 hole (the hole, its
    excised ring region, the atmosphere and the magnetised torus all on a
    32^3 grid of 16^3 blocks): one RK3 step against the oracle."""
    wl = W.fm_torus(nb=(2, 2, 2), n=16, extent=24.0)
    say("[torus rk3] FM torus 32^3, one step")
    ref, got, ctx = run_rk3_pair(wl, dev, steps=1)
    for arr in [got[0], got[2]] + list(got[1]):
        assert np.isfinite(arr).all()
    check_close(wl, ref, got, 1e-12)
    for v in (ga.RHO, ga.PRESS, ga.ZVECX, ga.ZVECY, ga.ZVECZ):
        assert linf_rel(got[2][:, v], ref[2][:, v]) < 1e-11, v
    bmax = max(np.abs(b).max() for b in got[1])
    assert np.max(np.abs(interior(wl.grid, got[2][:, ga.BDIV]))) * wl.extra["dx"] < 1e-12 * bmax

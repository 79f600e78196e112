"""GPU parity of the first-order flux correction (GS EMFs, ngz = 4): the
library's FOFC substep and RK3 steps against the oracle's (evolve.cpp:456-811)
on a blast wave where cells get flagged.  Tolerance as for the other
substeps: L-infinity relative <= 1e-12 per conserved variable and face-B
component over the interior, div B at round-off."""
import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O
from test_gpu_parity import check_cons, dev_arrays, to_host, interior, face_interior, linf_rel, say, run_rk3_pair, check_close

pytestmark = pytest.mark.gpu


def fofc_blast(**kw):
    return W.blast(nb=(2, 2, 2), n=16, ngz=4, emf=ga.EMF_GS, fofc=1, **kw)


def test_gpu_fofc_substep(dev):
    import torch
    wl = fofc_blast(rho=1e-6, p=1e-8)
    say("[fofc substep] low-density blast, GS, ngz 4")
    dt = 0.25 * wl.extra["dx"]
    st, B, aux = wl.copy_arrays()
    new_o, nB_o = st.copy(), [b.copy() for b in B]
    stepper = O.Stepper(wl)
    stepper.substep(dt, 1.0, new_o, st, nB_o, B, aux)
    say(f"  oracle flagged {stepper.fofc_flagged()} cells")
    assert stepper.fofc_flagged() > 0
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = dev_arrays(wl, dev)
    ng, nBg = sg.clone(), [b.clone() for b in Bg]
    dtd = torch.tensor([dt], dtype=torch.float64, device=dev)
    ctx.advance_substep(dtd, 1.0, ng, sg, nBg, Bg, ag)
    torch.cuda.synchronize()
    gn = to_host(ng)
    check_cons(wl, gn, new_o, aux, 1e-12)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, to_host(nBg[d]), d), face_interior(wl.grid, nB_o[d], d))
        assert e < 1e-12, (d, e)


def test_gpu_fofc_rk3(dev):
    wl = fofc_blast()
    say("[fofc rk3] blast, GS, ngz 4, 3 steps")
    ref, got, ctx = run_rk3_pair(wl, dev, steps=3)
    check_close(wl, ref, got, 1e-12)
    bdiv = interior(wl.grid, got[2][:, ga.BDIV])
    assert np.max(np.abs(bdiv)) < 1e-12


def test_gpu_fofc_rejects_uct_and_few_ghosts():
    wl = W.blast(nb=(1, 1, 1), n=8, ngz=4, fofc=1)        # UCT
    with pytest.raises(RuntimeError, match="UCT"):
        ga.Context(wl.params, wl.grid, wl.instances)
    wl = W.blast(nb=(1, 1, 1), n=8, ngz=3, emf=ga.EMF_GS, fofc=1)
    with pytest.raises(RuntimeError, match="ngz"):
        ga.Context(wl.params, wl.grid, wl.instances)


def test_gpu_fofc_torus_curved(dev):
    """FOFC in the curved metric: the Fishbone-Moncrief torus around the
    spin-0.9375 hole (GS, ngz 4), where the torus edge, the atmosphere and
    the excised region flag many cells; one substep and one RK3 step
    against the oracle."""
    import torch
    wl = W.fm_torus(nb=(2, 2, 2), n=16, ngz=4, extent=24.0, emf=ga.EMF_GS, fofc=1)
    say("[fofc torus] FM torus 32^3, GS, ngz 4")
    dt = 0.25 * wl.extra["dx"]
    st, B, aux = wl.copy_arrays()
    new_o, nB_o = st.copy(), [b.copy() for b in B]
    stepper = O.Stepper(wl)
    stepper.substep(dt, 1.0, new_o, st, nB_o, B, aux)
    say(f"  oracle flagged {stepper.fofc_flagged()} cells")
    assert stepper.fofc_flagged() > 1000
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    sg, Bg, ag = dev_arrays(wl, dev)
    ng, nBg = sg.clone(), [b.clone() for b in Bg]
    ctx.advance_substep(torch.tensor([dt], dtype=torch.float64, device=dev), 1.0, ng, sg, nBg, Bg, ag)
    torch.cuda.synchronize()
    gn = to_host(ng)
    check_cons(wl, gn, new_o, aux, 1e-12)
    for d in range(3):
        e = linf_rel(face_interior(wl.grid, to_host(nBg[d]), d), face_interior(wl.grid, nB_o[d], d))
        assert e < 1e-12, (d, e)
    ref, got, _ = run_rk3_pair(wl, dev, steps=1)
    check_close(wl, ref, got, 1e-12)
    bmax = max(np.abs(b).max() for b in got[1])
    assert np.max(np.abs(interior(wl.grid, got[2][:, ga.BDIV]))) * wl.extra["dx"] < 1e-12 * bmax

"""Autotuned launch shapes: grmhd_autotune reads but never modifies its
inputs, and neither workgroup sizes nor the face-flux kernel's stencil
prefetch change anything in the results (each thread's arithmetic is
independent of the launch shape and of when its loads are issued): RK3
steps under different tunings, eager and graph-replayed, are
bit-identical."""
import sys

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def dev_arrays(wl, dev):
    import torch
    return (torch.from_numpy(wl.state.copy()).to(dev), [torch.from_numpy(b.copy()).to(dev) for b in wl.B],
            torch.from_numpy(wl.aux.copy()).to(dev))


def test_gpu_autotune_bitexact(dev):
    import torch
    wl = W.blast(nb=(2, 2, 2), n=16)
    say("[autotune] blast 2x2x2 blocks of 16^3")
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    s0, B0, a0 = dev_arrays(wl, dev)
    keep = (s0.clone(), [b.clone() for b in B0], a0.clone())
    tune = ctx.autotune(s0, B0, a0, reps=2)
    torch.cuda.synchronize()
    say(f"  chosen {tune}")
    assert torch.equal(s0, keep[0]) and torch.equal(a0, keep[2])
    assert all(torch.equal(b, k) for b, k in zip(B0, keep[1]))
    vals = tune["flux"] + [tune[k] for k in ("emf", "update", "gather_faces", "gather_cells", "c2p")]
    assert all(v in (64, 128, 256) for v in vals)
    assert all(v in (0, 1) for v in tune["flux_prefetch"])
    dt = torch.tensor([0.25 * wl.extra["dx"]], dtype=torch.float64, device=dev)
    results = []
    stream = torch.cuda.Stream()
    for name, t in (("default", dict()), ("all64", dict(flux=(64, 64, 64), emf=64, update=64, gather_faces=64,
                                                         gather_cells=64, c2p=64)),
                    ("prefetch", dict(flux_prefetch=(1, 1, 1))),
                    ("prefetch_x_z_128", dict(flux=(128, 256, 128), flux_prefetch=(1, 0, 1))),
                    ("tuned", None)):
        if t is None:
            ctx.autotune(s0, B0, a0, reps=1)
        else:
            ctx.set_tuning(**t)
        for graph in (False, True):
            st, B, aux = dev_arrays(wl, dev)
            with torch.cuda.stream(stream):
                for _ in range(2):
                    ctx.rk3_step(dt, st, B, aux, stream=stream.cuda_stream, graph=graph)
            stream.synchronize()
            results.append((name, graph, st, B, aux))
    ref = results[0]
    for name, graph, st, B, aux in results[1:]:
        assert torch.equal(st, ref[2]), (name, graph)
        assert all(torch.equal(b, r) for b, r in zip(B, ref[3])), (name, graph)
        assert torch.equal(aux, ref[4]), (name, graph)

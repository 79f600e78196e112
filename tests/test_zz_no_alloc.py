"""grmhd_c2p_batch_noalloc (statistics records in the library's static device
array, no allocation) against grmhd_c2p_batch on the same Kerr-Schild states:
two calls in a row, the second with a smaller batch, bit-identical outputs and
identical statistics."""
import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _run(ctx, fn, st, n, dev):
    import torch
    cons = torch.from_numpy(np.ascontiguousarray(st["cons"][:, :n])).to(dev)
    met = torch.from_numpy(np.ascontiguousarray(st["met"][:, :n])).to(dev)
    prims = torch.full((13, n), -7.0, dtype=torch.float64, device=dev)
    err = torch.zeros(n, dtype=torch.int64, device=dev)
    fl = torch.full((n,), 9, dtype=torch.uint8, device=dev)
    roots = torch.zeros((4, n), dtype=torch.float64, device=dev)
    fn(n, cons, met, prims, err, fl, roots=roots)
    torch.cuda.synchronize()
    return ([x.cpu().numpy() for x in (prims, cons, err, fl, roots)], ctx.c2p_stats())


def test_gpu_c2p_batch_noalloc_matches_c2p_batch(dev):
    n0, n1 = 300_000, 70_001
    st = W.c2p_states(n0, seed=W.SEED + 7)
    params = st["params"]
    grid = ga.Grid(1, 1, 1, 3, 1)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))])
    ctx = ga.Context(params, grid, inst)
    for n in (n0, n1):
        say(f"[c2p noalloc] Kerr-Schild states n={n}")
        ref, sref = _run(ctx, ctx.c2p_batch, st, n, dev)
        got, sgot = _run(ctx, ctx.c2p_batch_noalloc, st, n, dev)
        say(f"  stats plain {sref}\n  stats noalloc {sgot}")
        for name, a, b in zip(("prims", "cons", "err", "floored", "roots"), ref, got):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
        assert sref == sgot
        assert sgot["n_cells"] == n and sgot["n_root_evals"] > 0


def test_gpu_c2p_batch_noalloc_rejects_oversize(dev):
    import torch
    grid = ga.Grid(1, 1, 1, 3, 1)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))])
    ctx = ga.Context(ga.default_params(), grid, inst)
    x = torch.zeros(16, dtype=torch.float64, device=dev)
    e = torch.zeros(16, dtype=torch.int64, device=dev)
    f = torch.zeros(16, dtype=torch.uint8, device=dev)
    with pytest.raises(RuntimeError, match="exceeds"):
        ctx.c2p_batch_noalloc((1 << 24) + 1, x, x, x, e, f)

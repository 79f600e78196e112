"""GPU tests of the learned Kastaun bracket (include/grmhd_mlp.h) against the
numpy restatement (oracle/mlp_oracle.py) and the CPU c2p oracle.

Tolerances: features within 4 FP32 ulp (+1e-6 absolute; device logf vs a correctly
rounded log);
forward outputs within 3e-2 (max) / 1e-3 (median) in normalised units --
the GPU accumulates in a different order, which can flip single bf16
roundings of hidden activations (a wrong MFMA
operand layout or a transposed weight gives O(1)); gradients within 2e-2
relative L2 per parameter block; learned c2p primitives within 1e-10 of the
FP64 oracle (state-normalised, test/test_c2p.cpp:87-103).
"""
import os
import sys

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W
import _oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_host(t):
    return t.detach().cpu().numpy()


def single_ctx(params):
    grid = ga.Grid(1, 1, 1, 3, 1)
    inst = ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))])
    return ga.Context(params, grid, inst)


def problem(n, seed):
    rng = np.random.default_rng(seed)
    st = W.c2p_states(n, seed=seed)
    feat, _, _, _ = M.features(st["cons"], st["met"], 1 + 1e-15)
    p = M.init_params(seed)
    p[M.B1:M.W2] = rng.normal(0, 0.3, M.H)
    p[M.B2:M.W3] = rng.normal(0, 0.3, M.H)
    p[M.B3:M.NP] = rng.normal(0, 0.3, M.NO)
    scale = np.zeros(M.NSCALE, dtype=np.float32)
    scale[:10] = feat.mean(1)
    scale[10:20] = 1.0 / feat.std(1)
    scale[20:24] = [0.2, -0.1, 1.5, 0.7]
    target = rng.normal(0, 1, (2, n)).astype(np.float32)
    target[1, ::5] = np.nan
    return st, feat, p, scale, target


def test_gpu_mlp_features(dev):
    import torch
    n = 5000
    say(f"[mlp features] n={n}")
    st = W.c2p_states(n, seed=31)
    ctx = single_ctx(st["params"])
    f = torch.zeros((10, n), dtype=torch.float32, device=dev)
    ga.mlp_features(ctx, n, to_dev(st["cons"], dev), to_dev(st["met"], dev), f)
    torch.cuda.synchronize()
    ref = M.features(st["cons"], st["met"], 1 + 1e-15)[0]
    got = to_host(f)
    # FP32 logf on the device vs a correctly rounded log: a few ulp
    ulp = np.spacing(np.abs(ref).astype(np.float32))
    d = np.abs(got.astype(np.float64) - ref)
    say(f"  max |diff|/ulp {np.max(d / ulp):.2f}")
    assert np.all(d <= 4 * ulp + 1e-6)


def test_gpu_mlp_forward_matches_oracle(dev):
    import torch
    n = 4099
    say(f"[mlp forward] n={n}")
    st, feat, p, scale, _ = problem(n, 5)
    mlp = ga.Mlp()
    mlp.set_params(p)
    mlp.set_scaling(scale)
    out = torch.zeros((2, n), dtype=torch.float32, device=dev)
    mlp.forward(n, to_dev(feat, dev), out)
    torch.cuda.synchronize()
    y, _ = M.forward(p, scale, feat, quant=True)
    ref = M.denormalise(y, scale)
    got = to_host(out).astype(np.float64)
    d = np.abs(got - ref) / scale[22:24, None]
    say(f"  normalised |diff| max {d.max():.3e} median {np.median(d):.3e}; output spread {y.std():.3f}")
    assert y.std() > 0.1
    assert np.median(d) < 1e-3 and d.max() < 3e-2


def test_gpu_mlp_grad_matches_oracle(dev):
    import torch
    n = 1000
    say(f"[mlp grad] n={n}")
    st, feat, p, scale, target = problem(n, 9)
    mlp = ga.Mlp()
    mlp.set_params(p)
    mlp.set_scaling(scale)
    g = torch.zeros(ga.MLP_NPARAM + 1, dtype=torch.float32, device=dev)
    mlp.grad(n, to_dev(feat, dev), to_dev(target, dev), g)
    torch.cuda.synchronize()
    gg = to_host(g).astype(np.float64)
    ref, loss = M.grad(p, scale, feat, target, quant=True)
    say(f"  loss gpu {gg[-1]:.6e} oracle {loss:.6e}")
    assert abs(gg[-1] - loss) < 1e-3 * loss
    for name, a, b in (("W1", M.W1, M.B1), ("b1", M.B1, M.W2), ("W2", M.W2, M.B2), ("b2", M.B2, M.W3),
                       ("W3", M.W3, M.B3), ("b3", M.B3, M.NP)):
        e = np.linalg.norm(gg[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        say(f"  {name}: rel L2 {e:.3e}")
        assert e < 2e-2, name
    assert np.all(gg[M.W1 + 10 * 64:M.B1] == 0)


def test_gpu_mlp_adam_step(dev):
    """One train_step equals Adam (0.9, 0.999, 1e-8, bias-corrected) applied
    to the gradient the library reports for the same batch."""
    import torch
    n = 512
    st, feat, p, scale, target = problem(n, 13)
    mlp = ga.Mlp()
    mlp.set_params(p)
    mlp.set_scaling(scale)
    fd, td = to_dev(feat, dev), to_dev(target, dev)
    g = torch.zeros(ga.MLP_NPARAM + 1, dtype=torch.float32, device=dev)
    mlp.grad(n, fd, td, g)
    mlp.train_step(n, fd, td, 1e-3)
    torch.cuda.synchronize()
    gh = to_host(g)[:-1].astype(np.float64)
    p1 = mlp.get_params().astype(np.float64)
    m = 0.1 * gh
    v = 0.001 * gh * gh
    exp = p.astype(np.float64) - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    exp[M.W1 + 10 * 64:M.B1] = 0.0
    assert np.max(np.abs(p1 - exp)) < 1e-6


def test_gpu_learned_c2p_trained(dev):
    """Train on states the plain c2p solves (seed A), then run the fused
    learned c2p on other states (seed B), with the Newton finish (default)
    and with brackets only: same error bits, primitives within 1e-10 of the
    FP64 oracle, fewer master-function evaluations than plain."""
    import torch
    ntr, n = 1 << 17, 20000
    say(f"[learned c2p] fit on {ntr} states")
    tr = W.c2p_states(ntr, seed=W.SEED + 7)
    ctx = single_ctx(tr["params"])
    mlp = ga.Mlp()
    rep = mlp.fit(ctx, ntr, to_dev(tr["cons"], dev), to_dev(tr["met"], dev), to_dev(tr["rad"], dev),
                  opts=ga.Mlp.default_fit_opts(epochs=60))
    say(f"  fit report {rep}")
    assert rep["loss_last"] < 0.1 * rep["loss_first"]
    assert rep["val_p50_e"] < 0.05
    te = W.c2p_states(n, seed=W.SEED + 8)
    res = {}
    assert mlp.get_newton() == 8
    for mode in ("plain", "learned", "bracket"):
        mlp.set_newton(0 if mode == "bracket" else 8)
        cons = to_dev(te["cons"], dev)
        met = to_dev(te["met"], dev)
        prims = torch.zeros((13, n), dtype=torch.float64, device=dev)
        err = torch.zeros(n, dtype=torch.int64, device=dev)
        fl = torch.zeros(n, dtype=torch.uint8, device=dev)
        if mode == "plain":
            ctx.c2p_batch(n, cons, met, prims, err, fl, r=to_dev(te["rad"], dev))
        else:
            ga.c2p_batch_learned(ctx, mlp, n, cons, met, prims, err, fl, r=to_dev(te["rad"], dev))
        torch.cuda.synchronize()
        res[mode] = (to_host(prims), to_host(err).astype(np.uint64), ctx.c2p_stats())
        say(f"  {mode}: {res[mode][2]}")
    op, oc, oe, of, ev, fatal = O.c2p_batch(te["params"], te["cons"], te["met"], r=te["rad"])
    from test_oracle_c2p import prim_residual
    pp, pe, ps = res["plain"]
    plain_evals = ps["n_root_evals"] + ps["n_bracket_evals"]
    for mode in ("learned", "bracket"):
        lp, le, ls = res[mode]
        same = le == oe
        assert same.mean() > 0.999
        r = prim_residual(lp, op)[same]
        say(f"  {mode} vs oracle: max state-normalised diff {r.max():.3e}")
        assert r.max() < 1e-10
        assert ls["n_fatal"] == fatal == 0
        assert ls["n_guess_used"] > 0.5 * n
        learned_evals = ls["n_root_evals"] + ls["n_bracket_evals"]
        say(f"  evals/state plain {plain_evals / n:.2f} {mode} {learned_evals / n:.2f}")
        assert learned_evals < plain_evals
    # the Newton finish needs fewer evaluations than the brackets
    assert res["learned"][2]["n_root_evals"] < 0.7 * res["bracket"][2]["n_root_evals"]


def test_gpu_learned_c2p_in_rk3_torus(dev):
    """The cell passes of the RK3 step with the learned guess (network
    trained on the torus's own cells through grmhd_mlp_fit_cells): one step
    of the Fishbone-Moncrief torus agrees with the plain-c2p step to the
    learned solve's root tolerance (1e-13 relative on mu, amplified through
    one step), with the guess used in most cells."""
    import torch
    from test_gpu_parity import linf_rel, interior, face_interior
    wl = W.fm_torus(nb=(2, 2, 2), n=16, extent=24.0)
    say("[learned rk3] FM torus 32^3")
    dt = torch.tensor([0.25 * wl.extra["dx"]], dtype=torch.float64, device=dev)
    out = {}
    mlp = ga.Mlp()
    for mode in ("plain", "learned"):
        ctx = ga.Context(wl.params, wl.grid, wl.instances)
        st = to_dev(wl.state, dev)
        B = [to_dev(b, dev) for b in wl.B]
        aux = to_dev(wl.aux, dev)
        if mode == "learned":
            rep = mlp.fit_cells(ctx, st, B, max_states=1 << 15, opts=ga.Mlp.default_fit_opts(epochs=40))
            say(f"  fit report {rep}")
            ctx.set_learned_c2p(mlp)
        ctx.rk3_step(dt, st, B, aux)
        torch.cuda.synchronize()
        res = (to_host(st), [to_host(b) for b in B], to_host(aux))
        ctx.compute_auxiliaries(st, B, aux)      # a counted c2p pass over every cell
        torch.cuda.synchronize()
        out[mode] = res + (ctx.c2p_stats(),)
        say(f"  {mode}: {out[mode][3]}")
    (sp, Bp, ap, _), (sl, Bl, al, stl) = out["plain"], out["learned"]
    for v in range(7):
        e = linf_rel(interior(wl.grid, sl[:, v]), interior(wl.grid, sp[:, v]))
        assert e < 1e-9, (v, e)
    for d in range(3):
        assert linf_rel(face_interior(wl.grid, Bl[d], d), face_interior(wl.grid, Bp[d], d)) < 1e-9, d
    assert np.array_equal(al[:, ga.C2P_ERR], ap[:, ga.C2P_ERR])
    assert stl["n_guess_used"] > 0.5 * stl["n_cells"]


@pytest.mark.parametrize("wl_name", ["torus", "blast"])
def test_gpu_warm_start_c2p_in_rk3(dev, wl_name):
    """The cell passes started from each cell's previous primitives
    (grmhd_set_c2p_warm_start): two RK3 steps (graph replay, so stages 2-3
    and the second step start from the library's own c2p results) agree
    with the plain-c2p steps to the Newton finish's root tolerance, with
    identical error bits, the guess used in most cells that solve, and fewer
    master-function evaluations than plain."""
    import torch
    from test_gpu_parity import linf_rel, interior, face_interior
    wl = W.fm_torus(nb=(2, 2, 2), n=16, extent=24.0) if wl_name == "torus" else W.blast(nb=(2, 2, 2), n=16)
    say(f"[warm rk3] {wl_name}")
    dt = torch.tensor([0.25 * wl.extra["dx"]], dtype=torch.float64, device=dev)
    out = {}
    for mode in ("plain", "warm"):
        ctx = ga.Context(wl.params, wl.grid, wl.instances)
        if mode == "warm":
            ctx.set_c2p_warm_start(8)
        st = to_dev(wl.state, dev)
        B = [to_dev(b, dev) for b in wl.B]
        aux = to_dev(wl.aux, dev)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            for _ in range(2):
                ctx.rk3_step(dt, st, B, aux, stream=s.cuda_stream, graph=True)
        s.synchronize()
        res = (to_host(st), [to_host(b) for b in B], to_host(aux))
        ctx.compute_auxiliaries(st, B, aux)      # a counted c2p pass over every cell
        torch.cuda.synchronize()
        out[mode] = res + (ctx.c2p_stats(),)
        say(f"  {mode}: {out[mode][3]}")
    (sp, Bp, ap, stp), (sw, Bw, aw, stw) = out["plain"], out["warm"]
    for v in range(7):
        e = linf_rel(interior(wl.grid, sw[:, v]), interior(wl.grid, sp[:, v]))
        assert e < 1e-9, (v, e)
    for d in range(3):
        assert linf_rel(face_interior(wl.grid, Bw[d], d), face_interior(wl.grid, Bp[d], d)) < 1e-9, d
    assert np.array_equal(aw[:, ga.C2P_ERR], ap[:, ga.C2P_ERR])
    solving = stw["n_cells"] - stw["n_atmo_reset"]
    assert stw["n_guess_used"] > 0.5 * solving
    assert stw["n_root_evals"] + stw["n_bracket_evals"] < stp["n_root_evals"] + stp["n_bracket_evals"]


def test_gpu_warm_start_rejected_for_hybrid_eos(dev):
    from test_oracle_c2p import hybrid_params
    ctx = ga.Context(hybrid_params(), ga.Grid(1, 1, 1, 3, 1),
                     ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))]))
    with pytest.raises(RuntimeError, match="status 4"):
        ctx.set_c2p_warm_start(8)
    ctx.set_c2p_warm_start(0)


@pytest.mark.parametrize("case", ["sigma", "cold", "dense", "near_hole"])
def test_gpu_learned_c2p_out_of_distribution(dev, case):
    """A network trained on the C4 distribution, run on states outside it
    (magnetisation to 1e4, eps down to 1e-10, rho to 1e2, r down to 1.1 M):
    the Newton finish either converges to the root or hands the state to
    the plain solve, so the error bits match and the primitives stay within
    1e-10 of the plain FP64 solve."""
    import torch
    from bench import state_residual
    kw = {"sigma": dict(sigma=(1e-2, 1e4)), "cold": dict(eps=(1e-10, 1e-4)), "dense": dict(rho=(1e-4, 1e2)),
          "near_hole": dict(r=(1.1, 2.0))}[case]
    ntr, n = 1 << 17, 1 << 15
    tr = W.c2p_states(ntr, seed=W.SEED + 7)
    ctx = single_ctx(tr["params"])
    mlp = ga.Mlp()
    mlp.fit(ctx, ntr, to_dev(tr["cons"], dev), to_dev(tr["met"], dev), to_dev(tr["rad"], dev),
            opts=ga.Mlp.default_fit_opts(epochs=40))
    te = W.c2p_states(n, seed=W.SEED + 77, **kw)
    say(f"[learned c2p, out of distribution] {case} {kw}")
    res = {}
    for mode in ("plain", "learned"):
        cons = to_dev(te["cons"], dev)
        prims = torch.zeros((13, n), dtype=torch.float64, device=dev)
        err = torch.zeros(n, dtype=torch.int64, device=dev)
        fl = torch.zeros(n, dtype=torch.uint8, device=dev)
        if mode == "plain":
            ctx.c2p_batch(n, cons, to_dev(te["met"], dev), prims, err, fl, r=to_dev(te["rad"], dev))
        else:
            ga.c2p_batch_learned(ctx, mlp, n, cons, to_dev(te["met"], dev), prims, err, fl, r=to_dev(te["rad"], dev))
        torch.cuda.synchronize()
        res[mode] = (prims, err, ctx.c2p_stats())
    same = res["plain"][1] == res["learned"][1]
    assert bool(same.all())
    r = state_residual(res["learned"][0], res["plain"][0])
    say(f"  max state-normalised difference {r.max().item():.3e}, stats {res['learned'][2]}")
    assert bool(torch.isfinite(res["learned"][0]).all())
    assert r.max().item() < 1e-10


def test_gpu_learned_rk3_graph_follows_settings(dev):
    """A captured RK3 graph with the learned c2p replays bit-identically to
    the eager step, and is captured again when the network's settings change
    (here the Newton evaluation limit: 0 switches to the brackets)."""
    import torch
    wl = W.fm_torus(nb=(2, 2, 2), n=16, extent=24.0)
    say("[learned rk3 graph] FM torus 32^3")
    dt = torch.tensor([0.25 * wl.extra["dx"]], dtype=torch.float64, device=dev)
    mlp = ga.Mlp()
    out = {}
    for mode in ("eager", "graph"):
        ctx = ga.Context(wl.params, wl.grid, wl.instances)
        st = to_dev(wl.state, dev)
        B = [to_dev(b, dev) for b in wl.B]
        aux = to_dev(wl.aux, dev)
        if mode == "eager":
            mlp.fit_cells(ctx, st, B, max_states=1 << 15, opts=ga.Mlp.default_fit_opts(epochs=20))
        ctx.set_learned_c2p(mlp)
        s = torch.cuda.Stream()
        res = []
        with torch.cuda.stream(s):
            for newton in (8, 0):
                mlp.set_newton(newton)
                ctx.rk3_step(dt, st, B, aux, stream=s.cuda_stream, graph=(mode == "graph"))
                s.synchronize()
                res.append((to_host(st), to_host(aux)))
        mlp.set_newton(8)
        out[mode] = res
    for (se, ae), (sg, ag) in zip(out["eager"], out["graph"]):
        assert np.array_equal(se, sg)
        assert np.array_equal(ae, ag)

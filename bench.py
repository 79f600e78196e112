#!/usr/bin/env python3
"""GRMHD zone-updates/s on one MI355X.

Workload (default): 3-D magnetised blast wave, 128^3 cells in 512 blocks of
16^3 with 3 ghost layers, ideal gas Gamma=4/3, WENO-Z + HLLE + UCT, FP64.
A "step" is one full SSP-RK3 step of the whole block list: device CFL dt
(cfl 0.25) + 3 x (fluxes, UCT EMFs, fused update + CT, ghost gather, c2p over
all cells), replayed from a hipGraph.  value = interior zones x steps / time.

Prints one JSON line.  Extra objects:
  roofline      dominant kernel (largest measured time per step; the three
                direction instantiations of the face flux kernel k_flux count
                as one kernel, averaged per launch, HIP events on its launch
                stream).  Its bound follows from its algorithmic intensity
                (operation count of the oracle per face over the algorithmic
                bytes per face) against the ridge point 78.6 TFLOP/s / 8 TB/s;
                both the FP64 and the HBM fractions are given, the executed
                FP64 rate and the counted HBM traffic from the committed
                rocprofv3 counter summaries under profiles/, and the step's
                algorithmic HBM and FP64 fractions
  cpu_baseline  the CPU oracle (contracted FMAs, built for the host CPU) timing
                RK3 steps of the same 128^3 workload on the host cores this
                process may use (child process, OMP_PROC_BIND=close,
                OMP_PLACES=cores): median of 5 steps, plus one step of the
                64^3 blast on one thread
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic bytes per interior zone for one launch of each kernel family
# (flat-metric variant; compulsory unique traffic of the algorithm's data that
# the kernel must move, counted once per zone).
#  flux   : per direction, reads the 9 reconstructed primitives (the
#           cell-centred normal B is not needed: the face value replaces it)
#           + the face-normal B (10 doubles), writes 7 fluxes + 4 UCT face
#           values (11 doubles)                                        -> 168 B
#  emf    : reads 2 face-B + 4 face values x 2 directions per edge family
#           (3 edge comps x (2 B + 2 vbar + 4 speeds)), writes 3 EMFs   -> 216 B
#  update : reads 7 cons x 2 levels + 21 fluxes + 3 EMF + 3 B x 2 levels,
#           writes 7 cons + 3 B                                        -> 456 B
#  c2p    : reads 7 cons + 3 face B + C2P_ERR, writes 12 prims + BDIV,
#           C2P_DENS_ERR, C2P_ERR and rewrites <= 7 cons                -> 264 B
ALG_BYTES = {"flux": 168, "emf": 216, "update": 456, "c2p": 264, "ghost": 0, "ghost_faces": 0}
# timer families of one kernel function (k_flux<EOS, REC, FLAT, D>, one launch per direction)
KERNEL_FAMILIES = {"flux": ("flux_x", "flux_y", "flux_z")}
# Whole-step algorithmic bytes per zone-update (flat metric; RHS 3x(38..48
# reads + 10 writes) + c2p 3x) -- 1552 B.
STEP_ALG_BYTES = 1552
# The same with the metric arrays read (curved metric: + 16 metric doubles per
# stage in the RHS and 10 in the c2p, x 3 stages) -- 2176 B.
STEP_ALG_BYTES_CURVED = 2176
RIDGE_FLOP_PER_BYTE = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
# Algorithmic FP64 operations, counted by running the CPU oracle with an
# operation-counting scalar on this 128^3 workload (oracle/count_flops.py;
# add/sub and mul, the oracle has no FMA; divisions and square roots counted
# apart).  Per face of one direction (WENO-Z of 10 variables on both sides,
# EOS, HLLE fluxes, UCT face values; flat metric: no face-metric set-up):
FALG_FACE_FLOPS = 2041
FALG_FACE_DIVSQRT = 122
# per edge (UCT: 4 WENO-Z reconstructions + 2 HLL fluxes):
FALG_EDGE_FLOPS = 565
# per zone-update of the RK3 step (3 x (3 faces, 3 edges, flux divergence,
# 3 CT faces, c2p) + 2 linear combinations), flat metric (no geometric
# sources and no metric set-up, which a flat-metric kernel skips) and as
# the reference evaluates it with the metric terms:
FALG_STEP_FLOPS_FLAT = 26034
FALG_STEP_FLOPS_REFERENCE = 30171
# the same per face direction launch of the flux kernel (one face per zone)
FALG_PER_ZONE = {"flux": FALG_FACE_FLOPS}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_profile_counters(kernel_tag):
    """FP64 VALU instruction counts and HBM traffic per launch from committed
    rocprofv3 --pmc summaries (profiles/<kernel_tag>_pmc.md), if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_pmc.md")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("|") and "=" in line:
                pass
            if line.startswith("- ") and ":" in line:
                k, v = line[2:].split(":", 1)
                try:
                    out[k.strip()] = float(v.split()[0])
                except ValueError:
                    pass
    return out


def _oracle_native_lib():
    """The oracle (contracted FMAs) compiled for this host's CPU
    (-O3 -march=native -fopenmp) into a temporary directory; the in-tree
    portable build (-march=x86-64-v3) if no compiler is available."""
    import shutil
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "oracle", "oracle_capi.cpp")
    cxx = shutil.which("g++")
    if cxx:
        out = os.path.join(tempfile.mkdtemp(prefix="grmhd_oracle_"), "liboracle_native.so")
        r = subprocess.run([cxx, "-std=c++17", "-O3", "-march=native", "-fPIC", "-shared", "-fopenmp",
                            "-ffp-contract=fast", "-o", out, src], capture_output=True, text=True)
        if r.returncode == 0:
            return out, "-O3 -march=native -fopenmp -ffp-contract=fast"
        log(f"[bench] native oracle build failed: {r.stderr[-500:]}")
    return os.path.join(ROOT, "oracle", "_build", "liboracle_fast.so"), "-O3 -march=x86-64-v3 -fopenmp (prebuilt)"


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_oracle_steps(L, wl, reps):
    """Seconds of each of `reps` RK3 steps (device-CFL dt) after one warm-up."""
    P = C.c_void_p
    st, B, aux = wl.copy_arrays()
    h = L.orc_stepper_create(C.byref(wl.params), C.byref(wl.grid), wl.instances, len(wl.inst_specs))
    p = lambda a: C.c_void_p(a.ctypes.data)
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        dt = L.orc_find_dt(h, p(st), p(aux), 0.25)
        L.orc_rk3_step(h, dt, p(st), p(B[0]), p(B[1]), p(B[2]), p(aux))
        if r:
            times.append(time.perf_counter() - t0)
    L.orc_stepper_destroy(h)
    return times


def cpu_baseline_child(reps=5):
    """Body of the CPU-baseline child process (no GPU): the oracle on the
    host cores, OMP_PROC_BIND=close OMP_PLACES=cores set by the parent.
    Median of `reps` RK3 steps of the full 128^3 workload after one warm-up
    on all threads, and one step of the 64^3 blast (64 blocks) on 1 thread."""
    from grace_amd import workloads as W
    path, flags = _oracle_native_lib()
    L = C.CDLL(path)
    P, d = C.c_void_p, C.c_double
    L.orc_num_threads.argtypes = [C.c_int]
    L.orc_num_threads.restype = C.c_int
    L.orc_stepper_create.argtypes = [P, P, P, C.c_int]
    L.orc_stepper_create.restype = P
    L.orc_stepper_destroy.argtypes = [P]
    L.orc_rk3_step.argtypes = [P, d, P, P, P, P, P]
    L.orc_find_dt.argtypes = [P, P, P, d]
    L.orc_find_dt.restype = d
    nthr = L.orc_num_threads(0)
    wl = W.blast(nb=(8, 8, 8), n=16)
    times = _time_oracle_steps(L, wl, reps)
    t = float(np.median(times))
    L.orc_num_threads(1)
    small = W.blast(nb=(4, 4, 4), n=16)
    t1 = _time_oracle_steps(L, small, 1)[0]
    return {"value": wl.zones / t, "unit": "zone-updates/s", "cores": nthr, "kind": "port",
            "cpu_model": _cpu_model(), "build": flags,
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "step_seconds": [round(x, 4) for x in times],
            "one_thread": {"value": small.zones / t1, "unit": "zone-updates/s",
                           "sample": f"one RK3 step of the 64^3 blast (64 blocks of 16^3), {t1:.2f}s"},
            "sample": f"median of {reps} RK3 steps (after 1 warm-up) of the full 128^3 blast "
                      f"({wl.zones} zones, device-CFL dt), C++ oracle, {nthr} OpenMP threads, "
                      f"median step {t:.3f}s"}


def cpu_baseline(threads=None):
    """Runs cpu_baseline_child in a child process (the OpenMP binding
    variables must be set before its runtime starts) on the host cores this
    process may use; returns its JSON or None."""
    import subprocess
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        ncpu = min(ncpu, int(os.environ["OMP_NUM_THREADS"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads or ncpu), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child"], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"[bench] cpu baseline child failed ({r.returncode}): {r.stderr[-800:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def state_residual(a, b):
    """max |a - b| over RHO, EPS, Z^i per state, over max(|rho|, |eps|, |z|)
    of the reference state (the norm of test/test_c2p.cpp:87-103)."""
    import torch
    z = torch.sqrt(b[2] ** 2 + b[3] ** 2 + b[4] ** 2)
    scale = torch.maximum(torch.maximum(b[0].abs(), b[7].abs()), torch.clamp(z, min=1e-30))
    e = torch.zeros_like(scale)
    for i in (0, 7, 2, 3, 4):
        e = torch.maximum(e, (a[i] - b[i]).abs())
    return e / scale


def c4_leg(ga, W, torch, dev, n=1 << 22, n_train=1 << 20, reps=5):
    """Kastaun c2p alone on the 4M Kerr-Schild states: plain FP64 solve vs the
    learned guess (MLP trained on the GPU on 1M states of another seed) with
    its Newton finish, and vs the learned brackets alone (width and moves
    picked by timing on a 1M-state slice)."""
    log(f"[bench] c2p leg: {n} states, training on {n_train}")
    st = W.c2p_states(n)
    tr = W.c2p_states(n_train, seed=W.SEED + 100)
    ctx = ga.Context(st["params"], ga.Grid(1, 1, 1, 3, 1),
                     ga.make_instances([dict(nb=(1, 1, 1), q0=0, bc=(0,) * 6, idx=(1.0, 1.0, 1.0))]))
    mlp = ga.Mlp()
    t0 = time.perf_counter()
    # 120 epochs (2.4 s): median log error 1.7e-3 instead of 2.7e-3 at the
    # default 30, 3.8 instead of 4.4 Newton evaluations per state
    rep = mlp.fit(ctx, n_train, torch.from_numpy(tr["cons"]).to(dev), torch.from_numpy(tr["met"]).to(dev),
                  torch.from_numpy(tr["rad"]).to(dev), opts=ga.Mlp.default_fit_opts(epochs=120))
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    del tr
    full = (torch.from_numpy(st["cons"]).to(dev), torch.from_numpy(st["met"]).to(dev),
            torch.from_numpy(st["rad"]).to(dev))
    m = n // 4
    part = tuple(x[..., :m].contiguous() for x in full)
    cons = torch.empty_like(full[0])
    prims = torch.zeros((13, n), dtype=torch.float64, device=dev)
    err = torch.zeros(n, dtype=torch.int64, device=dev)
    fl = torch.zeros(n, dtype=torch.uint8, device=dev)

    def once(learned, data, k, cons_k):
        c0, met, rad = data
        cons_k.copy_(c0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        if learned:
            ga.c2p_batch_learned(ctx, mlp, k, cons_k, met, prims, err, fl, r=rad)
        else:
            ctx.c2p_batch(k, cons_k, met, prims, err, fl, r=rad)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    def run(learned, data, k):
        cons_k = cons[..., :k] if k == n else torch.empty_like(data[0])
        once(learned, data, k, cons_k)
        ts = [once(learned, data, k, cons_k) for _ in range(reps)]
        return float(np.median(ts)), ctx.c2p_stats()

    def run_pair(newton_evals):
        """plain and learned alternated (clock and cache state shared)"""
        mlp.set_newton(newton_evals)
        cons_k = cons
        once(False, full, n, cons_k)
        once(True, full, n, cons_k)
        tp, tl = [], []
        for _ in range(2 * reps + 1):
            tp.append(once(False, full, n, cons_k))
            tl.append(once(True, full, n, cons_k))
        ratio = float(np.median(np.array(tp) / np.array(tl)))
        return float(np.median(tp)), float(np.median(tl)), ratio

    # bracket-only variant: width and moves picked by timing on a 1M-state slice
    de0, ds0, _ = mlp.get_bracket()
    newton = mlp.get_newton()
    mlp.set_newton(0)
    best = None
    for fac in (0.5, 1.0, 2.0, 4.0):
        for nexp in (2, 3, 4):
            mlp.set_bracket(de0 * fac, ds0, nexp)
            t, _ = run(True, part, m)
            if best is None or t < best[0]:
                best = (t, de0 * fac, nexp)
    mlp.set_bracket(best[1], ds0, best[2])
    t_pl, s_pl = run(False, full, n)
    p_pl, e_pl = prims.clone(), err.clone()

    def learned(newton_evals):
        t_p, t, ratio = run_pair(newton_evals)
        st_ = ctx.c2p_stats()
        same = e_pl == err
        res = state_residual(prims, p_pl)[same]
        solves = st_["n_cells"] + st_["n_ent_backup"]
        return t, {"ms": round(t, 4), "states_per_s": n / (t * 1e-3),
                   "root_evals_per_state": round(st_["n_root_evals"] / n, 3),
                   "bracket_evals_per_state": round(st_["n_bracket_evals"] / n, 3),
                   "max_root_evals": st_["max_root_evals"],
                   "fallback_rate": round(1.0 - st_["n_guess_used"] / max(solves, 1), 5),
                   "plain_ms_alternated": round(t_p, 4),
                   "speedup": round(ratio, 4),   # median over alternated plain/learned pairs
                   "error_vs_plain": {"max_state_rel": float(res.max()), "median_state_rel": float(res.median()),
                                      "same_error_bits": float(same.double().mean())}}

    counts = torch.zeros(n, dtype=torch.uint8, device=dev)

    def eval_dist(use_learned):
        """master-function + bracket evaluations per state (one untimed pass,
        grmhd_set_c2p_eval_counts): mean and percentiles"""
        ctx.set_c2p_eval_counts(counts)
        once(use_learned, full, n, cons)
        ctx.set_c2p_eval_counts(None)
        c = counts.to(torch.int32)
        pct = lambda f: int(torch.kthvalue(c, max(1, int(np.ceil(f * n)))).values.item())
        return {"mean": round(float(c.double().mean().item()), 3), "p50": pct(0.5), "p99": pct(0.99),
                "p999": pct(0.999), "max": int(c.max().item())}

    evals_plain = eval_dist(False)
    _, brk = learned(0)
    t_ln, lrn = learned(newton)
    lrn["evals_per_state"] = eval_dist(True)
    return {
        "workload": "c2p_4M_kerr_schild_a0.9_r1.5-50_ideal_gamma4/3",
        "states": n,
        "plain": {"ms": round(t_pl, 4), "states_per_s": n / (t_pl * 1e-3),
                  "root_evals_per_state": round(s_pl["n_root_evals"] / n, 3),
                  "bracket_evals_per_state": round(s_pl["n_bracket_evals"] / n, 3),
                  "max_root_evals": s_pl["max_root_evals"], "evals_per_state": evals_plain},
        "learned": dict(lrn, finish=f"Newton from the predicted roots, <= {newton} evaluations, then plain"),
        "learned_bracket_only": dict(brk, bracket={"delta_e": best[1], "delta_s": ds0, "nexp": best[2]}),
        "speedup": lrn["speedup"],
        "error_vs_plain": lrn["error_vs_plain"],
        "train": {"states": n_train, "seconds": round(fit_s, 3), "steps": rep["steps"],
                  "val_p50_log_err": rep["val_p50_e"], "val_p99_log_err": rep["val_p99_e"],
                  "loss_first": rep["loss_first"], "loss_last": rep["loss_last"]},
    }


def c3_leg(ga, W, torch, dev, steps=3):
    """64 uneven blast-wave instances (sides from {32,48,64,80,96}, 16^3 blocks)
    in one block list: every kernel of the RK3 step covers all instances in one
    launch; the step is replayed from a hipGraph with autotuned workgroups."""
    log("[bench] batched uneven instances: building")
    wl = W.multi_blast()
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    st = torch.from_numpy(wl.state).to(dev)
    B = [torch.from_numpy(b).to(dev) for b in wl.B]
    aux = torch.from_numpy(wl.aux).to(dev)
    zones, nq, sides = wl.zones, int(wl.grid.nq), [int(x) for x in wl.extra["sides"]]
    del wl
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream

    def timed(k):
        with torch.cuda.stream(stream):
            ctx.compute_dt(st, aux, 0.25, dt, stream=sp)
            ctx.rk3_step(dt, st, B, aux, stream=sp, graph=True)
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(k):
                ctx.compute_dt(st, aux, 0.25, dt, stream=sp)
                ctx.rk3_step(dt, st, B, aux, stream=sp, graph=True)
            stream.synchronize()
            return (time.perf_counter() - t0) / k

    t_default = timed(steps)
    with torch.cuda.stream(stream):
        tune = ctx.autotune(st, B, aux, reps=2, stream=sp)
        stream.synchronize()
    t_tuned = timed(steps)
    return {"workload": "64 blast instances, sides {32,48,64,80,96} (seeded), 16^3 blocks, ngz 3, "
                        "WENO-Z + HLLE + UCT, RK3 graph, device CFL dt",
            "zones": zones, "blocks": nq, "instance_sides": sides, "steps": steps,
            "zone_updates_per_s": zones / t_tuned, "ms_per_step": 1e3 * t_tuned,
            "ms_per_step_default_shapes": 1e3 * t_default, "tuning": tune}


def curved_leg(ga, W, torch, dev, steps=10, warmup=3, tune=True):
    """The 128^3 blast wave in the Kerr-Schild metric of a spin-0.9 hole
    (a cube of side 4 M centred at x = 8 M, 512 blocks of 16^3): the same
    step as the headline workload with the curved-metric kernels (face metric,
    geometric sources from the static-metric cache, metric read by the c2p),
    graph-replayed with the device CFL dt."""
    log("[bench] curved-metric 128^3 blast")
    wl = W.blast_ks(nb=(8, 8, 8), n=16)
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    st = torch.from_numpy(wl.state).to(dev)
    B = [torch.from_numpy(b).to(dev) for b in wl.B]
    aux = torch.from_numpy(wl.aux).to(dev)
    ctx.set_static_metric(st)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    tuning = None
    with torch.cuda.stream(stream):
        if tune:
            tuning = ctx.autotune(st, B, aux, reps=2, stream=sp)
        for _ in range(warmup):
            ctx.compute_dt(st, aux, 0.25, dt, stream=sp)
            ctx.rk3_step(dt, st, B, aux, stream=sp, graph=True)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.compute_dt(st, aux, 0.25, dt, stream=sp)
            ctx.rk3_step(dt, st, B, aux, stream=sp, graph=True)
        stream.synchronize()
        wall = time.perf_counter() - t0
        ctx.set_timing(1)
        ctx.rk3_step(dt, st, B, aux, stream=sp, graph=False)
        stream.synchronize()
    fam = {}
    for f in ("flux_x", "flux_y", "flux_z", "ring_x", "ring_y", "ring_z", "emf", "update", "ghost", "ghost_faces",
              "c2p"):
        ms, n = ctx.kernel_time_ms(f)
        if n:
            fam[f] = round(ms, 4)
    ctx.set_timing(0)
    ms_step = 1e3 * wall / steps
    return {"workload": "blast wave 128^3 (512 blocks of 16^3, ngz 3) in the Kerr-Schild metric, spin 0.9, cube of "
                        "side 4 M at x = 8 M, WENO-Z + HLLE + UCT, RK3 graph, device CFL dt, static-metric cache",
            "zones": wl.zones, "steps": steps, "ms_per_step": round(ms_step, 4),
            "zone_updates_per_s": wl.zones / (ms_step * 1e-3), "flat_metric": 0,
            "alg_bytes_per_zone_update": STEP_ALG_BYTES_CURVED, "alg_flops_per_zone_update": FALG_STEP_FLOPS_REFERENCE,
            "hbm_frac": round(STEP_ALG_BYTES_CURVED * wl.zones / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "fp64_frac": round(FALG_STEP_FLOPS_REFERENCE * wl.zones / (ms_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 5),
            "kernel_ms_per_step": fam, "tuning": tuning,
            "finite": bool(torch.isfinite(st).all().item() and torch.isfinite(aux).all().item())}


def torus_leg(ga, W, torch, dev, steps=5, warmup=2, tune=True):
    """Fishbone-Moncrief torus (r_in 6 M, r_max 12 M, Gamma 4/3, beta 100
    loop) around a spin-0.9375 hole in the Kerr-Schild metric on a two-level
    forest of 16^3 blocks: 6^3 coarse blocks on [-48, 48]^3 M (96^3, 1 M) with
    the central 4 x 4 x 2 refined to 0.5 M (the spacing of a 192^3 grid);
    coarse-fine prolongation / restriction and refluxing every substep, device
    CFL dt, graph RK3 steps with the plain c2p and with the learned guess
    (network trained on the torus's own cells)."""
    wl = W.fm_torus_forest()
    log(f"[bench] torus leg: two-level forest, {wl.grid.nq} blocks of 16^3")
    mlp = ga.Mlp()
    out = {"workload": "Fishbone-Moncrief torus a=0.9375 r_in=6 r_max=12 beta=100, Kerr-Schild (Cowling), two-level "
                       "forest on [-48,48]^3 M: 6^3 coarse blocks of 16^3 (1 M), central 4x4x2 refined (0.5 M, "
                       "192^3-equivalent), ngz 4, WENO-Z + HLLE + UCT, minmod / Toth-Roe prolongation, reflux, "
                       "rho_atm 1e-10",
           "blocks": int(wl.grid.nq), "fine_blocks": int(wl.extra["n_fine"]),
           "static_metric_cache": True,
           "zones": wl.zones, "steps": steps}
    runs = {}
    for mode in ("plain", "learned", "warm"):
        ctx = ga.Context(wl.params, wl.grid, wl.instances, forest=wl.forest)
        st = torch.from_numpy(wl.state).to(dev)
        B = [torch.from_numpy(b).to(dev) for b in wl.B]
        aux = torch.from_numpy(wl.aux).to(dev)
        ctx.fill_metric_ghosts(st)  # the metric's ghost exchange (Lagrange across coarse-fine faces), once
        ctx.set_static_metric(st)   # Cowling: metric derivatives taken once, outside the timed steps
        if tune and mode == "plain":
            tstream = torch.cuda.Stream()
            with torch.cuda.stream(tstream):
                out["tuning"] = ctx.autotune(st, B, aux, reps=2, stream=tstream.cuda_stream)
            tstream.synchronize()
        elif tune:
            ctx.set_tuning(**out["tuning"])
        if mode == "learned":
            t0 = time.perf_counter()
            mlp.fit_cells(ctx, st, B, max_states=1 << 20)
            out["fit_seconds"] = round(time.perf_counter() - t0, 3)
            ctx.set_learned_c2p(mlp)
        if mode == "warm":
            ctx.set_c2p_warm_start(8)
        runs[mode] = (ctx, st, B, aux, torch.zeros(1, dtype=torch.float64, device=dev), torch.cuda.Stream())

    def one_step(mode):
        ctx, st, B, aux, dtd, stream = runs[mode]
        sp = stream.cuda_stream
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
            ctx.compute_dt(st, aux, 0.25, dtd, stream=sp)
            ctx.rk3_step(dtd, st, B, aux, stream=sp, graph=True)
            e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1)

    # the two runs alternate step by step (same clocks and cache state)
    for _ in range(warmup):
        for mode in runs:
            one_step(mode)
    times = {m: [] for m in runs}
    for _ in range(steps):
        for mode in runs:
            times[mode].append(one_step(mode))
    for mode, (ctx, st, B, aux, dtd, stream) in runs.items():
        with torch.cuda.stream(stream):
            sp = stream.cuda_stream
            ctx.set_timing(1)
            ctx.rk3_step(dtd, st, B, aux, stream=sp, graph=False)
            stream.synchronize()
            c2p_ms = ctx.kernel_time_ms("c2p")[0]
            ctx.set_timing(0)
        ctx.compute_auxiliaries(st, B, aux)
        torch.cuda.synchronize()
        stt = ctx.c2p_stats()
        ms = float(np.median(times[mode]))
        out[mode] = {"ms_per_step": round(ms, 4), "zone_updates_per_s": wl.zones / (ms * 1e-3),
                     "c2p_ms_per_step": round(c2p_ms, 4), "dt": float(dtd.item()),
                     "finite": bool(torch.isfinite(st).all().item() and torch.isfinite(aux).all().item()),
                     "root_evals_per_cell": round(stt["n_root_evals"] / max(stt["n_cells"], 1), 3),
                     "bracket_evals_per_cell": round(stt["n_bracket_evals"] / max(stt["n_cells"], 1), 3),
                     "atmosphere_fraction": round(stt["n_atmo_reset"] / max(stt["n_cells"], 1), 4)}
    out["timing"] = "median of per-step HIP-event times, plain, learned and warm-start steps alternated"
    out["warm"]["guess"] = "Newton finish from each cell's previous primitives (grmhd_set_c2p_warm_start)"
    out["warm_vs_plain_step"] = round(out["plain"]["ms_per_step"] / out["warm"]["ms_per_step"], 4)
    del runs
    out["learned_vs_plain_step"] = round(out["plain"]["ms_per_step"] / out["learned"]["ms_per_step"], 4)
    return out


def c1_leg(ga, W, torch, dev, t_final=0.4, cfl=0.25):
    """Balsara-1 to t = 0.4 (400 cells, GS CT, device CFL dt read back every
    step); the oracle replays the same dt sequence for the parity figure."""
    log("[bench] Balsara-1 full run")
    wl = W.balsara1(N=400)
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    st = torch.from_numpy(wl.state.copy()).to(dev)
    B = [torch.from_numpy(b.copy()).to(dev) for b in wl.B]
    aux = torch.from_numpy(wl.aux.copy()).to(dev)
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    dts, t = [], 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while t < t_final:
        ctx.compute_dt(st, aux, cfl, dt)
        d = float(dt.item())
        if t + d > t_final:
            d = t_final - t
            dt.fill_(d)
        ctx.rk3_step(dt, st, B, aux)
        dts.append(d)
        t += d
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    out = {"workload": "Balsara-1, 400x1x1 cells, ngz 3, Gamma 2, WENO-Z + HLLE + GS, RK3, CFL 0.25, t = 0.4",
           "steps": len(dts), "gpu_seconds": gpu_s, "zone_updates_per_s_gpu": 400 * len(dts) / gpu_s}
    path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if os.path.exists(path):
        L = C.CDLL(path)
        L.orc_stepper_create.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_stepper_create.restype = C.c_void_p
        L.orc_stepper_destroy.argtypes = [C.c_void_p]
        L.orc_rk3_step.argtypes = [C.c_void_p, C.c_double] + [C.c_void_p] * 5
        s_, B_, a_ = wl.copy_arrays()
        h = L.orc_stepper_create(C.byref(wl.params), C.byref(wl.grid), wl.instances, len(wl.inst_specs))
        p = lambda a: C.c_void_p(a.ctypes.data)
        t0 = time.perf_counter()
        for d in dts:
            L.orc_rk3_step(h, d, p(s_), p(B_[0]), p(B_[1]), p(B_[2]), p(a_))
        cpu_s = time.perf_counter() - t0
        L.orc_stepper_destroy(h)
        g = wl.grid.ngz
        gs = st.cpu().numpy()
        rel = []
        for v in range(7):
            a_int, b_int = gs[0, v, g, g, g:-g], s_[0, v, g, g, g:-g]
            den = np.max(np.abs(b_int))
            rel.append(float(np.max(np.abs(a_int - b_int)) / (den if den > 0 else 1.0)))
        out.update({"cpu_seconds_oracle_1thread": cpu_s, "linf_rel_conserved_vs_oracle": max(rel)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-c2p", action="store_true", help="skip the learned-vs-plain c2p leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the batched uneven-instance leg")
    ap.add_argument("--no-torus", action="store_true", help="skip the Kerr-Schild torus leg")
    ap.add_argument("--no-curved", action="store_true", help="skip the curved-metric 128^3 blast leg")
    ap.add_argument("--no-tune", action="store_true", help="default workgroup sizes instead of grmhd_autotune")
    ap.add_argument("--tuning", default=None,
                    help="fixed tuning of the 128^3 blast instead of grmhd_autotune, as the JSON object the bench "
                         "line reports under config.tuning (profiling runs that must not contain the autotuner's "
                         "launches)")
    ap.add_argument("--no-c1", action="store_true", help="skip the Balsara-1 full-run leg")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--one-stream", action="store_true",
                    help="every launch on one stream (grmhd_params.stage_streams = 0), for kernel traces "
                         "whose launch durations are not overlapped")
    ap.add_argument("--only-c3", action="store_true",
                    help="only run the batched uneven-instance leg and print its JSON (for kernel traces)")
    ap.add_argument("--only-torus", action="store_true",
                    help="only run the two-level-forest torus leg and print its JSON (for profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.cpu_child:
        print(json.dumps(cpu_baseline_child(reps=args.cpu_reps)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported: the kernels run one "
              "GPU's share of the blocks, the multi-GPU halo exchange is not implemented", file=sys.stderr)
        sys.exit(2)
    cpu = None
    if not (args.no_cpu or args.only_torus or args.only_c3 or args.copy_probe):
        # before this process touches the GPU: a child process on the host cores
        log("[bench] CPU baseline (oracle, child process)")
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

    import torch
    import grace_amd as ga
    from grace_amd import workloads as W

    dev = torch.device("cuda:0")
    if args.only_c3:
        print(json.dumps(c3_leg(ga, W, torch, dev), default=float), flush=True)
        return
    if args.only_torus:
        print(json.dumps(torus_leg(ga, W, torch, dev, steps=args.steps, warmup=args.warmup, tune=not args.no_tune),
                         default=float), flush=True)
        return
    if args.copy_probe:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8, "launches": 3}), flush=True)
        return
    log("[bench] building 128^3 blast wave (512 blocks of 16^3)")
    wl = W.blast(nb=(8, 8, 8), n=16)
    if args.one_stream:
        wl.params.stage_streams = 0
    ctx = ga.Context(wl.params, wl.grid, wl.instances)
    st = torch.from_numpy(wl.state).to(dev)
    B = [torch.from_numpy(b).to(dev) for b in wl.B]
    aux = torch.from_numpy(wl.aux).to(dev)
    dt = torch.zeros(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    use_graph = not args.no_graph

    def step():
        ctx.compute_dt(st, aux, 0.25, dt, stream=sp)
        ctx.rk3_step(dt, st, B, aux, stream=sp, graph=use_graph)

    tuning = None
    with torch.cuda.stream(stream):
        if args.tuning:
            ctx.set_tuning(**json.loads(args.tuning))
            tuning = ctx.get_tuning()
        elif not args.no_tune:
            # workgroup sizes per kernel family picked by timing (inputs are
            # only read), before the warm-up and the timed steps
            tuning = ctx.autotune(st, B, aux, reps=7, stream=sp)
            stream.synchronize()
        log(f"[bench] warmup {args.warmup}")
        for _ in range(args.warmup):
            step()
        stream.synchronize()
        torch.cuda.synchronize()
        log(f"[bench] timing {args.steps} steps")
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = 1000.0 * wall / args.steps
    value = wl.zones * args.steps / wall

    # per-kernel durations: one more eager step with HIP events around each launch
    ctx.set_timing(1)
    with torch.cuda.stream(stream):
        for _ in range(2):
            ctx.compute_dt(st, aux, 0.25, dt, stream=sp)
            ctx.rk3_step(dt, st, B, aux, stream=sp, graph=False)
        stream.synchronize()
    fam = {}
    for f in ("flux_x", "flux_y", "flux_z", "ring_x", "ring_y", "ring_z", "emf", "update", "ghost", "ghost_faces",
              "c2p"):
        ms, n = ctx.kernel_time_ms(f)
        if n:
            fam[f] = (ms / n, n)
    ctx.set_timing(0)
    kern = {}   # kernel -> (total ms, launches, timer families)
    for f, (ms, n) in fam.items():
        k = next((g for g, fs in KERNEL_FAMILIES.items() if f in fs), f)
        t, c, fs = kern.get(k, (0.0, 0, ()))
        kern[k] = (t + ms * n, c + n, fs + (f,))
    dom = max((k for k in kern if k in ALG_BYTES), key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    alg = ALG_BYTES[dom] * wl.zones
    achieved = alg / (avg_ms * 1e-3) / 1e9
    # counters per launch: the mean over the kernel's timer families
    profs = [read_profile_counters(f"{f}_blast128") for f in kern[dom][2]]
    prof = {key: sum(p[key] for p in profs) / len(profs) for key in profs[0] if all(key in p for p in profs)}
    traffic = prof.get("hbm_bytes_per_launch")
    # Which roof bounds the kernel: its algorithmic intensity against the
    # ridge point (peak FP64 / peak HBM bandwidth)
    hbm = {"achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5), "alg_bytes_per_zone": ALG_BYTES[dom]}
    roof = {"traffic": traffic, "kernel": dom, "timer_families": list(kern[dom][2]),
            "avg_launch_ms": round(avg_ms, 4)}
    if dom in FALG_PER_ZONE:
        intensity = FALG_PER_ZONE[dom] / ALG_BYTES[dom]
        tf = FALG_PER_ZONE[dom] * wl.zones / (avg_ms * 1e-3) / 1e12
        fp64 = {"achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / FP64_PEAK_TFLOPS, 4), "alg_flops_per_zone": FALG_PER_ZONE[dom],
                "source": "oracle operation count (oracle/count_flops.py), add + mul"}
        bound = "fp64" if intensity > RIDGE_FLOP_PER_BYTE else "hbm"
        roof["intensity_flop_per_byte"] = round(intensity, 2)
        roof["ridge_flop_per_byte"] = round(RIDGE_FLOP_PER_BYTE, 2)
    else:
        fp64, bound = None, "hbm"
    top = fp64 if bound == "fp64" else hbm
    roof = {"bound": bound, "achieved": top["achieved"], "peak": top["peak"], "unit": top["unit"],
            "frac": top["frac"], **roof, "hbm": hbm, "fp64_algorithmic": fp64}
    if "fp64_flops_per_launch" in prof:
        tfx = prof["fp64_flops_per_launch"] / (avg_ms * 1e-3) / 1e12
        roof["fp64_executed"] = {"achieved": round(tfx, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(tfx / FP64_PEAK_TFLOPS, 4),
                                 "source": "rocprofv3 SQ_INSTS_VALU_*_F64 (FMA = 2)"}
    roof["counters_from"] = [f"profiles/{f}_blast128_pmc.md" for f in kern[dom][2]]
    roof["kernel_ms_per_step"] = {k: round(v[0] * v[1] / 2, 4) for k, v in fam.items()}
    roof["step"] = {"hbm_frac": round(STEP_ALG_BYTES * value / 1e9 / HBM_PEAK_GBS, 5),
                    "alg_bytes_per_zone_update": STEP_ALG_BYTES,
                    "fp64_frac": round(FALG_STEP_FLOPS_FLAT * value / 1e12 / FP64_PEAK_TFLOPS, 5),
                    "alg_flops_per_zone_update": FALG_STEP_FLOPS_FLAT,
                    "alg_flops_per_zone_update_with_metric_terms": FALG_STEP_FLOPS_REFERENCE}

    c2p_leg = None
    if not args.no_c2p:
        c2p_leg = c4_leg(ga, W, torch, dev)
    c1 = None if args.no_c1 else c1_leg(ga, W, torch, dev)
    c3 = None if args.no_c3 else c3_leg(ga, W, torch, dev)
    torus = None if args.no_torus else torus_leg(ga, W, torch, dev, tune=not args.no_tune)
    curved = None if args.no_curved else curved_leg(ga, W, torch, dev, tune=not args.no_tune)

    out = {
        "metric": "GRMHD zone-updates/s on 1 MI355X, learned vs plain c2p at <=1e-10 c2p err",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded blast-wave initial data, seed 20261202)",
        "config": {"workload": "blast3d_128cube_512x16cube_blocks_ngz3_wenoz_hlle_uct_rk3_plain_c2p",
                   "zones": wl.zones, "blocks": int(wl.grid.nq), "c2p": "plain Kastaun (FP64)",
                   "graph": use_graph, "stage_streams": int(wl.params.stage_streams),
                   "flat_metric": int(wl.params.flat_metric), "alg_bytes_variant": "flat metric, 1552 B",
                   "gpu_event_ms_per_step": gpu_ms / args.steps, "tuning": tuning},
        "roofline": roof,
        "cpu_baseline": cpu,
        "c2p_learned_vs_plain": c2p_leg,
        "balsara1_full_run": c1,
        "batched_uneven": c3,
        "fm_torus": torus,
        "curved_blast128": curved,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Algorithmic floating-point operation counts of one SSP-RK3 step of the
3-D magnetised blast wave (WENO-Z + HLLE + UCT, ideal gas, flat metric),
counted by running the CPU oracle with an operation-counting scalar
(liboracle_count.so, flop_count.hh).

TEST INFRASTRUCTURE.  The oracle evaluates some sites the algorithm does not
need (transverse ghost faces of the flux pass, ghost cells of the c2p pass);
the counts are therefore taken per site of each pass and credited per
interior zone: 3 faces, 3 edges, 1 cell of sources, 3 CT faces and 1 c2p
cell per zone and substep, plus the two RK3 linear combinations of the 7
conserved and 3 face-B values.

    python oracle/count_flops.py [--blocks 8] [--n 16]

prints one JSON object; the committed per-zone constants (bench.py) come
from the default 128^3 run.
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

KINDS = ("add", "mul", "div", "sqrt", "transcendental")
PASSES = {1: "face_flux", 2: "edge_emf", 3: "cell_flux_divergence", 4: "ct_face", 5: "c2p_cell"}
# sub-passes (flop_count.hh): geometric sources within pass 3, and metric
# set-up / metric derivatives (id 10 + enclosing pass), credited to the sites
# of the pass they run in
SUBS = {8: ("geometric_sources", 3), 11: ("face_flux_metric", 1), 13: ("cell_metric", 3),
        18: ("geometric_sources_metric_derivatives", 3), 15: ("c2p_cell_metric", 5)}
CREDIT = {1: 3, 2: 3, 3: 1, 4: 3, 5: 1}   # sites per interior zone and substep
NPASS = 20


def count_lib():
    path = os.path.join(HERE, "_build", "liboracle_count.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    P, d = C.c_void_p, C.c_double
    L.orc_stepper_create.argtypes = [P, P, P, C.c_int]
    L.orc_stepper_create.restype = P
    L.orc_stepper_destroy.argtypes = [P]
    L.orc_rk3_step.argtypes = [P, d, P, P, P, P, P]
    L.orc_flops_reset.argtypes = []
    L.orc_flops_read.argtypes = [P]
    return L


def count_rk3_step(nb=8, n=16):
    from grace_amd import workloads as W
    wl = W.blast(nb=(nb, nb, nb), n=n)
    g = wl.grid
    L = count_lib()
    st, B, aux = wl.copy_arrays()
    p = lambda a: C.c_void_p(a.ctypes.data)
    h = L.orc_stepper_create(C.byref(wl.params), C.byref(g), wl.instances, len(wl.inst_specs))
    assert h
    L.orc_flops_reset()
    L.orc_rk3_step(h, 0.25 * wl.extra["dx"], p(st), p(B[0]), p(B[1]), p(B[2]), p(aux))
    cnt = np.zeros((NPASS, 5), dtype=np.int64)
    L.orc_flops_read(p(cnt))
    L.orc_stepper_destroy(h)
    nq, N = g.nq, g.nx + 2 * g.ngz
    assert g.nx == g.ny == g.nz
    zones = nq * n ** 3
    # sites each pass of the oracle evaluates over one RK3 step (3 substeps)
    sites = {
        1: 3 * nq * 3 * (n + 1) * N * N,          # faces: normal [ng, n+ng], full transverse range
        2: 3 * nq * 3 * n * (n + 1) ** 2,         # edges of the interior
        3: 3 * nq * n ** 3,                       # interior cells
        4: 3 * nq * 3 * n * n * (n + 1),          # interior faces
        5: 3 * nq * N ** 3,                       # every cell, ghosts included
    }
    per_site, credit = {}, {}
    for pid, name in PASSES.items():
        per_site[name] = {k: cnt[pid, x] / sites[pid] for x, k in enumerate(KINDS)}
        credit[name] = CREDIT[pid]
    for sid, (name, pid) in SUBS.items():
        per_site[name] = {k: cnt[sid, x] / sites[pid] for x, k in enumerate(KINDS)}
        credit[name] = CREDIT[pid]
    unassigned = [int(i) for i in range(NPASS) if i not in PASSES and i not in SUBS and i != 6 and cnt[i].any()]

    def per_zone(names):
        out = {k: 0.0 for k in KINDS}
        for name in names:
            for k in KINDS:
                out[k] += 3 * credit[name] * per_site[name][k]   # 3 substeps
        # RK3 linear combinations: 2 per step of the 7 conserved + 3 face B,
        # each b*B + c*C (1 add, 2 mul)
        out["add"] += 2 * 10 * 1
        out["mul"] += 2 * 10 * 2
        out["flops"] = out["add"] + out["mul"]
        out["flops_incl_div_sqrt"] = out["flops"] + out["div"] + out["sqrt"]
        return out

    everything = list(per_site)
    flat = [nm for nm in everything if nm not in ("geometric_sources", "face_flux_metric", "cell_metric",
                                                  "geometric_sources_metric_derivatives", "c2p_cell_metric")]
    return dict(workload=f"blast {nb * n}^3 in {nq} blocks of {n}^3, WENO-Z + HLLE + UCT, ideal gas",
                zones=zones, per_site=per_site, sites_per_zone_substep=credit,
                per_zone_update_reference=per_zone(everything),
                per_zone_update_flat_metric=per_zone(flat),
                linop_ops_total=int(cnt[6, 0] + cnt[6, 1]), unassigned_pass_ids=unassigned)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=8)
    ap.add_argument("--n", type=int, default=16)
    a = ap.parse_args()
    print(json.dumps(count_rk3_step(a.blocks, a.n), indent=1))


if __name__ == "__main__":
    main()

"""grace_amd -- MI355X GRMHD RHS + Kastaun con2prim (Python view of the C ABI).

The compute path is libgrmhd_mi355x.so (hand-written gfx950 HIP kernels behind
the C ABI declared in include/grmhd_abi.h).  This module only binds that ABI
with ctypes and moves torch device tensors' pointers into it; it never falls
back to a CPU implementation: if the library is missing, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrmhd_mi355x.so")

# Variable indices (include/grmhd_abi.h)
DENS, SX, SY, SZ, TAU, YESTAR, ENTROPYSTAR = range(7)
GXX, GXY, GXZ, GYY, GYZ, GZZ, ALP, BETAX, BETAY, BETAZ = range(7, 17)
KXX, KXY, KXZ, KYY, KYZ, KZZ = range(17, 23)
NSTATE = 23
NHRSC = 7
(RHO, ZVECX, ZVECY, ZVECZ, BX, BY, BZ, YE, TEMP, ENTROPY, EPS, PRESS, BDIV, C2P_DENS_ERR,
 C2P_ERR) = range(15)
NAUX = 15
# c2p_err bits (c2p.hh:95-125)
C2P_BITS = ["RESET_DENS", "RESET_TAU", "RESET_STILDE", "RESET_ENTROPY", "RESET_YE",
            "SIG_RHO_TOO_LOW", "SIG_RHO_TOO_HIGH", "SIG_EPS_TOO_LOW", "SIG_EPS_TOO_HIGH",
            "SIG_YE_TOO_LOW", "SIG_YE_TOO_HIGH", "SIG_ENT_TOO_LOW", "SIG_ENT_TOO_HIGH",
            "SIG_TEMP_TOO_LOW", "SIG_TEMP_TOO_HIGH", "SIG_PRESS_TOO_LOW", "SIG_PRESS_TOO_HIGH",
            "SIG_VEL_TOO_HIGH", "SIG_SIGMA_TOO_HIGH", "ENT_BACKUP_USED", "ATMO_RESET", "T_FLOORED"]

EOS_IDEAL_GAS, EOS_HYBRID_PWPOLY = 0, 1
RECON_WENOZ, RECON_MC2, RECON_MINMOD, RECON_DONOR = 0, 1, 2, 3
RIEMANN_HLLE, RIEMANN_LLF = 0, 1
EMF_UCT, EMF_GS = 0, 1
THERMO_TEMP, THERMO_PRESS = 0, 1
BC_OUTFLOW, BC_PERIODIC = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("eos_kind", C.c_int32), ("gamma", C.c_double), ("eps_max", C.c_double),
        ("entropy_min", C.c_double), ("gamma_th", C.c_double), ("pwp_n", C.c_int32),
        ("pwp_kappa0", C.c_double), ("pwp_gamma", C.c_double * 8), ("pwp_rho", C.c_double * 8),
        ("rho_fl", C.c_double), ("temp_fl", C.c_double), ("ye_fl", C.c_double), ("atmo_tol", C.c_double),
        ("c2p_tol", C.c_double), ("max_lorentz", C.c_double), ("max_sigma", C.c_double),
        ("beta_fallback", C.c_double), ("use_entropy_backup", C.c_int32), ("bh_alp_thresh", C.c_double),
        ("excise_by_radius", C.c_int32), ("excision_radius", C.c_double), ("excision_lapse", C.c_double),
        ("excision_cks", C.c_int32), ("excision_spin", C.c_double),
        ("recon", C.c_int32), ("riemann", C.c_int32), ("emf", C.c_int32), ("thermo_recon", C.c_int32),
        ("fd_order", C.c_int32), ("rusanov_c_limit", C.c_double), ("flat_metric", C.c_int32),
        ("fofc", C.c_int32), ("fofc_dmp_enable", C.c_int32), ("fofc_dmp_M", C.c_double),
        ("fofc_ghost_emf", C.c_int32), ("stage_streams", C.c_int32),
    ]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ngz", C.c_int32), ("nq", C.c_int64)]


class Instance(C.Structure):
    _fields_ = [("nb", C.c_int32 * 3), ("q0", C.c_int64), ("bc", C.c_int32 * 6), ("idx", C.c_double * 3),
                ("x0", C.c_double * 3)]


class Forest(C.Structure):
    """grmhd_forest: two-level block forest (see include/grmhd_abi.h)."""
    _fields_ = [("nb", C.c_int32 * 3), ("bc", C.c_int32 * 6), ("idx", C.c_double * 3), ("n_refined", C.c_int32),
                ("refined", C.POINTER(C.c_int32)), ("x0", C.c_double * 3)]


def make_forest(nb, bc, idx, refined, x0=(0.0, 0.0, 0.0)):
    """A Forest struct plus the int32 array it points to (keep both alive);
    x0: lower corner of the domain (cell radii for radius excision)."""
    ref = np.ascontiguousarray(sorted(set(int(r) for r in refined)), dtype=np.int32)
    f = Forest()
    for d in range(3):
        f.nb[d] = int(nb[d]); f.idx[d] = float(idx[d]); f.x0[d] = float(x0[d])
    for d in range(6):
        f.bc[d] = int(bc[d])
    f.n_refined = len(ref)
    f.refined = ref.ctypes.data_as(C.POINTER(C.c_int32)) if len(ref) else C.POINTER(C.c_int32)()
    f._keep = ref
    return f


class C2PStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("n_cells", "n_fatal", "n_atmo_reset", "n_ent_backup", "n_guess_used",
                                          "n_root_evals", "n_bracket_evals", "max_root_evals")]


_lib = None

ABI_SYMBOLS = [
    "grmhd_default_params", "grmhd_ctx_create", "grmhd_ctx_destroy", "grmhd_last_error",
    "grmhd_ctx_create_forest", "grmhd_forest_block", "grmhd_setup_forest_wave", "grmhd_setup_fm_torus_forest",
    "grmhd_fill_metric_ghosts",
    "grmhd_advance_substep", "grmhd_fill_ghosts", "grmhd_compute_auxiliaries", "grmhd_c2p_batch",
    "grmhd_c2p_batch_noalloc", "grmhd_advance_substep_t", "grmhd_sync", "grmhd_find_stable_timestep",
    "grmhd_set_c2p_warm_start", "grmhd_set_c2p_eval_counts",
    "grmhd_prims_to_conservs_batch", "grmhd_c2p_stats", "grmhd_compute_dt", "grmhd_rk3_step",
    "grmhd_rk3_step_graph", "grmhd_set_timing", "grmhd_kernel_time_ms",
    "grmhd_autotune", "grmhd_get_tuning", "grmhd_set_tuning",
    "grmhd_setup_blast", "grmhd_setup_shocktube", "grmhd_kerr_schild", "grmhd_setup_c2p_states",
    "grmhd_kerr_schild_adm", "grmhd_setup_blast_ks", "grmhd_setup_fm_torus", "grmhd_fm_torus_point",
    # include/grmhd_mlp.h
    "grmhd_mlp_init_params", "grmhd_mlp_create", "grmhd_mlp_destroy", "grmhd_mlp_set_params",
    "grmhd_mlp_get_params", "grmhd_mlp_set_scaling", "grmhd_mlp_get_scaling", "grmhd_mlp_set_bracket",
    "grmhd_mlp_get_bracket", "grmhd_mlp_set_newton", "grmhd_mlp_get_newton",
    "grmhd_mlp_fit_cells", "grmhd_set_learned_c2p", "grmhd_set_static_metric", "grmhd_mlp_features", "grmhd_mlp_targets", "grmhd_mlp_forward", "grmhd_mlp_grad",
    "grmhd_mlp_train_step", "grmhd_mlp_default_fit_opts", "grmhd_mlp_fit", "grmhd_c2p_batch_learned",
]

MLP_NPARAM = 5378
MLP_NSCALE = 24


class Tuning(C.Structure):
    _fields_ = [("flux", C.c_int32 * 3), ("emf", C.c_int32), ("update", C.c_int32), ("gather_faces", C.c_int32),
                ("gather_cells", C.c_int32), ("c2p", C.c_int32), ("flux_prefetch", C.c_int32 * 3)]

    def as_dict(self):
        return {"flux": list(self.flux), "emf": self.emf, "update": self.update, "gather_faces": self.gather_faces,
                "gather_cells": self.gather_cells, "c2p": self.c2p, "flux_prefetch": list(self.flux_prefetch)}


class MlpFitOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("val_frac", C.c_float),
                ("delta_quantile", C.c_double), ("delta_scale", C.c_double), ("nexp", C.c_int32),
                ("seed", C.c_uint64)]


class MlpFitReport(C.Structure):
    _fields_ = [("n_train", C.c_int64), ("n_val", C.c_int64), ("steps", C.c_int64)] + \
               [(n, C.c_double) for n in ("loss_first", "loss_last", "val_mae_e", "val_p50_e", "val_p99_e",
                                           "val_max_e", "val_mae_s", "val_p99_s", "delta_e", "delta_s", "seconds")]


def lib():
    """Load libgrmhd_mi355x.so; raise if it is not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m grace_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    d = C.c_double
    i32, i64 = C.c_int32, C.c_int64
    L.grmhd_default_params.argtypes = [C.POINTER(Params)]
    L.grmhd_ctx_create.argtypes = [C.POINTER(Params), C.POINTER(Grid), C.POINTER(Instance), i32, C.POINTER(P)]
    L.grmhd_ctx_destroy.argtypes = [P]
    L.grmhd_ctx_create_forest.argtypes = [C.POINTER(Params), C.POINTER(Grid), C.POINTER(Forest), C.POINTER(P)]
    L.grmhd_forest_block.argtypes = [P, i64, C.POINTER(i32), P]
    L.grmhd_setup_forest_wave.argtypes = [C.POINTER(Grid), C.POINTER(Forest), d, P, d, d, d, d, d, P, P, P, d,
                                          P, P, P, P, P]
    L.grmhd_last_error.restype = C.c_char_p
    L.grmhd_advance_substep.argtypes = [P, P, d, P, P, P * 3, P * 3, P, P]
    L.grmhd_advance_substep_t.argtypes = [P, d, d, d, P, P, P * 3, P * 3, P, P]
    L.grmhd_sync.argtypes = [P, P]
    L.grmhd_find_stable_timestep.argtypes = [P, P, P, d, C.POINTER(d), P]
    L.grmhd_fill_ghosts.argtypes = [P, P, P * 3, P]
    L.grmhd_fill_metric_ghosts.argtypes = [P, P, P]
    L.grmhd_compute_auxiliaries.argtypes = [P, P, P * 3, P, P]
    L.grmhd_c2p_batch.argtypes = [P, i64, P, P, P, P, P, i32, P, P, P, P]
    L.grmhd_c2p_batch_noalloc.argtypes = [P, i64, P, P, P, P, P, i32, P, P, P, P]
    L.grmhd_prims_to_conservs_batch.argtypes = [P, i64, P, P, P, P]
    L.grmhd_c2p_stats.argtypes = [P, C.POINTER(C2PStats), P]
    L.grmhd_compute_dt.argtypes = [P, P, P, d, P, P]
    L.grmhd_rk3_step.argtypes = [P, P, P, P * 3, P, P]
    L.grmhd_rk3_step_graph.argtypes = [P, P, P, P * 3, P, P]
    L.grmhd_set_timing.argtypes = [P, i32]
    L.grmhd_kernel_time_ms.argtypes = [P, C.c_char_p, C.POINTER(d), C.POINTER(i64)]
    L.grmhd_autotune.argtypes = [P, P, P * 3, P, i32, C.POINTER(Tuning), P]
    L.grmhd_get_tuning.argtypes = [P, C.POINTER(Tuning)]
    L.grmhd_set_tuning.argtypes = [P, C.POINTER(Tuning)]
    L.grmhd_setup_blast.argtypes = [C.POINTER(Grid), C.POINTER(Instance), i32, d, P, P, d, d, d, d, P, d,
                                    C.c_uint64, P, P, P, P, P]
    L.grmhd_setup_shocktube.argtypes = [C.POINTER(Grid), C.POINTER(Instance), i32, d, d, d, P, P, P, P, P, P, P]
    L.grmhd_kerr_schild.argtypes = [d, d, d, d, P, P, P]
    L.grmhd_kerr_schild_adm.argtypes = [d, d, d, d, P, P, P, P]
    L.grmhd_setup_blast_ks.argtypes = [C.POINTER(Grid), C.POINTER(Instance), i32, d, P, P, d, d, d, d, d, P, d,
                                       C.c_uint64, P, P, P, P, P]
    L.grmhd_setup_c2p_states.argtypes = [i64, C.c_uint64, d, d, P, P, P, P, P, P, P, P, P]
    L.grmhd_setup_fm_torus.argtypes = [C.POINTER(Grid), C.POINTER(Instance), i32, d, P, d, d, d, d, d, d, d,
                                       C.c_uint64, P, P, P, P, P, P]
    L.grmhd_setup_fm_torus_forest.argtypes = [C.POINTER(Grid), C.POINTER(Forest), d, P, d, d, d, d, d, d, d,
                                              C.c_uint64, P, P, P, P, P, P]
    L.grmhd_fm_torus_point.argtypes = [d, d, d, d, d, d, d, P]
    f32 = C.c_float
    L.grmhd_mlp_init_params.argtypes = [C.c_uint64, P]
    L.grmhd_mlp_create.argtypes = [C.c_uint64, C.POINTER(P)]
    L.grmhd_mlp_destroy.argtypes = [P]
    L.grmhd_mlp_destroy.restype = None
    L.grmhd_mlp_set_params.argtypes = [P, P, P]
    L.grmhd_mlp_get_params.argtypes = [P, P, P]
    L.grmhd_mlp_set_scaling.argtypes = [P, P, P]
    L.grmhd_mlp_get_scaling.argtypes = [P, P]
    L.grmhd_mlp_set_bracket.argtypes = [P, d, d, i32]
    L.grmhd_mlp_get_bracket.argtypes = [P, C.POINTER(d), C.POINTER(d), C.POINTER(i32)]
    L.grmhd_mlp_set_newton.argtypes = [P, i32]
    L.grmhd_mlp_get_newton.argtypes = [P, C.POINTER(i32)]
    L.grmhd_mlp_features.argtypes = [P, i64, P, P, P, P]
    L.grmhd_mlp_targets.argtypes = [i64, P, P, P, P]
    L.grmhd_mlp_forward.argtypes = [P, i64, P, P, P]
    L.grmhd_mlp_grad.argtypes = [P, i64, P, P, P, P]
    L.grmhd_mlp_train_step.argtypes = [P, i64, P, P, f32, P, P]
    L.grmhd_mlp_default_fit_opts.argtypes = [C.POINTER(MlpFitOpts)]
    L.grmhd_mlp_default_fit_opts.restype = None
    L.grmhd_mlp_fit.argtypes = [P, P, i64, P, P, P, C.POINTER(MlpFitOpts), C.POINTER(MlpFitReport), P]
    L.grmhd_c2p_batch_learned.argtypes = [P, P, i64, P, P, P, P, P, P, P, P]
    L.grmhd_mlp_fit_cells.argtypes = [P, P, P, P * 3, i64, C.POINTER(MlpFitOpts), C.POINTER(MlpFitReport), P]
    L.grmhd_set_learned_c2p.argtypes = [P, P]
    L.grmhd_set_c2p_warm_start.argtypes = [P, i32]
    L.grmhd_set_c2p_eval_counts.argtypes = [P, P]
    L.grmhd_set_static_metric.argtypes = [P, P, i32, P]
    for name in ABI_SYMBOLS:
        getattr(L, name)  # raises AttributeError if an exported symbol is missing
    _lib = L
    return L


def default_params(**kw):
    p = Params()
    lib().grmhd_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        if k in ("pwp_gamma", "pwp_rho"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def check(status, what="call"):
    if status != 0:
        msg = lib().grmhd_last_error().decode(errors="replace")
        raise RuntimeError(f"grmhd {what} failed with status {status}: {msg}")


def _ptr(t):
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return C.c_void_p(t.data_ptr())
    if hasattr(t, "ctypes"):
        return C.c_void_p(t.ctypes.data)
    return C.c_void_p(t)


def ptr3(a, b, c):
    return (C.c_void_p * 3)(_ptr(a), _ptr(b), _ptr(c))


def make_instances(specs):
    """specs: list of dict(nb=(bx,by,bz), q0=int, bc=(6 ints), idx=(3 floats)
    [, x0=(3 floats): lower corner of the interior, default the origin])."""
    arr = (Instance * max(1, len(specs)))()
    for i, s in enumerate(specs):
        for d in range(3):
            arr[i].nb[d] = int(s["nb"][d])
            arr[i].idx[d] = float(s["idx"][d])
            arr[i].x0[d] = float(s.get("x0", (0.0, 0.0, 0.0))[d])
        arr[i].q0 = int(s["q0"])
        for d in range(6):
            arr[i].bc[d] = int(s["bc"][d])
    return arr


class Context:
    """Owns a grmhd_ctx; methods take torch device tensors."""

    def __init__(self, params, grid, instances, forest=None):
        self.L = lib()
        self.params = params
        self.grid = grid
        self.instances = instances
        self.forest = forest
        h = C.c_void_p()
        if forest is not None:
            check(self.L.grmhd_ctx_create_forest(C.byref(params), C.byref(grid), C.byref(forest), C.byref(h)),
                  "ctx_create_forest")
        else:
            check(self.L.grmhd_ctx_create(C.byref(params), C.byref(grid), instances, len(instances), C.byref(h)),
                  "ctx_create")
        self.h = h

    def forest_block(self, q):
        """(level, (cx, cy, cz)) of block q of a forest context."""
        lev = C.c_int32()
        crd = np.zeros(3, dtype=np.int32)
        check(self.L.grmhd_forest_block(self.h, int(q), C.byref(lev), crd.ctypes.data), "forest_block")
        return lev.value, tuple(int(x) for x in crd)

    def close(self):
        if self.h:
            self.L.grmhd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- entry points ---
    def advance_substep(self, dt_dev, dtfact, new_state, old_state, newB, oldB, aux, stream=None):
        check(self.L.grmhd_advance_substep(self.h, _ptr(dt_dev), dtfact, _ptr(new_state), _ptr(old_state),
                                           ptr3(*newB), ptr3(*oldB), _ptr(aux), _ptr(stream)), "advance_substep")

    def advance_substep_t(self, t, dt, dtfact, new_state, old_state, newB, oldB, aux, stream=None):
        check(self.L.grmhd_advance_substep_t(self.h, t, dt, dtfact, _ptr(new_state), _ptr(old_state), ptr3(*newB),
                                             ptr3(*oldB), _ptr(aux), _ptr(stream)), "advance_substep_t")

    def sync(self, stream=None):
        """Status of the sync point (0, or 5 = GRMHD_ERR_C2P_FATAL after fatal c2p cells)."""
        return self.L.grmhd_sync(self.h, _ptr(stream))

    def find_stable_timestep(self, state, aux, cfl, stream=None):
        dt = C.c_double()
        check(self.L.grmhd_find_stable_timestep(self.h, _ptr(state), _ptr(aux), cfl, C.byref(dt), _ptr(stream)),
              "find_stable_timestep")
        return dt.value

    def fill_ghosts(self, state, B, stream=None):
        check(self.L.grmhd_fill_ghosts(self.h, _ptr(state), ptr3(*B), _ptr(stream)), "fill_ghosts")

    def fill_metric_ghosts(self, state, stream=None):
        check(self.L.grmhd_fill_metric_ghosts(self.h, _ptr(state), _ptr(stream)), "fill_metric_ghosts")

    def compute_auxiliaries(self, state, B, aux, stream=None):
        check(self.L.grmhd_compute_auxiliaries(self.h, _ptr(state), ptr3(*B), _ptr(aux), _ptr(stream)),
              "compute_auxiliaries")

    def c2p_batch(self, n, cons, met, prims, err, floored, dry_run=0, r=None, guess=None, roots=None, stream=None):
        check(self.L.grmhd_c2p_batch(self.h, n, _ptr(cons), _ptr(met), _ptr(prims), _ptr(err), _ptr(floored),
                                     int(dry_run), _ptr(r), _ptr(guess), _ptr(roots), _ptr(stream)), "c2p_batch")

    def c2p_batch_noalloc(self, n, cons, met, prims, err, floored, dry_run=0, r=None, guess=None, roots=None,
                          stream=None):
        check(self.L.grmhd_c2p_batch_noalloc(self.h, n, _ptr(cons), _ptr(met), _ptr(prims), _ptr(err),
                                             _ptr(floored), int(dry_run), _ptr(r), _ptr(guess), _ptr(roots),
                                             _ptr(stream)), "c2p_batch_noalloc")

    def p2c_batch(self, n, prims, met, cons, stream=None):
        check(self.L.grmhd_prims_to_conservs_batch(self.h, n, _ptr(prims), _ptr(met), _ptr(cons), _ptr(stream)),
              "p2c_batch")

    def c2p_stats(self, stream=None):
        s = C2PStats()
        check(self.L.grmhd_c2p_stats(self.h, C.byref(s), _ptr(stream)), "c2p_stats")
        return {k: getattr(s, k) for k, _ in C2PStats._fields_}

    def compute_dt(self, state, aux, cfl, dt_dev, stream=None):
        check(self.L.grmhd_compute_dt(self.h, _ptr(state), _ptr(aux), float(cfl), _ptr(dt_dev), _ptr(stream)),
              "compute_dt")

    def rk3_step(self, dt_dev, state, B, aux, stream=None, graph=False):
        fn = self.L.grmhd_rk3_step_graph if graph else self.L.grmhd_rk3_step
        check(fn(self.h, _ptr(dt_dev), _ptr(state), ptr3(*B), _ptr(aux), _ptr(stream)), "rk3_step")

    def set_static_metric(self, state, enable=True, stream=None):
        check(self.L.grmhd_set_static_metric(self.h, _ptr(state) if enable else None, int(bool(enable)),
                                             _ptr(stream)), "set_static_metric")

    def set_learned_c2p(self, mlp):
        """Cell-pass c2p from the MLP's guess (mlp: Mlp or None for plain)."""
        check(self.L.grmhd_set_learned_c2p(self.h, mlp.h if mlp is not None else None), "set_learned_c2p")
        self._mlp = mlp   # keep it alive while the context uses it

    def set_c2p_warm_start(self, max_newton_evals=8):
        check(self.L.grmhd_set_c2p_warm_start(self.h, int(max_newton_evals)), "set_c2p_warm_start")

    def set_c2p_eval_counts(self, counts):
        """Per-state evaluation counts of the batch c2p into a uint8 device
        tensor of n elements (None: off)."""
        check(self.L.grmhd_set_c2p_eval_counts(self.h, _ptr(counts)), "set_c2p_eval_counts")
        self._eval_counts = counts   # keep it alive while the context writes it

    def autotune(self, state, B, aux, reps=3, stream=None):
        t = Tuning()
        check(self.L.grmhd_autotune(self.h, _ptr(state), ptr3(*B), _ptr(aux), int(reps), C.byref(t), _ptr(stream)),
              "autotune")
        return t.as_dict()

    def get_tuning(self):
        t = Tuning()
        check(self.L.grmhd_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, flux=(256, 256, 256), emf=256, update=256, gather_faces=256, gather_cells=256, c2p=256,
                   flux_prefetch=(0, 0, 0)):
        t = Tuning()
        for d in range(3):
            t.flux[d] = flux[d]
            t.flux_prefetch[d] = flux_prefetch[d]
        t.emf, t.update, t.gather_faces, t.gather_cells, t.c2p = emf, update, gather_faces, gather_cells, c2p
        check(self.L.grmhd_set_tuning(self.h, C.byref(t)), "set_tuning")

    def set_timing(self, on):
        check(self.L.grmhd_set_timing(self.h, int(on)), "set_timing")

    def kernel_time_ms(self, fam):
        ms = C.c_double()
        n = C.c_int64()
        check(self.L.grmhd_kernel_time_ms(self.h, fam.encode(), C.byref(ms), C.byref(n)), "kernel_time_ms")
        return ms.value, n.value


class Mlp:
    """Handle of the learned Kastaun bracket network (include/grmhd_mlp.h)."""

    def __init__(self, seed=20261202):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.grmhd_mlp_create(C.c_uint64(seed), C.byref(h)), "mlp_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.grmhd_mlp_destroy(self.h)
            self.h = None

    @staticmethod
    def init_params(seed):
        import numpy as np
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        check(lib().grmhd_mlp_init_params(C.c_uint64(seed), _ptr(p)), "mlp_init_params")
        return p

    def set_params(self, p, stream=None):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == MLP_NPARAM
        check(self.L.grmhd_mlp_set_params(self.h, _ptr(p), _ptr(stream)), "mlp_set_params")

    def get_params(self, stream=None):
        import numpy as np
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        check(self.L.grmhd_mlp_get_params(self.h, _ptr(p), _ptr(stream)), "mlp_get_params")
        return p

    def set_scaling(self, sc, stream=None):
        import numpy as np
        sc = np.ascontiguousarray(sc, dtype=np.float32)
        assert sc.size == MLP_NSCALE
        check(self.L.grmhd_mlp_set_scaling(self.h, _ptr(sc), _ptr(stream)), "mlp_set_scaling")

    def get_scaling(self):
        import numpy as np
        sc = np.zeros(MLP_NSCALE, dtype=np.float32)
        check(self.L.grmhd_mlp_get_scaling(self.h, _ptr(sc)), "mlp_get_scaling")
        return sc

    def set_bracket(self, delta_e, delta_s, nexp):
        check(self.L.grmhd_mlp_set_bracket(self.h, float(delta_e), float(delta_s), int(nexp)), "mlp_set_bracket")

    def set_newton(self, max_evals):
        check(self.L.grmhd_mlp_set_newton(self.h, int(max_evals)), "mlp_set_newton")

    def get_newton(self):
        v = C.c_int32()
        check(self.L.grmhd_mlp_get_newton(self.h, C.byref(v)), "mlp_get_newton")
        return v.value

    def get_bracket(self):
        a, b, c = C.c_double(), C.c_double(), C.c_int32()
        check(self.L.grmhd_mlp_get_bracket(self.h, C.byref(a), C.byref(b), C.byref(c)), "mlp_get_bracket")
        return a.value, b.value, c.value

    def forward(self, n, feat, out, stream=None):
        check(self.L.grmhd_mlp_forward(self.h, n, _ptr(feat), _ptr(out), _ptr(stream)), "mlp_forward")

    def grad(self, n, feat, target, grad_out, stream=None):
        check(self.L.grmhd_mlp_grad(self.h, n, _ptr(feat), _ptr(target), _ptr(grad_out), _ptr(stream)), "mlp_grad")

    def train_step(self, n, feat, target, lr, loss_out=None, stream=None):
        check(self.L.grmhd_mlp_train_step(self.h, n, _ptr(feat), _ptr(target), float(lr), _ptr(loss_out),
                                          _ptr(stream)), "mlp_train_step")

    @staticmethod
    def default_fit_opts(**kw):
        o = MlpFitOpts()
        lib().grmhd_mlp_default_fit_opts(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def fit_cells(self, ctx, state, B, max_states=1 << 20, opts=None, stream=None):
        """Train on the c2p inputs of up to max_states interior cells of ctx's blocks."""
        rep = MlpFitReport()
        o = opts or self.default_fit_opts()
        check(self.L.grmhd_mlp_fit_cells(ctx.h, self.h, _ptr(state), ptr3(*B), int(max_states), C.byref(o),
                                         C.byref(rep), _ptr(stream)), "mlp_fit_cells")
        return {k: getattr(rep, k) for k, _ in MlpFitReport._fields_}

    def fit(self, ctx, n, cons, met, r=None, opts=None, stream=None):
        rep = MlpFitReport()
        o = opts or self.default_fit_opts()
        check(self.L.grmhd_mlp_fit(ctx.h, self.h, n, _ptr(cons), _ptr(met), _ptr(r), C.byref(o), C.byref(rep),
                                   _ptr(stream)), "mlp_fit")
        return {k: getattr(rep, k) for k, _ in MlpFitReport._fields_}


def mlp_features(ctx, n, cons, met, feat, stream=None):
    check(lib().grmhd_mlp_features(ctx.h, n, _ptr(cons), _ptr(met), _ptr(feat), _ptr(stream)), "mlp_features")


def mlp_targets(n, roots, feat, target, stream=None):
    check(lib().grmhd_mlp_targets(n, _ptr(roots), _ptr(feat), _ptr(target), _ptr(stream)), "mlp_targets")


def c2p_batch_learned(ctx, mlp, n, cons, met, prims, err, floored, r=None, roots=None, stream=None):
    check(lib().grmhd_c2p_batch_learned(ctx.h, mlp.h, n, _ptr(cons), _ptr(met), _ptr(prims), _ptr(err),
                                        _ptr(floored), _ptr(r), _ptr(roots), _ptr(stream)), "c2p_batch_learned")


# ---- HDF5 state dumps (include/grmhd_io.h, libgrmhd_io.so) -------------------
IO_LIB_PATH = os.path.join(HERE, "libgrmhd_io.so")
_io = None


def io_lib():
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise RuntimeError(f"{IO_LIB_PATH} is not built: run `python -m grace_amd.build`")
    L = C.CDLL(IO_LIB_PATH)
    P = C.c_void_p
    L.grmhd_io_write.argtypes = [C.c_char_p, C.c_double, C.c_uint32, C.POINTER(Grid), C.c_int32, P, P, P * 3]
    L.grmhd_io_info.argtypes = [C.c_char_p, C.POINTER(Grid), C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                C.POINTER(C.c_uint32), P]
    L.grmhd_io_read.argtypes = [C.c_char_p, P, P * 3]
    L.grmhd_io_last_error.restype = C.c_char_p
    _io = L
    return L


def _io_check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {io_lib().grmhd_io_last_error().decode()}")


def write_state(path, grid, state, B, time=0.0, iteration=0, extents=(0, 1, 0, 1, 0, 1)):
    """Host numpy arrays (layout of the C ABI) -> HDF5 dump."""
    import numpy as np
    ext = np.ascontiguousarray(extents, dtype=np.float64)
    nv = state.shape[1]
    _io_check(io_lib().grmhd_io_write(os.fsencode(path), float(time), int(iteration), C.byref(grid), nv, _ptr(ext),
                                      _ptr(np.ascontiguousarray(state)), ptr3(*[np.ascontiguousarray(b) for b in B])),
              "write_state")


def read_state(path):
    """HDF5 dump -> (grid, state, B, time, iteration, extents) as host numpy arrays."""
    import numpy as np
    g, nv, t, it = Grid(), C.c_int32(), C.c_double(), C.c_uint32()
    ext = np.zeros(6)
    _io_check(io_lib().grmhd_io_info(os.fsencode(path), C.byref(g), C.byref(nv), C.byref(t), C.byref(it), _ptr(ext)),
              "read_state")
    N = [g.nx + 2 * g.ngz, g.ny + 2 * g.ngz, g.nz + 2 * g.ngz]
    state = np.zeros((g.nq, nv.value, N[2], N[1], N[0]))
    B = [np.zeros((g.nq, 1, N[2] + (d == 2), N[1] + (d == 1), N[0] + (d == 0))) for d in range(3)]
    _io_check(io_lib().grmhd_io_read(os.fsencode(path), _ptr(state), ptr3(*B)), "read_state")
    return g, state, B, t.value, it.value, ext

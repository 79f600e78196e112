"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) -- test-side only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_libs = {}


def olib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(ROOT, "oracle", "_build", name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    P, d, i32, i64 = C.c_void_p, C.c_double, C.c_int, C.c_int64
    L.orc_num_threads.argtypes = [i32]
    L.orc_num_threads.restype = i32
    L.orc_c2p_batch.argtypes = [P, i64, P, P, P, P, P, i32, P, P, P, P, P]
    L.orc_p2c_batch.argtypes = [P, i64, P, P, P]
    L.orc_eos_from_temp.argtypes = [P, d, d, P, P, P, P]
    L.orc_eos_from_temp.restype = d
    L.orc_W.argtypes = [P, P]
    L.orc_W.restype = d
    L.orc_smallb.argtypes = [d, P, P, P, P, d, P, P]
    L.orc_vtilde.argtypes = [d, P, P, d, P]
    L.orc_cm_cp.argtypes = [d, P, d, d, d, d, d, d, P, d, i32, P, P]
    L.orc_fluxes.argtypes = [d, P, P, d, d, d, P, d, d, d, P, P, i32, P, P]
    L.orc_geom_sources.argtypes = [d, P, P, P, P, P, P, P, d, d, d, P, P, d, P]
    L.orc_recon.argtypes = [i32, P, P]
    L.orc_metric.argtypes = [P, P, P]
    L.orc_brent_poly.argtypes = [d, d, d, d, d, P]
    L.orc_brent_poly.restype = d
    L.orc_kastaun_f.argtypes = [P, P, P, d, P]
    L.orc_kastaun_f.restype = d
    L.orc_stepper_create.argtypes = [P, P, P, i32]
    L.orc_stepper_create.restype = P
    L.orc_stepper_create_forest.argtypes = [P, P, P]
    L.orc_stepper_create_forest.restype = P
    L.orc_fill_metric_ghosts.argtypes = [P, P]
    L.orc_forest_block.argtypes = [P, i64, P, P]
    L.orc_forest_block.restype = i32
    L.orc_stepper_destroy.argtypes = [P]
    L.orc_rk3_step.argtypes = [P, d, P, P, P, P, P]
    L.orc_advance_substep.argtypes = [P, d, d, P, P, P, P, P, P, P, P, P]
    L.orc_fill_ghosts.argtypes = [P, P, P, P, P]
    L.orc_compute_aux.argtypes = [P, P, P, P, P, P]
    L.orc_find_dt.argtypes = [P, P, P, d]
    L.orc_find_dt.restype = d
    L.orc_fatal_count.argtypes = [P]
    L.orc_fatal_count.restype = i64
    L.orc_fofc_flagged.argtypes = [P]
    L.orc_fofc_flagged.restype = i64
    L.orc_face4.argtypes = [P]
    L.orc_face4.restype = d
    L.orc_fd_der.argtypes = [i32, P, d]
    L.orc_fd_der.restype = d
    L.orc_stepper_scratch.argtypes = [P, i32, i32, P]
    L.orc_stepper_scratch.restype = i64
    _libs[name] = L
    return L


def p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def arr(*vals):
    return np.array(vals, dtype=np.float64)


def c2p_batch(params, cons, met, dry_run=0, r=None, guess=None, fast=False, roots=None):
    """cons (10,n) densitised, met (10,n).  Returns prims, cons_out, err, floored, evals, fatal.
    guess: (4,n) energy/entropy brackets as fractions of mu0; roots: (4,n)
    array filled with mu0_e, t_e, mu0_s, t_s."""
    L = olib(fast)
    n = cons.shape[1]
    cons = np.ascontiguousarray(cons, dtype=np.float64).copy()
    met = np.ascontiguousarray(met, dtype=np.float64)
    prims = np.zeros((13, n))
    err = np.zeros(n, dtype=np.uint64)
    fl = np.zeros(n, dtype=np.uint8)
    ev = np.zeros((n, 2), dtype=np.int64)
    fatal = np.zeros(1, dtype=np.int64)
    L.orc_c2p_batch(C.byref(params), n, p(cons), p(met), p(prims), p(err), p(fl), int(dry_run),
                    p(r), p(guess), p(roots), p(ev), p(fatal))
    return prims, cons, err, fl, ev, int(fatal[0])


class Stepper:
    def __init__(self, wl):
        self.L = olib()
        self.wl = wl
        if getattr(wl, "forest", None) is not None:
            self.h = self.L.orc_stepper_create_forest(C.byref(wl.params), C.byref(wl.grid), C.byref(wl.forest))
        else:
            self.h = self.L.orc_stepper_create(C.byref(wl.params), C.byref(wl.grid), wl.instances,
                                               len(wl.inst_specs))
        assert self.h, "oracle stepper not created"

    def __del__(self):
        try:
            self.L.orc_stepper_destroy(self.h)
        except Exception:
            pass

    def rk3(self, dt, state, B, aux):
        self.L.orc_rk3_step(self.h, dt, p(state), p(B[0]), p(B[1]), p(B[2]), p(aux))

    def substep(self, dt, dtfact, new, old, nB, oB, aux):
        self.L.orc_advance_substep(self.h, dt, dtfact, p(new), p(old), p(nB[0]), p(nB[1]), p(nB[2]),
                                   p(oB[0]), p(oB[1]), p(oB[2]), p(aux))

    def ghosts(self, state, B):
        self.L.orc_fill_ghosts(self.h, p(state), p(B[0]), p(B[1]), p(B[2]))

    def aux(self, state, B, aux):
        self.L.orc_compute_aux(self.h, p(state), p(B[0]), p(B[1]), p(B[2]), p(aux))

    def dt(self, state, aux, cfl):
        return self.L.orc_find_dt(self.h, p(state), p(aux), cfl)

    def fatal(self):
        return self.L.orc_fatal_count(self.h)

    def metric_ghosts(self, state):
        self.L.orc_fill_metric_ghosts(self.h, p(state))

    def forest_block(self, q):
        lev = np.zeros(1, dtype=np.int32)
        crd = np.zeros(3, dtype=np.int32)
        assert self.L.orc_forest_block(self.h, int(q), lev.ctypes.data, crd.ctypes.data) == 0
        return int(lev[0]), tuple(int(x) for x in crd)

    def fofc_flagged(self):
        return self.L.orc_fofc_flagged(self.h)

    def scratch(self, which, d=0):
        """Copy of the last substep's scratch: which 0 fluxes [q][7][face d],
        1 face values [q][4][face d], 2 edge EMFs [q][3][edge], 3 cell E."""
        n = self.L.orc_stepper_scratch(self.h, which, d, None)
        assert n >= 0
        out = np.zeros(n)
        self.L.orc_stepper_scratch(self.h, which, d, out.ctypes.data)
        return out

"""Synthetic inputs of the benchmark configurations, built on host numpy arrays.

Every generator fills arrays in the library's layout (see include/grmhd_abi.h)
through the library's own host-side initial-data routines, so the GPU and any
CPU checker receive bit-identical inputs.  Seed 20261202 unless stated.
"""
import numpy as np

from . import (lib, Grid, make_instances, make_forest, default_params, NSTATE, NAUX, BC_OUTFLOW, BC_PERIODIC,
               EMF_GS, EMF_UCT)

SEED = 20261202


class Workload:
    """Grid + instances + host arrays (state, B[3], aux) + parameters."""

    def __init__(self, name, params, grid, inst_specs, state, B, aux, zones, extra=None):
        self.name = name
        self.params = params
        self.grid = grid
        self.inst_specs = inst_specs
        self.instances = make_instances(inst_specs)
        self.state = state
        self.B = B
        self.aux = aux
        self.zones = zones
        self.extra = extra or {}
        self.forest = None   # grmhd_forest of a two-level forest workload

    def copy_arrays(self):
        return self.state.copy(), [b.copy() for b in self.B], self.aux.copy()


def alloc(grid):
    Ni, Nj, Nk = grid.nx + 2 * grid.ngz, grid.ny + 2 * grid.ngz, grid.nz + 2 * grid.ngz
    nq = grid.nq
    state = np.zeros((nq, NSTATE, Nk, Nj, Ni))
    aux = np.zeros((nq, NAUX, Nk, Nj, Ni))
    B = [np.zeros((nq, 1, Nk, Nj, Ni + 1)), np.zeros((nq, 1, Nk, Nj + 1, Ni)), np.zeros((nq, 1, Nk + 1, Nj, Ni))]
    return state, B, aux


def _p(a):
    return a.ctypes.data


def forest_wave(nb=(4, 4, 4), n=8, ngz=4, refined=None, bc=BC_PERIODIC, extent=1.0, gamma=4.0 / 3.0,
                rho0=1.0, amp=0.2, p0=1.0, p_in=0.0, r_in=0.0, zvel=(0.3, 0.2, 0.1), Bvec=(0.2, 0.1, 0.05),
                b_loop=0.1, emf=EMF_UCT, recon=0, riemann=0):
    """Flat-space data on a two-level forest: nb coarse blocks of n^3 cubic
    cells of side extent / (nb[0] n), the box centred on the origin, the
    coarse blocks in `refined` (coarse ids
    bi + nb0 (bj + nb1 bk); default: the central 2x2x2, or the first block when
    nb has a side < 2) refined once.  A density wave rho0 (1 + amp sin sin),
    pressure p0 (p_in inside r < r_in at the centre), uniform z = W v, face B =
    Bvec plus a divergence-free loop of strength b_loop."""
    if refined is None:
        if min(nb) >= 2:
            c = [nb[d] // 2 - 1 for d in range(3)]
            refined = [(c[0] + a) + nb[0] * ((c[1] + b) + nb[1] * (c[2] + e))
                       for e in range(2) for b in range(2) for a in range(2)]
        else:
            refined = [0]
    bcs = (bc,) * 6 if isinstance(bc, int) else tuple(bc)
    dx = extent / (nb[0] * n)
    idx = (1.0 / dx,) * 3
    forest = make_forest(nb, bcs, idx, refined, x0=[-0.5 * nb[d] * n * dx for d in range(3)])
    nq = nb[0] * nb[1] * nb[2] + 7 * forest.n_refined
    grid = Grid(n, n, n, ngz, nq)
    state, B, aux = alloc(grid)
    lo = np.array([-0.5 * nb[d] * n * dx for d in range(3)])
    ctr = np.zeros(3)
    zv = np.array(zvel, dtype=np.float64)
    Bv = np.array(Bvec, dtype=np.float64)
    rc = lib().grmhd_setup_forest_wave(grid, forest, gamma, _p(lo), rho0, amp, p0, p_in, r_in, _p(ctr), _p(zv),
                                       _p(Bv), b_loop, _p(state), _p(B[0]), _p(B[1]), _p(B[2]), _p(aux))
    assert rc == 0, rc
    params = default_params(gamma=gamma, emf=emf, recon=recon, riemann=riemann, flat_metric=1)
    wl = Workload("forest_wave", params, grid, [], state, B, aux, zones=nq * n ** 3,
                  extra=dict(dx=dx, extent=extent))
    wl.forest = forest
    return wl


def blast(nb=(8, 8, 8), n=16, ngz=3, seed=SEED, gamma=4.0 / 3.0, noise=0.01, emf=EMF_UCT, recon=0, riemann=0,
          flat=True, rho=1e-2, p=1e-3, p_in=1.0, r_in=0.1, Bx=0.1, fofc=0):
    """Magnetised blast wave: 128^3 on [-0.5,0.5]^3 in 16^3 blocks by default.
    Ambient rho=1e-2, p=1e-3; p=1.0 for r<0.1; B=(0.1,0,0); +-1% noise."""
    nq = nb[0] * nb[1] * nb[2]
    grid = Grid(n, n, n, ngz, nq)
    N = [nb[d] * n for d in range(3)]
    spec = [dict(nb=nb, q0=0, bc=(BC_OUTFLOW,) * 6, idx=tuple(float(N[d]) for d in range(3)), x0=(-0.5,) * 3)]
    inst = make_instances(spec)
    state, B, aux = alloc(grid)
    lo = np.array([-0.5, -0.5, -0.5])
    hi = np.array([0.5, 0.5, 0.5])
    Bv = np.array([Bx, 0.0, 0.0])
    rc = lib().grmhd_setup_blast(grid, inst, 1, gamma, _p(lo), _p(hi), rho, p, p_in, r_in, _p(Bv), noise, seed,
                                 _p(state), _p(B[0]), _p(B[1]), _p(B[2]), _p(aux))
    assert rc == 0
    params = default_params(gamma=gamma, emf=emf, recon=recon, riemann=riemann, flat_metric=1 if flat else 0,
                            fofc=fofc)
    return Workload("blast", params, grid, spec, state, B, aux, zones=nq * n ** 3,
                    extra=dict(N=N, dx=1.0 / N[0]))


def blast_ks(nb=(2, 2, 2), n=16, ngz=3, spin=0.9, center=(8.0, 0.0, 0.0), size=4.0, seed=SEED, gamma=4.0 / 3.0,
             noise=0.01, rho=1e-2, p=1e-3, p_in=1.0, r_in=0.4, Bx=0.1, emf=EMF_UCT, fd_order=6):
    """The blast wave in the Kerr-Schild metric of a spin-a hole at the origin
    (Cowling, curved metric kernels): a cube of side `size` M centred at
    `center`, the blast at the cube's centre."""
    nq = nb[0] * nb[1] * nb[2]
    grid = Grid(n, n, n, ngz, nq)
    N = [nb[d] * n for d in range(3)]
    lo = np.array(center, dtype=np.float64) - 0.5 * size
    spec = [dict(nb=nb, q0=0, bc=(BC_OUTFLOW,) * 6, idx=tuple(N[d] / size for d in range(3)), x0=tuple(lo))]
    inst = make_instances(spec)
    state, B, aux = alloc(grid)
    hi = lo + size
    Bv = np.array([Bx, 0.0, 0.0])
    rc = lib().grmhd_setup_blast_ks(grid, inst, 1, gamma, _p(lo), _p(hi), spin, rho, p, p_in, r_in, _p(Bv), noise,
                                    seed, _p(state), _p(B[0]), _p(B[1]), _p(B[2]), _p(aux))
    assert rc == 0
    params = default_params(gamma=gamma, emf=emf, flat_metric=0, fd_order=fd_order)
    return Workload("blast_ks", params, grid, spec, state, B, aux, zones=nq * n ** 3,
                    extra=dict(N=N, dx=size / N[0]))


def fm_torus(nb=(4, 4, 4), n=16, ngz=3, extent=48.0, spin=0.9375, r_in=6.0, r_max=12.0, beta=100.0,
             rho_atm=1e-10, p_atm=1e-14, gamma=4.0 / 3.0, noise=0.01, seed=SEED, emf=EMF_UCT, excision_lapse=0.55,
             fofc=0):
    """Fishbone-Moncrief torus (r_in 6 M, pressure maximum 12 M, Gamma 4/3,
    single poloidal loop with max p / max p_mag = 100) around a spin-0.9375
    hole at the centre of the cube [-extent, extent]^3, Kerr-Schild metric
    (Cowling), atmosphere rho 1e-10; one uniform level of n^3 blocks.
    Cells near the ring singularity (lapse <= excision_lapse, all inside the
    horizon) are excised."""
    nq = nb[0] * nb[1] * nb[2]
    grid = Grid(n, n, n, ngz, nq)
    N = [nb[d] * n for d in range(3)]
    spec = [dict(nb=nb, q0=0, bc=(BC_OUTFLOW,) * 6, idx=tuple(N[d] / (2.0 * extent) for d in range(3)),
                 x0=(-extent,) * 3)]
    inst = make_instances(spec)
    state, B, aux = alloc(grid)
    lo = np.array([-extent] * 3, dtype=np.float64)
    info = np.zeros(4)
    rc = lib().grmhd_setup_fm_torus(grid, inst, 1, gamma, _p(lo), spin, r_in, r_max, beta, rho_atm, p_atm, noise,
                                    seed, _p(state), _p(B[0]), _p(B[1]), _p(B[2]), _p(aux), _p(info))
    assert rc == 0, rc
    params = default_params(gamma=gamma, emf=emf, flat_metric=0, fofc=fofc)
    params.rho_fl = rho_atm
    # atmosphere temperature p/rho: a zero floor would give the reset cells
    # eps = 0 and an entropy of -inf (ideal_gas_eos.hh:78-81)
    params.temp_fl = p_atm / rho_atm
    params.excision_lapse = excision_lapse
    return Workload("fm_torus", params, grid, spec, state, B, aux, zones=nq * n ** 3,
                    extra=dict(N=N, dx=2.0 * extent / N[0], ell=info[0], kappa=info[1], bscale=info[2],
                               pmag_max=info[3]))


def fm_torus_forest(nb=(6, 6, 6), n=16, ngz=4, extent=48.0, refined=None, spin=0.9375, r_in=6.0, r_max=12.0,
                    beta=100.0, rho_atm=1e-10, p_atm=1e-14, gamma=4.0 / 3.0, noise=0.01, seed=SEED, emf=EMF_UCT,
                    excision_lapse=0.55):
    """The Fishbone-Moncrief torus of fm_torus on a two-level forest: nb coarse
    blocks of n^3 cells on [-extent, extent]^3 (outflow faces), the coarse
    blocks in `refined` refined once.  Default: 6^3 blocks of 16^3 (96^3 coarse)
    with the central 4 x 4 x 2 coarse blocks -- the torus body -- refined, so
    the finest level has the spacing of a 192^3 grid."""
    if refined is None:
        lo_b = [(nb[0] - 4) // 2, (nb[1] - 4) // 2, (nb[2] - 2) // 2]
        refined = [(lo_b[0] + a) + nb[0] * ((lo_b[1] + b) + nb[1] * (lo_b[2] + e))
                   for e in range(2) for b in range(4) for a in range(4)]
    dx = 2.0 * extent / (nb[0] * n)
    forest = make_forest(nb, (BC_OUTFLOW,) * 6, (1.0 / dx,) * 3, refined, x0=(-0.5 * nb[0] * n * dx,) * 3)
    nq = nb[0] * nb[1] * nb[2] + 7 * forest.n_refined
    grid = Grid(n, n, n, ngz, nq)
    state, B, aux = alloc(grid)
    lo = np.array([-0.5 * nb[d] * n * dx for d in range(3)], dtype=np.float64)
    info = np.zeros(4)
    rc = lib().grmhd_setup_fm_torus_forest(grid, forest, gamma, _p(lo), spin, r_in, r_max, beta, rho_atm, p_atm,
                                           noise, seed, _p(state), _p(B[0]), _p(B[1]), _p(B[2]), _p(aux), _p(info))
    assert rc == 0, rc
    params = default_params(gamma=gamma, emf=emf, flat_metric=0)
    params.rho_fl = rho_atm
    params.temp_fl = p_atm / rho_atm
    params.excision_lapse = excision_lapse
    wl = Workload("fm_torus_forest", params, grid, [], state, B, aux, zones=nq * n ** 3,
                  extra=dict(dx=dx, ell=info[0], kappa=info[1], bscale=info[2], pmag_max=info[3],
                             n_fine=8 * forest.n_refined, equiv=nb[0] * n * 2))
    wl.forest = forest
    return wl


def fm_torus_point(x, y, z, spin=0.9375, gamma=4.0 / 3.0, r_in=6.0, r_max=12.0):
    """Torus primitives at one Kerr-Schild point: rho, p, z^i, u^t."""
    out = np.zeros(6)
    lib().grmhd_fm_torus_point(spin, gamma, r_in, r_max, x, y, z, _p(out))
    return out


def multi_blast(n_inst=64, sizes=(32, 48, 64, 80, 96), n=16, ngz=3, seed=SEED, gamma=4.0 / 3.0, noise=0.01):
    """Uneven instances: each a cubic blast-wave grid of side drawn from `sizes`,
    cut into n^3 blocks; all instances share one block list."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.array(sizes), size=n_inst)
    specs, q0 = [], 0
    for s in sides:
        b = int(s) // n
        specs.append(dict(nb=(b, b, b), q0=q0, bc=(BC_OUTFLOW,) * 6, idx=(float(s),) * 3, x0=(-0.5,) * 3))
        q0 += b ** 3
    grid = Grid(n, n, n, ngz, q0)
    inst = make_instances(specs)
    state, B, aux = alloc(grid)
    lo = np.array([-0.5, -0.5, -0.5])
    Bv = np.array([0.1, 0.0, 0.0])
    rc = lib().grmhd_setup_blast(grid, inst, len(specs), gamma, _p(lo), _p(-lo), 1e-2, 1e-3, 1.0, 0.1, _p(Bv), noise,
                                 seed, _p(state), _p(B[0]), _p(B[1]), _p(B[2]), _p(aux))
    assert rc == 0
    params = default_params(gamma=gamma, flat_metric=1)
    zones = int(sum(int(s) ** 3 for s in sides))
    return Workload("multi_blast", params, grid, specs, state, B, aux, zones=zones, extra=dict(sides=sides))


def balsara1(N=400, ngz=3, gamma=2.0):
    """Balsara-1 relativistic MHD shock tube: N x 1 x 1 on x in [0,1], jump at x=0.5.
    L: rho=1, p=1, B=(0.5,1,0); R: rho=0.125, p=0.1, B=(0.5,-1,0).  GS CT, WENO-Z, HLLE."""
    grid = Grid(N, 1, 1, ngz, 1)
    dx = 1.0 / N
    spec = [dict(nb=(1, 1, 1), q0=0, bc=(BC_OUTFLOW, BC_OUTFLOW, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC),
                 idx=(1.0 / dx,) * 3)]
    inst = make_instances(spec)
    state, B, aux = alloc(grid)
    L = np.array([1.0, 1.0, 0.5, 1.0, 0.0])
    R = np.array([0.125, 0.1, 0.5, -1.0, 0.0])
    rc = lib().grmhd_setup_shocktube(grid, inst, 1, gamma, 0.0, 0.5, _p(L), _p(R), _p(state), _p(B[0]), _p(B[1]),
                                     _p(B[2]), _p(aux))
    assert rc == 0
    params = default_params(gamma=gamma, emf=EMF_GS, flat_metric=1)
    return Workload("balsara1", params, grid, spec, state, B, aux, zones=N, extra=dict(dx=dx))


def c2p_states(n, seed=SEED, gamma=4.0 / 3.0, spin=0.9, rho=(1e-12, 1e-2), eps=(1e-6, 1.0), W=(1.0, 10.0),
               sigma=(1e-4, 1e2), r=(1.5, 50.0)):
    """Seeded c2p states in Kerr-Schild (spin a) at r in [1.5,50] M (SoA arrays)."""
    prims = np.zeros((13, n))
    met = np.zeros((10, n))
    cons = np.zeros((10, n))
    rad = np.zeros(n)
    rr = [np.array(x, dtype=np.float64) for x in (rho, eps, W, sigma, r)]
    rc = lib().grmhd_setup_c2p_states(n, seed, gamma, spin, *[_p(x) for x in rr], _p(prims), _p(met), _p(cons), _p(rad))
    assert rc == 0
    return dict(prims=prims, met=met, cons=cons, rad=rad, params=default_params(gamma=gamma))

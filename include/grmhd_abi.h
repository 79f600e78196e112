/*
 * grmhd_abi.h -- C ABI of the MI355X GRMHD RHS + con2prim library
 * (libgrmhd_mi355x.so).
 *
 * The reference (GRACE) has no FFI layer: its hot path is a set of C++
 * templates over the EOS type.  Each entry point below replaces one of those
 * calls (reference file:line given per function) and keeps its argument
 * meaning, its data layout and its error behaviour, with plain pointers and
 * sizes instead of Kokkos views.
 *
 * Data layout (identical to the reference's Kokkos LayoutLeft views,
 * include/grace/data_structures/variable_properties.hh:170-305):
 *   cell arrays   a[q][v][k][j][i]  i fastest, extents Ni=nx+2ngz, Nj, Nk
 *   face arrays   Bx[q][k][j][i] with extent Ni+1 along x (By: Nj+1, Bz: Nk+1)
 *   state vars    0..6  DENS SX SY SZ TAU YESTAR ENTROPYSTAR (densitised)
 *                 7..22 GXX GXY GXZ GYY GYZ GZZ ALP BETAX BETAY BETAZ
 *                       KXX KXY KXZ KYY KYZ KZZ   (Cowling metric)
 *   aux vars      RHO ZVECX ZVECY ZVECZ BX BY BZ YE TEMP ENTROPY EPS PRESS
 *                 BDIV C2P_DENS_ERR C2P_ERR  (variable_indices.hh:162-331)
 * All device pointers are caller-owned hipMalloc memory; the library never
 * frees them.  Every call is asynchronous on the given HIP stream and returns
 * a status code; nothing aborts across the ABI.  Per-cell c2p outcomes are
 * bit sets identical to the reference's c2p_err_enum_t
 * (include/grace/physics/eos/c2p.hh:95-125).
 */
#ifndef GRMHD_ABI_H
#define GRMHD_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* grmhd_stream_t; /* hipStream_t, 0 = default stream */
typedef struct grmhd_ctx grmhd_ctx;

typedef enum {
    GRMHD_OK = 0,
    GRMHD_ERR_INVALID = 1,      /* bad argument / size / layout */
    GRMHD_ERR_HIP = 2,          /* HIP runtime error (launch, memory) */
    GRMHD_ERR_ALLOC = 3,        /* device allocation failed */
    GRMHD_ERR_UNSUPPORTED = 4,  /* option combination not built */
    GRMHD_ERR_C2P_FATAL = 5     /* a cell hit a branch where the reference aborts */
} grmhd_status;

enum { GRMHD_EOS_IDEAL_GAS = 0, GRMHD_EOS_HYBRID_PWPOLY = 1 };
enum { GRMHD_RECON_WENOZ = 0, GRMHD_RECON_MC2 = 1, GRMHD_RECON_MINMOD = 2, GRMHD_RECON_DONOR = 3 };
enum { GRMHD_RIEMANN_HLLE = 0, GRMHD_RIEMANN_LLF = 1 };
enum { GRMHD_EMF_UCT = 0, GRMHD_EMF_GS = 1 };
enum { GRMHD_THERMO_TEMP = 0, GRMHD_THERMO_PRESS = 1 };
enum { GRMHD_BC_OUTFLOW = 0, GRMHD_BC_PERIODIC = 1 };
/* variables per block of the state and aux arrays (layout above) */
enum { GRMHD_NSTATE = 23, GRMHD_NAUX = 15, GRMHD_NHRSC = 7 };

/* Runtime + compile-time options of the reference's hot path.  Defaults
 * (grmhd_default_params) are the reference's parameter-schema defaults,
 * parameters/{grmhd,eos}.yaml; reference reader: grmhd_helpers.hh:165-292. */
typedef struct {
    /* EOS (src/physics/eos/eos_storage.cpp:160-220) */
    int32_t eos_kind;
    double gamma;             /* ideal gas Gamma */
    double eps_max;           /* eos.eps_maximum */
    double entropy_min;       /* eos.entropy_minimum (hybrid) */
    double gamma_th;          /* hybrid thermal Gamma */
    int32_t pwp_n;            /* hybrid: number of polytropic pieces (<= 8) */
    double pwp_kappa0;
    double pwp_gamma[8];
    double pwp_rho[8];        /* n-1 dividing densities */
    /* atmosphere (grmhd_helpers.hh:41-48) */
    double rho_fl, temp_fl, ye_fl, atmo_tol;
    /* c2p (grmhd_helpers.hh:52-59) */
    double c2p_tol, max_lorentz, max_sigma, beta_fallback;
    int32_t use_entropy_backup;
    double bh_alp_thresh;
    /* excision (grmhd_helpers.hh:89-96).  excise_by_radius = 1 resets cells
     * whose radius is <= excision_radius (c2p.cpp:345-348); the radius of a
     * cell comes from its centre x = x0 + (global index + 1/2) / idx of its
     * instance (or forest): |x|, or with excision_cks = 1 the Kerr-Schild
     * radius of a hole of spin excision_spin at the origin
     * (cart_to_sph, cartesian_coordinate_systems.hh:87-107).  Not accepted
     * with periodic faces, where the reference evaluates a ghost cell at its
     * unwrapped position while this library copies the periodic image. */
    int32_t excise_by_radius;
    double excision_radius, excision_lapse;
    int32_t excision_cks;
    double excision_spin;
    /* numerical scheme (CMakeLists.txt:206-256, cmake/code_modules.cmake:14-40) */
    int32_t recon, riemann, emf, thermo_recon, fd_order;
    double rusanov_c_limit;
    /* 1: the metric is Minkowski everywhere; kernels skip the 16 metric
     * arrays and the source terms (identically zero there). */
    int32_t flat_metric;
    /* first-order flux correction (evolve.cpp:456-811; grmhd.fofc.dmp_enable,
     * dmp_M, parameters/grmhd.yaml:162-174).  Needs the GS EMF scheme and
     * ngz >= 4, as in the reference. */
    int32_t fofc;
    int32_t fofc_dmp_enable;
    double fofc_dmp_M;
    /* 0 (default, as the reference): the GS edge EMFs are computed on the
     * interior edge range only, while the FOFC flag pass also advances face B
     * for the first ghost layer of cells (evolve.cpp:492-525 against
     * :1090-1104) and so reads edges no pass has written; the ghost-layer
     * cells' tentative states then flag spuriously and the two blocks beside
     * a shared face can tag it differently, which breaks conservation.
     * 1: compute the GS edge EMFs one ghost layer wider when FOFC is on, so
     * the ghost-layer flags equal the neighbour's interior flags. */
    int32_t fofc_ghost_emf;
    /* 1 (default): independent launches of an RK3 stage run side by side on
     * two library-owned streams (the three flux directions; ghost faces and
     * the c2p of ghosts that run their own next to the interior c2p), joined
     * back into the caller's stream by events (also inside the captured
     * graph); results are bit-identical to 0, which issues every launch on
     * the caller's stream (for kernel traces with one launch at a time).
     * FOFC runs use one stream either way. */
    int32_t stage_streams;
} grmhd_params;

/* Block grid: every block has the same interior extents. */
typedef struct { int32_t nx, ny, nz, ngz; int64_t nq; } grmhd_grid;

/* A uniform Cartesian instance cut into nb[0]*nb[1]*nb[2] blocks, stored at
 * block indices q = q0 + bi + nb0*(bj + nb1*bk).  bc[] per face
 * (-x,+x,-y,+y,-z,+z); periodicity is per dimension, as the reference's
 * periodic_x|y|z (connectivity.cpp:52-69): both faces of a dimension are
 * periodic or neither is (GRMHD_ERR_INVALID otherwise).  idx = inverse cell spacings.  x0 = coordinates of
 * the lower corner of the instance's interior (the reference's per-block
 * origin qx, cartesian_coordinate_systems.hh:56-66), used only for cell
 * radii (excise_by_radius).  With a curved metric the spacing must be equal
 * in x, y and z: the metric derivatives of the source terms use idx[0] in
 * every direction, as the reference (grmhd.hh:231-235) on its cubic cells. */
typedef struct { int32_t nb[3]; int64_t q0; int32_t bc[6]; double idx[3]; double x0[3]; } grmhd_instance;

/* Per-launch c2p statistics (reduced on device, read with grmhd_c2p_stats). */
typedef struct {
    int64_t n_cells, n_fatal, n_atmo_reset, n_ent_backup, n_guess_used;
    int64_t n_root_evals, n_bracket_evals, max_root_evals;
} grmhd_c2p_stats_t;

void grmhd_default_params(grmhd_params* p);

/* Context: validated parameters + grid + device scratch (fluxes, vbar/Eface,
 * edge EMFs, E_center, RK stage buffers).  Replaces the parameter readers and
 * the variable_list singleton of the reference. */
grmhd_status grmhd_ctx_create(const grmhd_params* p, const grmhd_grid* grid,
                              const grmhd_instance* inst, int32_t n_inst, grmhd_ctx** out);
/* Two-level block forest (the p4est forest of include/grace/amr/forest.hh
 * restricted to one refinement level): nb[0]*nb[1]*nb[2] coarse blocks of
 * spacing 1/idx, of which the n_refined listed (coarse ids
 * bi + nb0*(bj + nb1*bk), any order, no repeats) are refined once into 8
 * blocks of half the spacing.  bc[] per domain face as for an instance.
 * Block order q: the coarse leaves in ascending coarse id, then the 8
 * children (p4est Morton id cx + 2cy + 4cz) of each refined block, refined
 * blocks in ascending coarse id; grmhd_forest_block reports where each lies.
 * grid.nq must be nb0*nb1*nb2 + 7*n_refined; blocks cubic with even extent,
 * ngz even and idx equal in x, y, z (grace_initialize.cpp:243-248; the
 * divergence-preserving prolongation is written for cubic cells).
 * On such a context the ghost fill also prolongs coarse data into fine ghosts
 * and restricts fine data into coarse ghosts (minmod prolongation and 2x2x2
 * averages for the hydro conserved, prolongation_kernels.hh:59-106 /
 * pr_helpers.hh:71-212; Toth-Roe divergence-preserving prolongation and area
 * averages for face B, prolongation_kernels.hh:111-386 /
 * restrict_kernels.hh:125-164), and every substep refluxes: the coarse flux
 * of each coarse-fine face becomes the average of the four fine fluxes and
 * each coarse edge EMF touched by a fine cell the average of the two fine
 * EMFs along it (refluxing.cpp:547-806, 959-1173), after the FOFC
 * correction when FOFC is on (evolve.cpp:1685-1697).  The Cowling metric of
 * the ghost cells is the caller's (initial data), as on uniform grids. */
typedef struct {
    int32_t nb[3];
    int32_t bc[6];
    double idx[3];            /* inverse coarse cell spacings */
    int32_t n_refined;
    const int32_t* refined;   /* host array of coarse ids */
    double x0[3];             /* lower corner of the domain (cell radii only) */
} grmhd_forest;

grmhd_status grmhd_ctx_create_forest(const grmhd_params* p, const grmhd_grid* grid, const grmhd_forest* forest,
                                     grmhd_ctx** out);
/* Level (0 coarse, 1 fine) and block coordinates at that level of block q. */
grmhd_status grmhd_forest_block(const grmhd_ctx* ctx, int64_t q, int32_t* level, int32_t coords[3]);

void grmhd_ctx_destroy(grmhd_ctx* ctx);
const char* grmhd_last_error(void);

/* advance_substep<eos_t> (include/grace/evolution/evolve.hh:252-256,
 * src/evolution/evolve.cpp:1671-1717): new_state += dt*dtfact*L(old_state)
 * on interior cells, new_B += dt*dtfact*curl(E) on interior faces.  aux holds
 * the primitives of old_state.  dt is read from device memory (dt_dev). */
grmhd_status grmhd_advance_substep(grmhd_ctx* ctx, const double* dt_dev, double dtfact,
                                   double* new_state, const double* old_state,
                                   double* const new_B[3], const double* const old_B[3],
                                   const double* aux, grmhd_stream_t stream);

/* advance_substep<eos_t>(t, dt, dtfact, ...) with the reference's host
 * arguments (include/grace/evolution/evolve.hh:252-256): dt is a host value,
 * placed in a context scalar in stream order ahead of the substep; t is
 * accepted for the signature (the Cowling GRMHD substep does not depend on
 * it).  Otherwise as grmhd_advance_substep. */
grmhd_status grmhd_advance_substep_t(grmhd_ctx* ctx, double t, double dt, double dtfact,
                                     double* new_state, const double* old_state,
                                     double* const new_B[3], const double* const old_B[3],
                                     const double* aux, grmhd_stream_t stream);

/* Sync point: waits for the stream, then reports GRMHD_ERR_C2P_FATAL if any
 * cell of a c2p pass (grmhd_compute_auxiliaries, the RK3 step, the c2p
 * batches) took a branch where the reference calls Kokkos::abort
 * (c2p.hh:160-322: rho, eps, T or p above the EOS range outside the lenient
 * region) since the last sync point; the count is cleared.  The kernels
 * still write every cell (the caller decides), so an asynchronous step never
 * stops the stream. */
grmhd_status grmhd_sync(grmhd_ctx* ctx, grmhd_stream_t stream);

/* amr::apply_boundary_conditions (src/amr/boundary_conditions.cpp:55-81):
 * same-level ghost copy between blocks + outflow/periodic physical BC, for the
 * 7 hydro conserved variables and the 3 face-B arrays. */
grmhd_status grmhd_fill_ghosts(grmhd_ctx* ctx, double* state, double* const B[3],
                               grmhd_stream_t stream);

/* The ghost exchange of the Cowling metric (state vars GXX_..KZZ_), which the
 * reference runs with the evolved variables (they are registered
 * "fourth_order" with bc "none", variable_indices.cpp:442-446): copies from
 * same-level neighbours and periodic images; on a forest, 5-point Lagrange
 * restriction into coarse ghosts and prolongation into fine ghosts
 * (lagrange_restrict_op<4> / lagrange_prolong_op<4>, pr_helpers.hh:105-330,
 * with the reference's stencil choice and pair-symmetric sums); ghosts beyond
 * an outflow face keep their values.  The metric is static, so one call after
 * the initial data (before grmhd_set_static_metric) gives the ghost values the
 * reference's exchange settles on. */
grmhd_status grmhd_fill_metric_ghosts(grmhd_ctx* ctx, double* state, grmhd_stream_t stream);

/* compute_auxiliary_quantities<eos_t> (include/grace/evolution/auxiliaries.hh:49-54,
 * src/evolution/auxiliaries.cpp:81-180): c2p on every cell incl. ghosts,
 * conserved rewrite on reset bits, sticky-OR of C2P_ERR, BDIV, and (GS)
 * the cell-centred E field kept in the context. */
grmhd_status grmhd_compute_auxiliaries(grmhd_ctx* ctx, double* state, const double* const B[3],
                                       double* aux, grmhd_stream_t stream);

/* conservs_to_prims<eos_t> over a flat batch (include/grace/physics/eos/c2p.hh:342-353).
 * cons_soa[10*n]: DENS S_x S_y S_z TAU YE* S* B^x B^y B^z (densitised, SoA
 *   blocks of n); metric_soa[10*n]: gxx gxy gxz gyy gyz gzz alp betax betay betaz.
 * Outputs: prims_soa[13*n] (RHO PRESS ZX ZY ZZ YE TEMP EPS ENT BX BY BZ unused),
 *   cons_soa overwritten with the re-densitised conserved, err_bits[n]
 *   (c2p_err_t words), floored[n] (bool as uint8).  r_or_null: radii for
 *   radius-based excision.
 * guess_or_null (4*n, SoA): brackets [e_lo, e_hi] of the energy root and
 *   [s_lo, s_hi] of the entropy-backup root, in units of 1/h0 (mu h0; the
 *   master functions of kastaun_c2p.hh:121-301 and ent_based_c2p.hh:44-160
 *   have one root on (0, 1/h0]).  A bracket that changes sign replaces the
 *   Newton solve for mu+ and the default bracket [0, mu+]
 *   (kastaun_c2p.hh:355-372); any other bracket leaves the plain path as is.
 * roots_or_null (4*n, SoA): roots x_e = mu_e h0, x_s = mu_s h0, then the
 *   default bracket ends mu+ h0 of each solve; -1 where not computed.
 * Returns GRMHD_OK; per-state fatal branches are reported in stats. */
grmhd_status grmhd_c2p_batch(grmhd_ctx* ctx, int64_t n, double* cons_soa, const double* metric_soa,
                             double* prims_soa, uint64_t* err_bits, uint8_t* floored, int32_t dry_run,
                             const double* r_or_null, const double* guess_or_null, double* roots_or_null,
                             grmhd_stream_t stream);

/* grmhd_c2p_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as grmhd_c2p_batch, whose statistics records
 * grow a context buffer with the batch; here they live in a static device
 * array of the library sized for at most GRMHD_NOALLOC_MAX_STATES states
 * (4 MiB per GPU, held until the process ends).  Larger n returns
 * GRMHD_ERR_INVALID.  All calls on one GPU share that array: calls that
 * overlap on one GPU (different streams, different contexts) race, and
 * grmhd_c2p_stats then reports an undefined mix; serialise them. */
#define GRMHD_NOALLOC_MAX_STATES (1 << 24)
grmhd_status grmhd_c2p_batch_noalloc(grmhd_ctx* ctx, int64_t n, double* cons_soa, const double* metric_soa,
                                     double* prims_soa, uint64_t* err_bits, uint8_t* floored, int32_t dry_run,
                                     const double* r_or_null, const double* guess_or_null, double* roots_or_null,
                                     grmhd_stream_t stream);

/* prims_to_conservs (src/physics/eos/c2p.cpp:416-455) over a flat batch. */
grmhd_status grmhd_prims_to_conservs_batch(grmhd_ctx* ctx, int64_t n, const double* prims_soa,
                                           const double* metric_soa, double* cons_soa,
                                           grmhd_stream_t stream);

/* Statistics of the last c2p launch (host copy; synchronises the stream). */
grmhd_status grmhd_c2p_stats(grmhd_ctx* ctx, grmhd_c2p_stats_t* out, grmhd_stream_t stream);

/* Per-state evaluation counts of the batch c2p.  No counterpart in the
 * reference (its c2p reports no iteration counts): it gives a c2p benchmark
 * the distribution of the work per state (mean and percentiles) that the
 * block-reduced grmhd_c2p_stats cannot.  With a device array set, every
 * grmhd_c2p_batch / grmhd_c2p_batch_learned call of n states writes
 * min(255, master-function + bracket evaluations of state s) into
 * counts[s] (n bytes); null, the default, turns the output off.  The array
 * is a kernel argument: a captured graph keeps the setting it was captured
 * with. */
grmhd_status grmhd_set_c2p_eval_counts(grmhd_ctx* ctx, uint8_t* counts_or_null);

/* find_stable_timestep (src/evolution/find_stable_timestep.cpp:72-126):
 * dt = min over interior cells of cfl*dx/c_max, written to dt_dev. */
grmhd_status grmhd_compute_dt(grmhd_ctx* ctx, const double* state, const double* aux, double cfl,
                              double* dt_dev, grmhd_stream_t stream);

/* find_stable_timestep (src/evolution/find_stable_timestep.cpp:72-126) with
 * the reference's host result: dt (the reference's set_timestep value) into
 * *dt_out after the stream completes.  A sync point, as grmhd_sync. */
grmhd_status grmhd_find_stable_timestep(grmhd_ctx* ctx, const double* state, const double* aux, double cfl,
                                        double* dt_out, grmhd_stream_t stream);

/* One full SSP-RK3 step (evolve_impl, src/evolution/evolve.cpp:147-225):
 * reset of the per-step c2p diagnostics, three substeps each followed by the
 * ghost fill and the c2p pass.  state/B/aux are updated in place; the two
 * stage buffers live in the context.  The kernels of the step are fused
 * (linear combination + sources + flux divergence + CT in one pass).
 * With GS EMFs the first substep reads the cell-centred E the context's last
 * c2p pass left (grmhd_compute_auxiliaries or the previous step), as the
 * reference reads its E_center field: evolve one state per context. */
grmhd_status grmhd_rk3_step(grmhd_ctx* ctx, const double* dt_dev, double* state, double* const B[3],
                            double* aux, grmhd_stream_t stream);

/* Same step captured once into a hipGraph and replayed (pointers are bound
 * at the first call; a call with different pointers re-captures). */
grmhd_status grmhd_rk3_step_graph(grmhd_ctx* ctx, const double* dt_dev, double* state,
                                  double* const B[3], double* aux, grmhd_stream_t stream);

/* Cowling: the metric in `state` is static, so the 30 finite-difference
 * metric derivatives of the geometric source terms (grmhd.hh:232-235,
 * recomputed by the reference in every substep) and the interpolated face
 * metrics of the flux kernels (grmhd_helpers.hh:360-419, every face of every
 * substep in the reference) can be taken once: enable = 1 computes them for
 * `state` into context buffers (30 doubles per interior cell, 13 per y and
 * z face), and the update and flux kernels of calls whose metric
 * source is that same array then read them instead (results
 * bit-identical).  The cache is not refreshed automatically: call again
 * after changing the metric in that array, or enable = 0 to drop it.
 * No effect for flat_metric.  Invalidates a captured RK3 graph. */
grmhd_status grmhd_set_static_metric(grmhd_ctx* ctx, const double* state, int32_t enable, grmhd_stream_t stream);

/* Workgroup sizes (threads) of the kernel families, chosen by
 * grmhd_autotune; every size is one of 64, 128, 256. */
typedef struct {
    int32_t flux[3];        /* face-flux kernels, x/y/z */
    int32_t emf;            /* edge EMFs */
    int32_t update;         /* fused RK update + CT */
    int32_t gather_faces;   /* ghost gather of face B */
    int32_t gather_cells;   /* ghost gather of cell variables */
    int32_t c2p;            /* c2p over cells */
    int32_t flux_prefetch[3]; /* 1: face-flux kernel x/y/z loads its whole WENO-Z
                               * stencil before reconstructing (more loads in
                               * flight, fewer waves); 0 or 1, same results */
} grmhd_tuning;

/* Times each kernel family of one RK3 substep on the context's own block
 * list with every candidate workgroup size (the inputs are read, never
 * modified: updates, gathers and c2p run on internal copies) and keeps the
 * fastest; tuning_out_or_null receives the choice.  Invalidates any
 * captured RK3 graph.  The reference has no counterpart (its Kokkos launch
 * shapes are fixed at compile time). */
grmhd_status grmhd_autotune(grmhd_ctx* ctx, const double* state, const double* const B[3], const double* aux,
                            int32_t reps, grmhd_tuning* tuning_out_or_null, grmhd_stream_t stream);
grmhd_status grmhd_get_tuning(grmhd_ctx* ctx, grmhd_tuning* out);
grmhd_status grmhd_set_tuning(grmhd_ctx* ctx, const grmhd_tuning* in);

/* Total duration (ms) and number of launches of a kernel family since the
 * last query, measured with HIP events recorded on the launch stream around
 * each launch ("flux_x", "flux_y", "flux_z", "emf", "update", "ghost", "c2p",
 * "step").  Enable with grmhd_set_timing(ctx, 1); graph replays record only
 * "step". */
grmhd_status grmhd_set_timing(grmhd_ctx* ctx, int32_t on);
grmhd_status grmhd_kernel_time_ms(grmhd_ctx* ctx, const char* family, double* ms_total, int64_t* n_launch);

#ifdef __cplusplus
}
#endif
#endif /* GRMHD_ABI_H */

/*
 * grmhd_setup.h -- host-side input generators exported by libgrmhd_mi355x.so.
 *
 * Not part of the reference's hot-path interface: these build the synthetic
 * inputs of the benchmark configurations (blast wave, shock tube, Kerr-Schild
 * c2p states) on host arrays laid out as in grmhd_abi.h.  They follow the
 * reference's initial-data path (src/physics/grmhd.cpp:359-527, 676-759) and
 * its shock-tube data (include/grace/physics/id/shocktube.hh:46-100).
 */
#ifndef GRMHD_SETUP_H
#define GRMHD_SETUP_H
#include <stdint.h>
#include "grmhd_abi.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Magnetised blast wave on every block of every instance (cells incl. ghosts).
 * lo: lower domain corner; p = p_in for r < r_in, rho/p times (1 + noise*U[-1,1]). */
int grmhd_setup_blast(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                      const double lo[3], const double hi[3], double rho_amb, double p_amb, double p_in,
                      double r_in, const double Bvec[3], double noise, uint64_t seed, double* state,
                      double* Bx, double* By, double* Bz, double* aux);

/* Quasi-1D shock tube along x: L/R = {rho, p, Bx, By, Bz}, x <= x_jump is left. */
int grmhd_setup_shocktube(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                          double x0, double x_jump, const double L[5], const double R[5], double* state,
                          double* Bx, double* By, double* Bz, double* aux);

/* Kerr-Schild (M=1, spin a) lapse, shift and 3-metric at a Cartesian point. */
void grmhd_kerr_schild(double a, double x, double y, double z, double g[6], double* alp, double beta[3]);
/* The same plus the extrinsic curvature of the stationary slice (K_ij). */
void grmhd_kerr_schild_adm(double a, double x, double y, double z, double g[6], double* alp, double beta[3],
                           double K[6]);
/* grmhd_setup_blast centred in [lo, hi] with the Kerr-Schild metric (hole of
 * spin a at the origin) in every cell; face B = Bvec (densitised, so
 * discretely divergence free). */
int grmhd_setup_blast_ks(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                         const double lo[3], const double hi[3], double spin, double rho_amb, double p_amb,
                         double p_in, double r_in, const double Bvec[3], double noise, uint64_t seed, double* state,
                         double* Bx, double* By, double* Bz, double* aux);

/* Fishbone-Moncrief torus (inner edge r_in, pressure maximum r_max, rho = 1
 * there) around a spin-a hole at the origin, Kerr-Schild metric and K_ij in
 * every cell, atmosphere rho_atm/p_atm outside, p times (1 + noise U[-1,1]);
 * single poloidal loop from A_phi = max(rho - 0.2, 0) by the discrete curl
 * (divergence free), scaled to max p / max(b^2/2) = beta.
 * info (or null): l = u^t u_phi, kappa (p = kappa rho^gamma), B scale,
 * max b^2/2.  Returns 0, or nonzero for bad parameters / blocks. */
int grmhd_setup_fm_torus(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                         const double lo[3], double spin, double r_in, double r_max, double beta, double rho_atm,
                         double p_atm, double noise, uint64_t seed, double* state, double* Bx, double* By,
                         double* Bz, double* aux, double info[4]);
/* The same torus on a two-level forest (block order of
 * grmhd_ctx_create_forest), the box starting at lo; noise keyed by level and
 * cell index on the level. */
int grmhd_setup_fm_torus_forest(const grmhd_grid* g, const grmhd_forest* f, double gamma, const double lo[3],
                                double spin, double r_in, double r_max, double beta, double rho_atm, double p_atm,
                                double noise, uint64_t seed, double* state, double* Bx, double* By, double* Bz,
                                double* aux, double info[4]);
/* The torus primitives at one Cartesian Kerr-Schild point (no atmosphere):
 * out = {rho, p, z^x, z^y, z^z, u^t}. */
void grmhd_fm_torus_point(double spin, double gamma, double r_in, double r_max, double x, double y, double z,
                          double out[6]);

/* Flat-space data on a two-level forest (block order of
 * grmhd_ctx_create_forest; the forest's nb, idx and refined list) covering
 * the box from lo: rho = rho0 (1 + amp sin(2 pi x/Lx) sin(2 pi y/Ly)),
 * p = p0, or p_in inside the sphere of radius r_in around center, z = W v
 * = zvel, face B = Bvec plus the discrete curl of
 * A_z = b_loop Lx/(2 pi) sin(2 pi x/Lx) sin(2 pi y/Ly) sampled at the edges
 * (divergence free in every block, and a coarse face holds the average of
 * the four fine faces covering it).  Ghost cells included.  Ideal gas. */
int grmhd_setup_forest_wave(const grmhd_grid* g, const grmhd_forest* f, double gamma, const double lo[3],
                            double rho0, double amp, double p0, double p_in, double r_in, const double center[3],
                            const double zvel[3], const double Bvec[3], double b_loop, double* state, double* Bx,
                            double* By, double* Bz, double* aux);

/* Seeded c2p test states (SoA outputs, see grmhd_abi.h for the layouts). */
int grmhd_setup_c2p_states(int64_t n, uint64_t seed, double gamma, double spin, const double rho_rng[2],
                           const double eps_rng[2], const double W_rng[2], const double sig_rng[2],
                           const double r_rng[2], double* prims, double* met, double* cons, double* rad);

#ifdef __cplusplus
}
#endif
#endif

/* grmhd_mlp.h -- learned Kastaun guess: a 10 -> 64 -> 64 -> 2 MLP (ReLU,
 * BF16 MFMA, FP32 accumulation) that predicts the roots of the energy and
 * entropy-backup master functions, so that the FP64 solves start from them
 * (Newton finish, ideal gas) or from a narrow bracket around them, instead
 * of the Newton solve for mu+ and Brent on [0, mu+].
 *
 * The reference has no counterpart: the network only changes where the
 * FP64 root solves of kastaun_c2p.hh:355-372 and ent_based_c2p.hh:214-229
 * start.  Every primitive still comes from the FP64 Kastaun master
 * function, and a start that does not converge (or a bracket that misses
 * the root) falls back to the plain path, so results agree with
 * grmhd_c2p_batch to within the root tolerances (1e-13 relative for the
 * Newton finish, the plain solve's own for brackets).
 *
 * Parameters: FP32, GRMHD_MLP_NPARAM floats, row-major W[k][j]
 * (k input, j output):
 *   W1 [16][64] @0 (rows 10..15 unused, kept zero), b1 [64] @1024,
 *   W2 [64][64] @1088, b2 [64] @5184, W3 [64][2] @5248, b3 [2] @5376.
 * Scaling (GRMHD_MLP_NSCALE floats): feature mean[10], feature 1/std[10],
 *   output mean[2], output std[2].
 * Features (10 per state, SoA [10][n]) of the densitised conserved state:
 *   log q, log r^2, log B~^2, (r.B~)^2/(r^2 B~^2), log x_hi, log x_lo,
 *   log(1+q), r^2/(1+q)^2, B~^2/(1+q), log D
 *   (q, r, B~ as in kastaun_c2p.hh:306-338; x_lo, x_hi the closed-form
 *   bracket of the mu+ solve, kastaun_c2p.hh:101-117, times h0).
 * Outputs (SoA [2][n], de-normalised): log(x_e / x_hi) and log(x_s / x_e),
 *   x_e and x_s the energy and entropy roots times h0.
 */
#ifndef GRMHD_MLP_H
#define GRMHD_MLP_H

#include "grmhd_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GRMHD_MLP_NFEAT 10
#define GRMHD_MLP_NOUT 2
#define GRMHD_MLP_NPARAM 5378
#define GRMHD_MLP_NSCALE 24

typedef struct grmhd_mlp grmhd_mlp;

/* Glorot-uniform weights and zero biases from a SplitMix64 stream. */
grmhd_status grmhd_mlp_init_params(uint64_t seed, float* params_host);

/* Network on the current device, initialised with grmhd_mlp_init_params(seed),
 * identity scaling, brackets delta_e = 0.02, delta_s = 1e-6, nexp = 2. */
grmhd_status grmhd_mlp_create(uint64_t seed, grmhd_mlp** out);
void grmhd_mlp_destroy(grmhd_mlp* mlp);

/* Host FP32 parameters in/out; setting them resets the Adam state. */
grmhd_status grmhd_mlp_set_params(grmhd_mlp* mlp, const float* params_host, grmhd_stream_t stream);
grmhd_status grmhd_mlp_get_params(grmhd_mlp* mlp, float* params_host, grmhd_stream_t stream);
grmhd_status grmhd_mlp_set_scaling(grmhd_mlp* mlp, const float* scale_host, grmhd_stream_t stream);
grmhd_status grmhd_mlp_get_scaling(grmhd_mlp* mlp, float* scale_host);

/* Bracket half-widths (natural-log units) around the predicted energy root
 * and the predicted entropy/energy root ratio, and the number of times a
 * bracket that misses the root is moved towards it before the plain path
 * runs (0..8). */
grmhd_status grmhd_mlp_set_bracket(grmhd_mlp* mlp, double delta_e, double delta_s, int32_t nexp);
grmhd_status grmhd_mlp_get_bracket(grmhd_mlp* mlp, double* delta_e, double* delta_s, int32_t* nexp);

/* Newton finish (ideal gas): the FP64 solve first runs a safeguarded Newton
 * iteration on the Kastaun master function (analytic slope) from the
 * predicted root, at most max_evals evaluations, and keeps its root when the
 * residual divided by the slope is within 1e-12 of it; otherwise the
 * bracket above, then the plain path, run.  0 disables it; default 8.
 * Other EOS use the brackets only. */
grmhd_status grmhd_mlp_set_newton(grmhd_mlp* mlp, int32_t max_evals);
grmhd_status grmhd_mlp_get_newton(grmhd_mlp* mlp, int32_t* max_evals);

/* Device features of a batch of densitised conserved states (SoA as in
 * grmhd_c2p_batch); the EOS is the context's. */
grmhd_status grmhd_mlp_features(grmhd_ctx* ctx, int64_t n, const double* cons_soa, const double* metric_soa,
                                float* feat_soa, grmhd_stream_t stream);

/* Training targets from the roots output of grmhd_c2p_batch and the
 * features: [log(x_e) - feat4, log(x_s/x_e)], NaN where a root is absent. */
grmhd_status grmhd_mlp_targets(int64_t n, const double* roots_soa, const float* feat_soa, float* target_soa,
                               grmhd_stream_t stream);

/* Forward pass (device features in, device outputs out). */
grmhd_status grmhd_mlp_forward(grmhd_mlp* mlp, int64_t n, const float* feat_soa, float* out_soa,
                               grmhd_stream_t stream);

/* Gradient of loss = (1/n) sum_s sum_j (y_sj - t_sj)^2 (normalised units,
 * NaN targets skipped) w.r.t. the parameters: grad_out[GRMHD_MLP_NPARAM + 1]
 * on the device, the last entry the loss. */
grmhd_status grmhd_mlp_grad(grmhd_mlp* mlp, int64_t n, const float* feat_soa, const float* target_soa,
                            float* grad_out, grmhd_stream_t stream);

/* One Adam step (beta 0.9/0.999, eps 1e-8) on a minibatch; loss_out_or_null
 * (device float) receives the minibatch loss before the step. */
grmhd_status grmhd_mlp_train_step(grmhd_mlp* mlp, int64_t n, const float* feat_soa, const float* target_soa,
                                  float lr, float* loss_out_or_null, grmhd_stream_t stream);

typedef struct {
    int32_t epochs;          /* passes over the training split */
    int32_t batch;           /* minibatch states */
    float lr;                /* peak learning rate (cosine decay to lr/100) */
    float val_frac;          /* validation share of the states (0, 0.5] */
    double delta_quantile;   /* quantile of |error| on validation for delta_e */
    double delta_scale;      /* delta_e = scale * quantile */
    int32_t nexp;            /* bracket moves, see grmhd_mlp_set_bracket */
    uint64_t seed;           /* minibatch order */
} grmhd_mlp_fit_opts;

typedef struct {
    int64_t n_train, n_val, steps;
    double loss_first, loss_last;        /* minibatch losses */
    double val_mae_e, val_p50_e, val_p99_e, val_max_e;   /* |log x_e error| on validation */
    double val_mae_s, val_p99_s;
    double delta_e, delta_s;
    double seconds;
} grmhd_mlp_fit_report;

void grmhd_mlp_default_fit_opts(grmhd_mlp_fit_opts* opts);

/* Generates training pairs with the plain c2p on a copy of the states,
 * sets the scaling from them, trains on the GPU, and sets delta_e/delta_s
 * from the validation errors.  cons_soa is not modified. */
grmhd_status grmhd_mlp_fit(grmhd_ctx* ctx, grmhd_mlp* mlp, int64_t n, const double* cons_soa,
                           const double* metric_soa, const double* r_or_null, const grmhd_mlp_fit_opts* opts,
                           grmhd_mlp_fit_report* report, grmhd_stream_t stream);

/* Trains on the c2p inputs of up to max_states interior cells of the
 * context's block list (evenly spaced, state/B as for grmhd_rk3_step), as
 * grmhd_mlp_fit does on a batch.  Allocates and frees a temporary device
 * buffer of 20 doubles per state; synchronises the stream. */
grmhd_status grmhd_mlp_fit_cells(grmhd_ctx* ctx, grmhd_mlp* mlp, const double* state, const double* const B[3],
                                 int64_t max_states, const grmhd_mlp_fit_opts* opts,
                                 grmhd_mlp_fit_report* report, grmhd_stream_t stream);

/* Makes the c2p of the cell passes (grmhd_advance_substep,
 * grmhd_compute_auxiliaries, grmhd_rk3_step[_graph]) start from mlp's
 * prediction: the Newton finish for the ideal gas, brackets otherwise, the
 * plain solve where they fail (as grmhd_c2p_batch_learned).  mlp = NULL
 * restores the plain solve.  The context does not own mlp; a captured RK3
 * graph is re-captured. */
grmhd_status grmhd_set_learned_c2p(grmhd_ctx* ctx, grmhd_mlp* mlp);

/* Makes the c2p of the cell passes start the ideal-gas Newton finish from
 * each cell's previous primitives in aux (mu = 1/(h W) of the last c2p of
 * that cell: the previous RK stage inside grmhd_rk3_step, the caller's data
 * otherwise), at most max_newton_evals master-function evaluations before
 * the plain solve (hook: the initial guess of kastaun_c2p.hh:355-372; the
 * reference has no counterpart).  Takes precedence over an attached network
 * in the cell passes; 0 turns it off.  Ideal gas only (GRMHD_ERR_UNSUPPORTED
 * otherwise); a captured RK3 graph is re-captured. */
grmhd_status grmhd_set_c2p_warm_start(grmhd_ctx* ctx, int32_t max_newton_evals);

/* grmhd_c2p_batch with the learned guess: features, MLP and the FP64
 * Kastaun solves (Newton finish or brackets, grmhd_mlp_set_newton) in one
 * fused kernel.  Same outputs and statistics as grmhd_c2p_batch (no dry
 * run); roots_or_null as there. */
grmhd_status grmhd_c2p_batch_learned(grmhd_ctx* ctx, grmhd_mlp* mlp, int64_t n, double* cons_soa,
                                     const double* metric_soa, double* prims_soa, uint64_t* err_bits,
                                     uint8_t* floored, const double* r_or_null, double* roots_or_null,
                                     grmhd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif

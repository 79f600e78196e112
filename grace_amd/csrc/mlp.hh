// mlp.hh -- the bracket predictor for the Kastaun root solves: features,
// BF16 MFMA forward pass (FP32 accumulation) and the shared constants.
//
// Network: 10 features -> 64 -> 64 -> 2, ReLU hidden activations, linear
// output.  Outputs (normalised): y0 = log(x_e / x_hi), the energy root
// x_e = mu h0 relative to the closed-form upper end x_hi of the bracket of
// the mu+ Newton solve (kastaun_c2p.hh:44-118), and y1 = log(x_s / x_e), the
// entropy-backup root relative to the energy root.
//
// MFMA formulation (v_mfma_f32_16x16x16_bf16, 64 lanes): a layer computes
// OUT^T[j][s] = sum_k W[k][j] IN^T[k][s] for 16 states s at a time, with the
// weights as the A operand (16 j x 16 k) and the activations as the B
// operand (16 k x 16 s).  Lane l holds
//   A[i][k]: i = l%16, k = 4(l/16)+e      B[k][j]: j = l%16, k = 4(l/16)+e
//   D[i][j]: j = l%16, i = 4(l/16)+e      (e = 0..3)
// so the D fragment of one 16-row block of outputs is exactly the B fragment
// of the next layer's 16-row block of inputs: hidden activations never leave
// the registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cmath>
#include "phys.hh"

namespace gm {

constexpr int MLP_NF = 10;      // features
constexpr int MLP_KP = 16;      // padded feature count
constexpr int MLP_H = 64;       // hidden width
constexpr int MLP_NO = 2;       // outputs
// FP32 master parameter layout (row-major W[k][j], k = input, j = output)
constexpr int MLP_W1 = 0;                         // [16][64] (rows >= 10 stay zero)
constexpr int MLP_B1 = MLP_W1 + MLP_KP * MLP_H;   // [64]
constexpr int MLP_W2 = MLP_B1 + MLP_H;            // [64][64]
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;    // [64]
constexpr int MLP_W3 = MLP_B2 + MLP_H;            // [64][2]
constexpr int MLP_B3 = MLP_W3 + MLP_H * MLP_NO;   // [2]
constexpr int MLP_NP = MLP_B3 + MLP_NO;           // 5378
// BF16 A-operand fragments: layer 1 4 (jb), layer 2 16 (jb*4+kb), layer 3 4 (kb)
constexpr int MLP_NFRAG = 24;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 bits (finite inputs)
GHD uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
GHD float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// Everything a kernel needs to run the network (device pointers).
struct MlpDev {
    const uint2* frag;      // [24][64] x 4 bf16 (one uint2 per lane and fragment)
    const float* bias;      // b1[64], b2[64], b3[2]
    const float* norm;      // feat mean[10], feat inv_std[10], out mean[2], out std[2]
    double fe_lo, fe_hi;      // exp(-delta_e), exp(delta_e): energy bracket factors
    double fs_lo, fs_hi;      // the same for the entropy/energy root ratio
    int nexp;                 // bracket moves allowed before the plain path
    int newton;               // Newton evaluations from the predicted roots (0: brackets only)
};
constexpr int MLP_NORM = 24;

// Features of a densitised conserved state (kastaun_c2p.hh:306-338 variables):
//   0 log q  1 log r^2  2 log B~^2  3 (r.B~)^2/(r^2 B~^2)  4 log x_hi
//   5 log x_lo  6 log(1+q)  7 r^2/(1+q)^2  8 B~^2/(1+q)  9 log D
// x_lo, x_hi: closed-form ends of the mu+ bracket, times h0.  Also returns
// log x_hi (= feature 4) through lxhi.
template <typename EOS>
GHD void mlp_features(const EOS& eos, const Met& m, const double cons_dens[NCONS], float f[MLP_NF],
                      double& lxhi) {
    // products with reciprocals where the c2p divides: the features are
    // rounded to FP32 (and then bf16), far above the last-bit differences
    const double isg = 1.0 / m.sqrtg;
    double c[NCONS];
    for (int v = 0; v < NCONS; ++v) c[v] = cons_dens[v] * isg;
    double D = fmax(c[DENSL], 1e-300);
    const double iD = 1.0 / D, isD = sqrt(iD);
    double q = c[TAUL] * iD;
    double r0 = c[STXL] * iD, r1 = c[STYL] * iD, r2c = c[STZL] * iD;
    double b0 = c[BSXL] * isD, b1 = c[BSYL] * isD, b2 = c[BSZL] * isD;
    double r2 = m.sqcov(r0, r1, r2c);
    double Bt2 = m.sqv(b0, b1, b2);
    double rB = r0 * b0 + r1 * b1 + r2c * b2;
    double rB2 = rB * rB;
    double h0 = eos.hmin(), h02 = h0 * h0;
    double xlo = h0 / (h02 + r2);
    double mu0 = 1.0 / h0, x = 1.0 / (1.0 + mu0 * Bt2);
    double rb = x * (r2 * x + mu0 * (x + 1.0) * rB2);
    double xhi = h0 / sqrt(h02 + rb);
    double qp = fmax(q, 0.0);
    // the features are FP32 values: take the logs of FP32-rounded arguments
    // (clamped into the normal FP32 range)
    auto flog = [](double v) { return logf((float)fmin(fmax(v, 1e-37), 1e37)); };
    f[0] = flog(q);
    f[1] = flog(r2 + 1e-30);
    f[2] = flog(Bt2 + 1e-30);
    f[3] = (float)fmin(1.0, rB2 / (r2 * Bt2 + 1e-300));
    f[4] = flog(xhi);
    f[5] = flog(xlo);
    f[6] = log1pf((float)qp);
    const double i1q = 1.0 / (1.0 + qp);
    f[7] = (float)(r2 * i1q * i1q);
    f[8] = (float)(Bt2 * i1q);
    f[9] = flog(D);
    lxhi = (double)f[4];
}

// Hidden activation: ReLU (one v_max_f32; a tanh costs two quarter-rate
// transcendentals per unit, 128 units per state) and its slope in terms of
// the activation value.
__device__ __forceinline__ float mlp_act(float z) { return fmaxf(z, 0.0f); }
__device__ __forceinline__ float mlp_act_slope(float h) { return h > 0.0f ? 1.0f : 0.0f; }

// Round-to-nearest-even float -> bf16 pairs with v_cvt_pk_bf16_f32 (the same
// rounding as f2bf for finite inputs).
typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float f32v2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
    const bf16v2 h = __builtin_convertvector((f32v2){a, b}, bf16v2);
    return __builtin_bit_cast(uint32_t, h);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    const uint32_t lo = cvt_pk_bf16(a, b), hi = cvt_pk_bf16(c, d);
    bf16x4 r;
    r[0] = (short)(lo & 0xFFFF); r[1] = (short)(lo >> 16);
    r[2] = (short)(hi & 0xFFFF); r[3] = (short)(hi >> 16);
    return r;
}
__device__ __forceinline__ bf16x4 as_bf16x4(uint2 u) {
    bf16x4 r;
    r[0] = (short)(u.x & 0xFFFF); r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xFFFF); r[3] = (short)(u.y >> 16);
    return r;
}

// This lane's slices of the biases, in D layout: b1/b2 rows jb*16+4(l/16)+e,
// b3 rows 4(l/16)+e (zero beyond the two outputs).
struct MlpBias { float b1[16], b2[16], b3[4]; };

__device__ __forceinline__ void mlp_load_bias(const MlpDev& w, MlpBias& b) {
    const int rowoff = 4 * ((threadIdx.x & 63) >> 4);
    #pragma unroll
    for (int jb = 0; jb < 4; ++jb)
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
            b.b1[jb * 4 + e] = w.bias[jb * 16 + rowoff + e];
            b.b2[jb * 4 + e] = w.bias[MLP_H + jb * 16 + rowoff + e];
        }
    #pragma unroll
    for (int e = 0; e < 4; ++e) b.b3[e] = (rowoff + e < MLP_NO) ? w.bias[2 * MLP_H + rowoff + e] : 0.0f;
}

__device__ __forceinline__ void mlp_load_frags(const MlpDev& w, bf16x4 A[MLP_NFRAG]) {
    const int l = threadIdx.x & 63;
    #pragma unroll
    for (int f = 0; f < MLP_NFRAG; ++f) A[f] = as_bf16x4(w.frag[f * 64 + l]);
}

// Forward pass of one 16-state tile.  xin: this lane's B fragment of the
// normalised bf16 features (k = 4(l/16)+e of state l%16).  Returns the
// layer-3 D fragment (lanes 0-15: outputs 0 and 1 of state l in [0], [1]).
// With KEEP, h1/h2 receive the FP32 post-activation values in D layout.
template <bool KEEP>
__device__ __forceinline__ f32x4 mlp_tile(const bf16x4 A[MLP_NFRAG], const MlpBias& b, bf16x4 xin,
                                          f32x4* h1keep, f32x4* h2keep) {
    bf16x4 h1b[4], h2b[4];
    #pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
        f32x4 c;
        #pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = b.b1[jb * 4 + e];
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[jb], xin, c, 0, 0, 0);
        #pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = mlp_act(c[e]);
        if (KEEP) h1keep[jb] = c;
        h1b[jb] = pack4(c[0], c[1], c[2], c[3]);
    }
    #pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
        f32x4 c;
        #pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = b.b2[jb * 4 + e];
        #pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[4 + jb * 4 + kb], h1b[kb], c, 0, 0, 0);
        #pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = mlp_act(c[e]);
        if (KEEP) h2keep[jb] = c;
        h2b[jb] = pack4(c[0], c[1], c[2], c[3]);
    }
    f32x4 o;
    #pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = b.b3[e];
    #pragma unroll
    for (int kb = 0; kb < 4; ++kb)
        o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[20 + kb], h2b[kb], o, 0, 0, 0);
    return o;
}

// Normalised bf16 features of this lane's state into the wave's LDS rows.
__device__ __forceinline__ void mlp_stage_features(const MlpDev& w, const float feat[MLP_NF], uint16_t* xs) {
    const int l = threadIdx.x & 63;
    float v[MLP_KP];
    #pragma unroll
    for (int k = 0; k < MLP_KP; ++k) v[k] = (k < MLP_NF) ? (feat[k] - w.norm[k]) * w.norm[MLP_NF + k] : 0.0f;
    uint32_t* row = reinterpret_cast<uint32_t*>(xs + l * MLP_KP);
    #pragma unroll
    for (int k = 0; k < MLP_KP; k += 2) row[k / 2] = cvt_pk_bf16(v[k], v[k + 1]);
}

// Whole-block forward pass: each wave runs the 64 states of its lanes.
// feat: this lane's raw features.  xs: 64*16 bf16 and ys: 64*2 floats of
// LDS scratch per wave.  Every thread of the block must call it (it holds
// two block barriers).  Returns the de-normalised outputs of this lane.
__device__ __forceinline__ void mlp_block_forward(const MlpDev& w, const float feat[MLP_NF], uint16_t* xs,
                                                  float* ys, float& y0, float& y1) {
    const int l = threadIdx.x & 63;
    mlp_stage_features(w, feat, xs);
    __syncthreads();
    bf16x4 A[MLP_NFRAG];
    mlp_load_frags(w, A);
    MlpBias b;
    mlp_load_bias(w, b);
    const int rowoff = 4 * (l >> 4);
    #pragma unroll 1
    for (int t = 0; t < 4; ++t) {
        const uint2 xv = *reinterpret_cast<const uint2*>(xs + (t * 16 + (l & 15)) * MLP_KP + rowoff);
        f32x4 o = mlp_tile<false>(A, b, as_bf16x4(xv), nullptr, nullptr);
        if (l < 16) { ys[(t * 16 + l) * 2] = o[0]; ys[(t * 16 + l) * 2 + 1] = o[1]; }
    }
    __syncthreads();
    y0 = ys[l * 2] * w.norm[2 * MLP_NF + 2] + w.norm[2 * MLP_NF];
    y1 = ys[l * 2 + 1] * w.norm[2 * MLP_NF + 3] + w.norm[2 * MLP_NF + 1];
}

}  // namespace gm

// Host-side handle (opaque in the C ABI).
struct grmhd_mlp {
    int device = 0;
    float* master = nullptr;   // FP32 parameters, MLP_NP
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* wq = nullptr;       // bf16-rounded parameters as floats, MLP_NP
    float* w2t = nullptr;      // bf16-rounded W2 transposed, [64][64]
    uint2* frag = nullptr;     // A fragments, 24 x 64 lanes
    float* bias = nullptr;     // b1, b2, b3 (130 floats)
    float* norm = nullptr;     // gm::MLP_NORM floats
    float* grad = nullptr;     // MLP_NP gradients + loss
    float* part = nullptr;     // per-workgroup partial gradients
    int64_t part_cap = 0;      // workgroups the partial buffer holds
    float norm_h[gm::MLP_NORM];
    double delta_e = 0.02, delta_s = 1e-6;
    int nexp = 2;
    int newton = 8;
    int64_t step = 0;
    gm::MlpDev dev() const {
        return gm::MlpDev{frag, bias, norm, std::exp(-delta_e), std::exp(delta_e), std::exp(-delta_s),
                          std::exp(delta_s), nexp, newton};
    }
};

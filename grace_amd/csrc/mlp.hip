// mlp.hip -- training and inference kernels of the Kastaun bracket predictor
// (10 -> 64 -> 64 -> 2, BF16 MFMA forward with FP32 accumulation, FP32
// backward, Adam on FP32 master parameters) and its C ABI
// (include/grmhd_mlp.h).  The network has no counterpart in the reference:
// it narrows the root brackets of kastaun_c2p.hh:355-372 and
// ent_based_c2p.hh:214-229, whose FP64 solves still decide every result.
//
// Training data are (features, targets) pairs the plain c2p generates
// (grmhd_c2p_batch roots output); grmhd_mlp_fit below drives the training.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <vector>
#include "mlp.hh"
#include "../../include/grmhd_mlp.h"

using namespace gm;

void grmhd_set_error_text(const char* s);   // grmhd.hip

static void mset_err(const char* fmt, ...) {
    char buf[512];
    va_list ap; va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    grmhd_set_error_text(buf);
}

#define MHIP_TRY(x)                                                                 \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            mset_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GRMHD_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

static grmhd_status mlaunch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { mset_err("%s launch failed: %s", what, hipGetErrorString(e)); return GRMHD_ERR_HIP; }
    return GRMHD_OK;
}

// ---------------------------------------------------------------------------
// Parameter packing: FP32 master -> bf16-rounded copies in the layouts the
// kernels read (A fragments, transposed W2, bias vector).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ master, float* wq, float* w2t,
                                                  uint2* frag, float* bias) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_NP) {
        float v = bf2f(f2bf(master[t]));
        wq[t] = v;
        if (t >= MLP_W2 && t < MLP_B2) {
            int k = (t - MLP_W2) / MLP_H, j = (t - MLP_W2) % MLP_H;
            w2t[j * MLP_H + k] = v;
        }
    }
    if (t < MLP_H) { bias[t] = master[MLP_B1 + t]; bias[MLP_H + t] = master[MLP_B2 + t]; }
    if (t < MLP_NO) bias[2 * MLP_H + t] = master[MLP_B3 + t];
    if (t < MLP_NFRAG * 64) {
        const int f = t / 64, l = t % 64, i = l % 16, k0 = 4 * (l / 16);
        float v[4];
        for (int e = 0; e < 4; ++e) {
            const int k = k0 + e;
            if (f < 4) v[e] = master[MLP_W1 + k * MLP_H + f * 16 + i];
            else if (f < 20) {
                const int jb = (f - 4) / 4, kb = (f - 4) % 4;
                v[e] = master[MLP_W2 + (kb * 16 + k) * MLP_H + jb * 16 + i];
            } else {
                const int kb = f - 20;
                v[e] = (i < MLP_NO) ? master[MLP_W3 + (kb * 16 + k) * MLP_NO + i] : 0.0f;
            }
        }
        uint2 u;
        u.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        u.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        frag[t] = u;
    }
}

// ---------------------------------------------------------------------------
// Inference on precomputed raw features (SoA [10][n]) -> de-normalised
// outputs (SoA [2][n]).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_forward(int64_t n, MlpDev w, const float* __restrict__ feat,
                                                     float* __restrict__ out) {
    __shared__ __align__(16) uint16_t xs[4][64 * MLP_KP];
    __shared__ __align__(16) float ys[4][64 * 2];
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int wv = threadIdx.x >> 6;
    float f[MLP_NF];
    for (int k = 0; k < MLP_NF; ++k) f[k] = (s < n) ? feat[k * n + s] : w.norm[k];
    float y0, y1;
    mlp_block_forward(w, f, xs[wv], ys[wv], y0, y1);
    if (s < n) { out[s] = y0; out[n + s] = y1; }
}

// ---------------------------------------------------------------------------
// Loss gradient of one 64-state chunk per workgroup (one wave).
// loss = inv_n * sum_s sum_j m_sj (y_sj - t_sj)^2 in normalised units, with
// m_sj = 0 for NaN targets and states beyond n.  Forward: the MFMA tiles of
// the inference path (bf16 operands, FP32 accumulation).  Backward in FP32:
//   d3 = 2 m (y - t) inv_n
//   dz2 = (d3 W3q^T) [h2 > 0]        dz1 = (dz2 W2q^T) [h1 > 0]
//   dW3 = bf(h2)^T d3   dW2 = bf(h1)^T dz2   dW1 = bf(x)^T dz1   db = sums
// with Wq the bf16-rounded weights and bf() the bf16 rounding the forward
// applied to each layer input.  Partial sums go to part[wg][MLP_NP + 1]
// (last entry: loss).
// ---------------------------------------------------------------------------
constexpr int HS = MLP_H + 1;   // LDS row stride (floats) of the [64][64] tiles

__global__ __launch_bounds__(64) void k_mlp_grad(int64_t n, int64_t ld, MlpDev w, const float* __restrict__ wq,
                                                 const float* __restrict__ w2t, const float* __restrict__ feat,
                                                 const float* __restrict__ target, float inv_n,
                                                 float* __restrict__ part) {
    __shared__ __align__(16) uint16_t xs[64 * MLP_KP];
    __shared__ float H1[64 * HS];
    __shared__ float H2[64 * HS];     // reused for dz1 once dW3 is formed
    __shared__ float DZ2[64 * HS];
    __shared__ float D3[64 * 2];
    const int l = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.x * 64 + l;
    const bool act = s < n;
    float f[MLP_NF];
    for (int k = 0; k < MLP_NF; ++k) f[k] = act ? feat[k * ld + s] : w.norm[k];
    mlp_stage_features(w, f, xs);
    __syncthreads();
    bf16x4 A[MLP_NFRAG];
    mlp_load_frags(w, A);
    MlpBias b;
    mlp_load_bias(w, b);
    const int rowoff = 4 * (l >> 4);
    for (int t = 0; t < 4; ++t) {
        const uint2 xv = *reinterpret_cast<const uint2*>(xs + (t * 16 + (l & 15)) * MLP_KP + rowoff);
        f32x4 h1[4], h2[4];
        f32x4 o = mlp_tile<true>(A, b, as_bf16x4(xv), h1, h2);
        const int st = t * 16 + (l & 15);
        for (int jb = 0; jb < 4; ++jb)
            for (int e = 0; e < 4; ++e) {
                H1[st * HS + jb * 16 + rowoff + e] = h1[jb][e];
                H2[st * HS + jb * 16 + rowoff + e] = h2[jb][e];
            }
        if (l < 16) { D3[st * 2] = o[0]; D3[st * 2 + 1] = o[1]; }
    }
    __syncthreads();
    // per-state output error (lane = state)
    float loss = 0.0f, d3[2];
    for (int j = 0; j < 2; ++j) {
        float tv = act ? target[j * ld + s] : NAN;
        float y = D3[l * 2 + j];
        float tn = (tv - w.norm[2 * MLP_NF + j]) / w.norm[2 * MLP_NF + 2 + j];
        bool ok = act && (tv == tv);
        float d = ok ? (y - tn) : 0.0f;
        loss += d * d * inv_n;
        d3[j] = 2.0f * d * inv_n;
    }
    __syncthreads();
    D3[l * 2] = d3[0]; D3[l * 2 + 1] = d3[1];
    for (int k = 0; k < MLP_H; ++k) {
        float h = H2[l * HS + k];
        float dh = d3[0] * wq[MLP_W3 + k * MLP_NO] + d3[1] * wq[MLP_W3 + k * MLP_NO + 1];
        DZ2[l * HS + k] = dh * mlp_act_slope(h);
    }
    __syncthreads();
    float* P = part + (int64_t)blockIdx.x * (MLP_NP + 1);
    // thread l owns row l of W3/W2 and column l of W1
    {
        float g0 = 0.0f, g1 = 0.0f;
        for (int st = 0; st < 64; ++st) {
            float h = bf2f(f2bf(H2[st * HS + l]));
            g0 += h * D3[st * 2]; g1 += h * D3[st * 2 + 1];
        }
        P[MLP_W3 + l * MLP_NO] = g0; P[MLP_W3 + l * MLP_NO + 1] = g1;
        if (l < MLP_NO) {
            float gb = 0.0f;
            for (int st = 0; st < 64; ++st) gb += D3[st * 2 + l];
            P[MLP_B3 + l] = gb;
        }
        float gw[MLP_H];
        for (int j = 0; j < MLP_H; ++j) gw[j] = 0.0f;
        float gb2 = 0.0f;
        for (int st = 0; st < 64; ++st) {
            float h = bf2f(f2bf(H1[st * HS + l]));
            gb2 += DZ2[st * HS + l];
            for (int j = 0; j < MLP_H; ++j) gw[j] += h * DZ2[st * HS + j];
        }
        for (int j = 0; j < MLP_H; ++j) P[MLP_W2 + l * MLP_H + j] = gw[j];
        P[MLP_B2 + l] = gb2;
    }
    __syncthreads();
    // dz1 of this lane's state into the H2 buffer
    {
        float acc[MLP_H];
        for (int k = 0; k < MLP_H; ++k) acc[k] = 0.0f;
        for (int j = 0; j < MLP_H; ++j) {
            float dz = DZ2[l * HS + j];
            for (int k = 0; k < MLP_H; ++k) acc[k] += dz * w2t[j * MLP_H + k];
        }
        for (int k = 0; k < MLP_H; ++k) {
            float h = H1[l * HS + k];
            H2[l * HS + k] = acc[k] * mlp_act_slope(h);
        }
    }
    __syncthreads();
    {
        float gw[MLP_KP], gb1 = 0.0f;
        for (int k = 0; k < MLP_KP; ++k) gw[k] = 0.0f;
        for (int st = 0; st < 64; ++st) {
            float dz = H2[st * HS + l];
            gb1 += dz;
            for (int k = 0; k < MLP_NF; ++k) gw[k] += bf2f(xs[st * MLP_KP + k]) * dz;
        }
        for (int k = 0; k < MLP_KP; ++k) P[MLP_W1 + k * MLP_H + l] = gw[k];
        P[MLP_B1 + l] = gb1;
    }
    // loss: wave sum
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (l == 0) P[MLP_NP] = loss;
}

// grad[p] = sum over workgroups in index order (deterministic)
__global__ __launch_bounds__(256) void k_mlp_reduce(int64_t nwg, const float* __restrict__ part,
                                                    float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NP) return;
    float acc = 0.0f;
    for (int64_t w = 0; w < nwg; ++w) acc += part[w * (MLP_NP + 1) + p];
    grad[p] = acc;
}

__global__ __launch_bounds__(256) void k_mlp_adam(const float* __restrict__ grad, float* master, float* m, float* v,
                                                  float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    if (p < MLP_B1 && p / MLP_H >= MLP_NF) return;   // padded W1 rows stay zero
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm; v[p] = vv;
    master[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// (features, roots) -> targets: t0 = log(x_e) - log(x_hi) (feature 4),
// t1 = log(x_s / x_e) where the entropy solve ran, NaN otherwise.
__global__ __launch_bounds__(256) void k_mlp_targets(int64_t n, const double* __restrict__ roots,
                                                     const float* __restrict__ feat, float* __restrict__ target) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const double xe = roots[s], xs = roots[n + s];
    target[s] = xe > 0 ? (float)(log(xe) - (double)feat[4 * n + s]) : NAN;
    target[n + s] = (xe > 0 && xs > 0) ? (float)log(xs / xe) : NAN;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static grmhd_status pack(grmhd_mlp* m, hipStream_t s) {
    k_mlp_pack<<<(MLP_NP + 255) / 256, 256, 0, s>>>(m->master, m->wq, m->w2t, m->frag, m->bias);
    return mlaunch_check("mlp_pack");
}

static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" grmhd_status grmhd_mlp_init_params(uint64_t seed, float* params) {
    if (!params) { mset_err("null params"); return GRMHD_ERR_INVALID; }
    uint64_t st = seed;
    auto uni = [&]() { return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
    for (int i = 0; i < MLP_NP; ++i) params[i] = 0.0f;
    // Glorot-uniform weights, zero biases
    const double a1 = std::sqrt(6.0 / (MLP_NF + MLP_H)), a2 = std::sqrt(6.0 / (2 * MLP_H)),
                 a3 = std::sqrt(6.0 / (MLP_H + MLP_NO));
    for (int k = 0; k < MLP_NF; ++k)
        for (int j = 0; j < MLP_H; ++j) params[MLP_W1 + k * MLP_H + j] = (float)(a1 * uni());
    for (int k = 0; k < MLP_H * MLP_H; ++k) params[MLP_W2 + k] = (float)(a2 * uni());
    for (int k = 0; k < MLP_H * MLP_NO; ++k) params[MLP_W3 + k] = (float)(a3 * uni());
    return GRMHD_OK;
}

extern "C" void grmhd_mlp_destroy(grmhd_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    for (void* p : {(void*)m->master, (void*)m->adam_m, (void*)m->adam_v, (void*)m->wq, (void*)m->w2t,
                    (void*)m->frag, (void*)m->bias, (void*)m->norm, (void*)m->grad, (void*)m->part})
        if (p) (void)hipFree(p);
    delete m;
}

extern "C" grmhd_status grmhd_mlp_create(uint64_t seed, grmhd_mlp** out) {
    if (!out) { mset_err("null output handle"); return GRMHD_ERR_INVALID; }
    *out = nullptr;
    grmhd_mlp* m = new grmhd_mlp();
    MHIP_TRY(hipGetDevice(&m->device));
    hipError_t e = hipSuccess;
    e = e ? e : hipMalloc(&m->master, sizeof(float) * MLP_NP);
    e = e ? e : hipMalloc(&m->adam_m, sizeof(float) * MLP_NP);
    e = e ? e : hipMalloc(&m->adam_v, sizeof(float) * MLP_NP);
    e = e ? e : hipMalloc(&m->wq, sizeof(float) * MLP_NP);
    e = e ? e : hipMalloc(&m->w2t, sizeof(float) * MLP_H * MLP_H);
    e = e ? e : hipMalloc(&m->frag, sizeof(uint2) * MLP_NFRAG * 64);
    e = e ? e : hipMalloc(&m->bias, sizeof(float) * (2 * MLP_H + MLP_NO));
    e = e ? e : hipMalloc(&m->norm, sizeof(float) * MLP_NORM);
    e = e ? e : hipMalloc(&m->grad, sizeof(float) * (MLP_NP + 1));
    if (e != hipSuccess) {
        mset_err("mlp allocation failed: %s", hipGetErrorString(e));
        grmhd_mlp_destroy(m);
        return GRMHD_ERR_ALLOC;
    }
    for (int k = 0; k < MLP_NF; ++k) { m->norm_h[k] = 0.0f; m->norm_h[MLP_NF + k] = 1.0f; }
    m->norm_h[2 * MLP_NF] = m->norm_h[2 * MLP_NF + 1] = 0.0f;
    m->norm_h[2 * MLP_NF + 2] = m->norm_h[2 * MLP_NF + 3] = 1.0f;
    std::vector<float> p(MLP_NP);
    grmhd_mlp_init_params(seed, p.data());
    grmhd_status st = grmhd_mlp_set_params(m, p.data(), nullptr);
    if (st == GRMHD_OK) st = grmhd_mlp_set_scaling(m, m->norm_h, nullptr);
    if (st != GRMHD_OK) { grmhd_mlp_destroy(m); return st; }
    *out = m;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_set_params(grmhd_mlp* m, const float* params, grmhd_stream_t stream) {
    if (!m || !params) { mset_err("null argument"); return GRMHD_ERR_INVALID; }
    hipStream_t s = (hipStream_t)stream;
    MHIP_TRY(hipMemcpyAsync(m->master, params, sizeof(float) * MLP_NP, hipMemcpyHostToDevice, s));
    MHIP_TRY(hipMemsetAsync(m->adam_m, 0, sizeof(float) * MLP_NP, s));
    MHIP_TRY(hipMemsetAsync(m->adam_v, 0, sizeof(float) * MLP_NP, s));
    m->step = 0;
    grmhd_status st = pack(m, s);
    if (st != GRMHD_OK) return st;
    MHIP_TRY(hipStreamSynchronize(s));
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_get_params(grmhd_mlp* m, float* params, grmhd_stream_t stream) {
    if (!m || !params) { mset_err("null argument"); return GRMHD_ERR_INVALID; }
    MHIP_TRY(hipMemcpyAsync(params, m->master, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, (hipStream_t)stream));
    MHIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_set_scaling(grmhd_mlp* m, const float* norm, grmhd_stream_t stream) {
    if (!m || !norm) { mset_err("null argument"); return GRMHD_ERR_INVALID; }
    for (int k = 0; k < MLP_NORM; ++k) {
        if (!std::isfinite(norm[k])) { mset_err("non-finite scaling entry %d", k); return GRMHD_ERR_INVALID; }
    }
    if (norm[2 * MLP_NF + 2] <= 0 || norm[2 * MLP_NF + 3] <= 0) {
        mset_err("output scales must be positive"); return GRMHD_ERR_INVALID;
    }
    std::memcpy(m->norm_h, norm, sizeof(m->norm_h));
    MHIP_TRY(hipMemcpyAsync(m->norm, m->norm_h, sizeof(m->norm_h), hipMemcpyHostToDevice, (hipStream_t)stream));
    MHIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_get_scaling(grmhd_mlp* m, float* norm) {
    if (!m || !norm) { mset_err("null argument"); return GRMHD_ERR_INVALID; }
    std::memcpy(norm, m->norm_h, sizeof(m->norm_h));
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_set_bracket(grmhd_mlp* m, double delta_e, double delta_s, int32_t nexp) {
    if (!m || !(delta_e > 0) || !(delta_s > 0) || nexp < 0 || nexp > 8) {
        mset_err("invalid bracket settings"); return GRMHD_ERR_INVALID;
    }
    m->delta_e = delta_e; m->delta_s = delta_s; m->nexp = nexp;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_set_newton(grmhd_mlp* m, int32_t max_evals) {
    if (!m || max_evals < 0 || max_evals > 64) { mset_err("invalid Newton evaluation limit"); return GRMHD_ERR_INVALID; }
    m->newton = max_evals;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_get_newton(grmhd_mlp* m, int32_t* max_evals) {
    if (!m || !max_evals) { mset_err("null argument"); return GRMHD_ERR_INVALID; }
    *max_evals = m->newton;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_get_bracket(grmhd_mlp* m, double* delta_e, double* delta_s, int32_t* nexp) {
    if (!m) { mset_err("null argument"); return GRMHD_ERR_INVALID; }
    if (delta_e) *delta_e = m->delta_e;
    if (delta_s) *delta_s = m->delta_s;
    if (nexp) *nexp = m->nexp;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_mlp_forward(grmhd_mlp* m, int64_t n, const float* feat, float* out,
                                          grmhd_stream_t stream) {
    if (!m || n < 0 || (n > 0 && (!feat || !out))) { mset_err("invalid mlp forward arguments"); return GRMHD_ERR_INVALID; }
    if (n == 0) return GRMHD_OK;
    k_mlp_forward<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, m->dev(), feat, out);
    return mlaunch_check("mlp_forward");
}

extern "C" grmhd_status grmhd_mlp_targets(int64_t n, const double* roots, const float* feat, float* target,
                                          grmhd_stream_t stream) {
    if (n < 0 || (n > 0 && (!roots || !feat || !target))) { mset_err("invalid target arguments"); return GRMHD_ERR_INVALID; }
    if (n == 0) return GRMHD_OK;
    k_mlp_targets<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, roots, feat, target);
    return mlaunch_check("mlp_targets");
}

static grmhd_status grad_into(grmhd_mlp* m, int64_t n, int64_t ld, const float* feat, const float* target,
                              float* grad, hipStream_t s) {
    const int64_t nwg = (n + 63) / 64;
    if (nwg > m->part_cap) {
        if (m->part) (void)hipFree(m->part);
        m->part = nullptr; m->part_cap = 0;
        MHIP_TRY(hipMalloc(&m->part, sizeof(float) * (MLP_NP + 1) * nwg));
        m->part_cap = nwg;
    }
    k_mlp_grad<<<(unsigned)nwg, 64, 0, s>>>(n, ld, m->dev(), m->wq, m->w2t, feat, target, 1.0f / (float)n, m->part);
    k_mlp_reduce<<<(MLP_NP + 1 + 255) / 256, 256, 0, s>>>(nwg, m->part, grad);
    return mlaunch_check("mlp_grad");
}

extern "C" grmhd_status grmhd_mlp_grad(grmhd_mlp* m, int64_t n, const float* feat, const float* target,
                                       float* grad_out, grmhd_stream_t stream) {
    if (!m || n <= 0 || !feat || !target || !grad_out) { mset_err("invalid mlp grad arguments"); return GRMHD_ERR_INVALID; }
    return grad_into(m, n, n, feat, target, grad_out, (hipStream_t)stream);
}

static grmhd_status train_step(grmhd_mlp* m, int64_t n, int64_t ld, const float* feat, const float* target,
                               float lr, float* loss_out, hipStream_t s) {
    grmhd_status st = grad_into(m, n, ld, feat, target, m->grad, s);
    if (st != GRMHD_OK) return st;
    const double b1 = 0.9, b2 = 0.999;
    m->step++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, (double)m->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, (double)m->step)));
    k_mlp_adam<<<(MLP_NP + 255) / 256, 256, 0, s>>>(m->grad, m->master, m->adam_m, m->adam_v, lr, (float)b1,
                                                    (float)b2, 1e-8f, c1, c2);
    st = pack(m, s);
    if (st != GRMHD_OK) return st;
    if (loss_out) MHIP_TRY(hipMemcpyAsync(loss_out, m->grad + MLP_NP, sizeof(float), hipMemcpyDeviceToDevice, s));
    return mlaunch_check("mlp_train_step");
}

extern "C" grmhd_status grmhd_mlp_train_step(grmhd_mlp* m, int64_t n, const float* feat, const float* target,
                                             float lr, float* loss_out, grmhd_stream_t stream) {
    if (!m || n <= 0 || !feat || !target || !(lr > 0)) { mset_err("invalid mlp train arguments"); return GRMHD_ERR_INVALID; }
    return train_step(m, n, n, feat, target, lr, loss_out, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Training driver
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_gather(int64_t n_all, int64_t n, const int64_t* __restrict__ idx,
                                                    const float* __restrict__ feat, const float* __restrict__ target,
                                                    float* __restrict__ feat_out, float* __restrict__ target_out) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int64_t src = idx[s];
    for (int k = 0; k < MLP_NF; ++k) feat_out[k * n + s] = feat[k * n_all + src];
    for (int j = 0; j < MLP_NO; ++j) target_out[j * n + s] = target[j * n_all + src];
}

extern "C" void grmhd_mlp_default_fit_opts(grmhd_mlp_fit_opts* o) {
    if (!o) return;
    o->epochs = 30;
    o->batch = 4096;
    o->lr = 3e-3f;
    o->val_frac = 0.1f;
    o->delta_quantile = 0.9;
    o->delta_scale = 1.0;
    o->nexp = 3;
    o->seed = 20261202ull;
}

namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};
double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double pos = q * (double)(v.size() - 1);
    size_t i = (size_t)pos;
    double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}
}  // namespace

extern "C" grmhd_status grmhd_mlp_fit(grmhd_ctx* ctx, grmhd_mlp* m, int64_t n, const double* cons_soa,
                                      const double* metric_soa, const double* r, const grmhd_mlp_fit_opts* opts_in,
                                      grmhd_mlp_fit_report* rep, grmhd_stream_t stream) {
    if (!ctx || !m || n < 16 || !cons_soa || !metric_soa) { mset_err("invalid fit arguments"); return GRMHD_ERR_INVALID; }
    grmhd_mlp_fit_opts o;
    grmhd_mlp_default_fit_opts(&o);
    if (opts_in) o = *opts_in;
    if (o.epochs < 0 || o.batch < 64 || !(o.lr > 0) || !(o.val_frac > 0 && o.val_frac <= 0.5) ||
        !(o.delta_quantile > 0 && o.delta_quantile <= 1) || !(o.delta_scale > 0) || o.nexp < 0 || o.nexp > 8) {
        mset_err("invalid fit options"); return GRMHD_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    auto t0 = std::chrono::steady_clock::now();
    const int64_t n_val = std::max<int64_t>(1, (int64_t)(o.val_frac * (double)n));
    const int64_t n_tr = n - n_val;
    if (n_tr < o.batch) { mset_err("fewer training states than one minibatch"); return GRMHD_ERR_INVALID; }
    DevBuf cons, prims, err, fl, roots, feat, target, ftr, ttr, idx, out, loss;
    MHIP_TRY(hipMalloc(&cons.p, sizeof(double) * 10 * n));
    MHIP_TRY(hipMalloc(&prims.p, sizeof(double) * 13 * n));
    MHIP_TRY(hipMalloc(&err.p, sizeof(uint64_t) * n));
    MHIP_TRY(hipMalloc(&fl.p, n));
    MHIP_TRY(hipMalloc(&roots.p, sizeof(double) * 4 * n));
    MHIP_TRY(hipMalloc(&feat.p, sizeof(float) * MLP_NF * n));
    MHIP_TRY(hipMalloc(&target.p, sizeof(float) * MLP_NO * n));
    MHIP_TRY(hipMalloc(&ftr.p, sizeof(float) * MLP_NF * n_tr));
    MHIP_TRY(hipMalloc(&ttr.p, sizeof(float) * MLP_NO * n_tr));
    MHIP_TRY(hipMalloc(&idx.p, sizeof(int64_t) * n_tr));
    MHIP_TRY(hipMalloc(&out.p, sizeof(float) * MLP_NO * n_val));
    MHIP_TRY(hipMalloc(&loss.p, sizeof(float) * 2));
    // 1. training pairs from the plain c2p (on a copy: c2p rewrites cons)
    MHIP_TRY(hipMemcpyAsync(cons.p, cons_soa, sizeof(double) * 10 * n, hipMemcpyDeviceToDevice, s));
    grmhd_status st = grmhd_c2p_batch(ctx, n, (double*)cons.p, metric_soa, (double*)prims.p, (uint64_t*)err.p,
                                      (uint8_t*)fl.p, 0, r, nullptr, (double*)roots.p, stream);
    if (st == GRMHD_OK) st = grmhd_mlp_features(ctx, n, cons_soa, metric_soa, (float*)feat.p, stream);
    if (st == GRMHD_OK) st = grmhd_mlp_targets(n, (const double*)roots.p, (const float*)feat.p, (float*)target.p, stream);
    if (st != GRMHD_OK) return st;
    // 2. scaling from the training split (the first n_tr states)
    std::vector<float> fh((size_t)MLP_NF * n), th((size_t)MLP_NO * n);
    MHIP_TRY(hipMemcpyAsync(fh.data(), feat.p, sizeof(float) * fh.size(), hipMemcpyDeviceToHost, s));
    MHIP_TRY(hipMemcpyAsync(th.data(), target.p, sizeof(float) * th.size(), hipMemcpyDeviceToHost, s));
    MHIP_TRY(hipStreamSynchronize(s));
    float norm[MLP_NORM];
    for (int k = 0; k < MLP_NF + MLP_NO; ++k) {
        const float* col = (k < MLP_NF) ? &fh[(size_t)k * n] : &th[(size_t)(k - MLP_NF) * n];
        double sum = 0, sq = 0;
        int64_t cnt = 0;
        for (int64_t i = 0; i < n_tr; ++i) {
            double v = col[i];
            if (!std::isfinite(v)) continue;
            sum += v; sq += v * v; cnt++;
        }
        double mean = cnt ? sum / (double)cnt : 0.0;
        double sd = cnt ? std::sqrt(std::max(0.0, sq / (double)cnt - mean * mean)) : 1.0;
        if (k < MLP_NF) {
            if (!(sd > 1e-12)) sd = 1.0;
            norm[k] = (float)mean;
            norm[MLP_NF + k] = (float)(1.0 / sd);
        } else {
            // targets that barely vary (e.g. log(x_s/x_e) on consistent
            // states) are not blown up into unit-variance round-off noise
            if (!(sd > 1e-6)) sd = 1e-6;
            norm[2 * MLP_NF + (k - MLP_NF)] = (float)mean;
            norm[2 * MLP_NF + 2 + (k - MLP_NF)] = (float)sd;
        }
    }
    st = grmhd_mlp_set_scaling(m, norm, stream);
    if (st != GRMHD_OK) return st;
    // 3. Adam over minibatches of a fresh permutation each epoch, cosine decay
    uint64_t rs = o.seed;
    std::vector<int64_t> perm(n_tr);
    for (int64_t i = 0; i < n_tr; ++i) perm[i] = i;
    const int64_t nb = n_tr / o.batch;
    const int64_t total = (int64_t)o.epochs * nb;
    int64_t step = 0;
    for (int ep = 0; ep < o.epochs; ++ep) {
        for (int64_t i = n_tr - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(rs) % (uint64_t)(i + 1)]);
        MHIP_TRY(hipMemcpyAsync(idx.p, perm.data(), sizeof(int64_t) * n_tr, hipMemcpyHostToDevice, s));
        k_mlp_gather<<<(unsigned)((n_tr + 255) / 256), 256, 0, s>>>(n, n_tr, (const int64_t*)idx.p,
                                                                    (const float*)feat.p, (const float*)target.p,
                                                                    (float*)ftr.p, (float*)ttr.p);
        if ((st = mlaunch_check("mlp_gather")) != GRMHD_OK) return st;
        for (int64_t b = 0; b < nb; ++b, ++step) {
            const double c = std::cos(M_PI * (double)step / (double)std::max<int64_t>(1, total - 1));
            const float lr = (float)(o.lr * (0.01 + 0.99 * 0.5 * (1.0 + c)));
            float* lo = (step == 0) ? (float*)loss.p : (step == total - 1 ? (float*)loss.p + 1 : nullptr);
            st = train_step(m, o.batch, n_tr, (const float*)ftr.p + b * o.batch, (const float*)ttr.p + b * o.batch,
                            lr, lo, s);
            if (st != GRMHD_OK) return st;
        }
    }
    // 4. validation errors (the last n_val states) -> bracket widths
    const int64_t v0 = n_tr;
    // the forward kernel reads SoA with stride n_val: gather the validation columns first
    std::vector<int64_t> vidx(n_val);
    for (int64_t i = 0; i < n_val; ++i) vidx[i] = v0 + i;
    MHIP_TRY(hipMemcpyAsync(idx.p, vidx.data(), sizeof(int64_t) * n_val, hipMemcpyHostToDevice, s));
    k_mlp_gather<<<(unsigned)((n_val + 255) / 256), 256, 0, s>>>(n, n_val, (const int64_t*)idx.p,
                                                                 (const float*)feat.p, (const float*)target.p,
                                                                 (float*)ftr.p, (float*)ttr.p);
    if ((st = mlaunch_check("mlp_gather")) != GRMHD_OK) return st;
    st = grmhd_mlp_forward(m, n_val, (const float*)ftr.p, (float*)out.p, stream);
    if (st != GRMHD_OK) return st;
    std::vector<float> oh((size_t)MLP_NO * n_val), tv((size_t)MLP_NO * n_val);
    float lh[2];
    MHIP_TRY(hipMemcpyAsync(oh.data(), out.p, sizeof(float) * oh.size(), hipMemcpyDeviceToHost, s));
    MHIP_TRY(hipMemcpyAsync(tv.data(), ttr.p, sizeof(float) * tv.size(), hipMemcpyDeviceToHost, s));
    MHIP_TRY(hipMemcpyAsync(lh, loss.p, sizeof(lh), hipMemcpyDeviceToHost, s));
    MHIP_TRY(hipStreamSynchronize(s));
    std::vector<double> ee, es;
    for (int64_t i = 0; i < n_val; ++i) {
        if (std::isfinite(tv[i])) ee.push_back(std::fabs((double)oh[i] - (double)tv[i]));
        if (std::isfinite(tv[n_val + i])) es.push_back(std::fabs((double)oh[n_val + i] - (double)tv[n_val + i]));
    }
    auto mean = [](const std::vector<double>& v) {
        double a = 0; for (double x : v) a += x; return v.empty() ? 0.0 : a / (double)v.size();
    };
    const double de = std::max(1e-6, o.delta_scale * quantile(ee, o.delta_quantile));
    const double ds = std::max(1e-12, o.delta_scale * quantile(es, o.delta_quantile));
    st = grmhd_mlp_set_bracket(m, de, ds, o.nexp);
    if (st != GRMHD_OK) return st;
    if (rep) {
        rep->n_train = n_tr; rep->n_val = n_val; rep->steps = step;
        rep->loss_first = lh[0]; rep->loss_last = lh[1];
        rep->val_mae_e = mean(ee); rep->val_p50_e = quantile(ee, 0.5); rep->val_p99_e = quantile(ee, 0.99);
        rep->val_max_e = quantile(ee, 1.0);
        rep->val_mae_s = mean(es); rep->val_p99_s = quantile(es, 0.99);
        rep->delta_e = de; rep->delta_s = ds;
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return GRMHD_OK;
}

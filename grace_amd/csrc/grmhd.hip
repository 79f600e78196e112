// grmhd.hip -- gfx950 kernels and C ABI for the Cowling GRMHD RHS + con2prim.
//
// Kernel families (all FP64):
//   k_flux<D>      one thread per face of direction D: 10-variable
//                  reconstruction, face metric, EOS, HLLE/LLF fluxes, vbar/Eface
//                  (reference: evolve.cpp:813-1068, grmhd.hh:551-970)
//   k_emf_uct      one thread per edge triple: UCT edge EMFs (evolve.cpp:1130-1340)
//   k_emf_gs       Gardiner-Stone edge EMFs (grmhd_helpers.hh:492-627)
//   k_update       fused RK linear combination + geometric sources + flux
//                  divergence + CT face update (tstep_utils.hh:140-220,
//                  evolve.cpp:1342-1479)
//   k_gather       block ghost-zone gather (boundary_conditions.cpp:55-81)
//   k_amr_*        coarse-fine ghosts and refluxing of a two-level forest
//                  (forest.hh; prolongation/restriction kernels, refluxing.cpp)
//   k_c2p_cells    c2p over cells + conserved rewrite + BDIV + E_center
//                  (auxiliaries.cpp:81-180, grmhd.hh:306-434)
//   k_c2p_batch    c2p over a flat batch (c2p.cpp:188-414)
//   k_dt_*         CFL reduction (find_stable_timestep.cpp:72-126)
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <string>
#include <vector>
#include <map>
#include <unordered_map>
#include <mutex>
#include "phys.hh"
#include "mlp.hh"
#include "forest.hh"
#include "../../include/grmhd_abi.h"
#include "../../include/grmhd_mlp.h"

using namespace gm;

static thread_local char g_errbuf[512] = "";
static void set_err(const char* fmt, ...) {
    va_list ap; va_start(ap, fmt);
    vsnprintf(g_errbuf, sizeof(g_errbuf), fmt, ap);
    va_end(ap);
}
extern "C" const char* grmhd_last_error(void) { return g_errbuf; }
void grmhd_set_error_text(const char* t) { set_err("%s", t); }

#define HIP_TRY(x)                                                                  \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            set_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GRMHD_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

// ---------------------------------------------------------------------------
// Device-side geometry of the block grid
// ---------------------------------------------------------------------------
struct DevGrid {
    int n[3];          // interior cells per block
    int ng;            // ghost layers
    int N[3];          // padded extents
    int64_t ncell;     // cells per block (padded)
    int64_t nface[3];  // faces per block per direction
    int64_t nedge;     // edges per block per component
    int64_t nq;
};
struct DevInst { int nb[3]; int64_t q0; int bc[6]; double idx[3]; double x0[3]; };

// Radius of cell (i, j, k) (ghosts included) of block b for radius-based
// excision: centre x0 + (global index + 1/2) / idx, as the reference's
// get_physical_coordinates (cartesian_coordinate_systems.hh:56-66).
GHD double cell_radius(const ExcPar& ex, const DevInst& I, const int bb[3], const int n[3], int ng, int i, int j,
                       int k) {
    const int loc[3] = {i, j, k};
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = I.x0[d] + ((double)(bb[d] * n[d] + loc[d] - ng) + 0.5) / I.idx[d];
    return excision_radius_of(ex, x[0], x[1], x[2]);
}
struct DevBlk { int inst; int b[3]; };

struct Opts {
    Atmo atmo; C2PPar cp; ExcPar ex;
    int riemann, emf, thermo, fd_order;
    double c_limit;
    int fofc, dmp;      // first-order flux correction, its DMP test
    double dmp_M;
    // cells that took a branch where the reference aborts (c2p.hh:160-322),
    // counted by the c2p passes until a sync point reads and clears it
    unsigned long long* fatal_count;
    // batch c2p: per-state master-function + bracket evaluations (clamped to
    // 255) when set (grmhd_set_c2p_eval_counts)
    uint8_t* eval_counts;
};

GHD int64_t cidx(const DevGrid& G, int i, int j, int k) { return i + (int64_t)G.N[0] * (j + (int64_t)G.N[1] * k); }
template <int D>
GHD int64_t fidx(const DevGrid& G, int i, int j, int k) {
    int ei = G.N[0] + (D == 0), ej = G.N[1] + (D == 1);
    return i + (int64_t)ei * (j + (int64_t)ej * k);
}
GHD int64_t fidx_d(const DevGrid& G, int d, int i, int j, int k) {
    int ei = G.N[0] + (d == 0), ej = G.N[1] + (d == 1);
    return i + (int64_t)ei * (j + (int64_t)ej * k);
}
GHD int64_t eidx(const DevGrid& G, int i, int j, int k) {
    return i + (int64_t)(G.N[0] + 1) * (j + (int64_t)(G.N[1] + 1) * k);
}

// Position of this workgroup in a 1-D grid of wpb workgroups per block times
// the number of blocks.  The hardware deals workgroups to the 8 XCDs round
// robin; the remap hands each XCD a contiguous range of logical workgroups,
// so the workgroups of one block share that XCD's L2 instead of all eight
// L2s fetching the same stencil data.
// (A 1-D grid has fewer than 2^32 workgroups, so 32-bit unsigned
// arithmetic: one short scalar division instead of a 64-bit one.)
struct WgPos { int tile; int64_t q; };
__device__ __forceinline__ WgPos wg_pos(int wpb) {
    const uint32_t p = blockIdx.x, n = gridDim.x;
    const uint32_t x = p & 7u, i = p >> 3, base = n >> 3, rem = n & 7u;
    const uint32_t L = x * base + (x < rem ? x : rem) + i;
    const uint32_t w = (uint32_t)wpb, q = L / w;
    return WgPos{(int)(L - q * w), (int64_t)q};
}
// x = q d + r for x >= 0, d > 0 (unsigned: no sign fix-ups)
GHD void udivmod(int x, int d, int& q, int& r) {
    const uint32_t uq = (uint32_t)x / (uint32_t)d;
    q = (int)uq;
    r = x - (int)uq * d;
}

// Read the Cowling metric of a cell (state vars GXX_..BETAZ_).
GHD void load_metric(Met& m, const double* __restrict__ ms, int64_t ncell, int64_t c) {
    for (int a = 0; a < 6; ++a) m.g[a] = ms[(GXX_ + a) * ncell + c];
    m.alp = ms[ALP_ * ncell + c];
    for (int a = 0; a < 3; ++a) m.beta[a] = ms[(BETAX_ + a) * ncell + c];
    m.init();
}

// Metric of the face below cell c0 along stride st: 4-point interpolation of
// the 10 Cowling components (face_interp_4, grmhd_helpers.hh:360-380), their
// inverse and determinant, and the lapse test of the excision donor
// override (the lapse of either neighbouring cell at or below alp_ex).
// Shared by the flux kernel and the face-metric cache, so both give the
// same bits.
GHD void face_metric(const double* __restrict__ M, int64_t nc, int64_t c0, int64_t st, double alp_ex, Met& mf,
                     bool& excise) {
    double v[10];
    const int comps[10] = {GXX_, GXY_, GXZ_, GYY_, GYZ_, GZZ_, BETAX_, BETAY_, BETAZ_, ALP_};
#pragma unroll
    for (int a = 0; a < 10; ++a) {
        const double* P = M + comps[a] * nc + c0;
        double si = P[-st] + P[0];
        double so = P[-2 * st] + P[st];
        v[a] = 0.5625 * si + (-0.0625) * so;
    }
    for (int a = 0; a < 6; ++a) mf.g[a] = v[a];
    for (int a = 0; a < 3; ++a) mf.beta[a] = v[6 + a];
    mf.alp = v[9];
    mf.init();
    double alpL = M[ALP_ * nc + c0 - st], alpR = M[ALP_ * nc + c0];
    excise = (alpL <= alp_ex) || (alpR <= alp_ex);
}
// Face-metric cache (grmhd_set_static_metric): per face of direction D the
// values the flux kernel reads -- g (6), the diagonal inverse component of
// D, beta (3), alpha, sqrt(gamma), the excision flag -- as FMC arrays of
// nface[D] per block.
constexpr int FMC = 13;

// ---------------------------------------------------------------------------
// Face flux kernel
// ---------------------------------------------------------------------------
// Minimum waves per SIMD the compiler must fit (VGPR budget); unset: its own choice.
#ifdef GRMHD_FLUX_MINW
#define GRMHD_FLUX_BOUNDS __launch_bounds__(256, GRMHD_FLUX_MINW)
#else
#define GRMHD_FLUX_BOUNDS __launch_bounds__(256)
#endif
// PF: every stencil value of the face is loaded before the first
// reconstruction (54 loads in flight per lane instead of 6; about 150 VGPRs,
// 3 waves per SIMD instead of 4): a tuning choice per direction
// (grmhd_tuning.flux_prefetch), results bit-identical.
template <typename EOS, int REC, bool FLAT, int D, bool PF = false>
__global__ GRMHD_FLUX_BOUNDS void k_flux(DevGrid G, EOS eos, Opts o, const double* __restrict__ aux,
                                              const double* __restrict__ mstate, int mnv,
                                              const double* __restrict__ Bd, double* __restrict__ flux,
                                              double* __restrict__ vbar, int wpb,
                                              const uint8_t* __restrict__ only,
                                              const uint8_t* __restrict__ ring_own,
                                              const int32_t* __restrict__ push_nbr,
                                              const double* __restrict__ fmc) {
    // wpb workgroups per block (see wg_pos).  UCT: the first
    // (n_D+1) n_T0 n_T1 threads take the faces whose transverse indices are
    // both interior (full fluxes), the next ones the faces with exactly one
    // transverse index in the ghost layers ("ring": they feed only the
    // transverse reconstructions of the edge EMFs: vbar and wave speeds);
    // faces with both transverse indices in the ghost layers are never read
    // and get no thread.  GS reads fluxes and face EMFs one layer into the
    // transverse ghosts, so it takes every face (normal [ng, n+ng],
    // transverse [0, N)).  ring_own (per block, bit 2 d + side for the six
    // block faces): when given, a ring face is computed only if its ghost
    // side lies beyond an outflow face of the instance; the others are
    // interior faces of the neighbouring block, whose face threads write
    // them through push_nbr (the neighbour across each of the six block
    // faces; with periodicity per dimension the relation is symmetric).
    const WgPos wp = wg_pos(wpb);
    const int64_t q = wp.q;
    const int e = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    int i, j, k;
    bool ring = false;
    {
        const int ng = G.ng, Rn = G.n[D] + 1;
        constexpr int T0 = D == 0 ? 1 : 0, T1 = D == 2 ? 1 : 2;
        const int n0 = G.n[T0], n1 = G.n[T1];
        int a = 0, b = 0, c = 0;   // (normal, t0, t1)
        if (o.emf != 0) {
            // GS: every face; FOFC widens the normal range by one face on each side
            const int N0 = G.N[T0], N1 = G.N[T1];
            const int Rg = Rn + 2 * o.fofc;
            if (e >= Rg * N0 * N1) return;
            int r;
            if (D == 0) { udivmod(e, Rg, r, a); udivmod(r, N0, c, b); }
            else if (D == 1) { udivmod(e, N0, r, b); udivmod(r, Rg, c, a); }
            else { udivmod(e, N0, r, b); udivmod(r, N1, a, c); }
            a -= o.fofc;
        } else {
            const int nint = Rn * n0 * n1, g2 = 2 * ng;
            const int nA = Rn * g2 * n1, nB = Rn * n0 * g2;
            auto ghost = [&](int g, int nn) { return g < ng ? g : nn + g; };
            int r;
            if (e < nint) {
                if (D == 0) { udivmod(e, Rn, r, a); udivmod(r, n0, c, b); }
                else if (D == 1) { udivmod(e, n0, r, b); udivmod(r, Rn, c, a); }
                else { udivmod(e, n0, r, b); udivmod(r, n1, a, c); }
                b += ng; c += ng;
            } else if (e < nint + nA) {          // t0 ghost, t1 interior
                const int x = e - nint;
                ring = true;
                int gl;
                if (D == 0) { udivmod(x, Rn, r, a); udivmod(r, g2, c, gl); }
                else if (D == 1) { udivmod(x, g2, r, gl); udivmod(r, Rn, c, a); }
                else { udivmod(x, g2, r, gl); udivmod(r, n1, a, c); }
                c += ng;
                if (ring_own && !((ring_own[q] >> (2 * T0 + (gl >= ng))) & 1)) return;
                b = ghost(gl, n0);
            } else if (e < nint + nA + nB) {     // t0 interior, t1 ghost
                const int x = e - nint - nA;
                ring = true;
                int gl;
                if (D == 0) { udivmod(x, Rn, r, a); udivmod(r, n0, gl, b); }
                else if (D == 1) { udivmod(x, n0, r, b); udivmod(r, Rn, gl, a); }
                else { udivmod(x, n0, r, b); udivmod(r, g2, a, gl); }
                b += ng;
                if (ring_own && !((ring_own[q] >> (2 * T1 + (gl >= ng))) & 1)) return;
                c = ghost(gl, n1);
            } else {
                return;
            }
        }
        a += ng;
        if (D == 0) { i = a; j = b; k = c; }
        else if (D == 1) { i = b; j = a; k = c; }
        else { i = b; j = c; k = a; }
        // FOFC correction pass: only the tagged faces of the normal range
        // [ng, n+ng] (evolve.cpp:646-702)
        if (only && (a < ng || a > ng + G.n[D] || !only[q * G.nface[D] + fidx<D>(G, i, j, k)])) return;
    }
    const int64_t nc = G.ncell;
    const int64_t c0 = cidx(G, i, j, k);
    const int64_t st = D == 0 ? 1 : (D == 1 ? G.N[0] : (int64_t)G.N[0] * G.N[1]);
    const double* A = aux + q * NAUX * nc;
    // flat metric: the six stencil cells along D (var 0).  With push_nbr (the
    // second and third RK stages, whose gathers leave the face slabs alone)
    // a full face takes a cell beyond a block face from the neighbouring
    // block's interior (normal index shifted by one block extent); beyond an
    // outflow face (no neighbour) and on ring faces, from its own ghosts,
    // which run their own c2p or copy one that does.  (The curved kernels,
    // which never get push_nbr, keep one row pointer per variable: six
    // pointers cost them 7% on x faces.)
    const double* Pt[6];
    if constexpr (FLAT) {
        const int an = D == 0 ? i : (D == 1 ? j : k), nD = G.n[D];
        int32_t nl = -1, nh = -1;
        if (push_nbr && !ring) { nl = push_nbr[q * 6 + 2 * D]; nh = push_nbr[q * 6 + 2 * D + 1]; }
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int ap = an + t - 3;
            int64_t qb = q;
            int sh = 0;
            if (ap < G.ng && nl >= 0) { qb = nl; sh = nD; }
            else if (ap >= nD + G.ng && nh >= 0) { qb = nh; sh = -nD; }
            Pt[t] = aux + qb * NAUX * nc + c0 + (int64_t)(t - 3 + sh) * st;
        }
    }

    Met mf;
    bool excise_override = false;
    if constexpr (FLAT) {
        mf.flat();
    } else if (fmc) {
        // the face-metric cache: only the inverse component this direction
        // uses (cmcp, the LLF speed limit) is kept
        const int64_t nfd = G.nface[D];
        const double* FM = fmc + q * FMC * nfd + fidx<D>(G, i, j, k);
        for (int a = 0; a < 6; ++a) { mf.g[a] = FM[a * nfd]; mf.gi[a] = 0.0; }
        mf.gi[D == 0 ? 0 : (D == 1 ? 3 : 5)] = FM[6 * nfd];
        for (int a = 0; a < 3; ++a) mf.beta[a] = FM[(7 + a) * nfd];
        mf.alp = FM[10 * nfd];
        mf.sqrtg = FM[11 * nfd];
        excise_override = FM[12 * nfd] != 0.0;
    } else {
        face_metric(mstate + q * (int64_t)mnv * nc, nc, c0, st, o.ex.alp_ex, mf, excise_override);
    }
    // reconstruction of RHO Z(3) YE THERMO ENT B(3)
    const int th_aux = o.thermo ? PRESS_ : TEMP_;
    const int ridx[10] = {RHO_, ZVECX_, ZVECY_, ZVECZ_, YE_, th_aux, ENTROPY_, BX_, BY_, BZ_};
    const int rloc[10] = {RHOL, ZXL, ZYL, ZZL, YEL, o.thermo ? PRESSL : TEMPL, ENTL, BXL, BYL, BZL};
    double pL[NPRIM], pR[NPRIM];
    if constexpr (PF) {
        double us[10][6];
#pragma unroll
        for (int v = 0; v < 10; ++v) {
            if (v == 7 + D) continue;
            if constexpr (FLAT) {
                const int64_t ov = ridx[v] * nc;
#pragma unroll
                for (int t = 0; t < 6; ++t) us[v][t] = Pt[t][ov];
            } else {
                const double* P = A + ridx[v] * nc + c0;
#pragma unroll
                for (int t = 0; t < 6; ++t) us[v][t] = P[(t - 3) * st];
            }
        }
#pragma unroll
        for (int v = 0; v < 10; ++v) {
            if (v == 7 + D || (ring && (v == 4 || v == 6))) { pL[rloc[v]] = 0.0; pR[rloc[v]] = 0.0; continue; }
            double uL, uR;
            if (excise_override) { uL = us[v][2]; uR = us[v][3]; }
            else recon<REC>(us[v][0], us[v][1], us[v][2], us[v][3], us[v][4], us[v][5], uL, uR);
            pL[rloc[v]] = uL; pR[rloc[v]] = uR;
        }
    } else {
#pragma unroll
        for (int v = 0; v < 10; ++v) {
            // Work whose result is never read: the normal B component is replaced
            // by the face value below, and a ring face (UCT velocities and wave
            // speeds only) reads neither Ye nor the entropy.
            if (v == 7 + D || (ring && (v == 4 || v == 6))) { pL[rloc[v]] = 0.0; pR[rloc[v]] = 0.0; continue; }
            double uL, uR;
            if constexpr (FLAT) {
                const int64_t ov = ridx[v] * nc;
                recon<REC>(Pt[0][ov], Pt[1][ov], Pt[2][ov], Pt[3][ov], Pt[4][ov], Pt[5][ov], uL, uR);
            } else {
                const double* P = A + ridx[v] * nc + c0;
                if (excise_override) { uL = P[-st]; uR = P[0]; }
                else recon<REC>(P[-3 * st], P[-2 * st], P[-st], P[0], P[st], P[2 * st], uL, uR);
            }
            pL[rloc[v]] = uL; pR[rloc[v]] = uR;
        }
    }
    double Bn = Bd[q * G.nface[D] + fidx<D>(G, i, j, k)] / mf.sqrtg;
    pL[BXL + D] = Bn; pR[BXL + D] = Bn;
    unsigned ee = 0;
    if (o.thermo) {
        double h, s;
        pL[EPSL] = eos.eps_from_press(h, pL[CS2L], pL[TEMPL], s, pL[PRESSL], pL[RHOL], ee);
        pR[EPSL] = eos.eps_from_press(h, pR[CS2L], pR[TEMPL], s, pR[PRESSL], pR[RHOL], ee);
    } else {
        pL[PRESSL] = eos.press_from_temp(pL[EPSL], pL[CS2L], pL[TEMPL], pL[RHOL], ee);
        pR[PRESSL] = eos.press_from_temp(pR[EPSL], pR[CS2L], pR[TEMPL], pR[RHOL], ee);
    }
    // compute_mhd_fluxes (grmhd.hh:774-970)
    const double *zl = &pL[ZXL], *zr = &pR[ZXL], *Bl = &pL[BXL], *Br = &pR[BXL];
    double wl = lorentz<FLAT>(mf.g, zl), wr = lorentz<FLAT>(mf.g, zr);
    double bl[4], br[4], b2l, b2r;
    smallb<FLAT>(mf, Bl, zl, wl, bl, b2l);
    smallb<FLAT>(mf, Br, zr, wr, br, b2r);
    double vtl[3], vtr[3];
    vtilde(mf, zl, wl, vtl);
    vtilde(mf, zr, wr, vtr);
    double cml, cpl, cmr, cpr, cmin, cmax;
    cmcp<FLAT>(mf, pL[RHOL], pL[PRESSL], pL[EPSL], pL[CS2L], wl, b2l, vtl, D, cml, cpl);
    cmcp<FLAT>(mf, pR[RHOL], pR[PRESSL], pR[EPSL], pR[CS2L], wr, b2r, vtr, D, cmr, cpr);
    if (o.riemann == 1) {
        cmax = fmax(fmax(fabs(cml), fabs(cmr)), fmax(fabs(cpl), fabs(cpr)));
        if (fmin(pL[RHOL], pR[RHOL]) < o.c_limit) {
            double gdd = D == 0 ? mf.gi[0] : (D == 1 ? mf.gi[3] : mf.gi[5]);
            cmax = fmax(cmax, mf.alp * sqrt(gdd) + fabs(mf.beta[D]));
        }
        cmin = cmax;
    } else {
        cmin = -fmin(0., fmin(cml, cmr));
        cmax = fmax(0., fmax(cpl, cpr));
    }
    if (cmin < 1e-12 && cmax < 1e-12) { cmin = 1; cmax = 1; }
    const double ihll = 1.0 / (cmax + cmin);   // shared by the HLL fluxes of this face
    const int64_t fo = fidx<D>(G, i, j, k);
    const int64_t nf = G.nface[D];
    double* V = vbar + q * 4 * nf + fo;
    constexpr int J0 = D == 0 ? 1 : 0, J1 = D == 2 ? 1 : 2;
    double vb0 = 0.0, vb1 = 0.0;
    if (o.emf == 0) {  // UCT: HLL-averaged transverse vtilde + wave speeds, slots strided by nf
        vb0 = hll_r(vtl[J0], vtr[J0], 0, 0, cmin, cmax, ihll);
        vb1 = hll_r(vtl[J1], vtr[J1], 0, 0, cmin, cmax, ihll);
        V[0] = vb0; V[nf] = vb1;
        V[2 * nf] = cmin; V[3 * nf] = cmax;
        if (ring) return;
    }
    double Dl, taul, Sl[3], Esl, Dr, taur, Sr[3], Esr;
    conserved<FLAT>(mf, pL[RHOL], pL[PRESSL], pL[EPSL], zl, pL[ENTL], wl, b2l, bl, Dl, taul, Sl, Esl);
    conserved<FLAT>(mf, pR[RHOL], pR[PRESSL], pR[EPSL], zr, pR[ENTL], wr, b2r, br, Dr, taur, Sr, Esr);
    // fluxes along D (grmhd_subexpressions.hh:97-156)
    double bsh[3]; lower3(mf.g, mf.beta[0], mf.beta[1], mf.beta[2], bsh);
    double ia2 = 1. / (mf.alp * mf.alp);
    double Fl[7], Fr[7];
    {
        double rhoh = b2l + pL[EPSL] * pL[RHOL] + pL[PRESSL] + pL[RHOL];
        double ps = 0.5 * b2l + pL[PRESSL];
        double W2rh = wl * wl * rhoh;
        double fD = vtl[D] * Dl;
        Fl[0] = fD;
        double blo[3]; lower3<FLAT>(mf.g, bl[1], bl[2], bl[3], blo);
        double vlo[3]; lower3<FLAT>(mf.g, vtl[0], vtl[1], vtl[2], vlo);
        if constexpr (FLAT) {
            Fl[4] = W2rh * vtl[D] - bl[0] * bl[1 + D] - fD;
            const double c = vtl[D] * W2rh;
            for (int a = 0; a < 3; ++a)
                Fl[1 + a] = a == D ? -bl[1 + D] * blo[a] + ps + c * vlo[a] : -bl[1 + D] * blo[a] + c * vlo[a];
        } else {
            Fl[4] = W2rh * vtl[D] + ps * mf.beta[D] - mf.alp * mf.alp * bl[0] * bl[1 + D] - fD;
            const double c = vtl[D] * W2rh * ia2;
            for (int a = 0; a < 3; ++a)
                Fl[1 + a] = mf.alp * (-bl[1 + D] * (blo[a] + bl[0] * bsh[a]) + (a == D ? ps : 0.0) + c * (vlo[a] + bsh[a]));
        }
        Fl[6] = vtl[D] * Esl;
        Fl[5] = pL[YEL] * fD;
    }
    {
        double rhoh = b2r + pR[EPSL] * pR[RHOL] + pR[PRESSL] + pR[RHOL];
        double ps = 0.5 * b2r + pR[PRESSL];
        double W2rh = wr * wr * rhoh;
        double fD = vtr[D] * Dr;
        Fr[0] = fD;
        double blo[3]; lower3<FLAT>(mf.g, br[1], br[2], br[3], blo);
        double vlo[3]; lower3<FLAT>(mf.g, vtr[0], vtr[1], vtr[2], vlo);
        if constexpr (FLAT) {
            Fr[4] = W2rh * vtr[D] - br[0] * br[1 + D] - fD;
            const double c = vtr[D] * W2rh;
            for (int a = 0; a < 3; ++a)
                Fr[1 + a] = a == D ? -br[1 + D] * blo[a] + ps + c * vlo[a] : -br[1 + D] * blo[a] + c * vlo[a];
        } else {
            Fr[4] = W2rh * vtr[D] + ps * mf.beta[D] - mf.alp * mf.alp * br[0] * br[1 + D] - fD;
            const double c = vtr[D] * W2rh * ia2;
            for (int a = 0; a < 3; ++a)
                Fr[1 + a] = mf.alp * (-br[1 + D] * (blo[a] + br[0] * bsh[a]) + (a == D ? ps : 0.0) + c * (vlo[a] + bsh[a]));
        }
        Fr[6] = vtr[D] * Esr;
        Fr[5] = pR[YEL] * fD;
    }
    double* F = flux + q * NHRSC * nf + fo;
    const double sg = mf.sqrtg;
    F[DENS_ * nf] = sg * hll_r(Fl[0], Fr[0], Dl, Dr, cmin, cmax, ihll);
    F[ENTROPYSTAR_ * nf] = sg * hll_r(Fl[6], Fr[6], Esl, Esr, cmin, cmax, ihll);
    F[TAU_ * nf] = sg * hll_r(Fl[4], Fr[4], taul, taur, cmin, cmax, ihll);
    F[SX_ * nf] = sg * hll_r(Fl[1], Fr[1], Sl[0], Sr[0], cmin, cmax, ihll);
    F[SY_ * nf] = sg * hll_r(Fl[2], Fr[2], Sl[1], Sr[1], cmin, cmax, ihll);
    F[SZ_ * nf] = sg * hll_r(Fl[3], Fr[3], Sl[2], Sr[2], cmin, cmax, ihll);
    F[YESTAR_ * nf] = sg * hll_r(Fl[5], Fr[5], pL[YEL] * Dl, pR[YEL] * Dr, cmin, cmax, ihll);
    if (o.emf == 1) {  // GS face EMFs (grmhd.hh:937-969)
        if (D == 0) {
            double a = -Bl[0] * vtl[2] + Bl[2] * vtl[0], b = -Br[0] * vtr[2] + Br[2] * vtr[0];
            V[0] = sg * hll_r(a, b, Bl[2], Br[2], cmin, cmax, ihll);
            a = -Bl[0] * vtl[1] + Bl[1] * vtl[0]; b = -Br[0] * vtr[1] + Br[1] * vtr[0];
            V[nf] = -sg * hll_r(a, b, Bl[1], Br[1], cmin, cmax, ihll);
        } else if (D == 1) {
            double a = -Bl[1] * vtl[2] + Bl[2] * vtl[1], b = -Br[1] * vtr[2] + Br[2] * vtr[1];
            V[0] = -sg * hll_r(a, b, Bl[2], Br[2], cmin, cmax, ihll);
            a = Bl[0] * vtl[1] - Bl[1] * vtl[0]; b = Br[0] * vtr[1] - Br[1] * vtr[0];
            V[nf] = sg * hll_r(a, b, Bl[0], Br[0], cmin, cmax, ihll);
        } else {
            double a = Bl[1] * vtl[2] - Bl[2] * vtl[1], b = Br[1] * vtr[2] - Br[2] * vtr[1];
            V[0] = sg * hll_r(a, b, Bl[1], Br[1], cmin, cmax, ihll);
            a = Bl[0] * vtl[2] - Bl[2] * vtl[0]; b = Br[0] * vtr[2] - Br[2] * vtr[0];
            V[nf] = -sg * hll_r(a, b, Bl[0], Br[0], cmin, cmax, ihll);
        }
    }
    if (push_nbr) {
        // stages with shared ring faces: this interior face is a ring face of the
        // neighbour across each block face within ngz of it (transverse
        // index shifted by one block extent)
        const int ng = G.ng;
        constexpr int T0 = D == 0 ? 1 : 0, T1 = D == 2 ? 1 : 2;
        const int loc[3] = {i, j, k};
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int td = t == 0 ? T0 : T1;
            const int nt = G.n[td], it = loc[td] - ng;
#pragma unroll
            for (int hi = 0; hi < 2; ++hi) {
                if (hi ? it < nt - ng : it >= ng) continue;
                const int32_t qt = push_nbr[q * 6 + 2 * td + hi];
                if (qt < 0) continue;
                int l2[3] = {loc[0], loc[1], loc[2]};
                l2[td] = hi ? it - nt + ng : it + nt + ng;
                double* T = vbar + (int64_t)qt * 4 * nf + fidx<D>(G, l2[0], l2[1], l2[2]);
                T[0] = vb0; T[nf] = vb1; T[2 * nf] = cmin; T[3 * nf] = cmax;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// UCT edge EMFs.  Face arrays: B faces (1 var), vbar (4 vars, stride nf).
// ---------------------------------------------------------------------------
template <int REC>
GHD void rec6(const double* P, int64_t s, double& uL, double& uR) {
    recon<REC>(P[-3 * s], P[-2 * s], P[-s], P[0], P[s], P[2 * s], uL, uR);
}

template <int REC>
__global__ __launch_bounds__(256) void k_emf_uct(DevGrid G, const double* __restrict__ Bx,
                                                 const double* __restrict__ By, const double* __restrict__ Bz,
                                                 const double* __restrict__ vx, const double* __restrict__ vy,
                                                 const double* __restrict__ vz, double* __restrict__ emf, int wpb) {
    // wpb workgroups per block over its (n+1)^3 sites (see wg_pos)
    const int R0 = G.n[0] + 1, R1 = G.n[1] + 1, R2 = G.n[2] + 1;
    const WgPos wp = wg_pos(wpb);
    const int r = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (r >= R0 * R1 * R2) return;
    const int64_t q = wp.q;
    int kz, rxy, jy, ix0;
    udivmod(r, R0 * R1, kz, rxy);
    udivmod(rxy, R0, jy, ix0);
    const int i = G.ng + ix0, j = G.ng + jy, k = G.ng + kz;
    const int64_t nfx = G.nface[0], nfy = G.nface[1], nfz = G.nface[2];
    const double* bx = Bx + q * nfx; const double* by = By + q * nfy; const double* bz = Bz + q * nfz;
    const double* Vx = vx + q * 4 * nfx; const double* Vy = vy + q * 4 * nfy; const double* Vz = vz + q * 4 * nfz;
    // strides of the face arrays along each axis
    const int64_t sx_x = 1, sx_y = G.N[0] + 1, sx_z = (int64_t)(G.N[0] + 1) * G.N[1];
    const int64_t sy_x = 1, sy_y = G.N[0], sy_z = (int64_t)G.N[0] * (G.N[1] + 1);
    const int64_t sz_x = 1, sz_y = G.N[0], sz_z = (int64_t)G.N[0] * G.N[1];
    double* E = emf + q * 3 * G.nedge;
    const int64_t eo = eidx(G, i, j, k);
    if (i < G.n[0] + G.ng) {  // E^x at (i, j-1/2, k-1/2)
        int64_t fy = fidx<1>(G, i, j, k), fz = fidx<2>(G, i, j, k);
        double ByL, ByR, vzL, vzR, BzL, BzR, vyL, vyR;
        rec6<REC>(by + fy, sy_z, ByL, ByR);
        rec6<REC>(Vy + nfy + fy, sy_z, vzL, vzR);
        rec6<REC>(bz + fz, sz_y, BzL, BzR);
        rec6<REC>(Vz + nfz + fz, sz_y, vyL, vyR);
        double cmin_y = fmax(Vy[2 * nfy + fy], Vy[2 * nfy + fy - sy_z]);
        double cmax_y = fmax(Vy[3 * nfy + fy], Vy[3 * nfy + fy - sy_z]);
        double cmin_z = fmax(Vz[2 * nfz + fz], Vz[2 * nfz + fz - sz_y]);
        double cmax_z = fmax(Vz[3 * nfz + fz], Vz[3 * nfz + fz - sz_y]);
        E[eo] = hll(vzL * ByL, vzR * ByR, ByL, ByR, cmin_z, cmax_z) - hll(vyL * BzL, vyR * BzR, BzL, BzR, cmin_y, cmax_y);
    }
    if (j < G.n[1] + G.ng) {  // E^y at (i-1/2, j, k-1/2)
        int64_t fz = fidx<2>(G, i, j, k), fx = fidx<0>(G, i, j, k);
        double BzL, BzR, vxL, vxR, BxL, BxR, vzL, vzR;
        rec6<REC>(bz + fz, sz_x, BzL, BzR);
        rec6<REC>(Vz + fz, sz_x, vxL, vxR);
        rec6<REC>(bx + fx, sx_z, BxL, BxR);
        rec6<REC>(Vx + nfx + fx, sx_z, vzL, vzR);
        double cmin_z = fmax(Vz[2 * nfz + fz], Vz[2 * nfz + fz - sz_x]);
        double cmax_z = fmax(Vz[3 * nfz + fz], Vz[3 * nfz + fz - sz_x]);
        double cmin_x = fmax(Vx[2 * nfx + fx], Vx[2 * nfx + fx - sx_z]);
        double cmax_x = fmax(Vx[3 * nfx + fx], Vx[3 * nfx + fx - sx_z]);
        E[G.nedge + eo] = hll(vxL * BzL, vxR * BzR, BzL, BzR, cmin_x, cmax_x) - hll(vzL * BxL, vzR * BxR, BxL, BxR, cmin_z, cmax_z);
    }
    if (k < G.n[2] + G.ng) {  // E^z at (i-1/2, j-1/2, k)
        int64_t fx = fidx<0>(G, i, j, k), fy = fidx<1>(G, i, j, k);
        double BxL, BxR, vyL, vyR, ByL, ByR, vxL, vxR;
        rec6<REC>(bx + fx, sx_y, BxL, BxR);
        rec6<REC>(Vx + fx, sx_y, vyL, vyR);
        rec6<REC>(by + fy, sy_x, ByL, ByR);
        rec6<REC>(Vy + fy, sy_x, vxL, vxR);
        double cmin_x = fmax(Vx[2 * nfx + fx], Vx[2 * nfx + fx - sx_y]);
        double cmax_x = fmax(Vx[3 * nfx + fx], Vx[3 * nfx + fx - sx_y]);
        double cmin_y = fmax(Vy[2 * nfy + fy], Vy[2 * nfy + fy - sy_x]);
        double cmax_y = fmax(Vy[3 * nfy + fy], Vy[3 * nfy + fy - sy_x]);
        E[2 * G.nedge + eo] = hll(vyL * BxL, vyR * BxR, BxL, BxR, cmin_y, cmax_y) - hll(vxL * ByL, vxR * ByR, ByL, ByR, cmin_x, cmax_x);
    }
}

// Gardiner-Stone edge EMF (grmhd_helpers.hh:492-627).  Eface slots 0/1 of each
// face direction, Ecenter (3 comps, cell arrays), DENS fluxes.
__global__ __launch_bounds__(256) void k_emf_gs(DevGrid G, const double* __restrict__ ef0, const double* __restrict__ ef1,
                                                const double* __restrict__ ef2, const double* __restrict__ ec,
                                                const double* __restrict__ f0, const double* __restrict__ f1,
                                                const double* __restrict__ f2, double* __restrict__ emf, int wpb,
                                                const uint8_t* __restrict__ only, int wide) {
    // wpb workgroups per block over its (n+1+2 wide)^3 sites from ng - wide
    // (see wg_pos); only != null: FOFC correction of the tagged edges
    // (evolve.cpp:780-808); wide = 1: one ghost layer of edges more on every
    // side (grmhd_params.fofc_ghost_emf)
    const int R0 = G.n[0] + 1 + 2 * wide, R1 = G.n[1] + 1 + 2 * wide, R2 = G.n[2] + 1 + 2 * wide;
    const WgPos wp = wg_pos(wpb);
    const int r = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (r >= R0 * R1 * R2) return;
    const int64_t q = wp.q;
    int kz, rxy, jy, ix0;
    udivmod(r, R0 * R1, kz, rxy);
    udivmod(rxy, R0, jy, ix0);
    const int i = G.ng - wide + ix0, j = G.ng - wide + jy, k = G.ng - wide + kz;
    const int64_t nfx = G.nface[0], nfy = G.nface[1], nfz = G.nface[2], nc = G.ncell;
    auto EF = [&](int d, int c, int ii, int jj, int kk) {
        const double* base = d == 0 ? ef0 : (d == 1 ? ef1 : ef2);
        int64_t nf = G.nface[d];
        return base[q * 4 * nf + c * nf + fidx_d(G, d, ii, jj, kk)];
    };
    auto EC = [&](int c, int ii, int jj, int kk) { return ec[q * 3 * nc + c * nc + cidx(G, ii, jj, kk)]; };
    auto FD = [&](int d, int ii, int jj, int kk) {
        const double* base = d == 0 ? f0 : (d == 1 ? f1 : f2);
        int64_t nf = G.nface[d];
        return base[q * NHRSC * nf + DENS_ * nf + fidx_d(G, d, ii, jj, kk)];
    };
    (void)nfx; (void)nfy; (void)nfz;
    double* E = emf + q * 3 * G.nedge;
    const int64_t eo = eidx(G, i, j, k);
    if (i < G.n[0] + G.ng + wide && (!only || only[(q * 3 + 0) * G.nedge + eo])) {
        double Exzf = EF(2, 0, i, j, k), Exzfm = EF(2, 0, i, j - 1, k), Exyf = EF(1, 0, i, j, k), Exyfm = EF(1, 0, i, j, k - 1);
        double NE = EC(0, i, j, k), NW = EC(0, i, j - 1, k), SE = EC(0, i, j, k - 1), SW = EC(0, i, j - 1, k - 1);
        double Eavg = 0.25 * (Exzf + Exzfm + Exyf + Exyfm);
        double Sy = copysign(1., FD(1, i, j, k)), Sym = copysign(1., FD(1, i, j, k - 1));
        double dN = (1. - Sy) * (NE - Exzf) + (1. + Sy) * (NW - Exzfm);
        double dS = (1. - Sym) * (Exzf - SE) + (1. + Sym) * (Exzfm - SW);
        double Sz = copysign(1., FD(2, i, j, k)), Szm = copysign(1., FD(2, i, j - 1, k));
        double dE = (1. - Sz) * (NE - Exyf) + (1. + Sz) * (SE - Exyfm);
        double dW = (1. - Szm) * (Exyf - NW) + (1. + Szm) * (Exyfm - SW);
        E[eo] = Eavg + (1. / 8.) * (dS - dN) + (1. / 8.) * (dW - dE);
    }
    if (j < G.n[1] + G.ng + wide && (!only || only[(q * 3 + 1) * G.nedge + eo])) {
        double Eyzf = EF(2, 1, i, j, k), Eyzfm = EF(2, 1, i - 1, j, k), Eyxf = EF(0, 0, i, j, k), Eyxfm = EF(0, 0, i, j, k - 1);
        double NE = EC(1, i, j, k), NW = EC(1, i - 1, j, k), SE = EC(1, i, j, k - 1), SW = EC(1, i - 1, j, k - 1);
        double Eavg = 0.25 * (Eyzf + Eyzfm + Eyxf + Eyxfm);
        double Sx = copysign(1., FD(0, i, j, k)), Sxm = copysign(1., FD(0, i, j, k - 1));
        double dN = (1. - Sx) * (NE - Eyzf) + (1. + Sx) * (NW - Eyzfm);
        double dS = (1. - Sxm) * (Eyzf - SE) + (1. + Sxm) * (Eyzfm - SW);
        double Sz = copysign(1., FD(2, i, j, k)), Szm = copysign(1., FD(2, i - 1, j, k));
        double dE = (1. - Sz) * (NE - Eyxf) + (1. + Sz) * (SE - Eyxfm);
        double dW = (1. - Szm) * (Eyxf - NW) + (1. + Szm) * (Eyxfm - SW);
        E[G.nedge + eo] = Eavg + (1. / 8.) * (dS - dN) + (1. / 8.) * (dW - dE);
    }
    if (k < G.n[2] + G.ng + wide && (!only || only[(q * 3 + 2) * G.nedge + eo])) {
        double Ezyf = EF(1, 1, i, j, k), Ezyfm = EF(1, 1, i - 1, j, k), Ezxf = EF(0, 1, i, j, k), Ezxfm = EF(0, 1, i, j - 1, k);
        double NE = EC(2, i, j, k), NW = EC(2, i - 1, j, k), SE = EC(2, i, j - 1, k), SW = EC(2, i - 1, j - 1, k);
        double Eavg = 0.25 * (Ezyf + Ezyfm + Ezxf + Ezxfm);
        double Sx = copysign(1., FD(0, i, j, k)), Sxm = copysign(1., FD(0, i, j - 1, k));
        double dN = (1. - Sx) * (NE - Ezyf) + (1. + Sx) * (NW - Ezyfm);
        double dS = (1. - Sxm) * (Ezyf - SE) + (1. + Sxm) * (Ezyfm - SW);
        double Sy = copysign(1., FD(1, i, j, k)), Sym = copysign(1., FD(1, i - 1, j, k));
        double dE = (1. - Sy) * (NE - Ezxf) + (1. + Sy) * (SE - Ezxfm);
        double dW = (1. - Sym) * (Ezxf - NW) + (1. + Sym) * (Ezxfm - SW);
        E[2 * G.nedge + eo] = Eavg + (1. / 8.) * (dS - dN) + (1. / 8.) * (dW - dE);
    }
}

// ---------------------------------------------------------------------------
// Fused update: new = lin(A,B) + sqrtg dt dtfact S + dt dtfact divF (cells),
//               Bnew = lin(BA,BB) + dt dtfact curlE (faces).
// lin(A,B) = A when cb == 0, else ca A + cb B.
// ---------------------------------------------------------------------------
struct UpdArgs {
    double* nst; int nnv;            // new state, vars per block
    const double* ast; int anv;      // A
    const double* bst; int bnv;      // B (may be null)
    double ca, cb;
    const double* mstate; int mnv;   // metric source
    const double* aux;
    const double* flux[3];
    const double* emf;
    double* nB[3]; const double* aB[3]; const double* bB[3];
    const double* dt; double dtfact;
    const DevInst* inst; const DevBlk* blk;
    const double* mder;              // cached metric derivatives (30 per interior cell) or null
    int part;                        // 0: cells and faces; 1: cells only (fluxes, sources); 2: faces only (CT)
};

GHD double fdder(const double* P, int64_t s, double ih, int order) {
    if (order == 2) return (1.0 / 2.0) * ih * ((P[s] - P[-s]));
    if (order == 4) return (1.0 / 12.0) * ih * (+8.0 * (P[s] - P[-s]) - (P[2 * s] - P[-2 * s]));
    return (1.0 / 60.0) * ih * (+45.0 * (P[s] - P[-s]) - 9.0 * (P[2 * s] - P[-2 * s]) + (P[3 * s] - P[-3 * s]));
}

// geometric sources (same formulas as the reference generator output,
// grmhd_subexpressions.hh:158-316; derivation from T^{mu nu})
GHD void geom_sources(const Met& m, const double K[6], const double da[3], const double dg[18], const double db[9],
                      double rho, double press, double eps, const double B[3], const double z[3], double W,
                      double& stau, double ss[3]) {
    double b[4], b2; smallb(m, B, z, W, b, b2);
    double rhoh = b2 + eps * rho + press + rho, ps = press + 0.5 * b2;
    double ia2 = 1. / (m.alp * m.alp);
    double u[4] = {W / m.alp, z[0] - W * m.beta[0] / m.alp, z[1] - W * m.beta[1] / m.alp, z[2] - W * m.beta[2] / m.alp};
    const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double T00 = rhoh * u[0] * u[0] + ps * (-ia2) - b[0] * b[0];
    double T0i[3], Tij[6];
    for (int i = 0; i < 3; ++i) T0i[i] = rhoh * u[0] * u[1 + i] + ps * (m.beta[i] * ia2) - b[0] * b[1 + i];
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j)
            Tij[sidx[i][j]] = rhoh * u[1 + i] * u[1 + j] + ps * (m.gi[sidx[i][j]] - m.beta[i] * m.beta[j] * ia2) - b[1 + i] * b[1 + j];
    double Kb[3];
    for (int i = 0; i < 3; ++i) Kb[i] = K[sidx[i][0]] * m.beta[0] + K[sidx[i][1]] * m.beta[1] + K[sidx[i][2]] * m.beta[2];
    double bKb = m.beta[0] * Kb[0] + m.beta[1] * Kb[1] + m.beta[2] * Kb[2];
    double T0K = 0, TK = 0, Tda = 0;
    for (int i = 0; i < 3; ++i) {
        T0K += T0i[i] * Kb[i];
        for (int j = 0; j < 3; ++j) TK += Tij[sidx[i][j]] * K[sidx[i][j]];
        Tda += (T00 * m.beta[i] + T0i[i]) * da[i];
    }
    stau = m.alp * ((T00 * bKb + 2. * T0K + TK) - Tda);
    double bsh[3]; lower3(m.g, m.beta[0], m.beta[1], m.beta[2], bsh);
    double T0l[3];
    for (int i = 0; i < 3; ++i)
        T0l[i] = bsh[i] * T00 + m.g[sidx[i][0]] * T0i[0] + m.g[sidx[i][1]] * T0i[1] + m.g[sidx[i][2]] * T0i[2];
    for (int k = 0; k < 3; ++k) {
        const double* dgk = dg + 6 * k;
        double bdgb = 0, T0dgb = 0, Tdg = 0, T0db = 0;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                double d = dgk[sidx[i][j]];
                bdgb += m.beta[i] * m.beta[j] * d;
                T0dgb += T0i[i] * m.beta[j] * d;
                Tdg += Tij[sidx[i][j]] * d;
            }
            T0db += T0l[i] * db[i + 3 * k];
        }
        ss[k] = m.alp * (T00 * (0.5 * bdgb - m.alp * da[k]) + T0dgb + T0db + 0.5 * Tdg);
    }
}

// Derivatives of the Cowling metric at a cell (fill_deriv_* of
// grmhd.hh:232-235, with the reference's idx(0,q) in every direction):
// da[d], db[c + 3d], dg[b + 6d].
GHD void metric_ders(const double* M, int64_t nc, const int64_t sd[3], double ix, int fd_order, double da[3],
                     double db[9], double dg[18]) {
    for (int d = 0; d < 3; ++d) da[d] = fdder(M + ALP_ * nc, sd[d], ix, fd_order);
    for (int d = 0; d < 3; ++d)
        for (int cc = 0; cc < 3; ++cc) db[cc + 3 * d] = fdder(M + (BETAX_ + cc) * nc, sd[d], ix, fd_order);
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b < 6; ++b) dg[b + 6 * d] = fdder(M + (GXX_ + b) * nc, sd[d], ix, fd_order);
}

// The face-metric cache of direction D for every face whose 4-point stencil
// lies in the block array (normal index [2, N_D - 2]; wpb workgroups per
// block, see wg_pos).
template <int D>
__global__ __launch_bounds__(256) void k_face_metric(DevGrid G, const double* __restrict__ mstate, int mnv,
                                                     double alp_ex, double* __restrict__ out, int wpb) {
    const WgPos wp = wg_pos(wpb);
    const int64_t q = wp.q;
    const int e = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    const int64_t nf = G.nface[D];
    if (e >= nf) return;
    const int E0 = G.N[0] + (D == 0), E1 = G.N[1] + (D == 1);
    int r, i, j, k;
    udivmod(e, E0, r, i);
    udivmod(r, E1, k, j);
    const int a = D == 0 ? i : (D == 1 ? j : k);
    if (a < 2 || a > G.N[D] - 2) return;
    const int64_t nc = G.ncell;
    const int64_t st = D == 0 ? 1 : (D == 1 ? G.N[0] : (int64_t)G.N[0] * G.N[1]);
    Met mf;
    bool ex;
    face_metric(mstate + q * (int64_t)mnv * nc, nc, cidx(G, i, j, k), st, alp_ex, mf, ex);
    double* O = out + q * FMC * nf + e;
    for (int b = 0; b < 6; ++b) O[b * nf] = mf.g[b];
    O[6 * nf] = mf.gi[D == 0 ? 0 : (D == 1 ? 3 : 5)];
    for (int b = 0; b < 3; ++b) O[(7 + b) * nf] = mf.beta[b];
    O[10 * nf] = mf.alp;
    O[11 * nf] = mf.sqrtg;
    O[12 * nf] = ex ? 1.0 : 0.0;
}

// The 30 metric derivatives of every interior cell, [q][30][k][j][i].
__global__ __launch_bounds__(256) void k_metric_der(DevGrid G, const DevInst* __restrict__ inst,
                                                    const DevBlk* __restrict__ blk, const double* __restrict__ ms,
                                                    int mnv, int fd_order, double* __restrict__ out) {
    const int64_t per = (int64_t)G.n[0] * G.n[1] * G.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * G.nq) return;
    const int64_t q = t / per, e = t - q * per;
    const int i = G.ng + (int)(e % G.n[0]), j = G.ng + (int)((e / G.n[0]) % G.n[1]);
    const int k = G.ng + (int)(e / ((int64_t)G.n[0] * G.n[1]));
    const int64_t nc = G.ncell;
    const double* M = ms + q * (int64_t)mnv * nc + cidx(G, i, j, k);
    const int64_t sd[3] = {1, G.N[0], (int64_t)G.N[0] * G.N[1]};
    double da[3], db[9], dg[18];
    metric_ders(M, nc, sd, inst[blk[q].inst].idx[0], fd_order, da, db, dg);
    double* D = out + q * 30 * per + e;
    for (int d = 0; d < 3; ++d) D[d * per] = da[d];
    for (int b = 0; b < 9; ++b) D[(3 + b) * per] = db[b];
    for (int b = 0; b < 18; ++b) D[(12 + b) * per] = dg[b];
}

// MDER: the metric derivatives come from the static-metric cache a.mder
// (the finite differences are not compiled in, which frees registers).
template <bool FLAT, bool MDER = false>
__global__ __launch_bounds__(256) void k_update(DevGrid G, UpdArgs a, int fd_order, int wpb) {
    // wpb workgroups per block over its (n+1)^3 sites (see wg_pos)
    const int R0 = G.n[0] + 1, R1 = G.n[1] + 1, R2 = G.n[2] + 1;
    const WgPos wp = wg_pos(wpb);
    const int r = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (r >= R0 * R1 * R2) return;
    const int64_t q = wp.q;
    int kz, rxy, jy, ix0;
    udivmod(r, R0 * R1, kz, rxy);
    udivmod(rxy, R0, jy, ix0);
    const int i = G.ng + ix0, j = G.ng + jy, k = G.ng + kz;
    const DevInst& I = a.inst[a.blk[q].inst];
    const double ix = I.idx[0], iy = I.idx[1], iz = I.idx[2];
    const double dtf = (*a.dt) * a.dtfact;
    const bool ci = i < G.n[0] + G.ng, cj = j < G.n[1] + G.ng, ck = k < G.n[2] + G.ng;
    const int64_t nc = G.ncell;
    if (a.part != 2 && ci && cj && ck) {
        const int64_t c = cidx(G, i, j, k);
        double* Nn = a.nst + q * (int64_t)a.nnv * nc + c;
        const double* An = a.ast + q * (int64_t)a.anv * nc + c;
        const double* Bn = a.bst ? a.bst + q * (int64_t)a.bnv * nc + c : nullptr;
        double u[NHRSC];
#pragma unroll
        for (int v = 0; v < NHRSC; ++v) u[v] = (a.cb == 0.0) ? An[v * nc] : a.ca * An[v * nc] + a.cb * Bn[v * nc];
        if constexpr (!FLAT) {
            const double* M = a.mstate + q * (int64_t)a.mnv * nc + c;
            const double* X = a.aux + q * (int64_t)NAUX * nc + c;
            Met m;
            for (int b = 0; b < 6; ++b) m.g[b] = M[(GXX_ + b) * nc];
            m.alp = M[ALP_ * nc];
            for (int b = 0; b < 3; ++b) m.beta[b] = M[(BETAX_ + b) * nc];
            m.init();
            double z[3] = {X[ZVECX_ * nc], X[ZVECY_ * nc], X[ZVECZ_ * nc]};
            double B[3] = {X[BX_ * nc], X[BY_ * nc], X[BZ_ * nc]};
            double W = lorentz(m.g, z);
            double da[3], dg[18], db[9], K[6];
            if constexpr (MDER) {   // the static metric's derivatives, computed once by k_metric_der
                const int64_t per = (int64_t)G.n[0] * G.n[1] * G.n[2];
                const double* D = a.mder + q * 30 * per + ((k - G.ng) * (int64_t)G.n[1] + (j - G.ng)) * G.n[0] + (i - G.ng);
                for (int d = 0; d < 3; ++d) da[d] = D[d * per];
                for (int b = 0; b < 9; ++b) db[b] = D[(3 + b) * per];
                for (int b = 0; b < 18; ++b) dg[b] = D[(12 + b) * per];
            } else {
                const int64_t sd[3] = {1, G.N[0], (int64_t)G.N[0] * G.N[1]};
                metric_ders(M, nc, sd, ix, fd_order, da, db, dg);
            }
            for (int b = 0; b < 6; ++b) K[b] = M[(KXX_ + b) * nc];
            double stau, ss[3];
            geom_sources(m, K, da, dg, db, X[RHO_ * nc], X[PRESS_ * nc], X[EPS_ * nc], B, z, W, stau, ss);
            u[TAU_] += m.sqrtg * (*a.dt) * a.dtfact * stau;
            u[SX_] += m.sqrtg * (*a.dt) * a.dtfact * ss[0];
            u[SY_] += m.sqrtg * (*a.dt) * a.dtfact * ss[1];
            u[SZ_] += m.sqrtg * (*a.dt) * a.dtfact * ss[2];
        }
        const int64_t f0 = fidx<0>(G, i, j, k), f1 = fidx<1>(G, i, j, k), f2 = fidx<2>(G, i, j, k);
        const double* F0 = a.flux[0] + q * NHRSC * G.nface[0];
        const double* F1 = a.flux[1] + q * NHRSC * G.nface[1];
        const double* F2 = a.flux[2] + q * NHRSC * G.nface[2];
        const int64_t s1 = G.N[0], s2 = (int64_t)G.N[0] * (G.N[1] + 0);
#pragma unroll
        for (int v = 0; v < NHRSC; ++v) {
            double dFx = (F0[v * G.nface[0] + f0] - F0[v * G.nface[0] + f0 + 1]) * ix;
            double dFy = (F1[v * G.nface[1] + f1] - F1[v * G.nface[1] + f1 + s1]) * iy;
            double dFz = (F2[v * G.nface[2] + f2] - F2[v * G.nface[2] + f2 + s2]) * iz;
            u[v] += dtf * ((dFx + dFy) + dFz);
        }
#pragma unroll
        for (int v = 0; v < NHRSC; ++v) Nn[v * nc] = u[v];
    }
    if (a.part == 1) return;
    const double* E = a.emf + q * 3 * G.nedge;
    auto Ex = [&](int ii, int jj, int kk) { return E[eidx(G, ii, jj, kk)]; };
    auto Ey = [&](int ii, int jj, int kk) { return E[G.nedge + eidx(G, ii, jj, kk)]; };
    auto Ez = [&](int ii, int jj, int kk) { return E[2 * G.nedge + eidx(G, ii, jj, kk)]; };
    if (cj && ck) {  // x face (i in [ng, n+ng])
        int64_t f = q * G.nface[0] + fidx<0>(G, i, j, k);
        double b = (a.cb == 0.0) ? a.aB[0][f] : a.ca * a.aB[0][f] + a.cb * a.bB[0][f];
        a.nB[0][f] = b + dtf * ((Ey(i, j, k + 1) - Ey(i, j, k)) * iz + (Ez(i, j, k) - Ez(i, j + 1, k)) * iy);
    }
    if (ci && ck) {
        int64_t f = q * G.nface[1] + fidx<1>(G, i, j, k);
        double b = (a.cb == 0.0) ? a.aB[1][f] : a.ca * a.aB[1][f] + a.cb * a.bB[1][f];
        a.nB[1][f] = b + dtf * ((Ez(i + 1, j, k) - Ez(i, j, k)) * ix + (Ex(i, j, k) - Ex(i, j, k + 1)) * iz);
    }
    if (ci && cj) {
        int64_t f = q * G.nface[2] + fidx<2>(G, i, j, k);
        double b = (a.cb == 0.0) ? a.aB[2][f] : a.ca * a.aB[2][f] + a.cb * a.bB[2][f];
        a.nB[2][f] = b + dtf * ((Ex(i, j + 1, k) - Ex(i, j, k)) * iy + (Ey(i, j, k) - Ey(i + 1, j, k)) * ix);
    }
}

// Flat-metric update over pairs of sites along x (i0, i0 + 1 with i0 even).
// With even extents N[d] the cell arrays, the y- and z-face arrays (fluxes,
// face B) have even row widths, so the two values of a pair sit in one
// aligned 16-byte word and move with one access; the x-face and edge arrays
// (odd row widths) keep 8-byte accesses, three x faces serving the pair's
// two cells.  Per cell and face the arithmetic is that of k_update<true>,
// operation for operation.  All loads of a pair are issued before its first
// store (the stores may alias the loads: the third RK stage updates in
// place), which takes 194 VGPRs, 2 waves per SIMD: measured faster in the
// RK3 graph than k_update<true> and than a variant storing each variable as
// it completes (48 VGPRs, 8 waves).  pairs = pairs per row; wpb workgroups
// per block.
__global__ __launch_bounds__(256, 2) void k_update_pairs(DevGrid G, UpdArgs a, int pairs, int wpb) {
    const int R1 = G.n[1] + 1, R2 = G.n[2] + 1;
    const WgPos wp = wg_pos(wpb);
    const int r = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (r >= pairs * R1 * R2) return;
    const int64_t q = wp.q;
    int kz, rxy, jy, px;
    udivmod(r, pairs * R1, kz, rxy);
    udivmod(rxy, pairs, jy, px);
    const int ng = G.ng;
    const int i0 = (ng & ~1) + 2 * px, j = ng + jy, k = ng + kz;
    const DevInst& I = a.inst[a.blk[q].inst];
    const double ix = I.idx[0], iy = I.idx[1], iz = I.idx[2];
    const double dtf = (*a.dt) * a.dtfact;
    const int ilo = ng, ihi = G.n[0] + ng;             // cells [ilo, ihi), x faces [ilo, ihi]
    const bool c0 = i0 >= ilo && i0 < ihi, c1 = i0 + 1 >= ilo && i0 + 1 < ihi;
    const bool cj = j < G.n[1] + ng, ck = k < G.n[2] + ng;
    const int64_t nc = G.ncell;
    typedef HIP_vector_type<double, 2> d2;
    if (a.part != 2 && cj && ck && (c0 || c1)) {
        const int64_t c = cidx(G, i0, j, k);
        double* Nn = a.nst + q * (int64_t)a.nnv * nc + c;
        const double* An = a.ast + q * (int64_t)a.anv * nc + c;
        const double* Bn = a.bst ? a.bst + q * (int64_t)a.bnv * nc + c : nullptr;
        const int64_t f0 = fidx<0>(G, i0, j, k), f1 = fidx<1>(G, i0, j, k), f2 = fidx<2>(G, i0, j, k);
        const double* F0 = a.flux[0] + q * NHRSC * G.nface[0];
        const double* F1 = a.flux[1] + q * NHRSC * G.nface[1];
        const double* F2 = a.flux[2] + q * NHRSC * G.nface[2];
        const int64_t s1 = G.N[0], s2 = (int64_t)G.N[0] * G.N[1];
        double u0[NHRSC], u1[NHRSC];
#pragma unroll
        for (int v = 0; v < NHRSC; ++v) {
            const d2 x = *reinterpret_cast<const d2*>(An + v * nc);
            if (a.cb == 0.0) { u0[v] = x.x; u1[v] = x.y; }
            else {
                const d2 y = *reinterpret_cast<const d2*>(Bn + v * nc);
                u0[v] = a.ca * x.x + a.cb * y.x; u1[v] = a.ca * x.y + a.cb * y.y;
            }
        }
#pragma unroll
        for (int v = 0; v < NHRSC; ++v) {
            const double* X = F0 + v * G.nface[0] + f0;
            const double xa = X[0], xb = X[1], xc = X[2];
            const d2 ya = *reinterpret_cast<const d2*>(F1 + v * G.nface[1] + f1);
            const d2 yb = *reinterpret_cast<const d2*>(F1 + v * G.nface[1] + f1 + s1);
            const d2 za = *reinterpret_cast<const d2*>(F2 + v * G.nface[2] + f2);
            const d2 zb = *reinterpret_cast<const d2*>(F2 + v * G.nface[2] + f2 + s2);
            {
                double dFx = (xa - xb) * ix, dFy = (ya.x - yb.x) * iy, dFz = (za.x - zb.x) * iz;
                u0[v] += dtf * ((dFx + dFy) + dFz);
            }
            {
                double dFx = (xb - xc) * ix, dFy = (ya.y - yb.y) * iy, dFz = (za.y - zb.y) * iz;
                u1[v] += dtf * ((dFx + dFy) + dFz);
            }
        }
#pragma unroll
        for (int v = 0; v < NHRSC; ++v) {
            if (c0 && c1) *reinterpret_cast<d2*>(Nn + v * nc) = d2{u0[v], u1[v]};
            else if (c0) Nn[v * nc] = u0[v];
            else Nn[v * nc + 1] = u1[v];
        }
    }
    if (a.part == 1) return;
    const double* E = a.emf + q * 3 * G.nedge;
    auto Ex = [&](int ii, int jj, int kk) { return E[eidx(G, ii, jj, kk)]; };
    auto Ey = [&](int ii, int jj, int kk) { return E[G.nedge + eidx(G, ii, jj, kk)]; };
    auto Ez = [&](int ii, int jj, int kk) { return E[2 * G.nedge + eidx(G, ii, jj, kk)]; };
    if (cj && ck) {  // x faces i0, i0 + 1 within [ilo, ihi]
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int i = i0 + t;
            if (i < ilo || i > ihi) continue;
            int64_t f = q * G.nface[0] + fidx<0>(G, i, j, k);
            double b = (a.cb == 0.0) ? a.aB[0][f] : a.ca * a.aB[0][f] + a.cb * a.bB[0][f];
            a.nB[0][f] = b + dtf * ((Ey(i, j, k + 1) - Ey(i, j, k)) * iz + (Ez(i, j, k) - Ez(i, j + 1, k)) * iy);
        }
    }
    if (ck && (c0 || c1)) {  // y faces (cells' x positions, j in [ng, n+ng])
        const int64_t f = q * G.nface[1] + fidx<1>(G, i0, j, k);
        d2 b = *reinterpret_cast<const d2*>(a.aB[1] + f);
        if (a.cb != 0.0) {
            const d2 bb = *reinterpret_cast<const d2*>(a.bB[1] + f);
            b = d2{a.ca * b.x + a.cb * bb.x, a.ca * b.y + a.cb * bb.y};
        }
        const double ez0 = Ez(i0, j, k), ez1 = Ez(i0 + 1, j, k), ez2 = Ez(i0 + 2, j, k);
        const double n0v = b.x + dtf * ((ez1 - ez0) * ix + (Ex(i0, j, k) - Ex(i0, j, k + 1)) * iz);
        const double n1v = b.y + dtf * ((ez2 - ez1) * ix + (Ex(i0 + 1, j, k) - Ex(i0 + 1, j, k + 1)) * iz);
        if (c0 && c1) *reinterpret_cast<d2*>(a.nB[1] + f) = d2{n0v, n1v};
        else if (c0) a.nB[1][f] = n0v;
        else a.nB[1][f + 1] = n1v;
    }
    if (cj && (c0 || c1)) {  // z faces
        const int64_t f = q * G.nface[2] + fidx<2>(G, i0, j, k);
        d2 b = *reinterpret_cast<const d2*>(a.aB[2] + f);
        if (a.cb != 0.0) {
            const d2 bb = *reinterpret_cast<const d2*>(a.bB[2] + f);
            b = d2{a.ca * b.x + a.cb * bb.x, a.ca * b.y + a.cb * bb.y};
        }
        const double n0v = b.x + dtf * ((Ex(i0, j + 1, k) - Ex(i0, j, k)) * iy + (Ey(i0, j, k) - Ey(i0 + 1, j, k)) * ix);
        const double n1v =
            b.y + dtf * ((Ex(i0 + 1, j + 1, k) - Ex(i0 + 1, j, k)) * iy + (Ey(i0 + 1, j, k) - Ey(i0 + 2, j, k)) * ix);
        if (c0 && c1) *reinterpret_cast<d2*>(a.nB[2] + f) = d2{n0v, n1v};
        else if (c0) a.nB[2][f] = n0v;
        else a.nB[2][f + 1] = n1v;
    }
}

// ---------------------------------------------------------------------------
// Ghost gather.  Index space per block: [cells | x faces | y faces | z faces].
// Each ghost element copies the element at the same global position (clamped
// for outflow, wrapped for periodic) from the owning block.
// ---------------------------------------------------------------------------
// First-order flux correction (evolve.cpp:456-811)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* __restrict__ p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

struct FofcArgs {
    const double* ast; int anv;        // the stage's linear combination lin = ca A + cb B
    const double* bst; int bnv;
    double ca, cb;
    const double* mstate; int mnv;     // metric source
    const double* flux[3];
    const double* emf;
    const double* oB[3];               // stage input face B
    const double* dt; double dtfact;
    const DevInst* inst; const DevBlk* blk;
    uint8_t* ftag[3]; uint8_t* etag;
};

// flag_fofc_cells (evolve.cpp:456-633): every cell of [ng-1, n+ng+1)^3 of
// every block.  The tentative state (lin + dt dtfact div F, the stage input
// face B advanced by the edge EMFs) is formed with separately rounded
// products and sums, the way the scalar reference code evaluates it, so that
// the discrete tests below see the same values.  A cell whose tentative D or
// tau leaves [min/M, M max] of the 27-cell neighbourhood of lin, or whose
// lenient dry-run c2p floors, tags its 6 faces and 12 edges (plain byte
// stores: every writer stores 1).
template <typename EOS, bool FLAT>
__global__ __launch_bounds__(256) void k_fofc_flag(DevGrid G, EOS eos, Opts o, FofcArgs a, int wpb) {
    const WgPos wp = wg_pos(wpb);
    const int m0 = G.n[0] + 2, m1 = G.n[1] + 2, m2 = G.n[2] + 2;
    const int r = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (r >= m0 * m1 * m2) return;
    const int64_t q = wp.q;
    const int kz = r / (m0 * m1), rxy = r - kz * (m0 * m1);
    const int i = G.ng - 1 + rxy % m0, j = G.ng - 1 + rxy / m0, k = G.ng - 1 + kz;
    const int64_t nc = G.ncell;
    const int64_t c = cidx(G, i, j, k);
    const DevInst& I = a.inst[a.blk[q].inst];
    const double ix = I.idx[0], iy = I.idx[1], iz = I.idx[2];
    const double dt = *a.dt;
    const double dte = __dmul_rn(dt, a.dtfact);
    const double* A = a.ast + q * (int64_t)a.anv * nc;
    const double* Bs = a.bst ? a.bst + q * (int64_t)a.bnv * nc : nullptr;
    auto lin = [&](int v, int64_t cc) {
        return a.cb == 0.0 ? A[v * nc + cc] : __dadd_rn(__dmul_rn(a.ca, A[v * nc + cc]), __dmul_rn(a.cb, Bs[v * nc + cc]));
    };
    const int64_t f0 = fidx<0>(G, i, j, k), f1 = fidx<1>(G, i, j, k), f2 = fidx<2>(G, i, j, k);
    const int64_t nf0 = G.nface[0], nf1 = G.nface[1], nf2 = G.nface[2];
    const double* F0 = a.flux[0] + q * NHRSC * nf0;
    const double* F1 = a.flux[1] + q * NHRSC * nf1;
    const double* F2 = a.flux[2] + q * NHRSC * nf2;
    const int64_t s1 = G.N[0], s2 = (int64_t)G.N[0] * G.N[1];
    double cons[NCONS];
    for (int v = 0; v < NHRSC; ++v) {
        const double dFx = __dmul_rn(__dsub_rn(F0[v * nf0 + f0], F0[v * nf0 + f0 + 1]), ix);
        const double dFy = __dmul_rn(__dsub_rn(F1[v * nf1 + f1], F1[v * nf1 + f1 + s1]), iy);
        const double dFz = __dmul_rn(__dsub_rn(F2[v * nf2 + f2], F2[v * nf2 + f2 + s2]), iz);
        cons[v] = __dadd_rn(lin(v, c), __dmul_rn(dte, __dadd_rn(__dadd_rn(dFx, dFy), dFz)));
    }
    // ct_update_B{x,y,z}_face (grmhd_helpers.hh:441-475)
    const double* E = a.emf + q * 3 * G.nedge;
    auto Ex = [&](int ii, int jj, int kk) { return E[eidx(G, ii, jj, kk)]; };
    auto Ey = [&](int ii, int jj, int kk) { return E[G.nedge + eidx(G, ii, jj, kk)]; };
    auto Ez = [&](int ii, int jj, int kk) { return E[2 * G.nedge + eidx(G, ii, jj, kk)]; };
    const double* bx = a.oB[0] + q * nf0;
    const double* by = a.oB[1] + q * nf1;
    const double* bz = a.oB[2] + q * nf2;
    auto ctx = [&](int ii) {
        return __dadd_rn(bx[fidx<0>(G, ii, j, k)],
                         __dmul_rn(dte, __dadd_rn(__dmul_rn(__dsub_rn(Ey(ii, j, k + 1), Ey(ii, j, k)), iz),
                                                  __dmul_rn(__dsub_rn(Ez(ii, j, k), Ez(ii, j + 1, k)), iy))));
    };
    auto cty = [&](int jj) {
        return __dadd_rn(by[fidx<1>(G, i, jj, k)],
                         __dmul_rn(dte, __dadd_rn(__dmul_rn(__dsub_rn(Ez(i + 1, jj, k), Ez(i, jj, k)), ix),
                                                  __dmul_rn(__dsub_rn(Ex(i, jj, k), Ex(i, jj, k + 1)), iz))));
    };
    auto ctz = [&](int kk) {
        return __dadd_rn(bz[fidx<2>(G, i, j, kk)],
                         __dmul_rn(dte, __dadd_rn(__dmul_rn(__dsub_rn(Ex(i, j + 1, kk), Ex(i, j, kk)), iy),
                                                  __dmul_rn(__dsub_rn(Ey(i, j, kk), Ey(i + 1, j, kk)), ix))));
    };
    cons[BSXL] = 0.5 * __dadd_rn(ctx(i), ctx(i + 1));
    cons[BSYL] = 0.5 * __dadd_rn(cty(j), cty(j + 1));
    cons[BSZL] = 0.5 * __dadd_rn(ctz(k), ctz(k + 1));
    bool dmp = false;
    if (o.dmp) {
        const double inv_M = 1.0 / o.dmp_M;
        const int vars[2] = {DENS_, TAU_};
        for (int n = 0; n < 2 && !dmp; ++n) {
            double vmax = -1.7976931348623157e308, vmin = 1.7976931348623157e308;
            for (int kt = -1; kt <= 1; ++kt)
                for (int jt = -1; jt <= 1; ++jt)
                    for (int it = -1; it <= 1; ++it) {
                        const double v = lin(vars[n], cidx(G, i + it, j + jt, k + kt));
                        vmax = fmax(vmax, v);
                        vmin = fmin(vmin, v);
                    }
            const double test = cons[vars[n]];
            if (test > __dmul_rn(o.dmp_M, vmax) || test < __dmul_rn(vmin, inv_M)) dmp = true;
        }
    }
    Met m;
    if constexpr (FLAT) m.flat();
    else load_metric(m, a.mstate + q * (int64_t)a.mnv * nc, nc, c);
    C2PPar cp = o.cp;
    cp.alp_bh_thresh = 1e30;   // lenient everywhere during the dry run (evolve.cpp:496-500)
    double pr[NPRIM];
    uint64_t err = 0;
    C2PCount cnt;
    C2PGuess gs{0, 0, 0, 0, 0, 0};
    C2PRoots rt{-1, -1, -1, -1};
    const double rad = o.ex.by_radius ? cell_radius(o.ex, I, a.blk[q].b, G.n, G.ng, i, j, k) : 1e300;
    const C2POut out = c2p(cons, pr, m, eos, o.atmo, o.ex, cp, rad, err, true, cnt, gs, rt);
    if (!(out.floored || dmp)) return;
    for (int t = 0; t < 2; ++t) {
        a.ftag[0][q * nf0 + fidx<0>(G, i + t, j, k)] = 1;
        a.ftag[1][q * nf1 + fidx<1>(G, i, j + t, k)] = 1;
        a.ftag[2][q * nf2 + fidx<2>(G, i, j, k + t)] = 1;
    }
    uint8_t* T = a.etag + q * 3 * G.nedge;
    for (int u = 0; u < 2; ++u)
        for (int w = 0; w < 2; ++w) {
            T[eidx(G, i, j + u, k + w)] = 1;
            T[G.nedge + eidx(G, i + u, j, k + w)] = 1;
            T[2 * G.nedge + eidx(G, i + u, j + w, k)] = 1;
        }
}

// ---------------------------------------------------------------------------
GHD int map_cc(int g, int N, int lo, int hi) {
    if (g < 0) return lo == 1 ? ((g % N) + N) % N : 0;
    if (g >= N) return hi == 1 ? g % N : N - 1;
    return g;
}
GHD int map_fc(int f, int N, int lo, int hi) {
    if (f < 0) return lo == 1 ? ((f % N) + N) % N : 0;
    if (f > N) return hi == 1 ? f % N : N;
    return f;
}

// A ghost cell whose position lies outside the instance across an outflow
// face holds clamped copies of conserved and face B whose face average differs
// from any interior cell: it needs its own c2p.  Every other ghost cell (same
// level neighbour or periodic image) has the exact inputs of its source cell,
// so its c2p outputs are copies of the source's.
GHD bool cell_is_interior(const DevGrid& G, int i, int j, int k) {
    return i >= G.ng && i < G.n[0] + G.ng && j >= G.ng && j < G.n[1] + G.ng && k >= G.ng && k < G.n[2] + G.ng;
}
GHD bool ghost_needs_own_c2p(const DevGrid& G, const DevInst& I, const DevBlk& b, int i, int j, int k) {
    int loc[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        int Nt = I.nb[d] * G.n[d];
        int g = b.b[d] * G.n[d] + loc[d] - G.ng;
        if (g < 0 && I.bc[2 * d] == 0) return true;
        if (g >= Nt && I.bc[2 * d + 1] == 0) return true;
    }
    return false;
}

// Ghost gather over precomputed tables.  The ghost shell of a block is a
// list of local offsets shared by all blocks (entries [0,ncg) cells, then x,
// y, z faces); src[q][e] holds the source block and local offset of entry e
// of block q, found once on the host by the same global-index mapping.  Cell
// entries whose ghost needs its own c2p carry GSRC_OWN in the block field.
// cell_mode: 0 skip cells, 1 every ghost cell, 2 ghost cells that need their
// own c2p, 3 ghost cells that take copies of c2p outputs.
struct GatherArgs {
    double* st; int nv; int v0, vc;          // state vars [v0, v0+vc)
    double* aux; uint32_t aux_mask;          // aux vars selected by bit mask
    double* ecen;                            // GS cell-centred E (3 vars) or null
    double* B[3];                            // face arrays (null: skip faces)
    int cell_mode;
    bool own_copies;                         // mode 3: also the copying outflow ghosts (GSRC_COPY)
    bool copies_only;                        // only the copying outflow ghosts (all of them)
    bool faces_slab;                         // faces: the face slabs and the faces of ghosts beyond outflow faces
};
// GSRC_OWN / GSRC_AMR: forest.hh

// Source of ghost element `loc` (face direction fd, -1 for cells) of block b.
static void ghost_source(const DevGrid& G, const DevInst& I, const DevBlk& b, int fd, const int loc[3], int& qs,
                         int& off) {
    int sb[3], sl[3];
    for (int d = 0; d < 3; ++d) {
        int Nt = I.nb[d] * G.n[d];
        int g = b.b[d] * G.n[d] + loc[d] - G.ng;
        if (d == fd) {
            int m = map_fc(g, Nt, I.bc[2 * d], I.bc[2 * d + 1]);
            if (m == Nt) { sb[d] = I.nb[d] - 1; sl[d] = G.n[d] + G.ng; }
            else { sb[d] = m / G.n[d]; sl[d] = m % G.n[d] + G.ng; }
        } else {
            int m = map_cc(g, Nt, I.bc[2 * d], I.bc[2 * d + 1]);
            sb[d] = m / G.n[d]; sl[d] = m % G.n[d] + G.ng;
        }
    }
    qs = (int)(I.q0 + sb[0] + (int64_t)I.nb[0] * (sb[1] + (int64_t)I.nb[1] * sb[2]));
    off = (int)(fd < 0 ? cidx(G, sl[0], sl[1], sl[2]) : fidx_d(G, fd, sl[0], sl[1], sl[2]));
}

// Outflow ghosts that take a first-layer ghost's c2p outputs carry GSRC_COPY
// with that ghost as their source.
// One copy per thread.  A channel is one variable of one array
// (base, vars per block, var, elements per var per block).
struct GatherChan { double* base; int nv, v; int64_t n; };
constexpr int GATHER_MAXCH = 26;
struct GatherChans { GatherChan c[GATHER_MAXCH]; int nch; };

// Cell ghosts of every block (wpb workgroups per block, see wg_pos); each
// thread copies every channel of its cell, loads of four channels in flight
// before their stores.  mode: 1 every ghost cell, 2 only those that run their
// own c2p, 3 only those that take copies of c2p outputs.
__global__ __launch_bounds__(256) void k_gather_cells(const int32_t* __restrict__ list, const int2* __restrict__ src,
                                                      int per, int ncg, int wpb, int mode, int copies,
                                                      GatherChans ch) {
    const WgPos wp = wg_pos(wpb);
    const int e = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (e >= ncg) return;
    const int64_t q = wp.q;
    const int2 sr = src[q * per + e];
    const bool own = (sr.x & GSRC_OWN) != 0;
    if ((sr.x & GSRC_AMR) || (mode >= 2 && (mode == 2) != own)) return;
    // an outflow ghost that copies a first-layer ghost of its block (source
    // {block, first-layer cell}): its c2p outputs in mode 3 when asked; its
    // conserved in mode 1 come after this launch, from that filled cell
    if ((sr.x & GSRC_COPY) && (mode == 1 || !copies)) return;
    const int64_t qs = sr.x & (GSRC_COPY - 1);
    const int off = list[e];
    for (int c0 = 0; c0 < ch.nch; c0 += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (c0 + u < ch.nch) {
                const GatherChan& C = ch.c[c0 + u];
                v[u] = C.base[(qs * C.nv + C.v) * C.n + sr.y];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (c0 + u < ch.nch) {
                const GatherChan& C = ch.c[c0 + u];
                C.base[(q * C.nv + C.v) * C.n + off] = v[u];
            }
        }
    }
}

// Explicit {block, cell, source block, source cell} entries (the ghosts
// that run their own c2p), z over channels.
__global__ __launch_bounds__(256) void k_gather_list(const int4* __restrict__ ent, int n, GatherChans ch) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int4 t = ent[e];
    const GatherChan C = ch.c[blockIdx.z];
    C.base[((int64_t)t.x * C.nv + C.v) * C.n + t.y] = C.base[((int64_t)t.z * C.nv + C.v) * C.n + t.w];
}

// Face ghosts of the three face-B arrays (wpb workgroups per block): the
// first m0 / m1 / m2 of the nfg0 / nfg1 / nfg2 entries of each direction.
__global__ __launch_bounds__(256) void k_gather_faces(const int32_t* __restrict__ list, const int2* __restrict__ src,
                                                      int per, int ncg, int nfg0, int nfg1, int m0, int m1, int m2,
                                                      int wpb, double* B0, double* B1, double* B2, int64_t nf0,
                                                      int64_t nf1, int64_t nf2) {
    const WgPos wp = wg_pos(wpb);
    const int r = wp.tile * (int)blockDim.x + (int)threadIdx.x;
    if (r >= m0 + m1 + m2) return;
    const int64_t q = wp.q;
    double* B; int64_t nf; int e;
    if (r < m0) { B = B0; nf = nf0; e = ncg + r; }
    else if (r < m0 + m1) { B = B1; nf = nf1; e = ncg + nfg0 + (r - m0); }
    else { B = B2; nf = nf2; e = ncg + nfg0 + nfg1 + (r - m0 - m1); }
    const int2 sr = src[q * per + e];
    if (sr.x & GSRC_AMR) return;
    B[q * nf + list[e]] = B[(int64_t)sr.x * nf + sr.y];
}

// ---------------------------------------------------------------------------
// c2p statistics: one partial record per workgroup (no atomics), summed by a
// single-workgroup reduction.
// ---------------------------------------------------------------------------
struct DevStats {
    unsigned long long n_cells, n_fatal, n_atmo, n_ent, n_guess, n_root, n_bracket, max_root;
};
// Block-level reduction of the per-lane counters into part[blockIdx.x].
// Every thread of the (256-thread) workgroup must call it.
__device__ inline void stats_block(DevStats* part, bool active, bool fatal, uint64_t err, const C2PCount& cnt) {
    __shared__ unsigned long long sh[4][8];
    unsigned long long a[8] = {(unsigned long long)active, (unsigned long long)fatal,
                               (unsigned long long)((err >> E_ATMO_RESET) & 1),
                               (unsigned long long)((err >> E_ENT_BACKUP_USED) & 1), (unsigned long long)__popc((unsigned)cnt.guess_used),
                               (unsigned long long)cnt.root, (unsigned long long)cnt.bracket,
                               (unsigned long long)cnt.root};
    for (int off = 32; off > 0; off >>= 1) {
        for (int x = 0; x < 7; ++x) a[x] += __shfl_down(a[x], off, 64);
        unsigned long long o = __shfl_down(a[7], off, 64);
        a[7] = o > a[7] ? o : a[7];
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int x = 0; x < 8; ++x) sh[w][x] = a[x];
    __syncthreads();
    if (threadIdx.x == 0) {
        DevStats r{0, 0, 0, 0, 0, 0, 0, 0};
        const int nw = (blockDim.x + 63) >> 6;
        for (int v = 0; v < nw; ++v) {
            r.n_cells += sh[v][0]; r.n_fatal += sh[v][1]; r.n_atmo += sh[v][2]; r.n_ent += sh[v][3];
            r.n_guess += sh[v][4]; r.n_root += sh[v][5]; r.n_bracket += sh[v][6];
            r.max_root = sh[v][7] > r.max_root ? sh[v][7] : r.max_root;
        }
        part[blockIdx.x] = r;
    }
}
__global__ __launch_bounds__(256) void k_stats_final(const DevStats* __restrict__ part, int64_t np, DevStats* out) {
    __shared__ DevStats sh[256];
    DevStats r{0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t x = threadIdx.x; x < np; x += blockDim.x) {
        const DevStats& p = part[x];
        r.n_cells += p.n_cells; r.n_fatal += p.n_fatal; r.n_atmo += p.n_atmo; r.n_ent += p.n_ent;
        r.n_guess += p.n_guess; r.n_root += p.n_root; r.n_bracket += p.n_bracket;
        r.max_root = p.max_root > r.max_root ? p.max_root : r.max_root;
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        DevStats t{0, 0, 0, 0, 0, 0, 0, 0};
        for (int v = 0; v < (int)blockDim.x; ++v) {
            t.n_cells += sh[v].n_cells; t.n_fatal += sh[v].n_fatal; t.n_atmo += sh[v].n_atmo; t.n_ent += sh[v].n_ent;
            t.n_guess += sh[v].n_guess; t.n_root += sh[v].n_root; t.n_bracket += sh[v].n_bracket;
            t.max_root = sh[v].max_root > t.max_root ? sh[v].max_root : t.max_root;
        }
        *out = t;
    }
}

// Per-workgroup statistics records of grmhd_c2p_batch_noalloc: a file-scope
// device array instead of context memory that grows with the batch.  One
// record per 256-state workgroup for at most GRMHD_NOALLOC_MAX_STATES states:
// 65,536 records x 64 B = 4 MiB, held by every GPU the library runs on until
// the process ends.  Host code takes its device address from
// hipGetSymbolAddress (noalloc_scratch_ptr), never from the name.
static __device__ __align__(256) DevStats scratch[GRMHD_NOALLOC_MAX_STATES / 256];

// ---------------------------------------------------------------------------
// c2p over cells (compute_auxiliaries)
// ---------------------------------------------------------------------------
// Minimum resident 256-thread blocks per CU for the c2p kernels (caps their
// registers: 2 -> 256 VGPRs, 3 -> 168, 4 -> 128 per lane).
#ifndef GRMHD_C2P_MINB
#define GRMHD_C2P_MINB 3
#endif
// Cell sets: CS_INTERIOR (the interior cells of every block, wpb workgroups
// per block, see wg_pos), CS_SHELL (the ghost-shell cells of every block),
// CS_LIST (explicit {block, cell} pairs).
enum { CS_INTERIOR = 0, CS_SHELL = 1, CS_LIST = 2 };
// GUESS 1 (learned): the root solves start from the MLP's prediction (ideal
// gas: Newton finish, as k_c2p_batch<EOS, true, true>; other EOS: brackets).
// GUESS 2 (warm start, ideal gas): the Newton finish starts from the cell's
// previous primitives in aux, mu = 1/(h W) of the last c2p of this cell
// (the previous RK stage, or the caller's data), at most warm_newton
// evaluations before the plain path.
template <typename EOS, bool FLAT, int GUESS = 0>
__global__ __launch_bounds__(256, GRMHD_C2P_MINB) void k_c2p_cells(DevGrid G, EOS eos, Opts o, const DevInst* __restrict__ inst,
                                                   const DevBlk* __restrict__ blk, double* st, int snv,
                                                   const double* __restrict__ mstate, int mnv,
                                                   const double* __restrict__ Bx, const double* __restrict__ By,
                                                   const double* __restrict__ Bz, double* aux, double* ecen,
                                                   DevStats* stats, int cset, const int32_t* __restrict__ shell,
                                                   const int2* __restrict__ cl, int ne, int wpb, MlpDev mlp) {
    WgPos wp{0, 0};
    if (cset != CS_LIST) wp = wg_pos(wpb);
    const int e = cset == CS_LIST ? (int)(blockIdx.x * blockDim.x + threadIdx.x)
                                  : wp.tile * (int)blockDim.x + (int)threadIdx.x;
    bool active = e < ne;
    const int ee = active ? e : 0;
    int64_t q;
    int64_t c;
    int i, j, k;
    if (cset == CS_INTERIOR) {
        q = wp.q;
        const int n0 = G.n[0], n01 = G.n[0] * G.n[1];
        int r;
        udivmod(ee, n01, k, r);
        udivmod(r, n0, j, i);
        i += G.ng; j += G.ng; k += G.ng;
        c = cidx(G, i, j, k);
    } else {
        int off;
        if (cset == CS_SHELL) { q = wp.q; off = shell[ee]; }
        else { const int2 p = cl[ee]; q = p.x; off = p.y; }
        c = off;
        const int N0 = G.N[0], N01 = G.N[0] * G.N[1];
        int r;
        udivmod(off, N01, k, r);
        udivmod(r, N0, j, i);
    }
    const int64_t nc = G.ncell;
    Met m;
    if constexpr (FLAT) m.flat();
    else load_metric(m, mstate + q * (int64_t)mnv * nc, nc, c);
    double* S = st + q * (int64_t)snv * nc + c;
    double* A = aux + q * (int64_t)NAUX * nc + c;
    const double* bx = Bx + q * G.nface[0]; const double* by = By + q * G.nface[1]; const double* bz = Bz + q * G.nface[2];
    int64_t fx = fidx<0>(G, i, j, k), fy = fidx<1>(G, i, j, k), fz = fidx<2>(G, i, j, k);
    double bx0 = bx[fx], bx1 = bx[fx + 1], by0 = by[fy], by1 = by[fy + G.N[0]];
    double bz0 = bz[fz], bz1 = bz[fz + (int64_t)G.N[0] * G.N[1]];
    double cons[NCONS];
    for (int v = 0; v < NHRSC; ++v) cons[v] = S[v * nc];
    // cons order DENS SX SY SZ TAU YE* S* matches the state order
    cons[BSXL] = 0.5 * (bx0 + bx1); cons[BSYL] = 0.5 * (by0 + by1); cons[BSZL] = 0.5 * (bz0 + bz1);
    double Bc0 = cons[BSXL], Bc1 = cons[BSYL], Bc2 = cons[BSZL];
    uint64_t err = 0;
    double pr[NPRIM];
    C2PCount cnt;
    C2POut out{false, false};
    C2PGuess gs{0, 0, 0, 0, 0, 0};
    // The forward holds block barriers, so the block decides together: it
    // runs only if some cell of the block reaches a root solve (D / sqrt(gamma)
    // at or above the atmosphere threshold, the test of c2p); cells below it
    // are reset to atmosphere without a solve, so a block of them takes the
    // plain path and skips the network.
    bool use_net = false;
    if constexpr (GUESS == 1) {
        const bool solve = active && cons[DENSL] / m.sqrtg >= o.atmo.rho_fl * (1 + o.atmo.atmo_tol);
        use_net = __syncthreads_or(solve) != 0;
    }
    if (use_net) {
        // every thread of the block takes part (the forward holds barriers);
        // inactive threads run on the valid cell ee = 0 and discard the result
        __shared__ __align__(16) uint16_t xs[4][64 * MLP_KP];
        __shared__ __align__(16) float ys[4][64 * 2];
        float f[MLP_NF];
        double lxhi = 0.0;
        mlp_features(eos, m, cons, f, lxhi);
        float y0, y1;
        mlp_block_forward(mlp, f, xs[threadIdx.x >> 6], ys[threadIdx.x >> 6], y0, y1);
        const double xe = (double)__expf((float)lxhi + y0);
        const double z = (double)y1;
        const double rs = (fabs(z) < 1e-3) ? 1.0 + z * (1.0 + z * (0.5 + z * (1.0 / 6.0))) : exp(z);
        if constexpr (EOS::is_ideal) {
            gs.newton = mlp.newton;
            gs.e_c = xe;
            gs.s_c = rs;
        } else {
            gs.e_lo = xe * mlp.fe_lo;
            gs.e_hi = fmin(xe * mlp.fe_hi, 1.0);
            gs.s_lo = rs * mlp.fs_lo;
            gs.s_hi = rs * mlp.fs_hi;
            gs.nexp = mlp.nexp;
        }
        gs.s_rel = 1;
    }
    if constexpr (GUESS == 2 && EOS::is_ideal) {
        if (active) {
            // x = h0 / (h W) of the cell's previous primitives; both master
            // functions' roots lie in (0, 1] in this unit, the entropy root
            // near the energy root
            const double rho0 = A[RHO_ * nc], eps0 = A[EPS_ * nc], p0 = A[PRESS_ * nc];
            const double z0[3] = {A[ZVECX_ * nc], A[ZVECY_ * nc], A[ZVECZ_ * nc]};
            const double x = eos.hmin() / ((1.0 + eps0 + p0 / rho0) * lorentz(m.g, z0));
            if (x > 0.0 && x <= 1.0) {
                gs.newton = mlp.newton;
                gs.e_c = x;
                gs.s_c = 1.0;
                gs.s_rel = 1;
            }
        }
    }
    if (active) {
        A[BX_ * nc] = Bc0 / m.sqrtg; A[BY_ * nc] = Bc1 / m.sqrtg; A[BZ_ * nc] = Bc2 / m.sqrtg;
        C2PRoots rt{-1, -1, -1, -1};
        const double rad = o.ex.by_radius ? cell_radius(o.ex, inst[blk[q].inst], blk[q].b, G.n, G.ng, i, j, k) : 1e300;
        out = c2p(cons, pr, m, eos, o.atmo, o.ex, o.cp, rad, err, false, cnt, gs, rt);
        if (out.fatal && o.fatal_count) atomicAdd(o.fatal_count, 1ull);
        A[RHO_ * nc] = pr[RHOL]; A[EPS_ * nc] = pr[EPSL]; A[PRESS_ * nc] = pr[PRESSL];
        A[TEMP_ * nc] = pr[TEMPL]; A[ENTROPY_ * nc] = pr[ENTL]; A[YE_ * nc] = pr[YEL];
        A[ZVECX_ * nc] = pr[ZXL]; A[ZVECY_ * nc] = pr[ZYL]; A[ZVECZ_ * nc] = pr[ZZL];
        if (err & BIT(E_RESET_DENS)) { A[C2P_DENS_ERR_ * nc] += S[DENS_ * nc] - cons[DENSL]; S[DENS_ * nc] = cons[DENSL]; }
        if (err & BIT(E_RESET_STILDE)) { S[SX_ * nc] = cons[STXL]; S[SY_ * nc] = cons[STYL]; S[SZ_ * nc] = cons[STZL]; }
        if (err & BIT(E_RESET_TAU)) S[TAU_ * nc] = cons[TAUL];
        if (err & BIT(E_RESET_ENTROPY)) S[ENTROPYSTAR_ * nc] = cons[ENTSL];
        if (err & BIT(E_RESET_YE)) S[YESTAR_ * nc] = cons[YESL];
        uint64_t prev = (uint64_t)A[C2P_ERR_ * nc];
        A[C2P_ERR_ * nc] = (double)(prev | err);
        const DevInst& I = inst[blk[q].inst];
        A[BDIV_ * nc] = ((bx1 - bx0) * I.idx[0] + (by1 - by0) * I.idx[1] + (bz1 - bz0) * I.idx[2]) / m.sqrtg;
        if (ecen) {
            double z[3] = {pr[ZXL], pr[ZYL], pr[ZZL]};
            double W = sqrt(1. + m.sqv(z[0], z[1], z[2]));
            double vt[3] = {m.alp * z[0] / W - m.beta[0], m.alp * z[1] / W - m.beta[1], m.alp * z[2] / W - m.beta[2]};
            double* Ec = ecen + q * 3 * nc + c;
            Ec[0] = Bc1 * vt[2] - Bc2 * vt[1];
            Ec[nc] = -Bc0 * vt[2] + Bc2 * vt[0];
            Ec[2 * nc] = Bc0 * vt[1] - Bc1 * vt[0];
        }
    }
    if (!active) { cnt = C2PCount(); err = 0; }
    if (stats) stats_block(stats, active, active && out.fatal, err, cnt);
}

// ---------------------------------------------------------------------------
// c2p over a flat batch
// ---------------------------------------------------------------------------
// LEARNED: the brackets come from the MLP (features -> BF16 MFMA forward in
// LDS/registers, one pass per wave) instead of the guess array.  NEWTON
// (ideal gas): the predicted roots start Newton iterations, and where those
// do not converge the plain path runs (no bracket code in this variant).
template <typename EOS, bool LEARNED, bool NEWTON = false>
__global__ __launch_bounds__(256, GRMHD_C2P_MINB) void k_c2p_batch(int64_t n, EOS eos, Opts o, double* cons_soa,
                                                   const double* __restrict__ met, double* prims,
                                                   uint64_t* errb, uint8_t* floored, int dry_run,
                                                   const double* __restrict__ rad, const double* __restrict__ guess,
                                                   double* __restrict__ roots, DevStats* stats, MlpDev mlp) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = t < n;
    int64_t s = active ? t : 0;
    Met m;
    for (int a = 0; a < 6; ++a) m.g[a] = met[a * n + s];
    m.alp = met[6 * n + s];
    for (int a = 0; a < 3; ++a) m.beta[a] = met[(7 + a) * n + s];
    m.init();
    double cons[NCONS], pr[NPRIM];
    for (int v = 0; v < NCONS; ++v) cons[v] = cons_soa[v * n + s];
    for (int v = 0; v < NPRIM; ++v) pr[v] = 0.0;
    C2PGuess gs{0, 0, 0, 0, 0, 0};
    if (LEARNED) {
        __shared__ __align__(16) uint16_t xs[4][64 * MLP_KP];
        __shared__ __align__(16) float ys[4][64 * 2];
        float f[MLP_NF];
        double lxhi = 0.0;
        mlp_features(eos, m, cons, f, lxhi);
        if (!active) for (int k = 0; k < MLP_NF; ++k) f[k] = mlp.norm[k];
        float y0, y1;
        mlp_block_forward(mlp, f, xs[threadIdx.x >> 6], ys[threadIdx.x >> 6], y0, y1);
        // energy root: FP32 exp is ample for a bracket of half-width >= 1e-6;
        // the entropy/energy ratio is close to 1, so its bracket is formed in FP64
        const double xe = (double)__expf((float)lxhi + y0);
        const double z = (double)y1;
        const double rs = (fabs(z) < 1e-3) ? 1.0 + z * (1.0 + z * (0.5 + z * (1.0 / 6.0))) : exp(z);
        if constexpr (NEWTON) {
            gs.newton = mlp.newton;
            gs.e_c = xe;
            gs.s_c = rs;
        } else {
            gs.e_lo = xe * mlp.fe_lo;
            gs.e_hi = fmin(xe * mlp.fe_hi, 1.0);
            gs.s_lo = rs * mlp.fs_lo;
            gs.s_hi = rs * mlp.fs_hi;
            gs.nexp = mlp.nexp;
        }
        gs.s_rel = 1;
    } else if (guess) {
        gs.e_lo = guess[s]; gs.e_hi = guess[n + s]; gs.s_lo = guess[2 * n + s]; gs.s_hi = guess[3 * n + s];
    }
    C2PRoots rt{-1, -1, -1, -1};
    uint64_t err = 0;
    C2PCount cnt;
    C2POut out{false, false};
    if (active) {
        out = c2p(cons, pr, m, eos, o.atmo, o.ex, o.cp, rad ? rad[s] : 1e300, err, dry_run != 0, cnt, gs, rt);
        if (out.fatal && o.fatal_count) atomicAdd(o.fatal_count, 1ull);
        if (roots) { roots[s] = rt.x_e; roots[n + s] = rt.x_s; roots[2 * n + s] = rt.m_e; roots[3 * n + s] = rt.m_s; }
        for (int v = 0; v < NPRIM; ++v) prims[v * n + s] = pr[v];
        for (int v = 0; v < NCONS; ++v) cons_soa[v * n + s] = cons[v];
        errb[s] = err;
        floored[s] = out.floored ? 1 : 0;
        if (o.eval_counts) {
            const int ev = cnt.root + cnt.bracket;
            o.eval_counts[s] = (uint8_t)(ev < 255 ? ev : 255);
        }
    } else {
        cnt = C2PCount();
    }
    if (stats) stats_block(stats, active, active && out.fatal, active ? err : 0, cnt);
}

__global__ __launch_bounds__(256) void k_p2c_batch(int64_t n, const double* __restrict__ prims,
                                                   const double* __restrict__ met, double* cons_soa) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    Met m;
    for (int a = 0; a < 6; ++a) m.g[a] = met[a * n + s];
    m.alp = met[6 * n + s];
    for (int a = 0; a < 3; ++a) m.beta[a] = met[(7 + a) * n + s];
    m.init();
    double pr[NPRIM], cons[NCONS];
    for (int v = 0; v < NPRIM; ++v) pr[v] = prims[v * n + s];
    for (int v = 0; v < NCONS; ++v) cons[v] = cons_soa[v * n + s];
    p2c(pr, cons, m);
    for (int v = 0; v < 7; ++v) cons_soa[v * n + s] = cons[v];
}

// ---------------------------------------------------------------------------
// dt: per-WG minima then a single-WG reduce
// ---------------------------------------------------------------------------
template <typename EOS, bool FLAT>
__global__ __launch_bounds__(256) void k_dt_partial(DevGrid G, EOS eos, const DevInst* __restrict__ inst,
                                                    const DevBlk* __restrict__ blk, const double* __restrict__ mstate,
                                                    int mnv, const double* __restrict__ aux, double cfl,
                                                    double* __restrict__ part) {
    __shared__ double red[256];
    int64_t per = (int64_t)G.n[0] * G.n[1] * G.n[2];
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dtl = 1e300;
    if (t < per * G.nq) {
        int64_t q = t / per, r = t - q * per;
        int i = G.ng + (int)(r % G.n[0]), j = G.ng + (int)((r / G.n[0]) % G.n[1]), k = G.ng + (int)(r / ((int64_t)G.n[0] * G.n[1]));
        const int64_t nc = G.ncell, c = cidx(G, i, j, k);
        Met m;
        if constexpr (FLAT) m.flat(); else load_metric(m, mstate + q * (int64_t)mnv * nc, nc, c);
        const double* X = aux + q * (int64_t)NAUX * nc + c;
        double z[3] = {X[ZVECX_ * nc], X[ZVECY_ * nc], X[ZVECZ_ * nc]}, B[3] = {X[BX_ * nc], X[BY_ * nc], X[BZ_ * nc]};
        double rho = X[RHO_ * nc], T = X[TEMP_ * nc], eps = X[EPS_ * nc], press = X[PRESS_ * nc];
        double h, cs2; unsigned e = 0;
        eos.press_h_cs2_from_temp(h, cs2, T, rho, e);
        double W = lorentz(m.g, z);
        double b[4], b2; smallb(m, B, z, W, b, b2);
        double vt[3]; vtilde(m, z, W, vt);
        double cmax = 0;
        for (int d = 0; d < 3; ++d) {
            double cm, cp;
            cmcp(m, rho, press, eps, cs2, W, b2, vt, d, cm, cp);
            cmax = fmax(cmax, fmax(fabs(cp), fabs(cm)));
        }
        double dx0 = 1.0 / inst[blk[q].inst].idx[0];
        dtl = cfl / cmax * dx0;
    }
    red[threadIdx.x] = dtl;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_dt_final(const double* __restrict__ part, int64_t np, double* dt) {
    __shared__ double red[256];
    double v = 1e300;
    for (int64_t x = threadIdx.x; x < np; x += blockDim.x) v = fmin(v, part[x]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *dt = red[0];
}

// A host value into a device scalar, in stream order (host-dt entry points).
__global__ void k_set_scalar(double* p, double v) { *p = v; }

__global__ __launch_bounds__(256) void k_reset_diag(int64_t ncell, int64_t nq, double* aux) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * nq) return;
    int64_t q = t / ncell, c = t - q * ncell;
    aux[(q * NAUX + C2P_DENS_ERR_) * ncell + c] = 0.0;
    aux[(q * NAUX + C2P_ERR_) * ncell + c] = 0.0;
}

// ---------------------------------------------------------------------------
// Host side: context
// ---------------------------------------------------------------------------
struct TimerRec { std::string fam; hipEvent_t a, b; };

struct grmhd_ctx {
    grmhd_mlp* mlp = nullptr;                 // learned c2p guess for the cell passes (not owned)
    MlpDev gmlp{};                            // its kernel arguments in the captured graph
    double* mder = nullptr;                   // cached derivatives of the static metric (grmhd_set_static_metric)
    double* fmet[3] = {nullptr, nullptr, nullptr};   // cached y/z face metrics (FMC per face), same source
    const double* mder_src = nullptr;         // the state array they were taken from
    grmhd_params P;
    DevGrid G;
    Opts O;
    EosIdeal ideal;
    EosHybrid hybrid;
    std::vector<DevInst> inst_h;
    std::vector<DevBlk> blk_h;
    DevInst* inst_d = nullptr;
    DevBlk* blk_d = nullptr;
    double* flux[3] = {nullptr, nullptr, nullptr};
    double* vbar[3] = {nullptr, nullptr, nullptr};
    double* emf = nullptr;
    double* ecen = nullptr;
    double* U1 = nullptr; double* U2 = nullptr;        // stage buffers (7 vars)
    double* B1[3] = {nullptr, nullptr, nullptr};
    double* B2[3] = {nullptr, nullptr, nullptr};
    double* dtpart = nullptr; int64_t n_dtpart = 0;
    DevStats* stats = nullptr;
    DevStats* stats_part = nullptr;
    int64_t n_part_cap = 0;
    DevStats* noalloc_part = nullptr;         // device address of `scratch` on this context's device
    unsigned long long* fatal_d = nullptr;    // fatal c2p cells since the last sync point (Opts::fatal_count)
    double* dt_scalar = nullptr;              // dt of the host-dt entry points
    // ring faces written by the neighbouring block's face threads: per block
    // the face neighbours (-1 beyond an outflow face) and the mask of block
    // faces beyond which k_flux computes the ring itself; null when off
    int32_t* ring_nbr = nullptr;
    uint8_t* ring_own = nullptr;
    // set by the RK3 step around the flux launches of its second and third
    // stages: their ring faces come from the neighbours' pushes and their
    // full faces read the cells beyond a block face from the neighbour's
    // interior, so the gathers between stages leave the face slabs alone
    // (the first stage and grmhd_advance_substep read the caller's ghosts and
    // compute every ring face)
    bool ring_copy_now = false;
    int warm_newton = 0;                      // grmhd_set_c2p_warm_start (0: off)
    int32_t* ghost_list = nullptr;
    int2* ghost_src = nullptr;                // [nq][ncg + nfg0 + nfg1 + nfg2]
    uint8_t* ftag[3] = {nullptr, nullptr, nullptr};   // FOFC face tags (per face of every block)
    uint8_t* etag = nullptr;                  // FOFC edge tags, 3 components
    int2* own_cells = nullptr;                // {block, cell} of ghosts that run their own c2p
    int4* own_gather = nullptr;               // the same with their sources
    int64_t n_own = 0;
    // Ghosts beyond outflow faces whose c2p inputs are bit-identical to those
    // of the first ghost layer's cell of the same block (flat metric, no
    // radius excision): {block, cell, block, first-layer cell}; they take
    // copies of that cell's c2p outputs instead of running their own (the
    // stage-end gather, through GSRC_COPY in the ghost table) and, when every
    // ghost is filled, its conserved (this list, after the cell gather);
    // between RK stages with shared ring faces this list is the whole gather.
    int4* copy_gather = nullptr;
    // with shared ring faces: the ghost faces outside the face slabs that
    // bound a ghost cell beyond an outflow face {block, face, source block,
    // source face}, x then y then z (n_ofl each); between RK stages the face
    // gather is the face slabs and these
    int4* ofl_faces = nullptr;
    int64_t n_ofl[3] = {0, 0, 0};
    int64_t n_copy = 0;
    int64_t ncg = 0, nfg[3] = {0, 0, 0};
    int64_t nfg_slab[3] = {0, 0, 0};          // face ghosts of the face slabs (listed first per direction)
    int64_t last_c2p_n = 0;
    bool timing = false;
    std::vector<TimerRec> timers;
    std::vector<hipEvent_t> ev_pool;
    // side streams and fork/join events of the concurrent RK3 stages
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fev[8] = {};
    hipStream_t flux_s[3] = {nullptr, nullptr, nullptr};   // per-direction flux streams while set
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void* gkey[8] = {0};
    // workgroup sizes picked by grmhd_autotune (all 256 until then)
    grmhd_tuning tune{{256, 256, 256}, 256, 256, 256, 256, 256, {0, 0, 0}};
    // two-level forest (grmhd_ctx_create_forest)
    bool forest = false;
    FForest ff{};                             // device tables
    std::vector<int32_t> cmap_h;
    std::vector<int4> fblk_h;
    int32_t* cmap_d = nullptr;
    int4* fblk_d = nullptr;
    int2* amr_cells = nullptr; int64_t n_amr_cells = 0;
    int2* amr_faces = nullptr; int64_t n_amr_faces = 0;           // first pass: faces on coarse faces
    int2* amr_faces_mid = nullptr; int64_t n_amr_faces_mid = 0;   // second pass: interior fine faces
    RefluxFace* rflux = nullptr; int64_t n_rflux = 0;
    RefluxEdge* remf = nullptr; int64_t n_remf = 0;
    int64_t n_own_gather = 0;                 // own-c2p ghosts that copy a clamped source
    double* vcm = nullptr;                    // virtual coarse metric field (grmhd_fill_metric_ghosts)
};

// Timing-only events: no system-scope fence when recorded, so a timer does
// not write back and invalidate the caches between the launches it brackets
// (which would also slow the launch that follows it).
static hipEvent_t timing_event() {
    hipEvent_t e = nullptr;
    (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    return e;
}
static hipEvent_t pool_event(grmhd_ctx* c) {
    if (!c->ev_pool.empty()) { hipEvent_t e = c->ev_pool.back(); c->ev_pool.pop_back(); return e; }
    return timing_event();
}
struct Scope {
    grmhd_ctx* c; const char* fam; hipStream_t s; hipEvent_t a = nullptr, b = nullptr;
    Scope(grmhd_ctx* c_, const char* f, hipStream_t s_) : c(c_), fam(f), s(s_) {
        if (c->timing) { a = pool_event(c); (void)hipEventRecord(a, s); }
    }
    ~Scope() {
        if (c->timing) { b = pool_event(c); (void)hipEventRecord(b, s); c->timers.push_back({fam, a, b}); }
    }
};

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }
static inline unsigned nblk_bs(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

extern "C" void grmhd_default_params(grmhd_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->eos_kind = GRMHD_EOS_IDEAL_GAS;
    p->gamma = 2.0;
    p->eps_max = 1e5;
    p->entropy_min = 1e-13;
    p->gamma_th = 1.8;
    p->pwp_n = 1;
    p->pwp_kappa0 = 1.685819e+02;
    p->pwp_gamma[0] = 2.0;
    p->rho_fl = 1e-14; p->temp_fl = 0.0; p->ye_fl = 0.0; p->atmo_tol = 1e-3;
    p->c2p_tol = 1e-10; p->max_lorentz = 50; p->max_sigma = 100; p->beta_fallback = 1e-2;
    p->use_entropy_backup = 1; p->bh_alp_thresh = 0.2;
    p->excise_by_radius = 0; p->excision_radius = 1.0; p->excision_lapse = 0.05;
    p->recon = GRMHD_RECON_WENOZ; p->riemann = GRMHD_RIEMANN_HLLE; p->emf = GRMHD_EMF_UCT;
    p->thermo_recon = GRMHD_THERMO_TEMP; p->fd_order = 6; p->rusanov_c_limit = 0.0;
    p->flat_metric = 0;
    p->fofc = 0; p->fofc_dmp_enable = 1; p->fofc_dmp_M = 1.2; p->fofc_ghost_emf = 0;
    p->stage_streams = 1;
}

static void build_eos(grmhd_ctx* c) {
    const grmhd_params& p = c->P;
    c->ideal.gm1 = p.gamma - 1.0;
    c->ideal.gamma = c->ideal.gm1 + 1.0;
    c->ideal.ggm1 = c->ideal.gamma * c->ideal.gm1;
    c->ideal.eps_max = p.eps_max;
    c->ideal.temp_atm = p.temp_fl;
    c->ideal.ent_min = (p.temp_fl > 0.0) ? (p.temp_fl / c->ideal.gm1) : 1e-100;
    c->ideal.ls_min = std::log(c->ideal.ent_min) / c->ideal.gm1;
    c->ideal.ls_max = std::log(c->ideal.eps_max) / c->ideal.gm1;
    // hybrid over a piecewise polytrope (eos_storage.cpp:52-99)
    EosHybrid& h = c->hybrid;
    h.n = p.pwp_n;
    double rh[8] = {0};
    rh[0] = 0.0;
    for (int i = 1; i < h.n; ++i) rh[i] = p.pwp_rho[i - 1];
    h.K[0] = p.pwp_kappa0; h.ep[0] = 0.0;
    for (int i = 0; i < h.n; ++i) { h.G[i] = p.pwp_gamma[i]; h.rh[i] = rh[i]; }
    for (int i = 1; i < h.n; ++i) {
        h.K[i] = h.K[i - 1] * pow(rh[i], h.G[i - 1] - h.G[i]);
        h.ep[i] = h.ep[i - 1] + h.K[i - 1] * pow(rh[i], h.G[i - 1] - 1.) / (h.G[i - 1] - 1.)
                  - h.K[i] * pow(rh[i], h.G[i] - 1.) / (h.G[i] - 1.);
    }
    h.gth_m1 = p.gamma_th - 1.0;
    h.entropy_min = p.entropy_min;
    h.eps_max = p.eps_max;
    h.temp_atm = p.temp_fl;
}

static void free_ctx(grmhd_ctx* c) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    F(c->inst_d); F(c->blk_d); F(c->emf); F(c->ecen); F(c->U1); F(c->U2); F(c->dtpart); F(c->stats); F(c->stats_part); F(c->ghost_list); F(c->ghost_src); F(c->own_cells); F(c->own_gather); F(c->copy_gather); F(c->ofl_faces);
    F(c->ftag[0]); F(c->ftag[1]); F(c->ftag[2]); F(c->etag); F(c->mder);
    F(c->fmet[0]); F(c->fmet[1]); F(c->fmet[2]);
    F(c->cmap_d); F(c->fblk_d); F(c->amr_cells); F(c->amr_faces); F(c->amr_faces_mid); F(c->rflux); F(c->remf);
    F(c->vcm); F(c->fatal_d); F(c->dt_scalar); F(c->ring_nbr); F(c->ring_own);
    for (int d = 0; d < 3; ++d) { F(c->flux[d]); F(c->vbar[d]); F(c->B1[d]); F(c->B2[d]); }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto& t : c->timers) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto e : c->fev) if (e) (void)hipEventDestroy(e);
    for (auto t : c->side) if (t) (void)hipStreamDestroy(t);
}

// Host description of a two-level forest (grmhd_ctx_create_forest).
struct ForestIn {
    int nb[3], bc[6];
    std::vector<int32_t> cmap;
    std::vector<int4> fblk;
};

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
static grmhd_status noalloc_scratch_ptr(DevStats** out) {
    static std::mutex mu;
    static DevStats* ptr[64] = {};
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) { set_err("device ordinal %d out of range", dev); return GRMHD_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return GRMHD_ERR_HIP;
        }
        ptr[dev] = (DevStats*)p;
    }
    *out = ptr[dev];
    return GRMHD_OK;
}

static grmhd_status ctx_create_impl(const grmhd_params* p, const grmhd_grid* g, const grmhd_instance* inst,
                                    int32_t n_inst, const ForestIn* fin, grmhd_ctx** out) {
    if (!p || !g || !out) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    *out = nullptr;
    if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->ngz < 3 || g->nq < 1) {
        set_err("invalid grid (need n>=1, ngz>=3, nq>=1): %d %d %d ngz=%d nq=%lld", g->nx, g->ny, g->nz, g->ngz,
                (long long)g->nq);
        return GRMHD_ERR_INVALID;
    }
    if (g->nq >= GSRC_COPY) {   // the ghost tables keep flags above the block index
        set_err("too many blocks: %lld (at most %d)", (long long)g->nq, GSRC_COPY - 1);
        return GRMHD_ERR_INVALID;
    }
    if (p->eos_kind != GRMHD_EOS_IDEAL_GAS && p->eos_kind != GRMHD_EOS_HYBRID_PWPOLY) { set_err("eos_kind"); return GRMHD_ERR_UNSUPPORTED; }
    if (p->recon < 0 || p->recon > 3 || p->riemann < 0 || p->riemann > 1 || p->emf < 0 || p->emf > 1 ||
        p->thermo_recon < 0 || p->thermo_recon > 1 || (p->fd_order != 2 && p->fd_order != 4 && p->fd_order != 6)) {
        set_err("invalid scheme options"); return GRMHD_ERR_INVALID;
    }
    if (p->eos_kind == GRMHD_EOS_HYBRID_PWPOLY && (p->pwp_n < 1 || p->pwp_n > 8)) { set_err("pwp_n"); return GRMHD_ERR_INVALID; }
    if (p->eos_kind == GRMHD_EOS_IDEAL_GAS && !(p->gamma > 1.0)) { set_err("gamma must be > 1"); return GRMHD_ERR_INVALID; }
    if (p->fd_order / 2 > g->ngz) { set_err("fd_order needs more ghosts"); return GRMHD_ERR_INVALID; }
    // evolve.cpp:476-481: FOFC needs the GS EMF and one ghost layer beyond the 6-point stencil
    if (p->fofc && p->emf != GRMHD_EMF_GS) { set_err("FOFC is not supported with UCT"); return GRMHD_ERR_UNSUPPORTED; }
    if (p->fofc && g->ngz < 4) { set_err("FOFC requires ngz >= 4"); return GRMHD_ERR_INVALID; }
    if (p->fofc && !(p->fofc_dmp_M > 0)) { set_err("fofc_dmp_M must be > 0"); return GRMHD_ERR_INVALID; }
    if (p->fofc_ghost_emf != 0 && p->fofc_ghost_emf != 1) { set_err("fofc_ghost_emf must be 0 or 1"); return GRMHD_ERR_INVALID; }
    if (p->stage_streams != 0 && p->stage_streams != 1) { set_err("stage_streams must be 0 or 1"); return GRMHD_ERR_INVALID; }
    if (p->excise_by_radius != 0 && p->excise_by_radius != 1) { set_err("excise_by_radius must be 0 or 1"); return GRMHD_ERR_INVALID; }
    if (p->excision_cks != 0 && p->excision_cks != 1) { set_err("excision_cks must be 0 or 1"); return GRMHD_ERR_INVALID; }
    // instance / block tables
    std::vector<DevBlk> blk(g->nq, DevBlk{-1, {0, 0, 0}});
    for (int a = 0; a < n_inst; ++a) {
        const grmhd_instance& I = inst[a];
        int64_t nb = (int64_t)I.nb[0] * I.nb[1] * I.nb[2];
        if (I.nb[0] < 1 || I.nb[1] < 1 || I.nb[2] < 1 || I.q0 < 0 || I.q0 + nb > g->nq) {
            set_err("instance %d out of range", a); return GRMHD_ERR_INVALID;
        }
        for (int d = 0; d < 6; ++d) if (I.bc[d] != 0 && I.bc[d] != 1) { set_err("bc kind"); return GRMHD_ERR_INVALID; }
        // periodicity is a property of a dimension (periodic_x|y|z of the
        // brick connectivity, connectivity.cpp:52-69): a periodic face needs a
        // periodic opposite face
        for (int d = 0; d < 3; ++d)
            if ((I.bc[2 * d] == GRMHD_BC_PERIODIC) != (I.bc[2 * d + 1] == GRMHD_BC_PERIODIC)) {
                set_err("instance %d: the two faces of dimension %d must both be periodic or both not", a, d);
                return GRMHD_ERR_INVALID;
            }
        for (int d = 0; d < 3; ++d) if (!(I.idx[d] > 0)) { set_err("idx must be > 0"); return GRMHD_ERR_INVALID; }
        // the source terms' metric derivatives take idx[0] in every direction
        // (grmhd.hh:231-235, written for the reference's cubic cells)
        if (!p->flat_metric && (I.idx[1] != I.idx[0] || I.idx[2] != I.idx[0])) {
            set_err("instance %d: a curved metric needs equal spacing in x, y and z (idx %g %g %g)", a, I.idx[0],
                    I.idx[1], I.idx[2]);
            return GRMHD_ERR_UNSUPPORTED;
        }
        // radius excision at a ghost cell uses its unwrapped position in the
        // reference; this library copies a periodic image's c2p outcome
        if (p->excise_by_radius)
            for (int d = 0; d < 6; ++d)
                if (I.bc[d] == GRMHD_BC_PERIODIC) {
                    set_err("excise_by_radius is not supported with periodic faces (instance %d)", a);
                    return GRMHD_ERR_UNSUPPORTED;
                }
        for (int bk = 0; bk < I.nb[2]; ++bk)
            for (int bj = 0; bj < I.nb[1]; ++bj)
                for (int bi = 0; bi < I.nb[0]; ++bi) {
                    int64_t q = I.q0 + bi + (int64_t)I.nb[0] * (bj + (int64_t)I.nb[1] * bk);
                    if (blk[q].inst >= 0) { set_err("block %lld in two instances", (long long)q); return GRMHD_ERR_INVALID; }
                    blk[q] = DevBlk{a, {bi, bj, bk}};
                }
    }
    for (int64_t q = 0; q < g->nq; ++q) if (blk[q].inst < 0) { set_err("block %lld not covered", (long long)q); return GRMHD_ERR_INVALID; }
    grmhd_ctx* c = new grmhd_ctx();
    c->P = *p;
    DevGrid& G = c->G;
    G.n[0] = g->nx; G.n[1] = g->ny; G.n[2] = g->nz; G.ng = g->ngz; G.nq = g->nq;
    for (int d = 0; d < 3; ++d) G.N[d] = G.n[d] + 2 * G.ng;
    G.ncell = (int64_t)G.N[0] * G.N[1] * G.N[2];
    for (int d = 0; d < 3; ++d) G.nface[d] = (int64_t)(G.N[0] + (d == 0)) * (G.N[1] + (d == 1)) * (G.N[2] + (d == 2));
    G.nedge = (int64_t)(G.N[0] + 1) * (G.N[1] + 1) * (G.N[2] + 1);
    Opts& O = c->O;
    O.atmo = Atmo{p->rho_fl, p->temp_fl, p->ye_fl, p->atmo_tol};
    O.cp = C2PPar{p->c2p_tol, p->max_lorentz, p->max_sigma, p->beta_fallback, p->bh_alp_thresh, p->use_entropy_backup};
    O.ex = ExcPar{p->excision_radius, p->excision_lapse, p->excise_by_radius, p->excision_cks, p->excision_spin};
    O.riemann = p->riemann; O.emf = p->emf; O.thermo = p->thermo_recon; O.fd_order = p->fd_order;
    O.c_limit = p->rusanov_c_limit;
    O.fofc = p->fofc; O.dmp = p->fofc_dmp_enable; O.dmp_M = p->fofc_dmp_M;
    O.eval_counts = nullptr;
    build_eos(c);
    c->blk_h = blk;
    for (int a = 0; a < n_inst; ++a) {
        DevInst d;
        for (int x = 0; x < 3; ++x) { d.nb[x] = inst[a].nb[x]; d.idx[x] = inst[a].idx[x]; d.x0[x] = inst[a].x0[x]; }
        for (int x = 0; x < 6; ++x) d.bc[x] = inst[a].bc[x];
        d.q0 = inst[a].q0;
        c->inst_h.push_back(d);
    }
    auto fail = [&](const char* what, hipError_t e) {
        set_err("%s: %s", what, hipGetErrorString(e));
        free_ctx(c); delete c;
        return GRMHD_ERR_ALLOC;
    };
    hipError_t e;
    if (n_inst > 0) {
        if ((e = hipMalloc(&c->inst_d, sizeof(DevInst) * n_inst)) != hipSuccess) return fail("inst", e);
        if ((e = hipMemcpy(c->inst_d, c->inst_h.data(), sizeof(DevInst) * n_inst, hipMemcpyHostToDevice)) != hipSuccess) return fail("inst copy", e);
    }
    if ((e = hipMalloc(&c->blk_d, sizeof(DevBlk) * g->nq)) != hipSuccess) return fail("blk", e);
    if ((e = hipMemcpy(c->blk_d, blk.data(), sizeof(DevBlk) * g->nq, hipMemcpyHostToDevice)) != hipSuccess) return fail("blk copy", e);
    for (int d = 0; d < 3; ++d) {
        if ((e = hipMalloc(&c->flux[d], sizeof(double) * NHRSC * G.nface[d] * G.nq)) != hipSuccess) return fail("flux", e);
        if ((e = hipMalloc(&c->vbar[d], sizeof(double) * 4 * G.nface[d] * G.nq)) != hipSuccess) return fail("vbar", e);
    }
    if ((e = hipMalloc(&c->emf, sizeof(double) * 3 * G.nedge * G.nq)) != hipSuccess) return fail("emf", e);
    if (p->fofc) {
        for (int d = 0; d < 3; ++d)
            if ((e = hipMalloc(&c->ftag[d], G.nface[d] * G.nq)) != hipSuccess) return fail("fofc face tags", e);
        if ((e = hipMalloc(&c->etag, 3 * G.nedge * G.nq)) != hipSuccess) return fail("fofc edge tags", e);
    }
    if ((e = hipMalloc(&c->ecen, sizeof(double) * 3 * G.ncell * G.nq)) != hipSuccess) return fail("ecen", e);
    // E_center starts at zero, as the reference's freshly allocated view
    if ((e = hipMemset(c->ecen, 0, sizeof(double) * 3 * G.ncell * G.nq)) != hipSuccess) return fail("ecen clear", e);
    if ((e = hipMemset(c->emf, 0, sizeof(double) * 3 * G.nedge * G.nq)) != hipSuccess) return fail("emf clear", e);
    c->n_dtpart = nblk((int64_t)G.n[0] * G.n[1] * G.n[2] * G.nq);
    if ((e = hipMalloc(&c->dtpart, sizeof(double) * c->n_dtpart)) != hipSuccess) return fail("dtpart", e);
    if ((e = hipMalloc(&c->stats, sizeof(DevStats))) != hipSuccess) return fail("stats", e);
    if ((e = hipMemset(c->stats, 0, sizeof(DevStats))) != hipSuccess) return fail("stats clear", e);
    if ((e = hipMalloc(&c->fatal_d, sizeof(unsigned long long))) != hipSuccess) return fail("fatal counter", e);
    if ((e = hipMemset(c->fatal_d, 0, sizeof(unsigned long long))) != hipSuccess) return fail("fatal counter clear", e);
    c->O.fatal_count = c->fatal_d;
    if ((e = hipMalloc(&c->dt_scalar, sizeof(double))) != hipSuccess) return fail("dt scalar", e);
    // 2-D c2p grids with workgroups down to 64 threads, plus per-launch rounding
    c->n_part_cap = nblk_bs(G.ncell * G.nq, 64) + 2 * G.nq + 2;
    if ((e = hipMalloc(&c->stats_part, sizeof(DevStats) * c->n_part_cap)) != hipSuccess) return fail("stats part", e);
    if (!fin && p->emf == GRMHD_EMF_UCT && !p->fofc && p->flat_metric && G.n[0] >= G.ng && G.n[1] >= G.ng &&
        G.n[2] >= G.ng) {
        // Ring faces whose ghost side has a neighbouring block are that
        // block's interior faces; with a flat metric every input of those
        // faces is a copy on both sides (ghost metric values from the caller
        // could differ in the last bit from the neighbour's, hence flat only).
        std::vector<int32_t> nbr((size_t)g->nq * 6, -1);
        std::vector<uint8_t> own((size_t)g->nq, 0);
        for (int64_t q = 0; q < g->nq; ++q) {
            const DevBlk& bq = blk[q];
            const grmhd_instance& I = inst[bq.inst];
            for (int sd = 0; sd < 6; ++sd) {
                const int d = sd / 2, dir = (sd & 1) ? 1 : -1;
                int bc[3] = {bq.b[0], bq.b[1], bq.b[2]};
                bc[d] += dir;
                if (bc[d] < 0 || bc[d] >= I.nb[d]) {
                    if (I.bc[sd] != GRMHD_BC_PERIODIC) { own[q] |= (uint8_t)(1u << sd); continue; }
                    bc[d] = (bc[d] + I.nb[d]) % I.nb[d];
                }
                nbr[(size_t)q * 6 + sd] = (int32_t)(I.q0 + bc[0] + (int64_t)I.nb[0] * (bc[1] + (int64_t)I.nb[1] * bc[2]));
            }
        }
        if ((e = hipMalloc(&c->ring_nbr, sizeof(int32_t) * nbr.size())) != hipSuccess) return fail("ring neighbours", e);
        if ((e = hipMemcpy(c->ring_nbr, nbr.data(), sizeof(int32_t) * nbr.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail("ring neighbours copy", e);
        if ((e = hipMalloc(&c->ring_own, own.size())) != hipSuccess) return fail("ring mask", e);
        if ((e = hipMemcpy(c->ring_own, own.data(), own.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail("ring mask copy", e);
    }
    {
        grmhd_status sst = noalloc_scratch_ptr(&c->noalloc_part);
        if (sst != GRMHD_OK) { free_ctx(c); delete c; return sst; }
    }
    {   // ghost-shell tables (local offsets), cells then x/y/z faces
        std::vector<int32_t> lst;
        for (int fd = -1; fd < 3; ++fd) {
            int E[3] = {G.N[0] + (fd == 0), G.N[1] + (fd == 1), G.N[2] + (fd == 2)};
            int64_t cnt = 0;
            for (int k = 0; k < E[2]; ++k)
                for (int j = 0; j < E[1]; ++j)
                    for (int i = 0; i < E[0]; ++i) {
                        int loc[3] = {i, j, k};
                        bool ghost = false;
                        for (int d = 0; d < 3; ++d) ghost |= loc[d] < G.ng || loc[d] >= G.n[d] + G.ng + (d == fd ? 1 : 0);
                        if (!ghost) continue;
                        lst.push_back((int32_t)(i + (int64_t)E[0] * (j + (int64_t)E[1] * k)));
                        cnt++;
                    }
            if (fd < 0) {
                c->ncg = cnt;
            } else {
                // the face slabs first (face index in [ng, n+ng], exactly one
                // of the other two coordinates in the ghost layers): between
                // RK stages with shared ring faces the edge EMFs read no
                // other ghost face
                c->nfg[fd] = cnt;
                auto slab = [&](int32_t off) {
                    const int loc[3] = {off % E[0], (off / E[0]) % E[1], off / (E[0] * E[1])};
                    if (loc[fd] < G.ng || loc[fd] > G.n[fd] + G.ng) return false;
                    int ng_ = 0;
                    for (int d = 0; d < 3; ++d)
                        if (d != fd) ng_ += loc[d] < G.ng || loc[d] >= G.n[d] + G.ng;
                    return ng_ == 1;
                };
                auto mid = std::stable_partition(lst.end() - cnt, lst.end(), slab);
                c->nfg_slab[fd] = mid - (lst.end() - cnt);
            }
        }
        if ((e = hipMalloc(&c->ghost_list, sizeof(int32_t) * lst.size())) != hipSuccess) return fail("ghost list", e);
        if ((e = hipMemcpy(c->ghost_list, lst.data(), sizeof(int32_t) * lst.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail("ghost list copy", e);
        // per-block sources of every ghost element
        const int64_t per = (int64_t)lst.size();
        std::vector<int2> src((size_t)(per * G.nq));
        FForest ffh{};   // the forest's lookup over host copies of its tables
        if (fin) {
            c->forest = true;
            c->cmap_h = fin->cmap;
            c->fblk_h = fin->fblk;
            for (int d = 0; d < 3; ++d) { ffh.nb[d] = fin->nb[d]; ffh.n[d] = G.n[d]; ffh.N[d] = G.N[d]; ffh.nface[d] = G.nface[d]; }
            for (int d = 0; d < 6; ++d) ffh.bc[d] = fin->bc[d];
            ffh.ng = G.ng; ffh.ncell = G.ncell; ffh.nq = G.nq;
            ffh.cmap = c->cmap_h.data(); ffh.blk = c->fblk_h.data();
            if ((e = hipMalloc(&c->cmap_d, sizeof(int32_t) * c->cmap_h.size())) != hipSuccess) return fail("forest map", e);
            if ((e = hipMemcpy(c->cmap_d, c->cmap_h.data(), sizeof(int32_t) * c->cmap_h.size(), hipMemcpyHostToDevice)) != hipSuccess)
                return fail("forest map copy", e);
            if ((e = hipMalloc(&c->fblk_d, sizeof(int4) * c->fblk_h.size())) != hipSuccess) return fail("forest blocks", e);
            if ((e = hipMemcpy(c->fblk_d, c->fblk_h.data(), sizeof(int4) * c->fblk_h.size(), hipMemcpyHostToDevice)) != hipSuccess)
                return fail("forest blocks copy", e);
            c->ff = ffh;
            c->ff.cmap = c->cmap_d; c->ff.blk = c->fblk_d;
        }
        for (int64_t q = 0; q < G.nq; ++q) {
            const DevBlk& b = blk[q];
            const DevInst& I = c->inst_h[b.inst];
            for (int64_t x = 0; x < per; ++x) {
                int fd = x < c->ncg ? -1 : (x < c->ncg + c->nfg[0] ? 0 : (x < c->ncg + c->nfg[0] + c->nfg[1] ? 1 : 2));
                int E0 = G.N[0] + (fd == 0), E1 = G.N[1] + (fd == 1);
                int off = lst[x];
                int loc[3] = {off % E0, (off / E0) % E1, off / (E0 * E1)};
                int qs, so;
                if (fin) {
                    forest_source(ffh, (int)q, fd, loc, qs, so);
                } else {
                    ghost_source(G, I, b, fd, loc, qs, so);
                    if (fd < 0 && ghost_needs_own_c2p(G, I, b, loc[0], loc[1], loc[2])) qs |= GSRC_OWN;
                }
                src[(size_t)(q * per + x)] = make_int2(qs, so);
            }
        }
        // Beyond an outflow face every ghost layer holds the same clamped
        // conserved and, through the clamped face copies, the same cell-centred
        // B (both faces of a cell beyond the face copy the boundary face): with
        // a flat metric and no radius-based excision its c2p inputs are those
        // of the cell with every outflow-exceeding coordinate moved to the first
        // ghost layer.  Only that cell solves; the others copy its outputs.
        std::vector<int4> copyg;
        std::vector<uint8_t> is_copy((size_t)(per * G.nq), 0);
        if (!fin && c->P.flat_metric && !c->P.excise_by_radius) {
            std::unordered_map<int32_t, int64_t> pos;   // local cell offset -> entry
            for (int64_t x = 0; x < c->ncg; ++x) pos[lst[x]] = x;
            for (int64_t q = 0; q < G.nq; ++q) {
                const DevBlk& b = blk[q];
                const DevInst& I = c->inst_h[b.inst];
                for (int64_t x = 0; x < c->ncg; ++x) {
                    if (!(src[(size_t)(q * per + x)].x & GSRC_OWN)) continue;
                    const int off = lst[x];
                    int loc[3] = {off % G.N[0], (off / G.N[0]) % G.N[1], off / (G.N[0] * G.N[1])};
                    int rl[3] = {loc[0], loc[1], loc[2]};
                    for (int d = 0; d < 3; ++d) {
                        const int Nt = I.nb[d] * G.n[d];
                        const int g = b.b[d] * G.n[d] + loc[d] - G.ng;
                        if (g < 0 && I.bc[2 * d] == 0) rl[d] = -1 - b.b[d] * G.n[d] + G.ng;
                        if (g >= Nt && I.bc[2 * d + 1] == 0) rl[d] = Nt - b.b[d] * G.n[d] + G.ng;
                    }
                    if (rl[0] == loc[0] && rl[1] == loc[1] && rl[2] == loc[2]) continue;
                    bool inside = true;
                    for (int d = 0; d < 3; ++d) inside &= rl[d] >= 0 && rl[d] < G.N[d];
                    if (!inside) continue;
                    const int32_t roff = (int32_t)cidx(G, rl[0], rl[1], rl[2]);
                    auto it = pos.find(roff);
                    if (it == pos.end() || !(src[(size_t)(q * per + it->second)].x & GSRC_OWN)) continue;
                    is_copy[(size_t)(q * per + x)] = 1;
                    src[(size_t)(q * per + x)] = make_int2((int)q | GSRC_COPY, roff);
                    copyg.push_back(make_int4((int)q, off, (int)q, roff));
                }
            }
        }
        std::vector<int2> own, amr_c, amr_f, amr_fm;
        std::vector<int4> owng;
        for (int64_t q = 0; q < G.nq; ++q)
            for (int64_t x = 0; x < per; ++x) {
                const int2 sr = src[(size_t)(q * per + x)];
                if (is_copy[(size_t)(q * per + x)]) continue;
                if (x >= c->ncg) {
                    if (sr.x & GSRC_AMR) {
                        const int fd = x < c->ncg + c->nfg[0] ? 0 : (x < c->ncg + c->nfg[0] + c->nfg[1] ? 1 : 2);
                        const int E0 = G.N[0] + (fd == 0), E1 = G.N[1] + (fd == 1);
                        const int loc[3] = {lst[x] % E0, (lst[x] / E0) % E1, lst[x] / (E0 * E1)};
                        int lev, gpos[3];
                        ff_global(ffh, (int)q, fd, loc, &lev, gpos);
                        (lev == 1 && (gpos[fd] & 1) ? amr_fm : amr_f).push_back(make_int2((int)q, (fd << 24) | lst[x]));
                    }
                    continue;
                }
                if (sr.x & GSRC_OWN) own.push_back(make_int2((int)q, lst[x]));
                if (sr.x & GSRC_AMR) amr_c.push_back(make_int2((int)q, lst[x]));
                else if (sr.x & GSRC_OWN) owng.push_back(make_int4((int)q, lst[x], sr.x & (GSRC_AMR - 1), sr.y));
            }
        c->n_own = (int64_t)own.size();
        c->n_own_gather = (int64_t)owng.size();
        c->n_copy = (int64_t)copyg.size();
        std::vector<int4> ofl;
        if (c->ring_own) {
            int64_t base = c->ncg;
            for (int fd = 0; fd < 3; ++fd) {
                const int E0 = G.N[0] + (fd == 0), E1 = G.N[1] + (fd == 1);
                const size_t before = ofl.size();
                for (int64_t q = 0; q < G.nq; ++q) {
                    const DevBlk& b = blk[q];
                    const DevInst& I = c->inst_h[b.inst];
                    for (int64_t x = base + c->nfg_slab[fd]; x < base + c->nfg[fd]; ++x) {
                        // a face of a ghost cell beyond an outflow face (its
                        // c2p, or the one it copies, reads the cell's faces)
                        const int loc[3] = {lst[x] % E0, (lst[x] / E0) % E1, lst[x] / (E0 * E1)};
                        int lo[3] = {loc[0], loc[1], loc[2]};
                        lo[fd] -= 1;
                        const bool hi_ok = loc[fd] < G.N[fd], lo_ok = loc[fd] >= 1;
                        if (!(hi_ok && ghost_needs_own_c2p(G, I, b, loc[0], loc[1], loc[2])) &&
                            !(lo_ok && ghost_needs_own_c2p(G, I, b, lo[0], lo[1], lo[2])))
                            continue;
                        const int2 sr = src[(size_t)(q * per + x)];
                        ofl.push_back(make_int4((int)q, lst[x], sr.x & (GSRC_COPY - 1), sr.y));
                    }
                }
                c->n_ofl[fd] = (int64_t)(ofl.size() - before);
                base += c->nfg[fd];
            }
        }
        c->n_amr_cells = (int64_t)amr_c.size();
        c->n_amr_faces = (int64_t)amr_f.size();
        c->n_amr_faces_mid = (int64_t)amr_fm.size();
        auto upload = [&](auto** dst, const auto& v, const char* what) -> bool {
            if (v.empty()) return true;
            const size_t bytes = sizeof(v[0]) * v.size();
            if ((e = hipMalloc((void**)dst, bytes)) != hipSuccess) return false;
            if ((e = hipMemcpy(*dst, v.data(), bytes, hipMemcpyHostToDevice)) != hipSuccess) return false;
            (void)what;
            return true;
        };
        if (!upload(&c->own_cells, own, "own cells")) return fail("own cells", e);
        if (!upload(&c->own_gather, owng, "own gather")) return fail("own gather", e);
        if (!upload(&c->copy_gather, copyg, "copy gather")) return fail("outflow ghost copies", e);
        if (!upload(&c->ofl_faces, ofl, "outflow faces")) return fail("outflow ghost faces", e);
        if (!upload(&c->amr_cells, amr_c, "amr cells")) return fail("coarse-fine cells", e);
        if (!upload(&c->amr_faces, amr_f, "amr faces")) return fail("coarse-fine faces", e);
        if (!upload(&c->amr_faces_mid, amr_fm, "amr faces mid")) return fail("coarse-fine interior faces", e);
        if (fin) {
            std::vector<RefluxFace> rf;
            std::vector<RefluxEdge> re;
            forest_reflux(ffh, rf, re);
            c->n_rflux = (int64_t)rf.size();
            c->n_remf = (int64_t)re.size();
            if (!upload(&c->rflux, rf, "reflux faces")) return fail("reflux faces", e);
            if (!upload(&c->remf, re, "reflux edges")) return fail("reflux edges", e);
        }
        if ((e = hipMalloc(&c->ghost_src, sizeof(int2) * src.size())) != hipSuccess) return fail("ghost sources", e);
        if ((e = hipMemcpy(c->ghost_src, src.data(), sizeof(int2) * src.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail("ghost sources copy", e);
    }
    *out = c;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_ctx_create(const grmhd_params* p, const grmhd_grid* g, const grmhd_instance* inst,
                                         int32_t n_inst, grmhd_ctx** out) {
    return ctx_create_impl(p, g, inst, n_inst, nullptr, out);
}

extern "C" grmhd_status grmhd_ctx_create_forest(const grmhd_params* p, const grmhd_grid* g, const grmhd_forest* f,
                                                grmhd_ctx** out) {
    if (!p || !g || !f || !out) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    *out = nullptr;
    for (int d = 0; d < 3; ++d)
        if (f->nb[d] < 1 || !(f->idx[d] > 0)) { set_err("forest: nb must be >= 1 and idx > 0"); return GRMHD_ERR_INVALID; }
    for (int d = 0; d < 6; ++d) if (f->bc[d] != 0 && f->bc[d] != 1) { set_err("forest: bc kind"); return GRMHD_ERR_INVALID; }
    for (int d = 0; d < 3; ++d)
        if ((f->bc[2 * d] == GRMHD_BC_PERIODIC) != (f->bc[2 * d + 1] == GRMHD_BC_PERIODIC)) {
            set_err("forest: the two faces of dimension %d must both be periodic or both not", d);
            return GRMHD_ERR_INVALID;
        }
    const int64_t nc = (int64_t)f->nb[0] * f->nb[1] * f->nb[2];
    if (nc > (1 << 24)) { set_err("forest: too many coarse blocks"); return GRMHD_ERR_INVALID; }
    if (f->n_refined < 0 || f->n_refined > nc || (f->n_refined > 0 && !f->refined)) {
        set_err("forest: bad refined list"); return GRMHD_ERR_INVALID;
    }
    std::vector<char> ref((size_t)nc, 0);
    for (int r = 0; r < f->n_refined; ++r) {
        const int32_t id = f->refined[r];
        if (id < 0 || id >= nc || ref[id]) { set_err("forest: refined id %d out of range or repeated", id); return GRMHD_ERR_INVALID; }
        ref[id] = 1;
    }
    if (g->nx % 2 || g->ny % 2 || g->nz % 2 || g->ngz % 2) {
        set_err("forest: block extents and ngz must be even (coarse-fine operators work on 2x2x2 children)");
        return GRMHD_ERR_INVALID;
    }
    // grace_initialize.cpp:243-248; the divergence-preserving prolongation
    // (prolongation_kernels.hh:330-381) is written for cubic cells
    if (g->nx != g->ny || g->ny != g->nz || f->idx[0] != f->idx[1] || f->idx[1] != f->idx[2]) {
        set_err("forest: blocks must be cubic with equal spacing in x, y and z");
        return GRMHD_ERR_INVALID;
    }
    if (g->nq != nc + 7 * (int64_t)f->n_refined) {
        set_err("forest: grid.nq must be %lld (coarse blocks + 7 per refined block)", (long long)(nc + 7 * (int64_t)f->n_refined));
        return GRMHD_ERR_INVALID;
    }
    if ((int64_t)(g->nx + 2 * g->ngz + 1) * (g->ny + 2 * g->ngz + 1) * (g->nz + 2 * g->ngz + 1) >= (1 << 24)) {
        set_err("forest: blocks too large"); return GRMHD_ERR_INVALID;
    }
    ForestIn fin;
    for (int d = 0; d < 3; ++d) fin.nb[d] = f->nb[d];
    for (int d = 0; d < 6; ++d) fin.bc[d] = f->bc[d];
    fin.cmap.assign((size_t)nc, 0);
    int32_t nq = 0;
    for (int64_t b = 0; b < nc; ++b)
        if (!ref[b]) {
            fin.cmap[b] = nq++;
            fin.fblk.push_back(make_int4(0, (int)(b % f->nb[0]), (int)((b / f->nb[0]) % f->nb[1]), (int)(b / ((int64_t)f->nb[0] * f->nb[1]))));
        }
    for (int64_t b = 0; b < nc; ++b)
        if (ref[b]) {
            fin.cmap[b] = -(1 + nq);
            const int cb[3] = {(int)(b % f->nb[0]), (int)((b / f->nb[0]) % f->nb[1]), (int)(b / ((int64_t)f->nb[0] * f->nb[1]))};
            for (int ch = 0; ch < 8; ++ch, ++nq)
                fin.fblk.push_back(make_int4(1, 2 * cb[0] + (ch & 1), 2 * cb[1] + ((ch >> 1) & 1), 2 * cb[2] + ((ch >> 2) & 1)));
        }
    std::vector<grmhd_instance> inst((size_t)nq);
    for (int32_t q = 0; q < nq; ++q) {
        grmhd_instance& I = inst[q];
        const double s = fin.fblk[q].x ? 2.0 : 1.0;
        const int crd[3] = {fin.fblk[q].y, fin.fblk[q].z, fin.fblk[q].w};
        for (int d = 0; d < 3; ++d) {
            I.nb[d] = 1; I.idx[d] = f->idx[d] * s;
            I.x0[d] = f->x0[d] + (double)crd[d] * (d == 0 ? g->nx : (d == 1 ? g->ny : g->nz)) / I.idx[d];
        }
        for (int d = 0; d < 6; ++d) I.bc[d] = f->bc[d];
        I.q0 = q;
    }
    return ctx_create_impl(p, g, inst.data(), nq, &fin, out);
}

extern "C" grmhd_status grmhd_forest_block(const grmhd_ctx* c, int64_t q, int32_t* level, int32_t coords[3]) {
    if (!c || !level || !coords) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    if (!c->forest) { set_err("not a forest context"); return GRMHD_ERR_INVALID; }
    if (q < 0 || q >= c->G.nq) { set_err("block %lld out of range", (long long)q); return GRMHD_ERR_INVALID; }
    const int4 b = c->fblk_h[(size_t)q];
    *level = b.x; coords[0] = b.y; coords[1] = b.z; coords[2] = b.w;
    return GRMHD_OK;
}

extern "C" void grmhd_ctx_destroy(grmhd_ctx* c) {
    if (!c) return;
    (void)hipDeviceSynchronize();
    free_ctx(c);
    delete c;
}

static grmhd_status launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_err("%s launch: %s", what, hipGetErrorString(e)); return GRMHD_ERR_HIP; }
    return GRMHD_OK;
}

// ---- dispatch helpers -------------------------------------------------------
// only != null: FOFC correction pass over the tagged faces with the options O
// (donor cell + LLF); otherwise every face the later passes read.
template <typename EOS, int REC, bool FLAT>
static void launch_flux3(grmhd_ctx* c, const EOS& eos, const double* aux, const double* ms, int mnv,
                         const double* const B[3], hipStream_t s0, const Opts& O, uint8_t* const* only) {
    const DevGrid& G = c->G;
    const bool uct = c->P.emf != GRMHD_EMF_GS;
    int nf[3];
    for (int d = 0; d < 3; ++d) {
        const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
        const int Rn = G.n[d] + 1, n0 = G.n[t0], n1 = G.n[t1];
        nf[d] = uct ? Rn * (n0 * n1 + 2 * G.ng * (n0 + n1)) : (Rn + 2 * O.fofc) * G.N[t0] * G.N[t1];
    }
    static const char* fam[3] = {"flux_x", "flux_y", "flux_z"};
    for (int d = 0; d < 3; ++d) {
        const hipStream_t s = c->flux_s[d] ? c->flux_s[d] : s0;
        Scope sc(c, only ? "fofc" : fam[d], s);
        const int bs = c->tune.flux[d];
        const int wpb = (int)nblk_bs(nf[d], bs);
        const dim3 grid((unsigned)(wpb * G.nq));
        const uint8_t* od = only ? only[d] : nullptr;
        const uint8_t* ro = (only || !c->ring_copy_now) ? nullptr : c->ring_own;
        // the prefetching variant is built for WENO-Z only (the other
        // reconstructions load one or three values per side)
        bool pf = false;
        if constexpr (REC == 0) pf = c->tune.flux_prefetch[d] != 0 && !only;
        const int32_t* pn = ro ? c->ring_nbr : nullptr;
        // (x faces interpolate from four adjacent cells of a row, which is
        // cheaper than streaming 13 cached values: cache for y and z only)
        const double* fm = (!FLAT && d > 0 && c->fmet[d] && ms == c->mder_src && mnv == NSTATE) ? c->fmet[d] : nullptr;
#define FLUXL(DD, PFV) k_flux<EOS, REC, FLAT, DD, PFV><<<grid, bs, 0, s>>>(G, eos, O, aux, ms, mnv, B[DD], c->flux[DD], c->vbar[DD], wpb, od, ro, pn, fm)
        if constexpr (REC == 0) {
            if (pf) { if (d == 0) FLUXL(0, true); else if (d == 1) FLUXL(1, true); else FLUXL(2, true); }
            else { if (d == 0) FLUXL(0, false); else if (d == 1) FLUXL(1, false); else FLUXL(2, false); }
        } else {
            if (d == 0) FLUXL(0, false); else if (d == 1) FLUXL(1, false); else FLUXL(2, false);
        }
#undef FLUXL
    }
}
template <typename EOS, bool FLAT>
static void flux_rec(grmhd_ctx* c, const EOS& eos, const double* aux, const double* ms, int mnv,
                     const double* const B[3], hipStream_t s, int recon, const Opts& O, uint8_t* const* only) {
    switch (recon) {
        case 0: launch_flux3<EOS, 0, FLAT>(c, eos, aux, ms, mnv, B, s, O, only); break;
        case 1: launch_flux3<EOS, 1, FLAT>(c, eos, aux, ms, mnv, B, s, O, only); break;
        case 2: launch_flux3<EOS, 2, FLAT>(c, eos, aux, ms, mnv, B, s, O, only); break;
        default: launch_flux3<EOS, 3, FLAT>(c, eos, aux, ms, mnv, B, s, O, only); break;
    }
}
static void launch_fluxes(grmhd_ctx* c, const double* aux, const double* ms, int mnv, const double* const B[3],
                          hipStream_t s, bool fofc_pass = false) {
    const bool flat = c->P.flat_metric != 0;
    Opts O = c->O;
    int recon = c->P.recon;
    uint8_t* const* only = nullptr;
    if (fofc_pass) {   // apply_fofc_correction: donor cell + LLF (evolve.cpp:749-771)
        O.riemann = GRMHD_RIEMANN_LLF;
        recon = GRMHD_RECON_DONOR;
        only = c->ftag;
    }
    if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS) {
        if (flat) flux_rec<EosIdeal, true>(c, c->ideal, aux, ms, mnv, B, s, recon, O, only);
        else flux_rec<EosIdeal, false>(c, c->ideal, aux, ms, mnv, B, s, recon, O, only);
    } else {
        if (flat) flux_rec<EosHybrid, true>(c, c->hybrid, aux, ms, mnv, B, s, recon, O, only);
        else flux_rec<EosHybrid, false>(c, c->hybrid, aux, ms, mnv, B, s, recon, O, only);
    }
}
// 1-D grids of the per-site kernels: wpb workgroups per block over its
// (n+1)^3 sites, times the blocks
template <typename F>
static void site_launch(const DevGrid& G, int bs, F&& f) {
    const int sites = (G.n[0] + 1) * (G.n[1] + 1) * (G.n[2] + 1);
    const int wpb = (int)nblk_bs(sites, bs);
    f(dim3((unsigned)(wpb * G.nq)), wpb);
}
static void launch_emf(grmhd_ctx* c, const double* const B[3], hipStream_t s) {
    Scope sc(c, "emf", s);
    const DevGrid& G = c->G;
    if (c->P.emf == GRMHD_EMF_GS) {
        const int wide = (c->P.fofc && c->P.fofc_ghost_emf) ? 1 : 0;
        const int sites = (G.n[0] + 1 + 2 * wide) * (G.n[1] + 1 + 2 * wide) * (G.n[2] + 1 + 2 * wide);
        const int wpb = (int)nblk_bs(sites, c->tune.emf);
        k_emf_gs<<<dim3((unsigned)(wpb * G.nq)), c->tune.emf, 0, s>>>(G, c->vbar[0], c->vbar[1], c->vbar[2], c->ecen,
                                                                   c->flux[0], c->flux[1], c->flux[2], c->emf, wpb,
                                                                   nullptr, wide);
        return;
    }
    site_launch(G, c->tune.emf, [&](dim3 g, int wpb) {
        switch (c->P.recon) {
            case 0: k_emf_uct<0><<<g, c->tune.emf, 0, s>>>(G, B[0], B[1], B[2], c->vbar[0], c->vbar[1], c->vbar[2], c->emf, wpb); break;
            case 1: k_emf_uct<1><<<g, c->tune.emf, 0, s>>>(G, B[0], B[1], B[2], c->vbar[0], c->vbar[1], c->vbar[2], c->emf, wpb); break;
            case 2: k_emf_uct<2><<<g, c->tune.emf, 0, s>>>(G, B[0], B[1], B[2], c->vbar[0], c->vbar[1], c->vbar[2], c->emf, wpb); break;
            default: k_emf_uct<3><<<g, c->tune.emf, 0, s>>>(G, B[0], B[1], B[2], c->vbar[0], c->vbar[1], c->vbar[2], c->emf, wpb); break;
        }
    });
}
// k_update_pairs needs an even row width (the cell and y/z-face arrays then
// put a pair's two values in one aligned 16-byte word) and 16-byte aligned
// arrays
static bool update_pairs_ok(const DevGrid& G, const UpdArgs& a) {
    auto al = [](const void* p) { return p == nullptr || ((uintptr_t)p & 15) == 0; };
    bool ok = (G.N[0] % 2) == 0 && al(a.nst) && al(a.ast) && al(a.bst);
    for (int d = 1; d < 3; ++d) ok = ok && al(a.flux[d]) && al(a.nB[d]) && al(a.aB[d]) && al(a.bB[d]);
    return ok;
}
static void launch_update(grmhd_ctx* c, UpdArgs& a, hipStream_t s) {
    Scope sc(c, "update", s);
    const DevGrid& G = c->G;
    a.inst = c->inst_d; a.blk = c->blk_d;
    for (int d = 0; d < 3; ++d) a.flux[d] = c->flux[d];
    a.emf = c->emf;
    a.mder = (c->mder && a.mstate == c->mder_src && a.mnv == NSTATE) ? c->mder : nullptr;
    if (c->P.flat_metric && update_pairs_ok(G, a)) {
        const int pairs = (((G.n[0] + G.ng) & ~1) - (G.ng & ~1)) / 2 + 1;
        const int wpb = (int)nblk_bs((int64_t)pairs * (G.n[1] + 1) * (G.n[2] + 1), c->tune.update);
        k_update_pairs<<<dim3((unsigned)(wpb * G.nq)), c->tune.update, 0, s>>>(G, a, pairs, wpb);
        return;
    }
    site_launch(G, c->tune.update, [&](dim3 g, int wpb) {
        if (c->P.flat_metric) k_update<true><<<g, c->tune.update, 0, s>>>(G, a, c->P.fd_order, wpb);
        else if (a.mder) k_update<false, true><<<g, c->tune.update, 0, s>>>(G, a, c->P.fd_order, wpb);
        else k_update<false><<<g, c->tune.update, 0, s>>>(G, a, c->P.fd_order, wpb);
    });
}
static void launch_gather(grmhd_ctx* c, const GatherArgs& a, hipStream_t s) {
    const DevGrid& G = c->G;
    const int per = (int)(c->ncg + c->nfg[0] + c->nfg[1] + c->nfg[2]);
    if (a.B[0]) {
        Scope sc(c, "ghost_faces", s);
        const int64_t* m = a.faces_slab ? c->nfg_slab : c->nfg;
        const int wpb = (int)nblk_bs(m[0] + m[1] + m[2], c->tune.gather_faces);
        k_gather_faces<<<dim3((unsigned)(wpb * G.nq)), c->tune.gather_faces, 0, s>>>(
            c->ghost_list, c->ghost_src, per, (int)c->ncg, (int)c->nfg[0], (int)c->nfg[1], (int)m[0], (int)m[1],
            (int)m[2], wpb, a.B[0], a.B[1], a.B[2], G.nface[0], G.nface[1], G.nface[2]);
        if (a.faces_slab) {
            int64_t o = 0;
            for (int d = 0; d < 3; ++d) {
                if (c->n_ofl[d] > 0) {
                    GatherChans fc{};
                    fc.c[0] = GatherChan{a.B[d], 1, 0, G.nface[d]};
                    fc.nch = 1;
                    k_gather_list<<<dim3(nblk(c->n_ofl[d]), 1, 1), 256, 0, s>>>(c->ofl_faces + o, (int)c->n_ofl[d], fc);
                }
                o += c->n_ofl[d];
            }
        }
        // the second pass reads what the copies and the first pass wrote
        if (c->n_amr_faces > 0)
            k_amr_faces<<<nblk(c->n_amr_faces), 256, 0, s>>>(c->amr_faces, (int)c->n_amr_faces, a.B[0], a.B[1], a.B[2], c->ff);
        if (c->n_amr_faces_mid > 0)
            k_amr_faces_mid<<<nblk(c->n_amr_faces_mid), 256, 0, s>>>(c->amr_faces_mid, (int)c->n_amr_faces_mid, a.B[0],
                                                                     a.B[1], a.B[2], c->ff);
    }
    if (a.cell_mode == 0) return;
    Scope sc(c, "ghost", s);
    GatherChans ch{};
    auto add = [&](double* base, int nv, int v) { ch.c[ch.nch++] = GatherChan{base, nv, v, G.ncell}; };
    if (a.st)
        for (int v = a.v0; v < a.v0 + a.vc; ++v) add(a.st, a.nv, v);
    if (a.aux)
        for (int v = 0; v < NAUX; ++v)
            if (a.aux_mask & (1u << v)) add(a.aux, NAUX, v);
    if (a.ecen)
        for (int v = 0; v < 3; ++v) add(a.ecen, 3, v);
    if (ch.nch == 0) return;
    if (a.copies_only) {
        if (c->n_copy > 0)
            k_gather_list<<<dim3(nblk(c->n_copy), 1, ch.nch), 256, 0, s>>>(c->copy_gather, (int)c->n_copy, ch);
        return;
    }
    // coarse-fine ghost cells of the state (they run their own c2p)
    const bool amr = a.st && a.cell_mode <= 2 && c->n_amr_cells > 0;
    if (a.cell_mode == 2) {
        if (c->n_own_gather > 0)
            k_gather_list<<<dim3(nblk(c->n_own_gather), 1, ch.nch), 256, 0, s>>>(c->own_gather, (int)c->n_own_gather, ch);
        if (amr) k_amr_cells<<<nblk(c->n_amr_cells), 256, 0, s>>>(c->amr_cells, (int)c->n_amr_cells, a.st, a.nv, a.v0, a.vc, c->ff);
        return;
    }
    const int wpb = (int)nblk_bs(c->ncg, c->tune.gather_cells);
    k_gather_cells<<<dim3((unsigned)(wpb * G.nq)), c->tune.gather_cells, 0, s>>>(c->ghost_list, c->ghost_src, per,
                                                                                 (int)c->ncg, wpb, a.cell_mode,
                                                                                 a.own_copies ? 1 : 0, ch);
    // every ghost (mode 1): the copying outflow ghosts take the first-layer
    // cell the launch above filled
    if (a.cell_mode == 1 && c->n_copy > 0)
        k_gather_list<<<dim3(nblk(c->n_copy), 1, ch.nch), 256, 0, s>>>(c->copy_gather, (int)c->n_copy, ch);
    if (amr) k_amr_cells<<<nblk(c->n_amr_cells), 256, 0, s>>>(c->amr_cells, (int)c->n_amr_cells, a.st, a.nv, a.v0, a.vc, c->ff);
}

// reflux_correct_fluxes / reflux_correct_emfs (refluxing.cpp:547-806): after
// the edge EMFs, before the update that consumes both.
// part: 0 both, 1 the fluxes only, 2 the edge EMFs only.
static void launch_reflux(grmhd_ctx* c, hipStream_t s, int part = 0) {
    if (!c->forest) return;
    Scope sc(c, "reflux", s);
    const DevGrid& G = c->G;
    if (c->n_rflux > 0 && part != 2)
        k_reflux_flux<<<dim3(nblk(c->n_rflux), NHRSC), 256, 0, s>>>(c->rflux, (int)c->n_rflux, c->flux[0], c->flux[1],
                                                                    c->flux[2], G.nface[0], G.nface[1], G.nface[2]);
    if (c->n_remf > 0 && part != 1) k_reflux_emf<<<nblk(c->n_remf), 256, 0, s>>>(c->remf, (int)c->n_remf, c->emf);
}

// c2p over the cells of every block: the interior cells, plus either every
// ghost cell (all) or only the ghost cells that need their own c2p.
static void launch_c2p_cells(grmhd_ctx* c, double* st, int snv, const double* ms, int mnv, const double* const B[3],
                             double* aux, bool with_stats, bool selective, hipStream_t s,
                             hipStream_t s_list = nullptr) {
    Scope sc(c, "c2p", s);
    const DevGrid& G = c->G;
    double* ec = c->P.emf == GRMHD_EMF_GS ? c->ecen : nullptr;
    const bool flat = c->P.flat_metric != 0;
    const int ni = G.n[0] * G.n[1] * G.n[2];
    const int bs = c->tune.c2p;
    int64_t npart = 0;
    const bool warm = c->warm_newton > 0 && c->P.eos_kind == GRMHD_EOS_IDEAL_GAS;
    MlpDev md = (c->mlp && !warm) ? c->mlp->dev() : MlpDev{};
    if (warm) md.newton = c->warm_newton;
    // the explicit-list launch may go to its own stream (no statistics then:
    // k_stats_final below reads every partial record on s)
    const hipStream_t sl = (s_list && !with_stats) ? s_list : s;
    auto launch = [&](int cset, int ne, dim3 grid, int wpb) {
        DevStats* part = with_stats ? c->stats_part + npart : nullptr;
        npart += (int64_t)grid.x * grid.y;
        const hipStream_t ls = cset == CS_LIST ? sl : s;
#define C2PL(EOSV, FL, LR) k_c2p_cells<decltype(EOSV), FL, LR><<<grid, bs, 0, ls>>>(G, EOSV, c->O, c->inst_d, c->blk_d, st, snv, ms, mnv, B[0], B[1], B[2], aux, ec, part, cset, c->ghost_list, c->own_cells, ne, wpb, md)
        if (warm) {
            if (flat) C2PL(c->ideal, true, 2); else C2PL(c->ideal, false, 2);
        } else if (c->mlp) {
            if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS) { if (flat) C2PL(c->ideal, true, 1); else C2PL(c->ideal, false, 1); }
            else { if (flat) C2PL(c->hybrid, true, 1); else C2PL(c->hybrid, false, 1); }
        } else {
            if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS) { if (flat) C2PL(c->ideal, true, 0); else C2PL(c->ideal, false, 0); }
            else { if (flat) C2PL(c->hybrid, true, 0); else C2PL(c->hybrid, false, 0); }
        }
#undef C2PL
    };
    {
        const int wi = (int)nblk_bs(ni, bs), ws = (int)nblk_bs(c->ncg, bs);
        launch(CS_INTERIOR, ni, dim3((unsigned)(wi * G.nq)), wi);
        if (!selective) launch(CS_SHELL, (int)c->ncg, dim3((unsigned)(ws * G.nq)), ws);
    }
    if (selective && c->n_own > 0) launch(CS_LIST, (int)c->n_own, dim3(nblk_bs(c->n_own, bs), 1), 0);
    if (with_stats) {
        k_stats_final<<<1, 256, 0, s>>>(c->stats_part, npart, c->stats);
        c->last_c2p_n = G.ncell * G.nq;
    }
}

// flag_fofc_cells + apply_fofc_correction for one substep: clear the tags,
// flag, then recompute the tagged faces (donor + LLF) and GS edge EMFs.
// lin = ca A + cb B is the stage's linear combination (cb = 0: A alone).
static void launch_fofc(grmhd_ctx* c, const double* ast, int anv, const double* bst, int bnv, double ca, double cb,
                        const double* ms, const double* const oB[3], const double* aux, const double* dt_dev,
                        double dtfact, hipStream_t s) {
    Scope sc(c, "fofc", s);
    const DevGrid& G = c->G;
    for (int d = 0; d < 3; ++d)
        k_fill_u8<<<nblk(G.nface[d] * G.nq), 256, 0, s>>>(c->ftag[d], G.nface[d] * G.nq);
    k_fill_u8<<<nblk(3 * G.nedge * G.nq), 256, 0, s>>>(c->etag, 3 * G.nedge * G.nq);
    FofcArgs a{};
    a.ast = ast; a.anv = anv; a.bst = bst; a.bnv = bnv; a.ca = ca; a.cb = cb;
    a.mstate = ms; a.mnv = NSTATE;
    for (int d = 0; d < 3; ++d) { a.flux[d] = c->flux[d]; a.oB[d] = oB[d]; a.ftag[d] = c->ftag[d]; }
    a.emf = c->emf; a.etag = c->etag;
    a.dt = dt_dev; a.dtfact = dtfact;
    a.inst = c->inst_d; a.blk = c->blk_d;
    const int cells = (G.n[0] + 2) * (G.n[1] + 2) * (G.n[2] + 2);
    const int wpb = (int)nblk(cells);
    const dim3 grid((unsigned)(wpb * G.nq));
    const bool flat = c->P.flat_metric != 0;
    if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS) {
        if (flat) k_fofc_flag<EosIdeal, true><<<grid, 256, 0, s>>>(G, c->ideal, c->O, a, wpb);
        else k_fofc_flag<EosIdeal, false><<<grid, 256, 0, s>>>(G, c->ideal, c->O, a, wpb);
    } else {
        if (flat) k_fofc_flag<EosHybrid, true><<<grid, 256, 0, s>>>(G, c->hybrid, c->O, a, wpb);
        else k_fofc_flag<EosHybrid, false><<<grid, 256, 0, s>>>(G, c->hybrid, c->O, a, wpb);
    }
    launch_fluxes(c, aux, ms, NSTATE, oB, s, true);
    site_launch(G, c->tune.emf, [&](dim3 g, int w) {
        k_emf_gs<<<g, c->tune.emf, 0, s>>>(G, c->vbar[0], c->vbar[1], c->vbar[2], c->ecen, c->flux[0], c->flux[1],
                                           c->flux[2], c->emf, w, c->etag, 0);
    });
}

// ---- ABI entry points -------------------------------------------------------
extern "C" grmhd_status grmhd_advance_substep(grmhd_ctx* c, const double* dt_dev, double dtfact, double* new_state,
                                              const double* old_state, double* const new_B[3],
                                              const double* const old_B[3], const double* aux, grmhd_stream_t stream) {
    if (!c || !dt_dev || !new_state || !old_state || !new_B || !old_B || !aux) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    hipStream_t s = (hipStream_t)stream;
    launch_fluxes(c, aux, old_state, NSTATE, old_B, s);
    launch_emf(c, old_B, s);
    if (c->P.fofc)
        launch_fofc(c, new_state, NSTATE, nullptr, 0, 1.0, 0.0, old_state, old_B, aux, dt_dev, dtfact, s);
    launch_reflux(c, s);
    UpdArgs a{};
    a.nst = new_state; a.nnv = NSTATE;
    a.ast = new_state; a.anv = NSTATE; a.bst = nullptr; a.bnv = 0; a.ca = 1.0; a.cb = 0.0;
    a.mstate = old_state; a.mnv = NSTATE; a.aux = aux;
    for (int d = 0; d < 3; ++d) { a.nB[d] = new_B[d]; a.aB[d] = new_B[d]; a.bB[d] = nullptr; }
    a.dt = dt_dev; a.dtfact = dtfact;
    launch_update(c, a, s);
    return launch_check("advance_substep");
}

extern "C" grmhd_status grmhd_advance_substep_t(grmhd_ctx* c, double t, double dt, double dtfact, double* new_state,
                                                const double* old_state, double* const new_B[3],
                                                const double* const old_B[3], const double* aux,
                                                grmhd_stream_t stream) {
    (void)t;   // the Cowling GRMHD substep does not depend on t
    if (!c) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    if (!(dt >= 0.0) || !std::isfinite(dt)) { set_err("invalid dt %g", dt); return GRMHD_ERR_INVALID; }
    k_set_scalar<<<1, 1, 0, (hipStream_t)stream>>>(c->dt_scalar, dt);
    return grmhd_advance_substep(c, c->dt_scalar, dtfact, new_state, old_state, new_B, old_B, aux, stream);
}

// Reads and clears the fatal-cell counter; the stream has completed.
static grmhd_status read_fatal(grmhd_ctx* c, hipStream_t s) {
    unsigned long long n = 0;
    HIP_TRY(hipMemcpyAsync(&n, c->fatal_d, sizeof(n), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (n == 0) return GRMHD_OK;
    HIP_TRY(hipMemsetAsync(c->fatal_d, 0, sizeof(n), s));
    HIP_TRY(hipStreamSynchronize(s));
    set_err("%llu cell(s) took a c2p branch where the reference aborts (rho, eps, T or p above the EOS "
            "range outside the lenient region) since the last sync point", n);
    return GRMHD_ERR_C2P_FATAL;
}

extern "C" grmhd_status grmhd_sync(grmhd_ctx* c, grmhd_stream_t stream) {
    if (!c) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return read_fatal(c, (hipStream_t)stream);
}

extern "C" grmhd_status grmhd_fill_ghosts(grmhd_ctx* c, double* state, double* const B[3], grmhd_stream_t stream) {
    if (!c || !state || !B) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    GatherArgs a{};
    a.st = state; a.nv = NSTATE; a.v0 = 0; a.vc = NHRSC; a.cell_mode = 1;
    for (int d = 0; d < 3; ++d) a.B[d] = B[d];
    launch_gather(c, a, (hipStream_t)stream);
    return launch_check("fill_ghosts");
}

extern "C" grmhd_status grmhd_fill_metric_ghosts(grmhd_ctx* c, double* state, grmhd_stream_t stream) {
    if (!c || !state) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    hipStream_t s = (hipStream_t)stream;
    {   // same-level copies and periodic images; nothing beyond outflow faces
        GatherArgs a{};
        a.st = state; a.nv = NSTATE; a.v0 = GXX_; a.vc = 16; a.cell_mode = 3;
        launch_gather(c, a, s);
    }
    if (c->forest && c->n_amr_cells > 0) {
        const int64_t ncc = (int64_t)ff_ntot(c->ff, 0, 0) * ff_ntot(c->ff, 0, 1) * ff_ntot(c->ff, 0, 2);
        if (16 * ncc > (int64_t)1 << 29) { set_err("fill_metric_ghosts: coarse domain too large"); return GRMHD_ERR_UNSUPPORTED; }
        if (!c->vcm) {
            hipError_t e = hipMalloc(&c->vcm, sizeof(double) * 16 * ncc);
            if (e != hipSuccess) { c->vcm = nullptr; set_err("vcm: %s", hipGetErrorString(e)); return GRMHD_ERR_ALLOC; }
        }
        const LagW W = lagrange_weights();
        Scope sc(c, "ghost_metric", s);
        k_vc_metric<<<nblk(ncc), 256, 0, s>>>(c->ff, W, state, NSTATE, GXX_, 16, c->vcm, ncc);
        k_amr_metric<<<nblk(c->n_amr_cells), 256, 0, s>>>(c->amr_cells, (int)c->n_amr_cells, c->ff, W, state, NSTATE,
                                                         GXX_, 16, c->vcm, ncc);
    }
    return launch_check("fill_metric_ghosts");
}

extern "C" grmhd_status grmhd_compute_auxiliaries(grmhd_ctx* c, double* state, const double* const B[3], double* aux,
                                                  grmhd_stream_t stream) {
    if (!c || !state || !B || !aux) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    launch_c2p_cells(c, state, NSTATE, state, NSTATE, B, aux, true, false, (hipStream_t)stream);
    return launch_check("compute_auxiliaries");
}

// Batch c2p + statistics reduction with the per-workgroup records in `part`
// (nblk(n) records; every workgroup writes its own, nothing needs clearing).
static void launch_c2p_batch(grmhd_ctx* c, int64_t n, double* cons_soa, const double* metric_soa, double* prims_soa,
                             uint64_t* err_bits, uint8_t* floored, int32_t dry_run, const double* r,
                             const double* guess, double* roots, DevStats* part, hipStream_t s) {
    if (n > 0) {
        MlpDev none{};
        if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS)
            k_c2p_batch<EosIdeal, false><<<nblk(n), 256, 0, s>>>(n, c->ideal, c->O, cons_soa, metric_soa, prims_soa,
                                                                err_bits, floored, dry_run, r, guess, roots, part,
                                                                none);
        else
            k_c2p_batch<EosHybrid, false><<<nblk(n), 256, 0, s>>>(n, c->hybrid, c->O, cons_soa, metric_soa, prims_soa,
                                                                 err_bits, floored, dry_run, r, guess, roots, part,
                                                                 none);
        k_stats_final<<<1, 256, 0, s>>>(part, nblk(n), c->stats);
    }
    c->last_c2p_n = n;
}

extern "C" grmhd_status grmhd_c2p_batch(grmhd_ctx* c, int64_t n, double* cons_soa, const double* metric_soa,
                                        double* prims_soa, uint64_t* err_bits, uint8_t* floored, int32_t dry_run,
                                        const double* r, const double* guess, double* roots, grmhd_stream_t stream) {
    if (!c || n < 0 || (n > 0 && (!cons_soa || !metric_soa || !prims_soa || !err_bits || !floored))) {
        set_err("invalid c2p batch arguments"); return GRMHD_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    if (n > 0 && nblk(n) > c->n_part_cap) {
        (void)hipFree(c->stats_part);
        c->stats_part = nullptr;
        c->n_part_cap = 0;
        HIP_TRY(hipMalloc(&c->stats_part, sizeof(DevStats) * nblk(n)));
        c->n_part_cap = nblk(n);
    }
    Scope sc(c, "c2p", s);
    launch_c2p_batch(c, n, cons_soa, metric_soa, prims_soa, err_bits, floored, dry_run, r, guess, roots,
                     c->stats_part, s);
    return launch_check("c2p_batch");
}

extern "C" grmhd_status grmhd_c2p_batch_noalloc(grmhd_ctx* c, int64_t n, double* cons_soa, const double* metric_soa,
                                                double* prims_soa, uint64_t* err_bits, uint8_t* floored,
                                                int32_t dry_run, const double* r, const double* guess, double* roots,
                                                grmhd_stream_t stream) {
    if (!c || n < 0 || (n > 0 && (!cons_soa || !metric_soa || !prims_soa || !err_bits || !floored))) {
        set_err("invalid c2p batch arguments"); return GRMHD_ERR_INVALID;
    }
    if (n > GRMHD_NOALLOC_MAX_STATES) {
        set_err("c2p_batch_noalloc: n = %lld exceeds the %d states its static scratch holds", (long long)n,
                GRMHD_NOALLOC_MAX_STATES);
        return GRMHD_ERR_INVALID;
    }
    if (!c->noalloc_part) { set_err("c2p_batch_noalloc: no scratch address for this context"); return GRMHD_ERR_HIP; }
    // no Scope: kernel timing creates events
    launch_c2p_batch(c, n, cons_soa, metric_soa, prims_soa, err_bits, floored, dry_run, r, guess, roots,
                     c->noalloc_part, (hipStream_t)stream);
    return launch_check("c2p_batch_noalloc");
}

static grmhd_status ensure_stats_part(grmhd_ctx* c, int64_t n) {
    if (n > 0 && nblk(n) > c->n_part_cap) {
        (void)hipFree(c->stats_part);
        c->stats_part = nullptr;
        c->n_part_cap = 0;
        HIP_TRY(hipMalloc(&c->stats_part, sizeof(DevStats) * nblk(n)));
        c->n_part_cap = nblk(n);
    }
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_c2p_batch_learned(grmhd_ctx* c, grmhd_mlp* mlp, int64_t n, double* cons_soa,
                                                const double* metric_soa, double* prims_soa, uint64_t* err_bits,
                                                uint8_t* floored, const double* r, double* roots,
                                                grmhd_stream_t stream) {
    if (!c || !mlp || n < 0 || (n > 0 && (!cons_soa || !metric_soa || !prims_soa || !err_bits || !floored))) {
        set_err("invalid learned c2p batch arguments"); return GRMHD_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    grmhd_status st = ensure_stats_part(c, n);
    if (st != GRMHD_OK) return st;
    Scope sc(c, "c2p", s);
    if (n > 0) {
        if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS && mlp->newton > 0)
            k_c2p_batch<EosIdeal, true, true><<<nblk(n), 256, 0, s>>>(n, c->ideal, c->O, cons_soa, metric_soa,
                                                                     prims_soa, err_bits, floored, 0, r, nullptr,
                                                                     roots, c->stats_part, mlp->dev());
        else if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS)
            k_c2p_batch<EosIdeal, true><<<nblk(n), 256, 0, s>>>(n, c->ideal, c->O, cons_soa, metric_soa, prims_soa,
                                                               err_bits, floored, 0, r, nullptr, roots,
                                                               c->stats_part, mlp->dev());
        else
            k_c2p_batch<EosHybrid, true><<<nblk(n), 256, 0, s>>>(n, c->hybrid, c->O, cons_soa, metric_soa,
                                                                prims_soa, err_bits, floored, 0, r, nullptr, roots,
                                                                c->stats_part, mlp->dev());
        k_stats_final<<<1, 256, 0, s>>>(c->stats_part, nblk(n), c->stats);
    }
    c->last_c2p_n = n;
    return launch_check("c2p_batch_learned");
}

static void drop_graph(grmhd_ctx* c);

// c2p inputs (densitised conserved with cell-centred face-average B, and the
// metric) of n interior cells spread evenly over the block list, as SoA in
// the layout of grmhd_c2p_batch: sample s is interior cell s*stride.
__global__ __launch_bounds__(256) void k_c2p_inputs(DevGrid G, const double* __restrict__ st, int snv,
                                                   const double* __restrict__ Bx, const double* __restrict__ By,
                                                   const double* __restrict__ Bz, int flat, int64_t n,
                                                   int64_t stride, double* __restrict__ cons_soa,
                                                   double* __restrict__ met) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t per = (int64_t)G.n[0] * G.n[1] * G.n[2];
    const int64_t gi = s * stride, q = gi / per, e = gi - q * per;
    const int k = G.ng + (int)(e / ((int64_t)G.n[0] * G.n[1]));
    const int j = G.ng + (int)((e / G.n[0]) % G.n[1]);
    const int i = G.ng + (int)(e % G.n[0]);
    const int64_t nc = G.ncell, c = cidx(G, i, j, k);
    const double* S = st + q * (int64_t)snv * nc + c;
    for (int v = 0; v < NHRSC; ++v) cons_soa[v * n + s] = S[v * nc];
    const double* bx = Bx + q * G.nface[0]; const double* by = By + q * G.nface[1]; const double* bz = Bz + q * G.nface[2];
    const int64_t fx = fidx<0>(G, i, j, k), fy = fidx<1>(G, i, j, k), fz = fidx<2>(G, i, j, k);
    cons_soa[BSXL * n + s] = 0.5 * (bx[fx] + bx[fx + 1]);
    cons_soa[BSYL * n + s] = 0.5 * (by[fy] + by[fy + G.N[0]]);
    cons_soa[BSZL * n + s] = 0.5 * (bz[fz] + bz[fz + (int64_t)G.N[0] * G.N[1]]);
    const int comps[10] = {GXX_, GXY_, GXZ_, GYY_, GYZ_, GZZ_, ALP_, BETAX_, BETAY_, BETAZ_};
    const double flatv[10] = {1, 0, 0, 1, 0, 1, 1, 0, 0, 0};
    for (int a = 0; a < 10; ++a) met[a * n + s] = flat ? flatv[a] : S[comps[a] * nc];
}

extern "C" grmhd_status grmhd_mlp_fit_cells(grmhd_ctx* c, grmhd_mlp* mlp, const double* state,
                                            const double* const B[3], int64_t max_states,
                                            const grmhd_mlp_fit_opts* opts, grmhd_mlp_fit_report* report,
                                            grmhd_stream_t stream) {
    if (!c || !mlp || !state || !B || !B[0] || !B[1] || !B[2] || max_states < 16) {
        set_err("invalid mlp_fit_cells arguments"); return GRMHD_ERR_INVALID;
    }
    const DevGrid& G = c->G;
    const int64_t total = (int64_t)G.n[0] * G.n[1] * G.n[2] * G.nq;
    const int64_t stride = std::max<int64_t>(1, (total + max_states - 1) / max_states);
    const int64_t n = (total + stride - 1) / stride;
    hipStream_t s = (hipStream_t)stream;
    double* buf = nullptr;
    hipError_t e = hipMalloc(&buf, sizeof(double) * 20 * n);
    if (e != hipSuccess) { set_err("mlp_fit_cells: %s", hipGetErrorString(e)); return GRMHD_ERR_ALLOC; }
    k_c2p_inputs<<<nblk(n), 256, 0, s>>>(G, state, NSTATE, B[0], B[1], B[2], c->P.flat_metric, n, stride, buf,
                                          buf + 10 * n);
    grmhd_status st = launch_check("c2p_inputs");
    if (st == GRMHD_OK) st = grmhd_mlp_fit(c, mlp, n, buf, buf + 10 * n, nullptr, opts, report, stream);
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    return st;
}

extern "C" grmhd_status grmhd_set_static_metric(grmhd_ctx* c, const double* state, int32_t enable,
                                                grmhd_stream_t stream) {
    if (!c || (enable && !state)) { set_err("invalid set_static_metric arguments"); return GRMHD_ERR_INVALID; }
    drop_graph(c);
    if (!enable) { c->mder_src = nullptr; return GRMHD_OK; }
    if (c->P.flat_metric) { c->mder_src = nullptr; return GRMHD_OK; }   // flat kernels take no derivatives
    const DevGrid& G = c->G;
    const int64_t n = (int64_t)G.n[0] * G.n[1] * G.n[2] * G.nq;
    if (!c->mder) {
        hipError_t e = hipMalloc(&c->mder, sizeof(double) * 30 * n);
        if (e != hipSuccess) { c->mder = nullptr; set_err("static metric cache: %s", hipGetErrorString(e)); return GRMHD_ERR_ALLOC; }
    }
    k_metric_der<<<nblk(n), 256, 0, (hipStream_t)stream>>>(G, c->inst_d, c->blk_d, state, NSTATE, c->P.fd_order,
                                                            c->mder);
    for (int d = 1; d < 3; ++d) {
        if (!c->fmet[d]) {
            hipError_t e = hipMalloc(&c->fmet[d], sizeof(double) * FMC * G.nface[d] * G.nq);
            if (e != hipSuccess) { c->fmet[d] = nullptr; set_err("face metric cache: %s", hipGetErrorString(e)); return GRMHD_ERR_ALLOC; }
        }
        const int wpb = (int)nblk(G.nface[d]);
        const dim3 grid((unsigned)(wpb * G.nq));
        hipStream_t s = (hipStream_t)stream;
        if (d == 1) k_face_metric<1><<<grid, 256, 0, s>>>(G, state, NSTATE, c->O.ex.alp_ex, c->fmet[1], wpb);
        else k_face_metric<2><<<grid, 256, 0, s>>>(G, state, NSTATE, c->O.ex.alp_ex, c->fmet[2], wpb);
    }
    c->mder_src = state;
    return launch_check("metric_der");
}

extern "C" grmhd_status grmhd_set_c2p_eval_counts(grmhd_ctx* c, uint8_t* counts_or_null) {
    if (!c) { set_err("null context"); return GRMHD_ERR_INVALID; }
    c->O.eval_counts = counts_or_null;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_set_c2p_warm_start(grmhd_ctx* c, int32_t max_newton_evals) {
    if (!c) { set_err("null context"); return GRMHD_ERR_INVALID; }
    if (max_newton_evals < 0 || max_newton_evals > 64) { set_err("max_newton_evals %d out of [0, 64]", max_newton_evals); return GRMHD_ERR_INVALID; }
    if (max_newton_evals > 0 && c->P.eos_kind != GRMHD_EOS_IDEAL_GAS) {
        set_err("the warm-started Newton finish is built for the ideal-gas EOS"); return GRMHD_ERR_UNSUPPORTED;
    }
    if (c->warm_newton != max_newton_evals) { c->warm_newton = max_newton_evals; drop_graph(c); }
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_set_learned_c2p(grmhd_ctx* c, grmhd_mlp* mlp) {
    if (!c) { set_err("null context"); return GRMHD_ERR_INVALID; }
    if (c->mlp != mlp) { c->mlp = mlp; drop_graph(c); }
    return GRMHD_OK;
}

template <typename EOS>
__global__ __launch_bounds__(256) void k_mlp_features(int64_t n, EOS eos, const double* __restrict__ cons_soa,
                                                      const double* __restrict__ met, float* __restrict__ feat) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    Met m;
    for (int a = 0; a < 6; ++a) m.g[a] = met[a * n + s];
    m.alp = met[6 * n + s];
    for (int a = 0; a < 3; ++a) m.beta[a] = met[(7 + a) * n + s];
    m.init();
    double cons[NCONS];
    for (int v = 0; v < NCONS; ++v) cons[v] = cons_soa[v * n + s];
    float f[MLP_NF];
    double lxhi;
    mlp_features(eos, m, cons, f, lxhi);
    for (int k = 0; k < MLP_NF; ++k) feat[k * n + s] = f[k];
}

extern "C" grmhd_status grmhd_mlp_features(grmhd_ctx* c, int64_t n, const double* cons_soa, const double* metric_soa,
                                           float* feat_soa, grmhd_stream_t stream) {
    if (!c || n < 0 || (n > 0 && (!cons_soa || !metric_soa || !feat_soa))) {
        set_err("invalid feature arguments"); return GRMHD_ERR_INVALID;
    }
    if (n == 0) return GRMHD_OK;
    hipStream_t s = (hipStream_t)stream;
    if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS)
        k_mlp_features<EosIdeal><<<nblk(n), 256, 0, s>>>(n, c->ideal, cons_soa, metric_soa, feat_soa);
    else
        k_mlp_features<EosHybrid><<<nblk(n), 256, 0, s>>>(n, c->hybrid, cons_soa, metric_soa, feat_soa);
    return launch_check("mlp_features");
}

extern "C" grmhd_status grmhd_prims_to_conservs_batch(grmhd_ctx* c, int64_t n, const double* prims_soa,
                                                      const double* metric_soa, double* cons_soa, grmhd_stream_t stream) {
    if (!c || n < 0 || (n > 0 && (!prims_soa || !metric_soa || !cons_soa))) { set_err("invalid p2c arguments"); return GRMHD_ERR_INVALID; }
    if (n > 0) k_p2c_batch<<<nblk(n), 256, 0, (hipStream_t)stream>>>(n, prims_soa, metric_soa, cons_soa);
    return launch_check("p2c_batch");
}

extern "C" grmhd_status grmhd_c2p_stats(grmhd_ctx* c, grmhd_c2p_stats_t* out, grmhd_stream_t stream) {
    if (!c || !out) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    DevStats h;
    HIP_TRY(hipMemcpyAsync(&h, c->stats, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    out->n_cells = c->last_c2p_n;
    out->n_fatal = (int64_t)h.n_fatal;
    out->n_atmo_reset = (int64_t)h.n_atmo;
    out->n_ent_backup = (int64_t)h.n_ent;
    out->n_guess_used = (int64_t)h.n_guess;
    out->n_root_evals = (int64_t)h.n_root;
    out->n_bracket_evals = (int64_t)h.n_bracket;
    out->max_root_evals = (int64_t)h.max_root;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_compute_dt(grmhd_ctx* c, const double* state, const double* aux, double cfl,
                                         double* dt_dev, grmhd_stream_t stream) {
    if (!c || !state || !aux || !dt_dev || !(cfl > 0)) { set_err("invalid dt arguments"); return GRMHD_ERR_INVALID; }
    hipStream_t s = (hipStream_t)stream;
    const DevGrid& G = c->G;
    int64_t n = (int64_t)G.n[0] * G.n[1] * G.n[2] * G.nq;
    bool flat = c->P.flat_metric != 0;
#define DTL(EOSV, FL) k_dt_partial<decltype(EOSV), FL><<<nblk(n), 256, 0, s>>>(G, EOSV, c->inst_d, c->blk_d, state, NSTATE, aux, cfl, c->dtpart)
    if (c->P.eos_kind == GRMHD_EOS_IDEAL_GAS) { if (flat) DTL(c->ideal, true); else DTL(c->ideal, false); }
    else { if (flat) DTL(c->hybrid, true); else DTL(c->hybrid, false); }
#undef DTL
    k_dt_final<<<1, 256, 0, s>>>(c->dtpart, nblk(n), dt_dev);
    return launch_check("compute_dt");
}

extern "C" grmhd_status grmhd_find_stable_timestep(grmhd_ctx* c, const double* state, const double* aux, double cfl,
                                                   double* dt_out, grmhd_stream_t stream) {
    if (!c || !dt_out) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    hipStream_t s = (hipStream_t)stream;
    grmhd_status r = grmhd_compute_dt(c, state, aux, cfl, c->dt_scalar, stream);
    if (r != GRMHD_OK) return r;
    double dt = 0.0;
    HIP_TRY(hipMemcpyAsync(&dt, c->dt_scalar, sizeof(dt), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *dt_out = dt;
    return read_fatal(c, s);
}

static grmhd_status ensure_stage_buffers(grmhd_ctx* c) {
    if (c->U1) return GRMHD_OK;
    const DevGrid& G = c->G;
    for (auto& t : c->side)
        if (!t) HIP_TRY(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    for (auto& e : c->fev)
        if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipMalloc(&c->U1, sizeof(double) * NHRSC * G.ncell * G.nq));
    HIP_TRY(hipMalloc(&c->U2, sizeof(double) * NHRSC * G.ncell * G.nq));
    for (int d = 0; d < 3; ++d) {
        HIP_TRY(hipMalloc(&c->B1[d], sizeof(double) * G.nface[d] * G.nq));
        HIP_TRY(hipMalloc(&c->B2[d], sizeof(double) * G.nface[d] * G.nq));
    }
    return GRMHD_OK;
}

// The fused RK3 step.  Stage buffers hold only the 7 hydro conserved; the
// static metric is read from `state` throughout.
static void rk3_launches(grmhd_ctx* c, const double* dt_dev, double* state, double* const B[3], double* aux,
                         hipStream_t s) {
    const DevGrid& G = c->G;
    const double* const Bc[3] = {B[0], B[1], B[2]};
    const double* const B1c[3] = {c->B1[0], c->B1[1], c->B1[2]};
    const double* const B2c[3] = {c->B2[0], c->B2[1], c->B2[2]};
    struct Stage { double* nst; int nnv; const double* ast; int anv; const double* bst; int bnv; double ca, cb, dtf;
                   double* const* nB; const double* const* aB; const double* const* bB;
                   const double* ost; int onv; const double* const* oB; };
    // stage 1: U1 = U0 + dt L(U0)
    // stage 2: U2 = 3/4 U0 + 1/4 U1 + 1/4 dt L(U1)
    // stage 3: U0 = 1/3 U0 + 2/3 U2 + 2/3 dt L(U2)
    double* const nB1[3] = {c->B1[0], c->B1[1], c->B1[2]};
    double* const nB2[3] = {c->B2[0], c->B2[1], c->B2[2]};
    Stage st[3] = {
        {c->U1, NHRSC, state, NSTATE, nullptr, 0, 1.0, 0.0, 1.0, nB1, Bc, nullptr, state, NSTATE, Bc},
        {c->U2, NHRSC, state, NSTATE, c->U1, NHRSC, 0.75, 0.25, 0.25, nB2, Bc, B1c, c->U1, NHRSC, B1c},
        {state, NSTATE, state, NSTATE, c->U2, NHRSC, 1. / 3, 2. / 3., 2. / 3., B, Bc, B2c, c->U2, NHRSC, B2c},
    };
    // Without FOFC, independent launches of a stage run side by side: the
    // three flux directions; then the ghost-face passes (copies and, on a
    // forest, the coarse-fine face passes) and the c2p of the ghosts that run
    // their own (outflow and coarse-fine ghosts) next to the c2p of the
    // interior cells.  The conserved fills those ghost c2p read (copies and
    // coarse-fine prolongation/restriction from interior cells) stay on s
    // ahead of the interior c2p, which may rewrite interior conserved.
    // Ordering by events, so the same sequence is captured into the step's
    // graph.  (grmhd_params.stage_streams = 0 or per-family timing: every
    // launch on s.)
    const bool conc = c->P.stage_streams && !c->timing && !c->P.fofc && c->side[0] && c->side[1];
    auto fork = [&](int k, hipStream_t from, hipStream_t to) {
        (void)hipEventRecord(c->fev[k], from);
        (void)hipStreamWaitEvent(to, c->fev[k], 0);
    };
    // the c2p error fields are cleared once per step; only the c2p reads
    // them, so with side streams the clear runs beside the first fluxes
    if (!conc) k_reset_diag<<<nblk(G.ncell * G.nq), 256, 0, s>>>(G.ncell, G.nq, aux);
    for (int sidx = 0; sidx < 3; ++sidx) {
        Stage& S = st[sidx];
        if (conc) {
            fork(0, s, c->side[0]);
            fork(0, s, c->side[1]);
            if (sidx == 0) k_reset_diag<<<nblk(G.ncell * G.nq), 256, 0, c->side[1]>>>(G.ncell, G.nq, aux);
            c->flux_s[1] = c->side[0]; c->flux_s[2] = c->side[1];
            c->ring_copy_now = sidx > 0;
            launch_fluxes(c, aux, state, NSTATE, S.oB, s);
            c->ring_copy_now = false;
            c->flux_s[1] = c->flux_s[2] = nullptr;
            fork(1, c->side[0], s);
            fork(2, c->side[1], s);
        } else {
            c->ring_copy_now = sidx > 0;
            launch_fluxes(c, aux, state, NSTATE, S.oB, s);
            c->ring_copy_now = false;
        }
        // the cell part of the update needs only the fluxes (on a forest
        // after their reflux): it runs on side[0] beside the edge EMFs, the
        // face (CT) part after them (and their reflux) on s.  Not with GS
        // EMFs on a forest: those read the mass fluxes the reflux rewrites.
        const bool split_update = conc && !(c->forest && c->P.emf == GRMHD_EMF_GS);
        if (split_update) {
            fork(6, s, c->side[0]);
            launch_reflux(c, c->side[0], 1);
        }
        launch_emf(c, S.oB, s);
        if (c->P.fofc)
            launch_fofc(c, S.ast, S.anv, S.bst, S.bnv, S.ca, S.cb, state, S.oB, aux, dt_dev, S.dtf, s);
        launch_reflux(c, s, split_update ? 2 : 0);
        UpdArgs a{};
        a.nst = S.nst; a.nnv = S.nnv; a.ast = S.ast; a.anv = S.anv; a.bst = S.bst; a.bnv = S.bnv;
        a.ca = S.ca; a.cb = S.cb; a.mstate = state; a.mnv = NSTATE; a.aux = aux;
        for (int d = 0; d < 3; ++d) { a.nB[d] = S.nB[d]; a.aB[d] = S.aB[d]; a.bB[d] = S.bB ? S.bB[d] : nullptr; }
        a.dt = dt_dev; a.dtfact = S.dtf;
        if (split_update) {
            UpdArgs ac = a;
            ac.part = 1;
            launch_update(c, ac, c->side[0]);
            a.part = 2;
            launch_update(c, a, s);
            fork(7, c->side[0], s);
        } else {
            launch_update(c, a, s);
        }
        const double* const nBc[3] = {S.nB[0], S.nB[1], S.nB[2]};
        // ghost faces: every one after the last stage; between stages with
        // shared ring faces the next edge EMFs and the ring faces beyond
        // outflow faces read only the face slabs, the ghost cells beyond
        // outflow faces their own faces
        const bool faces_slab = sidx < 2 && c->ring_own != nullptr;
        if (conc) {
            // side[0]: ghost faces, then (after the conserved fills on s) the
            // c2p of the ghosts that run their own; s: conserved fills,
            // interior c2p.
            fork(3, s, c->side[0]);
            GatherArgs gf{};
            for (int d = 0; d < 3; ++d) gf.B[d] = S.nB[d];
            gf.cell_mode = 0;
            gf.faces_slab = faces_slab;
            launch_gather(c, gf, c->side[0]);
            GatherArgs gc{};
            gc.st = S.nst; gc.nv = S.nnv; gc.v0 = 0; gc.vc = NHRSC; gc.cell_mode = 2;
            launch_gather(c, gc, s);
            fork(4, s, c->side[0]);
            launch_c2p_cells(c, S.nst, S.nnv, state, NSTATE, nBc, aux, false, true, s, c->side[0]);
            fork(5, c->side[0], s);
        } else {
            {   // ghost faces everywhere; conserved only where a ghost runs its own c2p
                // (everywhere with FOFC: its next-stage tests read the ghost layer)
                GatherArgs g{};
                g.st = S.nst; g.nv = S.nnv; g.v0 = 0; g.vc = NHRSC; g.cell_mode = c->P.fofc ? 1 : 2;
                for (int d = 0; d < 3; ++d) g.B[d] = S.nB[d];
                g.faces_slab = faces_slab;
                launch_gather(c, g, s);
            }
            launch_c2p_cells(c, S.nst, S.nnv, state, NSTATE, nBc, aux, false, true, s);
        }
        {   // c2p outputs into the other ghost cells: the reconstructed
            // primitives (and the GS cell E, which the next stage's edge EMFs
            // read one cell into the ghosts) between stages, everything after
            // the last stage; with FOFC also the (possibly rewritten)
            // conserved at every stage, which the next stage's flag test
            // reads in the ghost layer.  (Pushing these from the c2p kernel
            // instead was measured slower: scattered stores from a
            // 3-wave/SIMD FP64 kernel.)
            GatherArgs g{};
            const bool last = sidx == 2;
            uint32_t mask = 0x3FFu | (c->P.thermo_recon ? (1u << PRESS_) : 0u);
            g.aux = aux; g.aux_mask = last ? 0x7FFFu : mask;
            g.st = (last || c->P.fofc) ? S.nst : nullptr; g.nv = S.nnv; g.v0 = 0; g.vc = NHRSC;
            g.ecen = c->P.emf == GRMHD_EMF_GS ? c->ecen : nullptr;
            g.cell_mode = 3;
            g.own_copies = true;
            // the next stage copies the ring faces' values from the
            // neighbouring blocks and its full faces read the cells beyond a
            // block face from the neighbour's interior: of the ghost
            // primitives it reads only those beyond outflow faces, which
            // run their own c2p or copy a first-layer ghost
            g.copies_only = !last && c->ring_own != nullptr;
            launch_gather(c, g, s);
        }
    }
}

extern "C" grmhd_status grmhd_rk3_step(grmhd_ctx* c, const double* dt_dev, double* state, double* const B[3],
                                       double* aux, grmhd_stream_t stream) {
    if (!c || !dt_dev || !state || !B || !aux) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    grmhd_status r = ensure_stage_buffers(c);
    if (r != GRMHD_OK) return r;
    hipStream_t s = (hipStream_t)stream;
    {
        Scope sc(c, "step", s);
        rk3_launches(c, dt_dev, state, B, aux, s);
    }
    return launch_check("rk3_step");
}

static bool same_mlp(const MlpDev& a, const MlpDev& b) {
    return a.frag == b.frag && a.bias == b.bias && a.norm == b.norm && a.fe_lo == b.fe_lo && a.fe_hi == b.fe_hi &&
           a.fs_lo == b.fs_lo && a.fs_hi == b.fs_hi && a.nexp == b.nexp && a.newton == b.newton;
}

extern "C" grmhd_status grmhd_rk3_step_graph(grmhd_ctx* c, const double* dt_dev, double* state, double* const B[3],
                                             double* aux, grmhd_stream_t stream) {
    if (!c || !dt_dev || !state || !B || !aux) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    if (!stream) { set_err("graph capture needs a non-default stream"); return GRMHD_ERR_INVALID; }
    grmhd_status r = ensure_stage_buffers(c);
    if (r != GRMHD_OK) return r;
    hipStream_t s = (hipStream_t)stream;
    const void* key[8] = {dt_dev, state, B[0], B[1], B[2], aux, nullptr, nullptr};
    // the learned c2p's bracket/Newton settings are kernel arguments, so a
    // change of them (grmhd_mlp_set_bracket/set_newton, a new fit) needs a
    // new capture; its weights are read through pointers
    const MlpDev md = c->mlp ? c->mlp->dev() : MlpDev{};
    bool same = c->gexec && std::memcmp(key, c->gkey, sizeof(key)) == 0 && same_mlp(md, c->gmlp);
    if (!same) {
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        bool tim = c->timing;
        c->timing = false;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rk3_launches(c, dt_dev, state, B, aux, s);
        hipError_t e = hipStreamEndCapture(s, &c->graph);
        c->timing = tim;
        if (e != hipSuccess) { set_err("capture: %s", hipGetErrorString(e)); return GRMHD_ERR_HIP; }
        HIP_TRY(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        std::memcpy(c->gkey, key, sizeof(key));
        c->gmlp = md;
    }
    Scope sc(c, "step", s);
    HIP_TRY(hipGraphLaunch(c->gexec, s));
    return GRMHD_OK;
}

static bool tuning_valid(const grmhd_tuning& t) {
    auto ok = [](int v) { return v == 64 || v == 128 || v == 256; };
    auto b = [](int v) { return v == 0 || v == 1; };
    return ok(t.flux[0]) && ok(t.flux[1]) && ok(t.flux[2]) && ok(t.emf) && ok(t.update) && ok(t.gather_faces) &&
           ok(t.gather_cells) && ok(t.c2p) && b(t.flux_prefetch[0]) && b(t.flux_prefetch[1]) &&
           b(t.flux_prefetch[2]);
}
static void drop_graph(grmhd_ctx* c) {
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    std::memset(c->gkey, 0, sizeof(c->gkey));
}

extern "C" grmhd_status grmhd_get_tuning(grmhd_ctx* c, grmhd_tuning* out) {
    if (!c || !out) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    *out = c->tune;
    return GRMHD_OK;
}
extern "C" grmhd_status grmhd_set_tuning(grmhd_ctx* c, const grmhd_tuning* in) {
    if (!c || !in || !tuning_valid(*in)) { set_err("invalid tuning"); return GRMHD_ERR_INVALID; }
    c->tune = *in;
    drop_graph(c);
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_autotune(grmhd_ctx* c, const double* state, const double* const B[3], const double* aux,
                                       int32_t reps, grmhd_tuning* out, grmhd_stream_t stream) {
    if (!c || !state || !B || !aux || reps < 1 || reps > 100) { set_err("invalid autotune arguments"); return GRMHD_ERR_INVALID; }
    grmhd_status st = ensure_stage_buffers(c);
    if (st != GRMHD_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const DevGrid& G = c->G;
    // scratch: a zero dt and a copy of aux for the c2p runs
    double* dt0 = nullptr;
    double* aux2 = nullptr;
    HIP_TRY(hipMalloc(&dt0, sizeof(double)));
    hipError_t e = hipMalloc(&aux2, sizeof(double) * NAUX * G.ncell * G.nq);
    if (e != hipSuccess) { (void)hipFree(dt0); set_err("autotune scratch: %s", hipGetErrorString(e)); return GRMHD_ERR_ALLOC; }
    (void)hipMemsetAsync(dt0, 0, sizeof(double), s);
    (void)hipMemcpyAsync(aux2, aux, sizeof(double) * NAUX * G.ncell * G.nq, hipMemcpyDeviceToDevice, s);
    const bool was_timing = c->timing;
    c->timing = false;
    hipEvent_t a = timing_event(), b = timing_event();
    auto time_it = [&](auto&& fn) {
        fn();   // warm-up
        std::vector<float> t;
        for (int r = 0; r < reps; ++r) {
            (void)hipEventRecord(a, s);
            fn();
            (void)hipEventRecord(b, s);
            (void)hipEventSynchronize(b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    };
    const int cand[3] = {64, 128, 256};
    auto pick = [&](int* slot, auto&& fn) {
        float best = 1e30f;
        int bv = *slot;
        for (int v : cand) {
            *slot = v;
            float ms = time_it(fn);
            if (ms < best) { best = ms; bv = v; }
        }
        *slot = bv;
    };
    const double* const Bc[3] = {B[0], B[1], B[2]};
    double* const B1[3] = {c->B1[0], c->B1[1], c->B1[2]};
    const double* const B1c[3] = {c->B1[0], c->B1[1], c->B1[2]};
    grmhd_tuning keep = c->tune;
    // the three flux launches as the RK3 stages run them: on three streams
    // when the stages use side streams, and in the form of the second and
    // third stages (ring faces pushed, reads across block faces) when the
    // context shares ring faces; the directions run concurrently, so each
    // choice is timed against the other two
    const bool conc = c->P.stage_streams && !c->P.fofc && c->side[0] && c->side[1];
    auto fluxes_as_in_step = [&] {
        if (conc) {
            for (int k = 0; k < 2; ++k) {
                (void)hipEventRecord(c->fev[k], s);
                (void)hipStreamWaitEvent(c->side[k], c->fev[k], 0);
            }
            c->flux_s[1] = c->side[0]; c->flux_s[2] = c->side[1];
        }
        c->ring_copy_now = c->ring_own != nullptr;
        launch_fluxes(c, aux, state, NSTATE, Bc, s);
        c->ring_copy_now = false;
        if (conc) {
            c->flux_s[1] = c->flux_s[2] = nullptr;
            for (int k = 0; k < 2; ++k) {
                (void)hipEventRecord(c->fev[2 + k], c->side[k]);
                (void)hipStreamWaitEvent(s, c->fev[2 + k], 0);
            }
        }
    };
    for (int d = 0; d < 3; ++d) {
        // one direction at a time (the other two run at their current
        // choice): workgroup size and, for WENO-Z, the prefetching variant
        float best = 1e30f;
        int bs = c->tune.flux[d], bp = c->tune.flux_prefetch[d];
        for (int pf = 0; pf < (c->P.recon == GRMHD_RECON_WENOZ ? 2 : 1); ++pf)
            for (int v : cand) {
                c->tune.flux[d] = v;
                c->tune.flux_prefetch[d] = pf;
                const float ms = time_it(fluxes_as_in_step);
                if (ms < best) { best = ms; bs = v; bp = pf; }
            }
        c->tune.flux[d] = bs;
        c->tune.flux_prefetch[d] = bp;
    }
    // the edge EMFs and the update as the stages run them: the update's cell
    // part beside the EMFs on a side stream, its face part after them
    const bool split = conc && !(c->forest && c->P.emf == GRMHD_EMF_GS);
    auto emf_update_as_in_step = [&] {
        UpdArgs u{};
        u.nst = c->U1; u.nnv = NHRSC; u.ast = state; u.anv = NSTATE; u.bst = nullptr; u.bnv = 0;
        u.ca = 1.0; u.cb = 0.0; u.mstate = state; u.mnv = NSTATE; u.aux = aux;
        for (int d = 0; d < 3; ++d) { u.nB[d] = B1[d]; u.aB[d] = Bc[d]; u.bB[d] = nullptr; }
        u.dt = dt0; u.dtfact = 1.0;
        if (split) {
            (void)hipEventRecord(c->fev[0], s);
            (void)hipStreamWaitEvent(c->side[0], c->fev[0], 0);
            UpdArgs uc = u;
            uc.part = 1;
            launch_update(c, uc, c->side[0]);
            launch_emf(c, Bc, s);
            u.part = 2;
            launch_update(c, u, s);
            (void)hipEventRecord(c->fev[1], c->side[0]);
            (void)hipStreamWaitEvent(s, c->fev[1], 0);
        } else {
            launch_emf(c, Bc, s);
            launch_update(c, u, s);
        }
    };
    pick(&c->tune.emf, emf_update_as_in_step);
    pick(&c->tune.update, emf_update_as_in_step);
    pick(&c->tune.gather_faces, [&] {
        GatherArgs g{};
        for (int d = 0; d < 3; ++d) g.B[d] = B1[d];
        launch_gather(c, g, s);
    });
    pick(&c->tune.gather_cells, [&] {
        GatherArgs g{};
        g.aux = aux2; g.aux_mask = 0x3FFu; g.cell_mode = 3; g.own_copies = true;
        launch_gather(c, g, s);
    });
    {   // ghosts of the stage copy, so the c2p sees valid inputs everywhere
        GatherArgs g{};
        g.st = c->U1; g.nv = NHRSC; g.v0 = 0; g.vc = NHRSC; g.cell_mode = 1;
        for (int d = 0; d < 3; ++d) g.B[d] = B1[d];
        launch_gather(c, g, s);
    }
    pick(&c->tune.c2p, [&] { launch_c2p_cells(c, c->U1, NHRSC, state, NSTATE, B1c, aux2, false, true, s); });
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(dt0);
    (void)hipFree(aux2);
    c->timing = was_timing;
    if (!tuning_valid(c->tune)) c->tune = keep;
    drop_graph(c);
    if (out) *out = c->tune;
    return launch_check("autotune");
}

extern "C" grmhd_status grmhd_set_timing(grmhd_ctx* c, int32_t on) {
    if (!c) { set_err("null ctx"); return GRMHD_ERR_INVALID; }
    c->timing = on != 0;
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_kernel_time_ms(grmhd_ctx* c, const char* fam, double* ms_total, int64_t* n_launch) {
    if (!c || !fam || !ms_total || !n_launch) { set_err("null argument"); return GRMHD_ERR_INVALID; }
    double tot = 0; int64_t n = 0;
    std::vector<TimerRec> keep;
    for (auto& t : c->timers) {
        HIP_TRY(hipEventSynchronize(t.b));
        if (t.fam == fam) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
            tot += ms; n++;
            c->ev_pool.push_back(t.a); c->ev_pool.push_back(t.b);
        } else keep.push_back(t);
    }
    c->timers.swap(keep);
    *ms_total = tot; *n_launch = n;
    return GRMHD_OK;
}

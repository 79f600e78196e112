// forest.hh -- two-level block forest: global-position lookup, the coarse-fine
// part of the ghost exchange (prolongation, restriction) and the refluxing
// entries.  Host + device (the host builds the gather tables and the reflux
// lists with the same lookup the kernels use).
//
// Forest: nb[0] x nb[1] x nb[2] coarse blocks, some refined once into 8
// children (p4est Morton child id c = cx + 2 cy + 4 cz).  Block order: the
// coarse leaves in ascending coarse id (bi + nb0 (bj + nb1 bk)), then the 8
// children of each refined block, refined blocks in ascending coarse id.
//
// A ghost element of a block is valued from its global position at the
// block's level, wrapped across periodic faces and clamped across outflow
// faces (the same-level exchange + physical BC of the uniform case):
//   same-level owner          -> copy (done by the ordinary gather tables)
//   coarse ghost, fine owner  -> restriction of the fine data
//   fine ghost, coarse owner  -> prolongation of the coarse data
// Both read the "virtual coarse field": a coarse leaf's own cell / face, or
// the restriction of the fine block that covers the position -- the contents
// of the reference's coarse buffers (restrict_op of the fine interior,
// restrict_kernels.hh:43-107; copies from coarse and same-level neighbours,
// amr_ghosts.cpp:293-420; physical BC on the buffer).
//   hydro:  slope_limited_prolong_op<minmod> (pr_helpers.hh:71-103);
//           second_order_restrict_op, pair-symmetric (pr_helpers.hh:198-212)
//   face B: area average (restrict_kernels.hh:125-164); Toth-Roe
//           divergence-preserving prolongation (prolongation_kernels.hh:111-386)
// Everything reads interior data only, so no element waits for another.
#pragma once

#include <vector>

#include "phys.hh"

// Gather-table flags (block field of a ghost source entry).
constexpr int GSRC_OWN = 1 << 30;   // the ghost runs its own c2p
constexpr int GSRC_AMR = 1 << 29;   // valued by prolongation / restriction, not copied
constexpr int GSRC_COPY = 1 << 28;  // takes the c2p outputs of a first-layer outflow ghost of its own block

struct FForest {
    int nb[3];
    int bc[6];             // 1 periodic, 0 outflow, per face -x +x -y +y -z +z
    int n[3], ng, N[3];    // interior cells, ghosts, padded extents per block
    int64_t ncell;         // padded cells per block
    int64_t nface[3];      // faces per block per direction
    int64_t nq;            // blocks
    const int32_t* cmap;   // per coarse block: q >= 0 leaf, or -(1 + q of child 0)
    const int4* blk;       // per block: level, coordinates at that level
};

struct FOwner { int q, cover, loc[3]; };

GHD int ff_ntot(const FForest& F, int lev, int d) { return (F.nb[d] * F.n[d]) << lev; }
GHD int ff_map_cc(const FForest& F, int lev, int d, int g) {
    const int N = ff_ntot(F, lev, d);
    if (g < 0) return F.bc[2 * d] == 1 ? ((g % N) + N) % N : 0;
    if (g >= N) return F.bc[2 * d + 1] == 1 ? g % N : N - 1;
    return g;
}
GHD int ff_map_fc(const FForest& F, int lev, int d, int f) {
    const int N = ff_ntot(F, lev, d);
    if (f < 0) return F.bc[2 * d] == 1 ? ((f % N) + N) % N : 0;
    if (f > N) return F.bc[2 * d + 1] == 1 ? f % N : N;
    return f;
}
// Owner of the mapped cell g at level lev: the block, its level (cover), and
// the local index of the cell (cover == lev), of the coarse cell holding it
// (lev 1, cover 0), or of the lower-corner fine child (lev 0, cover 1).
GHD FOwner ff_owner(const FForest& F, int lev, const int g[3]) {
    int cb[3];
    for (int d = 0; d < 3; ++d) cb[d] = lev == 0 ? g[d] / F.n[d] : (g[d] / F.n[d]) >> 1;
    const int m = F.cmap[cb[0] + F.nb[0] * (cb[1] + F.nb[1] * cb[2])];
    FOwner o;
    if (m >= 0) {
        o.q = m; o.cover = 0;
        for (int d = 0; d < 3; ++d) o.loc[d] = (lev == 0 ? g[d] : g[d] >> 1) % F.n[d] + F.ng;
    } else {
        o.cover = 1;
        int ch = 0;
        for (int d = 0; d < 3; ++d) {
            const int gf = lev == 1 ? g[d] : 2 * g[d];
            ch |= ((gf / F.n[d]) & 1) << d;
            o.loc[d] = gf % F.n[d] + F.ng;
        }
        o.q = -m - 1 + ch;
    }
    return o;
}
GHD int64_t ff_cidx(const FForest& F, int i, int j, int k) { return i + (int64_t)F.N[0] * (j + (int64_t)F.N[1] * k); }
GHD int64_t ff_fidx(const FForest& F, int d, int i, int j, int k) {
    return i + (int64_t)(F.N[0] + (d == 0)) * (j + (int64_t)(F.N[1] + (d == 1)) * k);
}
// Global position at the block's level of local element loc (face direction
// fd, -1 for cells), mapped.
GHD void ff_global(const FForest& F, int q, int fd, const int loc[3], int lev_out[1], int g[3]) {
    const int4 b = F.blk[q];
    const int lev = b.x, cr[3] = {b.y, b.z, b.w};
    lev_out[0] = lev;
    for (int d = 0; d < 3; ++d) {
        const int gg = cr[d] * F.n[d] + loc[d] - F.ng;
        g[d] = d == fd ? ff_map_fc(F, lev, d, gg) : ff_map_cc(F, lev, d, gg);
    }
}

// ---- cell-centred -------------------------------------------------------------
// A virtual coarse cell: block, local offset, and whether it is the
// restriction of the 8 fine cells from that (lower-corner) offset.
struct FCRef { int q; int off; int r8; };
GHD FCRef ff_vc_ref(const FForest& F, const int X[3]) {
    int g[3];
    for (int d = 0; d < 3; ++d) g[d] = ff_map_cc(F, 0, d, X[d]);
    const FOwner o = ff_owner(F, 0, g);
    return FCRef{o.q, (int)ff_cidx(F, o.loc[0], o.loc[1], o.loc[2]), o.cover};
}
// second_order_restrict_op: pair-symmetric sum of the 8 children.
GHD double ff_restrict8(const FForest& F, const double* p) {
    const int64_t sj = F.N[0], sk = (int64_t)F.N[0] * F.N[1];
    const double q00 = p[0] + p[1];
    const double q10 = p[sj] + p[sj + 1];
    const double q01 = p[sk] + p[sk + 1];
    const double q11 = p[sj + sk] + p[sj + sk + 1];
    const double r0 = q00 + q10, r1 = q01 + q11;
    return 0.125 * (r0 + r1);
}
GHD double ff_vc_val(const FForest& F, const double* st, int nv, int v, const FCRef& r) {
    const double* p = st + ((int64_t)r.q * nv + v) * F.ncell + r.off;
    return r.r8 ? ff_restrict8(F, p) : p[0];
}

// ---- face-staggered ------------------------------------------------------------
// Block and local index holding face (d; f along d; tangential cells t) of
// level lev (mapped), preferring the cell above the face; false when neither
// adjacent cell is covered at that level.
GHD bool ff_face_src(const FForest& F, int lev, int d, int f, const int t[3], int& q, int loc[3]) {
    const int N = ff_ntot(F, lev, d);
    int g[3] = {t[0], t[1], t[2]};
    for (int side = 1; side >= 0; --side) {
        int c = side ? f : f - 1;
        if (c < 0 || c >= N) {
            if (F.bc[2 * d + (c < 0 ? 0 : 1)] != 1) continue;
            c = ((c % N) + N) % N;
        }
        g[d] = c;
        const FOwner o = ff_owner(F, lev, g);
        if (o.cover != lev) continue;
        q = o.q;
        for (int e = 0; e < 3; ++e) loc[e] = o.loc[e];
        loc[d] += side ? 0 : 1;
        return true;
    }
    return false;
}
// Virtual coarse face field (d; F0 along d; tangential T), unmapped.
GHD double ff_vcf(const FForest& F, const double* Bd, int d, int F0, const int T[3]) {
    int t[3];
    for (int e = 0; e < 3; ++e) t[e] = e == d ? 0 : ff_map_cc(F, 0, e, T[e]);
    const int f = ff_map_fc(F, 0, d, F0);
    int q, loc[3];
    if (ff_face_src(F, 0, d, f, t, q, loc)) return Bd[(int64_t)q * F.nface[d] + ff_fidx(F, d, loc[0], loc[1], loc[2])];
    // both sides refined: some fine cell beside fine face 2f is covered
    int tf[3];
    for (int e = 0; e < 3; ++e) tf[e] = 2 * t[e];
    if (!ff_face_src(F, 1, d, 2 * f, tf, q, loc)) return 0.0;
    const double* p = Bd + (int64_t)q * F.nface[d];
    double val = 0;
    for (int ii = 0; ii <= (d != 0); ++ii)
        for (int jj = 0; jj <= (d != 1); ++jj)
            for (int kk = 0; kk <= (d != 2); ++kk) val += p[ff_fidx(F, d, loc[0] + ii, loc[1] + jj, loc[2] + kk)];
    return 0.25 * val;
}
// Fine face lying on a coarse face (even index f along d): the existing fine
// value, or the coarse face with its limited tangential slopes
// (fill_inside_face, prolongation_kernels.hh:130-160).
GHD double ff_fine_even(const FForest& F, const double* const B[3], int d, int f, const int t[3]) {
    int tm[3];
    for (int e = 0; e < 3; ++e) tm[e] = e == d ? 0 : ff_map_cc(F, 1, e, t[e]);
    const int fm = ff_map_fc(F, 1, d, f);
    int q, loc[3];
    if (ff_face_src(F, 1, d, fm, tm, q, loc)) return B[d][(int64_t)q * F.nface[d] + ff_fidx(F, d, loc[0], loc[1], loc[2])];
    int T[3];
    for (int e = 0; e < 3; ++e) T[e] = tm[e] >> 1;
    const int Fc = fm >> 1;
    const double U = ff_vcf(F, B[d], d, Fc, T);
    double r = U;
    for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        int Tm[3] = {T[0], T[1], T[2]}, Tp[3] = {T[0], T[1], T[2]};
        Tm[e] -= 1; Tp[e] += 1;
        const double s = 0.25 * gm::minmod(U - ff_vcf(F, B[d], d, Fc, Tm), ff_vcf(F, B[d], d, Fc, Tp) - U);
        r = r + (tm[e] & 1 ? 1 : -1) * s;
    }
    return r;
}
// Interior fine face of coarse cell C (Toth-Roe, prolongation_kernels.hh:330-381):
// direction d, fine index 2 C[d] + 1, offsets o0, o1 in {0,1} along the two
// other axes (lower axis first).  The 24 fine faces on the cell's six coarse
// faces are gathered once.
// Toth-Roe interior face from the 24 fine faces on the coarse cell's faces:
// u/v/w[normal side][first tangential][second tangential].
GHD double ff_toth_roe(const double u[2][2][2], const double v[2][2][2], const double w[2][2][2], int d, int o0, int o1) {
    // u(i_f+2ii, j_f+jj, k_f+kk) = u[ii][jj][kk]; v(i_f+ii, j_f+2jj, k_f+kk) = v[jj][ii][kk];
    // w(i_f+ii, j_f+jj, k_f+2kk) = w[kk][ii][jj]
    double Uxx = 0, Vyy = 0, Wzz = 0, Uxyz = 0, Vxyz = 0, Wxyz = 0;
    for (int ii = 0; ii <= 1; ++ii)
        for (int jj = 0; jj <= 1; ++jj)
            for (int kk = 0; kk <= 1; ++kk) {
                const int is = ii ? 1 : -1, js = jj ? 1 : -1, ks = kk ? 1 : -1;
                Uxx += is * js * v[jj][ii][kk] + is * ks * w[kk][ii][jj];
                Vyy += is * js * u[ii][jj][kk] + js * ks * w[kk][ii][jj];
                Wzz += is * ks * u[ii][jj][kk] + js * ks * v[jj][ii][kk];
                Uxyz += is * js * ks * u[ii][jj][kk];
                Vxyz += is * js * ks * v[jj][ii][kk];
                Wxyz += is * js * ks * w[kk][ii][jj];
            }
    Uxx *= 0.125; Vyy *= 0.125; Wzz *= 0.125;
    Uxyz *= 0.0625; Vxyz *= 0.0625; Wxyz *= 0.0625;
    const int s0 = o0 ? 1 : -1, s1 = o1 ? 1 : -1;
    if (d == 0) return 0.5 * (u[0][o0][o1] + u[1][o0][o1]) + Uxx + s1 * Vxyz + s0 * Wxyz;
    if (d == 1) return 0.5 * (v[0][o0][o1] + v[1][o0][o1]) + Vyy + s0 * Wxyz + s1 * Uxyz;
    return 0.5 * (w[0][o0][o1] + w[1][o0][o1]) + Wzz + s1 * Uxyz + s0 * Vxyz;
}
GHD double ff_fine_mid(const FForest& F, const double* const B[3], int d, const int C[3], int o0, int o1) {
    const int f0[3] = {2 * C[0], 2 * C[1], 2 * C[2]};
    double u[2][2][2], v[2][2][2], w[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                { const int t[3] = {0, f0[1] + b, f0[2] + c}; u[a][b][c] = ff_fine_even(F, B, 0, f0[0] + 2 * a, t); }
                { const int t[3] = {f0[0] + b, 0, f0[2] + c}; v[a][b][c] = ff_fine_even(F, B, 1, f0[1] + 2 * a, t); }
                { const int t[3] = {f0[0] + b, f0[1] + c, 0}; w[a][b][c] = ff_fine_even(F, B, 2, f0[2] + 2 * a, t); }
            }
    return ff_toth_roe(u, v, w, d, o0, o1);
}

// ---- the coarse-fine ghost elements (kernel bodies, also run on the host) ----
// Ghost cell `off` of block q (no same-level owner): state vars [v0, v0 + vc)
// of an array with nv vars per block.
GHD void ff_amr_cell(const FForest& F, int q, int off, double* st, int nv, int v0, int vc) {
    const int loc[3] = {off % F.N[0], (off / F.N[0]) % F.N[1], off / (F.N[0] * F.N[1])};
    int lev, g[3];
    ff_global(F, q, -1, loc, &lev, g);
    double* dst = st + (int64_t)q * nv * F.ncell + off;
    if (lev == 0) {   // coarse ghost inside the refined region
        const FOwner o = ff_owner(F, 0, g);
        const FCRef r{o.q, (int)ff_cidx(F, o.loc[0], o.loc[1], o.loc[2]), 1};
        for (int v = v0; v < v0 + vc; ++v) dst[v * F.ncell] = ff_vc_val(F, st, nv, v, r);
        return;
    }
    // fine ghost over a coarse leaf: slope_limited_prolong_op<minmod>
    const int X[3] = {g[0] >> 1, g[1] >> 1, g[2] >> 1};
    FCRef rc = ff_vc_ref(F, X), rm[3], rp[3];
    for (int d = 0; d < 3; ++d) {
        int Xm[3] = {X[0], X[1], X[2]}, Xp[3] = {X[0], X[1], X[2]};
        Xm[d] -= 1; Xp[d] += 1;
        rm[d] = ff_vc_ref(F, Xm); rp[d] = ff_vc_ref(F, Xp);
    }
    for (int v = v0; v < v0 + vc; ++v) {
        const double u0 = ff_vc_val(F, st, nv, v, rc);
        double uf = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double eta = (g[d] & 1 ? 1 : -1) * 0.25;
            const double sL = u0 - ff_vc_val(F, st, nv, v, rm[d]), sR = ff_vc_val(F, st, nv, v, rp[d]) - u0;
            uf += eta * gm::minmod(sL, sR);
        }
        dst[v * F.ncell] = u0 + uf;
    }
}
// Ghost face (fd << 24 | off) of block q with no data at its own level.
GHD void ff_amr_face(const FForest& F, int q, int fdoff, double* B0, double* B1, double* B2) {
    const int fd = fdoff >> 24, off = fdoff & 0xFFFFFF;
    const int E0 = F.N[0] + (fd == 0), E1 = F.N[1] + (fd == 1);
    const int loc[3] = {off % E0, (off / E0) % E1, off / (E0 * E1)};
    int lev, g[3];
    ff_global(F, q, fd, loc, &lev, g);
    const double* const B[3] = {B0, B1, B2};
    double val;
    if (lev == 0) val = ff_vcf(F, B[fd], fd, g[fd], g);
    else if ((g[fd] & 1) == 0) val = ff_fine_even(F, B, fd, g[fd], g);
    else {
        int C[3], o[2], m = 0;
        for (int d = 0; d < 3; ++d) {
            C[d] = d == fd ? (g[d] - 1) >> 1 : g[d] >> 1;
            if (d != fd) o[m++] = g[d] & 1;
        }
        val = ff_fine_mid(F, B, fd, C, o[0], o[1]);
    }
    double* Bd = fd == 0 ? B0 : (fd == 1 ? B1 : B2);
    Bd[(int64_t)q * F.nface[fd] + off] = val;
}

// Interior (odd) fine ghost face after the first pass has written every face
// lying on a coarse face: the 24 faces around its coarse cell are then in the
// block's own arrays (existing fine data, same-level copies or first-pass
// values, each equal to ff_fine_even at that position), so they are read
// instead of re-derived.  A face whose position was clamped across an outflow
// face takes the full derivation (the block's arrays there hold the clamped
// copies, not the faces around the clamped coarse cell).
GHD void ff_amr_face_mid(const FForest& F, int q, int fdoff, double* B0, double* B1, double* B2) {
    const int fd = fdoff >> 24, off = fdoff & 0xFFFFFF;
    const int E0 = F.N[0] + (fd == 0), E1 = F.N[1] + (fd == 1);
    const int loc[3] = {off % E0, (off / E0) % E1, off / (E0 * E1)};
    int lev, g[3];
    ff_global(F, q, fd, loc, &lev, g);
    const double* const B[3] = {B0, B1, B2};
    const int4 b = F.blk[q];
    const int cr[3] = {b.y, b.z, b.w};
    bool clamped = false;
    for (int d = 0; d < 3; ++d) {
        const int gg = cr[d] * F.n[d] + loc[d] - F.ng;
        const int N = ff_ntot(F, lev, d);
        clamped |= (gg < 0 || gg > N - (d == fd ? 0 : 1)) && F.bc[2 * d + (gg < 0 ? 0 : 1)] != 1;
    }
    int C[3], o[2], m = 0;
    for (int d = 0; d < 3; ++d) {
        C[d] = d == fd ? (g[d] - 1) >> 1 : g[d] >> 1;
        if (d != fd) o[m++] = g[d] & 1;
    }
    double val;
    if (clamped) {
        val = ff_fine_mid(F, B, fd, C, o[0], o[1]);
    } else {
        int lc[3];   // local index of the coarse cell's lower fine corner
        for (int d = 0; d < 3; ++d) lc[d] = loc[d] - (d == fd ? 1 : (g[d] & 1));
        double u[2][2][2], v[2][2][2], w[2][2][2];
        const double* P0 = B0 + (int64_t)q * F.nface[0];
        const double* P1 = B1 + (int64_t)q * F.nface[1];
        const double* P2 = B2 + (int64_t)q * F.nface[2];
        for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb)
                for (int c = 0; c < 2; ++c) {
                    u[a][bb][c] = P0[ff_fidx(F, 0, lc[0] + 2 * a, lc[1] + bb, lc[2] + c)];
                    v[a][bb][c] = P1[ff_fidx(F, 1, lc[0] + bb, lc[1] + 2 * a, lc[2] + c)];
                    w[a][bb][c] = P2[ff_fidx(F, 2, lc[0] + bb, lc[1] + c, lc[2] + 2 * a)];
                }
        val = ff_toth_roe(u, v, w, fd, o[0], o[1]);
    }
    double* Bd = fd == 0 ? B0 : (fd == 1 ? B1 : B2);
    Bd[(int64_t)q * F.nface[fd] + off] = val;
}

// ---- the Cowling metric (4th-order variables of the reference) --------------------
// 5-point Lagrange weights (lagrange_prolong_op<4> / lagrange_restrict_op<4>,
// pr_helpers.hh:105-330) from the Lagrange basis; numerator and denominator
// products are exact, so each weight is its exact dyadic value.  pro[r]:
// fine centre at -1/4 (r = 0) or +1/4 (r = 1) of a coarse cell from coarse
// nodes -2..2; res[s]: coarse centre from 5 fine cells starting s - 3.5 fine
// cells away (the L2, L1, CENTER, R2 stencils).
struct LagW { double pro[2][5], res[4][5]; };
inline LagW lagrange_weights() {
    auto w = [](const double* x, int d, double t) {
        double num = 1.0, den = 1.0;
        for (int m = 0; m < 5; ++m)
            if (m != d) { num *= t - x[m]; den *= x[d] - x[m]; }
        return num / den;
    };
    LagW W;
    const double xc[5] = {-2, -1, 0, 1, 2};
    for (int r = 0; r < 2; ++r)
        for (int d = 0; d < 5; ++d) W.pro[r][d] = w(xc, d, r ? 0.25 : -0.25);
    for (int st = 0; st < 4; ++st) {
        double xf[5];
        for (int d = 0; d < 5; ++d) xf[d] = st - 3.5 + d;
        for (int d = 0; d < 5; ++d) W.res[st][d] = w(xf, d, 0.0);
    }
    return W;
}
// contract1d_pairsym<5> (pr_helpers.hh:52-58): mirror pairs summed first, so
// mirrored stencils give bit-identical results.
GHD double ff_contract5(const double* c, const double u[5]) {
    const double so = c[0] * u[0] + c[4] * u[4];
    const double si = c[1] * u[1] + c[3] * u[3];
    const double sc = c[2] * u[2];
    return (so + si) + sc;
}
// lagrange_restrict_op<4>::compute_offset (pr_helpers.hh:309-330): stencil
// kept inside the fine block, mirrored between the two halves of the parent.
GHD int ff_ho_offset(int pos, int n, int ng, int half) {
    const int lb = pos - ng, ub = n + ng - pos - 1;
    if (ub == 1) return 0;
    if (lb == 0) return 3;
    if (ub == 2) return 1;
    return half ? 1 : 2;
}

// Virtual coarse metric field on every coarse cell of the domain (vars
// [v0, v0 + vc) of the state, vc-major): a coarse leaf's cell, or the
// 5-point Lagrange restriction of the fine block covering it.
__global__ __launch_bounds__(256) void k_vc_metric(FForest F, LagW W, const double* __restrict__ st, int nv, int v0,
                                                   int vc, double* __restrict__ vcm, int64_t ncc) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ncc) return;
    const int Nc0 = ff_ntot(F, 0, 0), Nc1 = ff_ntot(F, 0, 1);
    const int g[3] = {(int)(x % Nc0), (int)((x / Nc0) % Nc1), (int)(x / ((int64_t)Nc0 * Nc1))};
    const FOwner o = ff_owner(F, 0, g);
    if (o.cover == 0) {
        const double* p = st + ((int64_t)o.q * nv + v0) * F.ncell + ff_cidx(F, o.loc[0], o.loc[1], o.loc[2]);
        for (int v = 0; v < vc; ++v) vcm[v * ncc + x] = p[v * F.ncell];
        return;
    }
    const int4 b = F.blk[o.q];
    const int half[3] = {b.y & 1, b.z & 1, b.w & 1};
    int base[3];
    const double* c[3];
    for (int d = 0; d < 3; ++d) {
        const int s = ff_ho_offset(o.loc[d], F.n[d], F.ng, half[d]);
        c[d] = W.res[s];
        base[d] = o.loc[d] + s - 3;
    }
    const int64_t sj = F.N[0], sk = (int64_t)F.N[0] * F.N[1];
    for (int v = 0; v < vc; ++v) {
        const double* p = st + ((int64_t)o.q * nv + v0 + v) * F.ncell + ff_cidx(F, base[0], base[1], base[2]);
        double Sx[5][5];
        for (int dk = 0; dk < 5; ++dk)
            for (int dj = 0; dj < 5; ++dj) {
                double line[5];
                for (int di = 0; di < 5; ++di) line[di] = p[di + dj * sj + dk * sk];
                Sx[dj][dk] = ff_contract5(c[0], line);
            }
        double Sy[5];
        for (int dk = 0; dk < 5; ++dk) {
            double line[5];
            for (int dj = 0; dj < 5; ++dj) line[dj] = Sx[dj][dk];
            Sy[dk] = ff_contract5(c[1], line);
        }
        vcm[v * ncc + x] = ff_contract5(c[2], Sy);
    }
}
// Metric of the coarse-fine ghost cells (the amr cell entries) from the
// virtual coarse field: coarse ghosts take it, fine ghosts its 5-point
// Lagrange prolongation.  Ghosts beyond an outflow face keep their values
// (the metric's bc "none").
__global__ __launch_bounds__(256) void k_amr_metric(const int2* __restrict__ ent, int n, FForest F, LagW W,
                                                    double* st, int nv, int v0, int vc,
                                                    const double* __restrict__ vcm, int64_t ncc) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int2 t = ent[e];
    const int q = t.x, off = t.y;
    const int loc[3] = {off % F.N[0], (off / F.N[0]) % F.N[1], off / (F.N[0] * F.N[1])};
    const int4 b = F.blk[q];
    const int lev = b.x, cr[3] = {b.y, b.z, b.w};
    int g[3];
    for (int d = 0; d < 3; ++d) {
        const int gg = cr[d] * F.n[d] + loc[d] - F.ng;
        const int Nt = ff_ntot(F, lev, d);
        if ((gg < 0 && F.bc[2 * d] != 1) || (gg >= Nt && F.bc[2 * d + 1] != 1)) return;
        g[d] = ff_map_cc(F, lev, d, gg);
    }
    const int Nc0 = ff_ntot(F, 0, 0), Nc1 = ff_ntot(F, 0, 1);
    auto VC = [&](int v, int X0, int X1, int X2) {
        const int a0 = ff_map_cc(F, 0, 0, X0), a1 = ff_map_cc(F, 0, 1, X1), a2 = ff_map_cc(F, 0, 2, X2);
        return vcm[v * ncc + a0 + (int64_t)Nc0 * (a1 + (int64_t)Nc1 * a2)];
    };
    double* dst = st + (int64_t)q * nv * F.ncell + off;
    if (lev == 0) {
        for (int v = 0; v < vc; ++v) dst[(v0 + v) * F.ncell] = VC(v, g[0], g[1], g[2]);
        return;
    }
    const int X[3] = {g[0] >> 1, g[1] >> 1, g[2] >> 1};
    const double* c0 = W.pro[g[0] & 1];
    const double* c1 = W.pro[g[1] & 1];
    const double* c2 = W.pro[g[2] & 1];
    for (int v = 0; v < vc; ++v) {
        double Sx[5][5];
        for (int dk = 0; dk < 5; ++dk)
            for (int dj = 0; dj < 5; ++dj) {
                double line[5];
                for (int di = 0; di < 5; ++di) line[di] = VC(v, X[0] + di - 2, X[1] + dj - 2, X[2] + dk - 2);
                Sx[dj][dk] = ff_contract5(c0, line);
            }
        double Sy[5];
        for (int dk = 0; dk < 5; ++dk) {
            double line[5];
            for (int dj = 0; dj < 5; ++dj) line[dj] = Sx[dj][dk];
            Sy[dk] = ff_contract5(c1, line);
        }
        dst[(v0 + v) * F.ncell] = ff_contract5(c2, Sy);
    }
}

// ---- kernels ---------------------------------------------------------------------
// Entries {block, local cell offset}.
__global__ __launch_bounds__(256) void k_amr_cells(const int2* __restrict__ ent, int n, double* st, int nv, int v0,
                                                   int vc, FForest F) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int2 t = ent[e];
    ff_amr_cell(F, t.x, t.y, st, nv, v0, vc);
}
// Entries {block, fd << 24 | local face offset}: first pass (faces on coarse
// faces and restricted coarse ghosts), then the interior fine faces.
__global__ __launch_bounds__(256) void k_amr_faces(const int2* __restrict__ ent, int n, double* B0, double* B1,
                                                   double* B2, FForest F) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int2 t = ent[e];
    ff_amr_face(F, t.x, t.y, B0, B1, B2);
}
__global__ __launch_bounds__(256) void k_amr_faces_mid(const int2* __restrict__ ent, int n, double* B0, double* B1,
                                                       double* B2, FForest F) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int2 t = ent[e];
    ff_amr_face_mid(F, t.x, t.y, B0, B1, B2);
}

// Refluxing (refluxing.cpp:547-603): coarse face flux <- area average of the
// four fine fluxes, summed in the reference's loop order.  One thread per
// (entry, var); blockIdx.y is the var.
struct RefluxFace { int64_t dst, src; int32_t s0, s1, d, pad; };
__global__ __launch_bounds__(256) void k_reflux_flux(const RefluxFace* __restrict__ ent, int n, double* f0, double* f1,
                                                     double* f2, int64_t n0, int64_t n1, int64_t n2) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const RefluxFace r = ent[e];
    const int v = (int)blockIdx.y;
    double* f = r.d == 0 ? f0 : (r.d == 1 ? f1 : f2);
    const int64_t nf = r.d == 0 ? n0 : (r.d == 1 ? n1 : n2);
    const double* s = f + r.src + v * nf;
    const double sum = ((s[0] + s[r.s1]) + s[r.s0]) + s[r.s0 + r.s1];
    f[r.dst + v * nf] = 0.25 * sum;
}
// Coarse edge EMF <- average of the two fine edge EMFs along it
// (refluxing.cpp:708-806, 959-1173).
struct RefluxEdge { int64_t dst, src; int32_t stride, pad; };
__global__ __launch_bounds__(256) void k_reflux_emf(const RefluxEdge* __restrict__ ent, int n, double* emf) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const RefluxEdge r = ent[e];
    emf[r.dst] = 0.5 * (emf[r.src] + emf[r.src + r.stride]);
}

// ---- host: gather-table sources and refluxing entries ------------------------------
// Source of ghost element loc (face direction fd, -1 for cells) of block q of
// a forest: a same-level copy (qs, off), flagged GSRC_OWN where the position
// was clamped across an outflow face, or GSRC_AMR (prolongation /
// restriction, which also runs its own c2p).
inline void forest_source(const FForest& F, int q, int fd, const int loc[3], int& qs, int& off) {
    int lev, g[3];
    ff_global(F, q, fd, loc, &lev, g);
    if (fd < 0) {
        const FOwner o = ff_owner(F, lev, g);
        if (o.cover != lev) { qs = GSRC_AMR | GSRC_OWN; off = 0; return; }
        qs = o.q;
        off = (int)ff_cidx(F, o.loc[0], o.loc[1], o.loc[2]);
        const int4 b = F.blk[q];
        const int cr[3] = {b.y, b.z, b.w};
        for (int d = 0; d < 3; ++d) {
            const int gg = cr[d] * F.n[d] + loc[d] - F.ng;
            if ((gg < 0 && F.bc[2 * d] == 0) || (gg >= ff_ntot(F, lev, d) && F.bc[2 * d + 1] == 0)) qs |= GSRC_OWN;
        }
        return;
    }
    int sq, sl[3];
    if (ff_face_src(F, lev, fd, g[fd], g, sq, sl)) { qs = sq; off = (int)ff_fidx(F, fd, sl[0], sl[1], sl[2]); return; }
    qs = GSRC_AMR; off = 0;
}

// Refluxing entries (refluxing.cpp:547-603, 708-806, 959-1173): every face of
// a coarse leaf whose neighbour across is refined, and every edge on a coarse
// leaf's surface that a fine cell touches.
inline void forest_reflux(const FForest& F, std::vector<RefluxFace>& rf, std::vector<RefluxEdge>& re) {
    const int ng = F.ng;
    struct { int n[3], N[3]; int64_t nface[3], nedge; } G;
    for (int d = 0; d < 3; ++d) { G.n[d] = F.n[d]; G.N[d] = F.N[d]; G.nface[d] = F.nface[d]; }
    G.nedge = (int64_t)(F.N[0] + 1) * (F.N[1] + 1) * (F.N[2] + 1);
    const int64_t nq = F.nq;
    auto eix = [&](int i, int j, int k) { return i + (int64_t)(G.N[0] + 1) * (j + (int64_t)(G.N[1] + 1) * k); };
    for (int64_t q = 0; q < nq; ++q) {
        const int4 b = F.blk[q];
        if (b.x != 0) continue;
        const int cr[3] = {b.y, b.z, b.w};
        for (int d = 0; d < 3; ++d) {
            const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
            const int64_t sf[3] = {1, G.N[0] + (d == 0), (int64_t)(G.N[0] + (d == 0)) * (G.N[1] + (d == 1))};
            for (int side = 0; side < 2; ++side) {
                int cb[3] = {cr[0], cr[1], cr[2]};
                cb[d] += side ? 1 : -1;
                if (cb[d] < 0 || cb[d] >= F.nb[d]) {
                    if (F.bc[2 * d + side] != 1) continue;
                    cb[d] = (cb[d] + F.nb[d]) % F.nb[d];
                }
                const int m = F.cmap[cb[0] + F.nb[0] * (cb[1] + F.nb[1] * cb[2])];
                if (m >= 0) continue;
                const int c0 = -m - 1;
                for (int bb = 0; bb < G.n[t1]; ++bb)
                    for (int aa = 0; aa < G.n[t0]; ++aa) {
                        const int ch = ((side ? 0 : 1) << d) | (((2 * aa) / G.n[t0]) << t0) | (((2 * bb) / G.n[t1]) << t1);
                        int lc[3], lf[3];
                        lc[d] = side ? G.n[d] + ng : ng; lc[t0] = ng + aa; lc[t1] = ng + bb;
                        lf[d] = side ? ng : G.n[d] + ng; lf[t0] = ng + (2 * aa) % G.n[t0]; lf[t1] = ng + (2 * bb) % G.n[t1];
                        RefluxFace r;
                        r.dst = q * gm::NHRSC * G.nface[d] + ff_fidx(F, d, lc[0], lc[1], lc[2]);
                        r.src = (int64_t)(c0 + ch) * gm::NHRSC * G.nface[d] + ff_fidx(F, d, lf[0], lf[1], lf[2]);
                        r.s0 = (int32_t)sf[t0]; r.s1 = (int32_t)sf[t1]; r.d = d; r.pad = 0;
                        rf.push_back(r);
                    }
            }
            // edges along d on the block surface
            const int64_t se[3] = {1, G.N[0] + 1, (int64_t)(G.N[0] + 1) * (G.N[1] + 1)};
            for (int bb = ng; bb <= G.n[t1] + ng; ++bb)
                for (int aa = ng; aa <= G.n[t0] + ng; ++aa) {
                    if (!(aa == ng || aa == G.n[t0] + ng || bb == ng || bb == G.n[t1] + ng)) continue;
                    for (int ic = ng; ic < G.n[d] + ng; ++ic) {
                        const int gd = ff_map_cc(F, 0, d, cr[d] * G.n[d] + ic - ng);
                        const int A0 = cr[t0] * G.n[t0] + aa - ng, B0 = cr[t1] * G.n[t1] + bb - ng;
                        int qf = -1, lf[3] = {0, 0, 0};
                        for (int sb = 0; sb < 2 && qf < 0; ++sb)
                            for (int sa = 0; sa < 2 && qf < 0; ++sa) {
                                int ca = A0 - 1 + sa, cb2 = B0 - 1 + sb;
                                const int Na = ff_ntot(F, 0, t0), Nb = ff_ntot(F, 0, t1);
                                if ((ca < 0 || ca >= Na) && F.bc[2 * t0 + (ca < 0 ? 0 : 1)] != 1) continue;
                                if ((cb2 < 0 || cb2 >= Nb) && F.bc[2 * t1 + (cb2 < 0 ? 0 : 1)] != 1) continue;
                                ca = ((ca % Na) + Na) % Na; cb2 = ((cb2 % Nb) + Nb) % Nb;
                                int g[3]; g[d] = gd; g[t0] = ca; g[t1] = cb2;
                                const FOwner o = ff_owner(F, 0, g);
                                if (o.cover != 1) continue;
                                qf = o.q;
                                lf[0] = o.loc[0]; lf[1] = o.loc[1]; lf[2] = o.loc[2];
                                lf[t0] += sa ? 0 : 2;
                                lf[t1] += sb ? 0 : 2;
                            }
                        if (qf < 0) continue;
                        int lc[3]; lc[d] = ic; lc[t0] = aa; lc[t1] = bb;
                        RefluxEdge r;
                        r.dst = (q * 3 + d) * G.nedge + eix(lc[0], lc[1], lc[2]);
                        r.src = ((int64_t)qf * 3 + d) * G.nedge + eix(lf[0], lf[1], lf[2]);
                        r.stride = (int32_t)se[d]; r.pad = 0;
                        re.push_back(r);
                    }
                }
        }
    }
}


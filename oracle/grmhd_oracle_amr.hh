// grmhd_oracle_amr.hh -- CPU restatement of the coarse-fine ghost exchange and
// of refluxing on a two-level block forest.  TEST INFRASTRUCTURE: used only by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
//
// Forest: a uniform array of nb[0] x nb[1] x nb[2] coarse blocks, some of
// which are refined once into 8 children (p4est Morton child id
// c = cx + 2 cy + 4 cz).  Block order: the coarse leaves in linear order
// (bi + nb0 (bj + nb1 bk)), then the 8 children of every refined block, the
// refined blocks taken in ascending linear order.
//
// Ghost values, per element, from its global position at its own level
// (wrapped for periodic faces, clamped for outflow faces, as the same-level
// exchange and the physical BC of the uniform case do):
//   * same-level owner          -> copy                      (copy_kernels.hh)
//   * coarse ghost, fine owner  -> restriction of the fine data
//   * fine ghost, coarse owner  -> prolongation from the coarse data
// The coarse data seen by both is the "virtual coarse field": a coarse leaf's
// own cell / face, or the restriction of the fine block covering the position.
// That is what the reference's coarse buffers hold: restrict_op of the fine
// block's interior (restrict_kernels.hh:43-107, second_order_restrict_op,
// pr_helpers.hh:198-220), copies from coarse neighbours and from same-level
// neighbours' coarse buffers (amr_ghosts.cpp:293-420), phys BC on the buffer.
//   cell-centred hydro: slope_limited_prolong_op<minmod> (pr_helpers.hh:71-103),
//                       second-order restriction (pair-symmetric 8-average)
//   face-staggered B:   area-average restriction (restrict_kernels.hh:125-164),
//                       Toth-Roe divergence-preserving prolongation
//                       (prolongation_kernels.hh:111-386)
// Cells are cubic (grace_initialize.cpp:243-248): the Toth-Roe formula below
// is the one for equal spacings.
// Refluxing (refluxing.cpp:547-603, 708-806): the coarse-side flux of every
// coarse face on a coarse-fine interface becomes the area average of the four
// fine fluxes; every coarse edge touched by a fine cell takes the average of
// the two fine edge EMFs along it.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

struct Forest {
    int nb[3];
    int bc[6];
    int n[3], ng;
    std::vector<int64_t> cmap;           // per coarse block: q >= 0 leaf, or -(1 + q of child 0)
    std::vector<int> level;              // per block
    std::vector<std::array<int, 3>> crd; // per block, at its level
    int64_t nq = 0;

    Forest(const int nb_[3], const int bc_[6], const int n_[3], int ng_, const int32_t* refined, int n_ref) {
        for (int d = 0; d < 3; ++d) { nb[d] = nb_[d]; n[d] = n_[d]; }
        for (int d = 0; d < 6; ++d) bc[d] = bc_[d];
        ng = ng_;
        const int64_t nc = (int64_t)nb[0] * nb[1] * nb[2];
        std::vector<char> ref(nc, 0);
        for (int r = 0; r < n_ref; ++r) ref[refined[r]] = 1;
        cmap.assign(nc, 0);
        for (int64_t b = 0; b < nc; ++b)
            if (!ref[b]) {
                cmap[b] = nq++;
                level.push_back(0);
                crd.push_back({(int)(b % nb[0]), (int)((b / nb[0]) % nb[1]), (int)(b / ((int64_t)nb[0] * nb[1]))});
            }
        for (int64_t b = 0; b < nc; ++b)
            if (ref[b]) {
                cmap[b] = -(1 + nq);
                int cb[3] = {(int)(b % nb[0]), (int)((b / nb[0]) % nb[1]), (int)(b / ((int64_t)nb[0] * nb[1]))};
                for (int c = 0; c < 8; ++c) {
                    level.push_back(1);
                    crd.push_back({2 * cb[0] + (c & 1), 2 * cb[1] + ((c >> 1) & 1), 2 * cb[2] + ((c >> 2) & 1)});
                    nq++;
                }
            }
    }

    int ntot(int lev, int d) const { return nb[d] * n[d] << lev; }
    // cell index along d at level lev -> in [0, N)
    int map_cc(int lev, int d, int g) const {
        const int N = ntot(lev, d);
        if (g < 0) return bc[2 * d] == 1 ? ((g % N) + N) % N : 0;
        if (g >= N) return bc[2 * d + 1] == 1 ? g % N : N - 1;
        return g;
    }
    // face index along its normal d at level lev -> in [0, N]
    int map_fc(int lev, int d, int f) const {
        const int N = ntot(lev, d);
        if (f < 0) return bc[2 * d] == 1 ? ((f % N) + N) % N : 0;
        if (f > N) return bc[2 * d + 1] == 1 ? f % N : N;
        return f;
    }

    // Owner of the (mapped) cell g[] at level lev: block q, its level
    // (cover), and the block-local index of the cell (cover == lev), of the
    // coarse cell containing it (lev 1, cover 0), or of the lower-corner fine
    // child cell (lev 0, cover 1).
    struct Owner { int64_t q; int cover; int loc[3]; };
    Owner owner(int lev, const int g[3]) const {
        int cb[3];
        for (int d = 0; d < 3; ++d) cb[d] = lev == 0 ? g[d] / n[d] : (g[d] / n[d]) / 2;
        const int64_t m = cmap[cb[0] + (int64_t)nb[0] * (cb[1] + (int64_t)nb[1] * cb[2])];
        Owner o;
        if (m >= 0) {
            o.q = m; o.cover = 0;
            for (int d = 0; d < 3; ++d) o.loc[d] = (lev == 0 ? g[d] : g[d] / 2) % n[d] + ng;
        } else {
            o.cover = 1;
            int ch = 0;
            for (int d = 0; d < 3; ++d) {
                const int gf = lev == 1 ? g[d] : 2 * g[d];
                ch |= ((gf / n[d]) & 1) << d;
                o.loc[d] = gf % n[d] + ng;
            }
            o.q = -m - 1 + ch;
        }
        return o;
    }
};

// Grid accessors shared with the block stepper.
struct AmrArrays {
    int Ni, Nj, Nk, nv;
    int64_t cc(int i, int j, int k, int v, int64_t q) const {
        return i + (int64_t)Ni * (j + (int64_t)Nj * (k + (int64_t)Nk * (v + (int64_t)nv * q)));
    }
    int64_t fc(int d, int i, int j, int k, int64_t q) const {
        const int ei = Ni + (d == 0), ej = Nj + (d == 1), ek = Nk + (d == 2);
        return i + (int64_t)ei * (j + (int64_t)ej * (k + (int64_t)ek * q));
    }
};

inline double amr_minmod(double sL, double sR) {
    double a = std::copysign(1.0, sL), b = std::copysign(1.0, sR);
    return 0.5 * (a + b) * std::fmin(std::fabs(sL), std::fabs(sR));
}

// 4th-order (5-point) Lagrange weights, computed from the Lagrange basis
// (numerator and denominator products are exact, so each weight is the exact
// dyadic value): prolongation to a fine centre at -1/4 (row 0) or +1/4 (row 1)
// of a coarse cell from coarse nodes -2..2 (lagrange_prolong_op<4>,
// pr_helpers.hh:105-170); restriction to a coarse centre from 5 fine cells
// starting s - 3.5 fine cells away, s = 0..3 (the L2, L1, CENTER, R2 stencils
// of lagrange_restrict_op<4>, pr_helpers.hh:222-330).
struct LagrangeW {
    double pro[2][5], res[4][5];
    static double w(const double* x, int d, double t) {
        double num = 1.0, den = 1.0;
        for (int m = 0; m < 5; ++m)
            if (m != d) { num *= t - x[m]; den *= x[d] - x[m]; }
        return num / den;
    }
    LagrangeW() {
        const double xc[5] = {-2, -1, 0, 1, 2};
        for (int r = 0; r < 2; ++r)
            for (int d = 0; d < 5; ++d) pro[r][d] = w(xc, d, r ? 0.25 : -0.25);
        for (int st = 0; st < 4; ++st) {
            double xf[5];
            for (int d = 0; d < 5; ++d) xf[d] = st - 3.5 + d;
            for (int d = 0; d < 5; ++d) res[st][d] = w(xf, d, 0.0);
        }
    }
};
// contract1d_pairsym<5> (pr_helpers.hh:52-58): mirror pairs summed first.
inline double contract5(const double c[5], const double u[5]) {
    const double so = c[0] * u[0] + c[4] * u[4];
    const double si = c[1] * u[1] + c[3] * u[3];
    const double sc = c[2] * u[2];
    return (so + si) + sc;
}
// Tensor-product contraction over a 5x5x5 stencil, x then y then z.
template <typename U>
inline double contract555(const double cx[5], const double cy[5], const double cz[5], U u) {
    double Sx[5][5];
    for (int dk = 0; dk < 5; ++dk)
        for (int dj = 0; dj < 5; ++dj) {
            double line[5];
            for (int di = 0; di < 5; ++di) line[di] = u(di, dj, dk);
            Sx[dj][dk] = contract5(cx, line);
        }
    double Sy[5];
    for (int dk = 0; dk < 5; ++dk) {
        double line[5];
        for (int dj = 0; dj < 5; ++dj) line[dj] = Sx[dj][dk];
        Sy[dk] = contract5(cy, line);
    }
    return contract5(cz, Sy);
}

struct AmrExchange {
    const Forest& F;
    AmrArrays A;
    explicit AmrExchange(const Forest& f, int nv) : F(f) {
        A.Ni = f.n[0] + 2 * f.ng; A.Nj = f.n[1] + 2 * f.ng; A.Nk = f.n[2] + 2 * f.ng; A.nv = nv;
    }

    // ---- cell-centred -------------------------------------------------------
    // second_order_restrict_op (pr_helpers.hh:198-212): pair-symmetric sum of
    // the 8 children from the lower corner (i,j,k).
    double restrict8(const double* st, int v, int64_t q, int i, int j, int k) const {
        auto u = [&](int a, int b, int c) { return st[A.cc(i + a, j + b, k + c, v, q)]; };
        const double q00 = u(0, 0, 0) + u(1, 0, 0);
        const double q10 = u(0, 1, 0) + u(1, 1, 0);
        const double q01 = u(0, 0, 1) + u(1, 0, 1);
        const double q11 = u(0, 1, 1) + u(1, 1, 1);
        const double r0 = q00 + q10, r1 = q01 + q11;
        return 0.125 * (r0 + r1);
    }
    // Virtual coarse field at coarse cell X (unmapped).
    double vc(const double* st, int v, const int X[3]) const {
        int g[3];
        for (int d = 0; d < 3; ++d) g[d] = F.map_cc(0, d, X[d]);
        Forest::Owner o = F.owner(0, g);
        if (o.cover == 0) return st[A.cc(o.loc[0], o.loc[1], o.loc[2], v, o.q)];
        return restrict8(st, v, o.q, o.loc[0], o.loc[1], o.loc[2]);
    }
    // slope_limited_prolong_op<minmod> at the (mapped) fine cell P.
    double prolong(const double* st, int v, const int P[3]) const {
        int X[3];
        for (int d = 0; d < 3; ++d) X[d] = P[d] / 2;
        const double u0 = vc(st, v, X);
        double uf = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double eta = (P[d] & 1 ? 1 : -1) * 0.25;
            int Xm[3] = {X[0], X[1], X[2]}, Xp[3] = {X[0], X[1], X[2]};
            Xm[d] -= 1; Xp[d] += 1;
            const double sL = u0 - vc(st, v, Xm), sR = vc(st, v, Xp) - u0;
            uf += eta * amr_minmod(sL, sR);
        }
        return u0 + uf;
    }

    // ---- face-staggered ------------------------------------------------------
    // Block holding face (d; f along d, t[] tangential cells) of level lev
    // (mapped), preferring the cell above the face: returns false when neither
    // adjacent cell is covered at this level.
    bool face_src(int lev, int d, int f, const int t[3], int64_t& q, int loc[3]) const {
        const int N = F.ntot(lev, d);
        int g[3] = {t[0], t[1], t[2]};
        for (int side = 1; side >= 0; --side) {
            int c = side ? f : f - 1;
            if (c < 0 || c >= N) {
                if (F.bc[2 * d + (c < 0 ? 0 : 1)] != 1) continue;
                c = ((c % N) + N) % N;
            }
            g[d] = c;
            Forest::Owner o = F.owner(lev, g);
            if (o.cover != lev) continue;
            q = o.q;
            for (int e = 0; e < 3; ++e) loc[e] = o.loc[e];
            loc[d] += side ? 0 : 1;
            return true;
        }
        return false;
    }
    // Virtual coarse face field (d; F0 along d, T tangential), unmapped.
    double vcf(double* const B[3], int d, int F0, const int T[3]) const {
        int t[3];
        for (int e = 0; e < 3; ++e) t[e] = e == d ? 0 : F.map_cc(0, e, T[e]);
        const int f = F.map_fc(0, d, F0);
        int64_t q = 0; int loc[3] = {0, 0, 0};
        if (face_src(0, d, f, t, q, loc)) return B[d][A.fc(d, loc[0], loc[1], loc[2], q)];
        // both sides refined: div_free_restrict_op (restrict_kernels.hh:145-164)
        int tf[3];
        for (int e = 0; e < 3; ++e) tf[e] = 2 * t[e];
        face_src(1, d, 2 * f, tf, q, loc);
        double val = 0;
        for (int ii = 0; ii <= (d != 0); ++ii)
            for (int jj = 0; jj <= (d != 1); ++jj)
                for (int kk = 0; kk <= (d != 2); ++kk) val += B[d][A.fc(d, loc[0] + ii, loc[1] + jj, loc[2] + kk, q)];
        return 0.25 * val;
    }
    // Fine face on a coarse face (f even along d): existing fine data, or
    // fill_inside_face (prolongation_kernels.hh:130-160).
    double fine_even(double* const B[3], int d, int f, const int t[3]) const {
        int64_t q = 0; int loc[3] = {0, 0, 0};
        int tm[3];
        for (int e = 0; e < 3; ++e) tm[e] = e == d ? 0 : F.map_cc(1, e, t[e]);
        const int fm = F.map_fc(1, d, f);
        if (face_src(1, d, fm, tm, q, loc)) return B[d][A.fc(d, loc[0], loc[1], loc[2], q)];
        int T[3];
        for (int e = 0; e < 3; ++e) T[e] = tm[e] / 2;
        const int Fc = fm / 2;
        const double U = vcf(B, d, Fc, T);
        double r = U;
        for (int e = 0; e < 3; ++e) {
            if (e == d) continue;
            int Tm[3] = {T[0], T[1], T[2]}, Tp[3] = {T[0], T[1], T[2]};
            Tm[e] -= 1; Tp[e] += 1;
            const double s = 0.25 * amr_minmod(U - vcf(B, d, Fc, Tm), vcf(B, d, Fc, Tp) - U);
            r = r + (tm[e] & 1 ? 1 : -1) * s;
        }
        return r;
    }
    // Interior fine face of coarse cell C (Toth-Roe, prolongation_kernels.hh:330-381):
    // face (d; fine index 2 C[d] + 1; offsets o0, o1 in {0,1} along the two
    // other axes, lower axis first).
    double fine_mid(double* const B[3], int d, const int C[3], int o0, int o1) const {
        const int f0[3] = {2 * C[0], 2 * C[1], 2 * C[2]};
        auto u = [&](int a, int b, int c) { int t[3] = {0, f0[1] + b, f0[2] + c}; return fine_even(B, 0, f0[0] + a, t); };
        auto v = [&](int a, int b, int c) { int t[3] = {f0[0] + a, 0, f0[2] + c}; return fine_even(B, 1, f0[1] + b, t); };
        auto w = [&](int a, int b, int c) { int t[3] = {f0[0] + a, f0[1] + b, 0}; return fine_even(B, 2, f0[2] + c, t); };
        double Uxx = 0, Vyy = 0, Wzz = 0, Uxyz = 0, Vxyz = 0, Wxyz = 0;
        for (int ii = 0; ii <= 1; ++ii)
            for (int jj = 0; jj <= 1; ++jj)
                for (int kk = 0; kk <= 1; ++kk) {
                    const int is = ii ? 1 : -1, js = jj ? 1 : -1, ks = kk ? 1 : -1;
                    Uxx += is * js * v(ii, 2 * jj, kk) + is * ks * w(ii, jj, 2 * kk);
                    Vyy += is * js * u(2 * ii, jj, kk) + js * ks * w(ii, jj, 2 * kk);
                    Wzz += is * ks * u(2 * ii, jj, kk) + js * ks * v(ii, 2 * jj, kk);
                    Uxyz += is * js * ks * u(2 * ii, jj, kk);
                    Vxyz += is * js * ks * v(ii, 2 * jj, kk);
                    Wxyz += is * js * ks * w(ii, jj, 2 * kk);
                }
        Uxx *= 0.125; Vyy *= 0.125; Wzz *= 0.125;
        Uxyz *= 0.0625; Vxyz *= 0.0625; Wxyz *= 0.0625;
        const int s0 = o0 ? 1 : -1, s1 = o1 ? 1 : -1;
        if (d == 0) return 0.5 * (u(0, o0, o1) + u(2, o0, o1)) + Uxx + s1 * Vxyz + s0 * Wxyz;
        if (d == 1) return 0.5 * (v(o0, 0, o1) + v(o0, 2, o1)) + Vyy + s0 * Wxyz + s1 * Uxyz;
        return 0.5 * (w(o0, o1, 0) + w(o0, o1, 2)) + Wzz + s1 * Uxyz + s0 * Vxyz;
    }

    // ---- the exchange --------------------------------------------------------
    // Every ghost cell (vars [v0, v0+vc)) and ghost face of every block.
    // Reads interior data only, so the order of the writes does not matter.
    void fill(double* st, double* const B[3], int v0, int vcnt) const {
        const int ng = F.ng;
        const int N[3] = {A.Ni, A.Nj, A.Nk};
        for (int64_t q = 0; q < F.nq; ++q) {
            const int lev = F.level[q];
            const auto& cr = F.crd[q];
            for (int k = 0; k < N[2]; ++k)
                for (int j = 0; j < N[1]; ++j)
                    for (int i = 0; i < N[0]; ++i) {
                        const int loc[3] = {i, j, k};
                        bool ghost = false;
                        for (int d = 0; d < 3; ++d) ghost |= loc[d] < ng || loc[d] >= F.n[d] + ng;
                        if (!ghost) continue;
                        int g[3];
                        for (int d = 0; d < 3; ++d) g[d] = F.map_cc(lev, d, cr[d] * F.n[d] + loc[d] - ng);
                        Forest::Owner o = F.owner(lev, g);
                        for (int v = v0; v < v0 + vcnt; ++v) {
                            double val;
                            if (o.cover == lev) val = st[A.cc(o.loc[0], o.loc[1], o.loc[2], v, o.q)];
                            else if (lev == 0) val = restrict8(st, v, o.q, o.loc[0], o.loc[1], o.loc[2]);
                            else val = prolong(st, v, g);
                            st[A.cc(i, j, k, v, q)] = val;
                        }
                    }
            for (int fd = 0; fd < 3; ++fd) {
                const int E[3] = {N[0] + (fd == 0), N[1] + (fd == 1), N[2] + (fd == 2)};
                for (int k = 0; k < E[2]; ++k)
                    for (int j = 0; j < E[1]; ++j)
                        for (int i = 0; i < E[0]; ++i) {
                            const int loc[3] = {i, j, k};
                            bool ghost = false;
                            for (int d = 0; d < 3; ++d) ghost |= loc[d] < ng || loc[d] >= F.n[d] + ng + (d == fd);
                            if (!ghost) continue;
                            int t[3], f = 0;
                            for (int d = 0; d < 3; ++d) {
                                const int gg = cr[d] * F.n[d] + loc[d] - ng;
                                if (d == fd) { f = F.map_fc(lev, d, gg); t[d] = 0; }
                                else t[d] = F.map_cc(lev, d, gg);
                            }
                            double val;
                            int64_t qs = 0; int sl[3] = {0, 0, 0};
                            if (face_src(lev, fd, f, t, qs, sl)) val = B[fd][A.fc(fd, sl[0], sl[1], sl[2], qs)];
                            else if (lev == 0) {
                                int T[3] = {t[0], t[1], t[2]};
                                val = vcf(B, fd, f, T);
                            } else if ((f & 1) == 0) {
                                val = fine_even(B, fd, f, t);
                            } else {
                                int C[3];
                                for (int d = 0; d < 3; ++d) C[d] = d == fd ? (f - 1) / 2 : t[d] / 2;
                                int o[2], m = 0;
                                for (int d = 0; d < 3; ++d)
                                    if (d != fd) o[m++] = t[d] & 1;
                                val = fine_mid(B, fd, C, o[0], o[1]);
                            }
                            B[fd][A.fc(fd, i, j, k, q)] = val;
                        }
            }
        }
    }

    // ---- Cowling metric (variables registered "fourth_order",
    // variable_indices.cpp:442-446): exchanged with the evolved variables, so
    // copies between same-level blocks, 5-point Lagrange across coarse-fine
    // boundaries, and no physical BC (bc "none": ghosts beyond an outflow face
    // keep their values).  The metric is static, so one exchange after the
    // initial data gives the reference's ghost values.
    // lagrange_restrict_op<4>::compute_offset (pr_helpers.hh:309-330).
    int ho_offset(int pos, int n, int half) const {
        const int lb = pos - F.ng, ub = n + F.ng - pos - 1;
        if (ub == 1) return 0;
        if (lb == 0) return 3;
        if (ub == 2) return 1;
        return half ? 1 : 2;
    }
    // Restriction of fine block q at lower fine cell loc (child bits of q
    // select the half-dependent stencil).
    double restrict_ho(const LagrangeW& W, const double* st, int v, int64_t q, const int loc[3]) const {
        int off[3];
        const double* c[3];
        for (int d = 0; d < 3; ++d) {
            const int o = ho_offset(loc[d], F.n[d], F.crd[q][d] & 1);
            off[d] = o - 3;
            c[d] = W.res[o];
        }
        return contract555(c[0], c[1], c[2], [&](int a, int b, int cc) {
            return st[A.cc(loc[0] + off[0] + a, loc[1] + off[1] + b, loc[2] + off[2] + cc, v, q)];
        });
    }
    void fill_metric(double* st, int v0, int vcnt) const {
        LagrangeW W;
        const int ng = F.ng;
        const int N[3] = {A.Ni, A.Nj, A.Nk};
        // the virtual coarse metric field on every coarse cell of the domain
        const int Nc[3] = {F.ntot(0, 0), F.ntot(0, 1), F.ntot(0, 2)};
        const int64_t ncc = (int64_t)Nc[0] * Nc[1] * Nc[2];
        std::vector<double> vcm((size_t)(ncc * vcnt));
        for (int64_t x = 0; x < ncc; ++x) {
            const int g[3] = {(int)(x % Nc[0]), (int)((x / Nc[0]) % Nc[1]), (int)(x / ((int64_t)Nc[0] * Nc[1]))};
            const Forest::Owner o = F.owner(0, g);
            for (int v = 0; v < vcnt; ++v)
                vcm[(size_t)(v * ncc + x)] = o.cover == 0 ? st[A.cc(o.loc[0], o.loc[1], o.loc[2], v0 + v, o.q)]
                                                          : restrict_ho(W, st, v0 + v, o.q, o.loc);
        }
        auto VC = [&](int v, const int X[3]) {
            int g[3];
            for (int d = 0; d < 3; ++d) g[d] = F.map_cc(0, d, X[d]);
            return vcm[(size_t)(v * ncc + g[0] + (int64_t)Nc[0] * (g[1] + (int64_t)Nc[1] * g[2]))];
        };
        std::vector<double> out;
        std::vector<int64_t> where;
        for (int64_t q = 0; q < F.nq; ++q) {
            const int lev = F.level[q];
            const auto& cr = F.crd[q];
            for (int k = 0; k < N[2]; ++k)
                for (int j = 0; j < N[1]; ++j)
                    for (int i = 0; i < N[0]; ++i) {
                        const int loc[3] = {i, j, k};
                        bool ghost = false, outside = false;
                        int g[3];
                        for (int d = 0; d < 3; ++d) {
                            ghost |= loc[d] < ng || loc[d] >= F.n[d] + ng;
                            const int gg = cr[d] * F.n[d] + loc[d] - ng;
                            const int Nt = F.ntot(lev, d);
                            if ((gg < 0 && F.bc[2 * d] != 1) || (gg >= Nt && F.bc[2 * d + 1] != 1)) outside = true;
                            g[d] = F.map_cc(lev, d, gg);
                        }
                        if (!ghost || outside) continue;
                        const Forest::Owner o = F.owner(lev, g);
                        for (int v = 0; v < vcnt; ++v) {
                            double val;
                            if (o.cover == lev) val = st[A.cc(o.loc[0], o.loc[1], o.loc[2], v0 + v, o.q)];
                            else if (lev == 0) val = VC(v, g);
                            else {
                                const int X[3] = {g[0] / 2, g[1] / 2, g[2] / 2};
                                const double* c[3] = {W.pro[g[0] & 1], W.pro[g[1] & 1], W.pro[g[2] & 1]};
                                val = contract555(c[0], c[1], c[2], [&](int a, int b, int cc) {
                                    const int Y[3] = {X[0] + a - 2, X[1] + b - 2, X[2] + cc - 2};
                                    return VC(v, Y);
                                });
                            }
                            out.push_back(val);
                            where.push_back(A.cc(i, j, k, v0 + v, q));
                        }
                    }
        }
        for (size_t x = 0; x < out.size(); ++x) st[where[x]] = out[x];
    }

    // ---- refluxing -------------------------------------------------------------
    // Coarse neighbour across face (d, side) of coarse block q, -1 outside an
    // outflow domain; sets `refined` when that neighbour is refined and returns
    // the q of its child 0.
    int64_t across(int64_t q, int d, int side, bool& refined) const {
        int cb[3] = {F.crd[q][0], F.crd[q][1], F.crd[q][2]};
        cb[d] += side ? 1 : -1;
        refined = false;
        if (cb[d] < 0 || cb[d] >= F.nb[d]) {
            if (F.bc[2 * d + side] != 1) return -1;
            cb[d] = (cb[d] + F.nb[d]) % F.nb[d];
        }
        const int64_t m = F.cmap[cb[0] + (int64_t)F.nb[0] * (cb[1] + (int64_t)F.nb[1] * cb[2])];
        if (m < 0) { refined = true; return -m - 1; }
        return m;
    }
    // reflux_correct_fluxes (refluxing.cpp:547-603); flux[d] holds nvh vars.
    void reflux_fluxes(double* const flux[3], int nvh) const {
        const int ng = F.ng;
        for (int64_t q = 0; q < F.nq; ++q) {
            if (F.level[q] != 0) continue;
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    bool refined;
                    const int64_t c0 = across(q, d, side, refined);
                    if (!refined) continue;
                    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
                    for (int b = 0; b < F.n[t1]; ++b)
                        for (int a = 0; a < F.n[t0]; ++a) {
                            int ch = (side ? 0 : 1) << d;
                            ch |= ((2 * a) / F.n[t0]) << t0;
                            ch |= ((2 * b) / F.n[t1]) << t1;
                            const int64_t qf = c0 + ch;
                            int lc[3], lf[3];
                            lc[d] = side ? F.n[d] + ng : ng;
                            lc[t0] = ng + a; lc[t1] = ng + b;
                            lf[d] = side ? ng : F.n[d] + ng;
                            lf[t0] = ng + (2 * a) % F.n[t0]; lf[t1] = ng + (2 * b) % F.n[t1];
                            const int e0[3] = {t0 == 0, t0 == 1, t0 == 2}, e1[3] = {t1 == 0, t1 == 1, t1 == 2};
                            for (int v = 0; v < nvh; ++v) {
                                auto Fx = [&](int x, int y) {
                                    auto idx = [&](int i, int j, int k) {
                                        const int ei = A.Ni + (d == 0), ej = A.Nj + (d == 1), ek = A.Nk + (d == 2);
                                        return i + (int64_t)ei * (j + (int64_t)ej * (k + (int64_t)ek * (v + (int64_t)nvh * qf)));
                                    };
                                    return flux[d][idx(lf[0] + x * e0[0] + y * e1[0], lf[1] + x * e0[1] + y * e1[1],
                                                       lf[2] + x * e0[2] + y * e1[2])];
                                };
                                const double s = ((Fx(0, 0) + Fx(0, 1)) + Fx(1, 0)) + Fx(1, 1);
                                const int ei = A.Ni + (d == 0), ej = A.Nj + (d == 1), ek = A.Nk + (d == 2);
                                flux[d][lc[0] + (int64_t)ei * (lc[1] + (int64_t)ej * (lc[2] + (int64_t)ek * (v + (int64_t)nvh * q)))] = 0.25 * s;
                            }
                        }
                }
        }
    }
    // reflux_correct_emfs (refluxing.cpp:708-806 and the edge passes
    // :959-1173): emf[(c + 3 q)][edge (Ni+1)(Nj+1)(Nk+1)].
    void reflux_emfs(double* emf) const {
        const int ng = F.ng;
        const int64_t ne = (int64_t)(A.Ni + 1) * (A.Nj + 1) * (A.Nk + 1);
        auto eix = [&](int i, int j, int k) { return i + (int64_t)(A.Ni + 1) * (j + (int64_t)(A.Nj + 1) * k); };
        for (int64_t q = 0; q < F.nq; ++q) {
            if (F.level[q] != 0) continue;
            const auto& cr = F.crd[q];
            for (int d = 0; d < 3; ++d) {
                const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
                for (int b = ng; b <= F.n[t1] + ng; ++b)
                    for (int a = ng; a <= F.n[t0] + ng; ++a) {
                        const bool rim = a == ng || a == F.n[t0] + ng || b == ng || b == F.n[t1] + ng;
                        if (!rim) continue;
                        for (int ic = ng; ic < F.n[d] + ng; ++ic) {
                            // global coarse: cell along d, nodes along t0, t1
                            const int gd = F.map_cc(0, d, cr[d] * F.n[d] + ic - ng);
                            const int A0 = cr[t0] * F.n[t0] + a - ng, B0 = cr[t1] * F.n[t1] + b - ng;
                            int64_t qf = -1; int lf[3] = {0, 0, 0};
                            for (int sb = 0; sb < 2 && qf < 0; ++sb)
                                for (int sa = 0; sa < 2 && qf < 0; ++sa) {
                                    int ca = A0 - 1 + sa, cb2 = B0 - 1 + sb;
                                    const int Na = F.ntot(0, t0), Nb = F.ntot(0, t1);
                                    if ((ca < 0 || ca >= Na) && F.bc[2 * t0 + (ca < 0 ? 0 : 1)] != 1) continue;
                                    if ((cb2 < 0 || cb2 >= Nb) && F.bc[2 * t1 + (cb2 < 0 ? 0 : 1)] != 1) continue;
                                    ca = ((ca % Na) + Na) % Na; cb2 = ((cb2 % Nb) + Nb) % Nb;
                                    int g[3]; g[d] = gd; g[t0] = ca; g[t1] = cb2;
                                    Forest::Owner o = F.owner(0, g);
                                    if (o.cover != 1) continue;
                                    qf = o.q;
                                    lf[0] = o.loc[0]; lf[1] = o.loc[1]; lf[2] = o.loc[2];
                                    lf[t0] += sa ? 0 : 2;   // the node is on the high side of cell A0-1
                                    lf[t1] += sb ? 0 : 2;
                                }
                            if (qf < 0) continue;
                            const int ed[3] = {d == 0, d == 1, d == 2};
                            const double e1 = emf[(d + 3 * qf) * ne + eix(lf[0], lf[1], lf[2])];
                            const double e2 = emf[(d + 3 * qf) * ne + eix(lf[0] + ed[0], lf[1] + ed[1], lf[2] + ed[2])];
                            int lc[3]; lc[d] = ic; lc[t0] = a; lc[t1] = b;
                            emf[(d + 3 * q) * ne + eix(lc[0], lc[1], lc[2])] = 0.5 * (e1 + e2);
                        }
                    }
            }
        }
    }
};

}  // namespace oracle

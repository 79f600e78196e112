// grmhd_oracle_rhs.hh -- CPU oracle for the block-level GRMHD substep.
//
// TEST INFRASTRUCTURE ONLY (see grmhd_oracle.hh).  Restates the reference's
// multi-pass substep on the reference's own (i,j,k,var,q) LayoutLeft arrays:
//   compute_fluxes      src/evolution/evolve.cpp:813-1068, getflux
//                       include/grace/physics/grmhd.hh:551-970
//   compute_emfs        evolve.cpp:1069-1129 (GS), :1130-1340 (UCT)
//   add_fluxes_and_source_terms  evolve.cpp:1342-1404, grmhd.hh:190-297
//   update_CT           evolve.cpp:1406-1479
//   linop_apply         include/grace/utils/tstep_utils.hh:54-272
//   RK3                 evolve.cpp:190-225
//   ghost fill          src/amr/boundary_conditions.cpp:55-81 (same-level copy
//                       + outflow extrap_bc_t<0>, phys_bc_kernels.hh:76-89)
//   compute_auxiliaries src/evolution/auxiliaries.cpp:81-180, grmhd.hh:306-434
//   find_stable_timestep src/evolution/find_stable_timestep.cpp:72-126,
//                       grmhd.hh:444-519
// The flux pass evaluates faces on the normal range [ngz, n+ngz] with full
// transverse range, which is every face any later pass reads; the reference's
// wider normal range only adds faces nobody reads when FOFC is off.
#pragma once
#include "grmhd_oracle.hh"
#include "grmhd_oracle_amr.hh"
#include <vector>
#include <cstdio>

// Operation-count hooks: the counting build (flop_count.hh) tallies the
// operations of each pass under its id; no-ops in the checker build.
// ORC_PASS: 1 fluxes, 2 edge EMFs, 3 flux divergence, 4 CT update, 5 c2p
// over cells, 6 RK linear combinations, 7 FOFC, 0 everything else.
// ORC_SUB(8) tags the geometric source terms within pass 3; ORC_METRIC
// tags metric set-up and metric derivatives (id 10 + the enclosing pass),
// which a flat-metric kernel does not evaluate.
#ifndef ORC_PASS
#define ORC_PASS(id) ((void)0)
#define ORC_SUB(id) ((void)0)
#define ORC_METRIC() ((void)0)
#endif

namespace oracle {

enum Recon { REC_WENOZ = 0, REC_MC2 = 1, REC_MINMOD = 2, REC_DONOR = 3 };
enum Riemann { RS_HLLE = 0, RS_LLF = 1 };
enum EmfScheme { EMF_UCT = 0, EMF_GS = 1 };
enum Thermo { TH_TEMP = 0, TH_PRESS = 1 };
enum BCKind { BC_OUTFLOW = 0, BC_PERIODIC = 1 };

struct Options {
    int recon = REC_WENOZ, riemann = RS_HLLE, emf = EMF_UCT, thermo = TH_TEMP, fd_order = 6;
    double rusanov_c_limit = 0.0;
    // first-order flux correction (evolve.cpp:456-811, grmhd_helpers.hh:63-75,226-231)
    int fofc = 0, dmp_enable = 1;
    double dmp_M = 1.2;
    // 1: GS edge EMFs one ghost layer wider with FOFC, so the flag pass's
    // ghost-layer cells read computed edges (grmhd_params.fofc_ghost_emf)
    int fofc_ghost_emf = 0;
};

// Block grid.  All blocks share (nx,ny,nz,ngz).
struct Grid {
    int nx, ny, nz, ngz;
    int64_t nq;
    int Ni, Nj, Nk;
    Grid(int nx_, int ny_, int nz_, int ngz_, int64_t nq_) : nx(nx_), ny(ny_), nz(nz_), ngz(ngz_), nq(nq_) {
        Ni = nx + 2 * ngz; Nj = ny + 2 * ngz; Nk = nz + 2 * ngz;
    }
    // cell-centred, nv variables
    int64_t cc(int i, int j, int k, int v, int64_t q, int nv) const {
        return i + (int64_t)Ni * (j + (int64_t)Nj * (k + (int64_t)Nk * (v + (int64_t)nv * q)));
    }
    // face-staggered in direction d (extent +1 along d), nv variables
    int64_t fc(int d, int i, int j, int k, int v, int64_t q, int nv) const {
        int ei = Ni + (d == 0), ej = Nj + (d == 1), ek = Nk + (d == 2);
        return i + (int64_t)ei * (j + (int64_t)ej * (k + (int64_t)ek * (v + (int64_t)nv * q)));
    }
    int64_t fc_size(int d, int nv) const {
        return (int64_t)(Ni + (d == 0)) * (Nj + (d == 1)) * (Nk + (d == 2)) * nv * nq;
    }
    // edge arrays: extents (Ni+1, Nj+1, Nk+1), 3 components
    int64_t ed(int i, int j, int k, int c, int64_t q) const {
        return i + (int64_t)(Ni + 1) * (j + (int64_t)(Nj + 1) * (k + (int64_t)(Nk + 1) * (c + 3 * q)));
    }
    int64_t ed_size() const { return (int64_t)(Ni + 1) * (Nj + 1) * (Nk + 1) * 3 * nq; }
    int64_t cc_size(int nv) const { return (int64_t)Ni * Nj * Nk * nv * nq; }
};

// Uniform Cartesian instance cut into blocks; blocks of instance laid out
// q = q0 + bi + nb[0]*(bj + nb[1]*bk).
struct Instance {
    int nb[3];
    int64_t q0;
    int bc[6];       // -x,+x,-y,+y,-z,+z
    double idx[3];   // inverse spacings
    double x0[3] = {0.0, 0.0, 0.0};   // lower corner of the interior
};

struct BlockDesc { int inst; int b[3]; };

struct Layout {
    Grid grid;
    std::vector<Instance> inst;
    std::vector<BlockDesc> blk;
    const AmrExchange* amr = nullptr;   // two-level forest: coarse-fine exchange + reflux
    Layout(const Grid& g) : grid(g) {}
    double idx(int64_t q, int d) const { return inst[blk[q].inst].idx[d]; }
};

// The state arrays of one time level.
struct Fields {
    double* state;   // NSTATE vars
    double* B[3];    // face-staggered densitised B
};

// ---------------------------------------------------------------------------
// Reconstruction.  WENO-Z: include/grace/utils/weno_reconstruction.hh:172-278;
// MC2 / minmod / donor: reconstruction.hh:51-142, limiters.hh:42-80.
// Returns uL (i-1/2 - eps) and uR (i-1/2 + eps) from the 6-point stencil
// u[0..5] = u(i-3..i+2).
// ---------------------------------------------------------------------------
inline void recon_wenoz(const double u[6], double& uL, double& uR) {
    double um3 = u[0], um2 = u[1], um1 = u[2], u0 = u[3], up1 = u[4], up2 = u[5];
    const double c0 = 13.0 / 12.0, c1 = 0.25, eps = 1.0e-42;
    {
        double b0 = c0 * sqr(um2 + u0 - 2.0 * um1) + c1 * sqr(um2 + 3.0 * u0 - 4.0 * um1);
        double b1 = c0 * sqr(um1 + up1 - 2.0 * u0) + c1 * sqr(um1 - up1);
        double b2 = c0 * sqr(up2 + u0 - 2.0 * up1) + c1 * sqr(up2 + 3.0 * u0 - 4.0 * up1);
        double t5 = std::fabs(b0 - b2);
        double i0 = sqr(t5 / (b0 + eps)), i1 = sqr(t5 / (b1 + eps)), i2 = sqr(t5 / (b2 + eps));
        double f0 = (2.0 * up2 - 7.0 * up1 + 11.0 * u0);
        double f1 = (-1.0 * up1 + 5.0 * u0 + 2.0 * um1);
        double f2 = (2.0 * u0 + 5.0 * um1 - um2);
        double a0 = 0.1 * (1.0 + i2), a1 = 0.6 * (1.0 + i1), a2 = 0.3 * (1.0 + i0);
        double asum = 6.0 * (a0 + a1 + a2);
        uR = (f0 * a0 + f1 * a1 + f2 * a2) / asum;
    }
    {
        double b0 = c0 * sqr(um3 + um1 - 2.0 * um2) + c1 * sqr(um3 + 3.0 * um1 - 4.0 * um2);
        double b1 = c0 * sqr(um2 + u0 - 2.0 * um1) + c1 * sqr(um2 - u0);
        double b2 = c0 * sqr(up1 + um1 - 2.0 * u0) + c1 * sqr(up1 + 3.0 * um1 - 4.0 * u0);
        double t5 = std::fabs(b0 - b2);
        double i0 = sqr(t5 / (b0 + eps)), i1 = sqr(t5 / (b1 + eps)), i2 = sqr(t5 / (b2 + eps));
        double f0 = (2.0 * um3 - 7.0 * um2 + 11.0 * um1);
        double f1 = (-1.0 * um2 + 5.0 * um1 + 2.0 * u0);
        double f2 = (2.0 * um1 + 5.0 * u0 - up1);
        double a0 = 0.1 * (1.0 + i0), a1 = 0.6 * (1.0 + i1), a2 = 0.3 * (1.0 + i2);
        double asum = 6.0 * (a0 + a1 + a2);
        uL = (f0 * a0 + f1 * a1 + f2 * a2) / asum;
    }
}
inline double lim_mc2(double sL, double sR) {
    double sC = 0.5 * (sR + sL);
    double sgn = std::copysign(1.0, sR);
    return sgn * std::fmax(0., std::fmin(std::fmin(2. * std::fabs(sR), 2. * sgn * sL), sgn * sC));
}
inline double lim_minmod(double sL, double sR) {
    double a = std::copysign(1.0, sL), b = std::copysign(1.0, sR);
    return 0.5 * (a + b) * std::fmin(std::fabs(sL), std::fabs(sR));
}
inline void recon_any(int kind, const double u[6], double& uL, double& uR) {
    if (kind == REC_WENOZ) { recon_wenoz(u, uL, uR); return; }
    if (kind == REC_DONOR) { uL = u[2]; uR = u[3]; return; }
    // slope limited: u(imm)=u[1], u(im)=u[2], u(i)=u[3], u(ip)=u[4]
    double (*lim)(double, double) = (kind == REC_MC2) ? lim_mc2 : lim_minmod;
    double sL = u[3] - u[2], sR = u[4] - u[3];
    uR = u[3] - 0.5 * lim(sL, sR);
    sL = u[2] - u[1]; sR = u[3] - u[2];
    uL = u[2] + 0.5 * lim(sL, sR);
}

inline double hll(double fL, double fR, double uL, double uR, double cmin, double cmax) {
    return (cmax * fL + cmin * fR - cmax * cmin * (uR - uL)) / (cmax + cmin);   // riemann_solvers.hh:69
}

// face_interp_4 (grmhd_helpers.hh:360-380): value at the face between u1 and
// u2 from the four cells u0..u3 around it, the inner and outer pairs summed
// first (pair-symmetric, so mirrored data give bit-mirrored results).
inline double face_interp_4(double u0, double u1, double u2, double u3) {
    const double si = u1 + u2, so = u0 + u3;
    return 0.5625 * si + (-0.0625) * so;
}

// Centred first derivative of order 2, 4 or 6 (fd_subexpressions.hh:21-57)
// from u[0..6] = u(-3..3) at spacing 1/ih.
inline double fd_der_central(int order, const double u[7], double ih) {
    auto D = [&](int s) { return u[3 + s] - u[3 - s]; };
    if (order == 2) return (1.0 / 2.0) * ih * (D(1));
    if (order == 4) return (1.0 / 12.0) * ih * (+8.0 * D(1) - D(2));
    return (1.0 / 60.0) * ih * (+45.0 * D(1) - 9.0 * D(2) + D(3));
}

template <typename EOS>
struct Stepper {
    const Layout& L;
    const EOS& eos;
    AtmoParams atmo;
    ExcisionParams exc;
    C2PParams c2pp;
    Options opt;
    // scratch
    std::vector<double> flux[3], vbar[3], emf, ecen;
    // RK buffers
    std::vector<double> state_p, state_pp, B_p[3], B_pp[3];

    Stepper(const Layout& L_, const EOS& e) : L(L_), eos(e) {
        const Grid& G = L.grid;
        for (int d = 0; d < 3; ++d) {
            flux[d].assign(G.fc_size(d, NHRSC), 0.0);
            vbar[d].assign(G.fc_size(d, 4), 0.0);
            B_p[d].assign(G.fc_size(d, 1), 0.0);
            B_pp[d].assign(G.fc_size(d, 1), 0.0);
        }
        emf.assign(G.ed_size(), 0.0);
        ecen.assign(G.cc_size(3), 0.0);
        state_p.assign(G.cc_size(NSTATE), 0.0);
        state_pp.assign(G.cc_size(NSTATE), 0.0);
    }

    const Grid& G() const { return L.grid; }

    // spherical radius of the centre of cell (i,j,k) of block q, ghosts
    // included (get_physical_coordinates_sph, cartesian_coordinate_systems.hh:56-85)
    double cell_rtp0(int i, int j, int k, int64_t q) const {
        const Grid& g = G();
        const BlockDesc& b = L.blk[q];
        const Instance& I = L.inst[b.inst];
        const int loc[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
        double xyz[3];
        for (int d = 0; d < 3; ++d) xyz[d] = I.x0[d] + ((double)(b.b[d] * n[d] + loc[d] - g.ngz) + 0.5) / I.idx[d];
        return sph_radius(exc, xyz);
    }

    Metric metric_at(const double* st, int i, int j, int k, int64_t q) const {
        ORC_METRIC();
        const Grid& g = G();
        double gg[6], b[3];
        for (int a = 0; a < 6; ++a) gg[a] = st[g.cc(i, j, k, GXX_ + a, q, NSTATE)];
        for (int a = 0; a < 3; ++a) b[a] = st[g.cc(i, j, k, BETAX_ + a, q, NSTATE)];
        return Metric(gg, b, st[g.cc(i, j, k, ALP_, q, NSTATE)]);
    }

    // face_interp_4 (grmhd_helpers.hh:360-380) along d at face i-1/2
    double face4(const double* st, int i, int j, int k, int v, int64_t q, int d) const {
        const Grid& g = G();
        int di = d == 0, dj = d == 1, dk = d == 2;
        return face_interp_4(st[g.cc(i - 2 * di, j - 2 * dj, k - 2 * dk, v, q, NSTATE)],
                             st[g.cc(i - di, j - dj, k - dk, v, q, NSTATE)], st[g.cc(i, j, k, v, q, NSTATE)],
                             st[g.cc(i + di, j + dj, k + dk, v, q, NSTATE)]);
    }
    Metric face_metric(const double* st, int i, int j, int k, int64_t q, int d) const {
        ORC_METRIC();
        double gg[6], b[3];
        for (int a = 0; a < 6; ++a) gg[a] = face4(st, i, j, k, GXX_ + a, q, d);
        for (int a = 0; a < 3; ++a) b[a] = face4(st, i, j, k, BETAX_ + a, q, d);
        return Metric(gg, b, face4(st, i, j, k, ALP_, q, d));
    }

    // getflux<d> at face (i,j,k) = face i-1/2 of cell i (grmhd.hh:551-773).
    void getflux(int d, const Fields& old, const double* aux, int i, int j, int k, int64_t q) {
        const Grid& g = G();
        int di = d == 0, dj = d == 1, dk = d == 2;
        const double* st = old.state;
        Metric ml = metric_at(st, i - di, j - dj, k - dk, q);
        Metric mr = metric_at(st, i, j, k, q);
        Metric mf = face_metric(st, i, j, k, q, d);
        int th_aux = opt.thermo == TH_PRESS ? PRESS_ : TEMP_;
        int th_loc = opt.thermo == TH_PRESS ? PRESSL : TEMPL;
        const int ridx[10] = {RHO_, ZVECX_, ZVECY_, ZVECZ_, YE_, th_aux, ENTROPY_, BX_, BY_, BZ_};
        const int rloc[10] = {RHOL, ZXL, ZYL, ZZL, YEL, th_loc, ENTL, BXL, BYL, BZL};
        double pL[NPRIM] = {0}, pR[NPRIM] = {0};
        for (int v = 0; v < 10; ++v) {
            double u[6];
            for (int s = 0; s < 6; ++s)
                u[s] = aux[g.cc(i + (s - 3) * di, j + (s - 3) * dj, k + (s - 3) * dk, ridx[v], q, NAUX)];
            recon_any(opt.recon, u, pL[rloc[v]], pR[rloc[v]]);
        }
        if (ml.alp <= exc.alp_ex || mr.alp <= exc.alp_ex) {
            for (int v = 0; v < 10; ++v) {
                pL[rloc[v]] = aux[g.cc(i - di, j - dj, k - dk, ridx[v], q, NAUX)];
                pR[rloc[v]] = aux[g.cc(i, j, k, ridx[v], q, NAUX)];
            }
        }
        double Bn = old.B[d][g.fc(d, i, j, k, 0, q, 1)] / mf.sqrtg;
        pL[BXL + d] = pR[BXL + d] = Bn;
        Bits ee;
        if (opt.thermo == TH_PRESS) {
            double hL, hR, sL, sR;
            pL[EPSL] = eos.eps_h_csnd2_temp_entropy__press_rho_ye(hL, pL[CS2L], pL[TEMPL], sL, pL[PRESSL], pL[RHOL], ee);
            pR[EPSL] = eos.eps_h_csnd2_temp_entropy__press_rho_ye(hR, pR[CS2L], pR[TEMPL], sR, pR[PRESSL], pR[RHOL], ee);
        } else {
            pL[PRESSL] = eos.press_eps_csnd2__temp_rho_ye(pL[EPSL], pL[CS2L], pL[TEMPL], pL[RHOL], ee);
            pR[PRESSL] = eos.press_eps_csnd2__temp_rho_ye(pR[EPSL], pR[CS2L], pR[TEMPL], pR[RHOL], ee);
        }
        double f[NCONS], vb[4];
        mhd_fluxes(d, pL, pR, mf, f, vb);
        double* F = flux[d].data();
        F[g.fc(d, i, j, k, DENS_, q, NHRSC)] = f[DENSL];
        F[g.fc(d, i, j, k, YESTAR_, q, NHRSC)] = f[YESL];
        F[g.fc(d, i, j, k, ENTROPYSTAR_, q, NHRSC)] = f[ENTSL];
        F[g.fc(d, i, j, k, TAU_, q, NHRSC)] = f[TAUL];
        F[g.fc(d, i, j, k, SX_, q, NHRSC)] = f[STXL];
        F[g.fc(d, i, j, k, SY_, q, NHRSC)] = f[STYL];
        F[g.fc(d, i, j, k, SZ_, q, NHRSC)] = f[STZL];
        double* V = vbar[d].data();
        int nvb = opt.emf == EMF_GS ? 2 : 4;
        for (int c = 0; c < nvb; ++c) V[g.fc(d, i, j, k, c, q, 4)] = vb[c];
    }

    // compute_mhd_fluxes<d> (grmhd.hh:774-970)
    void mhd_fluxes(int d, double pL[NPRIM], double pR[NPRIM], const Metric& mf, double f[NCONS], double vb[4]) {
        const double *g = mf.g, *gu = mf.gi, *beta = mf.beta;
        double alp = mf.alp, sqrtg = mf.sqrtg;
        const double *zl = &pL[ZXL], *zr = &pR[ZXL], *Bl = &pL[BXL], *Br = &pR[BXL];
        double wl = get_W(g, zl), wr = get_W(g, zr);
        double bl[4], br[4], b2l, b2r;
        get_smallb(alp, beta, g, Bl, zl, wl, bl, b2l);
        get_smallb(alp, beta, g, Br, zr, wr, br, b2r);
        double vtl[3], vtr[3];
        get_vtilde(alp, beta, zl, wl, vtl);
        get_vtilde(alp, beta, zr, wr, vtr);
        const int mc[3] = {0, 3, 5};
        double cml, cpl, cmr, cpr, cmin, cmax;
        get_cm_cp(alp, beta, pL[RHOL], pL[PRESSL], pL[EPSL], pL[CS2L], wl, b2l, vtl, gu[mc[d]], d, cml, cpl);
        get_cm_cp(alp, beta, pR[RHOL], pR[PRESSL], pR[EPSL], pR[CS2L], wr, b2r, vtr, gu[mc[d]], d, cmr, cpr);
        if (opt.riemann == RS_LLF) {
            cmax = std::max(std::max(std::fabs(cml), std::fabs(cmr)), std::max(std::fabs(cpl), std::fabs(cpr)));
            if (std::min(pL[RHOL], pR[RHOL]) < opt.rusanov_c_limit) {
                double cg = alp * std::sqrt(gu[mc[d]]) + std::fabs(beta[d]);
                cmax = std::max(cmax, cg);
            }
            cmin = cmax;
        } else {
            cmin = -std::min(0., std::min(cml, cmr));
            cmax = std::max(0., std::max(cpl, cpr));
        }
        if (cmin < 1e-12 && cmax < 1e-12) { cmin = 1; cmax = 1; }
        if (opt.emf == EMF_UCT) {
            const int jk[3][2] = {{1, 2}, {0, 2}, {0, 1}};
            vb[0] = hll(vtl[jk[d][0]], vtr[jk[d][0]], 0, 0, cmin, cmax);
            vb[1] = hll(vtl[jk[d][1]], vtr[jk[d][1]], 0, 0, cmin, cmax);
            vb[2] = cmin; vb[3] = cmax;
        }
        double Dl, taul, Sl[3], Esl, fDl, ftl, fSl[3], fEl;
        double Dr, taur, Sr[3], Esr, fDr, ftr, fSr[3], fEr;
        get_fluxes(alp, beta, g, pL[RHOL], pL[PRESSL], pL[EPSL], zl, pL[ENTL], wl, b2l, bl, vtl, d,
                   Dl, taul, Sl, Esl, fDl, ftl, fSl, fEl);
        get_fluxes(alp, beta, g, pR[RHOL], pR[PRESSL], pR[EPSL], zr, pR[ENTL], wr, b2r, br, vtr, d,
                   Dr, taur, Sr, Esr, fDr, ftr, fSr, fEr);
        f[DENSL] = sqrtg * hll(fDl, fDr, Dl, Dr, cmin, cmax);
        f[ENTSL] = sqrtg * hll(fEl, fEr, Esl, Esr, cmin, cmax);
        f[TAUL] = sqrtg * hll(ftl, ftr, taul, taur, cmin, cmax);
        for (int a = 0; a < 3; ++a) f[STXL + a] = sqrtg * hll(fSl[a], fSr[a], Sl[a], Sr[a], cmin, cmax);
        double yel = pL[YEL], yer = pR[YEL];
        f[YESL] = sqrtg * hll(yel * fDl, yer * fDr, yel * Dl, yer * Dr, cmin, cmax);
        if (opt.emf == EMF_GS) {
            if (d == 0) {
                double Fbzl = -Bl[0] * vtl[2] + Bl[2] * vtl[0], Fbzr = -Br[0] * vtr[2] + Br[2] * vtr[0];
                vb[0] = sqrtg * hll(Fbzl, Fbzr, Bl[2], Br[2], cmin, cmax);
                double Fbyl = -Bl[0] * vtl[1] + Bl[1] * vtl[0], Fbyr = -Br[0] * vtr[1] + Br[1] * vtr[0];
                vb[1] = -sqrtg * hll(Fbyl, Fbyr, Bl[1], Br[1], cmin, cmax);
            } else if (d == 1) {
                double Fbzl = -Bl[1] * vtl[2] + Bl[2] * vtl[1], Fbzr = -Br[1] * vtr[2] + Br[2] * vtr[1];
                vb[0] = -sqrtg * hll(Fbzl, Fbzr, Bl[2], Br[2], cmin, cmax);
                double Fbxl = Bl[0] * vtl[1] - Bl[1] * vtl[0], Fbxr = Br[0] * vtr[1] - Br[1] * vtr[0];
                vb[1] = sqrtg * hll(Fbxl, Fbxr, Bl[0], Br[0], cmin, cmax);
            } else {
                double Fbyl = Bl[1] * vtl[2] - Bl[2] * vtl[1], Fbyr = Br[1] * vtr[2] - Br[2] * vtr[1];
                vb[0] = sqrtg * hll(Fbyl, Fbyr, Bl[1], Br[1], cmin, cmax);
                double Fbxl = Bl[0] * vtl[2] - Bl[2] * vtl[0], Fbxr = Br[0] * vtr[2] - Br[2] * vtr[0];
                vb[1] = -sqrtg * hll(Fbxl, Fbxr, Bl[0], Br[0], cmin, cmax);
            }
        }
    }

    void compute_fluxes(const Fields& old, const double* aux) {
        ORC_PASS(1);
        const Grid& g = G();
        for (int64_t q = 0; q < g.nq; ++q)
            for (int d = 0; d < 3; ++d) {
                int lo[3] = {0, 0, 0}, hi[3] = {g.Ni, g.Nj, g.Nk};
                int n[3] = {g.nx, g.ny, g.nz};
                // FOFC needs one more face on each side along the normal (evolve.cpp:895-920)
                lo[d] = g.ngz - (opt.fofc ? 1 : 0); hi[d] = n[d] + g.ngz + 1 + (opt.fofc ? 1 : 0);
#pragma omp parallel for collapse(2) schedule(static)
                for (int k = lo[2]; k < hi[2]; ++k)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int i = lo[0]; i < hi[0]; ++i) getflux(d, old, aux, i, j, k, q);
            }
    }

    // UCT edge EMFs (evolve.cpp:1130-1340)
    double rec_face(int kind, const double* arr, int fd, int nv, int v, int i, int j, int k, int64_t q, int rd) const {
        const Grid& g = G();
        int di = rd == 0, dj = rd == 1, dk = rd == 2;
        double u[6];
        for (int s = 0; s < 6; ++s) u[s] = arr[g.fc(fd, i + (s - 3) * di, j + (s - 3) * dj, k + (s - 3) * dk, v, q, nv)];
        double uL, uR;
        recon_any(kind, u, uL, uR);
        (void)uL;
        return 0;  // unused
    }
    void rec2(const double* arr, int fd, int nv, int v, int i, int j, int k, int64_t q, int rd, double& uL, double& uR) const {
        const Grid& g = G();
        int di = rd == 0, dj = rd == 1, dk = rd == 2;
        double u[6];
        for (int s = 0; s < 6; ++s) u[s] = arr[g.fc(fd, i + (s - 3) * di, j + (s - 3) * dj, k + (s - 3) * dk, v, q, nv)];
        recon_any(opt.recon, u, uL, uR);
    }
    double vb(int fd, int c, int i, int j, int k, int64_t q) const { return vbar[fd][G().fc(fd, i, j, k, c, q, 4)]; }

    void compute_emfs_uct(const Fields& old) {
        ORC_PASS(2);
        const Grid& g = G();
        int ng = g.ngz;
        for (int64_t q = 0; q < g.nq; ++q) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = ng; k < g.nz + ng + 1; ++k)
                for (int j = ng; j < g.ny + ng + 1; ++j)
                    for (int i = ng; i < g.nx + ng + 1; ++i) {
                        if (i < g.nx + ng) {  // E^x at (i, j-1/2, k-1/2)
                            double ByL, ByR, vzL, vzR, BzL, BzR, vyL, vyR;
                            rec2(old.B[1], 1, 1, 0, i, j, k, q, 2, ByL, ByR);
                            rec2(vbar[1].data(), 1, 4, 1, i, j, k, q, 2, vzL, vzR);
                            rec2(old.B[2], 2, 1, 0, i, j, k, q, 1, BzL, BzR);
                            rec2(vbar[2].data(), 2, 4, 1, i, j, k, q, 1, vyL, vyR);
                            double cmin_y = std::max(vb(1, 2, i, j, k, q), vb(1, 2, i, j, k - 1, q));
                            double cmax_y = std::max(vb(1, 3, i, j, k, q), vb(1, 3, i, j, k - 1, q));
                            double cmin_z = std::max(vb(2, 2, i, j, k, q), vb(2, 2, i, j - 1, k, q));
                            double cmax_z = std::max(vb(2, 3, i, j, k, q), vb(2, 3, i, j - 1, k, q));
                            emf[g.ed(i, j, k, 0, q)] = hll(vzL * ByL, vzR * ByR, ByL, ByR, cmin_z, cmax_z)
                                                       - hll(vyL * BzL, vyR * BzR, BzL, BzR, cmin_y, cmax_y);
                        }
                        if (j < g.ny + ng) {  // E^y at (i-1/2, j, k-1/2)
                            double BzL, BzR, vxL, vxR, BxL, BxR, vzL, vzR;
                            rec2(old.B[2], 2, 1, 0, i, j, k, q, 0, BzL, BzR);
                            rec2(vbar[2].data(), 2, 4, 0, i, j, k, q, 0, vxL, vxR);
                            rec2(old.B[0], 0, 1, 0, i, j, k, q, 2, BxL, BxR);
                            rec2(vbar[0].data(), 0, 4, 1, i, j, k, q, 2, vzL, vzR);
                            double cmin_z = std::max(vb(2, 2, i, j, k, q), vb(2, 2, i - 1, j, k, q));
                            double cmax_z = std::max(vb(2, 3, i, j, k, q), vb(2, 3, i - 1, j, k, q));
                            double cmin_x = std::max(vb(0, 2, i, j, k, q), vb(0, 2, i, j, k - 1, q));
                            double cmax_x = std::max(vb(0, 3, i, j, k, q), vb(0, 3, i, j, k - 1, q));
                            emf[g.ed(i, j, k, 1, q)] = hll(vxL * BzL, vxR * BzR, BzL, BzR, cmin_x, cmax_x)
                                                       - hll(vzL * BxL, vzR * BxR, BxL, BxR, cmin_z, cmax_z);
                        }
                        if (k < g.nz + ng) {  // E^z at (i-1/2, j-1/2, k)
                            double BxL, BxR, vyL, vyR, ByL, ByR, vxL, vxR;
                            rec2(old.B[0], 0, 1, 0, i, j, k, q, 1, BxL, BxR);
                            rec2(vbar[0].data(), 0, 4, 0, i, j, k, q, 1, vyL, vyR);
                            rec2(old.B[1], 1, 1, 0, i, j, k, q, 0, ByL, ByR);
                            rec2(vbar[1].data(), 1, 4, 0, i, j, k, q, 0, vxL, vxR);
                            double cmin_x = std::max(vb(0, 2, i, j, k, q), vb(0, 2, i, j - 1, k, q));
                            double cmax_x = std::max(vb(0, 3, i, j, k, q), vb(0, 3, i, j - 1, k, q));
                            double cmin_y = std::max(vb(1, 2, i, j, k, q), vb(1, 2, i - 1, j, k, q));
                            double cmax_y = std::max(vb(1, 3, i, j, k, q), vb(1, 3, i - 1, j, k, q));
                            emf[g.ed(i, j, k, 2, q)] = hll(vyL * BxL, vyR * BxR, BxL, BxR, cmin_y, cmax_y)
                                                       - hll(vxL * ByL, vxR * ByR, ByL, ByR, cmin_x, cmax_x);
                        }
                    }
        }
    }

    // GS edge EMFs (grmhd_helpers.hh:492-627, evolve.cpp:1069-1129)
    double Ef(int fd, int c, int i, int j, int k, int64_t q) const { return vbar[fd][G().fc(fd, i, j, k, c, q, 4)]; }
    double Ec(int c, int i, int j, int k, int64_t q) const { return ecen[G().cc(i, j, k, c, q, 3)]; }
    double Fd(int fd, int i, int j, int k, int64_t q) const { return flux[fd][G().fc(fd, i, j, k, DENS_, q, NHRSC)]; }

    // Gardiner-Stone edge EMF of component c at edge (i,j,k) (grmhd_helpers.hh:492-627)
    double gs_edge(int c, int i, int j, int k, int64_t q) const {
        if (c == 0) {
            double Exzf = Ef(2, 0, i, j, k, q), Exzfm = Ef(2, 0, i, j - 1, k, q);
            double Exyf = Ef(1, 0, i, j, k, q), Exyfm = Ef(1, 0, i, j, k - 1, q);
            double NE = Ec(0, i, j, k, q), NW = Ec(0, i, j - 1, k, q);
            double SE = Ec(0, i, j, k - 1, q), SW = Ec(0, i, j - 1, k - 1, q);
            double Eavg = 0.25 * (Exzf + Exzfm + Exyf + Exyfm);
            double Sy = std::copysign(1., Fd(1, i, j, k, q)), Sym = std::copysign(1., Fd(1, i, j, k - 1, q));
            double dN = (1. - Sy) * (NE - Exzf) + (1. + Sy) * (NW - Exzfm);
            double dS = (1. - Sym) * (Exzf - SE) + (1. + Sym) * (Exzfm - SW);
            double dEdz = (1. / 8.) * (dS - dN);
            double Sz = std::copysign(1., Fd(2, i, j, k, q)), Szm = std::copysign(1., Fd(2, i, j - 1, k, q));
            double dE = (1. - Sz) * (NE - Exyf) + (1. + Sz) * (SE - Exyfm);
            double dW = (1. - Szm) * (Exyf - NW) + (1. + Szm) * (Exyfm - SW);
            double dEdy = (1. / 8.) * (dW - dE);
            return Eavg + dEdz + dEdy;
        }
        if (c == 1) {
            double Eyzf = Ef(2, 1, i, j, k, q), Eyzfm = Ef(2, 1, i - 1, j, k, q);
            double Eyxf = Ef(0, 0, i, j, k, q), Eyxfm = Ef(0, 0, i, j, k - 1, q);
            double NE = Ec(1, i, j, k, q), NW = Ec(1, i - 1, j, k, q);
            double SE = Ec(1, i, j, k - 1, q), SW = Ec(1, i - 1, j, k - 1, q);
            double Eavg = 0.25 * (Eyzf + Eyzfm + Eyxf + Eyxfm);
            double Sx = std::copysign(1., Fd(0, i, j, k, q)), Sxm = std::copysign(1., Fd(0, i, j, k - 1, q));
            double dN = (1. - Sx) * (NE - Eyzf) + (1. + Sx) * (NW - Eyzfm);
            double dS = (1. - Sxm) * (Eyzf - SE) + (1. + Sxm) * (Eyzfm - SW);
            double dEdz = (1. / 8.) * (dS - dN);
            double Sz = std::copysign(1., Fd(2, i, j, k, q)), Szm = std::copysign(1., Fd(2, i - 1, j, k, q));
            double dE = (1. - Sz) * (NE - Eyxf) + (1. + Sz) * (SE - Eyxfm);
            double dW = (1. - Szm) * (Eyxf - NW) + (1. + Szm) * (Eyxfm - SW);
            double dEdx = (1. / 8.) * (dW - dE);
            return Eavg + dEdz + dEdx;
        }
        double Ezyf = Ef(1, 1, i, j, k, q), Ezyfm = Ef(1, 1, i - 1, j, k, q);
        double Ezxf = Ef(0, 1, i, j, k, q), Ezxfm = Ef(0, 1, i, j - 1, k, q);
        double NE = Ec(2, i, j, k, q), NW = Ec(2, i - 1, j, k, q);
        double SE = Ec(2, i, j - 1, k, q), SW = Ec(2, i - 1, j - 1, k, q);
        double Eavg = 0.25 * (Ezyf + Ezyfm + Ezxf + Ezxfm);
        double Sx = std::copysign(1., Fd(0, i, j, k, q)), Sxm = std::copysign(1., Fd(0, i, j - 1, k, q));
        double dN = (1. - Sx) * (NE - Ezyf) + (1. + Sx) * (NW - Ezyfm);
        double dS = (1. - Sxm) * (Ezyf - SE) + (1. + Sxm) * (Ezyfm - SW);
        double dEdy = (1. / 8.) * (dS - dN);
        double Sy = std::copysign(1., Fd(1, i, j, k, q)), Sym = std::copysign(1., Fd(1, i - 1, j, k, q));
        double dE = (1. - Sy) * (NE - Ezxf) + (1. + Sy) * (SE - Ezxfm);
        double dW = (1. - Sym) * (Ezxf - NW) + (1. + Sym) * (Ezxfm - SW);
        double dEdx = (1. / 8.) * (dW - dE);
        return Eavg + dEdy + dEdx;
    }

    // edge range [ngz, n+ngz] (evolve.cpp:1090-1104); one layer wider with
    // FOFC and fofc_ghost_emf
    void compute_emfs_gs() {
        ORC_PASS(2);
        const Grid& g = G();
        int ng = g.ngz;
        const int w = (opt.fofc && opt.fofc_ghost_emf) ? 1 : 0;
        for (int64_t q = 0; q < g.nq; ++q) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = ng - w; k < g.nz + ng + 1 + w; ++k)
                for (int j = ng - w; j < g.ny + ng + 1 + w; ++j)
                    for (int i = ng - w; i < g.nx + ng + 1 + w; ++i) {
                        if (i < g.nx + ng + w) emf[g.ed(i, j, k, 0, q)] = gs_edge(0, i, j, k, q);
                        if (j < g.ny + ng + w) emf[g.ed(i, j, k, 1, q)] = gs_edge(1, i, j, k, q);
                        if (k < g.nz + ng + w) emf[g.ed(i, j, k, 2, q)] = gs_edge(2, i, j, k, q);
                    }
        }
    }

    // Centered FD first derivative along d (fd_subexpressions.hh:21-57)
    double fd_der(const double* st, int i, int j, int k, int v, int64_t q, int d, double ih) const {
        ORC_METRIC();
        const Grid& g = G();
        int di = d == 0, dj = d == 1, dk = d == 2;
        double u[7];
        for (int s = -3; s <= 3; ++s) {
            const int w = (opt.fd_order / 2 >= (s < 0 ? -s : s)) ? s : 0;   // read only the stencil's points
            u[s + 3] = st[g.cc(i + w * di, j + w * dj, k + w * dk, v, q, NSTATE)];
        }
        return fd_der_central(opt.fd_order, u, ih);
    }

    // compute_source_terms (grmhd.hh:190-297) + divergence (evolve.cpp:1395-1401)
    void add_fluxes_and_sources(double dt, double dtfact, double* newst, const Fields& old, const double* aux) {
        ORC_PASS(3);
        const Grid& g = G();
        int ng = g.ngz;
        for (int64_t q = 0; q < g.nq; ++q) {
            double ix = L.idx(q, 0), iy = L.idx(q, 1), iz = L.idx(q, 2);
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = ng; k < g.nz + ng; ++k)
                for (int j = ng; j < g.ny + ng; ++j)
                    for (int i = ng; i < g.nx + ng; ++i) {
                        const double* st = old.state;
                        Metric m = metric_at(st, i, j, k, q);
                        {
                        ORC_SUB(8);
                        double z[3], B[3];
                        for (int a = 0; a < 3; ++a) {
                            z[a] = aux[g.cc(i, j, k, ZVECX_ + a, q, NAUX)];
                            B[a] = aux[g.cc(i, j, k, BX_ + a, q, NAUX)];
                        }
                        double W = get_W(m.g, z);
                        double da[3], dgd[18], dbd[9], K[6];
                        for (int d = 0; d < 3; ++d) da[d] = fd_der(st, i, j, k, ALP_, q, d, ix);
                        for (int d = 0; d < 3; ++d)
                            for (int c = 0; c < 3; ++c) dbd[c + 3 * d] = fd_der(st, i, j, k, BETAX_ + c, q, d, ix);
                        for (int d = 0; d < 3; ++d)
                            for (int a = 0; a < 6; ++a) dgd[a + 6 * d] = fd_der(st, i, j, k, GXX_ + a, q, d, ix);
                        for (int a = 0; a < 6; ++a) K[a] = st[g.cc(i, j, k, KXX_ + a, q, NSTATE)];
                        double stau, ss[3];
                        get_geom_sources(m.alp, m.beta, m.g, m.gi, K, da, dgd, dbd,
                                         aux[g.cc(i, j, k, RHO_, q, NAUX)], aux[g.cc(i, j, k, PRESS_, q, NAUX)],
                                         aux[g.cc(i, j, k, EPS_, q, NAUX)], B, z, W, stau, ss);
                        newst[g.cc(i, j, k, TAU_, q, NSTATE)] += m.sqrtg * dt * dtfact * stau;
                        newst[g.cc(i, j, k, SX_, q, NSTATE)] += m.sqrtg * dt * dtfact * ss[0];
                        newst[g.cc(i, j, k, SY_, q, NSTATE)] += m.sqrtg * dt * dtfact * ss[1];
                        newst[g.cc(i, j, k, SZ_, q, NSTATE)] += m.sqrtg * dt * dtfact * ss[2];
                        }
                        for (int v = 0; v < NHRSC; ++v) {
                            double dFx = (flux[0][g.fc(0, i, j, k, v, q, NHRSC)] - flux[0][g.fc(0, i + 1, j, k, v, q, NHRSC)]) * ix;
                            double dFy = (flux[1][g.fc(1, i, j, k, v, q, NHRSC)] - flux[1][g.fc(1, i, j + 1, k, v, q, NHRSC)]) * iy;
                            double dFz = (flux[2][g.fc(2, i, j, k, v, q, NHRSC)] - flux[2][g.fc(2, i, j, k + 1, v, q, NHRSC)]) * iz;
                            newst[g.cc(i, j, k, v, q, NSTATE)] += dt * dtfact * ((dFx + dFy) + dFz);
                        }
                    }
        }
    }

    // update_CT (evolve.cpp:1406-1479)
    void update_ct(double dt, double dtfact, double* nB[3]) {
        ORC_PASS(4);
        const Grid& g = G();
        int ng = g.ngz;
        for (int64_t q = 0; q < g.nq; ++q) {
            double ix = L.idx(q, 0), iy = L.idx(q, 1), iz = L.idx(q, 2);
            auto E = [&](int i, int j, int k, int c) { return emf[g.ed(i, j, k, c, q)]; };
            for (int k = ng; k < g.nz + ng; ++k)
                for (int j = ng; j < g.ny + ng; ++j)
                    for (int i = ng; i < g.nx + ng + 1; ++i)
                        nB[0][g.fc(0, i, j, k, 0, q, 1)] += dt * dtfact * ((E(i, j, k + 1, 1) - E(i, j, k, 1)) * iz + (E(i, j, k, 2) - E(i, j + 1, k, 2)) * iy);
            for (int k = ng; k < g.nz + ng; ++k)
                for (int j = ng; j < g.ny + ng + 1; ++j)
                    for (int i = ng; i < g.nx + ng; ++i)
                        nB[1][g.fc(1, i, j, k, 0, q, 1)] += dt * dtfact * ((E(i + 1, j, k, 2) - E(i, j, k, 2)) * ix + (E(i, j, k, 0) - E(i, j, k + 1, 0)) * iz);
            for (int k = ng; k < g.nz + ng + 1; ++k)
                for (int j = ng; j < g.ny + ng; ++j)
                    for (int i = ng; i < g.nx + ng; ++i)
                        nB[2][g.fc(2, i, j, k, 0, q, 1)] += dt * dtfact * ((E(i, j + 1, k, 0) - E(i, j, k, 0)) * iy + (E(i, j, k, 1) - E(i + 1, j, k, 1)) * ix);
        }
    }

    // ---------------- FOFC (evolve.cpp:456-811) ----------------
    // Per-block byte tags of faces and edges, cleared every substep.
    std::vector<uint8_t> ftag[3], etag;
    int64_t fofc_cells_flagged = 0;

    // flag_fofc_cells: tentative update of every cell of [ngz-1, n+ngz+1)^3
    // (interior widened by one ghost layer) with the high-order fluxes and
    // EMFs, a discrete-maximum-principle test on D and tau against the
    // 27-cell neighbourhood of `newst`, and a lenient dry-run c2p; a cell
    // that fails tags its 6 faces and 12 edges.
    void flag_fofc(double dt, double dtfact, const double* newst, const Fields& old) {
        ORC_PASS(7);
        const Grid& g = G();
        const int ng = g.ngz;
        for (int d = 0; d < 3; ++d) ftag[d].assign((size_t)g.fc_size(d, 1), 0);
        etag.assign((size_t)g.ed_size(), 0);
        C2PParams lenient = c2pp;
        lenient.alp_bh_thresh = 1e30;
        const double dte = dt * dtfact;
        int64_t nflag = 0;
        for (int64_t q = 0; q < g.nq; ++q) {
            const double ix = L.idx(q, 0), iy = L.idx(q, 1), iz = L.idx(q, 2);
            auto E = [&](int i, int j, int k, int c) { return emf[g.ed(i, j, k, c, q)]; };
            auto Bo = [&](int d, int i, int j, int k) { return old.B[d][g.fc(d, i, j, k, 0, q, 1)]; };
#pragma omp parallel for collapse(2) schedule(dynamic) reduction(+ : nflag)
            for (int k = ng - 1; k < g.nz + ng + 1; ++k)
                for (int j = ng - 1; j < g.ny + ng + 1; ++j)
                    for (int i = ng - 1; i < g.nx + ng + 1; ++i) {
                        double cons[NCONS];
                        for (int v = 0; v <= ENTROPYSTAR_; ++v) {
                            double dFx = (flux[0][g.fc(0, i, j, k, v, q, NHRSC)] - flux[0][g.fc(0, i + 1, j, k, v, q, NHRSC)]) * ix;
                            double dFy = (flux[1][g.fc(1, i, j, k, v, q, NHRSC)] - flux[1][g.fc(1, i, j + 1, k, v, q, NHRSC)]) * iy;
                            double dFz = (flux[2][g.fc(2, i, j, k, v, q, NHRSC)] - flux[2][g.fc(2, i, j, k + 1, v, q, NHRSC)]) * iz;
                            cons[v] = newst[g.cc(i, j, k, v, q, NSTATE)] + dt * dtfact * ((dFx + dFy) + dFz);
                        }
                        // ct_update_B{x,y,z}_face (grmhd_helpers.hh:441-475) on the stage input B
                        auto bx = [&](int ii) { return Bo(0, ii, j, k) + dte * ((E(ii, j, k + 1, 1) - E(ii, j, k, 1)) * iz + (E(ii, j, k, 2) - E(ii, j + 1, k, 2)) * iy); };
                        auto by = [&](int jj) { return Bo(1, i, jj, k) + dte * ((E(i + 1, jj, k, 2) - E(i, jj, k, 2)) * ix + (E(i, jj, k, 0) - E(i, jj, k + 1, 0)) * iz); };
                        auto bz = [&](int kk) { return Bo(2, i, j, kk) + dte * ((E(i, j + 1, kk, 0) - E(i, j, kk, 0)) * iy + (E(i, j, kk, 1) - E(i + 1, j, kk, 1)) * ix); };
                        cons[BSXL] = 0.5 * (bx(i) + bx(i + 1));
                        cons[BSYL] = 0.5 * (by(j) + by(j + 1));
                        cons[BSZL] = 0.5 * (bz(k) + bz(k + 1));
                        bool dmp = false;
                        if (opt.dmp_enable) {
                            const double inv_M = 1.0 / opt.dmp_M;
                            const int vars[2] = {DENS_, TAU_};
                            for (int n = 0; n < 2 && !dmp; ++n) {
                                double vmax = -std::numeric_limits<double>::max(), vmin = std::numeric_limits<double>::max();
                                for (int kt = -1; kt <= 1; ++kt)
                                    for (int jt = -1; jt <= 1; ++jt)
                                        for (int it = -1; it <= 1; ++it) {
                                            double v = newst[g.cc(i + it, j + jt, k + kt, vars[n], q, NSTATE)];
                                            vmax = std::fmax(vmax, v);
                                            vmin = std::fmin(vmin, v);
                                        }
                                double test = cons[vars[n]];
                                if (test > opt.dmp_M * vmax || test < vmin * inv_M) dmp = true;
                            }
                        }
                        Metric m = metric_at(old.state, i, j, k, q);
                        double prims[NPRIM] = {0};
                        Bits err;
                        double rtp[3] = {cell_rtp0(i, j, k, q), 0.0, 0.0};
                        C2PResult r = conservs_to_prims(cons, prims, m, eos, atmo, exc, lenient, rtp, err, true);
                        if (!(r.floored || dmp)) continue;
                        nflag++;
                        for (int t = 0; t < 2; ++t) {
                            ftag[0][g.fc(0, i + t, j, k, 0, q, 1)] = 1;
                            ftag[1][g.fc(1, i, j + t, k, 0, q, 1)] = 1;
                            ftag[2][g.fc(2, i, j, k + t, 0, q, 1)] = 1;
                        }
                        for (int a = 0; a < 2; ++a)
                            for (int b = 0; b < 2; ++b) {
                                etag[g.ed(i, j + a, k + b, 0, q)] = 1;
                                etag[g.ed(i + a, j, k + b, 1, q)] = 1;
                                etag[g.ed(i + a, j + b, k, 2, q)] = 1;
                            }
                    }
        }
        fofc_cells_flagged += nflag;
    }

    // apply_fofc_correction: donor-cell + LLF fluxes (and GS face EMFs) on the
    // tagged faces, then the GS edge EMFs on the tagged edges (the compaction
    // ranges of evolve.cpp:646-702).
    void apply_fofc(const Fields& old, const double* aux) {
        ORC_PASS(7);
        const Grid& g = G();
        const int ng = g.ngz;
        const int n[3] = {g.nx, g.ny, g.nz};
        Options keep = opt;
        opt.recon = REC_DONOR;
        opt.riemann = RS_LLF;
        for (int64_t q = 0; q < g.nq; ++q)
            for (int d = 0; d < 3; ++d) {
                int lo[3] = {ng - 1, ng - 1, ng - 1}, hi[3] = {n[0] + ng + 1, n[1] + ng + 1, n[2] + ng + 1};
                lo[d] = ng;
                for (int k = lo[2]; k < hi[2]; ++k)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int i = lo[0]; i < hi[0]; ++i)
                            if (ftag[d][g.fc(d, i, j, k, 0, q, 1)]) getflux(d, old, aux, i, j, k, q);
            }
        opt = keep;
        for (int64_t q = 0; q < g.nq; ++q)
            for (int k = ng; k < g.nz + ng + 1; ++k)
                for (int j = ng; j < g.ny + ng + 1; ++j)
                    for (int i = ng; i < g.nx + ng + 1; ++i) {
                        if (i < g.nx + ng && etag[g.ed(i, j, k, 0, q)]) emf[g.ed(i, j, k, 0, q)] = gs_edge(0, i, j, k, q);
                        if (j < g.ny + ng && etag[g.ed(i, j, k, 1, q)]) emf[g.ed(i, j, k, 1, q)] = gs_edge(1, i, j, k, q);
                        if (k < g.nz + ng && etag[g.ed(i, j, k, 2, q)]) emf[g.ed(i, j, k, 2, q)] = gs_edge(2, i, j, k, q);
                    }
    }

    // advance_substep (evolve.cpp:1671-1717); reflux on a two-level forest
    // (refluxing.cpp), a no-op on uniform grids
    void advance_substep(double dt, double dtfact, double* newst, double* nB[3], const Fields& old, const double* aux) {
        compute_fluxes(old, aux);
        if (opt.emf == EMF_GS) compute_emfs_gs();
        else compute_emfs_uct(old);
        if (opt.fofc) {
            flag_fofc(dt, dtfact, newst, old);
            apply_fofc(old, aux);
        }
        if (L.amr) {
            double* fl[3] = {flux[0].data(), flux[1].data(), flux[2].data()};
            L.amr->reflux_fluxes(fl, NHRSC);
            L.amr->reflux_emfs(emf.data());
        }
        add_fluxes_and_sources(dt, dtfact, newst, old, aux);
        update_ct(dt, dtfact, nB);
    }

    // ---------------- ghost fill (same-level copy + physical BC) ----------------
    // Global index along an axis of an instance -> (block coord, local index).
    // Cell-centred: g in [0, N) interior; outflow clamps, periodic wraps.
    static int map_cc(int g, int N, int bclo, int bchi) {
        if (g < 0) return bclo == BC_PERIODIC ? ((g % N) + N) % N : 0;
        if (g >= N) return bchi == BC_PERIODIC ? g % N : N - 1;
        return g;
    }
    // Face-centred along its normal: faces [0, N] interior.
    static int map_fc(int f, int N, int bclo, int bchi) {
        if (f < 0) return bclo == BC_PERIODIC ? ((f % N) + N) % N : 0;
        if (f > N) return bchi == BC_PERIODIC ? f % N : N;
        return f;
    }
    void fill_ghosts(double* st, double* Bf[3], int nvar_first, int nvar_count) {
        if (L.amr) { L.amr->fill(st, Bf, nvar_first, nvar_count); return; }
        const Grid& g = G();
        int n[3] = {g.nx, g.ny, g.nz};
        int Ne[3] = {g.Ni, g.Nj, g.Nk};
        int ng = g.ngz;
        for (int64_t q = 0; q < g.nq; ++q) {
            const BlockDesc& bd = L.blk[q];
            const Instance& I = L.inst[bd.inst];
            int Ntot[3] = {I.nb[0] * n[0], I.nb[1] * n[1], I.nb[2] * n[2]};
            // cell-centred
            for (int k = 0; k < Ne[2]; ++k)
                for (int j = 0; j < Ne[1]; ++j)
                    for (int i = 0; i < Ne[0]; ++i) {
                        int loc[3] = {i, j, k};
                        bool ghost = false;
                        for (int d = 0; d < 3; ++d) ghost |= (loc[d] < ng || loc[d] >= n[d] + ng);
                        if (!ghost) continue;
                        int sb[3], sl[3];
                        for (int d = 0; d < 3; ++d) {
                            int gg = bd.b[d] * n[d] + loc[d] - ng;
                            int m = map_cc(gg, Ntot[d], I.bc[2 * d], I.bc[2 * d + 1]);
                            sb[d] = m / n[d]; sl[d] = m % n[d] + ng;
                        }
                        int64_t qs = I.q0 + sb[0] + (int64_t)I.nb[0] * (sb[1] + (int64_t)I.nb[1] * sb[2]);
                        for (int v = nvar_first; v < nvar_first + nvar_count; ++v)
                            st[g.cc(i, j, k, v, q, NSTATE)] = st[g.cc(sl[0], sl[1], sl[2], v, qs, NSTATE)];
                    }
            // face-staggered
            for (int fd = 0; fd < 3; ++fd) {
                int E[3] = {Ne[0] + (fd == 0), Ne[1] + (fd == 1), Ne[2] + (fd == 2)};
                for (int k = 0; k < E[2]; ++k)
                    for (int j = 0; j < E[1]; ++j)
                        for (int i = 0; i < E[0]; ++i) {
                            int loc[3] = {i, j, k};
                            bool ghost = false;
                            for (int d = 0; d < 3; ++d) {
                                int hi = n[d] + ng + (d == fd ? 1 : 0);
                                ghost |= (loc[d] < ng || loc[d] >= hi);
                            }
                            if (!ghost) continue;
                            int sb[3], sl[3];
                            for (int d = 0; d < 3; ++d) {
                                int gg = bd.b[d] * n[d] + loc[d] - ng;
                                if (d == fd) {
                                    int m = map_fc(gg, Ntot[d], I.bc[2 * d], I.bc[2 * d + 1]);
                                    if (m == Ntot[d]) { sb[d] = I.nb[d] - 1; sl[d] = n[d] + ng; }
                                    else { sb[d] = m / n[d]; sl[d] = m % n[d] + ng; }
                                } else {
                                    int m = map_cc(gg, Ntot[d], I.bc[2 * d], I.bc[2 * d + 1]);
                                    sb[d] = m / n[d]; sl[d] = m % n[d] + ng;
                                }
                            }
                            int64_t qs = I.q0 + sb[0] + (int64_t)I.nb[0] * (sb[1] + (int64_t)I.nb[1] * sb[2]);
                            Bf[fd][g.fc(fd, i, j, k, 0, q, 1)] = Bf[fd][g.fc(fd, sl[0], sl[1], sl[2], 0, qs, 1)];
                        }
            }
        }
    }

    // ---------------- c2p over all cells (auxiliaries.cpp:81-180) ----------------
    C2PStats c2p_stats;
    int64_t n_fatal = 0;
    void compute_auxiliaries(double* st, double* Bf[3], double* aux) {
        ORC_PASS(5);
        const Grid& g = G();
        int64_t fatal = 0;
        for (int64_t q = 0; q < g.nq; ++q) {
            double ix = L.idx(q, 0), iy = L.idx(q, 1), iz = L.idx(q, 2);
#pragma omp parallel for collapse(2) schedule(dynamic) reduction(+ : fatal)
            for (int k = 0; k < g.Nk; ++k)
                for (int j = 0; j < g.Nj; ++j)
                    for (int i = 0; i < g.Ni; ++i) {
                        Metric m = metric_at(st, i, j, k, q);
                        double cons[NCONS];
                        cons[DENSL] = st[g.cc(i, j, k, DENS_, q, NSTATE)];
                        cons[STXL] = st[g.cc(i, j, k, SX_, q, NSTATE)];
                        cons[STYL] = st[g.cc(i, j, k, SY_, q, NSTATE)];
                        cons[STZL] = st[g.cc(i, j, k, SZ_, q, NSTATE)];
                        cons[TAUL] = st[g.cc(i, j, k, TAU_, q, NSTATE)];
                        cons[YESL] = st[g.cc(i, j, k, YESTAR_, q, NSTATE)];
                        cons[ENTSL] = st[g.cc(i, j, k, ENTROPYSTAR_, q, NSTATE)];
                        cons[BSXL] = 0.5 * (Bf[0][g.fc(0, i, j, k, 0, q, 1)] + Bf[0][g.fc(0, i + 1, j, k, 0, q, 1)]);
                        cons[BSYL] = 0.5 * (Bf[1][g.fc(1, i, j, k, 0, q, 1)] + Bf[1][g.fc(1, i, j + 1, k, 0, q, 1)]);
                        cons[BSZL] = 0.5 * (Bf[2][g.fc(2, i, j, k, 0, q, 1)] + Bf[2][g.fc(2, i, j, k + 1, 0, q, 1)]);
                        auto A = [&](int v) -> double& { return aux[g.cc(i, j, k, v, q, NAUX)]; };
                        A(BX_) = cons[BSXL] / m.sqrtg;
                        A(BY_) = cons[BSYL] / m.sqrtg;
                        A(BZ_) = cons[BSZL] / m.sqrtg;
                        Bits err;
                        double prims[NPRIM] = {0};
                        double rtp[3] = {cell_rtp0(i, j, k, q), 0.0, 0.0};
                        C2PResult r = conservs_to_prims(cons, prims, m, eos, atmo, exc, c2pp, rtp, err);
                        if (r.fatal) fatal++;
                        A(RHO_) = prims[RHOL]; A(EPS_) = prims[EPSL]; A(PRESS_) = prims[PRESSL];
                        A(TEMP_) = prims[TEMPL]; A(ENTROPY_) = prims[ENTL]; A(YE_) = prims[YEL];
                        A(ZVECX_) = prims[ZXL]; A(ZVECY_) = prims[ZYL]; A(ZVECZ_) = prims[ZZL];
                        auto S = [&](int v) -> double& { return st[g.cc(i, j, k, v, q, NSTATE)]; };
                        if (err.test(E_RESET_DENS)) { A(C2P_DENS_ERR_) += S(DENS_) - cons[DENSL]; S(DENS_) = cons[DENSL]; }
                        if (err.test(E_RESET_STILDE)) { S(SX_) = cons[STXL]; S(SY_) = cons[STYL]; S(SZ_) = cons[STZL]; }
                        if (err.test(E_RESET_TAU)) S(TAU_) = cons[TAUL];
                        if (err.test(E_RESET_ENTROPY)) S(ENTROPYSTAR_) = cons[ENTSL];
                        if (err.test(E_RESET_YE)) S(YESTAR_) = cons[YESL];
                        uint64_t prev = (uint64_t)A(C2P_ERR_);
                        A(C2P_ERR_) = (double)(prev | err.w);
                        A(BDIV_) = ((Bf[0][g.fc(0, i + 1, j, k, 0, q, 1)] - Bf[0][g.fc(0, i, j, k, 0, q, 1)]) * ix
                                    + (Bf[1][g.fc(1, i, j + 1, k, 0, q, 1)] - Bf[1][g.fc(1, i, j, k, 0, q, 1)]) * iy
                                    + (Bf[2][g.fc(2, i, j, k + 1, 0, q, 1)] - Bf[2][g.fc(2, i, j, k, 0, q, 1)]) * iz) / m.sqrtg;
                    }
            if (opt.emf == EMF_GS) {
                // compute_E_center (auxiliaries.cpp:133-162)
                for (int k = 0; k < g.Nk; ++k)
                    for (int j = 0; j < g.Nj; ++j)
                        for (int i = 0; i < g.Ni; ++i) {
                            Metric m = metric_at(st, i, j, k, q);
                            double Bx = 0.5 * (Bf[0][g.fc(0, i, j, k, 0, q, 1)] + Bf[0][g.fc(0, i + 1, j, k, 0, q, 1)]);
                            double By = 0.5 * (Bf[1][g.fc(1, i, j, k, 0, q, 1)] + Bf[1][g.fc(1, i, j + 1, k, 0, q, 1)]);
                            double Bz = 0.5 * (Bf[2][g.fc(2, i, j, k, 0, q, 1)] + Bf[2][g.fc(2, i, j, k + 1, 0, q, 1)]);
                            double z[3] = {aux[g.cc(i, j, k, ZVECX_, q, NAUX)], aux[g.cc(i, j, k, ZVECY_, q, NAUX)],
                                           aux[g.cc(i, j, k, ZVECZ_, q, NAUX)]};
                            double W = std::sqrt(1. + m.square_vec(z));
                            double vt[3] = {m.alp * z[0] / W - m.beta[0], m.alp * z[1] / W - m.beta[1], m.alp * z[2] / W - m.beta[2]};
                            ecen[g.cc(i, j, k, 0, q, 3)] = By * vt[2] - Bz * vt[1];
                            ecen[g.cc(i, j, k, 1, q, 3)] = -Bx * vt[2] + Bz * vt[0];
                            ecen[g.cc(i, j, k, 2, q, 3)] = Bx * vt[1] - By * vt[0];
                        }
            }
        }
        n_fatal += fatal;
    }

    // compute_max_eigenspeed (grmhd.hh:444-519) + dt reduce
    double find_dt(const double* st, const double* aux, double cfl) const {
        const Grid& g = G();
        double dtmin = 1e300;
        int ng = g.ngz;
        for (int64_t q = 0; q < g.nq; ++q) {
            double dx0 = 1.0 / L.idx(q, 0);
            for (int k = ng; k < g.nz + ng; ++k)
                for (int j = ng; j < g.ny + ng; ++j)
                    for (int i = ng; i < g.nx + ng; ++i) {
                        Metric m = metric_at(st, i, j, k, q);
                        auto A = [&](int v) { return aux[g.cc(i, j, k, v, q, NAUX)]; };
                        double z[3] = {A(ZVECX_), A(ZVECY_), A(ZVECZ_)}, B[3] = {A(BX_), A(BY_), A(BZ_)};
                        double rho = A(RHO_), T = A(TEMP_), eps = A(EPS_), press = A(PRESS_);
                        double cs2, h;
                        Bits e;
                        eos.press_h_csnd2__temp_rho_ye(h, cs2, T, rho, e);
                        double W = get_W(m.g, z);
                        double b[4], b2;
                        get_smallb(m.alp, m.beta, m.g, B, z, W, b, b2);
                        double vt[3];
                        get_vtilde(m.alp, m.beta, z, W, vt);
                        double cmax = 0;
                        const int mc[3] = {0, 3, 5};
                        for (int d = 0; d < 3; ++d) {
                            double cm, cp;
                            get_cm_cp(m.alp, m.beta, rho, press, eps, cs2, W, b2, vt, m.gi[mc[d]], d, cm, cp);
                            cmax = std::max(cmax, std::max(std::fabs(cp), std::fabs(cm)));
                        }
                        double dtl = cfl / cmax * dx0;
                        dtmin = dtmin > dtl ? dtl : dtmin;
                    }
        }
        return dtmin;
    }

    // linop_apply, 2-operand form (tstep_utils.hh:140-220): A = b B + c C
    void linop(double* A, const double* Bv, const double* Cv, double b, double c, int64_t n) {
        ORC_PASS(6);
        for (int64_t t = 0; t < n; ++t) A[t] = b * Bv[t] + c * Cv[t];
    }

    // One full RK3 step (evolve.cpp:147-225).  Includes the per-step reset of
    // the c2p diagnostics (:154-161).
    void rk3_step(double dt, double* st, double* Bf[3], double* aux) {
        const Grid& g = G();
        int64_t ncc = g.cc_size(NSTATE);
        std::memcpy(state_p.data(), st, ncc * sizeof(double));
        for (int d = 0; d < 3; ++d) std::memcpy(B_p[d].data(), Bf[d], g.fc_size(d, 1) * sizeof(double));
        for (int64_t t = 0; t < g.cc_size(1); ++t) {
            // reset C2P_DENS_ERR_ and C2P_ERR_ of every cell
            int64_t cell = t % ((int64_t)g.Ni * g.Nj * g.Nk), q = t / ((int64_t)g.Ni * g.Nj * g.Nk);
            aux[cell + (int64_t)g.Ni * g.Nj * g.Nk * (C2P_DENS_ERR_ + (int64_t)NAUX * q)] = 0.0;
            aux[cell + (int64_t)g.Ni * g.Nj * g.Nk * (C2P_ERR_ + (int64_t)NAUX * q)] = 0.0;
        }
        double* Bp[3] = {B_p[0].data(), B_p[1].data(), B_p[2].data()};
        double* Bpp[3] = {B_pp[0].data(), B_pp[1].data(), B_pp[2].data()};
        Fields F0{st, {Bf[0], Bf[1], Bf[2]}};
        Fields F1{state_p.data(), {Bp[0], Bp[1], Bp[2]}};
        Fields F2{state_pp.data(), {Bpp[0], Bpp[1], Bpp[2]}};
        // stage 1
        advance_substep(dt, 1.0, state_p.data(), Bp, F0, aux);
        fill_ghosts(state_p.data(), Bp, 0, NHRSC);
        compute_auxiliaries(state_p.data(), Bp, aux);
        // stage 2
        linop(state_pp.data(), st, state_p.data(), 0.75, 0.25, ncc);
        for (int d = 0; d < 3; ++d) linop(Bpp[d], Bf[d], Bp[d], 0.75, 0.25, g.fc_size(d, 1));
        advance_substep(dt, 0.25, state_pp.data(), Bpp, F1, aux);
        fill_ghosts(state_pp.data(), Bpp, 0, NHRSC);
        compute_auxiliaries(state_pp.data(), Bpp, aux);
        // stage 3
        linop(st, st, state_pp.data(), 1. / 3, 2. / 3., ncc);
        for (int d = 0; d < 3; ++d) linop(Bf[d], Bf[d], Bpp[d], 1. / 3, 2. / 3., g.fc_size(d, 1));
        advance_substep(dt, 2. / 3., st, Bf, F2, aux);
        fill_ghosts(st, Bf, 0, NHRSC);
        compute_auxiliaries(st, Bf, aux);
    }
};

}  // namespace oracle

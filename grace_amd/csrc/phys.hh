// phys.hh -- FP64 GRMHD point physics for the gfx950 kernels (host+device).
//
// EOS, metric algebra, Valencia conserved/flux/wave-speed formulas, Kastaun
// con2prim with entropy backup, reconstruction and HLL.  Reference
// counterparts (paths relative to the GRACE checkout):
//   ideal gas EOS        include/grace/physics/eos/ideal_gas_eos.hh:48-435
//   hybrid pwpoly EOS    include/grace/physics/eos/hybrid_eos.hh:49-484,
//                        piecewise_polytropic_eos.hh:51-193
//   3-metric             include/grace/utils/metric_utils.hh:48-553
//   GRMHD algebra        include/grace/physics/grmhd_subexpressions.hh:10-392
//   Kastaun c2p          include/grace/physics/eos/kastaun_c2p.hh:44-402
//   entropy backup       include/grace/physics/eos/ent_based_c2p.hh:43-256
//   c2p driver           src/physics/eos/c2p.cpp:40-455, c2p.hh:136-322
//   Brent / Newton       include/grace/utils/rootfinding.hh:73-306
//   WENO-Z / MC2 / etc.  include/grace/utils/weno_reconstruction.hh:172-278,
//                        reconstruction.hh:51-142, limiters.hh:42-80
//   HLL                  include/grace/utils/riemann_solvers.hh:42-99
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GHD __host__ __device__ __forceinline__

namespace gm {

// ---- index enums (same numbering as the reference) ----
enum { RHOL = 0, PRESSL, ZXL, ZYL, ZZL, YEL, TEMPL, EPSL, ENTL, BXL, BYL, BZL, CS2L, NPRIM };
enum { DENSL = 0, STXL, STYL, STZL, TAUL, YESL, ENTSL, BSXL, BSYL, BSZL, NCONS };
enum { DENS_ = 0, SX_, SY_, SZ_, TAU_, YESTAR_, ENTROPYSTAR_, GXX_, GXY_, GXZ_, GYY_, GYZ_, GZZ_,
       ALP_, BETAX_, BETAY_, BETAZ_, KXX_, KXY_, KXZ_, KYY_, KYZ_, KZZ_, NSTATE };
constexpr int NHRSC = 7;
enum { RHO_ = 0, ZVECX_, ZVECY_, ZVECZ_, BX_, BY_, BZ_, YE_, TEMP_, ENTROPY_, EPS_, PRESS_,
       BDIV_, C2P_DENS_ERR_, C2P_ERR_, NAUX };
enum { SIG_EPS_TOO_HIGH = 0, SIG_EPS_TOO_LOW, SIG_YE_TOO_LOW, SIG_YE_TOO_HIGH, SIG_RHO_TOO_HIGH,
       SIG_RHO_TOO_LOW, SIG_ENT_TOO_LOW, SIG_ENT_TOO_HIGH, SIG_VEL_TOO_HIGH, SIG_SIGMA_TOO_HIGH };
enum { E_RESET_DENS = 0, E_RESET_TAU, E_RESET_STILDE, E_RESET_ENTROPY, E_RESET_YE,
       E_SIG_RHO_TOO_LOW, E_SIG_RHO_TOO_HIGH, E_SIG_EPS_TOO_LOW, E_SIG_EPS_TOO_HIGH,
       E_SIG_YE_TOO_LOW, E_SIG_YE_TOO_HIGH, E_SIG_ENT_TOO_LOW, E_SIG_ENT_TOO_HIGH,
       E_SIG_TEMP_TOO_LOW, E_SIG_TEMP_TOO_HIGH, E_SIG_PRESS_TOO_LOW, E_SIG_PRESS_TOO_HIGH,
       E_SIG_VEL_TOO_HIGH, E_SIG_SIGMA_TOO_HIGH, E_ENT_BACKUP_USED, E_ATMO_RESET, E_T_FLOORED };
enum { EOS_RHO_TOO_HIGH = 0, EOS_RHO_TOO_LOW, EOS_EPS_TOO_HIGH, EOS_EPS_TOO_LOW, EOS_YE_TOO_LOW,
       EOS_YE_TOO_HIGH, EOS_TEMPERATURE_TOO_LOW, EOS_TEMPERATURE_TOO_HIGH, EOS_ENTROPY_TOO_LOW,
       EOS_ENTROPY_TOO_HIGH, EOS_PRESS_TOO_LOW, EOS_PRESS_TOO_HIGH };

#define BIT(x) (1u << (x))

GHD double sq(double x) { return x * x; }

// ---- 3-metric -------------------------------------------------------------
struct Met {
    double g[6], gi[6], beta[3], alp, sqrtg;
    GHD void init() {
        double det = -(sq(g[2]) * g[3]) + 2 * g[1] * g[2] * g[4] - g[0] * sq(g[4]) - sq(g[1]) * g[5]
                     + g[0] * g[3] * g[5];
        double id = 1.0 / det;
        gi[0] = (g[3] * g[5] - sq(g[4])) * id;
        gi[1] = (-g[1] * g[5] + g[2] * g[4]) * id;
        gi[2] = (-(g[2] * g[3]) + g[1] * g[4]) * id;
        gi[3] = (g[5] * g[0] - sq(g[2])) * id;
        gi[4] = (g[1] * g[2] - g[0] * g[4]) * id;
        gi[5] = (-sq(g[1]) + g[0] * g[3]) * id;
        sqrtg = sqrt(det);
    }
    GHD void flat() {
        g[0] = g[3] = g[5] = 1.0; g[1] = g[2] = g[4] = 0.0;
        gi[0] = gi[3] = gi[5] = 1.0; gi[1] = gi[2] = gi[4] = 0.0;
        beta[0] = beta[1] = beta[2] = 0.0; alp = 1.0; sqrtg = 1.0;
    }
    GHD double sqv(double x, double y, double z) const {
        return g[0] * x * x + g[3] * y * y + g[5] * z * z + 2. * (g[1] * x * y + g[2] * x * z + g[4] * y * z);
    }
    GHD double sqcov(double x, double y, double z) const {
        return gi[0] * x * x + gi[3] * y * y + gi[5] * z * z + 2. * (gi[1] * x * y + gi[2] * x * z + gi[4] * y * z);
    }
};

// ---- EOS ------------------------------------------------------------------
struct EosIdeal {
    static constexpr bool has_ye = false, is_ideal = true;
    double gm1, gamma, ggm1, eps_max, ent_min, temp_atm;
    double ls_min, ls_max;   // log(ent_min)/gm1, log(eps_max)/gm1 (entropy bounds are these minus log rho)
    GHD double hmin() const { return 1. + 1e-15; }
    GHD double rhomax() const { return 1e100; }
    GHD double rhomin() const { return 1e-100; }
    GHD double yemax() const { return 1e100; }
    GHD double yemin() const { return 0.0; }
    GHD double ent_of(double eps, double rho) const { return log(eps) / gm1 - log(rho); }
    GHD void limit_eps(double& eps, unsigned& e) const {
        if (eps > eps_max) { e |= BIT(EOS_EPS_TOO_HIGH); eps = eps_max; }
        if (eps < 0.0) { e |= BIT(EOS_EPS_TOO_LOW); eps = 0.0; }
    }
    GHD void limit_temp(double& t, unsigned& e) const {
        double tmax = eps_max * gm1;
        if (t < 0.0) { t = 0.0; e |= BIT(EOS_TEMPERATURE_TOO_LOW); }
        if (t > tmax) { e |= BIT(EOS_TEMPERATURE_TOO_HIGH); t = tmax; }
    }
    GHD void eps_range(double& emin, double& emax, double) const { emin = 0.0; emax = eps_max; }
    GHD double press_eps(double& eps, double rho, unsigned& e) const { limit_eps(eps, e); return gm1 * eps * rho; }
    GHD double press_full_eps(double& h, double& cs2, double& temp, double& ent, double& eps, double rho,
                              unsigned& e) const {
        limit_eps(eps, e);
        double p = gm1 * eps * rho;
        h = 1.0 + gamma * eps; cs2 = ggm1 * eps / h; temp = eps * gm1; ent = ent_of(eps, rho);
        return p;
    }
    GHD double press_full_ent(double& h, double& cs2, double& temp, double& eps, double& ent, double rho,
                              unsigned& e) const {
        const double lr = log(rho);
        double smin = ls_min - lr, smax = ls_max - lr;
        if (ent < smin) { ent = smin; e |= BIT(EOS_ENTROPY_TOO_LOW); }
        if (ent > smax) { ent = smax; e |= BIT(EOS_ENTROPY_TOO_HIGH); }
        eps = exp(gm1 * (ent + lr));
        double p = gm1 * eps * rho;
        h = 1.0 + gamma * eps; temp = eps * gm1; cs2 = ggm1 * eps / h;
        return p;
    }
    // press_eps_csnd2_entropy__temp_rho_ye
    GHD double press_from_temp_ent(double& eps, double& cs2, double& ent, double& temp, double rho,
                                   unsigned& e) const {
        limit_temp(temp, e);
        eps = temp / gm1;
        double p = gm1 * eps * rho;
        double h = 1.0 + gamma * eps;
        cs2 = ggm1 * eps / h; ent = ent_of(eps, rho);
        return p;
    }
    // press_eps_csnd2__temp_rho_ye (face EOS)
    GHD double press_from_temp(double& eps, double& cs2, double& temp, double rho, unsigned& e) const {
        limit_temp(temp, e);
        eps = temp / gm1;
        double p = gm1 * eps * rho;
        double h = 1.0 + gamma * eps;
        cs2 = ggm1 * eps / h;
        return p;
    }
    GHD double press_h_cs2_from_temp(double& h, double& cs2, double& temp, double rho, unsigned& e) const {
        limit_temp(temp, e);
        double eps = temp / gm1;
        double p = gm1 * eps * rho;
        h = 1.0 + gamma * eps; cs2 = ggm1 * eps / h;
        return p;
    }
    // eps_h_csnd2_temp_entropy__press_rho_ye
    GHD double eps_from_press(double& h, double& cs2, double& temp, double& ent, double& press, double rho,
                              unsigned&) const {
        double eps = press / (gm1 * rho);
        h = 1.0 + gamma * eps; cs2 = ggm1 * eps / h; ent = ent_of(eps, rho); temp = eps * gm1;
        return eps;
    }
};

struct EosHybrid {
    static constexpr bool has_ye = false, is_ideal = false;
    int n;
    double K[8], G[8], rh[8], ep[8];
    double gth_m1, entropy_min, eps_max, temp_atm;
    GHD double hmin() const { return 1. + 1e-10; }
    GHD double rhomax() const { return 1e10; }
    GHD double rhomin() const { return 1e-100; }
    GHD double yemax() const { return 1e99; }
    GHD double yemin() const { return 0.0; }
    GHD int idx(double& rho) const {
        rho = fmax(rhomin(), fmin(rhomax(), rho));
        for (int ii = 0; ii < n - 1; ++ii)
            if (rho > rh[ii] && rho <= rh[ii + 1]) return ii;
        return n - 1;
    }
    GHD double pcold(double& ec, double rho) const {
        int i = idx(rho);
        double pc = K[i] * pow(rho, G[i]);
        ec = ep[i] + pc / (rho * (G[i] - 1.));
        return pc;
    }
    GHD double dpcold(double rho) const { int i = idx(rho); return G[i] * (K[i] * pow(rho, G[i])) / rho; }
    GHD void chk_rho(double& rho, unsigned& e) const {
        if (rho < rhomin()) { e |= BIT(EOS_RHO_TOO_LOW); rho = rhomin(); }
        if (rho > rhomax()) { e |= BIT(EOS_RHO_TOO_HIGH); rho = rhomax(); }
    }
    GHD void chk_eps(double& eps, double ec, unsigned& e) const {
        double emin = fmin(ec, eps_max);
        if (eps < emin) { e |= BIT(EOS_EPS_TOO_LOW); eps = emin; }
        if (eps > eps_max) { e |= BIT(EOS_EPS_TOO_HIGH); eps = eps_max; }
    }
    GHD void chk_temp(double& t, unsigned& e) const {
        if (t < 0) { e |= BIT(EOS_TEMPERATURE_TOO_LOW); t = 0; }
        if (t > 1e99) { e |= BIT(EOS_TEMPERATURE_TOO_HIGH); t = 1e99; }
    }
    GHD double ent_of(double eth, double rho) const {
        double x = eth > entropy_min ? eth : entropy_min;
        return log(x * pow(rho, -gth_m1)) / gth_m1;
    }
    GHD double t_of(double eth) const { return fmax(0.0, eth * gth_m1); }
    GHD double eth_of_t(double t) const { return fmax(0.0, t / gth_m1); }
    GHD double cs2_of(double rho, double eth, double h) const {
        return (dpcold(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
    }
    GHD void eps_range(double& emin, double& emax, double rho) const {
        unsigned e = 0; chk_rho(rho, e);
        double ec; pcold(ec, rho);
        emin = fmin(ec, eps_max); emax = eps_max;
    }
    GHD double press_eps(double& eps, double rho, unsigned& e) const {
        chk_rho(rho, e);
        double ec; double pc = pcold(ec, rho);
        chk_eps(eps, ec, e);
        return pc + (eps - ec) * rho * gth_m1;
    }
    GHD double press_full_eps(double& h, double& cs2, double& temp, double& ent, double& eps, double rho,
                              unsigned& e) const {
        chk_rho(rho, e);
        double ec; double pc = pcold(ec, rho);
        chk_eps(eps, ec, e);
        double eth = eps - ec;
        double p = pc + eth * rho * gth_m1;
        h = 1. + eps + p / rho; cs2 = cs2_of(rho, eth, h); ent = ent_of(eth, rho); temp = t_of(eth);
        return p;
    }
    GHD double press_full_ent(double& h, double& cs2, double& temp, double& eps, double& ent, double rho,
                              unsigned& e) const {
        chk_rho(rho, e);
        {
            double r2 = rho; unsigned e2 = 0; chk_rho(r2, e2);
            double ec; pcold(ec, r2);
            double smin = ent_of(entropy_min, r2), smax = ent_of(eps_max - ec, r2);
            if (ent < smin) { e |= BIT(EOS_ENTROPY_TOO_LOW); ent = smin; }
            if (ent > smax) { e |= BIT(EOS_ENTROPY_TOO_HIGH); ent = smax; }
        }
        double ec; double pc = pcold(ec, rho);
        double eth = exp(gth_m1 * ent) * pow(rho, gth_m1);
        double p = pc + eth * rho * gth_m1;
        temp = t_of(eth); eps = ec + eth; h = 1. + eps + p / rho; cs2 = cs2_of(rho, eth, h);
        return p;
    }
    GHD double press_from_temp_ent(double& eps, double& cs2, double& ent, double& temp, double rho,
                                   unsigned& e) const {
        chk_rho(rho, e); chk_temp(temp, e);
        double ec; double pc = pcold(ec, rho);
        double eth = eth_of_t(temp);
        double p = pc + eth * rho * gth_m1;
        eps = eth + ec;
        double h = 1. + eps + p / rho;
        cs2 = cs2_of(rho, eth, h); ent = ent_of(eth, rho);
        return p;
    }
    GHD double press_from_temp(double& eps, double& cs2, double& temp, double rho, unsigned& e) const {
        chk_rho(rho, e); chk_temp(temp, e);
        double ec; double pc = pcold(ec, rho);
        double eth = eth_of_t(temp);
        double p = pc + eth * rho * gth_m1;
        eps = ec + eth;
        double h = 1. + eps + p / rho;
        cs2 = cs2_of(rho, eth, h);
        return p;
    }
    GHD double press_h_cs2_from_temp(double& h, double& cs2, double& temp, double rho, unsigned& e) const {
        chk_rho(rho, e); chk_temp(temp, e);
        double ec; double pc = pcold(ec, rho);
        double eth = eth_of_t(temp);
        double p = pc + eth * rho * gth_m1;
        h = 1. + ec + eth + p / rho; cs2 = cs2_of(rho, eth, h);
        return p;
    }
    GHD double eps_from_press(double& h, double& cs2, double& temp, double& ent, double& press, double rho,
                              unsigned& e) const {
        chk_rho(rho, e);
        double ec; double pc = pcold(ec, rho);
        press = fmax(press, pc);
        double eth = fmax(0, (press - pc) / (rho * gth_m1));
        double eps = eth + ec;
        h = 1. + eps + press / rho; cs2 = cs2_of(rho, eth, h); ent = ent_of(eth, rho); temp = t_of(eth);
        return eps;
    }
};

struct Atmo { double rho_fl, temp_fl, ye_fl, atmo_tol; };
struct C2PPar { double tol, max_w, max_sigma, beta_fallback, alp_bh_thresh; int use_ent_backup; };
struct ExcPar { double r_ex, alp_ex; int by_radius, cks; double spin; };

// Radius of a point for radius-based excision (cart_to_sph,
// cartesian_coordinate_systems.hh:87-107): |x|, or in Cartesian Kerr-Schild
// coordinates of a hole of spin a the spheroidal radius
// r^2 = (R^2 - a^2 + sqrt((R^2 - a^2)^2 + 4 a^2 z^2)) / 2, floored at 1e-6.
GHD double excision_radius_of(const ExcPar& ex, double x, double y, double z) {
    const double R2 = x * x + y * y + z * z;
    if (!ex.cks) return sqrt(R2);
    const double a2 = ex.spin * ex.spin, d = R2 - a2;
    return fmax(sqrt(d + sqrt(d * d + 4.0 * a2 * z * z)) / sqrt(2.0), 1.0e-6);
}

// ---- GRMHD algebra ----------------------------------------------------------
// FL = true: the metric is the identity (lapse 1, shift 0, gamma_ij =
// delta_ij) and the terms that multiply a zero metric entry are left out.
// IEEE arithmetic cannot drop them by itself (0 x is not 0 for infinite or
// NaN x, and x + 0 turns a zero x of negative sign positive), so the
// flat-metric kernels ask for it; on finite inputs the values are those of
// the general formulas up to the sign of a zero.
template <bool FL = false>
GHD void lower3(const double g[6], double x, double y, double z, double o[3]) {
    if constexpr (FL) {
        o[0] = x; o[1] = y; o[2] = z;
    } else {
        o[0] = g[0] * x + g[1] * y + g[2] * z;
        o[1] = g[1] * x + g[3] * y + g[4] * z;
        o[2] = g[2] * x + g[4] * y + g[5] * z;
    }
}
template <bool FL = false>
GHD double lorentz(const double g[6], const double z[3]) {
    double zl[3]; lower3<FL>(g, z[0], z[1], z[2], zl);
    return sqrt(z[0] * zl[0] + z[1] * zl[1] + z[2] * zl[2] + 1);
}
template <bool FL = false>
GHD void smallb(const Met& m, const double B[3], const double z[3], double W, double b[4], double& b2) {
    double zl[3]; lower3<FL>(m.g, z[0], z[1], z[2], zl);
    double Bz = B[0] * zl[0] + B[1] * zl[1] + B[2] * zl[2];
    double iW = 1. / W;
    if constexpr (FL) {
        b[0] = Bz;
        for (int i = 0; i < 3; ++i) b[1 + i] = iW * (B[i] + Bz * z[i]);
    } else {
        b[0] = Bz / m.alp;
        for (int i = 0; i < 3; ++i) b[1 + i] = iW * (B[i] + Bz * (z[i] - W * m.beta[i] / m.alp));
    }
    double Bl[3]; lower3<FL>(m.g, B[0], B[1], B[2], Bl);
    b2 = (B[0] * Bl[0] + B[1] * Bl[1] + B[2] * Bl[2] + Bz * Bz) / (W * W);
}
GHD void vtilde(const Met& m, const double z[3], double W, double vt[3]) {
    double iW = 1. / W;
    for (int i = 0; i < 3; ++i) vt[i] = m.alp * iW * z[i] - m.beta[i];
}
template <bool FL = false>
GHD void cmcp(const Met& m, double rho, double press, double eps, double cs2, double W, double b2,
              const double vt[3], int d, double& cm, double& cp) {
    const double gdd = d == 0 ? m.gi[0] : (d == 1 ? m.gi[3] : m.gi[5]);
    double ia2 = 1. / (m.alp * m.alp);
    double vA2 = b2 / (b2 + eps * rho + press + rho);
    double w2 = vA2 + cs2 * (1. - vA2);
    double W2_1mw2 = W * W * (1. - w2);
    double A = (W2_1mw2 + w2) * ia2;
    double Bc;
    if constexpr (FL) Bc = 2. * (-(vt[d] * W2_1mw2));
    else Bc = 2. * ia2 * (m.beta[d] * w2 - vt[d] * W2_1mw2);
    double g4dd = gdd - m.beta[d] * m.beta[d] * ia2;
    double C = vt[d] * vt[d] * W2_1mw2 * ia2 - w2 * g4dd;
    double disc = sqrt(fmax(0., Bc * Bc - 4. * A * C));
    double i2A = 1. / (2. * A);
    double l1 = (-Bc + disc) * i2A, l2 = (-Bc - disc) * i2A;
    cm = fmin(l1, l2); cp = fmax(l1, l2);
}
// conserved D, tau, S_j, S* (undensitised)
template <bool FL = false>
GHD void conserved(const Met& m, double rho, double press, double eps, const double z[3], double s, double W,
                   double b2, const double b[4], double& D, double& tau, double S[3], double& Ss) {
    double rhoh = b2 + eps * rho + press + rho;
    D = W * rho;
    tau = W * W * rhoh - press - 0.5 * b2 - m.alp * m.alp * b[0] * b[0] - D;
    double zl[3], bsh[3], bl[3];
    lower3<FL>(m.g, z[0], z[1], z[2], zl);
    lower3<FL>(m.g, b[1], b[2], b[3], bl);
    if constexpr (FL) {
        for (int j = 0; j < 3; ++j) S[j] = rhoh * W * zl[j] - b[0] * bl[j];
    } else {
        lower3(m.g, m.beta[0], m.beta[1], m.beta[2], bsh);
        for (int j = 0; j < 3; ++j) S[j] = rhoh * W * zl[j] - m.alp * b[0] * (bl[j] + b[0] * bsh[j]);
    }
    Ss = s * D;
}

// ---- roots --------------------------------------------------------------------
// Brent's zero-in (Brent 1973, ch. 4): same step sequence as rootfinding.hh:73-231.
constexpr int BRENT_MAX_ITER = 1000;
template <typename F>
__host__ __device__ inline double brent_fab(F& f, double a, double b, double fa, double fb, double t) {
    double sa = a, sb = b;
    double c = sa, fc = fa, e = sb - sa, d = e;
    const double macheps = 2.220446049250313e-16;
    // Finite inputs converge in far fewer steps (the interval at least halves
    // every few iterations); the cap only ends the loop on NaN input.
    for (int it = 0; it < BRENT_MAX_ITER; ++it) {
        if (fabs(fc) < fabs(fb)) { sa = sb; sb = c; c = sa; fa = fb; fb = fc; fc = fa; }
        double tol = 2.0 * macheps * fabs(sb) + t;
        double mm = 0.5 * (c - sb);
        if (fabs(mm) <= tol || fb == 0.0) break;
        if (fabs(e) < tol || fabs(fa) <= fabs(fb)) { e = mm; d = e; }
        else {
            double s = fb / fa, p, q;
            if (sa == c) { p = 2.0 * mm * s; q = 1.0 - s; }
            else {
                double qq = fa / fc, r = fb / fc;
                p = s * (2.0 * mm * qq * (qq - r) - (sb - sa) * (r - 1.0));
                q = (qq - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (0.0 < p) q = -q; else p = -p;
            s = e; e = d;
            if (2.0 * p < 3.0 * mm * q - fabs(tol * q) && p < fabs(0.5 * s * q)) d = p / q;
            else { e = mm; d = e; }
        }
        sa = sb; fa = fb;
        if (tol < fabs(d)) sb = sb + d; else if (0.0 < mm) sb = sb + tol; else sb = sb - tol;
        fb = f(sb);
        if ((0.0 < fb && 0.0 < fc) || (fb <= 0.0 && fc <= 0.0)) { c = sa; fc = fa; e = sb - sa; d = e; }
    }
    return sb;
}
template <typename F>
__host__ __device__ inline double brent(F& f, double a, double b, double t) {
    double fa = f(a), fb = f(b);
    return brent_fab(f, a, b, fa, fb, t);
}
template <typename F>
__host__ __device__ inline double newton_br(double sa, double sb, F& f, int maxiter, double tol, int& err) {
    const double macheps = 2.220446049250313e-16;
    double a = sa, b = sb, fa, fb, dd;
    f(a, fa, dd); f(b, fb, dd);
    if (fa == 0) { err = 0; return a; }
    if (fb == 0) { err = 0; return b; }
    if (fa * fb > 0) { err = 1; return 0; }
    double x = (fb * a - fa * b) / (fb - fa);
    int it = 0;
    double xold, fx, dfx, t;
    do {
        xold = x; f(x, fx, dfx);
        if (fx * fb > 0) { fb = fx; b = x; } else if (fx * fa > 0) { fa = fx; a = x; }
        x -= fx / dfx;
        t = 2. * macheps * fabs(x) + tol;
        if (fabs(x - xold) < t || fx == 0) { err = 0; return x; }
        if (x > b || x < a) x = 0.5 * (b + a);
        it++;
    } while (it < maxiter);
    err = 1;
    return 0;
}

// ---- Kastaun master functions ----------------------------------------------------
struct FBr {
    double rsqr, bsqr, rbsqr, h0sqr; int* cnt;
    GHD double xm(double mu) const { return 1. / (1. + mu * bsqr); }
    GHD double rb(double mu, double x) const { return x * (rsqr * x + mu * (x + 1.0) * rbsqr); }
    GHD void operator()(double mu, double& f, double& df) {
        (*cnt)++;
        double x = xm(mu), x2 = x * x;
        double hw = sqrt(h0sqr + rb(mu, x));
        double bb = x * (x2 * rsqr + mu * (1 + x + x2) * rbsqr);
        f = mu * hw - 1.; df = (h0sqr + bb) / hw;
    }
    GHD double operator()(double mu) { (*cnt)++; double x = xm(mu); return mu * sqrt(h0sqr + rb(mu, x)) - 1.; }
    GHD void bracket(double& lo, double& hi) const {
        lo = 1. / (h0sqr + rsqr);
        double mu0 = 1. / sqrt(h0sqr);
        hi = 1.0 / sqrt(h0sqr + rb(mu0, xm(mu0)));
        lo *= (1 - 1e-10); hi *= (1 + 1e-10);
        if (hi <= lo) { lo = 0; hi = mu0 * (1.0 + 1e-10); }
    }
};

template <typename EOS>
struct FMu {
    const EOS& eos;
    double d, qtot, rsqr, bsqr, rbsqr, brosqr, vsqrmax, wmax; int* cnt;
    GHD FMu(const EOS& e, double d_, double q_, double r2, double rb2, double b2, double h0, int* c)
        : eos(e), d(d_), qtot(q_), rsqr(r2), bsqr(b2), rbsqr(rb2), brosqr(r2 * b2 - rb2), cnt(c) {
        double zs = rsqr / (h0 * h0), ws = 1 + zs;
        wmax = sqrt(ws); vsqrmax = zs / ws;
    }
    GHD double xm(double mu) const { return 1. / (1. + mu * bsqr); }
    GHD double rf(double mu, double x) const { return x * (rsqr * x + mu * (x + 1.0) * rbsqr); }
    GHD double qf(double mu, double x) const { double mx = mu * x; return qtot - 0.5 * (bsqr + mx * mx * brosqr); }
    GHD double epsraw(double mu, double q, double r, double w) const { return w * (q - mu * r * (1.0 - mu * w / (1 + w))); }
    GHD double operator()(double mu) {
        (*cnt)++;
        double x = xm(mu), r = rf(mu, x), q = qf(mu, x);
        double vs = r * mu * mu, w;
        if (vs > vsqrmax) { w = wmax; } else w = 1 / sqrt(1 - vs);
        double rho = fmin(eos.rhomax(), fmax(eos.rhomin(), d / w));
        double eps = epsraw(mu, q, r, w);
        double emin, emax; eos.eps_range(emin, emax, rho);
        eps = fmin(emax, fmax(emin, eps));
        unsigned ee = 0;
        double p = eos.press_eps(eps, rho, ee);
        double a = p / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        double hbw = fmax((1 + a) * (1 + q - mu * r), h / w);
        return mu - 1. / (hbw + r * mu);
    }
    // Value and derivative in mu of the function above, ideal gas only (a =
    // (gamma-1) eps/(1+eps) depends on eps alone); the eps, velocity and
    // density clamps contribute zero slope where they are active.
    GHD void deriv(double mu, double& f, double& df) {
        (*cnt)++;
        const double x = xm(mu), dx = -bsqr * x * x;
        const double r = rf(mu, x), dr = 2.0 * rsqr * x * dx + rbsqr * (x * (x + 1.0) + mu * dx * (2.0 * x + 1.0));
        const double mx = mu * x, q = qf(mu, x), dq = -mx * (x + mu * dx) * brosqr;
        const double vs = r * mu * mu, dvs = dr * mu * mu + 2.0 * r * mu;
        double w, dw;
        if (vs > vsqrmax) { w = wmax; dw = 0.0; } else { w = 1 / sqrt(1 - vs); dw = 0.5 * w * w * w * dvs; }
        const double rho = fmin(eos.rhomax(), fmax(eos.rhomin(), d / w));
        const double i1w = 1.0 / (1.0 + w);
        const double u = mu * w * i1w, du = (w * (1.0 + w) + mu * dw) * i1w * i1w;
        const double g = mu * r * (1.0 - u), dg = (r + mu * dr) * (1.0 - u) - mu * r * du;
        double eps = w * (q - g), deps = dw * (q - g) + w * (dq - dg);
        double emin, emax; eos.eps_range(emin, emax, rho);
        if (eps > emax || eps < emin) deps = 0.0;
        eps = fmin(emax, fmax(emin, eps));
        unsigned ee = 0;
        const double p = eos.press_eps(eps, rho, ee);
        const double i1e = 1.0 / (1 + eps);
        const double a = p / (rho * (1 + eps)), da = eos.gm1 * deps * i1e * i1e;
        const double h = (1 + eps) * (1 + a), dh = deps * (1 + a) + (1 + eps) * da;
        const double A = (1 + a) * (1 + q - mu * r), B = h / w;
        double hbw, dhbw;
        if (A >= B) { hbw = A; dhbw = da * (1 + q - mu * r) + (1 + a) * (dq - r - mu * dr); }
        else { hbw = B; dhbw = (dh - h * dw / w) / w; }
        const double den = hbw + r * mu, iden = 1.0 / den;
        f = mu - iden;
        df = 1.0 + (dhbw + r + mu * dr) * iden * iden;
    }
    GHD double prims(double mu, unsigned& sig, double& x, double& w, double& rho, double& press, double& eps,
                     double& temp, double& ent) {
        x = xm(mu);
        double r = rf(mu, x), q = qf(mu, x);
        double vs = r * mu * mu;
        if (vs > vsqrmax) { w = wmax; sig |= BIT(SIG_VEL_TOO_HIGH); } else w = 1 / sqrt(1 - vs);
        rho = d / w;
        if (rho >= eos.rhomax()) { sig |= BIT(SIG_RHO_TOO_HIGH); rho = eos.rhomax(); }
        else if (rho <= eos.rhomin()) { sig |= BIT(SIG_RHO_TOO_LOW); rho = eos.rhomin(); }
        eps = epsraw(mu, q, r, w);
        double emin, emax; eos.eps_range(emin, emax, rho);
        emax = fmin(emax, eos.eps_max);
        if (eps >= emax) { sig |= BIT(SIG_EPS_TOO_HIGH); eps = emax * 0.999; }
        else if (eps < emin) { sig |= BIT(SIG_EPS_TOO_LOW); eps = emin; }
        double hh, cs2; unsigned ee = 0;
        press = eos.press_full_eps(hh, cs2, temp, ent, eps, rho, ee);
        double a = press / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        double hbw = fmax((1 + a) * (1 + q - mu * r), h / w);
        return mu - 1. / (hbw + r * mu);
    }
};

template <typename EOS>
struct FMuEnt {
    const EOS& eos;
    double d, s, rsqr, bsqr, rbsqr, vsqrmax, wmax; int* cnt;
    GHD FMuEnt(const EOS& e, double d_, double r2, double rb2, double b2, double s_, double h0, int* c)
        : eos(e), d(d_), s(s_), rsqr(r2), bsqr(b2), rbsqr(rb2), cnt(c) {
        double zs = rsqr / (h0 * h0), ws = 1 + zs;
        wmax = sqrt(ws); vsqrmax = zs / ws;
    }
    GHD double xm(double mu) const { return 1. / (1. + mu * bsqr); }
    GHD double rf(double mu, double x) const { return x * (rsqr * x + mu * (x + 1.0) * rbsqr); }
    GHD double operator()(double mu) {
        (*cnt)++;
        double x = xm(mu), r = rf(mu, x);
        double vs = r * mu * mu, w = 1 / sqrt(1 - vs);
        if (vs > vsqrmax) w = wmax;
        double rho = fmin(eos.rhomax(), fmax(eos.rhomin(), d / w));
        double hh, cs2, temp, eps, sl = s; unsigned ee = 0;
        double p = eos.press_full_ent(hh, cs2, temp, eps, sl, rho, ee);
        double a = p / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        return mu - 1. / (h / w + r * mu);
    }
    // Value and derivative in mu, ideal gas only: eps = exp((gamma-1)(s + log rho))
    // gives d eps/d rho = (gamma-1) eps/rho while s is inside its clamps.
    GHD void deriv(double mu, double& f, double& df) {
        (*cnt)++;
        const double x = xm(mu), dx = -bsqr * x * x;
        const double r = rf(mu, x), dr = 2.0 * rsqr * x * dx + rbsqr * (x * (x + 1.0) + mu * dx * (2.0 * x + 1.0));
        const double vs = r * mu * mu, dvs = dr * mu * mu + 2.0 * r * mu;
        double w = 1 / sqrt(1 - vs), dw = 0.5 * w * w * w * dvs;
        if (vs > vsqrmax) { w = wmax; dw = 0.0; }
        const double rr = d / w;
        const double rho = fmin(eos.rhomax(), fmax(eos.rhomin(), rr));
        const double drho = (rr == rho) ? -rho * dw / w : 0.0;
        double hh, cs2, temp, eps, sl = s; unsigned ee = 0;
        const double p = eos.press_full_ent(hh, cs2, temp, eps, sl, rho, ee);
        const double deps = (ee & (BIT(EOS_ENTROPY_TOO_LOW) | BIT(EOS_ENTROPY_TOO_HIGH))) ? 0.0 : eos.gm1 * eps * drho / rho;
        const double i1e = 1.0 / (1 + eps);
        const double a = p / (rho * (1 + eps)), da = eos.gm1 * deps * i1e * i1e;
        const double h = (1 + eps) * (1 + a), dh = deps * (1 + a) + (1 + eps) * da;
        const double den = h / w + r * mu, iden = 1.0 / den;
        f = mu - iden;
        df = 1.0 + ((dh - h * dw / w) / w + r + mu * dr) * iden * iden;
    }
    GHD double prims(double mu, unsigned& sig, double& x, double& w, double& rho, double& press, double& eps,
                     double& temp, double& ent) {
        ent = s;
        x = xm(mu);
        double r = rf(mu, x);
        double vs = r * mu * mu;
        w = 1 / sqrt(1 - vs);
        if (vs > vsqrmax) { w = wmax; sig |= BIT(SIG_VEL_TOO_HIGH); }
        rho = d / w;
        if (rho >= eos.rhomax()) { sig |= BIT(SIG_RHO_TOO_HIGH); rho = eos.rhomax(); }
        else if (rho <= eos.rhomin()) { sig |= BIT(SIG_RHO_TOO_LOW); rho = eos.rhomin(); }
        double hh, cs2; unsigned ee = 0;
        press = eos.press_full_ent(hh, cs2, temp, eps, ent, rho, ee);
        if (ee & BIT(EOS_ENTROPY_TOO_LOW)) sig |= BIT(SIG_ENT_TOO_LOW);
        if (ee & BIT(EOS_ENTROPY_TOO_HIGH)) sig |= BIT(SIG_ENT_TOO_HIGH);
        double a = press / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        return mu - 1. / (h / w + r * mu);
    }
};

// Invariants of the undensitised conserved state shared by both inversions.
struct C2PIn {
    double D, q, s, ye, r2, Bt2, rBt, rBt2, h0;
    double r[3], Bt[3], B[3];
};
template <typename EOS>
GHD void c2p_setup(const EOS& eos, const Met& m, double cons[NCONS], C2PIn& c) {
    c.h0 = eos.hmin();
    cons[DENSL] = fmax(0., cons[DENSL]);
    c.D = cons[DENSL];
    double iD = 1.0 / c.D;
    c.q = cons[TAUL] / c.D;
    c.s = cons[ENTSL] / c.D;
    double rl0 = cons[STXL] / c.D, rl1 = cons[STYL] / c.D, rl2 = cons[STZL] / c.D;
    (void)iD;
    double sD = sqrt(c.D);
    for (int i = 0; i < 3; ++i) { c.Bt[i] = cons[BSXL + i] / sD; c.B[i] = cons[BSXL + i]; }
    c.r2 = m.sqcov(rl0, rl1, rl2);
    c.Bt2 = m.sqv(c.Bt[0], c.Bt[1], c.Bt[2]);
    c.rBt = rl0 * c.Bt[0] + rl1 * c.Bt[1] + rl2 * c.Bt[2];
    c.rBt2 = c.rBt * c.rBt;
    c.r[0] = m.gi[0] * rl0 + m.gi[1] * rl1 + m.gi[2] * rl2;
    c.r[1] = m.gi[1] * rl0 + m.gi[3] * rl1 + m.gi[4] * rl2;
    c.r[2] = m.gi[2] * rl0 + m.gi[4] * rl1 + m.gi[5] * rl2;
    c.ye = cons[YESL] / c.D;
}

struct C2PCount { int root = 0, bracket = 0; int guess_used = 0; };
// Optional brackets for the two root solves, as fractions of 1/h0 (the
// master functions have one root on (0, 1/h0]): energy [e_lo, e_hi], entropy
// [s_lo, s_hi].  A bracket is used when 0 <= lo < hi <= 1 and the master
// function changes sign on it; the Newton solve for the upper bound mu+ and
// the default bracket [0, mu+] are then skipped.  Otherwise the plain path
// runs unchanged.
// nexp > 0 lets a bracket that misses the root be moved towards it (see
// guess_bracket); s_rel makes [s_lo, s_hi] factors on the energy root.
// newton > 0 (ideal gas only) first runs a safeguarded Newton iteration on
// the master function from e_c (entropy: s_c, a factor on the energy root
// when s_rel), at most `newton` evaluations; where it does not converge, or
// its root fails the residual check, the bracket path above follows.
struct C2PGuess { double e_lo, e_hi, s_lo, s_hi; int nexp, s_rel; int newton; double e_c, s_c; };
// Diagnostics of the root solves, all scaled by h0: the roots (x_e, x_s) and
// the default bracket ends mu+ h0 (m_e, m_s); -1 when not computed.
struct C2PRoots { double x_e, x_s, m_e, m_s; };

GHD bool guess_ok(double lo, double hi) { return lo < hi && lo >= 0.0 && hi <= 1.0; }

// Sign change of a master function (negative left of its one root on
// (0, umax]) starting from [lo, hi].  While the bracket misses the root, up
// to nexp times, it is replaced by the adjacent interval on the root's side,
// with its log-width squared-and-doubled each time (factor g -> g^4), reusing
// the end already evaluated.  On success a, b, fa, fb hold the bracket.
template <typename F>
GHD bool guess_bracket(F& f, double lo, double hi, double umax, int nexp, double& a, double& b, double& fa,
                       double& fb) {
    fa = f(lo); fb = f(hi);
    double g = (lo > 0.0) ? hi / lo : 2.0;
    for (int k = 0;; ++k) {
        if (fa * fb < 0) { a = lo; b = hi; return true; }
        if (k == nexp || fa == 0.0 || fb == 0.0) return false;
        g = g * g; g = g * g;
        if (fb < 0) {                 // both ends left of the root
            if (hi >= umax) return false;
            lo = hi; fa = fb; hi = fmin(hi * g, umax); fb = f(hi);
        } else {                      // both ends right of the root
            hi = lo; fb = fa; lo = lo / g; fa = f(lo);
        }
    }
}

// Newton iteration for the one root of a master function on (0, umax], which
// is negative left of the root and positive right of it: every evaluation
// narrows the bracket [lo, hi] (initially (0, umax]), and a step that leaves
// it (or a non-finite step) is replaced by bisection.  It stops after a
// Newton step dx after which the new iterate's error |f''/(2 f')| dx^2
// (Taylor) is within acc of it: either |dx| <= NEWTON_TINY x, or
// |dx| <= NEWTON_RTOL x with f'' estimated from the slopes of the last two
// iterates (both reached by Newton steps).  Otherwise it iterates on, up to
// maxit evaluations.  The callers pass acc = NEWTON_ACCEPT / (1 + r^2/h0^2):
// a relative error e of mu moves W by up to W^2 e, and W^2 <= 1 + r^2/h0^2
// on the master function's domain.
constexpr double NEWTON_RTOL = 1e-7;
constexpr double NEWTON_TINY = 1e-10;
constexpr double NEWTON_ACCEPT = 1e-13;
template <typename F>
GHD bool newton_guess(F& f, double x, double umax, int maxit, double acc, double& root) {
    double lo = 0.0, hi = umax;
    if (!(x > lo && x < hi)) x = 0.5 * (lo + hi);
    double xp = 0.0, dfp = 0.0;
    bool prev_newton = false;
    for (int it = 0; it < maxit; ++it) {
        double fx, df;
        f.deriv(x, fx, df);
        if (fx == 0.0) { root = x; return true; }
        if (fx < 0.0) lo = x; else hi = x;
        const double dx = fx / df;
        double xn = x - dx;
        const bool newton = xn > lo && xn < hi;
        if (!newton) xn = 0.5 * (lo + hi);
        if (newton) {
            if (fabs(dx) <= NEWTON_TINY * xn) { root = xn; return true; }
            if (prev_newton && fabs(dx) <= NEWTON_RTOL * xn &&
                fabs(0.5 * (df - dfp) / ((x - xp) * df)) * dx * dx <= acc * xn) {
                root = xn;
                return true;
            }
        }
        prev_newton = newton;
        xp = x; dfp = df; x = xn;
    }
    return false;
}

template <typename EOS>
GHD double kastaun(const EOS& eos, const C2PIn& c, double pr[NPRIM], unsigned& sig, C2PCount& cnt,
                   const C2PGuess& gs, C2PRoots& rt) {
    double h0 = c.h0;
    FMu<EOS> f(eos, c.D, c.q, c.r2, c.rBt2, c.Bt2, h0, &cnt.root);
    double mu = -1.0;
    if constexpr (EOS::is_ideal) {
        double r;
        if (gs.newton > 0 && gs.e_c > 0.0 &&
            newton_guess(f, gs.e_c / h0, 1.0 / h0, gs.newton, NEWTON_ACCEPT / (1.0 + c.r2 / (h0 * h0)), r)) {
            mu = r; cnt.guess_used |= 1;
        }
    }
    if (mu < 0.0 && guess_ok(gs.e_lo, gs.e_hi)) {
        double a, b, fa, fb;
        if (guess_bracket(f, gs.e_lo / h0, gs.e_hi / h0, 1.0 / h0, gs.nexp, a, b, fa, fb)) {
            mu = brent_fab(f, a, b, fa, fb, 1e-15);
            cnt.guess_used |= 1;
        }
    }
    rt.m_e = -1.0;
    if (mu < 0.0) {
        double mu0 = 1 / h0;
        if (c.r2 >= h0 * h0) {
            FBr g{c.r2, c.Bt2, c.rBt2, h0 * h0, &cnt.bracket};
            double lo, hi; g.bracket(lo, hi);
            int rerr;
            mu0 = newton_br(lo, hi, g, 30, 1e-10, rerr);
            if (rerr == 0) mu0 *= 1 + 1e-10; else mu0 = 1.0 / h0;
        }
        mu = brent(f, 0., mu0, 1e-15);
        rt.m_e = mu0 * h0;
    }
    rt.x_e = mu * h0;
    double x, w, eps, rho, press, temp, ent;
    double res = f.prims(mu, sig, x, w, rho, press, eps, temp, ent);
    pr[EPSL] = eps; pr[RHOL] = rho; pr[PRESSL] = press; pr[TEMPL] = temp; pr[ENTL] = ent;
    pr[BXL] = c.B[0]; pr[BYL] = c.B[1]; pr[BZL] = c.B[2];
    for (int i = 0; i < 3; ++i) pr[ZXL + i] = w * mu * x * (c.r[i] + mu * c.rBt * c.Bt[i]);
    return fabs(res);
}

template <typename EOS>
GHD double kastaun_ent(const EOS& eos, const C2PIn& c, double pr[NPRIM], unsigned& sig, C2PCount& cnt,
                       const C2PGuess& gs, C2PRoots& rt) {
    double h0 = c.h0;
    FMuEnt<EOS> f(eos, c.D, c.r2, c.rBt2, c.Bt2, c.s, h0, &cnt.root);
    double mu = -1.0;
    if constexpr (EOS::is_ideal) {
        double r;
        if (gs.newton > 0 && gs.s_c > 0.0 &&
            newton_guess(f, gs.s_c / h0, 1.0 / h0, gs.newton, NEWTON_ACCEPT / (1.0 + c.r2 / (h0 * h0)), r)) {
            mu = r; cnt.guess_used |= 2;
        }
    }
    if (mu < 0.0 && guess_ok(gs.s_lo, gs.s_hi)) {
        double a, b, fa, fb;
        if (guess_bracket(f, gs.s_lo / h0, gs.s_hi / h0, 1.0 / h0, gs.nexp, a, b, fa, fb)) {
            mu = brent_fab(f, a, b, fa, fb, 1e-15);
            cnt.guess_used |= 2;
        }
    }
    rt.m_s = -1.0;
    if (mu < 0.0) {
        double mu0 = 1 / h0;
        if (c.r2 >= h0 * h0) {
            FBr g{c.r2, c.Bt2, c.rBt2, h0 * h0, &cnt.bracket};
            int rerr;
            mu0 = newton_br(0., 1. / h0, g, 30, 1e-10, rerr);
            if (rerr == 0) mu0 *= 1 + 1e-10; else mu0 = brent(g, 0., 1. / h0, 1e-10) * (1 + 1e-10);
        }
        mu = brent(f, 0., mu0, 1e-15);
        rt.m_s = mu0 * h0;
    }
    rt.x_s = mu * h0;
    double x, w, eps, rho, press, temp, ent;
    double res = f.prims(mu, sig, x, w, rho, press, eps, temp, ent);
    pr[EPSL] = eps; pr[RHOL] = rho; pr[PRESSL] = press; pr[TEMPL] = temp; pr[ENTL] = ent;
    pr[BXL] = c.B[0]; pr[BYL] = c.B[1]; pr[BZL] = c.B[2];
    for (int i = 0; i < 3; ++i) pr[ZXL + i] = w * mu * x * (c.r[i] + mu * c.rBt * c.Bt[i]);
    return fabs(res);
}

GHD void set_all_resets(uint64_t& e) {
    e |= BIT(E_RESET_DENS) | BIT(E_RESET_TAU) | BIT(E_RESET_STILDE) | BIT(E_RESET_ENTROPY) | BIT(E_RESET_YE);
}
// c2p.hh:160-229; returns false where the reference aborts
template <typename EOS>
GHD bool handle_sig(unsigned sig, bool lenient, uint64_t& err) {
    bool ok = true;
    const uint64_t core = BIT(E_RESET_DENS) | BIT(E_RESET_TAU) | BIT(E_RESET_STILDE) | BIT(E_RESET_ENTROPY)
                          | (EOS::has_ye ? BIT(E_RESET_YE) : 0u);
    if (sig & BIT(SIG_RHO_TOO_HIGH)) { err |= BIT(E_SIG_RHO_TOO_HIGH); if (lenient) set_all_resets(err); else ok = false; }
    if (sig & BIT(SIG_RHO_TOO_LOW)) err |= core | BIT(E_SIG_RHO_TOO_LOW);
    if (EOS::has_ye) {
        if (sig & BIT(SIG_YE_TOO_LOW)) err |= BIT(E_RESET_YE) | BIT(E_SIG_YE_TOO_LOW);
        if (sig & BIT(SIG_YE_TOO_HIGH)) err |= BIT(E_RESET_YE) | BIT(E_SIG_YE_TOO_HIGH);
    }
    if (sig & BIT(SIG_ENT_TOO_LOW)) err |= BIT(E_RESET_ENTROPY) | BIT(E_SIG_ENT_TOO_LOW);
    if (sig & BIT(SIG_ENT_TOO_HIGH)) err |= BIT(E_RESET_ENTROPY) | BIT(E_SIG_ENT_TOO_HIGH);
    if (sig & BIT(SIG_EPS_TOO_HIGH)) err |= BIT(E_RESET_STILDE) | BIT(E_RESET_TAU) | BIT(E_SIG_EPS_TOO_HIGH);
    if (sig & BIT(SIG_EPS_TOO_LOW)) err |= BIT(E_RESET_STILDE) | BIT(E_RESET_TAU) | BIT(E_SIG_EPS_TOO_LOW);
    if (sig & BIT(SIG_VEL_TOO_HIGH)) err |= core | BIT(E_SIG_VEL_TOO_HIGH);
    if (sig & BIT(SIG_SIGMA_TOO_HIGH)) err |= core | BIT(E_SIG_SIGMA_TOO_HIGH);
    return ok;
}
// c2p.hh:242-322
template <typename EOS>
GHD bool handle_eos_sig(unsigned e, bool lenient, uint64_t& err) {
    bool ok = true;
    const uint64_t core = BIT(E_RESET_DENS) | BIT(E_RESET_TAU) | BIT(E_RESET_STILDE) | BIT(E_RESET_ENTROPY)
                          | (EOS::has_ye ? BIT(E_RESET_YE) : 0u);
    const uint64_t mt = BIT(E_RESET_STILDE) | BIT(E_RESET_TAU);
    if (e & BIT(EOS_RHO_TOO_HIGH)) { err |= BIT(E_SIG_RHO_TOO_HIGH); ok = false; }
    if (e & BIT(EOS_RHO_TOO_LOW)) err |= core | BIT(E_SIG_RHO_TOO_LOW);
    if (EOS::has_ye) {
        if (e & BIT(EOS_YE_TOO_LOW)) err |= BIT(E_RESET_YE) | BIT(E_SIG_YE_TOO_LOW);
        if (e & BIT(EOS_YE_TOO_HIGH)) err |= BIT(E_RESET_YE) | BIT(E_SIG_YE_TOO_HIGH);
    }
    if (e & BIT(EOS_ENTROPY_TOO_HIGH)) err |= BIT(E_RESET_ENTROPY) | BIT(E_SIG_ENT_TOO_HIGH);
    if (e & BIT(EOS_ENTROPY_TOO_LOW)) err |= BIT(E_RESET_ENTROPY) | BIT(E_SIG_ENT_TOO_LOW);
    if (e & BIT(EOS_EPS_TOO_HIGH)) { err |= BIT(E_SIG_EPS_TOO_HIGH); if (lenient) err |= mt; else ok = false; }
    if (e & BIT(EOS_EPS_TOO_LOW)) err |= mt | BIT(E_SIG_EPS_TOO_LOW);
    if (e & BIT(EOS_TEMPERATURE_TOO_HIGH)) { err |= BIT(E_SIG_TEMP_TOO_HIGH); if (lenient) err |= mt; else ok = false; }
    if (e & BIT(EOS_TEMPERATURE_TOO_LOW)) err |= mt | BIT(E_SIG_TEMP_TOO_LOW);
    if (e & BIT(EOS_PRESS_TOO_HIGH)) { err |= BIT(E_SIG_PRESS_TOO_HIGH); if (lenient) err |= mt; else ok = false; }
    if (e & BIT(EOS_PRESS_TOO_LOW)) err |= mt | BIT(E_SIG_PRESS_TOO_LOW);
    return ok;
}

// prims_to_conservs with W and b^mu of pr already at hand
GHD void p2c_wb(const double pr[NPRIM], double cons[NCONS], const Met& m, double W, const double b[4], double b2) {
    const double* z = &pr[ZXL];
    double D, tau, S[3], Ss;
    conserved(m, pr[RHOL], pr[PRESSL], pr[EPSL], z, pr[ENTL], W, b2, b, D, tau, S, Ss);
    cons[DENSL] = m.sqrtg * D;
    cons[STXL] = m.sqrtg * S[0]; cons[STYL] = m.sqrtg * S[1]; cons[STZL] = m.sqrtg * S[2];
    cons[TAUL] = m.sqrtg * tau; cons[ENTSL] = m.sqrtg * Ss;
    cons[YESL] = cons[DENSL] * pr[YEL];
}
GHD void p2c(const double pr[NPRIM], double cons[NCONS], const Met& m) {
    double W = lorentz(m.g, &pr[ZXL]);
    double b[4], b2; smallb(m, &pr[BXL], &pr[ZXL], W, b, b2);
    p2c_wb(pr, cons, m, W, b, b2);
}

template <typename EOS>
GHD void atmo_reset(double cons[NCONS], double pr[NPRIM], const Met& m, const Atmo& at, const EOS& eos, uint64_t& err) {
    pr[RHOL] = at.rho_fl; pr[YEL] = at.ye_fl; pr[TEMPL] = at.temp_fl;
    pr[ZXL] = pr[ZYL] = pr[ZZL] = 0.;
    unsigned ee = 0; double cs2;
    pr[PRESSL] = eos.press_from_temp_ent(pr[EPSL], cs2, pr[ENTL], pr[TEMPL], pr[RHOL], ee);
    p2c(pr, cons, m);
    set_all_resets(err);
    err |= BIT(E_ATMO_RESET);
}

struct C2POut { bool floored, fatal; };

// conservs_to_prims (c2p.cpp:188-414).  cons densitised in, re-densitised out.
template <typename EOS>
GHD C2POut c2p(double cons[NCONS], double pr[NPRIM], const Met& m, const EOS& eos, const Atmo& at,
               const ExcPar& ex, const C2PPar& cp, double radius, uint64_t& err, bool dry_run, C2PCount& cnt,
               const C2PGuess& gs, C2PRoots& rt) {
    C2POut out{false, false};
    bool failed = false, floored = false;
    unsigned ret = 0;
    err = BIT(E_RESET_ENTROPY);
    double isg = 1.0 / m.sqrtg;
    (void)isg;
    for (int i = 0; i < NCONS; ++i) cons[i] /= m.sqrtg;
    pr[BXL] = cons[BSXL]; pr[BYL] = cons[BSYL]; pr[BZL] = cons[BSZL];
    if (cons[DENSL] < 0) { atmo_reset(cons, pr, m, at, eos, err); out.floored = true; return out; }
    pr[YEL] = cons[YESL] / cons[DENSL];
    if (pr[YEL] > eos.yemax()) { pr[YEL] = eos.yemax(); err |= BIT(E_RESET_YE); }
    if (pr[YEL] < eos.yemin()) { pr[YEL] = eos.yemin(); err |= BIT(E_RESET_YE); }
    bool lenient = m.alp < cp.alp_bh_thresh;
    {   // limit_conservatives (c2p.cpp:117-157)
        double emin, emax; eos.eps_range(emin, emax, cons[DENSL]);
        double B2 = m.sqv(cons[BSXL], cons[BSYL], cons[BSZL]);
        if (cons[TAUL] / cons[DENSL] - 0.5 * B2 / cons[DENSL] < fmin(0.0, emin)) {
            cons[TAUL] = emin * cons[DENSL] + 0.5 * B2;
            err |= BIT(E_RESET_TAU); floored = true;
        }
        double st2max = sq(cons[TAUL] / cons[DENSL] + 1);
        double st2 = m.sqcov(cons[STXL], cons[STYL], cons[STZL]) / sq(cons[DENSL]);
        if (st2 > st2max) {
            double f = sqrt(st2max / st2);
            cons[STXL] *= f; cons[STYL] *= f; cons[STZL] *= f;
            err |= BIT(E_RESET_STILDE); floored = true;
        }
    }
    if (floored && dry_run) { out.floored = true; return out; }
    if (cons[DENSL] >= at.rho_fl * (1 + at.atmo_tol)) {
        // both inversions start from the same invariants of the (limited)
        // conserved state, so they are formed once
        C2PIn ci; c2p_setup(eos, m, cons, ci);
        double res = kastaun(eos, ci, pr, ret, cnt, gs, rt);
        failed = (fabs(res) > cp.tol) || (ret & BIT(SIG_RHO_TOO_HIGH)) || (ret & BIT(SIG_EPS_TOO_HIGH)) || !isfinite(res);
        double beta;
        {
            double W = lorentz(m.g, &pr[ZXL]);
            double b[4], b2; smallb(m, &pr[BXL], &pr[ZXL], W, b, b2);
            beta = 2.0 * pr[PRESSL] / fmax(b2, 1e-50);
        }
        bool distrust = failed || (ret & BIT(SIG_EPS_TOO_LOW)) || (beta <= cp.beta_fallback);
        if (distrust && cp.use_ent_backup && !dry_run) {
            double save[NPRIM];
            for (int i = 0; i < NPRIM; ++i) save[i] = pr[i];
            unsigned bret = 0;
            C2PGuess gse = gs;
            if (gs.s_rel) {
                gse.s_lo = gs.s_lo * rt.x_e; gse.s_hi = fmin(gs.s_hi * rt.x_e, 1.0);
                gse.s_c = gs.s_c * rt.x_e;
            }
            double rb = kastaun_ent(eos, ci, pr, bret, cnt, gse, rt);
            bool ok = (fabs(rb) <= cp.tol) && isfinite(rb) && (pr[EPSL] < eos.eps_max);
            if (ok) { ret = bret; failed = false; err |= BIT(E_ENT_BACKUP_USED); }
            else if (!failed) { for (int i = 0; i < NPRIM; ++i) pr[i] = save[i]; }
        }
        if (ret & (BIT(SIG_EPS_TOO_HIGH) | BIT(SIG_RHO_TOO_LOW) | BIT(SIG_RHO_TOO_HIGH) | BIT(SIG_VEL_TOO_HIGH))) floored = true;
        if (EOS::is_ideal && (ret & BIT(SIG_EPS_TOO_LOW))) floored = true;
        if (!handle_sig<EOS>(ret, lenient, err)) out.fatal = true;
    } else {
        failed = true;
    }
    bool excise = ex.by_radius ? (radius <= ex.r_ex) : (m.alp <= ex.alp_ex);
    if ((pr[RHOL] < (1. + at.atmo_tol) * at.rho_fl) || failed || excise) {
        atmo_reset(cons, pr, m, at, eos, err);
        floored = true;
    } else if (pr[TEMPL] < at.temp_fl) {
        pr[TEMPL] = at.temp_fl;
        unsigned ee = 0; double cs2;
        pr[PRESSL] = eos.press_from_temp_ent(pr[EPSL], cs2, pr[ENTL], pr[TEMPL], pr[RHOL], ee);
        if (!handle_eos_sig<EOS>(ee, lenient, err)) out.fatal = true;
        err |= BIT(E_RESET_TAU) | BIT(E_RESET_STILDE) | BIT(E_RESET_ENTROPY) | BIT(E_T_FLOORED);
        if (EOS::is_ideal) floored = true;
    } else {   // limit_primitives (c2p.cpp:63-115)
        unsigned sg = 0;
        double z2max = sq(cp.max_w) - 1.;
        double z2 = m.sqv(pr[ZXL], pr[ZYL], pr[ZZL]);
        if (z2 >= z2max) {
            double zn = 0.99 * sqrt(z2max / z2);
            pr[ZXL] *= zn; pr[ZYL] *= zn; pr[ZZL] *= zn;
            sg |= BIT(SIG_VEL_TOO_HIGH);
        }
        const double W = lorentz(m.g, &pr[ZXL]);
        double b[4], b2; smallb(m, &pr[BXL], &pr[ZXL], W, b, b2);
        double sigma = b2 / pr[RHOL];
        if (sigma >= cp.max_sigma) {
            pr[RHOL] *= 1.001 * sigma / cp.max_sigma;
            unsigned ee = 0; double cs2;
            pr[PRESSL] = eos.press_from_temp_ent(pr[EPSL], cs2, pr[ENTL], pr[TEMPL], pr[RHOL], ee);
            sg |= BIT(SIG_SIGMA_TOO_HIGH);
        }
        if (!handle_sig<EOS>(sg, lenient, err)) out.fatal = true;
        if (sg & (BIT(SIG_VEL_TOO_HIGH) | BIT(SIG_SIGMA_TOO_HIGH))) floored = true;
        // the sigma limit changes rho and p only: W and b^mu still hold
        p2c_wb(pr, cons, m, W, b, b2);
        out.floored = floored;
        return out;
    }
    p2c(pr, cons, m);
    out.floored = floored;
    return out;
}

// ---- reconstruction ----------------------------------------------------------
// u0..u5 = u(i-3..i+2); uL at i-1/2-eps, uR at i-1/2+eps
// WENO-Z (weno_reconstruction.hh:172-278).  Per side, the three ratios
// t5/(b_k+eps) come from one reciprocal of d0 d1 d2 (d_k >= 1e-42, so the
// product cannot underflow), and the two sides of a face share their
// reciprocals: 1/(P_R P_L) gives 1/P_R and 1/P_L, 1/(S_R S_L) the two
// normalisations (guarded against overflow of the products, which falls back
// to separate divisions).  This differs from the reference's three divisions
// per weight by a few ulp.
struct WenoSide { double q0, q1, q2; double P; };
GHD void weno_betas(double b0, double b1, double b2, double& d0, double& d1, double& d2, double& t5) {
    const double eps = 1.0e-42;
    d0 = b0 + eps; d1 = b1 + eps; d2 = b2 + eps;
    t5 = fabs(b0 - b2);
}
// 1/b for a normal, finite b: the hardware reciprocal refined by two
// Newton-Raphson steps r <- r + r (1 - b r) (within 1 ulp of 1/b; about half
// the instructions of an IEEE division, which also handles zero, infinity and
// denormal operands).  Host builds divide.
GHD double nr_rcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
#else
    double r = 1.0 / b;
#endif
    double e = fma(-b, r, 1.0);
    r = fma(e, r, r);
    e = fma(-b, r, 1.0);
    return fma(e, r, r);
}
GHD void wenoz(double um3, double um2, double um1, double u0, double up1, double up2, double& uL, double& uR) {
    const double c0 = 13.0 / 12.0, c1 = 0.25;
    // right state (cell i stencils; a0 pairs with i2, a2 with i0 on this side)
    const double rb0 = c0 * sq(um2 + u0 - 2.0 * um1) + c1 * sq(um2 + 3.0 * u0 - 4.0 * um1);
    const double rb1 = c0 * sq(um1 + up1 - 2.0 * u0) + c1 * sq(um1 - up1);
    const double rb2 = c0 * sq(up2 + u0 - 2.0 * up1) + c1 * sq(up2 + 3.0 * u0 - 4.0 * up1);
    // left state (cell i-1 stencils)
    const double lb0 = c0 * sq(um3 + um1 - 2.0 * um2) + c1 * sq(um3 + 3.0 * um1 - 4.0 * um2);
    const double lb1 = c0 * sq(um2 + u0 - 2.0 * um1) + c1 * sq(um2 - u0);
    const double lb2 = c0 * sq(up1 + um1 - 2.0 * u0) + c1 * sq(up1 + 3.0 * um1 - 4.0 * u0);
    double rd0, rd1, rd2, rt, ld0, ld1, ld2, lt;
    weno_betas(rb2, rb1, rb0, rd0, rd1, rd2, rt);
    weno_betas(lb0, lb1, lb2, ld0, ld1, ld2, lt);
    const double rP01 = rd0 * rd1, rP = rP01 * rd2, lP01 = ld0 * ld1, lP = lP01 * ld2;
    double rr, lr;
    const double PP = rP * lP;
    // PP >= 1e-252 (every d >= 1e-42): a normal, finite reciprocal
    if (PP < 1e300) { const double inv = nr_rcp(PP); rr = rt * lP * inv; lr = lt * rP * inv; }
    else { rr = rt / rP; lr = lt / lP; }
    const double rq0 = rd1 * rd2 * rr, rq1 = rd0 * rd2 * rr, rq2 = rP01 * rr;
    const double lq0 = ld1 * ld2 * lr, lq1 = ld0 * ld2 * lr, lq2 = lP01 * lr;
    const double ra0 = 0.1 * (1.0 + rq0 * rq0), ra1 = 0.6 * (1.0 + rq1 * rq1), ra2 = 0.3 * (1.0 + rq2 * rq2);
    const double la0 = 0.1 * (1.0 + lq0 * lq0), la1 = 0.6 * (1.0 + lq1 * lq1), la2 = 0.3 * (1.0 + lq2 * lq2);
    const double rn = (2.0 * up2 - 7.0 * up1 + 11.0 * u0) * ra0 + (-1.0 * up1 + 5.0 * u0 + 2.0 * um1) * ra1 +
                      (2.0 * u0 + 5.0 * um1 - um2) * ra2;
    const double ln = (2.0 * um3 - 7.0 * um2 + 11.0 * um1) * la0 + (-1.0 * um2 + 5.0 * um1 + 2.0 * u0) * la1 +
                      (2.0 * um1 + 5.0 * u0 - up1) * la2;
    const double rS = 6.0 * (ra0 + ra1 + ra2), lS = 6.0 * (la0 + la1 + la2);
    const double SS = rS * lS;
    // SS >= 36 (every a >= its linear weight)
    if (SS < 1e300) { const double inv = nr_rcp(SS); uR = rn * lS * inv; uL = ln * rS * inv; }
    else { uR = rn / rS; uL = ln / lS; }
}
GHD double mc2(double sL, double sR) {
    double sC = 0.5 * (sR + sL), sg = copysign(1.0, sR);
    return sg * fmax(0., fmin(fmin(2. * fabs(sR), 2. * sg * sL), sg * sC));
}
GHD double minmod(double sL, double sR) {
    return 0.5 * (copysign(1.0, sL) + copysign(1.0, sR)) * fmin(fabs(sL), fabs(sR));
}
template <int REC>
GHD void recon(double um3, double um2, double um1, double u0, double up1, double up2, double& uL, double& uR) {
    if constexpr (REC == 0) {
        wenoz(um3, um2, um1, u0, up1, up2, uL, uR);
    } else if constexpr (REC == 3) {
        uL = um1; uR = u0;
    } else {
        double sL = u0 - um1, sR = up1 - u0;
        uR = u0 - 0.5 * (REC == 1 ? mc2(sL, sR) : minmod(sL, sR));
        sL = um1 - um2; sR = u0 - um1;
        uL = um1 + 0.5 * (REC == 1 ? mc2(sL, sR) : minmod(sL, sR));
    }
}
GHD double hll(double fL, double fR, double uL, double uR, double cmin, double cmax) {
    return (cmax * fL + cmin * fR - cmax * cmin * (uR - uL)) / (cmax + cmin);
}
// hll with the reciprocal of cmax + cmin shared by the fluxes of one face
GHD double hll_r(double fL, double fR, double uL, double uR, double cmin, double cmax, double inv) {
    return (cmax * fL + cmin * fR - cmax * cmin * (uR - uL)) * inv;
}

}  // namespace gm

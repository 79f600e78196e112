// grmhd_oracle.hh -- CPU oracle for the Cowling GRMHD RHS + Kastaun con2prim.
//
// TEST INFRASTRUCTURE ONLY.  This header is a plain C++ restatement of the
// reference algorithm (GRACE) for the hot path.  It is used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// The product library (grace_amd/csrc) never includes, links or calls it.
//
// Every function cites the reference file:line it restates.  Paths are
// relative to the reference checkout.
//
// Parity status per function (what pins it) is listed in the header comment
// of each section; every function has a check independent of its own
// formulas (golden vectors, closed forms, invariants, convergence).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <algorithm>

namespace oracle {

// ---------------------------------------------------------------------------
// Index enums.  Values follow include/grace/physics/grmhd_helpers.hh:104-149
// and include/grace/data_structures/variable_indices.hh:162-331.
// ---------------------------------------------------------------------------
enum PrimLoc { RHOL = 0, PRESSL, ZXL, ZYL, ZZL, YEL, TEMPL, EPSL, ENTL, BXL, BYL, BZL, CS2L, NPRIM };
enum ConsLoc { DENSL = 0, STXL, STYL, STZL, TAUL, YESL, ENTSL, BSXL, BSYL, BSZL, NCONS };
enum StateVar {
    DENS_ = 0, SX_, SY_, SZ_, TAU_, YESTAR_, ENTROPYSTAR_,
    GXX_, GXY_, GXZ_, GYY_, GYZ_, GZZ_, ALP_, BETAX_, BETAY_, BETAZ_,
    KXX_, KXY_, KXZ_, KYY_, KYZ_, KZZ_, NSTATE
};
constexpr int NHRSC = 7;
enum AuxVar {
    RHO_ = 0, ZVECX_, ZVECY_, ZVECZ_, BX_, BY_, BZ_, YE_, TEMP_, ENTROPY_, EPS_, PRESS_,
    BDIV_, C2P_DENS_ERR_, C2P_ERR_, NAUX
};

// c2p signal bits (include/grace/physics/eos/c2p.hh:44-56)
enum C2PSig {
    SIG_EPS_TOO_HIGH = 0, SIG_EPS_TOO_LOW, SIG_YE_TOO_LOW, SIG_YE_TOO_HIGH, SIG_RHO_TOO_HIGH,
    SIG_RHO_TOO_LOW, SIG_ENT_TOO_LOW, SIG_ENT_TOO_HIGH, SIG_VEL_TOO_HIGH, SIG_SIGMA_TOO_HIGH
};
// c2p error/outcome bits (include/grace/physics/eos/c2p.hh:95-125)
enum C2PErr {
    E_RESET_DENS = 0, E_RESET_TAU, E_RESET_STILDE, E_RESET_ENTROPY, E_RESET_YE,
    E_SIG_RHO_TOO_LOW, E_SIG_RHO_TOO_HIGH, E_SIG_EPS_TOO_LOW, E_SIG_EPS_TOO_HIGH,
    E_SIG_YE_TOO_LOW, E_SIG_YE_TOO_HIGH, E_SIG_ENT_TOO_LOW, E_SIG_ENT_TOO_HIGH,
    E_SIG_TEMP_TOO_LOW, E_SIG_TEMP_TOO_HIGH, E_SIG_PRESS_TOO_LOW, E_SIG_PRESS_TOO_HIGH,
    E_SIG_VEL_TOO_HIGH, E_SIG_SIGMA_TOO_HIGH, E_ENT_BACKUP_USED, E_ATMO_RESET, E_T_FLOORED
};
// EOS error bits (include/grace/physics/eos/eos_base.hh:40-54)
enum EosErr {
    EOS_RHO_TOO_HIGH = 0, EOS_RHO_TOO_LOW, EOS_EPS_TOO_HIGH, EOS_EPS_TOO_LOW, EOS_YE_TOO_LOW,
    EOS_YE_TOO_HIGH, EOS_TEMPERATURE_TOO_LOW, EOS_TEMPERATURE_TOO_HIGH, EOS_ENTROPY_TOO_LOW,
    EOS_ENTROPY_TOO_HIGH, EOS_PRESS_TOO_LOW, EOS_PRESS_TOO_HIGH
};

struct Bits {
    uint64_t w = 0;
    void set(int i) { w |= (uint64_t(1) << i); }
    bool test(int i) const { return (w >> i) & 1u; }
    void reset() { w = 0; }
};

// Optional operation counter, used to count residual evaluations of the
// root solves (iterations statistic of the c2p benchmarks).
struct C2PStats { int64_t n_eval = 0; int64_t n_bracket_eval = 0; };

inline double sqr(double x) { return x * x; }

// ---------------------------------------------------------------------------
// 3-metric helper.  Restates include/grace/utils/metric_utils.hh:48-82
// (determinant, inverse, sqrt(det)) and :180-230 (square_vec/covec, raise).
// Pinned by: inverse identity g*ginv = 1 (tests), Minkowski values
// (test/test_metric.cpp:68-186).
// ---------------------------------------------------------------------------
struct Metric {
    double g[6], gi[6], beta[3], alp, sqrtg;
    Metric() {}
    Metric(const double g_[6], const double b_[3], double a_) {
        for (int i = 0; i < 6; ++i) g[i] = g_[i];
        for (int i = 0; i < 3; ++i) beta[i] = b_[i];
        alp = a_;
        // det = gxx gyy gzz + 2 gxy gxz gyz - gxx gyz^2 - gyy gxz^2 - gzz gxy^2
        double det = -(sqr(g[2]) * g[3]) + 2 * g[1] * g[2] * g[4] - g[0] * sqr(g[4])
                     - sqr(g[1]) * g[5] + g[0] * g[3] * g[5];
        gi[0] = (g[3] * g[5] - sqr(g[4])) / det;
        gi[1] = (-g[1] * g[5] + g[2] * g[4]) / det;
        gi[2] = (-(g[2] * g[3]) + g[1] * g[4]) / det;
        gi[3] = (g[5] * g[0] - sqr(g[2])) / det;
        gi[4] = (g[1] * g[2] - g[0] * g[4]) / det;
        gi[5] = (-sqr(g[1]) + g[0] * g[3]) / det;
        sqrtg = std::sqrt(det);
    }
    double square_vec(const double v[3]) const {
        return g[0] * v[0] * v[0] + g[3] * v[1] * v[1] + g[5] * v[2] * v[2]
               + 2. * (g[1] * v[0] * v[1] + g[2] * v[0] * v[2] + g[4] * v[1] * v[2]);
    }
    double square_covec(const double v[3]) const {
        return gi[0] * v[0] * v[0] + gi[3] * v[1] * v[1] + gi[5] * v[2] * v[2]
               + 2. * (gi[1] * v[0] * v[1] + gi[2] * v[0] * v[2] + gi[4] * v[1] * v[2]);
    }
    void raise(const double v[3], double out[3]) const {
        out[0] = gi[0] * v[0] + gi[1] * v[1] + gi[2] * v[2];
        out[1] = gi[1] * v[0] + gi[3] * v[1] + gi[4] * v[2];
        out[2] = gi[2] * v[0] + gi[4] * v[1] + gi[5] * v[2];
    }
    void lower(const double v[3], double out[3]) const {
        out[0] = g[0] * v[0] + g[1] * v[1] + g[2] * v[2];
        out[1] = g[1] * v[0] + g[3] * v[1] + g[4] * v[2];
        out[2] = g[2] * v[0] + g[4] * v[1] + g[5] * v[2];
    }
};

// ---------------------------------------------------------------------------
// Ideal-gas EOS.  Restates include/grace/physics/eos/ideal_gas_eos.hh:48-435
// with the base-class bounds set at construction (:56-85; eos_base.hh:99-111).
// Pinned by: closed forms p=(G-1)rho eps, h=1+G eps, cs2=G(G-1)eps/h (tests).
// ---------------------------------------------------------------------------
struct IdealGas {
    static constexpr bool has_ye = false;
    static constexpr bool is_ideal = true;
    double gm1, gamma, ggm1, eps_max, eps_min, ent_min, temp_atm;
    double rhomax = 1e100, rhomin = 1e-100, tempmin = 0.0, yemax = 1e100, yemin = 0.0;
    double hmin = 1. + 1e-15;
    IdealGas() {}
    IdealGas(double gamma_, double temp_atm_, double eps_max_) {
        gm1 = gamma_ - 1.;
        gamma = gm1 + 1.;
        ggm1 = gamma * gm1;
        eps_max = eps_max_;
        eps_min = 0.0;
        temp_atm = temp_atm_;
        ent_min = (temp_atm_ > 0.0) ? (temp_atm_ / gm1) : 1e-100;
    }
    double enthalpy_minimum() const { return hmin; }
    double density_minimum() const { return rhomin; }
    double density_maximum() const { return rhomax; }
    double c2p_eps_max() const { return eps_max; }
    double c2p_ye_max() const { return yemax; }
    double c2p_ye_min() const { return yemin; }
    double temp_atmosphere() const { return temp_atm; }
    double ye_atmosphere() const { return 0.0; }

    double entropy__eps_rho(double eps, double rho) const { return std::log(eps) / gm1 - std::log(rho); }
    double eps__entropy_rho(double s, double rho) const { return std::exp(gm1 * (s + std::log(rho))); }
    double temp__eps(double eps) const { return eps * gm1; }
    double eps__temp(double t) const { return t / gm1; }
    double h__eps(double eps) const { return 1.0 + gamma * eps; }
    double csnd2__eps_h(double eps, double h) const { return ggm1 * eps / h; }
    double press__eps_rho(double eps, double rho) const { return gm1 * eps * rho; }
    void limit_eps(double& eps, Bits& e) const {
        if (eps > eps_max) { e.set(EOS_EPS_TOO_HIGH); eps = eps_max; }
        if (eps < eps_min) { e.set(EOS_EPS_TOO_LOW); eps = eps_min; }
    }
    void limit_temp(double& t, Bits& e) const {
        double tmax = temp__eps(eps_max);
        if (t < tempmin) { t = tempmin; e.set(EOS_TEMPERATURE_TOO_LOW); }
        if (t > tmax) { e.set(EOS_TEMPERATURE_TOO_HIGH); t = tmax; }
    }
    void limit_entropy(double& s, double rho, Bits& e) const {
        double smin = entropy__eps_rho(ent_min, rho), smax = entropy__eps_rho(eps_max, rho);
        if (s < smin) { s = smin; e.set(EOS_ENTROPY_TOO_LOW); }
        if (s > smax) { s = smax; e.set(EOS_ENTROPY_TOO_HIGH); }
    }
    void eps_range__rho_ye(double& emin, double& emax, double /*rho*/, Bits&) const { emin = 0; emax = eps_max; }
    double press__eps_rho_ye(double& eps, double rho, Bits& e) const { limit_eps(eps, e); return press__eps_rho(eps, rho); }
    double press_h_csnd2_temp_entropy__eps_rho_ye(double& h, double& cs2, double& temp, double& ent,
                                                  double& eps, double rho, Bits& e) const {
        limit_eps(eps, e);
        double p = press__eps_rho(eps, rho);
        h = h__eps(eps); cs2 = csnd2__eps_h(eps, h); temp = temp__eps(eps); ent = entropy__eps_rho(eps, rho);
        return p;
    }
    double press_h_csnd2_temp_eps__entropy_rho_ye(double& h, double& cs2, double& temp, double& eps,
                                                  double& ent, double rho, Bits& e) const {
        limit_entropy(ent, rho, e);
        eps = eps__entropy_rho(ent, rho);
        double p = press__eps_rho(eps, rho);
        h = h__eps(eps); temp = temp__eps(eps); cs2 = csnd2__eps_h(eps, h);
        return p;
    }
    double press_eps_csnd2_entropy__temp_rho_ye(double& eps, double& cs2, double& ent, double& temp,
                                                double rho, Bits& e) const {
        limit_temp(temp, e);
        eps = eps__temp(temp);
        double p = press__eps_rho(eps, rho);
        double h = h__eps(eps);
        cs2 = csnd2__eps_h(eps, h); ent = entropy__eps_rho(eps, rho);
        return p;
    }
    double press_eps_csnd2__temp_rho_ye(double& eps, double& cs2, double& temp, double rho, Bits& e) const {
        limit_temp(temp, e);
        eps = eps__temp(temp);
        double p = press__eps_rho(eps, rho);
        double h = h__eps(eps);
        cs2 = csnd2__eps_h(eps, h);
        return p;
    }
    double press_h_csnd2__temp_rho_ye(double& h, double& cs2, double& temp, double rho, Bits& e) const {
        limit_temp(temp, e);
        double eps = eps__temp(temp);
        double p = press__eps_rho(eps, rho);
        h = h__eps(eps); cs2 = csnd2__eps_h(eps, h);
        return p;
    }
    // ideal_gas_eos.hh:307-318 (pressure-reconstruction / initial-data hook)
    double eps_h_csnd2_temp_entropy__press_rho_ye(double& h, double& cs2, double& temp, double& ent,
                                                  double& press, double rho, Bits&) const {
        double eps = press / (gm1 * rho);
        h = h__eps(eps); cs2 = csnd2__eps_h(eps, h); ent = entropy__eps_rho(eps, rho); temp = temp__eps(eps);
        return eps;
    }
};

// ---------------------------------------------------------------------------
// Hybrid EOS over a piecewise polytrope.  Restates
// include/grace/physics/eos/hybrid_eos.hh:49-484 and
// include/grace/physics/eos/piecewise_polytropic_eos.hh:51-193; construction
// from src/physics/eos/eos_storage.cpp:52-99,164-186.  Only needed to run the
// reference's own c2p round-trip sweeps (test/test_c2p.cpp).
// Pinned by: those sweeps (tests/test_oracle_c2p.py).
// ---------------------------------------------------------------------------
struct HybridPwpoly {
    static constexpr bool has_ye = false;
    static constexpr bool is_ideal = false;
    int n = 1;
    double K[8], G[8], rh[8], ep[8], pr[8];
    double gth_m1, entropy_min, hmin, eps_max, temp_atm;
    double rhomax = 1e10, rhomin = 1e-100, tempmax = 1e99, tempmin = 0, yemax = 1e99, yemin = 0;
    HybridPwpoly() {}
    // Single-piece constructor (the configuration of test/configs/c2p_test_gamma2.yaml)
    HybridPwpoly(double K0, double G0, double gamma_th, double ent_min, double eps_max_, double temp_atm_) {
        n = 1; K[0] = K0; G[0] = G0; rh[0] = 0.; ep[0] = 0.; pr[0] = 0.;
        gth_m1 = gamma_th - 1.; entropy_min = ent_min; hmin = 1. + 1e-10;
        eps_max = eps_max_; temp_atm = temp_atm_;
    }
    double enthalpy_minimum() const { return hmin; }
    double density_minimum() const { return rhomin; }
    double density_maximum() const { return rhomax; }
    double c2p_eps_max() const { return eps_max; }
    double c2p_ye_max() const { return yemax; }
    double c2p_ye_min() const { return yemin; }
    double temp_atmosphere() const { return temp_atm; }
    double ye_atmosphere() const { return 0.0; }

    int find_index_rho(double& rho) const {
        rho = std::fmax(rhomin, std::fmin(rhomax, rho));
        for (int ii = 0; ii < n - 1; ++ii)
            if (rho > rh[ii] && rho <= rh[ii + 1]) return ii;
        return n - 1;
    }
    double press_cold_eps_cold__rho(double& eps_cold, double rho) const {
        int idx = find_index_rho(rho);
        double pc = K[idx] * std::pow(rho, G[idx]);
        eps_cold = ep[idx] + pc / (rho * (G[idx] - 1.));
        return pc;
    }
    double dpress_cold_drho__rho(double rho) const {
        int idx = find_index_rho(rho);
        double pc = K[idx] * std::pow(rho, G[idx]);
        return G[idx] * pc / rho;
    }
    void check_rho_range(double& rho, Bits& e) const {
        if (rho < rhomin) { e.set(EOS_RHO_TOO_LOW); rho = rhomin; }
        if (rho > rhomax) { e.set(EOS_RHO_TOO_HIGH); rho = rhomax; }
    }
    void check_eps_range(double& eps, double eps_cold, Bits& e) const {
        double emin = std::min(eps_cold, eps_max), emax = eps_max;
        if (eps < emin) { e.set(EOS_EPS_TOO_LOW); eps = emin; }
        if (eps > emax) { e.set(EOS_EPS_TOO_HIGH); eps = emax; }
    }
    void check_temp_range(double& t, Bits& e) const {
        if (t < tempmin) { e.set(EOS_TEMPERATURE_TOO_LOW); t = tempmin; }
        if (t > tempmax) { e.set(EOS_TEMPERATURE_TOO_HIGH); t = tempmax; }
    }
    double entropy__eps_th_rho(double eps_th, double rho) const {
        double e = std::max(eps_th, entropy_min);
        return std::log(e * std::pow(rho, -gth_m1)) / gth_m1;
    }
    double temp__eps_th(double eth) const { return std::fmax(0.0, eth * gth_m1); }
    double eps_th__temp(double t) const { return std::fmax(0.0, t / gth_m1); }
    void eps_range__rho_ye(double& emin, double& emax, double rho, Bits& e) const {
        check_rho_range(rho, e);
        double ec; press_cold_eps_cold__rho(ec, rho);
        emin = std::min(ec, eps_max); emax = eps_max;
    }
    void entropy_range__rho_ye(double& smin, double& smax, double rho, Bits& e) const {
        check_rho_range(rho, e);
        double ec; press_cold_eps_cold__rho(ec, rho);
        double ethmax = eps_max - ec;
        smin = entropy__eps_th_rho(entropy_min, rho);
        smax = entropy__eps_th_rho(ethmax, rho);
    }
    void check_entropy_range(double& s, double rho, Bits& e) const {
        double smin, smax; entropy_range__rho_ye(smin, smax, rho, e);
        if (s < smin) { e.set(EOS_ENTROPY_TOO_LOW); s = smin; }
        if (s > smax) { e.set(EOS_ENTROPY_TOO_HIGH); s = smax; }
    }
    double press__eps_rho_ye(double& eps, double rho, Bits& e) const {
        check_rho_range(rho, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        check_eps_range(eps, ec, e);
        return pc + (eps - ec) * rho * gth_m1;
    }
    double press_h_csnd2_temp_entropy__eps_rho_ye(double& h, double& cs2, double& temp, double& ent,
                                                  double& eps, double rho, Bits& e) const {
        check_rho_range(rho, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        check_eps_range(eps, ec, e);
        double eth = eps - ec;
        double p = pc + eth * rho * gth_m1;
        h = 1. + eps + p / rho;
        cs2 = (dpress_cold_drho__rho(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
        ent = entropy__eps_th_rho(eth, rho);
        temp = temp__eps_th(eth);
        return p;
    }
    double press_h_csnd2_temp_eps__entropy_rho_ye(double& h, double& cs2, double& temp, double& eps,
                                                  double& ent, double rho, Bits& e) const {
        check_rho_range(rho, e);
        check_entropy_range(ent, rho, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        double eth = std::exp(gth_m1 * ent) * std::pow(rho, gth_m1);
        double p = pc + eth * rho * gth_m1;
        temp = temp__eps_th(eth);
        eps = ec + eth;
        h = 1. + eps + p / rho;
        cs2 = (dpress_cold_drho__rho(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
        return p;
    }
    double press_eps_csnd2_entropy__temp_rho_ye(double& eps, double& cs2, double& ent, double& temp,
                                                double rho, Bits& e) const {
        check_rho_range(rho, e);
        check_temp_range(temp, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        double eth = eps_th__temp(temp);
        double p = pc + eth * rho * gth_m1;
        eps = eth + ec;
        double h = 1. + eps + p / rho;
        cs2 = (dpress_cold_drho__rho(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
        ent = entropy__eps_th_rho(eth, rho);
        return p;
    }
    double press_eps_csnd2__temp_rho_ye(double& eps, double& cs2, double& temp, double rho, Bits& e) const {
        check_rho_range(rho, e);
        check_temp_range(temp, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        double eth = eps_th__temp(temp);
        double p = pc + eth * rho * gth_m1;
        eps = ec + eth;
        double h = 1. + eps + p / rho;
        cs2 = (dpress_cold_drho__rho(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
        return p;
    }
    double press_h_csnd2__temp_rho_ye(double& h, double& cs2, double& temp, double rho, Bits& e) const {
        check_rho_range(rho, e);
        check_temp_range(temp, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        double eth = eps_th__temp(temp);
        double p = pc + eth * rho * gth_m1;
        h = 1. + ec + eth + p / rho;
        cs2 = (dpress_cold_drho__rho(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
        return p;
    }
    double eps_h_csnd2_temp_entropy__press_rho_ye(double& h, double& cs2, double& temp, double& ent,
                                                  double& press, double rho, Bits& e) const {
        check_rho_range(rho, e);
        double ec; double pc = press_cold_eps_cold__rho(ec, rho);
        press = std::max(press, pc);
        double eth = std::fmax(0, (press - pc) / (rho * gth_m1));
        double eps = eth + ec;
        h = 1. + eps + press / rho;
        cs2 = (dpress_cold_drho__rho(rho) + gth_m1 * (gth_m1 + 1) * eth) / h;
        ent = entropy__eps_th_rho(eth, rho);
        temp = temp__eps_th(eth);
        return eps;
    }
};

// ---------------------------------------------------------------------------
// Parameter blocks.  Restate include/grace/physics/grmhd_helpers.hh:41-97;
// defaults from parameters/grmhd.yaml (atmosphere :27-61, excision :73-85,
// c2p :96-125, riemann :136).
// ---------------------------------------------------------------------------
struct AtmoParams { double ye_fl = 0, rho_fl = 1e-14, temp_fl = 0, atmo_tol = 1e-3; };
struct C2PParams {
    double tol = 1e-10, max_w = 50, max_sigma = 100, beta_fallback = 1e-2;
    bool use_ent_backup = true;
    double alp_bh_thresh = 0.2;
};
struct ExcisionParams {
    double r_ex = 1.0, alp_ex = 0.05;
    bool excise_by_radius = false;
    bool cks = false;        // Cartesian Kerr-Schild coordinates: spheroidal radius
    double spin = 0.0;
};

// rtp[0] of cart_to_sph (include/grace/coordinates/cartesian_coordinate_systems.hh:87-107)
inline double sph_radius(const ExcisionParams& e, const double xyz[3]) {
    const double rad = std::sqrt(xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2]);
    if (!e.cks) return rad;
    const double a = e.spin;
    return std::fmax(std::sqrt(rad * rad - a * a + std::sqrt((rad * rad - a * a) * (rad * rad - a * a)
                                                             + 4.0 * a * a * xyz[2] * xyz[2])) / std::sqrt(2.0),
                     1.0e-6);
}

// ---------------------------------------------------------------------------
// GRMHD algebra (Valencia variables, z^i = W v^i).  The reference generates
// these with SymPy (include/grace/physics/grmhd_subexpressions.hh:10-392); they
// are re-derived here from the stress-energy tensor.
//   W, b^mu, b^2            : grmhd_subexpressions.hh:10-44
//     pinned by test/test_grmhd_kernels.cpp:80-116 closed forms
//   vtilde^i = alp z^i/W - beta^i : :46-58, pinned by :119-160
//   conserved (D,S_i,tau,S*) : :318-350, pinned by :163-215 closed forms
//   fluxes                   : :97-156, pinned by :218-270 (hydro); B != 0
//                              and v != 0 in generic curved metrics pinned
//                              against fluxes built from T^{mu nu} with numpy
//                              (tests/test_oracle_pins.py)
//   cm/cp                    : :60-95 (Gammie-style fast-speed quadratic),
//                              pinned by :273-318 (hydro +-c_s); B != 0 by the
//                              roots of the covariant dispersion relation and
//                              relativistic velocity addition
//                              (tests/test_oracle_pins.py)
//   geometric sources        : :158-316, pinned by the two golden vectors of
//                              test/test_grmhd_kernels.cpp:393-469 (rel 1e-12)
// ---------------------------------------------------------------------------
inline double get_W(const double g[6], const double z[3]) {
    double zl0 = g[0] * z[0] + g[1] * z[1] + g[2] * z[2];
    double zl1 = g[1] * z[0] + g[3] * z[1] + g[4] * z[2];
    double zl2 = g[2] * z[0] + g[4] * z[1] + g[5] * z[2];
    return std::sqrt(z[0] * zl0 + z[1] * zl1 + z[2] * zl2 + 1);
}

// b^mu (4 comps) and b^2 given alp, beta, gamma_ij, B^i (undensitised), z^i, W.
inline void get_smallb(double alp, const double beta[3], const double g[6], const double B[3],
                       const double z[3], double W, double b[4], double& b2) {
    double zl0 = g[0] * z[0] + g[1] * z[1] + g[2] * z[2];
    double zl1 = g[1] * z[0] + g[3] * z[1] + g[4] * z[2];
    double zl2 = g[2] * z[0] + g[4] * z[1] + g[5] * z[2];
    double Bz = B[0] * zl0 + B[1] * zl1 + B[2] * zl2;     // B^i z_i = alp W b^0 / W ... (= alp b^0)
    double b0 = Bz / alp;
    double iW = 1. / W;
    // b^i = (B^i + alp b^0 u^i)/W with u^i = z^i - W beta^i/alp
    for (int i = 0; i < 3; ++i) b[1 + i] = iW * (B[i] + Bz * (z[i] - W * beta[i] / alp));
    b[0] = b0;
    double Bl0 = g[0] * B[0] + g[1] * B[1] + g[2] * B[2];
    double Bl1 = g[1] * B[0] + g[3] * B[1] + g[4] * B[2];
    double Bl2 = g[2] * B[0] + g[4] * B[1] + g[5] * B[2];
    double B2 = B[0] * Bl0 + B[1] * Bl1 + B[2] * Bl2;
    b2 = (B2 + Bz * Bz) / (W * W);
}

inline void get_vtilde(double alp, const double beta[3], const double z[3], double W, double vt[3]) {
    double iW = 1. / W;
    for (int i = 0; i < 3; ++i) vt[i] = alp * iW * z[i] - beta[i];
}

// Fast magnetosonic speeds along direction d: roots of
//   A l^2 + B l + C = 0,  A = (u^0)^2 - (g^00 + (u^0)^2) w2,
//   B = -2 (u^0 u^d - (g^0d + u^0 u^d) w2),  C = (u^d)^2 - (g^dd + (u^d)^2) w2,
// with w2 = vA^2 + cs^2 (1 - vA^2), vA^2 = b^2/(rho h + b^2), u^0 = W/alp,
// u^d = W vtilde^d/alp  (grmhd_subexpressions.hh:60-95).
inline void get_cm_cp(double alp, const double beta[3], double rho, double press, double eps,
                      double cs2, double W, double b2, const double vt[3], double gdd_inv, int d,
                      double& cm, double& cp) {
    double ia2 = 1. / (alp * alp);
    double vA2 = b2 / (b2 + eps * rho + press + rho);
    double w2 = vA2 + cs2 * (1. - vA2);
    double W2 = W * W;
    double W2_1mw2 = W2 * (1. - w2);
    double A = (W2_1mw2 + w2) * ia2;
    double Bc = 2. * ia2 * (beta[d] * w2 - vt[d] * W2_1mw2);
    double g4dd = gdd_inv - beta[d] * beta[d] * ia2;
    double C = vt[d] * vt[d] * W2_1mw2 * ia2 - w2 * g4dd;
    double disc = std::sqrt(std::fmax(0., Bc * Bc - 4. * A * C));
    double l1 = (-Bc + disc) / (2. * A);
    double l2 = (-Bc - disc) / (2. * A);
    cm = std::fmin(l1, l2);
    cp = std::fmax(l1, l2);
}

// Conserved variables (undensitised) D, tau, S_i, S* = s D.
inline void get_conserved(double alp, const double beta[3], const double g[6], double rho,
                          double press, double eps, const double z[3], double s, double W,
                          double b2, const double b[4], double& D, double& tau, double S[3],
                          double& Sstar) {
    double rhoh = b2 + eps * rho + press + rho;   // rho h + b^2
    D = W * rho;
    tau = W * W * rhoh - press - 0.5 * b2 - alp * alp * b[0] * b[0] - D;
    double zl[3], bl[3], bsh[3];
    double tmp[3] = {b[1], b[2], b[3]};
    double Mz[3]; for (int i = 0; i < 3; ++i) Mz[i] = z[i];
    (void)Mz;
    // z_j lowered with gamma; b_j = beta_j b^0 + gamma_jk b^k
    zl[0] = g[0] * z[0] + g[1] * z[1] + g[2] * z[2];
    zl[1] = g[1] * z[0] + g[3] * z[1] + g[4] * z[2];
    zl[2] = g[2] * z[0] + g[4] * z[1] + g[5] * z[2];
    bsh[0] = g[0] * beta[0] + g[1] * beta[1] + g[2] * beta[2];
    bsh[1] = g[1] * beta[0] + g[3] * beta[1] + g[4] * beta[2];
    bsh[2] = g[2] * beta[0] + g[4] * beta[1] + g[5] * beta[2];
    bl[0] = g[0] * tmp[0] + g[1] * tmp[1] + g[2] * tmp[2] + b[0] * bsh[0];
    bl[1] = g[1] * tmp[0] + g[3] * tmp[1] + g[4] * tmp[2] + b[0] * bsh[1];
    bl[2] = g[2] * tmp[0] + g[4] * tmp[1] + g[5] * tmp[2] + b[0] * bsh[2];
    for (int j = 0; j < 3; ++j) S[j] = rhoh * W * zl[j] - alp * b[0] * bl[j];
    Sstar = s * D;
}

// Conserved + fluxes along direction d (undensitised).
inline void get_fluxes(double alp, const double beta[3], const double g[6], double rho, double press,
                       double eps, const double z[3], double s, double W, double b2, const double b[4],
                       const double vt[3], int d, double& D, double& tau, double S[3], double& Sstar,
                       double& fD, double& ftau, double fS[3], double& fSstar) {
    get_conserved(alp, beta, g, rho, press, eps, z, s, W, b2, b, D, tau, S, Sstar);
    double rhoh = b2 + eps * rho + press + rho;
    double pstar = 0.5 * b2 + press;
    double W2rhoh = W * W * rhoh;
    fD = vt[d] * D;
    ftau = W2rhoh * vt[d] + pstar * beta[d] - alp * alp * b[0] * b[1 + d] - fD;
    double bsh[3], bl[3], vl[3];
    bsh[0] = g[0] * beta[0] + g[1] * beta[1] + g[2] * beta[2];
    bsh[1] = g[1] * beta[0] + g[3] * beta[1] + g[4] * beta[2];
    bsh[2] = g[2] * beta[0] + g[4] * beta[1] + g[5] * beta[2];
    bl[0] = g[0] * b[1] + g[1] * b[2] + g[2] * b[3] + b[0] * bsh[0];
    bl[1] = g[1] * b[1] + g[3] * b[2] + g[4] * b[3] + b[0] * bsh[1];
    bl[2] = g[2] * b[1] + g[4] * b[2] + g[5] * b[3] + b[0] * bsh[2];
    vl[0] = g[0] * vt[0] + g[1] * vt[1] + g[2] * vt[2] + bsh[0];
    vl[1] = g[1] * vt[0] + g[3] * vt[1] + g[4] * vt[2] + bsh[1];
    vl[2] = g[2] * vt[0] + g[4] * vt[1] + g[5] * vt[2] + bsh[2];
    double c = vt[d] * W2rhoh / (alp * alp);
    for (int j = 0; j < 3; ++j)
        fS[j] = alp * (-b[1 + d] * bl[j] + (j == d ? pstar : 0.0) + c * vl[j]);
    fSstar = vt[d] * Sstar;
}

// Geometric source terms (without the sqrt(gamma) factor).
//   S_tau = alp [ (T^00 b^i b^j + 2 T^0i b^j + T^ij) K_ij - (T^00 b^i + T^0i) d_i alp ]
//   S_Sk  = alp [ T^00 (1/2 b^i b^j d_k g_ij - alp d_k alp) + T^0i b^j d_k g_ij
//                 + T^0_i d_k b^i + 1/2 T^ij d_k g_ij ]
// (b = shift here).  dg[a + 6 k] = d_k g_a, db[c + 3 k] = d_k beta^c.
inline void get_geom_sources(double alp, const double beta[3], const double g[6], const double gu[6],
                             const double K[6], const double da[3], const double dg[18],
                             const double db[9], double rho, double press, double eps,
                             const double B[3], const double z[3], double W, double& stau,
                             double ss[3]) {
    double b[4], b2;
    get_smallb(alp, beta, g, B, z, W, b, b2);
    double rhoh = b2 + eps * rho + press + rho;
    double pstar = press + 0.5 * b2;
    double ia2 = 1. / (alp * alp);
    double u[4];
    u[0] = W / alp;
    for (int i = 0; i < 3; ++i) u[1 + i] = z[i] - W * beta[i] / alp;
    // 4-inverse metric
    double g4[4][4];
    g4[0][0] = -ia2;
    for (int i = 0; i < 3; ++i) { g4[0][1 + i] = beta[i] * ia2; g4[1 + i][0] = beta[i] * ia2; }
    const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) g4[1 + i][1 + j] = gu[sidx[i][j]] - beta[i] * beta[j] * ia2;
    double T[4][4];
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) T[m][n] = rhoh * u[m] * u[n] + pstar * g4[m][n] - b[m] * b[n];
    double Kb[3], tK = 0;
    for (int i = 0; i < 3; ++i) Kb[i] = K[sidx[i][0]] * beta[0] + K[sidx[i][1]] * beta[1] + K[sidx[i][2]] * beta[2];
    double bKb = beta[0] * Kb[0] + beta[1] * Kb[1] + beta[2] * Kb[2];
    double T0K = 0, TK = 0, Tda = 0;
    for (int i = 0; i < 3; ++i) {
        T0K += T[0][1 + i] * Kb[i];
        for (int j = 0; j < 3; ++j) TK += T[1 + i][1 + j] * K[sidx[i][j]];
        Tda += (T[0][0] * beta[i] + T[0][1 + i]) * da[i];
    }
    (void)tK;
    stau = alp * ((T[0][0] * bKb + 2. * T0K + TK) - Tda);
    double bsh[3] = {g[0] * beta[0] + g[1] * beta[1] + g[2] * beta[2],
                     g[1] * beta[0] + g[3] * beta[1] + g[4] * beta[2],
                     g[2] * beta[0] + g[4] * beta[1] + g[5] * beta[2]};
    double T0l[3];
    for (int i = 0; i < 3; ++i)
        T0l[i] = bsh[i] * T[0][0] + g[sidx[i][0]] * T[0][1] + g[sidx[i][1]] * T[0][2] + g[sidx[i][2]] * T[0][3];
    for (int k = 0; k < 3; ++k) {
        const double* dgk = dg + 6 * k;
        double bdgb = 0, T0dgb = 0, Tdg = 0, T0db = 0;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                double dgij = dgk[sidx[i][j]];
                bdgb += beta[i] * beta[j] * dgij;
                T0dgb += T[0][1 + i] * beta[j] * dgij;
                Tdg += T[1 + i][1 + j] * dgij;
            }
            T0db += T0l[i] * db[i + 3 * k];
        }
        ss[k] = alp * (T[0][0] * (0.5 * bdgb - alp * da[k]) + T0dgb + T0db + 0.5 * Tdg);
    }
}

// ---------------------------------------------------------------------------
// Root finders.  Restate include/grace/utils/rootfinding.hh:73-231 (Brent's
// zero-in, Brent 1973, "Algorithms for Minimization Without Derivatives",
// ch. 4; no sign check on the bracket, as in the reference) and :255-306
// (Newton-Raphson safeguarded by bisection on a bracket).
// Pinned by: closed-form roots and agreement with bisection (tests).
// ---------------------------------------------------------------------------
template <typename F>
inline double brent_with_ends(F& f, double a, double b, double fa, double fb, double t) {
    double sa = a, sb = b;
    double c = sa, fc = fa;
    double e = sb - sa, d = e;
    const double macheps = std::numeric_limits<double>::epsilon();
    // the reference loops without a cap; finite inputs never reach this one,
    // which only ends the loop on NaN input
    for (int it = 0; it < 1000; ++it) {
        if (std::fabs(fc) < std::fabs(fb)) {
            sa = sb; sb = c; c = sa;
            fa = fb; fb = fc; fc = fa;
        }
        double tol = 2.0 * macheps * std::fabs(sb) + t;
        double m = 0.5 * (c - sb);
        if (std::fabs(m) <= tol || fb == 0.0) break;
        if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) {
            e = m; d = e;                        // bisection
        } else {
            double s = fb / fa, p, q;
            if (sa == c) {                       // secant
                p = 2.0 * m * s;
                q = 1.0 - s;
            } else {                             // inverse quadratic interpolation
                double qq = fa / fc, r = fb / fc;
                p = s * (2.0 * m * qq * (qq - r) - (sb - sa) * (r - 1.0));
                q = (qq - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (0.0 < p) q = -q; else p = -p;
            s = e; e = d;
            if (2.0 * p < 3.0 * m * q - std::fabs(tol * q) && p < std::fabs(0.5 * s * q)) {
                d = p / q;
            } else {
                e = m; d = e;
            }
        }
        sa = sb; fa = fb;
        if (tol < std::fabs(d)) sb = sb + d;
        else if (0.0 < m) sb = sb + tol;
        else sb = sb - tol;
        fb = f(sb);
        if ((0.0 < fb && 0.0 < fc) || (fb <= 0.0 && fc <= 0.0)) {
            c = sa; fc = fa; e = sb - sa; d = e;
        }
    }
    return sb;
}
template <typename F>
inline double brent(F& f, double a, double b, double t) {
    double fa = f(a), fb = f(b);
    return brent_with_ends(f, a, b, fa, fb, t);
}

template <typename F>
inline double newton_bracketed(double sa, double sb, F& f, unsigned long maxiter, double tol, int& err) {
    const double macheps = std::numeric_limits<double>::epsilon();
    double a = sa, b = sb, fa, fb, dummy;
    f(a, fa, dummy);
    f(b, fb, dummy);
    if (fa == 0) { err = 0; return a; }
    if (fb == 0) { err = 0; return b; }
    if (fa * fb > 0) { err = 1; return 0; }
    double x = (fb * a - fa * b) / (fb - fa);
    unsigned long iter = 0;
    double xold, fx, dfx, t;
    do {
        xold = x;
        f(x, fx, dfx);
        if (fx * fb > 0) { fb = fx; b = x; }
        else if (fx * fa > 0) { fa = fx; a = x; }
        x -= fx / dfx;
        t = 2. * macheps * std::fabs(x) + tol;
        if (std::fabs(x - xold) < t || fx == 0) { err = 0; return x; }
        if (x > b || x < a) x = 0.5 * (b + a);
        iter++;
    } while (iter < maxiter);
    err = 1;
    return 0;
}

// ---------------------------------------------------------------------------
// Kastaun et al. (2021) primitive recovery.  Restates
// include/grace/physics/eos/kastaun_c2p.hh:44-402 (fbrack_t, froot_t,
// kastaun_c2p_t) and ent_based_c2p.hh:43-256 (entropy backup).
// Pinned by: test/test_c2p.cpp round-trip sweeps; agreement with an
// independent bisection on the master function (tests).
// ---------------------------------------------------------------------------
struct FBrack {  // kastaun_c2p.hh:44-115
    double rsqr, bsqr, rbsqr, h0sqr;
    C2PStats* st = nullptr;
    FBrack(double r2, double b2, double rb2, double h0) : rsqr(r2), bsqr(b2), rbsqr(rb2), h0sqr(h0 * h0) {}
    double x__mu(double mu) const { return 1. / (1. + mu * bsqr); }
    double rbsq__mu_x(double mu, double x) const { return x * (rsqr * x + mu * (x + 1.0) * rbsqr); }
    void operator()(double mu, double& f, double& df) {
        if (st) st->n_bracket_eval++;
        double x = x__mu(mu), xsqr = x * x;
        double hw = std::sqrt(h0sqr + rbsq__mu_x(mu, x));
        double bb = x * (xsqr * rsqr + mu * (1 + x + xsqr) * rbsqr);
        f = mu * hw - 1.;
        df = (h0sqr + bb) / hw;
    }
    double operator()(double mu) {
        if (st) st->n_bracket_eval++;
        double x = x__mu(mu);
        return mu * std::sqrt(h0sqr + rbsq__mu_x(mu, x)) - 1.;
    }
    void bracket(double& mu_min, double& mu_max) const {
        mu_min = 1. / (h0sqr + rsqr);
        double mu0 = 1. / std::sqrt(h0sqr);
        double rfsqr_min = rbsq__mu_x(mu0, x__mu(mu0));
        mu_max = 1.0 / std::sqrt(h0sqr + rfsqr_min);
        mu_min *= (1 - 1e-10);
        mu_max *= (1 + 1e-10);
        if (mu_max <= mu_min) { mu_min = 0; mu_max = mu0 * (1.0 + 1e-10); }
    }
};

template <typename EOS>
struct FRoot {  // kastaun_c2p.hh:120-301
    const EOS& eos;
    double d, qtot, ye, rsqr, bsqr, rbsqr, brosqr, vsqrmax, wmax;
    C2PStats* st = nullptr;
    FRoot(const EOS& e, double d_, double q_, double r2, double rb2, double b2, double ye_, double h0)
        : eos(e), d(d_), qtot(q_), ye(ye_), rsqr(r2), bsqr(b2), rbsqr(rb2), brosqr(r2 * b2 - rb2) {
        double zsqrmax = rsqr / (h0 * h0);
        double wsqrmax = 1 + zsqrmax;
        wmax = std::sqrt(wsqrmax);
        vsqrmax = zsqrmax / wsqrmax;
    }
    double x__mu(double mu) const { return 1. / (1. + mu * bsqr); }
    double rfsqr__mu_x(double mu, double x) const { return x * (rsqr * x + mu * (x + 1.0) * rbsqr); }
    double qf__mu_x(double mu, double x) const { double mux = mu * x; return qtot - 0.5 * (bsqr + mux * mux * brosqr); }
    double eps_raw(double mu, double qf, double rfsqr, double w) const {
        return w * (qf - mu * rfsqr * (1.0 - mu * w / (1 + w)));
    }
    double compute_primitives(double mu, Bits& err, double& x, double& w, double& rho, double& press,
                              double& eps, double& temp, double& entropy) {
        x = x__mu(mu);
        const double rfsqr = rfsqr__mu_x(mu, x);
        const double qf = qf__mu_x(mu, x);
        double vsqr = rfsqr * mu * mu;
        if (vsqr > vsqrmax) { vsqr = vsqrmax; w = wmax; err.set(SIG_VEL_TOO_HIGH); }
        else w = 1 / std::sqrt(1 - vsqr);
        double rhomax = eos.density_maximum(), rhomin = eos.density_minimum();
        rho = d / w;
        if (rho >= rhomax) { err.set(SIG_RHO_TOO_HIGH); rho = rhomax; }
        else if (rho <= rhomin) { err.set(SIG_RHO_TOO_LOW); rho = rhomin; }
        eps = eps_raw(mu, qf, rfsqr, w);
        double epsmin, epsmax;
        Bits ee;
        eos.eps_range__rho_ye(epsmin, epsmax, rho, ee);
        epsmax = std::fmin(epsmax, eos.c2p_eps_max());
        if (eps >= epsmax) { err.set(SIG_EPS_TOO_HIGH); eps = epsmax * 0.999; }
        else if (eps < epsmin) { err.set(SIG_EPS_TOO_LOW); eps = epsmin; }
        double hh, cs2;
        Bits eerr;
        press = eos.press_h_csnd2_temp_entropy__eps_rho_ye(hh, cs2, temp, entropy, eps, rho, eerr);
        double a = press / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        double hbw_raw = (1 + a) * (1 + qf - mu * rfsqr);
        double hbw = std::fmax(hbw_raw, h / w);
        double newmu = 1. / (hbw + rfsqr * mu);
        return mu - newmu;
    }
    double operator()(double mu) {
        if (st) st->n_eval++;
        const double x = x__mu(mu);
        const double rfsqr = rfsqr__mu_x(mu, x);
        const double qf = qf__mu_x(mu, x);
        double vsqr = rfsqr * mu * mu, w;
        if (vsqr > vsqrmax) { vsqr = vsqrmax; w = wmax; }
        else w = 1 / std::sqrt(1 - vsqr);
        double rho = std::fmin(eos.density_maximum(), std::fmax(eos.density_minimum(), d / w));
        double eps = eps_raw(mu, qf, rfsqr, w);
        double epsmin, epsmax;
        Bits ee;
        eos.eps_range__rho_ye(epsmin, epsmax, rho, ee);
        eps = std::fmin(epsmax, std::fmax(epsmin, eps));
        Bits eerr;
        double press = eos.press__eps_rho_ye(eps, rho, eerr);
        double a = press / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        double hbw_raw = (1 + a) * (1 + qf - mu * rfsqr);
        double hbw = std::fmax(hbw_raw, h / w);
        double newmu = 1. / (hbw + rfsqr * mu);
        return mu - newmu;
    }
};

template <typename EOS>
struct EntFRoot {  // ent_based_c2p.hh:43-160
    const EOS& eos;
    double d, s, ye, rsqr, bsqr, rbsqr, vsqrmax, wmax;
    C2PStats* st = nullptr;
    EntFRoot(const EOS& e, double d_, double r2, double rb2, double b2, double s_, double ye_, double h0)
        : eos(e), d(d_), s(s_), ye(ye_), rsqr(r2), bsqr(b2), rbsqr(rb2) {
        double zsqrmax = rsqr / (h0 * h0);
        double wsqrmax = 1 + zsqrmax;
        wmax = std::sqrt(wsqrmax);
        vsqrmax = zsqrmax / wsqrmax;
    }
    double x__mu(double mu) const { return 1. / (1. + mu * bsqr); }
    double rfsqr__mu_x(double mu, double x) const { return x * (rsqr * x + mu * (x + 1.0) * rbsqr); }
    double operator()(double mu) {
        if (st) st->n_eval++;
        const double x = x__mu(mu);
        const double rfsqr = rfsqr__mu_x(mu, x);
        double vsqr = rfsqr * mu * mu;
        double w = 1 / std::sqrt(1 - vsqr);
        if (vsqr > vsqrmax) { vsqr = vsqrmax; w = wmax; }
        double rho = std::fmin(eos.density_maximum(), std::fmax(eos.density_minimum(), d / w));
        double hh, cs2, temp, eps;
        double sl = s;
        Bits eerr;
        double press = eos.press_h_csnd2_temp_eps__entropy_rho_ye(hh, cs2, temp, eps, sl, rho, eerr);
        double a = press / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        double hbw = h / w;
        double newmu = 1. / (hbw + rfsqr * mu);
        return mu - newmu;
    }
    double compute_primitives(double mu, Bits& err, double& x, double& w, double& rho, double& press,
                              double& eps, double& temp, double& entropy) {
        entropy = s;
        x = x__mu(mu);
        const double rfsqr = rfsqr__mu_x(mu, x);
        double vsqr = rfsqr * mu * mu;
        w = 1 / std::sqrt(1 - vsqr);
        if (vsqr > vsqrmax) { vsqr = vsqrmax; w = wmax; err.set(SIG_VEL_TOO_HIGH); }
        double rhomax = eos.density_maximum(), rhomin = eos.density_minimum();
        rho = d / w;
        if (rho >= rhomax) { err.set(SIG_RHO_TOO_HIGH); rho = rhomax; }
        else if (rho <= rhomin) { err.set(SIG_RHO_TOO_LOW); rho = rhomin; }
        double hh, cs2;
        Bits eerr;
        press = eos.press_h_csnd2_temp_eps__entropy_rho_ye(hh, cs2, temp, eps, entropy, rho, eerr);
        if (eerr.test(EOS_ENTROPY_TOO_LOW)) err.set(SIG_ENT_TOO_LOW);
        if (eerr.test(EOS_ENTROPY_TOO_HIGH)) err.set(SIG_ENT_TOO_HIGH);
        double a = press / (rho * (1 + eps));
        double h = (1 + eps) * (1 + a);
        double hbw = h / w;
        double newmu = 1. / (hbw + rfsqr * mu);
        return mu - newmu;
    }
};

// Shared set-up of the two inversions (kastaun_c2p.hh:307-337,
// ent_based_c2p.hh:166-196).  cons are undensitised on entry.
struct C2PSetup {
    double D, q, s, ye, r2, Bt2, rBt, rBt2, h0;
    double r[3], Bt[3], B[3];
    template <typename EOS>
    C2PSetup(const EOS& eos, const Metric& m, double cons[NCONS]) {
        h0 = eos.enthalpy_minimum();
        cons[DENSL] = std::fmax(0, cons[DENSL]);
        D = cons[DENSL];
        q = cons[TAUL] / D;
        s = cons[ENTSL] / D;
        double rl[3] = {cons[STXL] / D, cons[STYL] / D, cons[STZL] / D};
        double sD = std::sqrt(D);
        for (int i = 0; i < 3; ++i) { Bt[i] = cons[BSXL + i] / sD; B[i] = cons[BSXL + i]; }
        r2 = m.square_covec(rl);
        Bt2 = m.square_vec(Bt);
        rBt = rl[0] * Bt[0] + rl[1] * Bt[1] + rl[2] * Bt[2];
        rBt2 = rBt * rBt;
        m.raise(rl, r);
        ye = cons[YESL] / D;
    }
};

// Optional narrowed brackets for the energy root (e_lo, e_hi) and the
// entropy-backup root (s_lo, s_hi), in units of 1/h0: the master functions
// have one root on (0, 1/h0] (Kastaun, Kalinani & Ciolfi 2021, PRD 103,
// 023018, sec. III; checked on a dense grid in the tests).  A bracket with
// 0 <= lo < hi <= 1 on which f changes sign is searched directly, skipping
// the Newton solve for mu+ and the bracket [0, mu+]; otherwise the plain
// path runs.  The roots x = mu h0 and the bracket ends mu+ h0 are written
// back (-1 where not computed).
struct MuGuess {
    double e_lo = 0, e_hi = 0, s_lo = 0, s_hi = 0;
    int used = 0;                    // bit 0: energy bracket, bit 1: entropy bracket
    double x_e = -1, x_s = -1, m_e = -1, m_s = -1;
};

// Brent on the guessed bracket [lo h0^-1, hi h0^-1] when it changes sign;
// returns -1 when the bracket is not usable.
template <typename F>
inline double solve_in_guess(F& f, double h0, double lo_frac, double hi_frac, bool& used) {
    used = false;
    if (lo_frac < hi_frac && lo_frac >= 0.0 && hi_frac <= 1.0) {
        double lo = lo_frac / h0, hi = hi_frac / h0;
        double flo = f(lo), fhi = f(hi);
        if (flo * fhi < 0) { used = true; return brent_with_ends(f, lo, hi, flo, fhi, 1e-15); }
    }
    return -1.0;
}

template <typename EOS>
inline double kastaun_invert(const EOS& eos, const Metric& m, double cons[NCONS], double prims[NPRIM],
                             Bits& sig, C2PStats* st = nullptr, MuGuess* guess = nullptr) {
    C2PSetup cs(eos, m, cons);
    double h0 = cs.h0;
    FRoot<EOS> fmu(eos, cs.D, cs.q, cs.r2, cs.rBt2, cs.Bt2, cs.ye, h0);
    fmu.st = st;
    bool used = false;
    double mu = guess ? solve_in_guess(fmu, h0, guess->e_lo, guess->e_hi, used) : -1.0;
    if (guess) { guess->used |= used ? 1 : 0; guess->m_e = -1; }
    if (!used) {
        double mu0 = 1 / h0;
        if (cs.r2 >= h0 * h0) {
            FBrack g(cs.r2, cs.Bt2, cs.rBt2, h0);
            g.st = st;
            double lo, hi;
            g.bracket(lo, hi);
            int rerr;
            mu0 = newton_bracketed(lo, hi, g, 30, 1e-10, rerr);
            if (rerr == 0) mu0 *= 1 + 1e-10;
            else mu0 = 1.0 / h0;
        }
        mu = brent(fmu, 0, mu0, 1e-15);
        if (guess) guess->m_e = mu0 * h0;
    }
    if (guess) guess->x_e = mu * h0;
    double x, w, eps, rho, press, temp, entropy;
    double residual = fmu.compute_primitives(mu, sig, x, w, rho, press, eps, temp, entropy);
    prims[EPSL] = eps; prims[RHOL] = rho; prims[PRESSL] = press; prims[TEMPL] = temp; prims[ENTL] = entropy;
    prims[BXL] = cs.B[0]; prims[BYL] = cs.B[1]; prims[BZL] = cs.B[2];
    for (int ii = 0; ii < 3; ++ii) prims[ZXL + ii] = w * mu * x * (cs.r[ii] + mu * cs.rBt * cs.Bt[ii]);
    return std::fabs(residual);
}

template <typename EOS>
inline double entropy_invert(const EOS& eos, const Metric& m, double cons[NCONS], double prims[NPRIM],
                             Bits& sig, C2PStats* st = nullptr, MuGuess* guess = nullptr) {
    C2PSetup cs(eos, m, cons);
    double h0 = cs.h0;
    EntFRoot<EOS> fmu(eos, cs.D, cs.r2, cs.rBt2, cs.Bt2, cs.s, cs.ye, h0);
    fmu.st = st;
    bool used = false;
    double mu = guess ? solve_in_guess(fmu, h0, guess->s_lo, guess->s_hi, used) : -1.0;
    if (guess) { guess->used |= used ? 2 : 0; guess->m_s = -1; }
    if (!used) {
        double mu0 = 1 / h0;
        if (cs.r2 >= h0 * h0) {
            FBrack g(cs.r2, cs.Bt2, cs.rBt2, h0);
            g.st = st;
            int rerr;
            mu0 = newton_bracketed(0, 1. / h0, g, 30, 1e-10, rerr);
            if (rerr == 0) mu0 *= 1 + 1e-10;
            else mu0 = brent(g, 0, 1. / h0, 1e-10) * (1 + 1e-10);
        }
        mu = brent(fmu, 0, mu0, 1e-15);
        if (guess) guess->m_s = mu0 * h0;
    }
    if (guess) guess->x_s = mu * h0;
    double x, w, eps, rho, press, temp, entropy;
    double residual = fmu.compute_primitives(mu, sig, x, w, rho, press, eps, temp, entropy);
    prims[EPSL] = eps; prims[RHOL] = rho; prims[PRESSL] = press; prims[TEMPL] = temp; prims[ENTL] = entropy;
    prims[BXL] = cs.B[0]; prims[BYL] = cs.B[1]; prims[BZL] = cs.B[2];
    for (int ii = 0; ii < 3; ++ii) prims[ZXL + ii] = w * mu * x * (cs.r[ii] + mu * cs.rBt * cs.Bt[ii]);
    return std::fabs(residual);
}

// ---------------------------------------------------------------------------
// con2prim driver.  Restates src/physics/eos/c2p.cpp:40-455 and the signal
// handlers of include/grace/physics/eos/c2p.hh:136-322 (aborting branches are
// reported through the returned status instead of aborting).
// ---------------------------------------------------------------------------
inline void c2p_set_all_resets(Bits& e) {
    e.set(E_RESET_DENS); e.set(E_RESET_TAU); e.set(E_RESET_STILDE); e.set(E_RESET_ENTROPY); e.set(E_RESET_YE);
}

// returns false if the reference would have aborted (fatal, non-lenient)
template <typename EOS>
inline bool c2p_handle_signals(const Bits& sig, bool lenient, Bits& err) {
    bool ok = true;
    if (sig.test(SIG_RHO_TOO_HIGH)) {
        err.set(E_SIG_RHO_TOO_HIGH);
        if (lenient) c2p_set_all_resets(err); else ok = false;
    }
    if (sig.test(SIG_RHO_TOO_LOW)) {
        err.set(E_RESET_DENS); err.set(E_RESET_TAU); err.set(E_RESET_STILDE); err.set(E_RESET_ENTROPY);
        if (EOS::has_ye) err.set(E_RESET_YE);
        err.set(E_SIG_RHO_TOO_LOW);
    }
    if (EOS::has_ye) {
        if (sig.test(SIG_YE_TOO_LOW)) { err.set(E_RESET_YE); err.set(E_SIG_YE_TOO_LOW); }
        if (sig.test(SIG_YE_TOO_HIGH)) { err.set(E_RESET_YE); err.set(E_SIG_YE_TOO_HIGH); }
    }
    if (sig.test(SIG_ENT_TOO_LOW)) { err.set(E_RESET_ENTROPY); err.set(E_SIG_ENT_TOO_LOW); }
    if (sig.test(SIG_ENT_TOO_HIGH)) { err.set(E_RESET_ENTROPY); err.set(E_SIG_ENT_TOO_HIGH); }
    if (sig.test(SIG_EPS_TOO_HIGH)) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); err.set(E_SIG_EPS_TOO_HIGH); }
    if (sig.test(SIG_EPS_TOO_LOW)) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); err.set(E_SIG_EPS_TOO_LOW); }
    if (sig.test(SIG_VEL_TOO_HIGH)) {
        err.set(E_RESET_DENS); err.set(E_RESET_TAU); err.set(E_RESET_STILDE); err.set(E_RESET_ENTROPY);
        if (EOS::has_ye) err.set(E_RESET_YE);
        err.set(E_SIG_VEL_TOO_HIGH);
    }
    if (sig.test(SIG_SIGMA_TOO_HIGH)) {
        err.set(E_RESET_DENS); err.set(E_RESET_TAU); err.set(E_RESET_STILDE); err.set(E_RESET_ENTROPY);
        if (EOS::has_ye) err.set(E_RESET_YE);
        err.set(E_SIG_SIGMA_TOO_HIGH);
    }
    return ok;
}

template <typename EOS>
inline bool c2p_handle_eos_signals(const Bits& e, bool lenient, Bits& err) {
    bool ok = true;
    if (e.test(EOS_RHO_TOO_HIGH)) { err.set(E_SIG_RHO_TOO_HIGH); ok = false; }
    if (e.test(EOS_RHO_TOO_LOW)) {
        err.set(E_RESET_DENS); err.set(E_RESET_TAU); err.set(E_RESET_STILDE); err.set(E_RESET_ENTROPY);
        if (EOS::has_ye) err.set(E_RESET_YE);
        err.set(E_SIG_RHO_TOO_LOW);
    }
    if (EOS::has_ye) {
        if (e.test(EOS_YE_TOO_LOW)) { err.set(E_RESET_YE); err.set(E_SIG_YE_TOO_LOW); }
        if (e.test(EOS_YE_TOO_HIGH)) { err.set(E_RESET_YE); err.set(E_SIG_YE_TOO_HIGH); }
    }
    if (e.test(EOS_ENTROPY_TOO_HIGH)) { err.set(E_RESET_ENTROPY); err.set(E_SIG_ENT_TOO_HIGH); }
    if (e.test(EOS_ENTROPY_TOO_LOW)) { err.set(E_RESET_ENTROPY); err.set(E_SIG_ENT_TOO_LOW); }
    if (e.test(EOS_EPS_TOO_HIGH)) {
        err.set(E_SIG_EPS_TOO_HIGH);
        if (lenient) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); } else ok = false;
    }
    if (e.test(EOS_EPS_TOO_LOW)) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); err.set(E_SIG_EPS_TOO_LOW); }
    if (e.test(EOS_TEMPERATURE_TOO_HIGH)) {
        err.set(E_SIG_TEMP_TOO_HIGH);
        if (lenient) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); } else ok = false;
    }
    if (e.test(EOS_TEMPERATURE_TOO_LOW)) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); err.set(E_SIG_TEMP_TOO_LOW); }
    if (e.test(EOS_PRESS_TOO_HIGH)) {
        err.set(E_SIG_PRESS_TOO_HIGH);
        if (lenient) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); } else ok = false;
    }
    if (e.test(EOS_PRESS_TOO_LOW)) { err.set(E_RESET_STILDE); err.set(E_RESET_TAU); err.set(E_SIG_PRESS_TOO_LOW); }
    return ok;
}

// c2p.cpp:416-455
inline void prims_to_conservs(const double prims[NPRIM], double cons[NCONS], const Metric& m) {
    const double* z = &prims[ZXL];
    const double* B = &prims[BXL];
    double W = get_W(m.g, z);
    double b[4], b2;
    get_smallb(m.alp, m.beta, m.g, B, z, W, b, b2);
    double D, tau, S[3], Ss;
    get_conserved(m.alp, m.beta, m.g, prims[RHOL], prims[PRESSL], prims[EPSL], z, prims[ENTL], W, b2, b,
                  D, tau, S, Ss);
    cons[DENSL] = m.sqrtg * D;
    cons[STXL] = m.sqrtg * S[0];
    cons[STYL] = m.sqrtg * S[1];
    cons[STZL] = m.sqrtg * S[2];
    cons[TAUL] = m.sqrtg * tau;
    cons[ENTSL] = m.sqrtg * Ss;
    cons[YESL] = cons[DENSL] * prims[YEL];
}

// c2p.cpp:40-61
inline double compute_beta(const double prims[NPRIM], const Metric& m) {
    double W = get_W(m.g, &prims[ZXL]);
    double b[4], b2;
    get_smallb(m.alp, m.beta, m.g, &prims[BXL], &prims[ZXL], W, b, b2);
    return 2.0 * prims[PRESSL] / std::fmax(b2, 1e-50);
}

// c2p.cpp:63-115
template <typename EOS>
inline void limit_primitives(double prims[NPRIM], const Metric& m, const EOS& eos, double max_w,
                             double max_sigma, Bits& sig) {
    double z2max = sqr(max_w) - 1.;
    double z2 = m.square_vec(&prims[ZXL]);
    if (z2 >= z2max) {
        double zn = 0.99 * std::sqrt(z2max / z2);
        prims[ZXL] *= zn; prims[ZYL] *= zn; prims[ZZL] *= zn;
        sig.set(SIG_VEL_TOO_HIGH);
    }
    double W = get_W(m.g, &prims[ZXL]);
    double b[4], b2;
    get_smallb(m.alp, m.beta, m.g, &prims[BXL], &prims[ZXL], W, b, b2);
    double sigma = b2 / prims[RHOL];
    if (sigma >= max_sigma) {
        prims[RHOL] *= 1.001 * sigma / max_sigma;
        Bits ee; double cs2;
        prims[PRESSL] = eos.press_eps_csnd2_entropy__temp_rho_ye(prims[EPSL], cs2, prims[ENTL], prims[TEMPL],
                                                                 prims[RHOL], ee);
        sig.set(SIG_SIGMA_TOO_HIGH);
    }
}

// c2p.cpp:117-157
template <typename EOS>
inline bool limit_conservatives(double cons[NCONS], const Metric& m, const EOS& eos, Bits& err) {
    bool any = false;
    double rhoL = cons[DENSL];
    double epsmin, epsmax;
    Bits ee;
    eos.eps_range__rho_ye(epsmin, epsmax, rhoL, ee);
    double B2L = m.square_vec(&cons[BSXL]);
    if (cons[TAUL] / cons[DENSL] - 0.5 * B2L / cons[DENSL] < std::fmin(0.0, epsmin)) {
        cons[TAUL] = epsmin * cons[DENSL] + 0.5 * B2L;
        err.set(E_RESET_TAU);
        any = true;
    }
    double st2_max = sqr(cons[TAUL] / cons[DENSL] + 1);
    double st2L = m.square_covec(&cons[STXL]) / sqr(cons[DENSL]);
    if (st2L > st2_max) {
        double f = std::sqrt(st2_max / st2L);
        cons[STXL] *= f; cons[STYL] *= f; cons[STZL] *= f;
        err.set(E_RESET_STILDE);
        any = true;
    }
    return any;
}

// c2p.cpp:159-186
template <typename EOS>
inline void reset_to_atmosphere(double cons[NCONS], double prims[NPRIM], const Metric& m, const AtmoParams& atmo,
                                const EOS& eos, Bits& err) {
    prims[RHOL] = atmo.rho_fl; prims[YEL] = atmo.ye_fl; prims[TEMPL] = atmo.temp_fl;
    prims[ZXL] = prims[ZYL] = prims[ZZL] = 0.;
    Bits ee; double cs2;
    prims[PRESSL] = eos.press_eps_csnd2_entropy__temp_rho_ye(prims[EPSL], cs2, prims[ENTL], prims[TEMPL],
                                                             prims[RHOL], ee);
    prims_to_conservs(prims, cons, m);
    c2p_set_all_resets(err);
    err.set(E_ATMO_RESET);
}

struct C2PResult { bool floored = false; bool fatal = false; };

// c2p.cpp:188-414.  cons[] densitised on entry; on exit it holds the
// re-densitised conserved recomputed from the final primitives.
template <typename EOS>
inline C2PResult conservs_to_prims(double cons[NCONS], double prims[NPRIM], const Metric& m, const EOS& eos,
                                   const AtmoParams& atmo, const ExcisionParams& exc, const C2PParams& cp,
                                   const double* rtp, Bits& err, bool dry_run = false,
                                   C2PStats* st = nullptr, MuGuess* guess = nullptr) {
    C2PResult res;
    bool failed = false;
    bool floored = false;
    Bits ret;
    err.reset();
    err.set(E_RESET_ENTROPY);
    for (int i = 0; i < NCONS; ++i) cons[i] /= m.sqrtg;
    prims[BXL] = cons[BSXL]; prims[BYL] = cons[BSYL]; prims[BZL] = cons[BSZL];
    if (cons[DENSL] < 0) {
        reset_to_atmosphere(cons, prims, m, atmo, eos, err);
        res.floored = true;
        return res;
    }
    prims[YEL] = cons[YESL] / cons[DENSL];
    double yemax = eos.c2p_ye_max(), yemin = eos.c2p_ye_min();
    if (prims[YEL] > yemax) { prims[YEL] = yemax; err.set(E_RESET_YE); }
    if (prims[YEL] < yemin) { prims[YEL] = yemin; err.set(E_RESET_YE); }
    bool lenient = m.alp < cp.alp_bh_thresh;
    floored |= limit_conservatives(cons, m, eos, err);
    if (floored && dry_run) { res.floored = floored; return res; }
    if (cons[DENSL] >= atmo.rho_fl * (1 + atmo.atmo_tol)) {
        double residual = kastaun_invert(eos, m, cons, prims, ret, st, guess);
        failed = (std::fabs(residual) > cp.tol) || ret.test(SIG_RHO_TOO_HIGH) || ret.test(SIG_EPS_TOO_HIGH)
                 || (!std::isfinite(residual));
        double beta = compute_beta(prims, m);
        bool distrust = failed || ret.test(SIG_EPS_TOO_LOW) || (beta <= cp.beta_fallback);
        if (distrust && cp.use_ent_backup && !dry_run) {
            double save[NPRIM];
            std::memcpy(save, prims, sizeof(save));
            Bits bret;
            double rb = entropy_invert(eos, m, cons, prims, bret, st, guess);
            bool ok = (std::fabs(rb) <= cp.tol) && std::isfinite(rb) && (prims[EPSL] < eos.c2p_eps_max());
            if (ok) {
                ret = bret;
                failed = false;
                err.set(E_ENT_BACKUP_USED);
            } else if (!failed) {
                std::memcpy(prims, save, sizeof(save));
            }
        }
        if (ret.test(SIG_EPS_TOO_HIGH) || ret.test(SIG_RHO_TOO_LOW) || ret.test(SIG_RHO_TOO_HIGH)
            || ret.test(SIG_VEL_TOO_HIGH))
            floored = true;
        if (EOS::is_ideal && ret.test(SIG_EPS_TOO_LOW)) floored = true;
        if (!c2p_handle_signals<EOS>(ret, lenient, err)) res.fatal = true;
    } else {
        failed = true;
    }
    bool excise = exc.excise_by_radius ? (rtp ? rtp[0] <= exc.r_ex : false) : (m.alp <= exc.alp_ex);
    if ((prims[RHOL] < (1. + atmo.atmo_tol) * atmo.rho_fl) || failed || excise) {
        reset_to_atmosphere(cons, prims, m, atmo, eos, err);
        floored = true;
    } else if (prims[TEMPL] < atmo.temp_fl) {
        prims[TEMPL] = atmo.temp_fl;
        Bits ee; double cs2;
        prims[PRESSL] = eos.press_eps_csnd2_entropy__temp_rho_ye(prims[EPSL], cs2, prims[ENTL], prims[TEMPL],
                                                                 prims[RHOL], ee);
        if (!c2p_handle_eos_signals<EOS>(ee, lenient, err)) res.fatal = true;
        err.set(E_RESET_TAU); err.set(E_RESET_STILDE); err.set(E_RESET_ENTROPY); err.set(E_T_FLOORED);
        if (EOS::is_ideal) floored = true;
    } else {
        Bits sg;
        limit_primitives(prims, m, eos, cp.max_w, cp.max_sigma, sg);
        if (!c2p_handle_signals<EOS>(sg, lenient, err)) res.fatal = true;
        if (sg.test(SIG_VEL_TOO_HIGH) || sg.test(SIG_SIGMA_TOO_HIGH)) floored = true;
    }
    prims_to_conservs(prims, cons, m);
    res.floored = floored;
    return res;
}

}  // namespace oracle

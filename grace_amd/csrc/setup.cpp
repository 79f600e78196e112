// setup.cpp -- host-side input generators for the GRMHD workloads.
//
// Initial data follow the reference's initial-data path: primitives at every
// cell incl. ghosts (src/physics/grmhd.cpp:359-527), face B = B * sqrt(gamma)
// at the staggered positions ("direct" initialisation, :457-490), cell B from
// the face average and conserved from prims_to_conservs
// (set_conservs_from_prims, :676-759).  Noise is drawn from a counter-based
// SplitMix64 keyed by (seed, instance, padded global cell index) so any block
// order gives bit-identical inputs.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include "phys.hh"
#include "../../include/grmhd_setup.h"

using namespace gm;

namespace {

inline uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform in [0,1) from (seed, a, b, c)
inline double u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = splitmix(seed ^ splitmix(a + 0x632BE59BD9B4E019ull * (b + 1)) ^ splitmix(c * 0x8CB92BA72F3D8DD7ull + 7));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// f(q0, q1, t) over [0, nq) split across host threads; t = thread index.
template <typename F>
void parallel_blocks(int64_t nq, F f) {
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nq, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back([&, t] { f(nq * t / nt, nq * (t + 1) / nt, t); });
    for (auto& x : th) x.join();
}

struct Idx {
    int N[3]; int64_t ncell, nface[3];
    Idx(const grmhd_grid* g) {
        N[0] = g->nx + 2 * g->ngz; N[1] = g->ny + 2 * g->ngz; N[2] = g->nz + 2 * g->ngz;
        ncell = (int64_t)N[0] * N[1] * N[2];
        for (int d = 0; d < 3; ++d) nface[d] = (int64_t)(N[0] + (d == 0)) * (N[1] + (d == 1)) * (N[2] + (d == 2));
    }
    int64_t c(int i, int j, int k) const { return i + (int64_t)N[0] * (j + (int64_t)N[1] * k); }
    int64_t f(int d, int i, int j, int k) const {
        int e0 = N[0] + (d == 0), e1 = N[1] + (d == 1);
        return i + (int64_t)e0 * (j + (int64_t)e1 * k);
    }
};

// Block of instance containing q
bool find_block(const grmhd_instance* inst, int n_inst, int64_t q, int& a, int b[3]) {
    for (a = 0; a < n_inst; ++a) {
        const grmhd_instance& I = inst[a];
        int64_t nb = (int64_t)I.nb[0] * I.nb[1] * I.nb[2];
        if (q >= I.q0 && q < I.q0 + nb) {
            int64_t r = q - I.q0;
            b[0] = (int)(r % I.nb[0]); b[1] = (int)((r / I.nb[0]) % I.nb[1]); b[2] = (int)(r / ((int64_t)I.nb[0] * I.nb[1]));
            return true;
        }
    }
    return false;
}

void flat_metric(double* st, const Idx& X, int64_t q, int64_t c) {
    const double v[16] = {1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 16; ++a) st[(q * NSTATE + GXX_ + a) * X.ncell + c] = v[a];
}

// cell B from faces, conserved from primitives (set_conservs_from_prims)
void finish_cells(const grmhd_grid* g, const Idx& X, double* st, double* const B[3], double* aux, const double* idx_q,
                  int64_t q) {
    for (int k = 0; k < X.N[2]; ++k)
        for (int j = 0; j < X.N[1]; ++j)
            for (int i = 0; i < X.N[0]; ++i) {
                int64_t c = X.c(i, j, k);
                Met m;
                for (int a = 0; a < 6; ++a) m.g[a] = st[(q * NSTATE + GXX_ + a) * X.ncell + c];
                m.alp = st[(q * NSTATE + ALP_) * X.ncell + c];
                for (int a = 0; a < 3; ++a) m.beta[a] = st[(q * NSTATE + BETAX_ + a) * X.ncell + c];
                m.init();
                double bx0 = B[0][q * X.nface[0] + X.f(0, i, j, k)], bx1 = B[0][q * X.nface[0] + X.f(0, i + 1, j, k)];
                double by0 = B[1][q * X.nface[1] + X.f(1, i, j, k)], by1 = B[1][q * X.nface[1] + X.f(1, i, j + 1, k)];
                double bz0 = B[2][q * X.nface[2] + X.f(2, i, j, k)], bz1 = B[2][q * X.nface[2] + X.f(2, i, j, k + 1)];
                auto A = [&](int v) -> double& { return aux[(q * NAUX + v) * X.ncell + c]; };
                A(BX_) = 0.5 * (bx0 + bx1) / m.sqrtg;
                A(BY_) = 0.5 * (by0 + by1) / m.sqrtg;
                A(BZ_) = 0.5 * (bz0 + bz1) / m.sqrtg;
                A(BDIV_) = ((bx1 - bx0) * idx_q[0] + (by1 - by0) * idx_q[1] + (bz1 - bz0) * idx_q[2]) / m.sqrtg;
                double pr[NPRIM] = {0};
                pr[RHOL] = A(RHO_); pr[PRESSL] = A(PRESS_); pr[ZXL] = A(ZVECX_); pr[ZYL] = A(ZVECY_); pr[ZZL] = A(ZVECZ_);
                pr[YEL] = A(YE_); pr[TEMPL] = A(TEMP_); pr[EPSL] = A(EPS_); pr[ENTL] = A(ENTROPY_);
                pr[BXL] = A(BX_); pr[BYL] = A(BY_); pr[BZL] = A(BZ_);
                double cons[NCONS];
                p2c(pr, cons, m);
                for (int v = 0; v < NHRSC; ++v) st[(q * NSTATE + v) * X.ncell + c] = cons[v];
            }
    (void)g;
}

int setup_blast_impl(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                     const double lo[3], const double center[3], double rho_amb, double p_amb, double p_in,
                     double r_in, const double Bvec[3], double noise, uint64_t seed, double* state, double* Bx,
                     double* By, double* Bz, double* aux) {
    Idx X(g);
    double* B[3] = {Bx, By, Bz};
    std::memset(aux, 0, sizeof(double) * NAUX * X.ncell * g->nq);
    const double gm1 = gamma - 1.0;
    int n[3] = {g->nx, g->ny, g->nz};
    for (int64_t q = 0; q < g->nq; ++q) {
        int a, b[3];
        if (!find_block(inst, n_inst, q, a, b)) return 1;
        const grmhd_instance& I = inst[a];
        double dx[3];
        int Ntot[3];
        for (int d = 0; d < 3; ++d) { dx[d] = 1.0 / I.idx[d]; Ntot[d] = I.nb[d] * n[d]; }
        for (int k = 0; k < X.N[2]; ++k)
            for (int j = 0; j < X.N[1]; ++j)
                for (int i = 0; i < X.N[0]; ++i) {
                    int loc[3] = {i, j, k};
                    int gl[3];
                    double x[3];
                    for (int d = 0; d < 3; ++d) {
                        gl[d] = b[d] * n[d] + loc[d] - g->ngz;
                        x[d] = lo[d] + (gl[d] + 0.5) * dx[d];
                    }
                    int64_t key = (int64_t)(gl[0] + g->ngz) + (int64_t)(Ntot[0] + 2 * g->ngz) *
                                  ((gl[1] + g->ngz) + (int64_t)(Ntot[1] + 2 * g->ngz) * (gl[2] + g->ngz));
                    const double xc[3] = {x[0] - center[0], x[1] - center[1], x[2] - center[2]};
                    double r = std::sqrt(xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
                    double rho = rho_amb, p = r < r_in ? p_in : p_amb;
                    rho *= 1.0 + noise * (2.0 * u01(seed, a, (uint64_t)key, 0) - 1.0);
                    p *= 1.0 + noise * (2.0 * u01(seed, a, (uint64_t)key, 1) - 1.0);
                    int64_t c = X.c(i, j, k);
                    flat_metric(state, X, q, c);
                    auto A = [&](int v) -> double& { return aux[(q * NAUX + v) * X.ncell + c]; };
                    double eps = p / (gm1 * rho);
                    A(RHO_) = rho; A(PRESS_) = p; A(EPS_) = eps; A(TEMP_) = eps * gm1;
                    A(ENTROPY_) = std::log(eps) / gm1 - std::log(rho);
                    A(ZVECX_) = A(ZVECY_) = A(ZVECZ_) = 0.0; A(YE_) = 0.0;
                }
        for (int d = 0; d < 3; ++d)
            for (int64_t f = 0; f < X.nface[d]; ++f) B[d][q * X.nface[d] + f] = Bvec[d];
        finish_cells(g, X, state, B, aux, I.idx, q);
    }
    return 0;
}
// ---------------------------------------------------------------------------
// Fishbone-Moncrief torus (Fishbone & Moncrief 1976, ApJ 207, 962) around a
// Kerr hole (M = 1, spin a) in Boyer-Lindquist (r, theta), with the constant
// l = u^t u_phi.  With Sigma = r^2 + a^2 cos^2, Delta = r^2 - 2r + a^2,
// A = (r^2+a^2)^2 - a^2 Delta sin^2 and e = Sigma^2 Delta / (A^2 sin^2)
// (= alpha^2 / g_phiphi):
//   ln h = 1/2 ln[(1 + sqrt(1 + 4 l^2 e)) A / (Sigma Delta)]
//          - 1/2 sqrt(1 + 4 l^2 e) - 2 a r l / A  - (the same at r_in, pi/2)
// and the azimuthal flow: U = u_(phi) in the zero-angular-momentum frame
// solves U^2 (1 + U^2) = l^2 e, u^t = sqrt(1 + U^2)/alpha,
// u^phi = u^t omega + U / sqrt(g_phiphi), omega = 2 a r / A.  (These satisfy
// u.u = -1, u^t u_phi = l and the relativistic Euler equation of a
// stationary azimuthal flow, d ln h = -d ln(-u_t) + Omega/(1 - Omega l_s) d l_s
// with l_s = -u_phi/u_t; tests/test_torus_setup.py checks all three.)
// ---------------------------------------------------------------------------
struct FMTorus {
    double a, l;
    void parts(double r, double th, double& Sig, double& Del, double& A, double& s2) const {
        const double st = std::sin(th), ct = std::cos(th);
        s2 = st * st;
        Sig = r * r + a * a * ct * ct;
        Del = r * r - 2.0 * r + a * a;
        A = (r * r + a * a) * (r * r + a * a) - a * a * Del * s2;
    }
    double lnh_raw(double r, double th) const {
        double Sig, Del, A, s2;
        parts(r, th, Sig, Del, A, s2);
        const double t = std::sqrt(1.0 + 4.0 * l * l * Sig * Sig * Del / (A * A * s2));
        return 0.5 * std::log((1.0 + t) * A / (Sig * Del)) - 0.5 * t - 2.0 * a * r * l / A;
    }
    // BL u^t and u^phi of the torus flow
    void vel(double r, double th, double& ut, double& uph) const {
        double Sig, Del, A, s2;
        parts(r, th, Sig, Del, A, s2);
        const double e = Sig * Sig * Del / (A * A * s2);
        const double U = std::sqrt(0.5 * (-1.0 + std::sqrt(1.0 + 4.0 * l * l * e)));
        const double alpha = std::sqrt(Sig * Del / A), om = 2.0 * a * r / A, gpp = A * s2 / Sig;
        ut = std::sqrt(1.0 + U * U) / alpha;
        uph = ut * om + U / std::sqrt(gpp);
    }
};

// l such that the equatorial enthalpy peaks at r_max: d(ln h)/dr = 0 there
// (bisection on a central difference; the slope decreases with l).
double fm_ell(double a, double r_max) {
    auto slope = [&](double l) {
        FMTorus T{a, l};
        const double h = 1e-5 * r_max;
        return (T.lnh_raw(r_max + h, 0.5 * M_PI) - T.lnh_raw(r_max - h, 0.5 * M_PI)) / (2.0 * h);
    };
    double lo = 1.0, hi = 50.0, flo = slope(lo), fhi = slope(hi);
    if (!(flo * fhi < 0.0)) return -1.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi), fm = slope(mid);
        if ((fm < 0.0) == (flo < 0.0)) { lo = mid; flo = fm; } else hi = mid;
    }
    return 0.5 * (lo + hi);
}

// Boyer-Lindquist r and theta of a Cartesian Kerr-Schild point
// (x + iy = (r + ia) e^{i phi} sin(theta), z = r cos(theta)).
void ks_r_theta(double a, double x, double y, double z, double& r, double& th) {
    const double R2 = x * x + y * y + z * z, a2 = a * a;
    const double r2 = 0.5 * (R2 - a2 + std::sqrt((R2 - a2) * (R2 - a2) + 4.0 * a2 * z * z));
    r = std::sqrt(r2);
    th = std::acos(std::fmax(-1.0, std::fmin(1.0, z / r)));
}

struct FMPoint { double rho, p, z[3], ut; bool in; };
// Primitives of the torus at a Cartesian Kerr-Schild point (ideal gas, rho
// normalised to 1 at the pressure maximum; atmosphere outside).
FMPoint fm_point(const FMTorus& T, double lnh_in, double kappa, double gamma, double r_in, double rho_atm,
                 double p_atm, double x, double y, double z, const double beta_up[3]) {
    FMPoint P{rho_atm, p_atm, {0.0, 0.0, 0.0}, 0.0, false};
    double r, th;
    ks_r_theta(T.a, x, y, z, r, th);
    const double rh = 1.0 + std::sqrt(1.0 - T.a * T.a);
    if (r < r_in || r <= rh || std::sin(th) < 1e-12) return P;
    const double lnh = T.lnh_raw(r, th) - lnh_in;
    if (!(lnh > 0.0)) return P;
    const double h = std::exp(lnh);
    const double rho = std::pow((h - 1.0) * (gamma - 1.0) / (kappa * gamma), 1.0 / (gamma - 1.0));
    if (!(rho > rho_atm)) return P;
    double ut, uph;
    T.vel(r, th, ut, uph);
    // Kerr-Schild u^t, u^phi equal the Boyer-Lindquist ones for u^r = 0;
    // u^x = dx/dphi u^phi = -y u^phi, u^y = x u^phi; Valencia z^i = u^i + u^t beta^i
    const double u[3] = {-y * uph, x * uph, 0.0};
    P.rho = rho;
    P.p = kappa * std::pow(rho, gamma);
    for (int i = 0; i < 3; ++i) P.z[i] = u[i] + ut * beta_up[i];
    P.ut = ut;
    P.in = true;
    return P;
}

// Where a block lies: pos(d, loc, off) = org[d] + (loc - ngz + off) dx[d];
// goff + loc - ngz is the block's cell index on its level (Ntot cells wide),
// which with `a` (instance, or level of a forest) keys the noise.
struct BlockGeom { double org[3], dx[3], idx[3]; int goff[3], Ntot[3], a; };

bool geoms_of_instances(const grmhd_grid* g, const grmhd_instance* inst, int n_inst, const double lo[3],
                        std::vector<BlockGeom>& out) {
    const int n[3] = {g->nx, g->ny, g->nz};
    out.resize((size_t)g->nq);
    for (int64_t q = 0; q < g->nq; ++q) {
        int a, b[3];
        if (!find_block(inst, n_inst, q, a, b)) return false;
        BlockGeom& G = out[(size_t)q];
        G.a = a;
        for (int d = 0; d < 3; ++d) {
            G.idx[d] = inst[a].idx[d]; G.dx[d] = 1.0 / G.idx[d];
            G.goff[d] = b[d] * n[d]; G.Ntot[d] = inst[a].nb[d] * n[d];
            G.org[d] = lo[d] + G.goff[d] * G.dx[d];
        }
    }
    return true;
}

// Block list of a two-level forest in grmhd_ctx_create_forest order.
bool geoms_of_forest(const grmhd_grid* g, const grmhd_forest* f, const double lo[3], std::vector<BlockGeom>& out) {
    const int64_t nc = (int64_t)f->nb[0] * f->nb[1] * f->nb[2];
    if (g->nq != nc + 7 * (int64_t)f->n_refined) return false;
    std::vector<char> ref((size_t)nc, 0);
    for (int r = 0; r < f->n_refined; ++r) {
        if (f->refined[r] < 0 || f->refined[r] >= nc) return false;
        ref[f->refined[r]] = 1;
    }
    const int n[3] = {g->nx, g->ny, g->nz};
    out.clear();
    auto add = [&](int lev, const int crd[3]) {
        BlockGeom G;
        G.a = lev;
        for (int d = 0; d < 3; ++d) {
            G.idx[d] = f->idx[d] * (lev ? 2.0 : 1.0); G.dx[d] = 1.0 / G.idx[d];
            G.goff[d] = crd[d] * n[d]; G.Ntot[d] = (f->nb[d] * n[d]) << lev;
            G.org[d] = lo[d] + G.goff[d] * G.dx[d];
        }
        out.push_back(G);
    };
    for (int64_t b = 0; b < nc; ++b)
        if (!ref[b]) {
            const int crd[3] = {(int)(b % f->nb[0]), (int)((b / f->nb[0]) % f->nb[1]), (int)(b / ((int64_t)f->nb[0] * f->nb[1]))};
            add(0, crd);
        }
    for (int64_t b = 0; b < nc; ++b)
        if (ref[b]) {
            const int cb[3] = {(int)(b % f->nb[0]), (int)((b / f->nb[0]) % f->nb[1]), (int)(b / ((int64_t)f->nb[0] * f->nb[1]))};
            for (int ch = 0; ch < 8; ++ch) {
                const int crd[3] = {2 * cb[0] + (ch & 1), 2 * cb[1] + ((ch >> 1) & 1), 2 * cb[2] + ((ch >> 2) & 1)};
                add(1, crd);
            }
        }
    return true;
}

int setup_fm_torus_impl(const grmhd_grid* g, const std::vector<BlockGeom>& geo, double gamma, double spin, double r_in,
                        double r_max, double beta, double rho_atm, double p_atm, double noise, uint64_t seed,
                        double* state, double* Bx, double* By, double* Bz, double* aux, double info[4]);

}  // namespace

extern "C" {

int grmhd_setup_forest_wave(const grmhd_grid* g, const grmhd_forest* f, double gamma, const double lo[3],
                            double rho0, double amp, double p0, double p_in, double r_in, const double center[3],
                            const double zvel[3], const double Bvec[3], double b_loop, double* state, double* Bx,
                            double* By, double* Bz, double* aux) {
    if (!g || !f || !(gamma > 1.0) || !(rho0 > 0) || !(p0 > 0)) return 1;
    std::vector<BlockGeom> geo;
    if (!geoms_of_forest(g, f, lo, geo)) return 1;
    Idx X(g);
    double* B[3] = {Bx, By, Bz};
    const int n[3] = {g->nx, g->ny, g->nz};
    const double L[3] = {f->nb[0] * n[0] / f->idx[0], f->nb[1] * n[1] / f->idx[1], f->nb[2] * n[2] / f->idx[2]};
    const double kx = 2.0 * M_PI / L[0], ky = 2.0 * M_PI / L[1];
    const double gm1 = gamma - 1.0;
    auto Az = [&](double x, double y) { return b_loop / kx * std::sin(kx * (x - lo[0])) * std::sin(ky * (y - lo[1])); };
    parallel_blocks(g->nq, [&](int64_t q0, int64_t q1, int) {
        for (int64_t q = q0; q < q1; ++q) {
            std::memset(aux + q * NAUX * X.ncell, 0, sizeof(double) * NAUX * X.ncell);
            const BlockGeom& G = geo[(size_t)q];
            const double* idq = G.idx;
            auto pos = [&](int d, int l, double off) { return G.org[d] + (l - g->ngz + off) * G.dx[d]; };
            for (int k = 0; k < X.N[2]; ++k)
                for (int j = 0; j < X.N[1]; ++j)
                    for (int i = 0; i < X.N[0]; ++i) {
                        const double x = pos(0, i, 0.5), y = pos(1, j, 0.5), z = pos(2, k, 0.5);
                        const double r = std::sqrt((x - center[0]) * (x - center[0]) + (y - center[1]) * (y - center[1]) +
                                                   (z - center[2]) * (z - center[2]));
                        const double rho = rho0 * (1.0 + amp * std::sin(kx * (x - lo[0])) * std::sin(ky * (y - lo[1])));
                        const double p = (r_in > 0 && r < r_in) ? p_in : p0;
                        const int64_t c = X.c(i, j, k);
                        flat_metric(state, X, q, c);
                        auto A = [&](int v) -> double& { return aux[(q * NAUX + v) * X.ncell + c]; };
                        const double eps = p / (gm1 * rho);
                        A(RHO_) = rho; A(PRESS_) = p; A(EPS_) = eps; A(TEMP_) = eps * gm1;
                        A(ENTROPY_) = std::log(eps) / gm1 - std::log(rho);
                        A(ZVECX_) = zvel[0]; A(ZVECY_) = zvel[1]; A(ZVECZ_) = zvel[2]; A(YE_) = 0.0;
                    }
            // B_x = dA_z/dy, B_y = -dA_z/dx on the faces, from A_z at the z-edges
            for (int k = 0; k < X.N[2]; ++k)
                for (int j = 0; j < X.N[1] + 1; ++j)
                    for (int i = 0; i < X.N[0] + 1; ++i) {
                        if (j < X.N[1])
                            B[0][q * X.nface[0] + X.f(0, i, j, k)] =
                                Bvec[0] + (Az(pos(0, i, 0.0), pos(1, j + 1, 0.0)) - Az(pos(0, i, 0.0), pos(1, j, 0.0))) * idq[1];
                        if (i < X.N[0])
                            B[1][q * X.nface[1] + X.f(1, i, j, k)] =
                                Bvec[1] - (Az(pos(0, i + 1, 0.0), pos(1, j, 0.0)) - Az(pos(0, i, 0.0), pos(1, j, 0.0))) * idq[0];
                    }
            for (int64_t fz = 0; fz < X.nface[2]; ++fz) B[2][q * X.nface[2] + fz] = Bvec[2];
            finish_cells(g, X, state, B, aux, idq, q);
        }
    });
    return 0;
}

int grmhd_setup_blast(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                      const double lo[3], const double hi[3], double rho_amb, double p_amb, double p_in, double r_in,
                      const double Bvec[3], double noise, uint64_t seed, double* state, double* Bx, double* By,
                      double* Bz, double* aux) {
    const double origin[3] = {0.0, 0.0, 0.0};
    (void)hi;
    return setup_blast_impl(g, inst, n_inst, gamma, lo, origin, rho_amb, p_amb, p_in, r_in, Bvec, noise, seed, state,
                            Bx, By, Bz, aux);
}

int grmhd_setup_shocktube(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma, double x0,
                          double x_jump, const double L[5], const double R[5], double* state, double* Bx, double* By,
                          double* Bz, double* aux) {
    // L/R = {rho, p, Bx, By, Bz}; discontinuity along x at x_jump (x <= x_jump is left)
    Idx X(g);
    double* B[3] = {Bx, By, Bz};
    std::memset(aux, 0, sizeof(double) * NAUX * X.ncell * g->nq);
    const double gm1 = gamma - 1.0;
    for (int64_t q = 0; q < g->nq; ++q) {
        int a, b[3];
        if (!find_block(inst, n_inst, q, a, b)) return 1;
        const grmhd_instance& I = inst[a];
        double dx = 1.0 / I.idx[0];
        for (int k = 0; k < X.N[2]; ++k)
            for (int j = 0; j < X.N[1]; ++j)
                for (int i = 0; i < X.N[0]; ++i) {
                    int gl = b[0] * g->nx + i - g->ngz;
                    double x = x0 + (gl + 0.5) * dx;
                    const double* S = x <= x_jump ? L : R;
                    int64_t c = X.c(i, j, k);
                    flat_metric(state, X, q, c);
                    auto A = [&](int v) -> double& { return aux[(q * NAUX + v) * X.ncell + c]; };
                    double eps = S[1] / (gm1 * S[0]);
                    A(RHO_) = S[0]; A(PRESS_) = S[1]; A(EPS_) = eps; A(TEMP_) = eps * gm1;
                    A(ENTROPY_) = std::log(eps) / gm1 - std::log(S[0]);
                }
        // staggered B: face coordinate along x is the lower face of cell i
        for (int d = 0; d < 3; ++d) {
            int e0 = X.N[0] + (d == 0), e1 = X.N[1] + (d == 1), e2 = X.N[2] + (d == 2);
            for (int k = 0; k < e2; ++k)
                for (int j = 0; j < e1; ++j)
                    for (int i = 0; i < e0; ++i) {
                        int gl = b[0] * g->nx + i - g->ngz;
                        double x = d == 0 ? x0 + gl * dx : x0 + (gl + 0.5) * dx;
                        const double* S = x <= x_jump ? L : R;
                        B[d][q * X.nface[d] + X.f(d, i, j, k)] = S[2 + d];
                    }
        }
        finish_cells(g, X, state, B, aux, I.idx, q);
    }
    return 0;
}

// Kerr-Schild ADM quantities at Cartesian (x,y,z) for mass 1, spin a.
void grmhd_kerr_schild(double a, double x, double y, double z, double g[6], double* alp, double beta[3]) {
    double R2 = x * x + y * y + z * z;
    double a2 = a * a;
    double r2 = 0.5 * (R2 - a2 + std::sqrt((R2 - a2) * (R2 - a2) + 4.0 * a2 * z * z));
    double r = std::sqrt(r2);
    double H = r2 * r / (r2 * r2 + a2 * z * z);
    double l[3] = {(r * x + a * y) / (r2 + a2), (r * y - a * x) / (r2 + a2), z / r};
    const int sidx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int s = 0; s < 6; ++s) g[s] = (sidx[s][0] == sidx[s][1] ? 1.0 : 0.0) + 2.0 * H * l[sidx[s][0]] * l[sidx[s][1]];
    *alp = 1.0 / std::sqrt(1.0 + 2.0 * H);
    for (int i = 0; i < 3; ++i) beta[i] = 2.0 * H * l[i] / (1.0 + 2.0 * H);
}

// Kerr-Schild ADM data with the extrinsic curvature of the stationary slice,
// K_ij = (D_i beta_j + D_j beta_i) / (2 alpha) (d_t gamma_ij = 0 in
// d_t gamma_ij = -2 alpha K_ij + D_i beta_j + D_j beta_i), with the spatial
// derivatives of the analytic gamma_ij and beta_i by 4th-order central
// differences (step 1e-3 M).
void grmhd_kerr_schild_adm(double a, double x, double y, double z, double g[6], double* alp, double beta[3],
                           double K[6]) {
    grmhd_kerr_schild(a, x, y, z, g, alp, beta);
    const int sidx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    auto sym = [&](const double* v, int i, int j) {
        for (int s = 0; s < 6; ++s)
            if ((sidx[s][0] == i && sidx[s][1] == j) || (sidx[s][0] == j && sidx[s][1] == i)) return v[s];
        return 0.0;
    };
    // dg[l][s]: d_l gamma_s; db[l][j]: d_l beta_j (lowered)
    double dg[3][6], db[3][3];
    const double h = 1e-3;
    for (int l = 0; l < 3; ++l) {
        double gs[4][6], bl[4][3];
        const double off[4] = {2 * h, h, -h, -2 * h};
        for (int t = 0; t < 4; ++t) {
            double p[3] = {x, y, z};
            p[l] += off[t];
            double al, bu[3];
            grmhd_kerr_schild(a, p[0], p[1], p[2], gs[t], &al, bu);
            for (int j = 0; j < 3; ++j) bl[t][j] = sym(gs[t], j, 0) * bu[0] + sym(gs[t], j, 1) * bu[1] + sym(gs[t], j, 2) * bu[2];
        }
        for (int s = 0; s < 6; ++s) dg[l][s] = (-gs[0][s] + 8 * gs[1][s] - 8 * gs[2][s] + gs[3][s]) / (12 * h);
        for (int j = 0; j < 3; ++j) db[l][j] = (-bl[0][j] + 8 * bl[1][j] - 8 * bl[2][j] + bl[3][j]) / (12 * h);
    }
    Met m;
    for (int s = 0; s < 6; ++s) m.g[s] = g[s];
    m.init();
    double bl[3];
    for (int j = 0; j < 3; ++j) bl[j] = sym(g, j, 0) * beta[0] + sym(g, j, 1) * beta[1] + sym(g, j, 2) * beta[2];
    // Christoffels of the first kind G_lij = (d_i g_jl + d_j g_il - d_l g_ij)/2, D_i b_j = d_i b_j - G^k_ij b_k
    double Db[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double corr = 0;
            for (int l = 0; l < 3; ++l) {
                const double Gl = 0.5 * (sym(dg[i], j, l) + sym(dg[j], i, l) - sym(dg[l], i, j));
                double bu_l = 0;   // raise b_k with the inverse metric: G^k_ij b_k = G_lij b^l
                bu_l = beta[l];
                corr += Gl * bu_l;
            }
            Db[i][j] = db[i][j] - corr;
        }
    for (int s = 0; s < 6; ++s) {
        const int i = sidx[s][0], j = sidx[s][1];
        K[s] = (Db[i][j] + Db[j][i]) / (2.0 * *alp);
    }
}

// Blast wave of grmhd_setup_blast centred in the box [lo, hi] placed in the
// Kerr-Schild metric of a spin-a hole at the coordinate origin (Cowling: the
// metric is static data).  Primitives as there; the face arrays hold the
// constant densitised flux sqrt(gamma) B^i = Bvec (divergence free), cell B
// and conserved follow from the metric (set_conservs_from_prims).
int grmhd_setup_blast_ks(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                         const double lo[3], const double hi[3], double spin, double rho_amb, double p_amb,
                         double p_in, double r_in, const double Bvec[3], double noise, uint64_t seed, double* state,
                         double* Bx, double* By, double* Bz, double* aux) {
    const double c0[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    int rc = setup_blast_impl(g, inst, n_inst, gamma, lo, c0, rho_amb, p_amb, p_in, r_in, Bvec, noise, seed, state,
                              Bx, By, Bz, aux);
    if (rc) return rc;
    Idx X(g);
    double* B[3] = {Bx, By, Bz};
    int n[3] = {g->nx, g->ny, g->nz};
    for (int64_t q = 0; q < g->nq; ++q) {
        int a, b[3];
        if (!find_block(inst, n_inst, q, a, b)) return 1;
        const grmhd_instance& I = inst[a];
        double dx[3];
        for (int d = 0; d < 3; ++d) dx[d] = 1.0 / I.idx[d];
        auto pos = [&](int d, int l, bool face) { return lo[d] + (b[d] * n[d] + l - g->ngz + (face ? 0.0 : 0.5)) * dx[d]; };
        for (int k = 0; k < X.N[2]; ++k)
            for (int j = 0; j < X.N[1]; ++j)
                for (int i = 0; i < X.N[0]; ++i) {
                    const int64_t c = X.c(i, j, k);
                    double gg[6], al, be[3], K[6];
                    grmhd_kerr_schild_adm(spin, pos(0, i, false), pos(1, j, false), pos(2, k, false), gg, &al, be, K);
                    auto S = [&](int v) -> double& { return state[(q * NSTATE + v) * X.ncell + c]; };
                    for (int s = 0; s < 6; ++s) { S(GXX_ + s) = gg[s]; S(KXX_ + s) = K[s]; }
                    S(ALP_) = al;
                    for (int d = 0; d < 3; ++d) S(BETAX_ + d) = be[d];
                }
        // face B keeps the constant densitised flux sqrt(gamma) B^i = Bvec of
        // grmhd_setup_blast: discretely divergence free in any metric
        finish_cells(g, X, state, B, aux, I.idx, q);
    }
    return 0;
}

// FM torus in the Kerr-Schild metric (Cowling): metric + K_ij in every cell,
// torus primitives (atmosphere rho_atm, p_atm outside, p times
// (1 + noise U[-1,1])), and a single poloidal loop from the vector potential
// A_phi = max(rho/rho_max - 0.2, 0) (A = A_phi dphi with phi = atan2(y, x)),
// whose edge values give the densitised face fluxes by the discrete curl
// (divergence free to round-off), scaled so that max p / max(b^2/2) = beta
// over the interior cells.  info (or null): l, kappa, B scale, max b^2/2.
int grmhd_setup_fm_torus(const grmhd_grid* g, const grmhd_instance* inst, int32_t n_inst, double gamma,
                         const double lo[3], double spin, double r_in, double r_max, double beta, double rho_atm,
                         double p_atm, double noise, uint64_t seed, double* state, double* Bx, double* By,
                         double* Bz, double* aux, double info[4]) {
    std::vector<BlockGeom> geo;
    if (!geoms_of_instances(g, inst, n_inst, lo, geo)) return 1;
    return setup_fm_torus_impl(g, geo, gamma, spin, r_in, r_max, beta, rho_atm, p_atm, noise, seed, state, Bx, By, Bz,
                               aux, info);
}

int grmhd_setup_fm_torus_forest(const grmhd_grid* g, const grmhd_forest* f, double gamma, const double lo[3],
                                double spin, double r_in, double r_max, double beta, double rho_atm, double p_atm,
                                double noise, uint64_t seed, double* state, double* Bx, double* By, double* Bz,
                                double* aux, double info[4]) {
    std::vector<BlockGeom> geo;
    if (!g || !f || !geoms_of_forest(g, f, lo, geo)) return 1;
    return setup_fm_torus_impl(g, geo, gamma, spin, r_in, r_max, beta, rho_atm, p_atm, noise, seed, state, Bx, By, Bz,
                               aux, info);
}

}  // extern "C"

namespace {
int setup_fm_torus_impl(const grmhd_grid* g, const std::vector<BlockGeom>& geo, double gamma, double spin, double r_in,
                        double r_max, double beta, double rho_atm, double p_atm, double noise, uint64_t seed,
                        double* state, double* Bx, double* By, double* Bz, double* aux, double info[4]) {
    const double l = fm_ell(spin, r_max);
    if (!(l > 0.0)) return 2;
    FMTorus T{spin, l};
    const double lnh_in = T.lnh_raw(r_in, 0.5 * M_PI);
    const double hmax = std::exp(T.lnh_raw(r_max, 0.5 * M_PI) - lnh_in);
    if (!(hmax > 1.0)) return 3;
    const double kappa = (hmax - 1.0) * (gamma - 1.0) / gamma;   // rho(r_max) = 1
    const double gm1 = gamma - 1.0;
    Idx X(g);
    double* B[3] = {Bx, By, Bz};
    std::memset(aux, 0, sizeof(double) * NAUX * X.ncell * g->nq);
    const int n[3] = {g->nx, g->ny, g->nz};
    std::vector<double> pmax_t(16, 0.0), pmag_t(16, 0.0);
    std::vector<int> bad(16, 0);
    // pass 1: metric, primitives, unscaled face fluxes
    parallel_blocks(g->nq, [&](int64_t qa, int64_t qb, int t) {
    double& pmax = pmax_t[t];
    for (int64_t q = qa; q < qb; ++q) {
        const BlockGeom& G = geo[(size_t)q];
        const double* dx = G.dx;
        const int* Ntot = G.Ntot;
        const int a = G.a;
        auto pos = [&](int d, int loc, double off) { return G.org[d] + (loc - g->ngz + off) * dx[d]; };
        for (int k = 0; k < X.N[2]; ++k)
            for (int j = 0; j < X.N[1]; ++j)
                for (int i = 0; i < X.N[0]; ++i) {
                    const int64_t c = X.c(i, j, k);
                    const double x = pos(0, i, 0.5), y = pos(1, j, 0.5), z = pos(2, k, 0.5);
                    double gg[6], al, be[3], K[6];
                    grmhd_kerr_schild_adm(spin, x, y, z, gg, &al, be, K);
                    auto S = [&](int v) -> double& { return state[(q * NSTATE + v) * X.ncell + c]; };
                    for (int s2 = 0; s2 < 6; ++s2) { S(GXX_ + s2) = gg[s2]; S(KXX_ + s2) = K[s2]; }
                    S(ALP_) = al;
                    for (int d = 0; d < 3; ++d) S(BETAX_ + d) = be[d];
                    FMPoint P = fm_point(T, lnh_in, kappa, gamma, r_in, rho_atm, p_atm, x, y, z, be);
                    const int gl[3] = {G.goff[0] + i - g->ngz, G.goff[1] + j - g->ngz, G.goff[2] + k - g->ngz};
                    const int64_t key = (int64_t)(gl[0] + g->ngz) + (int64_t)(Ntot[0] + 2 * g->ngz) *
                                        ((gl[1] + g->ngz) + (int64_t)(Ntot[1] + 2 * g->ngz) * (gl[2] + g->ngz));
                    double p = P.p * (1.0 + noise * (2.0 * u01(seed, a, (uint64_t)key, 1) - 1.0));
                    const bool interior = i >= g->ngz && i < g->ngz + n[0] && j >= g->ngz && j < g->ngz + n[1] &&
                                          k >= g->ngz && k < g->ngz + n[2];
                    if (interior) pmax = std::fmax(pmax, p);
                    auto A = [&](int v) -> double& { return aux[(q * NAUX + v) * X.ncell + c]; };
                    const double eps = p / (gm1 * P.rho);
                    A(RHO_) = P.rho; A(PRESS_) = p; A(EPS_) = eps; A(TEMP_) = eps * gm1;
                    A(ENTROPY_) = std::log(eps) / gm1 - std::log(P.rho);
                    A(ZVECX_) = P.z[0]; A(ZVECY_) = P.z[1]; A(ZVECZ_) = P.z[2]; A(YE_) = 0.0;
                }
        // A_phi at a point (zero outside the torus), A_x = -y A_phi/R^2, A_y = x A_phi/R^2
        auto Aphi = [&](double x, double y, double z) {
            double gg[6], al, be[3];
            grmhd_kerr_schild(spin, x, y, z, gg, &al, be);
            const FMPoint P = fm_point(T, lnh_in, kappa, gamma, r_in, rho_atm, p_atm, x, y, z, be);
            return P.in ? std::fmax(P.rho - 0.2, 0.0) : 0.0;
        };
        auto Ax = [&](double x, double y, double z) { const double R2 = x * x + y * y; return R2 > 0 ? -y * Aphi(x, y, z) / R2 : 0.0; };
        auto Ay = [&](double x, double y, double z) { const double R2 = x * x + y * y; return R2 > 0 ? x * Aphi(x, y, z) / R2 : 0.0; };
        // densitised face fluxes = discrete curl of the edge potential (A_z = 0):
        // Bx = -dAy/dz, By = dAx/dz, Bz = dAy/dx - dAx/dy
        for (int d = 0; d < 3; ++d) {
            const int E[3] = {X.N[0] + (d == 0), X.N[1] + (d == 1), X.N[2] + (d == 2)};
            for (int k = 0; k < E[2]; ++k)
                for (int j = 0; j < E[1]; ++j)
                    for (int i = 0; i < E[0]; ++i) {
                        double v = 0.0;
                        if (d == 0) {   // face (x_i, y_j+1/2, z_k+1/2): A_y on the z-edges
                            const double x = pos(0, i, 0.0), y = pos(1, j, 0.5);
                            v = -(Ay(x, y, pos(2, k, 1.0)) - Ay(x, y, pos(2, k, 0.0))) / dx[2];
                        } else if (d == 1) {
                            const double x = pos(0, i, 0.5), y = pos(1, j, 0.0);
                            v = (Ax(x, y, pos(2, k, 1.0)) - Ax(x, y, pos(2, k, 0.0))) / dx[2];
                        } else {
                            const double z = pos(2, k, 0.0);
                            v = (Ay(pos(0, i, 1.0), pos(1, j, 0.5), z) - Ay(pos(0, i, 0.0), pos(1, j, 0.5), z)) / dx[0] -
                                (Ax(pos(0, i, 0.5), pos(1, j, 1.0), z) - Ax(pos(0, i, 0.5), pos(1, j, 0.0), z)) / dx[1];
                        }
                        B[d][q * X.nface[d] + X.f(d, i, j, k)] = v;
                    }
        }
    }
    });
    for (int t = 0; t < 16; ++t) if (bad[t]) return 1;
    // b^2/2 maximum over the interior cells (cell B from the face average)
    parallel_blocks(g->nq, [&](int64_t qa, int64_t qb, int t) {
    double& pmag = pmag_t[t];
    for (int64_t q = qa; q < qb; ++q)
        for (int k = g->ngz; k < g->ngz + n[2]; ++k)
            for (int j = g->ngz; j < g->ngz + n[1]; ++j)
                for (int i = g->ngz; i < g->ngz + n[0]; ++i) {
                    const int64_t c = X.c(i, j, k);
                    Met m;
                    for (int s2 = 0; s2 < 6; ++s2) m.g[s2] = state[(q * NSTATE + GXX_ + s2) * X.ncell + c];
                    m.alp = state[(q * NSTATE + ALP_) * X.ncell + c];
                    for (int s2 = 0; s2 < 3; ++s2) m.beta[s2] = state[(q * NSTATE + BETAX_ + s2) * X.ncell + c];
                    m.init();
                    const double Bc[3] = {
                        0.5 * (B[0][q * X.nface[0] + X.f(0, i, j, k)] + B[0][q * X.nface[0] + X.f(0, i + 1, j, k)]) / m.sqrtg,
                        0.5 * (B[1][q * X.nface[1] + X.f(1, i, j, k)] + B[1][q * X.nface[1] + X.f(1, i, j + 1, k)]) / m.sqrtg,
                        0.5 * (B[2][q * X.nface[2] + X.f(2, i, j, k)] + B[2][q * X.nface[2] + X.f(2, i, j, k + 1)]) / m.sqrtg};
                    const double zz[3] = {aux[(q * NAUX + ZVECX_) * X.ncell + c], aux[(q * NAUX + ZVECY_) * X.ncell + c],
                                          aux[(q * NAUX + ZVECZ_) * X.ncell + c]};
                    const double W = lorentz(m.g, zz);
                    double bb[4], b2;
                    smallb(m, Bc, zz, W, bb, b2);
                    pmag = std::fmax(pmag, 0.5 * b2);
                }
    });
    double pmax = 0.0, pmag = 0.0;
    for (int t = 0; t < 16; ++t) { pmax = std::fmax(pmax, pmax_t[t]); pmag = std::fmax(pmag, pmag_t[t]); }
    const double scale = (pmag > 0.0 && beta > 0.0) ? std::sqrt(pmax / (beta * pmag)) : 0.0;
    parallel_blocks(g->nq, [&](int64_t qa, int64_t qb, int) {
        for (int64_t q = qa; q < qb; ++q) {
            for (int d = 0; d < 3; ++d)
                for (int64_t f = 0; f < X.nface[d]; ++f) B[d][q * X.nface[d] + f] *= scale;
            finish_cells(g, X, state, B, aux, geo[(size_t)q].idx, q);
        }
    });
    if (info) { info[0] = l; info[1] = kappa; info[2] = scale; info[3] = pmag * scale * scale; }
    return 0;
}
}  // namespace

extern "C" {

void grmhd_fm_torus_point(double spin, double gamma, double r_in, double r_max, double x, double y, double z,
                          double out[6]) {
    const double l = fm_ell(spin, r_max);
    FMTorus T{spin, l};
    const double lnh_in = T.lnh_raw(r_in, 0.5 * M_PI);
    const double hmax = std::exp(T.lnh_raw(r_max, 0.5 * M_PI) - lnh_in);
    const double kappa = (hmax - 1.0) * (gamma - 1.0) / gamma;
    double gg[6], al, be[3];
    grmhd_kerr_schild(spin, x, y, z, gg, &al, be);
    const FMPoint P = fm_point(T, lnh_in, kappa, gamma, r_in, 0.0, 0.0, x, y, z, be);
    out[0] = P.rho; out[1] = P.p; out[2] = P.z[0]; out[3] = P.z[1]; out[4] = P.z[2]; out[5] = P.ut;
}

// Config-4 c2p states: rho log-U[1e-12,1e-2], eps log-U[1e-6,1], W U[1,10]
// isotropic, b^2/rho log-U[1e-4,1e2], Kerr-Schild spin a at r U[1.5,50]
// with isotropic angles; ideal gas.  Outputs SoA prims (13n), metric (10n),
// densitised conserved (10n) and radius (n).
int grmhd_setup_c2p_states(int64_t n, uint64_t seed, double gamma, double spin, const double rho_rng[2],
                           const double eps_rng[2], const double W_rng[2], const double sig_rng[2],
                           const double r_rng[2], double* prims, double* met, double* cons, double* rad) {
    const double gm1 = gamma - 1.0;
    for (int64_t s = 0; s < n; ++s) {
        auto U = [&](int k) { return u01(seed, 0x4C32, (uint64_t)s, (uint64_t)k); };
        double rho = std::exp(std::log(rho_rng[0]) + (std::log(rho_rng[1]) - std::log(rho_rng[0])) * U(0));
        double eps = std::exp(std::log(eps_rng[0]) + (std::log(eps_rng[1]) - std::log(eps_rng[0])) * U(1));
        double W = W_rng[0] + (W_rng[1] - W_rng[0]) * U(2);
        double sig = std::exp(std::log(sig_rng[0]) + (std::log(sig_rng[1]) - std::log(sig_rng[0])) * U(3));
        double r = r_rng[0] + (r_rng[1] - r_rng[0]) * U(4);
        double ct = 2.0 * U(5) - 1.0, ph = 2.0 * M_PI * U(6);
        double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
        double rr = std::sqrt(r * r + spin * spin);
        double x = rr * st * std::cos(ph), y = rr * st * std::sin(ph), z = r * ct;
        Met m;
        grmhd_kerr_schild(spin, x, y, z, m.g, &m.alp, m.beta);
        m.init();
        // velocity direction (isotropic) scaled to gamma_ij z^i z^j = W^2 - 1
        double cv = 2.0 * U(7) - 1.0, pv = 2.0 * M_PI * U(8), sv = std::sqrt(std::fmax(0.0, 1 - cv * cv));
        double nv[3] = {sv * std::cos(pv), sv * std::sin(pv), cv};
        double zn = std::sqrt(m.sqv(nv[0], nv[1], nv[2]));
        double zm = std::sqrt(W * W - 1.0) / zn;
        double zv[3] = {nv[0] * zm, nv[1] * zm, nv[2] * zm};
        // B direction (isotropic), magnitude from b^2 = sigma rho
        double cb = 2.0 * U(9) - 1.0, pb = 2.0 * M_PI * U(10), sb = std::sqrt(std::fmax(0.0, 1 - cb * cb));
        double nb[3] = {sb * std::cos(pb), sb * std::sin(pb), cb};
        double Wc = lorentz(m.g, zv);
        double bb[4], b2u;
        smallb(m, nb, zv, Wc, bb, b2u);
        double Bm = std::sqrt(sig * rho / b2u);
        double pr[NPRIM] = {0};
        pr[RHOL] = rho; pr[EPSL] = eps; pr[PRESSL] = gm1 * rho * eps; pr[TEMPL] = gm1 * eps;
        pr[ENTL] = std::log(eps) / gm1 - std::log(rho); pr[YEL] = 0.0;
        for (int i = 0; i < 3; ++i) { pr[ZXL + i] = zv[i]; pr[BXL + i] = Bm * nb[i]; }
        double cs[NCONS];
        p2c(pr, cs, m);
        for (int i = 0; i < 3; ++i) cs[BSXL + i] = m.sqrtg * pr[BXL + i];
        for (int v = 0; v < NPRIM; ++v) prims[v * n + s] = pr[v];
        for (int v = 0; v < NCONS; ++v) cons[v * n + s] = cs[v];
        for (int a2 = 0; a2 < 6; ++a2) met[a2 * n + s] = m.g[a2];
        met[6 * n + s] = m.alp;
        for (int a2 = 0; a2 < 3; ++a2) met[(7 + a2) * n + s] = m.beta[a2];
        if (rad) rad[s] = r;
    }
    return 0;
}

}  // extern "C"

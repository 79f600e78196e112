// oracle_capi.cpp -- C entry points of the CPU oracle (TEST INFRASTRUCTURE).
//
// Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg, through ctypes.  Takes the same parameter/grid structs and the same
// SoA / (i,j,k,var,q) layouts as the product ABI (include/grmhd_abi.h) so the
// same host arrays can be fed to both.
#include "grmhd_oracle_rhs.hh"
#include "../include/grmhd_abi.h"
#include <memory>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

namespace {

IdealGas make_ideal(const grmhd_params* p) { return IdealGas(p->gamma, p->temp_fl, p->eps_max); }

HybridPwpoly make_hybrid(const grmhd_params* p) {
    HybridPwpoly h(p->pwp_kappa0, p->pwp_gamma[0], p->gamma_th, p->entropy_min, p->eps_max, p->temp_fl);
    h.n = p->pwp_n;
    // piecewise construction (src/physics/eos/eos_storage.cpp:52-99)
    double rh[8] = {0};
    for (int i = 1; i < h.n; ++i) rh[i] = p->pwp_rho[i - 1];
    h.K[0] = p->pwp_kappa0; h.ep[0] = 0; h.pr[0] = 0;
    for (int i = 0; i < h.n; ++i) { h.G[i] = p->pwp_gamma[i]; h.rh[i] = rh[i]; }
    for (int i = 1; i < h.n; ++i) {
        h.K[i] = h.K[i - 1] * std::pow(rh[i], h.G[i - 1] - h.G[i]);
        h.ep[i] = h.ep[i - 1] + h.K[i - 1] * std::pow(rh[i], h.G[i - 1] - 1.) / (h.G[i - 1] - 1.)
                  - h.K[i] * std::pow(rh[i], h.G[i] - 1.) / (h.G[i] - 1.);
        h.pr[i] = h.K[i] * std::pow(rh[i], h.G[i]);
    }
    return h;
}

AtmoParams make_atmo(const grmhd_params* p) {
    AtmoParams a; a.rho_fl = p->rho_fl; a.temp_fl = p->temp_fl; a.ye_fl = p->ye_fl; a.atmo_tol = p->atmo_tol; return a;
}
C2PParams make_c2p(const grmhd_params* p) {
    C2PParams c; c.tol = p->c2p_tol; c.max_w = p->max_lorentz; c.max_sigma = p->max_sigma;
    c.beta_fallback = p->beta_fallback; c.use_ent_backup = p->use_entropy_backup != 0; c.alp_bh_thresh = p->bh_alp_thresh;
    return c;
}
ExcisionParams make_exc(const grmhd_params* p) {
    ExcisionParams e; e.r_ex = p->excision_radius; e.alp_ex = p->excision_lapse; e.excise_by_radius = p->excise_by_radius != 0;
    e.cks = p->excision_cks != 0; e.spin = p->excision_spin;
    return e;
}
Options make_opt(const grmhd_params* p) {
    Options o; o.recon = p->recon; o.riemann = p->riemann; o.emf = p->emf; o.thermo = p->thermo_recon;
    o.fofc = p->fofc; o.dmp_enable = p->fofc_dmp_enable; o.dmp_M = p->fofc_dmp_M;
    o.fofc_ghost_emf = p->fofc_ghost_emf;
    o.fd_order = p->fd_order; o.rusanov_c_limit = p->rusanov_c_limit; return o;
}

Metric met_from_soa(const double* met, int64_t n, int64_t s) {
    double g[6], b[3];
    for (int a = 0; a < 6; ++a) g[a] = met[a * n + s];
    for (int a = 0; a < 3; ++a) b[a] = met[(7 + a) * n + s];
    return Metric(g, b, met[6 * n + s]);
}

template <typename EOS>
void c2p_batch_impl(const grmhd_params* p, const EOS& eos, int64_t n, double* cons_soa, const double* met,
                    double* prims, uint64_t* errb, uint8_t* floored, int dry_run, const double* r,
                    const double* guess, double* roots, int64_t* evals, int64_t* fatal_out) {
    AtmoParams atmo = make_atmo(p);
    C2PParams cp = make_c2p(p);
    ExcisionParams ex = make_exc(p);
    int64_t fatal = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : fatal)
    for (int64_t s = 0; s < n; ++s) {
        Metric m = met_from_soa(met, n, s);
        double cons[NCONS], pr[NPRIM] = {0};
        for (int v = 0; v < NCONS; ++v) cons[v] = cons_soa[v * n + s];
        Bits err;
        C2PStats st;
        MuGuess g;
        if (guess) { g.e_lo = guess[s]; g.e_hi = guess[n + s]; g.s_lo = guess[2 * n + s]; g.s_hi = guess[3 * n + s]; }
        double rr = r ? r[s] : 1e300;
        C2PResult res = conservs_to_prims(cons, pr, m, eos, atmo, ex, cp, &rr, err, dry_run != 0, &st,
                                          (guess || roots) ? &g : nullptr);
        if (roots) { roots[s] = g.x_e; roots[n + s] = g.x_s; roots[2 * n + s] = g.m_e; roots[3 * n + s] = g.m_s; }
        if (res.fatal) fatal++;
        for (int v = 0; v < NPRIM; ++v) prims[v * n + s] = pr[v];
        for (int v = 0; v < NCONS; ++v) cons_soa[v * n + s] = cons[v];
        errb[s] = err.w;
        floored[s] = res.floored ? 1 : 0;
        if (evals) { evals[2 * s] = st.n_eval; evals[2 * s + 1] = st.n_bracket_eval; }
    }
    if (fatal_out) *fatal_out = fatal;
}

struct StepperBase {
    virtual ~StepperBase() {}
    virtual void rk3(double dt, double* st, double* B[3], double* aux) = 0;
    virtual void substep(double dt, double dtfact, double* ns, const double* os, double* nB[3], const double* oB[3],
                         const double* aux) = 0;
    virtual void ghosts(double* st, double* B[3]) = 0;
    virtual void aux(double* st, double* B[3], double* aux) = 0;
    virtual double dt(const double* st, const double* aux, double cfl) = 0;
    virtual int64_t fatal() = 0;
    virtual int64_t fofc_flagged() = 0;
    virtual const Forest* forest_of() const = 0;
    virtual void metric_ghosts(double* st) = 0;
    // the substep's scratch arrays: which 0 face fluxes, 1 face values
    // (UCT vbar / GS face EMFs), 2 edge EMFs, 3 cell E (GS); d = direction
    virtual const std::vector<double>* scratch(int which, int d) const = 0;
};

template <typename EOS>
struct StepperT : StepperBase {
    Layout L;
    EOS eos;
    std::unique_ptr<Stepper<EOS>> S;
    std::unique_ptr<Forest> forest;
    std::unique_ptr<AmrExchange> exch;
    StepperT(const grmhd_params* p, const grmhd_grid* g, const grmhd_instance* inst, int n_inst, const EOS& e)
        : L(Grid(g->nx, g->ny, g->nz, g->ngz, g->nq)), eos(e) {
        L.blk.assign(g->nq, BlockDesc{-1, {0, 0, 0}});
        for (int a = 0; a < n_inst; ++a) {
            Instance I;
            for (int d = 0; d < 3; ++d) { I.nb[d] = inst[a].nb[d]; I.idx[d] = inst[a].idx[d]; I.x0[d] = inst[a].x0[d]; }
            for (int d = 0; d < 6; ++d) I.bc[d] = inst[a].bc[d];
            I.q0 = inst[a].q0;
            L.inst.push_back(I);
            for (int bk = 0; bk < I.nb[2]; ++bk)
                for (int bj = 0; bj < I.nb[1]; ++bj)
                    for (int bi = 0; bi < I.nb[0]; ++bi)
                        L.blk[I.q0 + bi + (int64_t)I.nb[0] * (bj + (int64_t)I.nb[1] * bk)] = BlockDesc{a, {bi, bj, bk}};
        }
        S.reset(new Stepper<EOS>(L, eos));
        S->atmo = make_atmo(p); S->exc = make_exc(p); S->c2pp = make_c2p(p); S->opt = make_opt(p);
    }
    void rk3(double dt, double* st, double* B[3], double* aux) override { S->rk3_step(dt, st, B, aux); }
    void substep(double dt, double dtfact, double* ns, const double* os, double* nB[3], const double* oB[3],
                 const double* aux) override {
        Fields F{const_cast<double*>(os), {const_cast<double*>(oB[0]), const_cast<double*>(oB[1]), const_cast<double*>(oB[2])}};
        S->advance_substep(dt, dtfact, ns, nB, F, aux);
    }
    void ghosts(double* st, double* B[3]) override { S->fill_ghosts(st, B, 0, NHRSC); }
    void aux(double* st, double* B[3], double* aux) override { S->compute_auxiliaries(st, B, aux); }
    double dt(const double* st, const double* aux, double cfl) override { return S->find_dt(st, aux, cfl); }
    int64_t fatal() override { return S->n_fatal; }
    int64_t fofc_flagged() override { return S->fofc_cells_flagged; }
    const Forest* forest_of() const override { return forest.get(); }
    const std::vector<double>* scratch(int which, int d) const override {
        if (d < 0 || d > 2) return nullptr;
        switch (which) {
            case 0: return &S->flux[d];
            case 1: return &S->vbar[d];
            case 2: return &S->emf;
            case 3: return &S->ecen;
            default: return nullptr;
        }
    }
    // The Cowling metric's ghost exchange: on a forest copies + 5-point
    // Lagrange (AmrExchange::fill_metric); on uniform instances copies from
    // the same-level neighbour or periodic image, none beyond outflow faces.
    void metric_ghosts(double* st) override {
        if (exch) { exch->fill_metric(st, GXX_, 16); return; }
        const Grid& g = L.grid;
        const int n[3] = {g.nx, g.ny, g.nz}, Ne[3] = {g.Ni, g.Nj, g.Nk}, ng = g.ngz;
        std::vector<double> out;
        std::vector<int64_t> where;
        for (int64_t q = 0; q < g.nq; ++q) {
            const BlockDesc& bd = L.blk[q];
            const Instance& I = L.inst[bd.inst];
            for (int k = 0; k < Ne[2]; ++k)
                for (int j = 0; j < Ne[1]; ++j)
                    for (int i = 0; i < Ne[0]; ++i) {
                        const int loc[3] = {i, j, k};
                        bool ghost = false, outside = false;
                        int sb[3], sl[3];
                        for (int d = 0; d < 3; ++d) {
                            ghost |= loc[d] < ng || loc[d] >= n[d] + ng;
                            const int Nt = I.nb[d] * n[d];
                            int gg = bd.b[d] * n[d] + loc[d] - ng;
                            if ((gg < 0 && I.bc[2 * d] != BC_PERIODIC) || (gg >= Nt && I.bc[2 * d + 1] != BC_PERIODIC)) outside = true;
                            gg = ((gg % Nt) + Nt) % Nt;
                            sb[d] = gg / n[d]; sl[d] = gg % n[d] + ng;
                        }
                        if (!ghost || outside) continue;
                        const int64_t qs = I.q0 + sb[0] + (int64_t)I.nb[0] * (sb[1] + (int64_t)I.nb[1] * sb[2]);
                        for (int v = GXX_; v < GXX_ + 16; ++v) {
                            out.push_back(st[g.cc(sl[0], sl[1], sl[2], v, qs, NSTATE)]);
                            where.push_back(g.cc(i, j, k, v, q, NSTATE));
                        }
                    }
        }
        for (size_t x = 0; x < out.size(); ++x) st[where[x]] = out[x];
    }
};

// Two-level forest: one 1x1x1 instance per block at its level's spacing, the
// coarse-fine exchange and reflux switched on (grmhd_oracle_amr.hh).
template <typename EOS>
StepperBase* make_forest_stepper(const grmhd_params* p, const grmhd_grid* g, const grmhd_forest* f, const EOS& e) {
    const int n[3] = {g->nx, g->ny, g->nz};
    std::unique_ptr<Forest> F(new Forest(f->nb, f->bc, n, g->ngz, f->refined, f->n_refined));
    if (F->nq != g->nq) return nullptr;
    std::vector<grmhd_instance> inst(F->nq);
    for (int64_t q = 0; q < F->nq; ++q) {
        for (int d = 0; d < 3; ++d) {
            inst[q].nb[d] = 1; inst[q].idx[d] = f->idx[d] * (F->level[q] ? 2.0 : 1.0);
            inst[q].x0[d] = f->x0[d] + (double)F->crd[q][d] * n[d] / inst[q].idx[d];
        }
        for (int d = 0; d < 6; ++d) inst[q].bc[d] = f->bc[d];
        inst[q].q0 = q;
    }
    auto* st = new StepperT<EOS>(p, g, inst.data(), (int)F->nq, e);
    st->exch.reset(new AmrExchange(*F, NSTATE));
    st->forest = std::move(F);
    st->L.amr = st->exch.get();
    return st;
}
}  // namespace

extern "C" {

int orc_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n; return 1;
#endif
}

int orc_c2p_batch(const grmhd_params* p, int64_t n, double* cons_soa, const double* met, double* prims,
                  uint64_t* errb, uint8_t* floored, int dry_run, const double* r, const double* guess,
                  double* roots, int64_t* evals, int64_t* fatal) {
    if (p->eos_kind == GRMHD_EOS_IDEAL_GAS)
        c2p_batch_impl(p, make_ideal(p), n, cons_soa, met, prims, errb, floored, dry_run, r, guess, roots, evals, fatal);
    else
        c2p_batch_impl(p, make_hybrid(p), n, cons_soa, met, prims, errb, floored, dry_run, r, guess, roots, evals, fatal);
    return 0;
}

int orc_p2c_batch(const grmhd_params* p, int64_t n, const double* prims, const double* met, double* cons_soa) {
    (void)p;
    for (int64_t s = 0; s < n; ++s) {
        Metric m = met_from_soa(met, n, s);
        double pr[NPRIM], cons[NCONS];
        for (int v = 0; v < NPRIM; ++v) pr[v] = prims[v * n + s];
        for (int v = 0; v < NCONS; ++v) cons[v] = cons_soa[v * n + s];
        prims_to_conservs(pr, cons, m);
        for (int v = 0; v < 7; ++v) cons_soa[v * n + s] = cons[v];
    }
    return 0;
}

// EOS point call used by tests (press_eps_csnd2_entropy__temp_rho_ye)
double orc_eos_from_temp(const grmhd_params* p, double temp, double rho, double* eps, double* cs2, double* ent,
                         uint64_t* eos_err) {
    Bits e;
    double pr;
    if (p->eos_kind == GRMHD_EOS_IDEAL_GAS) pr = make_ideal(p).press_eps_csnd2_entropy__temp_rho_ye(*eps, *cs2, *ent, temp, rho, e);
    else pr = make_hybrid(p).press_eps_csnd2_entropy__temp_rho_ye(*eps, *cs2, *ent, temp, rho, e);
    if (eos_err) *eos_err = e.w;
    return pr;
}

// Point kernels for unit tests ------------------------------------------------
double orc_W(const double g[6], const double z[3]) { return get_W(g, z); }
void orc_smallb(double alp, const double beta[3], const double g[6], const double B[3], const double z[3], double W,
                double b[4], double* b2) { get_smallb(alp, beta, g, B, z, W, b, *b2); }
void orc_vtilde(double alp, const double beta[3], const double z[3], double W, double vt[3]) { get_vtilde(alp, beta, z, W, vt); }
void orc_cm_cp(double alp, const double beta[3], double rho, double press, double eps, double cs2, double W, double b2,
               const double vt[3], double guu_dd, int d, double* cm, double* cp) {
    get_cm_cp(alp, beta, rho, press, eps, cs2, W, b2, vt, guu_dd, d, *cm, *cp);
}
void orc_fluxes(double alp, const double beta[3], const double g[6], double rho, double press, double eps,
                const double z[3], double s, double W, double b2, const double b[4], const double vt[3], int d,
                double cons[6], double flux[6]) {
    double S[3], fS[3];
    get_fluxes(alp, beta, g, rho, press, eps, z, s, W, b2, b, vt, d, cons[0], cons[4], S, cons[5], flux[0], flux[4], fS, flux[5]);
    for (int i = 0; i < 3; ++i) { cons[1 + i] = S[i]; flux[1 + i] = fS[i]; }
}
void orc_geom_sources(double alp, const double beta[3], const double g[6], const double gu[6], const double K[6],
                      const double da[3], const double dg[18], const double db[9], double rho, double press, double eps,
                      const double B[3], const double z[3], double W, double out[4]) {
    double ss[3];
    get_geom_sources(alp, beta, g, gu, K, da, dg, db, rho, press, eps, B, z, W, out[0], ss);
    out[1] = ss[0]; out[2] = ss[1]; out[3] = ss[2];
}
void orc_recon(int kind, const double u[6], double out[2]) { recon_any(kind, u, out[0], out[1]); }
// face_interp_4 at the face between u[1] and u[2]; the centred FD first
// derivative of order 2/4/6 from u(-3..3)
double orc_face4(const double u[4]) { return face_interp_4(u[0], u[1], u[2], u[3]); }
double orc_fd_der(int order, const double u[7], double ih) { return fd_der_central(order, u, ih); }
void orc_metric(const double g[6], double gi[6], double* sqrtg) {
    double b[3] = {0, 0, 0};
    Metric m(g, b, 1.0);
    for (int i = 0; i < 6; ++i) gi[i] = m.gi[i];
    *sqrtg = m.sqrtg;
}
// Brent on a test polynomial family: f(x) = (x - r0)(x^2 + c)  -> root r0
double orc_brent_poly(double r0, double c, double a, double b, double tol, int64_t* nev) {
    struct F { double r0, c; int64_t n = 0; double operator()(double x) { ++n; return (x - r0) * (x * x + c); } } f{r0, c};
    double x = brent(f, a, b, tol);
    if (nev) *nev = f.n;
    return x;
}
// Kastaun master function (energy) at mu for one state, for bisection checks.
double orc_kastaun_f(const grmhd_params* p, const double cons_in[10], const double met[10], double mu, double* mu0_out) {
    double g[6], b[3];
    for (int a = 0; a < 6; ++a) g[a] = met[a];
    for (int a = 0; a < 3; ++a) b[a] = met[7 + a];
    Metric m(g, b, met[6]);
    double cons[NCONS];
    for (int v = 0; v < NCONS; ++v) cons[v] = cons_in[v] / m.sqrtg;
    IdealGas eos = make_ideal(p);
    C2PSetup cs(eos, m, cons);
    double h0 = cs.h0, mu0 = 1 / h0;
    if (cs.r2 >= h0 * h0) {
        FBrack gb(cs.r2, cs.Bt2, cs.rBt2, h0);
        double lo, hi; gb.bracket(lo, hi);
        int rerr;
        mu0 = newton_bracketed(lo, hi, gb, 30, 1e-10, rerr);
        if (rerr == 0) mu0 *= 1 + 1e-10; else mu0 = 1.0 / h0;
    }
    if (mu0_out) *mu0_out = mu0;
    FRoot<IdealGas> f(eos, cs.D, cs.q, cs.r2, cs.rBt2, cs.Bt2, cs.ye, h0);
    return f(mu);
}

// Master function of the energy (which=0) or entropy (which=1) solve on a
// grid of mu values, for one densitised state (ideal gas).
void orc_master_grid(const grmhd_params* p, const double cons_in[10], const double met[10], int which,
                     int64_t nmu, const double* mu, double* out, double* h0_out) {
    double g[6], b[3];
    for (int a = 0; a < 6; ++a) g[a] = met[a];
    for (int a = 0; a < 3; ++a) b[a] = met[7 + a];
    Metric m(g, b, met[6]);
    double cons[NCONS];
    for (int v = 0; v < NCONS; ++v) cons[v] = cons_in[v] / m.sqrtg;
    IdealGas eos = make_ideal(p);
    C2PSetup cs(eos, m, cons);
    if (h0_out) *h0_out = cs.h0;
    if (which == 0) {
        FRoot<IdealGas> f(eos, cs.D, cs.q, cs.r2, cs.rBt2, cs.Bt2, cs.ye, cs.h0);
        for (int64_t i = 0; i < nmu; ++i) out[i] = f(mu[i]);
    } else {
        EntFRoot<IdealGas> f(eos, cs.D, cs.r2, cs.rBt2, cs.Bt2, cs.s, cs.ye, cs.h0);
        for (int64_t i = 0; i < nmu; ++i) out[i] = f(mu[i]);
    }
}

// Block stepper ----------------------------------------------------------------
void* orc_stepper_create(const grmhd_params* p, const grmhd_grid* g, const grmhd_instance* inst, int n_inst) {
    if (p->eos_kind == GRMHD_EOS_IDEAL_GAS) return new StepperT<IdealGas>(p, g, inst, n_inst, make_ideal(p));
    return new StepperT<HybridPwpoly>(p, g, inst, n_inst, make_hybrid(p));
}
void* orc_stepper_create_forest(const grmhd_params* p, const grmhd_grid* g, const grmhd_forest* f) {
    if (p->eos_kind == GRMHD_EOS_IDEAL_GAS) return make_forest_stepper(p, g, f, make_ideal(p));
    return make_forest_stepper(p, g, f, make_hybrid(p));
}
// Level and coordinates of block q of a forest stepper (-1 if not a forest).
int orc_forest_block(void* h, int64_t q, int32_t* level, int32_t coords[3]);

void orc_stepper_destroy(void* h) { delete static_cast<StepperBase*>(h); }
void orc_rk3_step(void* h, double dt, double* st, double* Bx, double* By, double* Bz, double* aux) {
    double* B[3] = {Bx, By, Bz};
    static_cast<StepperBase*>(h)->rk3(dt, st, B, aux);
}
void orc_advance_substep(void* h, double dt, double dtfact, double* ns, const double* os, double* nBx, double* nBy,
                         double* nBz, const double* oBx, const double* oBy, const double* oBz, const double* aux) {
    double* nB[3] = {nBx, nBy, nBz};
    const double* oB[3] = {oBx, oBy, oBz};
    static_cast<StepperBase*>(h)->substep(dt, dtfact, ns, os, nB, oB, aux);
}
void orc_fill_ghosts(void* h, double* st, double* Bx, double* By, double* Bz) {
    double* B[3] = {Bx, By, Bz};
    static_cast<StepperBase*>(h)->ghosts(st, B);
}
void orc_compute_aux(void* h, double* st, double* Bx, double* By, double* Bz, double* aux) {
    double* B[3] = {Bx, By, Bz};
    static_cast<StepperBase*>(h)->aux(st, B, aux);
}
double orc_find_dt(void* h, const double* st, const double* aux, double cfl) {
    return static_cast<StepperBase*>(h)->dt(st, aux, cfl);
}
// Copy of a scratch array of the last substep (layouts of the product's
// context scratch: fluxes [q][7][face], face values [q][4][face], edge EMFs
// [q][3][edge], cell E [q][3][cell]); returns its length, copies when out != 0.
int64_t orc_stepper_scratch(void* h, int which, int d, double* out) {
    const std::vector<double>* v = static_cast<StepperBase*>(h)->scratch(which, d);
    if (!v) return -1;
    if (out) std::copy(v->begin(), v->end(), out);
    return (int64_t)v->size();
}
int64_t orc_fatal_count(void* h) { return static_cast<StepperBase*>(h)->fatal(); }
int64_t orc_fofc_flagged(void* h) { return static_cast<StepperBase*>(h)->fofc_flagged(); }
void orc_fill_metric_ghosts(void* h, double* st) { static_cast<StepperBase*>(h)->metric_ghosts(st); }
int orc_forest_block(void* h, int64_t q, int32_t* level, int32_t coords[3]) {
    const Forest* F = static_cast<StepperBase*>(h)->forest_of();
    if (!F || q < 0 || q >= F->nq) return -1;
    *level = F->level[q];
    for (int d = 0; d < 3; ++d) coords[d] = F->crd[q][d];
    return 0;
}

#ifdef ORACLE_COUNT_FLOPS
// Operation counts (liboracle_count.so only, see flop_count.hh):
// out[pass * 5 + kind], 20 pass ids x {add, mul, div, sqrt, pow/log/exp}.
// Reset also pins the library to one thread (the counters are plain).
void orc_flops_reset(void) {
    std::memset(orcflop::g_counts, 0, sizeof(orcflop::g_counts));
    orcflop::g_pass = 0;
    omp_set_num_threads(1);
}
void orc_flops_read(int64_t* out) { std::memcpy(out, orcflop::g_counts, sizeof(orcflop::g_counts)); }
#endif

}  // extern "C"

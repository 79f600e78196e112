// test_host_api.cpp -- C++ tests of the host interface above the C ABI
// (grace_amd/host/grace_mi355x.hh), written like the reference's own C++
// tests: the c2p round-trip sweep of test/test_c2p.cpp:130-216 (hybrid
// piecewise polytrope, Gamma=2, Gamma_th=1.8, K=168.5819, Minkowski,
// 576 samples, state-normalised error <= 1e-10, nothing floored), one
// substep of a magnetised blast wave against the CPU oracle (L-infinity
// relative <= 1e-12 on the 7 conserved and the 3 face-B arrays over the
// interior), and the error behaviour of the interface (an option
// combination the reference rejects throws grace_mi355x::error).
//
// Needs a GPU; built by grace_amd/build.py, run by tests/test_gpu_cpp_host.py.
// Prints one "[PASS] name" / "[FAIL] name: why" line per test and returns the
// number of failures.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../grace_amd/host/grace_mi355x.hh"
#include "../../include/grmhd_setup.h"

// CPU oracle (oracle/oracle_capi.cpp): the checker, linked by this test only.
extern "C" {
double orc_eos_from_temp(const grmhd_params* p, double temp, double rho, double* eps, double* cs2, double* ent,
                         uint64_t* eos_err);
int orc_p2c_batch(const grmhd_params* p, int64_t n, const double* prims, const double* met, double* cons_soa);
void* orc_stepper_create(const grmhd_params* p, const grmhd_grid* g, const grmhd_instance* inst, int n_inst);
void orc_stepper_destroy(void* h);
void orc_advance_substep(void* h, double dt, double dtfact, double* ns, const double* os, double* nBx, double* nBy,
                         double* nBz, const double* oBx, const double* oBy, const double* oBz, const double* aux);
}

using namespace grace_mi355x;

static int g_fail = 0;
static void report(const char* name, bool ok, const char* why = "") {
    std::printf("[%s] %s%s%s\n", ok ? "PASS" : "FAIL", name, ok ? "" : ": ", ok ? "" : why);
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

#define HIPCHECK(x)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) { std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); std::exit(99); } \
    } while (0)

template <typename T>
struct dev_buf {
    T* p = nullptr;
    size_t n = 0;
    explicit dev_buf(size_t n_) : n(n_) { HIPCHECK(hipMalloc(&p, sizeof(T) * (n ? n : 1))); HIPCHECK(hipMemset(p, 0, sizeof(T) * (n ? n : 1))); }
    ~dev_buf() { (void)hipFree(p); }
    void put(const std::vector<T>& h) { HIPCHECK(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice)); }
    std::vector<T> get() const { std::vector<T> h(n); HIPCHECK(hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost)); return h; }
};

// test/test_c2p.cpp:87-103: max |a-b| over RHO, EPS, z^i / max(|rho|, |eps|, |z|)
static double state_residual(const std::vector<double>& a, const std::vector<double>& b, int64_t n, int64_t s) {
    auto P = [&](const std::vector<double>& v, int i) { return v[i * n + s]; };
    const double zm = std::sqrt(P(b, 2) * P(b, 2) + P(b, 3) * P(b, 3) + P(b, 4) * P(b, 4));
    const double scale = std::max(std::max(std::fabs(P(b, 0)), std::fabs(P(b, 7))), std::max(zm, 1e-30));
    double e = 0;
    for (int i : {0, 7, 2, 3, 4}) e = std::max(e, std::fabs(P(a, i) - P(b, i)));
    return e / scale;
}

static void test_c2p_minkowski_hydro_sweep() {
    grmhd_params p;
    grmhd_default_params(&p);
    p.eos_kind = GRMHD_EOS_HYBRID_PWPOLY; p.pwp_n = 1; p.pwp_kappa0 = 168.5819; p.pwp_gamma[0] = 2.0;
    p.gamma_th = 1.8; p.eps_max = 1e5; p.entropy_min = 1e-13;
    grmhd_grid g{1, 1, 1, 3, 1};
    std::vector<grmhd_instance> inst(1);
    grmhd_instance& I = inst[0];
    std::memset(&I, 0, sizeof(I));
    I.nb[0] = I.nb[1] = I.nb[2] = 1; I.idx[0] = I.idx[1] = I.idx[2] = 1.0;
    grmhd_evolution ev(p, g, inst);
    // the sweep of test_c2p.cpp: 6 directions x 4 Lorentz factors x 4 temperatures x 6 densities
    const double dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const double Ws[4] = {1.05, 1.5, 2.0, 3.0}, Ts[4] = {1e-3, 1e-1, 1.0, 10.0};
    const double lrho[6] = {-12.0, -8.0, -5.0, -3.0, -1.5, -0.5};
    const int64_t n = 6 * 4 * 4 * 6;
    std::vector<double> prims(13 * n, 0.0), met(10 * n, 0.0), cons(10 * n, 0.0);
    int64_t s = 0;
    for (int d = 0; d < 6; ++d)
        for (int w = 0; w < 4; ++w)
            for (int t = 0; t < 4; ++t)
                for (int r = 0; r < 6; ++r, ++s) {
                    const double rho = std::pow(10.0, lrho[r]), zm = std::sqrt(Ws[w] * Ws[w] - 1.0);
                    double eps, cs2, ent;
                    const double pr = orc_eos_from_temp(&p, Ts[t], rho, &eps, &cs2, &ent, nullptr);
                    const double v[13] = {rho, pr, zm * dirs[d][0], zm * dirs[d][1], zm * dirs[d][2], 0.0, Ts[t],
                                          eps, ent, 0.0, 0.0, 0.0, 0.0};
                    for (int k = 0; k < 13; ++k) prims[k * n + s] = v[k];
                    const double m[10] = {1, 0, 0, 1, 0, 1, 1, 0, 0, 0};
                    for (int k = 0; k < 10; ++k) met[k * n + s] = m[k];
                }
    orc_p2c_batch(&p, n, prims.data(), met.data(), cons.data());
    dev_buf<double> dc(10 * n), dm(10 * n), dp(13 * n);
    dev_buf<uint64_t> de(n);
    dev_buf<uint8_t> df(n);
    dc.put(cons); dm.put(met);
    ev.conservs_to_prims(n, dc.p, dm.p, dp.p, de.p, df.p);
    HIPCHECK(hipDeviceSynchronize());
    const std::vector<double> got = dp.get();
    const std::vector<uint8_t> fl = df.get();
    double worst = 0;
    int floored = 0;
    for (int64_t k = 0; k < n; ++k) { worst = std::max(worst, state_residual(got, prims, n, k)); floored += fl[k]; }
    char why[128];
    std::snprintf(why, sizeof(why), "max residual %.3e, %d floored", worst, floored);
    report("c2p_minkowski_hydro_sweep (576 samples, <= 1e-10)", worst <= 1e-10 && floored == 0, why);
}

static void test_blast_substep_against_oracle() {
    grmhd_params p;
    grmhd_default_params(&p);
    p.gamma = 4.0 / 3.0; p.flat_metric = 1;
    const int n = 16, ngz = 3;
    grmhd_grid g{n, n, n, ngz, 8};
    std::vector<grmhd_instance> inst(1);
    grmhd_instance& I = inst[0];
    std::memset(&I, 0, sizeof(I));
    I.nb[0] = I.nb[1] = I.nb[2] = 2; I.idx[0] = I.idx[1] = I.idx[2] = 2.0 * n;   // 32^3 on [-0.5, 0.5]^3
    const int N = n + 2 * ngz;
    const int64_t ncell = (int64_t)N * N * N, nface = (int64_t)(N + 1) * N * N;
    std::vector<double> st(g.nq * GRMHD_NSTATE * ncell), aux(g.nq * GRMHD_NAUX * ncell);
    std::vector<double> B[3];
    for (auto& b : B) b.assign(g.nq * nface, 0.0);
    const double lo[3] = {-0.5, -0.5, -0.5}, hi[3] = {0.5, 0.5, 0.5}, Bv[3] = {0.1, 0.0, 0.0};
    if (grmhd_setup_blast(&g, inst.data(), 1, p.gamma, lo, hi, 1e-2, 1e-3, 1.0, 0.1, Bv, 0.01, 20261202ull,
                          st.data(), B[0].data(), B[1].data(), B[2].data(), aux.data()) != 0) {
        report("blast_substep_against_oracle", false, "setup failed");
        return;
    }
    const double dt = 0.25 / (2.0 * n);
    // oracle
    std::vector<double> ns_o = st;
    std::vector<double> nB_o[3] = {B[0], B[1], B[2]};
    void* orc = orc_stepper_create(&p, &g, inst.data(), 1);
    orc_advance_substep(orc, dt, 1.0, ns_o.data(), st.data(), nB_o[0].data(), nB_o[1].data(), nB_o[2].data(),
                        B[0].data(), B[1].data(), B[2].data(), aux.data());
    orc_stepper_destroy(orc);
    // library through the C++ interface
    grmhd_evolution ev(p, g, inst);
    dev_buf<double> d_st(st.size()), d_ns(st.size()), d_aux(aux.size()), d_dt(1);
    dev_buf<double> d_B0(nface * g.nq), d_B1(nface * g.nq), d_B2(nface * g.nq);
    dev_buf<double> d_nB0(nface * g.nq), d_nB1(nface * g.nq), d_nB2(nface * g.nq);
    d_st.put(st); d_ns.put(st); d_aux.put(aux); d_dt.put(std::vector<double>{dt});
    d_B0.put(B[0]); d_B1.put(B[1]); d_B2.put(B[2]); d_nB0.put(B[0]); d_nB1.put(B[1]); d_nB2.put(B[2]);
    ev.advance_substep(d_dt.p, 1.0, var_array_t{d_ns.p}, var_array_t{d_st.p},
                       staggered_variable_arrays_t{d_nB0.p, d_nB1.p, d_nB2.p},
                       staggered_variable_arrays_t{d_B0.p, d_B1.p, d_B2.p}, var_array_t{d_aux.p});
    HIPCHECK(hipDeviceSynchronize());
    const std::vector<double> ns_g = d_ns.get();
    const std::vector<double> nB_g[3] = {d_nB0.get(), d_nB1.get(), d_nB2.get()};
    // L-infinity relative over the interior, per conserved variable / face array
    double worst = 0;
    for (int v = 0; v < 7; ++v) {
        double num = 0, den = 0;
        for (int64_t q = 0; q < g.nq; ++q)
            for (int k = ngz; k < ngz + n; ++k)
                for (int j = ngz; j < ngz + n; ++j)
                    for (int i = ngz; i < ngz + n; ++i) {
                        const int64_t idx = (q * GRMHD_NSTATE + v) * ncell + i + (int64_t)N * (j + (int64_t)N * k);
                        num = std::max(num, std::fabs(ns_g[idx] - ns_o[idx]));
                        den = std::max(den, std::fabs(ns_o[idx]));
                    }
        worst = std::max(worst, den > 0 ? num / den : num);
    }
    for (int d = 0; d < 3; ++d) {
        const int E[3] = {N + (d == 0), N + (d == 1), N + (d == 2)};
        double num = 0, den = 0;
        for (int64_t q = 0; q < g.nq; ++q)
            for (int k = ngz; k < ngz + n + (d == 2); ++k)
                for (int j = ngz; j < ngz + n + (d == 1); ++j)
                    for (int i = ngz; i < ngz + n + (d == 0); ++i) {
                        const int64_t idx = q * nface + i + (int64_t)E[0] * (j + (int64_t)E[1] * k);
                        num = std::max(num, std::fabs(nB_g[d][idx] - nB_o[d][idx]));
                        den = std::max(den, std::fabs(nB_o[d][idx]));
                    }
        worst = std::max(worst, den > 0 ? num / den : num);
    }
    char why[96];
    std::snprintf(why, sizeof(why), "max L-inf relative %.3e", worst);
    report("blast_substep_against_oracle (<= 1e-12)", worst <= 1e-12, why);
}

static void test_rejected_options_throw() {
    grmhd_params p;
    grmhd_default_params(&p);
    p.fofc = 1;   // FOFC with the UCT EMF: the reference rejects it (evolve.cpp:479-481)
    grmhd_grid g{8, 8, 8, 4, 1};
    std::vector<grmhd_instance> inst(1);
    std::memset(&inst[0], 0, sizeof(grmhd_instance));
    inst[0].nb[0] = inst[0].nb[1] = inst[0].nb[2] = 1;
    inst[0].idx[0] = inst[0].idx[1] = inst[0].idx[2] = 8.0;
    bool threw = false;
    std::string msg;
    try {
        grmhd_evolution ev(p, g, inst);
    } catch (const error& e) {
        threw = true;
        msg = e.what();
    }
    report("rejected_options_throw", threw && msg.find("UCT") != std::string::npos, msg.c_str());
}

// Two-level forest through the C++ layer: the block order (coarse leaves, then
// the 8 Morton children of each refined block) and the reference's
// constraints (cubic cells) enforced with an exception.
static void test_forest_layout_and_checks() {
    grmhd_params p;
    grmhd_default_params(&p);
    grmhd_grid g{8, 8, 8, 4, 8 + 7 * 2};
    const int32_t refined[2] = {7, 0};
    grmhd_forest f;
    std::memset(&f, 0, sizeof(f));
    f.nb[0] = f.nb[1] = f.nb[2] = 2;
    for (int d = 0; d < 6; ++d) f.bc[d] = GRMHD_BC_PERIODIC;
    f.idx[0] = f.idx[1] = f.idx[2] = 16.0;
    f.n_refined = 2;
    f.refined = refined;
    bool ok = true;
    std::string why;
    try {
        grmhd_evolution ev(p, g, f);
        int32_t lev, c[3];
        ev.forest_block(0, lev, c);
        ok &= lev == 0 && c[0] == 1 && c[1] == 0 && c[2] == 0;
        ev.forest_block(6, lev, c);
        ok &= lev == 1 && c[0] == 0 && c[1] == 0 && c[2] == 0;     // child 0 of coarse block 0
        ev.forest_block(13, lev, c);
        ok &= lev == 1 && c[0] == 1 && c[1] == 1 && c[2] == 1;     // child 7 of coarse block 0
        ev.forest_block(21, lev, c);
        ok &= lev == 1 && c[0] == 3 && c[1] == 3 && c[2] == 3;     // child 7 of coarse block 7
        if (!ok) why = "block order";
    } catch (const error& e) {
        ok = false;
        why = e.what();
    }
    f.idx[2] = 8.0;   // anisotropic cells
    bool threw = false;
    try {
        grmhd_evolution ev(p, g, f);
    } catch (const error& e) {
        threw = std::string(e.what()).find("cubic") != std::string::npos;
    }
    if (!threw && why.empty()) why = "anisotropic forest accepted";
    report("forest_layout_and_checks", ok && threw, why.c_str());
}

int main() {
    test_c2p_minkowski_hydro_sweep();
    test_blast_substep_against_oracle();
    test_rejected_options_throw();
    test_forest_layout_and_checks();
    std::printf("%d failure(s)\n", g_fail);
    return g_fail;
}

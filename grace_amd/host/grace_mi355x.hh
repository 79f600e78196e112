// grace_mi355x.hh -- C++ host interface above the C ABI of libgrmhd_mi355x.so,
// with the names and argument meaning of the reference's evolution and c2p
// entry points, for C++ callers (GRACE's evolution module, the C++ tests).
//
//   grmhd_evolution::advance_substep        evolve.hh:252-256 (advance_substep<eos_t>)
//   grmhd_evolution::compute_auxiliary_quantities
//                                           auxiliaries.hh:49-54
//   grmhd_evolution::find_stable_timestep   find_stable_timestep.cpp:72-126
//   grmhd_evolution::evolve_rk3             the rk3 branch of evolve_impl, evolve.cpp:190-225
//   grmhd_evolution::conservs_to_prims      c2p.hh:342-353 over a batch of cells
//   grmhd_evolution::prims_to_conservs      c2p.hh:355-358 over a batch of cells
//
// Arrays are device pointers in the reference's layout: var_array_t is
// (i,j,k,var,q) LayoutLeft with the context's block extents (n+2ngz per
// direction), the face arrays of staggered_variable_arrays_t are
// (n+2ngz+1 along their normal, n+2ngz otherwise, q).  Where the reference
// aborts (ERROR / Kokkos::abort) this layer throws grace_mi355x::error with
// the library's message; per-cell c2p outcomes are error bits, not errors.
#ifndef GRACE_MI355X_HH
#define GRACE_MI355X_HH

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grmhd_abi.h"
#include "../../include/grmhd_mlp.h"

namespace grace_mi355x {

struct error : std::runtime_error {
    grmhd_status status;
    error(const std::string& what, grmhd_status s) : std::runtime_error(what), status(s) {}
};

inline void check(grmhd_status s, const char* what) {
    if (s != GRMHD_OK) throw error(std::string(what) + ": " + grmhd_last_error(), s);
}

// Device views in the reference's layout (see the file comment).
struct var_array_t {
    double* data;
};
struct staggered_variable_arrays_t {
    double* face_staggered_fields_x;
    double* face_staggered_fields_y;
    double* face_staggered_fields_z;
    double* const* faces() const { return &face_staggered_fields_x; }
};

// One context: the parameters, the block list and the device scratch of the
// evolution (fluxes, EMFs, stage buffers, graph).  Not copyable.
class grmhd_evolution {
public:
    grmhd_evolution(const grmhd_params& p, const grmhd_grid& g, const std::vector<grmhd_instance>& inst)
        : grid_(g) {
        check(grmhd_ctx_create(&p, &g, inst.data(), (int32_t)inst.size(), &ctx_), "grmhd_ctx_create");
    }
    // A two-level forest (coarse-fine ghosts and refluxing in every call).
    grmhd_evolution(const grmhd_params& p, const grmhd_grid& g, const grmhd_forest& f) : grid_(g) {
        check(grmhd_ctx_create_forest(&p, &g, &f, &ctx_), "grmhd_ctx_create_forest");
    }
    ~grmhd_evolution() { grmhd_ctx_destroy(ctx_); }
    grmhd_evolution(const grmhd_evolution&) = delete;
    grmhd_evolution& operator=(const grmhd_evolution&) = delete;

    grmhd_ctx* handle() const { return ctx_; }
    const grmhd_grid& grid() const { return grid_; }

    // new_state = dtfact-weighted substep of old_state (and the face B), the
    // metric read from old_state; dt_dev is a device double.
    void advance_substep(const double* dt_dev, double dtfact, var_array_t new_state, var_array_t old_state,
                         staggered_variable_arrays_t new_stag, staggered_variable_arrays_t old_stag, var_array_t aux,
                         hipStream_t stream = nullptr) {
        const double* ob[3] = {old_stag.face_staggered_fields_x, old_stag.face_staggered_fields_y,
                               old_stag.face_staggered_fields_z};
        check(grmhd_advance_substep(ctx_, dt_dev, dtfact, new_state.data, old_state.data, new_stag.faces(), ob,
                                    aux.data, (grmhd_stream_t)stream),
              "advance_substep");
    }

    // The reference's signature advance_substep(t, dt, dtfact, ...) with a
    // host dt (t unused by the Cowling GRMHD substep).
    void advance_substep(double t, double dt, double dtfact, var_array_t new_state, var_array_t old_state,
                         staggered_variable_arrays_t new_stag, staggered_variable_arrays_t old_stag, var_array_t aux,
                         hipStream_t stream = nullptr) {
        const double* ob[3] = {old_stag.face_staggered_fields_x, old_stag.face_staggered_fields_y,
                               old_stag.face_staggered_fields_z};
        check(grmhd_advance_substep_t(ctx_, t, dt, dtfact, new_state.data, old_state.data, new_stag.faces(), ob,
                                      aux.data, (grmhd_stream_t)stream),
              "advance_substep");
    }

    // Waits for the stream; throws (status GRMHD_ERR_C2P_FATAL) if a c2p
    // pass met a cell where the reference aborts since the last sync point.
    void sync(hipStream_t stream = nullptr) { check(grmhd_sync(ctx_, (grmhd_stream_t)stream), "sync"); }

    // c2p over every cell (ghosts included), cell-centred B from the faces.
    void compute_auxiliary_quantities(var_array_t state, staggered_variable_arrays_t stag, var_array_t aux,
                                      hipStream_t stream = nullptr) {
        const double* b[3] = {stag.face_staggered_fields_x, stag.face_staggered_fields_y, stag.face_staggered_fields_z};
        check(grmhd_compute_auxiliaries(ctx_, state.data, b, aux.data, (grmhd_stream_t)stream),
              "compute_auxiliary_quantities");
    }

    // The Cowling metric's ghost exchange (copies; 5-point Lagrange across
    // coarse-fine faces of a forest), once after the initial data.
    void exchange_metric_ghosts(var_array_t state, hipStream_t stream = nullptr) {
        check(grmhd_fill_metric_ghosts(ctx_, state.data, (grmhd_stream_t)stream), "exchange_metric_ghosts");
    }

    // dt = cfl * dx / max characteristic speed, left on the device.
    void find_stable_timestep(var_array_t state, var_array_t aux, double cfl, double* dt_dev,
                              hipStream_t stream = nullptr) {
        check(grmhd_compute_dt(ctx_, state.data, aux.data, cfl, dt_dev, (grmhd_stream_t)stream), "find_stable_timestep");
    }

    // The reference's find_stable_timestep(): dt returned on the host (a
    // sync point, as sync()).
    double find_stable_timestep(var_array_t state, var_array_t aux, double cfl, hipStream_t stream = nullptr) {
        double dt = 0.0;
        check(grmhd_find_stable_timestep(ctx_, state.data, aux.data, cfl, &dt, (grmhd_stream_t)stream),
              "find_stable_timestep");
        return dt;
    }

    // One SSP-RK3 step in place; graph = true replays a captured hipGraph
    // (needs a non-default stream).
    void evolve_rk3(const double* dt_dev, var_array_t state, staggered_variable_arrays_t stag, var_array_t aux,
                    hipStream_t stream = nullptr, bool graph = false) {
        if (graph)
            check(grmhd_rk3_step_graph(ctx_, dt_dev, state.data, stag.faces(), aux.data, (grmhd_stream_t)stream),
                  "evolve_rk3");
        else
            check(grmhd_rk3_step(ctx_, dt_dev, state.data, stag.faces(), aux.data, (grmhd_stream_t)stream),
                  "evolve_rk3");
    }

    // conservs_to_prims over n cells (SoA device arrays: cons[10][n]
    // densitised, metric[10][n] = gxx gxy gxz gyy gyz gzz alp betax betay
    // betaz, prims[13][n]); cons may be rewritten by resets, as the
    // reference's by-reference argument.
    void conservs_to_prims(int64_t n, double* cons_soa, const double* metric_soa, double* prims_soa,
                           uint64_t* err_bits, uint8_t* floored, bool dry_run = false,
                           const double* radius_or_null = nullptr, hipStream_t stream = nullptr) {
        check(grmhd_c2p_batch(ctx_, n, cons_soa, metric_soa, prims_soa, err_bits, floored, dry_run ? 1 : 0,
                              radius_or_null, nullptr, nullptr, (grmhd_stream_t)stream),
              "conservs_to_prims");
    }

    void prims_to_conservs(int64_t n, const double* prims_soa, const double* metric_soa, double* cons_soa,
                           hipStream_t stream = nullptr) {
        check(grmhd_prims_to_conservs_batch(ctx_, n, prims_soa, metric_soa, cons_soa, (grmhd_stream_t)stream),
              "prims_to_conservs");
    }

    // Level (0 coarse, 1 fine) and coordinates at that level of block q of a forest.
    void forest_block(int64_t q, int32_t& level, int32_t coords[3]) const {
        check(grmhd_forest_block(ctx_, q, &level, coords), "forest_block");
    }

    grmhd_c2p_stats_t c2p_stats(hipStream_t stream = nullptr) {
        grmhd_c2p_stats_t s;
        check(grmhd_c2p_stats(ctx_, &s, (grmhd_stream_t)stream), "c2p_stats");
        return s;
    }

private:
    grmhd_ctx* ctx_ = nullptr;
    grmhd_grid grid_;
};

}  // namespace grace_mi355x

#endif

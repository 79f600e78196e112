// state_io.cpp -- HDF5 state dumps in the reference's checkpoint-data layout
// (include/grmhd_io.h).  Host code only.
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/grmhd_io.h"

namespace {

thread_local char g_err[512] = "";
void set_err(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int64_t cells(const grmhd_grid& g) {
    return (int64_t)(g.nx + 2 * g.ngz) * (g.ny + 2 * g.ngz) * (g.nz + 2 * g.ngz);
}
int64_t faces(const grmhd_grid& g, int d) {
    const int64_t e[3] = {g.nx + 2 * g.ngz + (d == 0), g.ny + 2 * g.ngz + (d == 1), g.nz + 2 * g.ngz + (d == 2)};
    return e[0] * e[1] * e[2];
}

// RAII for HDF5 handles
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H() { if (id >= 0) close(id); }
    operator hid_t() const { return id; }
};

bool write_attr(hid_t loc, const char* name, hid_t type, const void* v) {
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    H at(H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return at >= 0 && H5Awrite(at, type, v) >= 0;
}
bool read_attr(hid_t loc, const char* name, hid_t type, void* v) {
    if (H5Aexists(loc, name) <= 0) return false;
    H at(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    return at >= 0 && H5Aread(at, type, v) >= 0;
}

// 1-D double dataset, chunked like write_data_hdf5 (8 MB chunks)
bool write_1d(hid_t file, const char* name, const double* data, int64_t n) {
    hsize_t dims[1] = {(hsize_t)n};
    H sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
    H pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    hsize_t chunk[1] = {(hsize_t)(n < (8 << 20) / 8 ? n : (8 << 20) / 8)};
    if (n > 0) H5Pset_chunk(pl, 1, chunk);
    H ds(H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, pl, H5P_DEFAULT), H5Dclose);
    return ds >= 0 && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}
bool read_1d(hid_t file, const char* name, double* data, int64_t n) {
    H ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (ds < 0) return false;
    H sp(H5Dget_space(ds), H5Sclose);
    if (H5Sget_simple_extent_npoints(sp) != n) return false;
    return H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

const char* kFaceNames[3] = {"FaceCenteredDataX", "FaceCenteredDataY", "FaceCenteredDataZ"};
const char* kExtentNames[3] = {"grid_extent_x", "grid_extent_y", "grid_extent_z"};

}  // namespace

extern "C" const char* grmhd_io_last_error(void) { return g_err; }

extern "C" grmhd_status grmhd_io_write(const char* path, double time, uint32_t iteration, const grmhd_grid* g,
                                       int32_t nvars, const double extents[6], const double* state,
                                       const double* const B[3]) {
    if (!path || !g || !extents || !state || !B || nvars < 1 || g->nx < 1 || g->ny < 1 || g->nz < 1 ||
        g->ngz < 0 || g->nq < 1) {
        set_err("invalid arguments");
        return GRMHD_ERR_INVALID;
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file < 0) { set_err("cannot create %s", path); return GRMHD_ERR_INVALID; }
    bool ok = write_attr(file, "Time", H5T_NATIVE_DOUBLE, &time) &&
              write_attr(file, "Iteration", H5T_NATIVE_UINT, &iteration);
    const int64_t meta[6] = {g->nx, g->ny, g->nz, g->ngz, g->nq, nvars};
    const char* mnames[6] = {"nx", "ny", "nz", "ngz", "nq", "nvars"};
    for (int a = 0; a < 6 && ok; ++a) ok = write_attr(file, mnames[a], H5T_NATIVE_INT64, &meta[a]);
    for (int d = 0; d < 3 && ok; ++d) ok = write_1d(file, kExtentNames[d], extents + 2 * d, 2);
    ok = ok && write_1d(file, "CellCenteredData", state, cells(*g) * nvars * g->nq);
    for (int d = 0; d < 3 && ok; ++d) ok = write_1d(file, kFaceNames[d], B[d], faces(*g, d) * g->nq);
    if (!ok) { set_err("HDF5 write of %s failed", path); return GRMHD_ERR_INVALID; }
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_io_info(const char* path, grmhd_grid* g, int32_t* nvars, double* time,
                                      uint32_t* iteration, double extents[6]) {
    if (!path || !g) { set_err("invalid arguments"); return GRMHD_ERR_INVALID; }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file < 0) { set_err("cannot open %s", path); return GRMHD_ERR_INVALID; }
    int64_t meta[6];
    const char* mnames[6] = {"nx", "ny", "nz", "ngz", "nq", "nvars"};
    for (int a = 0; a < 6; ++a)
        if (!read_attr(file, mnames[a], H5T_NATIVE_INT64, &meta[a])) {
            set_err("%s: attribute %s missing", path, mnames[a]);
            return GRMHD_ERR_INVALID;
        }
    g->nx = (int32_t)meta[0]; g->ny = (int32_t)meta[1]; g->nz = (int32_t)meta[2]; g->ngz = (int32_t)meta[3];
    g->nq = meta[4];
    if (nvars) *nvars = (int32_t)meta[5];
    double t = 0;
    uint32_t it = 0;
    if ((time && !read_attr(file, "Time", H5T_NATIVE_DOUBLE, &t)) ||
        (iteration && !read_attr(file, "Iteration", H5T_NATIVE_UINT, &it))) {
        set_err("%s: Time/Iteration missing", path);
        return GRMHD_ERR_INVALID;
    }
    if (time) *time = t;
    if (iteration) *iteration = it;
    if (extents)
        for (int d = 0; d < 3; ++d)
            if (!read_1d(file, kExtentNames[d], extents + 2 * d, 2)) {
                set_err("%s: %s missing", path, kExtentNames[d]);
                return GRMHD_ERR_INVALID;
            }
    return GRMHD_OK;
}

extern "C" grmhd_status grmhd_io_read(const char* path, double* state, double* const B[3]) {
    if (!path || !state || !B) { set_err("invalid arguments"); return GRMHD_ERR_INVALID; }
    grmhd_grid g;
    int32_t nv = 0;
    grmhd_status st = grmhd_io_info(path, &g, &nv, nullptr, nullptr, nullptr);
    if (st != GRMHD_OK) return st;
    H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file < 0) { set_err("cannot open %s", path); return GRMHD_ERR_INVALID; }
    bool ok = read_1d(file, "CellCenteredData", state, cells(g) * nv * g.nq);
    for (int d = 0; d < 3 && ok; ++d) ok = read_1d(file, kFaceNames[d], B[d], faces(g, d) * g.nq);
    if (!ok) { set_err("%s: data sets missing or of the wrong size", path); return GRMHD_ERR_INVALID; }
    return GRMHD_OK;
}

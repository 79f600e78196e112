/* grmhd_io.h -- state dumps in the layout of the reference's checkpoint data
 * file (src/system/checkpoint_handler.cpp:421-640, write_data_hdf5 at
 * :308-400): an HDF5 file with the scalar attributes "Time" (double) and
 * "Iteration" (uint32), the datasets "grid_extent_x|y|z" ([min, max]), and the
 * 1-D double datasets "CellCenteredData" (the (i,j,k,var,q) state array,
 * i fastest, q slowest) and "FaceCenteredDataX|Y|Z" (the face-B arrays).
 * The block grid (nx, ny, nz, ngz, nq) and the state variable count are
 * stored as int64 attributes so that a dump can be read without the forest
 * file the reference keeps next to it.  Host buffers only: the caller copies
 * device arrays in and out.  Built as a separate library (libgrmhd_io.so,
 * linked against HDF5) so the compute library has no I/O dependency.
 */
#ifndef GRMHD_IO_H
#define GRMHD_IO_H

#include "grmhd_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* state: nvars * (nx+2ngz)(ny+2ngz)(nz+2ngz) * nq doubles;
 * B[d]: (n_d+2ngz+1) * (others+2ngz) * nq doubles. */
grmhd_status grmhd_io_write(const char* path, double time, uint32_t iteration, const grmhd_grid* grid,
                            int32_t nvars, const double extents[6], const double* state,
                            const double* const B[3]);

/* Grid, variable count, time, iteration and extents of a dump. */
grmhd_status grmhd_io_info(const char* path, grmhd_grid* grid, int32_t* nvars, double* time, uint32_t* iteration,
                           double extents[6]);

/* Reads a dump into caller buffers sized for the grid grmhd_io_info reports. */
grmhd_status grmhd_io_read(const char* path, double* state, double* const B[3]);

const char* grmhd_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

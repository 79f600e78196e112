"""HDF5 state dumps (include/grmhd_io.h) in the reference's checkpoint-data
layout: scalar attributes Time/Iteration, grid_extent_{x,y,z} and the 1-D
datasets CellCenteredData / FaceCenteredData{X,Y,Z} holding the
(i,j,k,var,q) arrays as they lie in memory (checkpoint_handler.cpp:287-400)."""
import os

import numpy as np
import pytest

import grace_amd as ga
from grace_amd import workloads as W


def test_state_dump_round_trip(tmp_path):
    wl = W.blast(nb=(2, 1, 1), n=8)
    path = str(tmp_path / "checkpoint_data_000042.h5")
    ext = (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    ga.write_state(path, wl.grid, wl.state, wl.B, time=0.125, iteration=42, extents=ext)
    g, st, B, t, it, e = ga.read_state(path)
    assert (g.nx, g.ny, g.nz, g.ngz, g.nq) == (8, 8, 8, 3, 2)
    assert t == 0.125 and it == 42 and np.array_equal(e, ext)
    assert st.shape == wl.state.shape and np.array_equal(st, wl.state)
    for d in range(3):
        assert np.array_equal(B[d], wl.B[d])


def test_state_dump_errors(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        ga.read_state(str(tmp_path / "missing.h5"))
    bad = tmp_path / "not_hdf5.h5"
    bad.write_bytes(b"not an hdf5 file")
    with pytest.raises(RuntimeError):
        ga.read_state(str(bad))

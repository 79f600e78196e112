"""The outflow-ghost c2p copies (grmhd.hip, copy_gather): beyond an outflow
face the ghost fill clamps cell indices to the boundary cell and face indices
to the boundary face (boundary_conditions.cpp:55-81; the library's map_cc /
map_fc), so every ghost cell's c2p inputs -- conserved and the cell-centred B
averaged from its two faces per direction -- equal those of the cell with each
outflow-exceeding coordinate moved to the first ghost layer.  Checked here on
random data for every ghost position of grids with outflow, periodic and mixed
faces; a planted mistake (the B average taken over a face pair that is not
clamped) fails."""
import itertools

import numpy as np
import pytest


def map_cc(g, N, lo, hi):
    if g < 0:
        return g % N if lo == 1 else 0
    if g >= N:
        return g % N if hi == 1 else N - 1
    return g


def map_fc(f, N, lo, hi):
    if f < 0:
        return f % N if lo == 1 else 0
    if f > N:
        return f % N if hi == 1 else N
    return f


def inputs(C, Bf, Nt, bc, p, shift_hi=1):
    src = tuple(map_cc(p[d], Nt[d], bc[2 * d], bc[2 * d + 1]) for d in range(3))

    def bface(d, q):
        idx = tuple(map_fc(q[e], Nt[e], bc[2 * e], bc[2 * e + 1]) if e == d
                    else map_cc(q[e], Nt[e], bc[2 * e], bc[2 * e + 1]) for e in range(3))
        return Bf[d][idx]
    bcc = []
    for d in range(3):
        q = list(p)
        q[d] += shift_hi
        bcc.append(0.5 * (bface(d, p) + bface(d, q)))
    return C[src], tuple(bcc)


def first_layer(p, Nt, bc):
    r = list(p)
    for d in range(3):
        if p[d] < 0 and bc[2 * d] == 0:
            r[d] = -1
        if p[d] >= Nt[d] and bc[2 * d + 1] == 0:
            r[d] = Nt[d]
    return tuple(r)


def copies(Nt, ng, bc, shift_hi=1):
    rng = np.random.default_rng(7)
    C = rng.random(Nt)
    Bf = [rng.random(tuple(Nt[e] + (e == d) for e in range(3))) for d in range(3)]
    n = bad = 0
    for p in itertools.product(*[range(-ng, Nt[d] + ng) for d in range(3)]):
        r = first_layer(p, Nt, bc)
        if r == p:
            continue
        n += 1
        if inputs(C, Bf, Nt, bc, p, shift_hi) != inputs(C, Bf, Nt, bc, r, shift_hi):
            bad += 1
    return n, bad


@pytest.mark.parametrize("Nt,ng,bc", [((8, 8, 8), 3, (0,) * 6), ((8, 6, 4), 3, (0, 0, 1, 1, 0, 0)),
                                      ((5, 4, 6), 4, (1, 1, 0, 0, 0, 0)), ((4, 4, 4), 3, (0, 0, 1, 1, 0, 0))])
def test_outflow_ghost_layers_share_c2p_inputs(Nt, ng, bc):
    n, bad = copies(Nt, ng, bc)
    assert n > 0 and bad == 0


def test_planted_wrong_face_pair_fails():
    # a cell-centred B averaged over faces p and p+2 (a wrong stencil) is not
    # layer-invariant, and the check sees it
    n, bad = copies((6, 6, 6), 3, (0,) * 6, shift_hi=2)
    assert bad > 0

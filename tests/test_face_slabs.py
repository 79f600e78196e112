"""Ghost faces between RK stages (grmhd.hip launch_gather, faces_slab): with
shared UCT ring faces the gather between stages copies only the face slabs
(face index in [ng, n+ng], exactly one other coordinate in the ghost layers)
and the other faces of the ghost cells beyond outflow faces.  The readers of
ghost faces in the next stage are restated here from the kernels -- the transverse WENO stencils of the UCT edge EMFs
(k_emf_uct), the six faces of every ghost cell that runs its own c2p (beyond
an outflow face, ghost_needs_own_c2p), the normal face of every ring face
computed beyond an outflow face (k_flux with ring_own) -- and every ghost
face they read must be one the reduced gather fills."""
import pytest

GRIDS = [
    (8, 3, (3, 3, 3), (0,) * 6),
    (8, 3, (2, 2, 2), (1,) * 6),
    (8, 4, (1, 2, 3), (1, 1, 0, 0, 1, 1)),
    (4, 3, (2, 2, 2), (1, 1, 0, 0, 1, 1)),
    (5, 3, (3, 2, 1), (0, 0, 1, 1, 0, 0)),
]


def is_ghost_face(n, ng, fd, loc):
    return any(loc[d] < ng or loc[d] >= n + ng + (d == fd) for d in range(3))


def is_slab(n, ng, fd, loc):
    if loc[fd] < ng or loc[fd] > n + ng:
        return False
    return sum(loc[d] < ng or loc[d] >= n + ng for d in range(3) if d != fd) == 1


def bounds_outflow_ghost(n, ng, nb, bc, b, fd, loc):
    """a face of a ghost cell beyond an outflow face (the extra list)"""
    N = n + 2 * ng
    lo = list(loc)
    lo[fd] -= 1
    return ((loc[fd] < N and cell_beyond_outflow(n, ng, nb, bc, b, loc))
            or (loc[fd] >= 1 and cell_beyond_outflow(n, ng, nb, bc, b, lo)))


def cell_beyond_outflow(n, ng, nb, bc, b, loc):
    for d in range(3):
        g = b[d] * n + loc[d] - ng
        if (g < 0 and bc[2 * d] == 0) or (g >= nb[d] * n and bc[2 * d + 1] == 0):
            return True
    return False


def reads(n, ng, nb, bc, b):
    """(face direction, local face index) of every face the next stage reads"""
    N = n + 2 * ng
    out = set()
    rng = range(ng, n + ng + 1)
    # k_emf_uct: E^x reads By along z and Bz along y, E^y Bz along x and Bx
    # along z, E^z Bx along y and By along x (sites [ng, n+ng]^3, the edge
    # direction's index below n+ng)
    for i in rng:
        for j in rng:
            for k in rng:
                for t in range(6):
                    o = t - 3
                    if i < n + ng:
                        out.add((1, (i, j, k + o)))
                        out.add((2, (i, j + o, k)))
                    if j < n + ng:
                        out.add((2, (i + o, j, k)))
                        out.add((0, (i, j, k + o)))
                    if k < n + ng:
                        out.add((0, (i, j + o, k)))
                        out.add((1, (i + o, j, k)))
    # ghost cells with their own c2p: the six faces of the cell
    for k in range(N):
        for j in range(N):
            for i in range(N):
                loc = (i, j, k)
                if all(ng <= x < n + ng for x in loc) or not cell_beyond_outflow(n, ng, nb, bc, b, loc):
                    continue
                for d in range(3):
                    for h in (0, 1):
                        f = list(loc)
                        f[d] += h
                        out.add((d, tuple(f)))
    # ring faces computed beyond outflow faces: their normal face
    for D in range(3):
        T = [d for d in range(3) if d != D]
        for a in rng:
            for t0 in range(N):
                for t1 in range(N):
                    g0, g1 = not (ng <= t0 < n + ng), not (ng <= t1 < n + ng)
                    if g0 == g1:
                        continue
                    loc = [0, 0, 0]
                    loc[D], loc[T[0]], loc[T[1]] = a, t0, t1
                    td, tv = (T[0], t0) if g0 else (T[1], t1)
                    side = 2 * td + (tv >= ng)
                    edge = (b[td] == 0) if side % 2 == 0 else (b[td] == nb[td] - 1)
                    if edge and bc[side] == 0:
                        out.add((D, tuple(loc)))
    return out


@pytest.mark.parametrize("n,ng,nb,bc", GRIDS)
def test_stage_gather_covers_every_ghost_face_read(n, ng, nb, bc):
    for bk in range(nb[2]):
        for bj in range(nb[1]):
            for bi in range(nb[0]):
                b = (bi, bj, bk)
                for fd, loc in reads(n, ng, nb, bc, b):
                    if not is_ghost_face(n, ng, fd, loc):
                        continue
                    assert is_slab(n, ng, fd, loc) or bounds_outflow_ghost(n, ng, nb, bc, b, fd, loc), (b, fd, loc)


def test_slabs_are_about_half_of_the_ghost_faces():
    n, ng = 16, 3
    N = n + 2 * ng
    tot = slab = 0
    for fd in range(3):
        E = [N + (d == fd) for d in range(3)]
        for k in range(E[2]):
            for j in range(E[1]):
                for i in range(E[0]):
                    loc = (i, j, k)
                    if is_ghost_face(n, ng, fd, loc):
                        tot += 1
                        slab += is_slab(n, ng, fd, loc)
    assert tot == 20340 and slab == 3 * 17 * 16 * 2 * ng * 2

"""UCT ring values between RK stages (grmhd.hip k_flux push_nbr): the pushes
of the interior faces within ngz of a block face write exactly the ring faces
a copy from the neighbouring block would fill -- faces with one transverse
index in the ghost layers whose side has a neighbour -- each from the same
source face and once, whenever the neighbour relation is symmetric.  The index maps are
restated here from the kernels (block-local face indices, the six-neighbour
table with periodic wrap) and compared on outflow, periodic, self-periodic
and mixed grids."""
import pytest


def fidx(N, D, i, j, k):
    return i + (N[0] + (D == 0)) * (j + (N[1] + (D == 1)) * k)


def neighbours(insts):
    blk, nq = {}, 0
    for a, (nb, q0, bc) in enumerate(insts):
        for bk in range(nb[2]):
            for bj in range(nb[1]):
                for bi in range(nb[0]):
                    q = q0 + bi + nb[0] * (bj + nb[1] * bk)
                    blk[q] = (a, (bi, bj, bk))
                    nq = max(nq, q + 1)
    nbr = {}
    for q in range(nq):
        a, b = blk[q]
        nb, q0, bc = insts[a]
        for sd in range(6):
            d, dr = sd // 2, (1 if sd & 1 else -1)
            c = list(b)
            c[d] += dr
            if c[d] < 0 or c[d] >= nb[d]:
                if bc[sd] != 1:
                    nbr[(q, sd)] = -1
                    continue
                c[d] = (c[d] + nb[d]) % nb[d]
            nbr[(q, sd)] = q0 + c[0] + nb[0] * (c[1] + nb[1] * c[2])
    return nq, nbr


def copies(n, ng, nq, nbr, D):
    """ring face -> the neighbour's interior face it equals"""
    N = [n + 2 * ng] * 3
    T0, T1 = (1 if D == 0 else 0), (1 if D == 2 else 2)
    want = {}
    for q in range(nq):
        for a in range(ng, n + ng + 1):
            for t0 in range(n + 2 * ng):
                for t1 in range(n + 2 * ng):
                    g0, g1 = not (ng <= t0 < n + ng), not (ng <= t1 < n + ng)
                    if g0 == g1:
                        continue
                    td, tv = (T0, t0) if g0 else (T1, t1)
                    qs = nbr[(q, 2 * td + (tv >= ng))]
                    if qs < 0:
                        continue
                    loc = [0, 0, 0]
                    loc[D], loc[T0], loc[T1] = a, t0, t1
                    src = list(loc)
                    src[td] = tv + n if tv < ng else tv - n
                    want[(q, fidx(N, D, *loc))] = (qs, fidx(N, D, *src))
    return want


def pushes(n, ng, nq, nbr, D):
    """ring face -> pushing face, as the interior face threads of k_flux write them"""
    N = [n + 2 * ng] * 3
    T0, T1 = (1 if D == 0 else 0), (1 if D == 2 else 2)
    got = {}
    for p in range(nq):
        for a in range(ng, n + ng + 1):
            for t0 in range(ng, n + ng):
                for t1 in range(ng, n + ng):
                    loc = [0, 0, 0]
                    loc[D], loc[T0], loc[T1] = a, t0, t1
                    for td in (T0, T1):
                        it = loc[td] - ng
                        for hi in (0, 1):
                            if (it < n - ng) if hi else (it >= ng):
                                continue
                            qt = nbr[(p, 2 * td + hi)]
                            if qt < 0:
                                continue
                            l2 = list(loc)
                            l2[td] = it - n + ng if hi else it + n + ng
                            key = (qt, fidx(N, D, *l2))
                            assert key not in got, ("written twice", key)
                            got[key] = (p, fidx(N, D, *loc))
    return got


GRIDS = [
    (8, 3, [((3, 3, 3), 0, (0,) * 6)]),
    (8, 3, [((2, 2, 2), 0, (1,) * 6)]),
    (8, 4, [((1, 2, 3), 0, (1, 1, 0, 0, 1, 1)), ((3, 1, 1), 6, (0,) * 6)]),
    (4, 3, [((2, 2, 2), 0, (1, 1, 0, 0, 1, 1))]),
    (5, 3, [((3, 2, 1), 0, (1,) * 6)]),
]


@pytest.mark.parametrize("n,ng,insts", GRIDS)
def test_ring_pushes_fill_what_the_copies_fill(n, ng, insts):
    nq, nbr = neighbours(insts)
    assert all(nbr[(q, s)] < 0 or nbr[(nbr[(q, s)], s ^ 1)] == q for q in range(nq) for s in range(6))
    for D in range(3):
        assert pushes(n, ng, nq, nbr, D) == copies(n, ng, nq, nbr, D)


def test_per_dimension_periodicity_gives_symmetric_neighbours():
    # the contexts accept periodicity per dimension only (a periodic face
    # needs a periodic opposite face, as the reference's brick connectivity):
    # then q's neighbour across a face has q across the opposite face, which
    # the pushes rely on; a one-sided periodic face would break it
    for px in (0, 1):
        for py in (0, 1):
            for pz in (0, 1):
                for nb in ((1, 1, 1), (2, 1, 3), (3, 2, 2)):
                    nq, nbr = neighbours([(nb, 0, (px, px, py, py, pz, pz))])
                    assert all(nbr[(q, s)] < 0 or nbr[(nbr[(q, s)], s ^ 1)] == q
                               for q in range(nq) for s in range(6))
    nq, nbr = neighbours([((1, 2, 3), 0, (1, 0, 0, 0, 1, 1))])
    assert not all(nbr[(q, s)] < 0 or nbr[(nbr[(q, s)], s ^ 1)] == q for q in range(nq) for s in range(6))


def stencil_reads(n, ng, nq, nbr, D, q, a):
    """(block, normal index) of the six stencil cells of a full face at normal
    index a of block q, as k_flux reads them in the second and third RK
    stages: beyond a block face with a neighbour, that neighbour's interior"""
    out = []
    for t in range(6):
        ap = a + t - 3
        if ap < ng and nbr[(q, 2 * D)] >= 0:
            out.append((nbr[(q, 2 * D)], ap + n))
        elif ap >= n + ng and nbr[(q, 2 * D + 1)] >= 0:
            out.append((nbr[(q, 2 * D + 1)], ap - n))
        else:
            out.append((q, ap))
    return out


@pytest.mark.parametrize("n,ng,insts", GRIDS)
def test_full_face_stencils_read_what_the_gather_copies(n, ng, insts):
    # the source of a ghost cell by global position in its instance (the
    # ghost table's mapping, periodic wrap): every cell a full face reads
    # beyond a block face must be that source; a cell beyond an outflow face
    # is read from the block's own ghost layer, which runs its own c2p (or
    # copies a first-layer ghost that does)
    nq, nbr = neighbours(insts)
    blk = {}
    for a_, (nb, q0, bc) in enumerate(insts):
        for bk in range(nb[2]):
            for bj in range(nb[1]):
                for bi in range(nb[0]):
                    blk[q0 + bi + nb[0] * (bj + nb[1] * bk)] = (nb, q0, bc, (bi, bj, bk))
    for D in range(3):
        for q in range(nq):
            nb, q0, bc, b = blk[q]
            Nt = nb[D] * n
            for a in range(ng, n + ng + 1):
                for t, (qb, ib) in enumerate(stencil_reads(n, ng, nq, nbr, D, q, a)):
                    ap = a + t - 3
                    g = b[D] * n + ap - ng
                    outside = g < 0 or g >= Nt
                    if outside and bc[2 * D + (g >= Nt)] != 1:
                        assert (qb, ib) == (q, ap)
                        continue
                    g %= Nt
                    c = list(b)
                    c[D] = g // n
                    owner = q0 + c[0] + nb[0] * (c[1] + nb[1] * c[2])
                    assert (qb, ib) == (owner, g % n + ng), (D, q, a, t)
                    assert ng <= ib < n + ng

"""Tile bookkeeping of the LDS-staged CTU face kernels (k_ctu_x / k_ctu_y /
k_ctu_z in camr_amd/csrc/hydro_fused.hip), restated on the host.

For tiles of 32 x 8 points (the x-face stage merging a single leftover
point into the last full tile) this mirrors, thread by thread, which
transdo pair each thread solves in phase 1 and where it stores it in LDS,
which transdd differences phase 2 forms, and which face phase 3 finishes.
It checks that every LDS slot a thread reads holds the record of the face
or cell it expects, written in the same phase; that every face of the stage
range is finished exactly once; that every pair face lies in the transdo
stage range nodes(grow(bx, 1), D); and that every LDS position is inside
its array.  A change to the kernels' predicates must be made here too.
"""
import pytest

CX, CY = 32, 8
NB = CX * CY

def ctu_ntx(n, merge):
    return n // CX if (merge and n > CX and n % CX == 1) else (n + CX - 1) // CX

def tiles(sb, merge):
    (l0, l1, l2), (h0, h1, h2) = sb
    n0, n1 = h0 - l0 + 1, h1 - l1 + 1
    ntx, nty = ctu_ntx(n0, merge), (n1 + CY - 1) // CY
    for by in range(nty):
        for bx in range(ntx):
            wx = n0 - bx * CX if bx == ntx - 1 else CX
            yield l0 + bx * CX, l1 + by * CY, wx

def in_t1(bx, D, i, j, k):   # nodes(grow(bx,1), D)
    (l, h) = bx
    p = (i, j, k)
    for d in range(3):
        lo, hi = l[d] - 1, h[d] + 1 + (1 if d == D else 0)
        if not (lo <= p[d] <= hi):
            return False
    return True

def check(bx):
    (l, h) = bx
    k = l[2]
    nout = {}
    for D in range(3):
        hi = list(h); hi[D] += 1
        sb = (tuple(l), tuple(hi))
        hx, hy = hi[0], hi[1]
        for (i0, j0, wx) in tiles(sb, D == 0):
            if D == 2:
                NA, NBB = CY * (CX + 1), (CY + 1) * CX
                A, Bw = {}, {}
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if i <= hx + 1 and j <= hy:
                        p = ty * (CX + 1) + tx; assert p not in A; A[p] = (i, j); assert in_t1(bx, 0, i, j, k)
                    if i <= hx and j <= hy + 1:
                        p = ty * CX + tx; assert p not in Bw; Bw[p] = (i, j); assert in_t1(bx, 1, i, j, k)
                    eB, eA = t, t - 64
                    if eB < CX and i0 + eB <= hx and j0 + CY <= hy + 1:
                        p = CY * CX + eB; assert p not in Bw; Bw[p] = (i0 + eB, j0 + CY); assert in_t1(bx, 1, i0 + eB, j0 + CY, k)
                    if 0 <= eA < CY and i0 + CX <= hx + 1 and j0 + eA <= hy:
                        p = eA * (CX + 1) + CX; assert p not in A; A[p] = (i0 + CX, j0 + eA); assert in_t1(bx, 0, i0 + CX, j0 + eA, k)
                assert all(0 <= p < NA for p in A) and all(0 <= p < NBB for p in Bw)
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if i <= hx and j <= hy:
                        assert A[ty * (CX + 1) + tx] == (i, j), (D, i, j)
                        assert A[ty * (CX + 1) + tx + 1] == (i + 1, j)
                        assert Bw[ty * CX + tx] == (i, j)
                        assert Bw[(ty + 1) * CX + tx] == (i, j + 1)
                        key = (D, i, j); assert key not in nout; nout[key] = 1
            elif D == 0:
                NC = CX + 2
                NY, NT = (CY + 1) * NC, CY * NC
                w = min(wx, CX)
                Y, Tw = {}, {}
                halo = {}
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    face = tx < w and j <= hy
                    if tx < w and j <= hy + 1:
                        p = ty * NC + tx + 1; assert p not in Y; Y[p] = (i, j); assert in_t1(bx, 1, i, j, k)
                    if face:
                        assert in_t1(bx, 2, i, j, k) and in_t1(bx, 2, i, j, k + 1)
                    ext = None
                    if t < CX:
                        if t < w and j0 + CY <= hy + 1: ext = (i0 + t, j0 + CY, CY * NC + t + 1)
                    elif t < CX + CY + 1:
                        r = t - CX
                        if j0 + r <= hy + 1: ext = (i0 - 1, j0 + r, r * NC)
                    elif t < CX + 2 * (CY + 1):
                        r = t - CX - (CY + 1)
                        if wx == CX + 1 and j0 + r <= hy + 1: ext = (i0 + CX, j0 + r, r * NC + CX + 1)
                    if ext:
                        assert t < 64
                        ei, ej, p = ext; assert p not in Y; Y[p] = (ei, ej); assert in_t1(bx, 1, ei, ej, k)
                    hh = t - 64
                    if 0 <= hh < 2 * CY and (T := (hh < CY and j0 + hh <= hy) or (hh >= CY and wx == CX + 1 and j0 + hh - CY <= hy)):
                        ci = i0 - 1 if hh < CY else i0 + CX
                        cj = j0 + (hh if hh < CY else hh - CY)
                        col = 0 if hh < CY else CX + 1
                        halo[t] = (ci, cj, col)
                        assert in_t1(bx, 2, ci, cj, k) and in_t1(bx, 2, ci, cj, k + 1)
                assert all(0 <= p < NY for p in Y)
                # td phase
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if tx < w and j <= hy:
                        assert Y[ty * NC + tx + 1] == (i, j) and Y[(ty + 1) * NC + tx + 1] == (i, j + 1)
                        p = ty * NC + tx + 1; assert p not in Tw and p < NT; Tw[p] = (i, j)
                for t, (ci, cj, col) in halo.items():
                    r = cj - j0
                    assert Y[r * NC + col] == (ci, cj) and Y[(r + 1) * NC + col] == (ci, cj + 1)
                    p = r * NC + col; assert p not in Tw and p < NT; Tw[p] = (ci, cj)
                # faces
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if tx < w and j <= hy:
                        assert Tw[ty * NC + tx] == (i - 1, j), (i0, j0, i, j)
                        key = (D, i, j); assert key not in nout; nout[key] = 1
                for t, (ci, cj, col) in halo.items():
                    if t - 64 >= CY:
                        r = cj - j0
                        assert Tw[r * NC + CX] == (ci - 1, cj) and Tw[r * NC + CX + 1] == (ci, cj)
                        key = (D, ci, cj); assert key not in nout; nout[key] = 1
            else:
                NC = CX + 1
                NX, NT = (CY + 1) * NC, (CY + 1) * CX
                X, Tw, halo = {}, {}, {}
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if i <= hx + 1 and j <= hy:
                        p = (ty + 1) * NC + tx; assert p not in X; X[p] = (i, j); assert in_t1(bx, 0, i, j, k)
                    if i <= hx and j <= hy:
                        assert in_t1(bx, 2, i, j, k) and in_t1(bx, 2, i, j, k + 1)
                    ext = None
                    if t < CY:
                        if i0 + CX <= hx + 1 and j0 + t <= hy: ext = (i0 + CX, j0 + t, (t + 1) * NC + CX)
                    elif t < CY + CX + 1:
                        c = t - CY
                        if i0 + c <= hx + 1: ext = (i0 + c, j0 - 1, c)
                    if ext:
                        ei, ej, p = ext; assert p not in X; X[p] = (ei, ej); assert in_t1(bx, 0, ei, ej, k)
                    hh = t - 64
                    if 0 <= hh < CX and i0 + hh <= hx:
                        halo[t] = (i0 + hh, j0 - 1)
                        assert in_t1(bx, 2, i0 + hh, j0 - 1, k) and in_t1(bx, 2, i0 + hh, j0 - 1, k + 1)
                assert all(0 <= p < NX for p in X)
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if i <= hx and j <= hy:
                        assert X[(ty + 1) * NC + tx] == (i, j) and X[(ty + 1) * NC + tx + 1] == (i + 1, j)
                        p = (ty + 1) * CX + tx; assert p not in Tw and p < NT; Tw[p] = (i, j)
                for t, (ci, cj) in halo.items():
                    hh = t - 64
                    assert X[hh] == (ci, cj) and X[hh + 1] == (ci + 1, cj)
                    p = hh; assert p not in Tw; Tw[p] = (ci, cj)
                for t in range(NB):
                    tx, ty = t % CX, t // CX
                    i, j = i0 + tx, j0 + ty
                    if i <= hx and j <= hy:
                        assert Tw[ty * CX + tx] == (i, j - 1)
                        key = (D, i, j); assert key not in nout; nout[key] = 1
        # every face of the stage range exactly once
        want = {(D, i, j) for i in range(l[0], hi[0] + 1) for j in range(l[1], hi[1] + 1)}
        got = {kk for kk in nout if kk[0] == D}
        assert got == want, (D, len(got), len(want))
    return True

SHAPES = [(256, 256, 4), (32, 32, 3), (33, 40, 2), (8, 8, 2), (17, 9, 2), (64, 31, 2),
          (1, 1, 1), (31, 33, 1), (96, 48, 1), (65, 65, 1)]


@pytest.mark.parametrize("n", SHAPES)
@pytest.mark.parametrize("lo", [(0, 0, 0), (-5, 7, 3)])
def test_ctu_tile_coverage(n, lo):
    bx = (lo, tuple(lo[d] + n[d] - 1 for d in range(3)))
    assert check(bx)

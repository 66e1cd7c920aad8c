"""Embedded-boundary geometry of an implicit body (host set-up, numpy).

Stands in for AMReX's ``EB2::SphereIF`` + ``EB2::Build`` as called by
``make_eb_sphere`` (Source/EB/eb_sphere.cpp:14-44): the cut-cell data that
``hydro_umdrv_eb`` (Source/Hydro/Hydro_umdrv_eb.cpp:14-191) consumes --
cell flags with neighbour connectivity, volume fractions, face area
fractions (apertures) and face centroids.  AMReX itself is not available;
this is a restatement of the published construction:

* the level set is sampled at nodes; phi < 0 is fluid (SphereIF with
  ``internal_flow = 0``: fluid outside the sphere);
* an edge whose end nodes differ in sign is cut where the exact sphere
  crosses it (the root of the quadratic along the edge);
* a face with cut edges is cut by the straight line through its cut points:
  its aperture and centroid are those of the fluid polygon;
* a cell with mixed node signs is cut by a plane with normal
  (a_lo - a_hi) per direction (the divergence theorem applied to a
  constant field); its volume fraction follows from the divergence theorem
  applied to x: V = (1/2 sum of the six apertures + A_b s) / 3, with s the
  mean signed offset of the cell's cut points along the plane normal;
* a cell is connected to a neighbour when a path of open faces (aperture
  > 0) leads there, one axis step at a time.

The numbers therefore differ from AMReX's in the last digits of cut cells;
the hydro path takes them as given inputs, so GPU and oracle parity does
not depend on them.

Flag word (int32) per cell: bits 0..26 connectivity, bit
(ii+1) + 3 (jj+1) + 9 (kk+1) for the neighbour at offset (ii,jj,kk);
bits 27..28 the type (0 regular, 1 single-valued cut cell, 2 covered).
"""
from __future__ import annotations

import itertools

import numpy as np

REGULAR, SINGLE, COVERED = 0, 1, 2
TYPE_SHIFT = 27
ALL_CONNECTED = (1 << 27) - 1
GEOM_GROW = 7      # ghost width of the geometry arrays around each box
BOX_REGULAR, BOX_EB, BOX_COVERED = 0, 1, 2


def conn_bit(ii, jj, kk):
    return (ii + 1) + 3 * (jj + 1) + 9 * (kk + 1)


class Sphere:
    """Level set r^2 - |x - c|^2 (fluid outside) or its negative."""

    def __init__(self, center, radius, fluid_inside=False):
        self.c = np.asarray(center, dtype=np.float64)
        self.r = float(radius)
        self.sign = -1.0 if fluid_inside else 1.0

    def phi(self, x, y, z):
        d2 = (x - self.c[0]) ** 2 + (y - self.c[1]) ** 2 + (z - self.c[2]) ** 2
        return self.sign * (self.r * self.r - d2)

    def cut(self, a, b):
        """Fraction t in [0,1] along a->b where the sphere crosses."""
        a = np.asarray(a, dtype=np.float64)
        d = np.asarray(b, dtype=np.float64) - a
        e = a - self.c
        A = float(d @ d)
        B = 2.0 * float(d @ e)
        C = float(e @ e) - self.r * self.r
        disc = max(B * B - 4 * A * C, 0.0)
        sq = np.sqrt(disc)
        for t in ((-B - sq) / (2 * A), (-B + sq) / (2 * A)):
            if -1e-12 <= t <= 1 + 1e-12:
                return min(max(t, 0.0), 1.0)
        return 0.5


def _polygon(pts):
    """Area and centroid of a simple polygon (shoelace)."""
    if len(pts) < 3:
        return 0.0, 0.0, 0.0
    P = np.asarray(pts)
    x, y = P[:, 0], P[:, 1]
    xn, yn = np.roll(x, -1), np.roll(y, -1)
    cr = x * yn - xn * y
    a = 0.5 * cr.sum()
    if abs(a) < 1e-300:
        return 0.0, 0.0, 0.0
    cx = ((x + xn) * cr).sum() / (6 * a)
    cy = ((y + yn) * cr).sum() / (6 * a)
    return abs(a), cx, cy


class Geometry:
    """Cut-cell geometry of `body` on the index space of `prob`.

    ``flags`` covers the domain grown by GEOM_GROW cells (x fastest,
    array order [k][j][i]); ``box_geom(b)`` returns the volume fractions,
    apertures and face centroids on grow(box b, GEOM_GROW)."""

    def __init__(self, prob, body):
        self.prob = prob
        self.body = body
        g = GEOM_GROW
        self.lo = np.array(prob.domlo) - g             # first cell
        self.hi = np.array(prob.domhi) + g             # last cell
        self.dx = np.array(prob.dx, dtype=np.float64)
        self.plo = np.array(prob.prob_lo, dtype=np.float64)
        n = self.hi - self.lo + 1
        # node coordinates (node i sits at prob_lo + i dx)
        xs = [self.plo[d] + (np.arange(self.lo[d], self.hi[d] + 2)) * self.dx[d] for d in range(3)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        self.phi = body.phi(X, Y, Z)                   # [k][j][i] on nodes
        self.fluid = self.phi < 0.0
        f = self.fluid
        # face classification from corner nodes: x-face (i,j,k) has nodes
        # (i, j..j+1, k..k+1)
        self.ap = []
        self.fc = []
        for d in range(3):
            ax = [0, 1, 2]
            ax.remove(2 - d)       # array axes of the two tangential node pairs
            sl = []
            corners = []
            for a in (0, 1):
                for b in (0, 1):
                    s = [slice(None)] * 3
                    s[ax[0]] = slice(a, f.shape[ax[0]] - 1 + a)
                    s[ax[1]] = slice(b, f.shape[ax[1]] - 1 + b)
                    corners.append(f[tuple(s)])
            allf = corners[0] & corners[1] & corners[2] & corners[3]
            anyf = corners[0] | corners[1] | corners[2] | corners[3]
            ap = np.where(allf, 1.0, 0.0)
            fc = np.zeros(ap.shape + (2,))
            cut = np.argwhere(anyf & ~allf)
            for kji in cut:
                a_, c0, c1 = self._cut_face(d, tuple(int(v) for v in kji))
                ap[tuple(kji)] = a_
                fc[tuple(kji)] = (c0, c1)
            self.ap.append(ap)     # shape: nodes in d, cells in the others
            self.fc.append(fc)
        self._cells()

    # ---- faces -------------------------------------------------------------
    def _node_pos(self, i, j, k):
        return self.plo + (np.array([i, j, k]) + self.lo) * self.dx

    def _cut_face(self, d, kji):
        """Fluid polygon of face `kji` (array index) normal to d."""
        k, j, i = kji
        t = [e for e in range(3) if e != d]     # tangential directions (u, v)
        base = [i, j, k]
        # corners in perimeter order in (u, v) local coordinates
        loop = [(0, 0), (1, 0), (1, 1), (0, 1)]
        pts = []
        def node(c):
            n = list(base)
            n[t[0]] += c[0]
            n[t[1]] += c[1]
            return n
        for m in range(4):
            ca, cb = loop[m], loop[(m + 1) % 4]
            na, nb = node(ca), node(cb)
            fa = self.fluid[na[2], na[1], na[0]]
            fb = self.fluid[nb[2], nb[1], nb[0]]
            if fa:
                pts.append((ca[0] - 0.5, ca[1] - 0.5))
            if fa != fb:
                s = self.body.cut(self._node_pos(*na), self._node_pos(*nb))
                pts.append((ca[0] + s * (cb[0] - ca[0]) - 0.5,
                            ca[1] + s * (cb[1] - ca[1]) - 0.5))
        a, cx, cy = _polygon(pts)
        if a <= 0.0:
            return 0.0, 0.0, 0.0
        return min(a, 1.0), cx, cy

    # ---- cells -------------------------------------------------------------
    def _cells(self):
        f = self.fluid
        n8 = [f[a:f.shape[0] - 1 + a, b:f.shape[1] - 1 + b, c:f.shape[2] - 1 + c]
              for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        allf = np.logical_and.reduce(n8)
        anyf = np.logical_or.reduce(n8)
        typ = np.where(allf, REGULAR, np.where(anyf, SINGLE, COVERED)).astype(np.int32)
        vf = np.where(allf, 1.0, 0.0)
        apx, apy, apz = self.ap
        for kji in np.argwhere(typ == SINGLE):
            k, j, i = (int(v) for v in kji)
            a = [apx[k, j, i], apx[k, j, i + 1], apy[k, j, i], apy[k, j + 1, i],
                 apz[k, j, i], apz[k + 1, j, i]]
            nb = np.array([a[0] - a[1], a[2] - a[3], a[4] - a[5]])
            barea = float(np.sqrt(nb @ nb))
            s = 0.0
            if barea > 0.0:
                nrm = nb / barea
                offs = []
                for e in range(3):
                    for u in (0, 1):
                        for v in (0, 1):
                            tt = [x for x in range(3) if x != e]
                            na = [i, j, k]
                            na[tt[0]] += u
                            na[tt[1]] += v
                            nbn = list(na)
                            nbn[e] += 1
                            fa = self.fluid[na[2], na[1], na[0]]
                            fb = self.fluid[nbn[2], nbn[1], nbn[0]]
                            if fa != fb:
                                w = self.body.cut(self._node_pos(*na), self._node_pos(*nbn))
                                p = np.array(na, dtype=np.float64) - [i, j, k] - 0.5
                                p[e] += w
                                offs.append(float(nrm @ p))
                s = float(np.mean(offs)) if offs else 0.0
            v = (0.5 * sum(a) + barea * s) / 3.0
            vf[k, j, i] = min(max(v, 0.0), 1.0)
            if vf[k, j, i] == 0.0:
                typ[k, j, i] = COVERED
        self.vfrac = vf
        # a cut cell all of whose faces are closed holds no fluid
        self.type = typ
        self.flags = self._connect(typ)

    def _open(self, d, k, j, i):
        """Aperture > 0 of the face of direction d at node index (array)."""
        ap = self.ap[d]
        if not (0 <= k < ap.shape[0] and 0 <= j < ap.shape[1] and 0 <= i < ap.shape[2]):
            return True   # beyond the sampled region: regular far field
        return ap[k, j, i] > 0.0

    def _step_open(self, c, d, s):
        """Face crossed stepping from cell c (array [k,j,i]) by s in d."""
        k, j, i = c
        if d == 0:
            return self._open(0, k, j, i + (1 if s > 0 else 0))
        if d == 1:
            return self._open(1, k, j + (1 if s > 0 else 0), i)
        return self._open(2, k + (1 if s > 0 else 0), j, i)

    def _connect(self, typ):
        flags = np.full(typ.shape, ALL_CONNECTED, dtype=np.int64)
        flags = flags | (typ.astype(np.int64) << TYPE_SHIFT)
        flags[typ == COVERED] = COVERED << TYPE_SHIFT
        # cells whose 3^3 neighbourhood holds a non-regular cell need paths
        nonreg = typ != REGULAR
        near = np.zeros_like(nonreg)
        K, J, I = nonreg.shape
        for a, b, c in itertools.product((-1, 0, 1), repeat=3):
            src = nonreg[max(a, 0):K + min(a, 0), max(b, 0):J + min(b, 0), max(c, 0):I + min(c, 0)]
            near[max(-a, 0):K + min(-a, 0), max(-b, 0):J + min(-b, 0), max(-c, 0):I + min(-c, 0)] |= src
        for kji in np.argwhere(near & (typ != COVERED)):
            c0 = tuple(int(v) for v in kji)
            word = 1 << conn_bit(0, 0, 0)
            for off in itertools.product((-1, 0, 1), repeat=3):
                kk, jj, ii = off
                if off == (0, 0, 0):
                    continue
                steps = [(0, ii), (1, jj), (2, kk)]
                steps = [s for s in steps if s[1] != 0]
                ok = False
                for perm in itertools.permutations(steps):
                    c = list(c0)
                    good = True
                    for d, s in perm:
                        if not self._step_open(tuple(c), d, s):
                            good = False
                            break
                        c[2 - d] += s
                    if good:
                        ok = True
                        break
                if ok:
                    word |= 1 << conn_bit(ii, jj, kk)
            flags[c0] = word | (int(typ[c0]) << TYPE_SHIFT)
        return flags.astype(np.int32)

    # ---- per-box views -----------------------------------------------------
    def _cells_of(self, lo, hi):
        s = tuple(slice(lo[d] - self.lo[d], hi[d] - self.lo[d] + 1) for d in (2, 1, 0))
        return s

    def box_geom(self, b):
        """vfrac, apx, apy, apz, fcx, fcy, fcz on grow(box b, GEOM_GROW)."""
        bx = self.prob.boxes[b]
        lo = np.array(bx[:3]) - GEOM_GROW
        hi = np.array(bx[3:]) + GEOM_GROW
        c = self._cells_of(lo, hi)
        out = {"vfrac": np.ascontiguousarray(self.vfrac[c])}
        for d, name in enumerate("xyz"):
            s = list(c)
            s[2 - d] = slice(c[2 - d].start, c[2 - d].stop + 1)
            out["ap" + name] = np.ascontiguousarray(self.ap[d][tuple(s)])
            out["fc" + name] = np.ascontiguousarray(np.moveaxis(self.fc[d][tuple(s)], -1, 0))
        return out

    def box_kind(self, b):
        """BOX_COVERED if every valid cell is covered, BOX_EB if a
        non-regular cell lies within 2 cells (the hydro_umdrv_eb test of
        CAMR_construct_hydro_source.cpp:80,150-151), else BOX_REGULAR."""
        bx = self.prob.boxes[b]
        t = self.type[self._cells_of(bx[:3], bx[3:])]
        if np.all(t == COVERED):
            return BOX_COVERED
        g = self.type[self._cells_of(np.array(bx[:3]) - 2, np.array(bx[3:]) + 2)]
        return BOX_EB if np.any(g != REGULAR) else BOX_REGULAR

    def covered_mask(self, lo, hi):
        return self.type[self._cells_of(lo, hi)] == COVERED

"""Box layouts of an AMR level: regular chopping and tag clustering.

The reference gets its grids from AMReX (``BoxArray::maxSize`` for level 0,
``TagBoxArray`` + Berger-Rigoutsos clustering under ``blocking_factor`` and
``max_grid_size`` for refined levels).  This module builds the same kind of
layouts for the batched level advance:

* ``chop(domain, max_size)``: the domain cut into boxes of at most
  ``max_size`` cells per side (level 0).
* ``cluster(tags, blocking_factor, max_size, efficiency)``: boxes covering
  every tagged cell, aligned to ``blocking_factor``, no side above
  ``max_size``.  Tags are first coarsened to blocking-factor blocks, then
  the tagged blocks are split recursively (Berger & Rigoutsos 1991: split at
  a hole of the signature, else at the strongest inflection of its second
  difference, else in the middle) until each box is at least ``efficiency``
  full and within ``max_size``.
"""
from __future__ import annotations

import numpy as np


def chop(lo, hi, max_size):
    """Boxes (lo..., hi...) tiling [lo, hi] with sides <= max_size."""
    out = []
    rng = [list(range(lo[d], hi[d] + 1, max_size)) for d in range(3)]
    for k in rng[2]:
        for j in rng[1]:
            for i in rng[0]:
                out.append((i, j, k, min(i + max_size - 1, hi[0]),
                            min(j + max_size - 1, hi[1]), min(k + max_size - 1, hi[2])))
    return out


def _split_point(sig):
    """Split index (in 1..len-1) of a 1-D signature, or None."""
    n = len(sig)
    if n < 2:
        return None
    holes = np.nonzero(sig == 0)[0]
    if holes.size:
        # the hole closest to the centre
        c = (n - 1) / 2.0
        h = int(holes[np.argmin(np.abs(holes - c))])
        return h if h > 0 else h + 1
    if n >= 4:
        lap = np.zeros(n)
        lap[1:-1] = sig[2:] - 2 * sig[1:-1] + sig[:-2]
        best, where = 0, None
        for i in range(1, n - 2):
            if lap[i] * lap[i + 1] < 0:
                jump = abs(lap[i + 1] - lap[i])
                if jump > best:
                    best, where = jump, i + 1
        if where is not None:
            return where
    return n // 2


def _cluster_blocks(t, off, max_blocks, efficiency, out):
    idx = np.argwhere(t)
    if idx.size == 0:
        return
    lo = idx.min(axis=0)
    hi = idx.max(axis=0) + 1
    sub = t[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    size = np.array(sub.shape)
    eff = sub.sum() / sub.size
    if eff >= efficiency and size.max() <= max_blocks:
        out.append((off + lo, off + hi))
        return
    # split along the longest side that can be split
    for d in np.argsort(-size, kind="stable"):
        if size[d] < 2:
            continue
        sig = sub.sum(axis=tuple(a for a in range(3) if a != d))
        s = _split_point(sig)
        if size[d] > max_blocks and (s is None or eff >= efficiency):
            s = size[d] // 2
        if s is None:
            continue
        sl = [slice(None)] * 3
        sl[d] = slice(0, s)
        a = sub[tuple(sl)]
        sl[d] = slice(s, None)
        b = sub[tuple(sl)]
        ob = off + lo
        ob2 = ob.copy()
        ob2[d] += s
        _cluster_blocks(a, ob, max_blocks, efficiency, out)
        _cluster_blocks(b, ob2, max_blocks, efficiency, out)
        return
    out.append((off + lo, off + hi))


def cluster(tags, blocking_factor=8, max_size=64, efficiency=0.7, domlo=(0, 0, 0)):
    """Boxes covering the tagged cells of `tags` (bool array indexed
    [i, j, k] from domlo), aligned to blocking_factor, sides <= max_size."""
    bf = blocking_factor
    n = np.array(tags.shape)
    assert np.all(n % bf == 0), "domain must be a multiple of the blocking factor"
    nb = n // bf
    blocks = tags.reshape(nb[0], bf, nb[1], bf, nb[2], bf).any(axis=(1, 3, 5))
    raw = []
    _cluster_blocks(blocks, np.zeros(3, dtype=int), max_size // bf, efficiency, raw)
    out = []
    for blo, bhi in raw:
        lo = [int(domlo[d] + blo[d] * bf) for d in range(3)]
        hi = [int(domlo[d] + bhi[d] * bf - 1) for d in range(3)]
        out.append((lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))
    return out


def sphere_tags(n, radius=0.3, centre=(0.5, 0.5, 0.5)):
    """Cells of an n^3 grid on [0,1]^3 whose centre lies within `radius`."""
    x = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return ((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) <= radius ** 2


def covers(boxes, tags):
    """Every tagged cell lies in exactly one box (and boxes are disjoint)."""
    cnt = np.zeros(tags.shape, dtype=np.int32)
    for b in boxes:
        cnt[b[0]:b[3] + 1, b[1]:b[4] + 1, b[2]:b[5] + 1] += 1
    return bool(np.all(cnt <= 1) and np.all(cnt[tags] == 1))

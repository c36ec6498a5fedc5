"""Oracle restatement of the tag clustering (TEST INFRASTRUCTURE ONLY: used by
tests/ to check ampe_cluster_mask / ampe_cluster_tags).

Berger & Rigoutsos, IEEE Trans. Syst. Man Cybern. 21 (1991) 1278-1286, in the
role SAMRAI's BergerRigoutsos plays for the reference
(source/PFModel.cc:258,565-590): shrink a box to its tags' bounding box,
accept it at efficiency >= efficiency_tol or when it cannot be cut into two
pieces of min_size, else cut at the empty slab nearest the middle, else at the
strongest sign change of the signature's second difference, else in the
middle of the longest cuttable direction; then chop accepted boxes to
max_size.  Tags are first grown by tag_buffer cells (box neighbourhood,
clipped).  SAMRAI's own variant is absent from the reference tree, so the
boxes are "parity unpinned" against it; the tests pin this restatement with
closed-form cases and the invariants every clustering must meet (coverage,
disjointness, efficiency, minimum size)."""
import numpy as np


def _buffer(m, b):
    if b <= 0:
        return m.copy()
    g = m.copy()
    for ax in range(m.ndim):
        h = np.zeros_like(g)
        n = g.shape[ax]
        for s in range(-b, b + 1):
            src = [slice(None)] * g.ndim
            dst = [slice(None)] * g.ndim
            if s >= 0:
                src[ax] = slice(0, n - s)
                dst[ax] = slice(s, n)
            else:
                src[ax] = slice(-s, n)
                dst[ax] = slice(0, n + s)
            h[tuple(dst)] |= g[tuple(src)]
        g = h
    return g


def cluster(mask, efficiency_tol=0.8, min_size=(1, 1, 1), max_size=(0, 0, 0), tag_buffer=0):
    """Boxes [(lo, hi)] (inclusive, x first) for a mask ([nz,] ny, nx)."""
    m = _buffer(np.asarray(mask) != 0, tag_buffer)
    ndim = m.ndim
    mx = np.transpose(m)  # x first
    out = []

    def sig(lo, hi):
        sub = mx[tuple(slice(lo[d], hi[d] + 1) for d in range(ndim))].astype(np.int64)
        s = []
        for d in range(ndim):
            axes = tuple(a for a in range(ndim) if a != d)
            s.append(sub.sum(axis=axes))
        return int(sub.sum()), s

    def run(lo, hi):
        lo, hi = list(lo), list(hi)
        n, s = sig(lo, hi)
        if n == 0:
            return
        for d in range(ndim):
            nz = np.nonzero(s[d])[0]
            lo[d], hi[d] = lo[d] + int(nz[0]), lo[d] + int(nz[-1])
        n, s = sig(lo, hi)
        cells = int(np.prod([hi[d] - lo[d] + 1 for d in range(ndim)]))
        if n >= efficiency_tol * cells:
            out.append((tuple(lo), tuple(hi)))
            return
        ext = [hi[d] - lo[d] + 1 for d in range(ndim)]
        cut = None
        best = None
        for d in range(ndim):
            e = ext[d]
            if e < 2 * min_size[d]:
                continue
            for c in range(min_size[d], e - min_size[d] + 1):
                if s[d][c] != 0 and s[d][c - 1] != 0:
                    continue
                dist = abs(c - 0.5 * e) / e
                if best is None or dist < best:
                    best, cut = dist, (d, c)
        if cut is None:
            bests, bestd = 0, None
            for d in range(ndim):
                e = ext[d]
                if e < 2 * min_size[d] or e < 4:
                    continue
                L = np.zeros(e, np.int64)
                L[1:e - 1] = s[d][:e - 2] - 2 * s[d][1:e - 1] + s[d][2:]
                for i in range(1, e - 2):
                    if not ((L[i] < 0 < L[i + 1]) or (L[i] > 0 > L[i + 1])):
                        continue
                    c = i + 1
                    if c < min_size[d] or c > e - min_size[d]:
                        continue
                    st = abs(int(L[i + 1] - L[i]))
                    dist = abs(c - 0.5 * e) / e
                    if st > bests or (st == bests and bestd is not None and dist < bestd):
                        bests, bestd, cut = st, dist, (d, c)
        if cut is None:
            cand = [d for d in range(ndim) if ext[d] >= 2 * min_size[d]]
            if not cand:
                out.append((tuple(lo), tuple(hi)))
                return
            d = max(cand, key=lambda a: (ext[a], -a))
            cut = (d, ext[d] // 2)
        d, c = cut
        h1 = list(hi)
        h1[d] = lo[d] + c - 1
        l2 = list(lo)
        l2[d] = lo[d] + c
        run(lo, h1)
        run(l2, hi)

    run([0] * ndim, [mx.shape[d] - 1 for d in range(ndim)])
    res = []
    for lo, hi in out:
        npc = [1] * ndim
        for d in range(ndim):
            if max_size[d] > 0:
                npc[d] = -(-(hi[d] - lo[d] + 1) // max_size[d])
        for idx in np.ndindex(*npc[::-1]):
            idx = idx[::-1]
            plo, phi = [], []
            for d in range(ndim):
                e = hi[d] - lo[d] + 1
                plo.append(lo[d] + e * idx[d] // npc[d])
                phi.append(lo[d] + e * (idx[d] + 1) // npc[d] - 1)
            res.append((tuple(plo), tuple(phi)))
    return res

"""Initial conditions as the reference's generator writes them
(utils/make_nuclei.py): spherical nuclei of phase 1 in a phase-0 matrix,
interior / exterior compositions with the exterior one set so that the
nominal composition is met, and single-precision fields (the generator's
default, 'f' NetCDF variables) that the reference reads back as doubles.

Restated behaviour, with the generator's quirks kept:
  * centres: one nucleus at --center0 (the z coordinate defaults to 1) or at
    the box centre; two nuclei at (nx/3, ny/3, nz/2) and (2nx/3, 2ny/3, nz/2)
    with radii r and 1.5 r (make_nuclei.py:520-560);
  * distances are plain Euclidean: the periodic-image loops compare the
    option strings with the integer 1 (make_nuclei.py:238-252) and so never
    run;
  * a cell centre (i + 0.5, j + 0.5, k + 0.5) gets phase_in where its
    distance is below the radius (width 0), or the tanh profile
    phase_in (1 + tanh(-d / 2w)) / 2 where d < 8 w (make_nuclei.py:626-660);
    vs accumulates the stored (single-precision) values, vl = nx ny nz - vs;
  * with a nominal and an interior composition the exterior one is
    (c vol - c_in vs) / vl, stored in single precision
    (make_nuclei.py:715-721); c = c_in phi + c_out (1 - phi).
  * orientations (qlen 4): after random.seed(21361) each nucleus draws
    t = pi U(-1, 1) and Shoemake's (u1, u2, u3), each u floored to a multiple
    of 0.1 / 0.05 / 0.05 (make_nuclei.py:175-230, utils/quat.py:522-545,
    Python's Mersenne Twister); every cell takes the quaternion of its current
    nucleus, which changes only when the cell is farther than 2 r from it
    (squared distance above (4 r)^2 / 4), then to the first nearer nucleus,
    the search stopping once within 2 r (make_nuclei.py:673-700).
More than two nuclei are not restated."""
import math
import random

import numpy as np


def _random_quats(n):
    """Orientations of make_nuclei.py's setRandomQinSpheres for qlen 4."""
    rng = random.Random(21361)
    out = []
    for _ in range(n):
        rng.uniform(-1, 1.)  # the discrete angle t, used only for qlen 1 / 2
        u1, u2, u3 = rng.uniform(0, 1.), rng.uniform(0, 1.), rng.uniform(0, 1.)
        h1 = 0.1
        h2 = 0.5 * h1
        h3 = h2
        u1 = math.floor(u1 / h1) * h1
        u2 = math.floor(u2 / h2) * h2
        u3 = math.floor(u3 / h3) * h3
        out.append([math.sqrt(1. - u1) * math.sin(2. * math.pi * u2),
                    math.sqrt(1. - u1) * math.cos(2. * math.pi * u2),
                    math.sqrt(u1) * math.sin(2. * math.pi * u3),
                    math.sqrt(u1) * math.cos(2. * math.pi * u3)])
    return out


def make_nuclei(nx, ny, nz=1, radius=None, ngrains=1, phase_in=1.0, phase_out=0.0,
                nomconc=None, conc_in=None, conc_out=None, width=0.0, center0=None,
                temperature_in=None, temperature_out=None, double_precision=False,
                qlen=0):
    if ngrains not in (1, 2):
        raise NotImplementedError("make_nuclei: one or two nuclei")
    dt = np.float64 if double_precision else np.float32
    if radius is None:
        radius = nx / 10
    if ngrains == 1:
        r = [radius]
        if center0 is not None:
            cx, cy = [center0[0]], [center0[1]]
            cz = [center0[2] if len(center0) > 2 else 1]
        else:
            cx, cy = [nx / 2], [ny / 2]
            cz = [nz / 2 if nz > 1 else 1]
    else:
        cx, cy, cz = [nx / 3, 2 * nx / 3], [ny / 3, 2 * ny / 3], [nz / 2, nz / 2]
        r = [radius, radius * 1.5]
    phase = np.zeros((nz, ny, nx), dt)
    if phase_out:
        phase[...] = phase_out
    z = np.arange(nz) + 0.5
    y = np.arange(ny) + 0.5
    x = np.arange(nx) + 0.5
    vs = 0.0
    for g in range(ngrains):
        thr = (r[g] + 5.0 * width) ** 2
        for k in range(nz):
            if (z[k] - cz[g]) ** 2 >= thr:
                continue
            for j in range(ny):
                if (y[j] - cy[g]) ** 2 >= thr:
                    continue
                d2 = (x - cx[g]) ** 2 + (y[j] - cy[g]) ** 2 + (z[k] - cz[g]) ** 2
                d = np.sqrt(d2) - np.sqrt(r[g] ** 2)
                if width > 0.0:
                    sel = d < 8.0 * width
                    phase[k, j, sel] = phase_in * 0.5 * (1.0 + np.tanh(-1.0 * d[sel] / (2.0 * width)))
                else:
                    sel = d < 0.0
                    phase[k, j, sel] = phase_in
                # vs sums the stored values cell by cell, i ascending
                for v in phase[k, j, sel]:
                    vs = vs + float(v)
    vol = nx * ny * nz
    vl = vol - vs
    out = dict(phase=phase.astype(np.float64), vs=vs, vl=vl)
    if nomconc is not None or conc_in is not None:
        c = list(nomconc) if nomconc is not None else None
        ci = list(conc_in) if conc_in is not None else list(c)
        ns = len(set(c)) if c is not None else len(ci)
        co = np.zeros(ns, dt) if conc_out is None else np.asarray(conc_out, np.float64)
        if conc_in is not None and c is not None and vl > 0:
            for s in range(ns):
                co[s] = (c[s] * vol - ci[s] * vs) / vl
        conc = []
        ph = phase.astype(np.float64)
        for s in range(ns):
            conc.append((ci[s] * ph + float(co[s]) * (1.0 - ph)).astype(dt).astype(np.float64))
        out["conc"] = conc
        out["conc_out"] = [float(v) for v in co]
    if qlen:
        if qlen != 4:
            raise NotImplementedError("make_nuclei: qlen 4 orientations")
        qs = _random_quats(ngrains)
        mind2 = 0.25 * (4.0 * radius) ** 2
        quat = np.zeros((4, nz, ny, nx), dt)
        gmin = 0
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    d2min = ((x[i] - cx[gmin]) ** 2 + (y[j] - cy[gmin]) ** 2 +
                             (z[k] - cz[gmin]) ** 2)
                    if d2min > mind2:
                        for g in range(ngrains):
                            d2 = ((x[i] - cx[g]) ** 2 + (y[j] - cy[g]) ** 2 +
                                  (z[k] - cz[g]) ** 2)
                            if d2 < d2min:
                                d2min = d2
                                gmin = g
                                if d2min < mind2:
                                    break
                    quat[:, k, j, i] = qs[gmin]
        out["quat"] = quat.astype(np.float64)
        out["grain_quats"] = qs
    if temperature_in is not None:
        T = np.full((nz, ny, nx), temperature_out, dt)
        for g in range(ngrains):
            d2 = ((x[None, None, :] - cx[g]) ** 2 + (y[None, :, None] - cy[g]) ** 2 +
                  (z[:, None, None] - cz[g]) ** 2)
            T[d2 - r[g] ** 2 < 0.0] = temperature_in
        out["temperature"] = T.astype(np.float64)
    return out


def make_multi_spheres(nx, ny, spheres, conc_a, conc_b, conc_out, nz=1, width=0.0):
    """Restatement of utils/make_multi_spheres.py: one order parameter per
    sphere (x, y, z, radius, phase 'A' or 'B' of the csv rows, cell units),
    1 inside (distance to the centre below the radius), else 0, or
    0.5 (1 + tanh(-d / 2w)) within 8w of the surface when width > 0; the
    composition takes the sphere phase's values inside and conc_out elsewhere.
    As in the generator, the periodicity option never takes effect (it
    compares text with a number), so distances are plain Euclidean ones, and
    for nz = 1 the centre sits at z = 1 while the cells sit at z = 0.5.
    Single precision fields, as the generator writes them.

    Returns dict(phase=[nspheres, nz, ny, nx], conc=[nspecies, nz, ny, nx])."""
    nspecies = len(set(conc_a))
    conc = np.zeros((nspecies, nz, ny, nx), np.float32)
    phase = np.zeros((len(spheres), nz, ny, nx), np.float32)
    for s in range(nspecies):
        conc[s] = conc_out[s]
    z = (np.arange(nz) + 0.5)[:, None, None]
    y = (np.arange(ny) + 0.5)[None, :, None]
    x = (np.arange(nx) + 0.5)[None, None, :]
    for g, (cx, cy, cz, r, ph) in enumerate(spheres):
        if nz == 1:
            cz = 1
        cs = conc_a if ph == "A" else conc_b
        d2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2
        d = np.sqrt(d2) - np.sqrt(float(r) ** 2)
        inside = d < 0.0
        phase[g][inside] = 1.0
        for s in range(nspecies):
            conc[s][inside] = cs[s]
        if width > 0.0:
            band = np.abs(d) < 8.0 * width
            phase[g][band] = (0.5 * (1.0 + np.tanh(-1.0 * d / (2.0 * width))))[band]
    return dict(phase=phase, conc=conc)

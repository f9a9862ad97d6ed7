"""Synthetic workloads of the lattice update, shared by the benchmark and the tests.

Each builder returns plain descriptions (lbm_desc, obstacle masks, link lists) so that
the same lattice can be built in several storage forms and compared:

* ``cavity(n)``            lid-driven cavity n x n (src/cavity.cpp:26-34, 61-84)
* ``ensemble_descs()``     the seeded ensemble of 256 uneven channel lattices
* ``large_domain()``       the 32768 x 16384 LES domain with 128 seeded obstacles
* ``q_circle``             exact-circle q of one link (src/lbm_types.hpp:118-138)

Everything here runs on the host with numpy; the lattices step on the GPU through
lbm2d.Ensemble.
"""
from __future__ import annotations

import numpy as np

import lbm2d as G

# Reference defaults used by the workloads (src/lbm_types.hpp:32, src/main.cpp:32-43).
U_IN = 0.1
CS_DEFAULT = 0.12
AUTO_LES_TAU = 0.55  # the cylinder driver turns LES on below this tau (src/main.cpp:174-182)


def q_circle(ccx, ccy, r, xf, yf, i) -> float:
    """q along link i from fluid node (xf, yf) to the circle of centre (ccx, ccy) and
    radius r: the first root t in (0, 1] of |x + t e_i - c| = r, 1.0 when the ray misses
    (BounceBackGeometry::q_cylinder, src/lbm_types.hpp:118-138)."""
    dx, dy = xf - ccx, yf - ccy
    ex, ey = float(G.CXV[i]), float(G.CYV[i])
    a = ex * ex + ey * ey
    b = 2.0 * (dx * ex + dy * ey)
    c = dx * dx + dy * dy - r * r
    disc = b * b - 4.0 * a * c
    if disc < 0.0:
        return 1.0
    sd = np.sqrt(disc)
    t1, t2 = (-b - sd) / (2 * a), (-b + sd) / (2 * a)
    if 0.0 < t1 <= 1.0:
        return float(t1)
    if 0.0 < t2 <= 1.0:
        return float(t2)
    return 1.0


# ---- C1: lid-driven cavity ---------------------------------------------------------

def cavity(n=256, tau=1.268, u_lid=0.1):
    """The cavity driver's lattice (src/cavity.cpp:26-34, 61-84): walls at x = 0,
    x = n-1 and y = 0 (place_walls, src/lbm.hpp:79-87), lid u_lid on row n-1, feq(1, 0, 0)
    everywhere.  Returns (desc, mask)."""
    d = G.desc_for_case(G.CAVITY, n, n, tau=tau, u_in=u_lid)
    return d, G.walls_mask(n, n)


# ---- C4: the batched ensemble ------------------------------------------------------

def ensemble_descs(seed=42, n=256):
    """The ensemble's lattices: NX ~ U[256,1024], NY ~ U[96,384], 1-4 circles r ~ U[8,40],
    centres uniform with a 2r margin, bounding circles 2 cells apart; Re log-uniform in
    [50,400]; tau = 0.5 + 3 u D / Re with D the largest diameter and u = 0.1.  A lattice
    whose tau falls below 0.55 runs Smagorinsky LES with c_s = 0.12, the cylinder driver's
    auto-LES rule (src/main.cpp:174-182).  numpy Generator(PCG64(seed)), drawn in the
    order nx, ny, k, then per circle r, cx, cy, then Re.  Returns a list of
    (desc, mask, circles) with circles as (cx, cy, r)."""
    rng = np.random.default_rng(seed)
    lats = []
    for _ in range(n):
        nx = int(rng.integers(256, 1025))
        ny = int(rng.integers(96, 385))
        k = int(rng.integers(1, 5))
        circles = []
        tries = 0
        while len(circles) < k and tries < 200:
            tries += 1
            r = int(rng.integers(8, 41))
            if 4 * r >= ny - 2 or 4 * r >= nx - 2:
                continue
            cx = int(rng.integers(2 * r, nx - 2 * r))
            cy = int(rng.integers(2 * r, ny - 2 * r))
            if all((cx - a) ** 2 + (cy - b) ** 2 > (r + c + 2) ** 2 for a, b, c in circles):
                circles.append((cx, cy, r))
        re = float(np.exp(rng.uniform(np.log(50), np.log(400))))
        dmax = 2 * max(c[2] for c in circles)
        tau = 0.5 + 3 * U_IN * dmax / re
        les = tau < AUTO_LES_TAU
        d = G.desc_for_case(G.CYLINDER, nx, ny, tau=tau, u_in=U_IN, les=les, cs=CS_DEFAULT)
        mask = np.zeros((ny, nx), np.uint8)
        for cx, cy, r in circles:
            mask |= G.circle_mask(nx, ny, cx, cy, r)
        lats.append((d, mask, circles))
    return lats


def ensemble_links(mask, circles, nx, ny):
    """Per-obstacle exact-circle q for every boundary link of a lattice (the circle
    nearest to the solid node owns the link), Mei scheme, force slot 0."""
    links = []
    ys, xs = np.nonzero(mask == 0)
    solid = mask.astype(bool)
    for i in range(1, 9):
        tx = xs + G.CXV[i]
        ty = (ys + G.CYV[i]) % ny
        ok = (tx >= 0) & (tx < nx)
        sel = np.nonzero(ok)[0]
        hit = sel[solid[ty[sel], tx[sel]]]
        for k in hit:
            x, y = int(xs[k]), int(ys[k])
            sx, sy = x + G.CXV[i], y + G.CYV[i]
            c = min(circles, key=lambda c: (sx - c[0]) ** 2 + (sy - c[1]) ** 2)
            q = q_circle(c[0], c[1], c[2], x, y, i)
            links.append((y * nx + x, i, G.BB_MEI, 0, q, 0.0, 0.0))
    return links


def build_ensemble(lats, storage="f32", stream=None, scale_exp=8, links=None):
    """An Ensemble of the lattices of ensemble_descs(), each with its per-obstacle links
    (links[l] when given, else computed) and the uniform start feq(1, 0.1, 0) built on
    the device."""
    ens = G.Ensemble([d for d, _, _ in lats], storage=storage, stream=stream, scale_exp=scale_exp)
    # every geometry first, so that the first initialisation builds them in one commit
    for l, (d, mask, circles) in enumerate(lats):
        ens.set_obstacle(l, mask)
        ens.set_links(l, links[l] if links is not None else ensemble_links(mask, circles, d.nx, d.ny))
    for l in range(len(lats)):
        ens.init_equilibrium(l, 1.0, U_IN, 0.0)
    return ens


# ---- C5: the large LES domain ------------------------------------------------------

LARGE = dict(nx=32768, ny=16384, tau=0.505, u=U_IN, cs=CS_DEFAULT, n_circles=64, n_plates=64, seed=42)


def large_domain(seed=42, nx=32768, ny=16384, n_circles=64, n_plates=64, with_shapes=False, scheme=G.BB_MEI):
    """Obstacles of the large domain: n_circles circles r ~ U[20,60] (exact-circle Mei
    links, force slots 1..n_circles) and n_plates flat plates 155 x 2 at an angle of
    attack ~ U[-15, 15] deg (polygon staircase, simple bounce-back, force slot 0).
    Centres uniform in x in [nx/16, 15nx/16) and in y clear of the periodic seam,
    bounding circles at least 8 cells apart; numpy Generator(PCG64(seed)), circles
    first.  Returns the mask (ny, nx) and the explicit link list (and, with_shapes, the
    obstacles: ("circle", cx, cy, r) or ("plate", cx, cy, aoa) in placement order)."""
    rng = np.random.default_rng(seed)
    mask = np.zeros((ny, nx), np.uint8)
    placed, links, shapes = [], [], []
    half_diag = float(np.hypot(77.5, 1.0))
    for k in range(n_circles + n_plates):
        circle = k < n_circles
        while True:
            if circle:
                r = int(rng.integers(20, 61))
                br = float(r)
            else:
                aoa = float(rng.uniform(-15.0, 15.0))
                br = half_diag
            cx = int(rng.integers(nx // 16, nx - nx // 16))
            cy = int(rng.integers(int(br) + 4, ny - int(br) - 4))
            if all((cx - a) ** 2 + (cy - b) ** 2 > (br + c + 8) ** 2 for a, b, c in placed):
                break
        placed.append((cx, cy, br))
        shapes.append(("circle", cx, cy, r) if circle else ("plate", cx, cy, aoa))
        x0, y0 = cx - int(br) - 2, cy - int(br) - 2
        w = h = 2 * int(br) + 5
        if circle:
            yy, xx = np.mgrid[y0:y0 + h, x0:x0 + w]
            dx, dy = (xx - cx).astype(float), (yy - cy).astype(float)
            mask[y0:y0 + h, x0:x0 + w] |= (np.sqrt(dx * dx + dy * dy) < r).astype(np.uint8)
            sub = mask[y0:y0 + h, x0:x0 + w]
            ys, xs = np.nonzero(sub[1:-1, 1:-1] == 0)
            ys, xs = ys + 1, xs + 1
            for i in range(1, 9):
                sx, sy = (xs + G.CXV[i] + x0 - cx).astype(float), (ys + G.CYV[i] + y0 - cy).astype(float)
                hit = (sub[ys + G.CYV[i], xs + G.CXV[i]] != 0) & (np.sqrt(sx * sx + sy * sy) < r)
                for yl, xl in zip(ys[hit], xs[hit]):
                    x, y = x0 + int(xl), y0 + int(yl)
                    links.append((y * nx + x, i, scheme, 1 + k, q_circle(cx, cy, r, x, y, i), 0.0, 0.0))
        else:
            xy = G.plate_polygon(float(cx), float(cy), 155.0, aoa, 2.0)
            mask[y0:y0 + h, x0:x0 + w] |= G.polygon_mask(w, h, xy - np.array([x0, y0], float))
    return (mask, links, shapes) if with_shapes else (mask, links)


def large_lattice(mask, links, storage, stream=None, sr=False, sr_seed=1, scale_exp=8, nx=None, ny=None,
                  tau=LARGE["tau"], les=True):
    """One handle of the large domain (periodic y, Zou/He inlet u = 0.1, outflow copy,
    LES c_s = 0.12, tau = 0.505) from large_domain()'s mask and links, started at
    feq(1, 0.1, 0) on the device."""
    ny_, nx_ = mask.shape
    nx = nx or nx_
    ny = ny or ny_
    d = G.desc_for_case(G.CYLINDER, nx, ny, tau=tau, u_in=U_IN, les=les, cs=CS_DEFAULT)
    e = G.Ensemble([d], storage=storage, stream=stream, scale_exp=scale_exp)
    if sr:
        e.set_stochastic_rounding(True, sr_seed)
    e.set_obstacle(0, mask)
    e.set_force_slots(0, 1 + max(k[3] for k in links))
    e.set_links(0, links)
    e.init_equilibrium(0, 1.0, U_IN, 0.0)
    return e

"""Bounds of the step kernels' speculative pull addresses against the population buffers.

The kernels form a node's nine pull addresses before they know whether the node lies
on an edge (lbm_kernels.hip: cell_update, node_values of the two-step kernel, the
thermal kernel's entry loads), the general path gathers sources one row off the
lattice on open edges (preload), and the grouped bulk path moves aligned vectors of V
nodes with the wave's outer lanes fetching one node further (group_update).  Every
population buffer -- the two pools, the thermal populations' two, the autotuner's copies
-- has the layout lbm_population_layout reports: planes of every lattice, then a guard.
Here every address those access patterns can form, over every launch shape the library
compiles (a superset: block 64..512, rows 1..8, 1..8 nodes per thread within 16 bytes
per plane), is checked to lie inside [0, pool + guard), for exact-size pools (1024 x 512
FP32 fills nine 2 MiB pages) and the 32768 x 16384 domain among others.  For a fixed
direction and column an address grows with the row, so large lattices are enumerated
on their first two and last two rows; small ones on every row.  No device is needed.
"""
import numpy as np
import pytest

import lbm2d as G

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
SHAPES = [(bx, rows, v) for bx in (64, 128, 256, 512) for rows in (1, 2, 4, 8) for v in (1, 2, 4, 8)]


def _rows(ny):
    return np.arange(ny) if ny <= 64 else np.array([0, 1, ny - 2, ny - 1])


def _extremes(d, f_off, bpp):
    """(min, max) element index over every address the kernels form for lattice d."""
    nx, ny = d.nx, d.ny
    pitch = (nx + 63) // 64 * 64
    plane = ny * pitch
    lo, hi = np.inf, -np.inf

    def upd(a):
        nonlocal lo, hi
        lo = min(lo, int(a.min()))
        hi = max(hi, int(a.max()))

    ys = _rows(ny)
    xs = np.arange(nx)
    X, Y = np.meshgrid(xs, ys)
    idx = Y * pitch + X
    for j in range(9):
        # cell_update / node_values / thermal entry: unwrapped pulls of every node; the
        # outlet column pulls the copied node nx - 2's sources
        shift = np.where((X == nx - 1) & bool(d.outlet), 1, 0)
        upd(f_off + j * plane + idx - shift - (CY[j] * pitch + CX[j]))
        upd(f_off + j * plane + idx)
        # preload: sources wrapped on periodic edges only
        sx, sy = X - CX[j], Y - CY[j]
        if d.x_edge == G.EDGE_PERIODIC:
            sx = np.where(sx < 0, sx + nx, np.where(sx >= nx, sx - nx, sx))
        if d.y_edge == G.EDGE_PERIODIC:
            sy = np.where(sy < 0, sy + ny, np.where(sy >= ny, sy - ny, sy))
        upd(f_off + j * plane + sy * pitch + sx)
    # group_update: interior rows, groups of V nodes at x = multiples of V below pitch
    inner = ys[(ys >= 1) & (ys <= ny - 2)]
    for bx, rows, v in SHAPES:
        if v == 1 or bpp * v > 16 or len(inner) == 0:
            continue
        gx = np.arange(0, pitch, v)
        GX, GY = np.meshgrid(gx, inner)
        gidx = GY * pitch + GX
        for j in range(9):
            base = f_off + j * plane + gidx - CY[j] * pitch
            upd(base)
            upd(base + v - 1)
            if CX[j] == 1:
                upd(base - 1)   # lane 0's node before the wave
            if CX[j] == -1:
                upd(base + v)   # lane 63's node after the wave
    return lo, hi


def _descs(sizes, edges=(G.EDGE_OPEN, G.EDGE_PERIODIC), outlet=1):
    out = []
    for nx, ny in sizes:
        d = G.desc_for_case(G.CYLINDER, nx, ny, tau=0.6, u_in=0.1)
        d.x_edge, d.y_edge = edges
        d.outlet = outlet
        out.append(d)
    return out


CASES = {
    "exact_1024x512": [(1024, 512)],
    "large_32768x16384": [(32768, 16384)],
    "tiny": [(3, 3)],
    "thin_rows_past_nx": [(65, 3)],
    "thin_200x3": [(200, 3)],
    "ensemble_uneven": [(300, 100), (1024, 96), (257, 384), (700, 200), (64, 64)],
}


@pytest.mark.parametrize("storage,bpp", [("f64", 8), ("f32", 4), ("f16s", 2)])
@pytest.mark.parametrize("edges", [(G.EDGE_OPEN, G.EDGE_PERIODIC), (G.EDGE_OPEN, G.EDGE_OPEN),
                                   (G.EDGE_WALL, G.EDGE_WALL), (G.EDGE_PERIODIC, G.EDGE_PERIODIC)])
@pytest.mark.parametrize("name", sorted(CASES))
def test_pull_addresses_stay_inside_the_buffer(name, edges, storage, bpp):
    descs = _descs(CASES[name], edges)
    f_off, pool, guard = G.population_layout(descs)
    assert guard >= max((d.nx + 63) // 64 * 64 for d in descs) + 256
    for d, off in zip(descs, f_off):
        lo, hi = _extremes(d, off, bpp)
        assert lo >= 0, (name, d.nx, d.ny, lo)
        assert hi < pool + guard, (name, d.nx, d.ny, hi, pool, guard)


@pytest.mark.parametrize("name", ["exact_1024x512", "large_32768x16384"])
def test_without_the_guard_the_last_lattice_reads_past_its_pool(name):
    """The check above has teeth: an exact-size buffer (pool only, the autotuner's old
    copies) is overrun by the last lattice's pulls."""
    descs = _descs(CASES[name])
    f_off, pool, _ = G.population_layout(descs)
    _, hi = _extremes(descs[-1], f_off[-1], 4)
    assert hi >= pool


def test_layout_matches_the_stated_rule():
    """f_off: blocks of 9 ny pitch values rounded up to 256, in lattice order."""
    descs = _descs([(300, 100), (1024, 96), (3, 3)])
    f_off, pool, guard = G.population_layout(descs)
    sizes = [-(-9 * d.ny * ((d.nx + 63) // 64 * 64) // 256) * 256 for d in descs]
    assert f_off == [0, sizes[0], sizes[0] + sizes[1]]
    assert pool == sum(sizes)
    assert guard == 1024 + 256
    assert (9 * 1024 * 512 * 4) % (2 << 20) == 0  # the exact-size case fills whole pages
    assert G.population_layout(_descs([(1024, 512)]))[1] == 9 * 1024 * 512


# ---- the two-step kernel's ring ------------------------------------------------------
#
# step2_kernel computes its tile's 66 x (rows + 2) positions (tile and one-node ring) at
# lattice nodes: positions off a periodic edge wrap once, positions off another edge are
# skipped, and positions further than one node off the lattice are skipped (no tile node
# has a neighbour there).  Those nodes' own loads are the in-lattice nodes checked above,
# so the ring is safe when every position it computes is a lattice node.

def _ring_nodes(d, rows, ensemble, skip_far=True):
    nx, ny = d.nx, d.ny
    x_start = -63 if (ensemble and d.outlet and nx % 64 == 1) else 0
    LX, LY = np.meshgrid(np.arange(66), np.arange(rows + 2))
    gxs, gys = [], []
    for y0 in range(0, ny, rows):
        for x0 in range(x_start, nx, 64):
            ux, uy = x0 - 1 + LX, y0 - 1 + LY
            keep = np.ones(ux.shape, bool)
            if skip_far:
                keep &= (ux >= -1) & (ux <= nx) & (uy >= -1) & (uy <= ny)
            gx, gy = ux.copy(), uy.copy()
            if d.x_edge == G.EDGE_PERIODIC:
                gx = np.where(gx < 0, gx + nx, np.where(gx >= nx, gx - nx, gx))
            else:
                keep &= (gx >= 0) & (gx < nx)
            if d.y_edge == G.EDGE_PERIODIC:
                gy = np.where(gy < 0, gy + ny, np.where(gy >= ny, gy - ny, gy))
            else:
                keep &= (gy >= 0) & (gy < ny)
            gxs.append(gx[keep])
            gys.append(gy[keep])
    return np.concatenate(gxs), np.concatenate(gys)


RING_SIZES = [(3, 3), (65, 3), (200, 3), (3, 200), (64, 17), (129, 48), (130, 9), (1, 1)]


@pytest.mark.parametrize("ensemble", [False, True])
@pytest.mark.parametrize("rows", [4, 8, 16])
@pytest.mark.parametrize("edges", [(G.EDGE_OPEN, G.EDGE_PERIODIC), (G.EDGE_PERIODIC, G.EDGE_PERIODIC),
                                   (G.EDGE_WALL, G.EDGE_WALL)])
def test_two_step_ring_positions_are_lattice_nodes(edges, rows, ensemble):
    for nx, ny in RING_SIZES:
        if nx < 3 or ny < 3:
            continue  # below the library's minimum lattice
        (d,) = _descs([(nx, ny)], edges)
        gx, gy = _ring_nodes(d, rows, ensemble)
        assert gx.size > 0
        assert gx.min() >= 0 and gx.max() < nx, (nx, ny, rows, edges)
        assert gy.min() >= 0 and gy.max() < ny, (nx, ny, rows, edges)


def test_two_step_ring_needs_the_far_positions_skipped():
    """Wrapping once is not enough on a periodic lattice lower or narrower than the tile
    and its ring: without the skip, positions fall outside the lattice."""
    (d,) = _descs([(200, 3)], (G.EDGE_OPEN, G.EDGE_PERIODIC))
    _, gy = _ring_nodes(d, 8, False, skip_far=False)
    assert gy.max() >= 3
    (d,) = _descs([(3, 3)], (G.EDGE_PERIODIC, G.EDGE_PERIODIC))
    gx, _ = _ring_nodes(d, 4, False, skip_far=False)
    assert gx.max() >= 3


# ---- bulk tiles of two-step launches (lbm_kernels.hip step2_bulk_kernel) -------------
#
# A bulk tile's ring (66 x (rows + 2) nodes from (x0 - 1, y0 - 1)) lies inside the lattice
# and holds no general-path node; every edge node is a general-path node, so each ring
# node lies off the edges and its nine pulls are lattice nodes.  The kernel forms those
# addresses as the ring origin plus a 32-bit offset (ly * pitch + lx, ly * pitch by a
# 24-bit multiply) and finds a ring position's row by a multiply-shift.

def _bulk_tiles(special, rows):
    ny, nx = special.shape
    out = []
    for y0 in range(0, ny, rows):
        for x0 in range(0, nx, 64):
            if (x0 >= 1 and x0 + 64 <= nx - 1 and y0 >= 1 and y0 + rows <= ny - 1
                    and not special[y0 - 1:y0 + rows + 1, x0 - 1:x0 + 65].any()):
                out.append((x0, y0))
    return out


@pytest.mark.parametrize("rows", [8, 16, 32])
def test_bulk_tile_pulls_are_lattice_nodes(rows):
    rng = np.random.default_rng(rows)
    for nx, ny in ((300, 140), (1000, 333), (129, 97)):
        solid = rng.random((ny, nx)) < 0.0005
        special = solid.copy()
        for cx, cy in zip(CX[1:], CY[1:]):
            special |= np.roll(np.roll(solid, -cy, 0), -cx, 1)
        special[0, :] = special[-1, :] = special[:, 0] = special[:, -1] = True
        tiles = _bulk_tiles(special, rows)
        assert tiles or nx < 200
        for x0, y0 in tiles:
            ly, lx = np.meshgrid(np.arange(rows + 2), np.arange(66), indexing="ij")
            x, y = x0 - 1 + lx, y0 - 1 + ly
            assert not special[y, x].any()
            for j in range(9):
                sx, sy = x - CX[j], y - CY[j]
                assert sx.min() >= 0 and sx.max() < nx and sy.min() >= 0 and sy.max() < ny


def test_bulk_ring_row_and_offsets():
    """floor(r * 993 / 2^16) = r // 66 over every ring size compiled (up to 66 x 34); the
    host admits a lattice to the bulk kernel only while (rows + 3) rows of population bytes
    fit in 31 bits and pitch < 2^24 (lbm_host.hip bulk_eligible)."""
    r = np.arange(66 * 34)
    assert np.array_equal((r * 993) >> 16, r // 66)
    r = np.arange(32768)
    assert np.array_equal((r * 993) >> 16, r // 66)
    pitch, bpp, rows = (32768 + 63) // 64 * 64, 4, 32
    assert (rows + 3) * pitch * bpp < 2 ** 31 and pitch < 2 ** 24
    # the largest offset a ring load forms, in bytes
    assert ((rows + 1) * pitch + 65 + pitch + 1) * bpp < 2 ** 32


# ---- edge tiles of two-step launches (step2_bulk_kernel's edge form) ---------------------
#
# An edge tile starts on a node of its lattice and may be cut off at its last column and
# row; its ring positions wrap around every edge, those off a non-periodic edge or more than
# one node off the lattice are skipped in the first step (nothing written to their LDS
# slots).  A ring node's pulls come from the gathered column (x - 1 on the outlet column)
# and the rows and columns around it, wrapped, so every pull is a lattice node and every
# address a node offset within a plane.  In the second step a tile node reads the LDS ring
# at (column - cx, row - cy), the outlet node at the column before it: those reads must land
# on ring positions written in the first step, except an inlet node's slots whose sources
# lie off the lattice (the three slots Zou/He replaces; every slot for the equilibrium
# inlet).

def _edge_cases():
    from cases import _hills, _step_geometry
    yy, xx = np.mgrid[0:160, 0:448]
    out = [(G.desc_for_case(G.CYLINDER, 448, 160), ((xx - 112) ** 2 + (yy - 81) ** 2 <= 144).astype(np.uint8)),
           (G.desc_for_case(G.STEP, 384, 144), _step_geometry(384, 144)),
           (G.desc_for_case(G.PERIODIC_HILLS, 384, 144), _hills(384, 144))]
    # the equilibrium inlet, periodic y edges
    d = G.desc_for_case(G.CYLINDER, 320, 96)
    d.inlet = G.INLET_EQUILIBRIUM
    out.append((d, np.zeros((96, 320), np.uint8)))
    # widths and heights that cut the last tiles off, narrower and lower than a tile and
    # its ring, with an outlet column alone in the last tile (nx % 64 == 1)
    for nx, ny in ((300, 100), (65, 30), (129, 41), (40, 7)):
        yy, xx = np.mgrid[0:ny, 0:nx]
        out.append((G.desc_for_case(G.CYLINDER, nx, ny), ((xx - nx // 3) ** 2 + (yy - ny // 2) ** 2 <= 4).astype(np.uint8)))
        out.append((G.desc_for_case(G.PERIODIC_HILLS, nx, ny), np.zeros((ny, nx), np.uint8)))
    return out


@pytest.mark.parametrize("rows", [8, 16, 32])
def test_edge_tile_pulls_and_ring_reads(rows):
    from cases import node_kinds, ring_nodes, tile_classes
    n_edge = 0
    for d, m in _edge_cases():
        ny, nx = m.shape
        special, allowed = node_kinds(m, d)
        for x0, y0, cls in tile_classes(m, d, rows):
            if cls != 2:
                continue
            n_edge += 1
            assert x0 >= 0 and y0 >= 0
            ys, xs, inm = ring_nodes(d, nx, ny, x0, y0, rows)
            # the kernel's wrap (once) and clamp of the ring positions: inside the lattice
            uy, ux = np.meshgrid(y0 - 1 + np.arange(rows + 2), x0 - 1 + np.arange(66), indexing="ij")
            gx = np.minimum(np.where(ux < 0, ux + nx, np.where(ux >= nx, ux - nx, ux)), nx - 1)
            gy = np.minimum(np.where(uy < 0, uy + ny, np.where(uy >= ny, uy - ny, uy)), ny - 1)
            assert gx.min() >= 0 and gy.min() >= 0
            assert np.array_equal(gx[inm], np.broadcast_to(xs, gx.shape)[inm])
            assert np.array_equal(gy[inm], np.broadcast_to(ys[:, None], gy.shape)[inm])
            # step 1: the pulls of every ring position (skipped ones included: their loads
            # are issued) are lattice nodes
            px = np.where(d.outlet & (gx == nx - 1), gx - 1, gx)
            for j in range(9):
                sx = np.where(px - CX[j] < 0, nx - 1, np.where(px - CX[j] >= nx, 0, px - CX[j]))
                sy = (gy - CY[j]) % ny
                assert sx.min() >= 0 and sx.max() < nx and sy.min() >= 0 and sy.max() < ny
                # off-lattice sources are read only by inlet nodes (replaced slots)
                off = (px - CX[j] < 0) | (px - CX[j] >= nx)
                if d.x_edge != G.EDGE_PERIODIC:
                    assert not (off & inm & (gx != 0)).any()
            # step 2: LDS reads of the tile nodes
            for ty in range(rows):
                for tx in range(64):
                    x, y = x0 + tx, y0 + ty
                    if x >= nx or y >= ny:
                        continue  # not a node: the kernel skips it
                    lx, ly = tx + 1, ty + 1
                    slx = lx - 1 if (d.outlet and x == nx - 1) else lx
                    inlet = x == 0 and d.inlet != G.INLET_NONE
                    for j in range(9):
                        rx, ry = slx - CX[j], ly - CY[j]
                        assert 0 <= rx <= 65 and 0 <= ry <= rows + 1
                        if not inm[ry, rx]:
                            assert inlet and (d.inlet == G.INLET_EQUILIBRIUM or j in (1, 5, 8)), (x, y, j)
    assert n_edge > 0

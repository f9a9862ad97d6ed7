"""GPU parity: the HIP step against the FP64 oracle on identical inputs.

Tolerances (absolute, on populations f_i ~ O(0.5); forces relative to |F|+1e-3):
  F64 storage  : 1e-12 populations, 1e-10 forces   (same algorithm in FP64)
  F32 storage  : 2e-6 populations, 1e-4 forces     (FP32 arithmetic on f - w)
  F16S (2^8)   : 2e-3 populations                  (binary16 deviations)
  BF16S (2^8)  : 1.5e-2 populations                (bfloat16 deviations)
"""
import numpy as np
import pytest

import lbm2d as G
from cases import CASES, build_gpu, build_oracle

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 2e-6, "f16s": 2e-3, "bf16s": 1.5e-2}
FTOL = {"f64": 1e-10, "f32": 1e-4}


def _run(name, storage, steps=None, graph_steps=None):
    c = CASES[name]
    n = steps or c.steps
    ens, o = build_gpu(c, storage)
    if graph_steps:
        ens.set_launch(graph_steps=graph_steps)
    fx_o, fy_o = [], []
    for _ in range(n):
        o.step(c.tau, c.u)
        fx, fy = o.total_force()
        fx_o.append(fx)
        fy_o.append(fy)
    ens.step(n)
    f_gpu = ens.download(0)
    gx, gy = ens.forces(0)
    return c, o, ens, f_gpu, np.array(fx_o), np.array(fy_o), gx[:, 0], gy[:, 0]


@pytest.mark.parametrize("name", sorted(CASES))
def test_f64_matches_oracle(name):
    c, o, ens, f, fx_o, fy_o, gx, gy = _run(name, "f64")
    print(f"[{name} f64] max|df| = {np.abs(f - o.f).max():.3e}")
    np.testing.assert_allclose(f, o.f, rtol=0, atol=TOL["f64"])
    if c.case != G.CAVITY:
        assert len(gx) == c.steps
        np.testing.assert_allclose(gx, fx_o, rtol=0, atol=FTOL["f64"] * (np.abs(fx_o).max() + 1e-3))
        np.testing.assert_allclose(gy, fy_o, rtol=0, atol=FTOL["f64"] * (np.abs(fx_o).max() + 1e-3))
    m = ens.mass(0)
    assert abs(m - o.mass()) <= 1e-10 * abs(o.mass())


@pytest.mark.parametrize("name", sorted(CASES))
def test_f32_matches_oracle(name):
    c, o, ens, f, fx_o, fy_o, gx, gy = _run(name, "f32")
    err = np.abs(f - o.f).max()
    print(f"[{name} f32] max|df| = {err:.3e}")
    assert err <= TOL["f32"]
    if c.case != G.CAVITY:
        scale = np.abs(fx_o).max() + 1e-3
        assert np.abs(gx - fx_o).max() <= FTOL["f32"] * scale
        assert np.abs(gy - fy_o).max() <= FTOL["f32"] * scale


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
@pytest.mark.parametrize("name", ["cylinder", "cavity", "les"])
def test_16bit_storage_error(name, storage):
    c, o, ens, f, *_ = _run(name, storage)
    err = np.abs(f - o.f).max()
    print(f"[{name} {storage}] max|df| = {err:.3e}")
    assert err <= TOL[storage]


def test_download_before_any_step_returns_upload():
    c = CASES["cylinder"]
    ens, o = build_gpu(c, "f64")
    np.testing.assert_array_equal(ens.download(0), o.f)


def test_graph_equals_single_launches():
    c = CASES["cylinder"]
    a, _ = build_gpu(c, "f32")
    b, _ = build_gpu(c, "f32")
    a.set_launch(graph_steps=20)
    a.step(40)  # two graph replays
    for _ in range(40):
        b.step_lattice(0, 1)
    np.testing.assert_array_equal(a.download(0), b.download(0))
    fa, _ = a.forces(0)
    fb, _ = b.forces(0)
    np.testing.assert_array_equal(fa, fb)  # force totals are summed in a fixed order


def test_odd_step_counts_and_interleaving():
    c = CASES["cylinder"]
    ens, o = build_gpu(c, "f64")
    ens.set_launch(graph_steps=4)
    done = 0
    for n in (1, 3, 5, 2, 7):
        ens.step(n)
        o.step(c.tau, c.u, n)
        done += n
        np.testing.assert_allclose(ens.download(0), o.f, rtol=0, atol=TOL["f64"])
    assert ens.steps_done(0) == done
    fx, _ = ens.forces(0)
    assert fx.shape == (done, 1)


def test_ensemble_of_uneven_lattices_matches_each_alone():
    names = ["cylinder", "cavity", "step", "plate"]
    from cases import apply_gpu_geometry, gpu_desc
    cs = [CASES[n] for n in names]
    orcs = [build_oracle(c) for c in cs]
    ens = G.Ensemble([gpu_desc(c) for c in cs], storage="f64")
    for l, (c, o) in enumerate(zip(cs, orcs)):
        apply_gpu_geometry(ens, l, c, o.obstacle.copy())
        ens.upload(l, o.f.copy())
    ens.set_launch(graph_steps=10)
    ens.step(30)
    for l, (c, o) in enumerate(zip(cs, orcs)):
        alone, _ = build_gpu(c, "f64", o=build_oracle(c))
        alone.step(30)
        o.step(c.tau, c.u, 30)
        np.testing.assert_allclose(ens.download(l), o.f, rtol=0, atol=TOL["f64"])
        # the ensemble kernel and the one-lattice kernel do the same arithmetic
        np.testing.assert_array_equal(ens.download(l), alone.download(0))
        np.testing.assert_array_equal(ens.forces(l)[0], alone.forces(0)[0])


def test_autotune_keeps_state():
    c = CASES["cylinder"]
    ens, o = build_gpu(c, "f32")
    ens.step(3)
    before = ens.download(0)
    (bx, rows, nodes, options), ms = ens.autotune()
    assert bx in (64, 128, 256, 512) and rows in (1, 2, 4, 8, 16) and nodes in (1, 2, 4, 8) and ms > 0
    assert options in range(8)  # XCD order, streaming stores, two steps per launch
    if options & 4:
        assert (bx, nodes) == (64, 1) and rows in (4, 8, 16)
    np.testing.assert_array_equal(ens.download(0), before)
    ens.step(5)
    o.step(c.tau, c.u, 8)
    assert np.abs(ens.download(0) - o.f).max() <= TOL["f32"]


@pytest.mark.parametrize("storage,nx,ny", [("f32", 1024, 512), ("f16s", 1024, 1024)])
def test_autotune_on_an_exactly_sized_pool(storage, nx, ny):
    """The autotuner steps two copies of the population buffers; the step kernel's pulls
    of a lattice's last rows reach past its pool, so the copies carry the buffers' guard.
    The populations fill exactly nine 2 MiB pages (1024 x 512 FP32, 1024 x 1024 binary16:
    no row padding, no slack after the pool).  Tuning leaves the state alone, and stepping
    afterwards matches, bit for bit, a handle that never tuned."""
    assert (9 * nx * ny * G.BYTES_PER_POP[storage]) % (2 << 20) == 0
    assert G.population_layout([G.desc_for_case(G.CYLINDER, nx, ny)])[1] == 9 * nx * ny
    desc = G.desc_for_case(G.CYLINDER, nx, ny, tau=0.6, u_in=0.1)
    mask = G.circle_mask(nx, ny, 256, 256, 40)
    runs = []
    for tune in (True, False):
        e = G.Ensemble([desc], storage=storage)
        e.set_obstacle(0, mask)
        e.set_bb_circle(0, 256, 256, 40, G.BB_MEI)
        e.init_equilibrium(0, 1.0, 0.1, 0.0)
        e.step(4)
        if tune:
            before = e.download(0)
            e.autotune()
            np.testing.assert_array_equal(e.download(0), before)
        e.step(20)
        runs.append((e.download(0), e.forces(0)[0]))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_nan_flag_trips():
    c = CASES["cylinder"]
    o = build_oracle(c)
    f = o.f.copy()
    f[20, 60, 3] = np.nan
    ens, _ = build_gpu(c, "f32", o=o)
    ens.upload(0, f)
    assert not ens.nan_flag(0)
    ens.step(2)
    assert ens.nan_flag(0)
    assert not ens.nan_flag(0)  # reading clears it


def test_params_change_between_steps():
    c = CASES["cylinder"]
    ens, o = build_gpu(c, "f64")
    ens.step(5)
    o.step(c.tau, c.u, 5)
    ens.set_params(0, 0.62, 0.08)
    ens.step(5)
    o.step(0.62, 0.08, 5)
    np.testing.assert_allclose(ens.download(0), o.f, rtol=0, atol=TOL["f64"])


def test_wall_distance_matches_oracle():
    import oracle as O
    c = CASES["les"]
    ens, o = build_gpu(c, "f32")
    np.testing.assert_array_equal(ens.wall_distance(0), O.wall_distance(o.obstacle))


@pytest.mark.parametrize("kind", ["random", "none", "corner", "walls", "wide"])
def test_device_wall_distance_equals_bfs(kind):
    """The device distance transform against the oracle's BFS (the fixpoint of the
    reference's sweeps) on lattices of several shapes -- no solid, one far corner, walls,
    rows longer than one chunk per thread -- in an ensemble, so lattice offsets count."""
    import oracle as O
    rng = np.random.default_rng({"random": 1, "none": 2, "corner": 3, "walls": 4, "wide": 5}[kind])
    shapes = {"random": (70, 45), "none": (33, 20), "corner": (300, 7), "walls": (64, 64), "wide": (1100, 9)}
    nx, ny = shapes[kind]
    m = np.zeros((ny, nx), np.uint8)
    if kind == "random":
        m[rng.random((ny, nx)) < 0.04] = 1
    elif kind == "corner":
        m[ny - 1, 0] = 1
    elif kind == "walls":
        m[0, :] = m[:, 0] = m[:, nx - 1] = 1
    elif kind == "wide":
        m[3, 517] = m[0, 1099] = 1
    other = G.desc_for_case(G.CYLINDER, 40, 30, tau=0.6)
    d = G.desc_for_case(G.RIBS, nx, ny, tau=0.6, les=True, cs=0.1)
    ens = G.Ensemble([other, d], storage="f64")
    om = np.zeros((30, 40), np.uint8)
    om[10:14, 10:14] = 1
    ens.set_obstacle(0, om)
    ens.set_obstacle(1, m)
    np.testing.assert_array_equal(ens.wall_distance(1), O.wall_distance(m))
    np.testing.assert_array_equal(ens.wall_distance(0), O.wall_distance(om))


@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_every_launch_shape_and_option_is_bitwise_identical(storage):
    """Each compiled (block, rows, nodes per thread) shape, with and without the XCD
    tile order and streaming stores, computes the same populations bit for bit."""
    c = CASES["cylinder"]
    shapes = [(256, 1, 1), (64, 1, 1), (256, 2, 1), (128, 1, 2), (64, 1, 4), (128, 2, 4),
              (64, 1, 8)]
    ref = None
    for bx, rows, v in shapes:
        for opt in (0, 1, 2, 3):
            ens, _ = build_gpu(c, storage)
            try:
                ens.set_launch(bx, rows, v, opt)
            except G.LbmError:
                assert G.BYTES_PER_POP[storage] * v > 16
                continue
            ens.step(7)
            f = ens.download(0)
            if ref is None:
                ref = f
            else:
                np.testing.assert_array_equal(f, ref, err_msg=f"shape {(bx, rows, v)} options {opt}")


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s"])
@pytest.mark.parametrize("name", ["cylinder", "les"])
def test_device_equilibrium_init_equals_upload(name, storage):
    """lbm_init_equilibrium builds the drivers' uniform initial state on the device:
    bit for bit what uploading compute_equilibrium's array gives, before and after steps."""
    from cases import apply_gpu_geometry, gpu_desc
    c = CASES[name]
    o = build_oracle(c)
    nx, ny = o.nx, o.ny
    one = np.ones((ny, nx))
    feq = G.equilibrium(one, 0.1 * one, 0.0 * one)
    a = G.Ensemble([gpu_desc(c)], storage=storage)
    b = G.Ensemble([gpu_desc(c)], storage=storage)
    for e in (a, b):
        apply_gpu_geometry(e, 0, c, o.obstacle.copy())
    a.init_equilibrium(0, 1.0, 0.1, 0.0)
    b.upload(0, feq)
    np.testing.assert_array_equal(a.download(0), b.download(0))
    a.step(7)
    b.step(7)
    np.testing.assert_array_equal(a.download(0), b.download(0))
    np.testing.assert_array_equal(a.forces(0)[0], b.forces(0)[0])
    assert a.steps_done(0) == 7


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s"])
@pytest.mark.parametrize("name", ["cylinder", "near_wall"])
def test_device_perturbation_matches_oracle(name, storage):
    """lbm_perturb_velocity applies the drivers' seed-42 trigger (src/main.cpp:67-83) to
    the device-built equilibrium: the state, and the run after it, match the oracle's
    host-side perturbation to the storage's tolerance (F64 to rounding of f - w)."""
    from cases import apply_gpu_geometry, gpu_desc
    c = CASES[name]
    o = build_oracle(c)
    x0, x1 = c.pert
    ens = G.Ensemble([gpu_desc(c)], storage=storage)
    apply_gpu_geometry(ens, 0, c, o.obstacle.copy())
    ens.init_equilibrium(0, *c.init)
    dv = G.seed42_draws(o.obstacle, x0, x1)
    ens.perturb_velocity(0, x0, min(x1, o.nx), dv)
    tol = {"f64": 1e-15, "f32": 2e-7, "f16s": 2e-3}[storage]
    err0 = np.abs(ens.download(0) - o.f).max()
    assert err0 <= tol, err0
    ens.step(c.steps)
    for _ in range(c.steps):
        o.step(c.tau, c.u)
    err = np.abs(ens.download(0) - o.f).max()
    print(f"[{name} {storage}] perturbed start {err0:.3e}, after {c.steps} steps {err:.3e}")
    assert err <= {"f64": 1e-12, "f32": TOL["f32"], "f16s": TOL["f16s"]}[storage]
    with pytest.raises(G.LbmError):
        ens.perturb_velocity(0, x0, min(x1, o.nx), dv)


def test_device_rasterisation_matches_place_functions():
    """lbm_place_obstacles is place_cylinder / place_polygon (src/lbm.hpp:931-964) cell
    for cell: integer circles (some crossing the lattice edges) and polygons (rotated
    plates, a concave one, one- and two-vertex ones, one crossing the edge) against the
    oracle's restatement; a fractional circle against the same FP64 expression in numpy;
    cells already solid stay solid; the mask then drives the link build as an uploaded
    one does."""
    import oracle as O
    nx, ny = 300, 200
    desc = G.desc_for_case(G.CYLINDER, nx, ny, tau=0.6, u_in=0.1)
    pre = np.zeros((ny, nx), np.uint8)
    pre[5:9, 250:260] = 1
    circles = [(60, 100, 30), (0, 0, 12), (299, 150, 7), (150, 199, 40), (20, 60, 1)]
    fcircle = (200.5, 60.25, 17.75)
    polys = [G.plate_polygon(120.0, 40.0, 155.0, 10.0, 2.0), G.plate_polygon(230.3, 120.7, 60.0, -13.0, 1.0),
             np.array([(30.0, 150.0), (90.0, 150.0), (90.0, 160.0), (45.5, 160.0), (45.5, 190.0), (30.0, 190.0)]),
             np.array([(10.0, 10.0)]), np.array([(20.0, 20.0), (40.0, 25.0)]),
             np.array([(280.0, 180.0), (330.0, 210.0), (270.0, 230.0)])]
    ens = G.Ensemble([desc], storage="f32")
    ens.set_obstacle(0, pre)
    ens.place_obstacles(0, circles + [fcircle], polys)
    got = ens.obstacle(0)
    o = O.Oracle(nx, ny, O.CYLINDER)
    o.set_obstacle(pre)
    for cx, cy, r in circles:
        o.place_cylinder(cx, cy, r)
    for xy in polys:
        o.place_polygon(xy)
    want = o.obstacle.copy()
    y, x = np.mgrid[0:ny, 0:nx]
    dx, dy = x - fcircle[0], y - fcircle[1]
    want |= (np.sqrt(dx * dx + dy * dy) < fcircle[2]).astype(np.uint8)
    print(f"solid cells: {int(got.sum())}, differing: {int((got != want).sum())}")
    np.testing.assert_array_equal(got, want)
    # the same lattice stepped from the device-built mask and from the uploaded one
    ref = G.Ensemble([desc], storage="f32")
    ref.set_obstacle(0, want)
    for e in (ens, ref):
        e.set_bb_circle(0, 60, 100, 30, G.BB_MEI)
        e.init_equilibrium(0, 1.0, 0.1, 0.0)
        e.step(9)
    np.testing.assert_array_equal(ens.download(0), ref.download(0))
    np.testing.assert_array_equal(ens.forces(0)[0], ref.forces(0)[0])
    # rejected on the host, before any device work: a vertex-less polygon, non-finite
    # values, an unknown lattice; the mask is unchanged by the failed calls
    before = ens.obstacle(0)
    with pytest.raises(G.LbmError):
        ens.place_obstacles(0, [], [np.zeros((0, 2))])
    with pytest.raises(G.LbmError):
        ens.place_obstacles(0, [(10.0, 10.0, np.nan)])
    with pytest.raises(G.LbmError):
        ens.place_obstacles(0, [], [np.array([(0.0, 0.0), (np.inf, 5.0), (3.0, 9.0)])])
    with pytest.raises(G.LbmError):
        ens.place_obstacles(1, [(10.0, 10.0, 3.0)])
    np.testing.assert_array_equal(ens.obstacle(0), before)
    # an empty call changes nothing but still resets the link set-up, as lbm_set_obstacle
    ens.place_obstacles(0)
    np.testing.assert_array_equal(ens.obstacle(0), before)


@pytest.mark.parametrize("case", ["cylinder", "hills", "cavity"])
def test_geometry_pass_on_a_porous_mask(case):
    """The device geometry pass on a random porous mask (25% solid nodes, runs across the
    periodic seam and along the edges, rows padded to 192): fluid and link-node counts as
    numpy counts them under the case's edge rules, and ten FP64 steps with simple
    bounce-back on every link against the oracle."""
    import oracle as O
    kind = {"cylinder": (G.CYLINDER, O.CYLINDER, 0.05), "hills": (G.PERIODIC_HILLS, O.PERIODIC_HILLS, 0.0),
            "cavity": (G.CAVITY, O.CAVITY, 0.05)}[case]
    nx, ny, tau = 131, 77, 0.8
    rng = np.random.default_rng(5)
    mask = (rng.random((ny, nx)) < 0.25).astype(np.uint8)
    mask[10:14, :3] = 1
    mask[10:14, -3:] = 1
    mask[0, 40:50] = 1
    mask[-1, 40:50] = 1
    # the outflow copy (src/lbm.hpp:63-73) of a solid node would give a fluid outlet node
    # zero density (u = 0/0, in the reference too): the outlet column follows its source
    mask[:, -1] = mask[:, -2]
    desc = G.desc_for_case(kind[0], nx, ny, tau=tau, u_in=kind[2])
    ens = G.Ensemble([desc], storage="f64")
    ens.set_obstacle(0, mask)
    o = O.Oracle(nx, ny, kind[1])
    o.set_obstacle(mask)
    o.init_equilibrium(1.0, kind[2], 0.0)
    ens.upload(0, o.f.copy())
    # fluid nodes with a solid neighbour along some e_i, neighbours wrapped on periodic edges
    solid = mask.astype(bool)
    near = np.zeros_like(solid)
    for i in range(1, 9):
        cx, cy = int(G.CXV[i]), int(G.CYV[i])
        sh = np.roll(solid, (-cy, -cx), axis=(0, 1))
        if desc.x_edge != G.EDGE_PERIODIC:
            if cx == 1:
                sh[:, -1] = False
            elif cx == -1:
                sh[:, 0] = False
        if desc.y_edge != G.EDGE_PERIODIC:
            if cy == 1:
                sh[-1, :] = False
            elif cy == -1:
                sh[0, :] = False
        near |= sh
    info = ens.info()
    assert info.fluid_cells == int((~solid).sum())
    assert info.link_cells == int((near & ~solid).sum())
    fx_o = []
    for _ in range(10):
        o.step(tau, kind[2])
        fx_o.append(o.total_force()[0])
    ens.step(10)
    err = np.abs(ens.download(0) - o.f).max()
    print(f"[{case}] link nodes {info.link_cells}, max|df| = {err:.3e}")
    assert err <= TOL["f64"]
    if kind[0] != G.CAVITY:
        fx = ens.forces(0)[0][:, 0]
        np.testing.assert_allclose(fx, fx_o, rtol=0, atol=FTOL["f64"] * (np.abs(fx_o).max() + 1e-3))


@pytest.mark.parametrize("shape", [(3, 3), (5, 200), (200, 3), (65, 4), (130, 7)])
@pytest.mark.parametrize("case", ["cylinder", "hills"])
def test_tiny_and_thin_lattices_match_oracle(case, shape):
    """Lattices made mostly of edge nodes (the list blocks' edge lanes, rows padded past
    nx, columns of two nodes): twelve FP64 steps against the oracle, with a solid node
    where the lattice has an interior."""
    import oracle as O
    nx, ny = shape
    kind = {"cylinder": (G.CYLINDER, O.CYLINDER, 0.05), "hills": (G.PERIODIC_HILLS, O.PERIODIC_HILLS, 0.0)}[case]
    mask = np.zeros((ny, nx), np.uint8)
    if nx >= 5 and ny >= 5:
        mask[ny // 2, nx // 2] = 1
    desc = G.desc_for_case(kind[0], nx, ny, tau=0.7, u_in=kind[2])
    o = O.Oracle(nx, ny, kind[1])
    o.set_obstacle(mask)
    o.init_equilibrium(1.0, kind[2], 0.0)
    rng = np.random.default_rng(nx * 1000 + ny)
    f0 = o.f * (1.0 + 0.01 * rng.standard_normal(o.f.shape)) * (mask == 0)[..., None]
    o.f[...] = f0
    ens = G.Ensemble([desc], storage="f64")
    ens.set_obstacle(0, mask)
    ens.upload(0, f0.copy())
    for _ in range(12):
        o.step(0.7, kind[2])
    ens.step(12)
    err = np.abs(ens.download(0) - o.f).max()
    print(f"[{case} {nx}x{ny}] max|df| = {err:.3e}")
    assert err <= TOL["f64"]


TINY_SHAPES = [(3, 3), (5, 200), (200, 3), (65, 4), (130, 7)]


def _tiny_lattice(case, nx, ny):
    import oracle as O
    kind = {"cylinder": (G.CYLINDER, O.CYLINDER, 0.05), "hills": (G.PERIODIC_HILLS, O.PERIODIC_HILLS, 0.0)}[case]
    mask = np.zeros((ny, nx), np.uint8)
    if nx >= 5 and ny >= 5:
        mask[ny // 2, nx // 2] = 1
    desc = G.desc_for_case(kind[0], nx, ny, tau=0.7, u_in=kind[2])
    o = O.Oracle(nx, ny, kind[1])
    o.set_obstacle(mask)
    o.init_equilibrium(1.0, kind[2], 0.0)
    rng = np.random.default_rng(nx * 1000 + ny)
    o.f[...] = o.f * (1.0 + 0.01 * rng.standard_normal(o.f.shape)) * (mask == 0)[..., None]
    return desc, mask, o, kind[2]


@pytest.mark.parametrize("rows", [8, 16])
@pytest.mark.parametrize("case", ["cylinder", "hills"])
def test_tiny_and_thin_lattices_two_steps_per_launch(case, rows):
    """Lattices lower or narrower than a two-step tile and its ring (the ring's positions
    more than one node off a periodic edge are skipped): each alone where the one-lattice
    kernel takes it, and all in one ensemble, two steps per launch, equal bit for bit to
    one step per launch and to the FP64 oracle within its bound."""
    lats = [_tiny_lattice(case, nx, ny) for nx, ny in TINY_SHAPES]
    ref = []
    for desc, mask, o, u in lats:
        e = G.Ensemble([desc], storage="f64")
        e.set_obstacle(0, mask)
        e.upload(0, o.f.copy())
        e.step(12)
        ref.append(e.download(0))
        e.close()
    alone = 0
    for (desc, mask, o, u), r in zip(lats, ref):
        e = G.Ensemble([desc], storage="f64")
        e.set_obstacle(0, mask)
        e.upload(0, o.f.copy())
        try:
            e.set_launch(64, rows, 1, 4, graph_steps=6)
        except G.LbmError:
            continue  # an outlet with nx % 64 == 1: the ensemble below takes it
        e.step(5)
        e.step(7)
        np.testing.assert_array_equal(e.download(0), r, err_msg=f"{desc.nx}x{desc.ny} alone")
        alone += 1
        e.close()
    assert alone >= 3
    ens = G.Ensemble([d for d, _, _, _ in lats], storage="f64")
    for l, (desc, mask, o, u) in enumerate(lats):
        ens.set_obstacle(l, mask)
        ens.upload(l, o.f.copy())
    ens.set_launch(64, rows, 1, 5, graph_steps=6)
    ens.step(5)
    ens.step(7)
    for l, ((desc, mask, o, u), r) in enumerate(zip(lats, ref)):
        np.testing.assert_array_equal(ens.download(l), r, err_msg=f"{desc.nx}x{desc.ny} in the ensemble")
        for _ in range(12):
            o.step(0.7, u)
        assert np.abs(r - o.f).max() <= TOL["f64"]
    ens.close()


def _q_circle(ccx, ccy, r, xf, yf, i):
    """q of a link to a circle (q_cylinder, src/lbm_types.hpp:118-138), FP64."""
    dx, dy = xf - ccx, yf - ccy
    ex, ey = float(G.CXV[i]), float(G.CYV[i])
    a = ex * ex + ey * ey
    b = 2.0 * (dx * ex + dy * ey)
    c = dx * dx + dy * dy - r * r
    disc = b * b - 4.0 * a * c
    if disc < 0.0:
        return 1.0
    sd = float(np.sqrt(disc))
    t1, t2 = (-b - sd) / (2.0 * a), (-b + sd) / (2.0 * a)
    if 0.0 < t1 <= 1.0:
        return t1
    if 0.0 < t2 <= 1.0:
        return t2
    return 1.0


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_explicit_links_equal_the_geometry_they_restate(storage):
    """lbm_set_links with the q of the circle and the rotating wall's velocity at each
    solid node, link by link, steps bit for bit as lbm_set_bb_circle + lbm_set_moving_wall
    (which match the oracle); splitting the links over two force slots leaves the state
    unchanged and the slots sum to the one-slot totals; links the mask does not have, bad
    schemes and slots beyond the lattice's are rejected."""
    nx, ny, cx, cy, r, omega, u = 96, 48, 24, 25, 6, 0.01, 0.05
    mask = G.circle_mask(nx, ny, cx, cy, r)
    desc = G.desc_for_case(G.CYLINDER, nx, ny, tau=0.6, u_in=u)
    f0 = G.equilibrium(np.ones((ny, nx)), np.full((ny, nx), u), np.zeros((ny, nx)))
    G.perturb_seed42(f0, mask, cx + 5, cx + 60)
    links, split = [], []
    for y in range(ny):
        for x in range(nx):
            if mask[y, x]:
                continue
            for i in range(1, 9):
                tx, ty = x + int(G.CXV[i]), (y + int(G.CYV[i])) % ny
                if not (0 <= tx < nx) or not mask[ty, tx]:
                    continue
                px, py = float(x + G.CXV[i]), float(y + G.CYV[i])
                k = (y * nx + x, i, G.BB_MEI, 0, _q_circle(cx, cy, r, x, y, i), -omega * (py - cy), omega * (px - cx))
                links.append(k)
                split.append(k[:3] + (1 if y < cy else 0,) + k[4:])
    runs = {}
    for name in ("geometry", "links", "split"):
        e = G.Ensemble([desc], storage=storage)
        e.set_obstacle(0, mask)
        if name == "geometry":
            e.set_bb_circle(0, cx, cy, r, G.BB_MEI)
            e.set_moving_wall(0, omega, cx, cy)
        elif name == "links":
            e.set_links(0, links)
        else:
            e.set_force_slots(0, 2)
            e.set_links(0, split)
        e.upload(0, f0.copy())
        e.step(30)
        runs[name] = (e, e.download(0), e.forces(0, n_slots=2 if name == "split" else 1))
    g, l, sp = runs["geometry"], runs["links"], runs["split"]
    np.testing.assert_array_equal(l[1], g[1])
    np.testing.assert_array_equal(l[2][0], g[2][0])
    np.testing.assert_array_equal(l[2][1], g[2][1])
    np.testing.assert_array_equal(sp[1], g[1])
    fx, fy = sp[2]
    scale = np.abs(g[2][0]).max() + 1e-3
    np.testing.assert_allclose(fx.sum(axis=1), g[2][0][:, 0], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(fy.sum(axis=1), g[2][1][:, 0], rtol=0, atol=1e-12 * scale)
    assert np.abs(fx[:, 0]).max() > 0 and np.abs(fx[:, 1]).max() > 0
    # the links are checked when the geometry is committed (here by the upload)
    e = runs["links"][0]
    fluid_without_links = (cy * nx + 2, 1, G.BB_MEI, 0, 0.5, 0.0, 0.0)
    for bad, why in ((fluid_without_links, "fluid-to-solid"), (links[0][:2] + (3,) + links[0][3:], "scheme"),
                     (links[0][:3] + (5,) + links[0][4:], "slot")):
        e.set_links(0, [bad])
        with pytest.raises(G.LbmError, match=why):
            e.upload(0, f0.copy())


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16s"])
def test_macros_equal_compute_macros_of_download(storage):
    """lbm_macros = compute_macros (src/lbm_types.hpp:278-288) of the downloaded state, bit
    for bit, on a whole lattice and on a row range; before any step and after steps."""
    c = CASES["cylinder"]
    ens, o = build_gpu(c, storage)
    for steps in (0, 5):
        ens.step(steps)
        f = ens.download(0)
        rho = np.zeros(f.shape[:2])
        mx = np.zeros(f.shape[:2])
        my = np.zeros(f.shape[:2])
        for i in range(9):
            rho = rho + f[..., i]
            mx = mx + f[..., i] * G.CXV[i]
            my = my + f[..., i] * G.CYV[i]
        with np.errstate(invalid="ignore", divide="ignore"):
            u, v = mx / rho, my / rho
        r2, u2, v2 = ens.macros(0)
        np.testing.assert_array_equal(r2, rho)
        np.testing.assert_array_equal(u2, u)
        np.testing.assert_array_equal(v2, v)
        r3, u3, v3 = ens.macros(0, 7, 19)
        np.testing.assert_array_equal(u3, u[7:19])
        np.testing.assert_array_equal(r3, rho[7:19])


def test_storage_shift_round_trip_and_errors():
    """The storage shift changes the encoding, not the state: download after upload is the
    upload to 16-bit rounding, steps stay within the 16-bit bound of the oracle, and the
    FP32 / FP64 forms refuse it."""
    c = CASES["cylinder"]
    o = build_oracle(c)
    from cases import apply_gpu_geometry, gpu_desc
    for storage, sexp in (("f16s", 8), ("f16s", 14), ("bf16s", 8)):
        e = G.Ensemble([gpu_desc(c)], storage=storage, scale_exp=sexp)
        apply_gpu_geometry(e, 0, c, o.obstacle.copy())
        e.set_storage_shift(0, c.u, 0.0)
        e.upload(0, o.f.copy())
        err0 = np.abs(e.download(0) - o.f).max()
        assert err0 <= (1e-5 if storage == "f16s" else 2e-4), (storage, sexp, err0)
    with pytest.raises(G.LbmError):
        G.Ensemble([gpu_desc(c)], storage="f32").set_storage_shift(0, 0.1, 0.0)
    # stepping: same bound as the unshifted form against the oracle
    ens = G.Ensemble([gpu_desc(c)], storage="f16s", scale_exp=14)
    apply_gpu_geometry(ens, 0, c, o.obstacle.copy())
    ens.set_storage_shift(0, c.u, 0.0)
    ens.upload(0, o.f.copy())
    for _ in range(c.steps):
        o.step(c.tau, c.u)
    ens.step(c.steps)
    assert np.abs(ens.download(0) - o.f).max() <= TOL["f16s"]


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16s"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_two_steps_per_launch_is_bitwise_identical(name, storage):
    """Two steps per launch (tile and ring in LDS) against one step per launch, on every
    case's boundaries: the same populations and force totals bit for bit after odd and
    even step counts (graph and single launches)."""
    c = CASES[name]
    a, _ = build_gpu(c, storage)
    b, _ = build_gpu(c, storage)
    a.set_launch(64, 1, 1, 0)
    for rows, opt in ((4, 4), (8, 5), (8, 6), (4, 7), (16, 5)):
        try:
            b.set_launch(64, rows, 1, opt, graph_steps=10)
        except G.LbmError:
            pytest.skip("two steps per launch do not apply to this lattice")
        for n in (1, 5, 12):
            a.step(n)
            b.step(n)
            np.testing.assert_array_equal(b.download(0), a.download(0), err_msg=f"rows {rows} options {opt} n {n}")
    fa, fya = a.forces(0)
    fb, fyb = b.forces(0)
    np.testing.assert_array_equal(fb, fa)
    np.testing.assert_array_equal(fyb, fya)
    assert a.steps_done(0) == b.steps_done(0)


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
def test_stochastic_rounding_repeats_in_every_launch_form(storage):
    """With stochastic rounding the bits come from (seed, step, node): launch shapes,
    two steps per launch, graphs and single launches all give the same populations bit
    for bit; a different seed gives different ones, still within the form's bound."""
    c = CASES["cylinder"]
    ref = None
    for launch in ((256, 1, 1, 0, 100), (64, 1, 4, 3, 100), (64, 8, 1, 5, 6), (128, 1, 1, 2, 0)):
        ens, o = build_gpu(c, storage)
        ens.set_stochastic_rounding(True, 1234)
        ens.set_launch(*launch[:4], graph_steps=launch[4]) if launch[4] else ens.set_launch(*launch[:4])
        for n in (3, 8):
            ens.step(n)
        f = ens.download(0)
        if ref is None:
            ref = f
            for _ in range(11):
                o.step(c.tau, c.u)
            assert np.abs(f - o.f).max() <= TOL[storage]
        else:
            np.testing.assert_array_equal(f, ref, err_msg=f"launch {launch}")
    other, _ = build_gpu(c, storage)
    other.set_stochastic_rounding(True, 99)
    other.step(11)
    assert not np.array_equal(other.download(0), ref)
    with pytest.raises(G.LbmError):
        G.Ensemble([G.desc_for_case(G.CYLINDER, 64, 32)], storage="f32").set_stochastic_rounding(True)

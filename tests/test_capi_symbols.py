"""The C-ABI library loads and exports every entry point include/lbm_gpu.h declares.

No compute calls: these run on a host without a GPU.
"""
import ctypes

import lbm2d as G


def test_library_exports_every_declared_function():
    names = G.declared_functions()
    assert len(names) >= 20, names
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_function():
    L = G.lib()
    for n in G.declared_functions():
        assert getattr(L, n).argtypes is not None or n == "lbm_last_error", n


def test_desc_for_case_mirrors_reference_branches():
    # src/lbm.hpp:408-484 (edges), 530-539 (boundaries), 543 (forces)
    cav = G.desc_for_case(G.CAVITY, 16, 16)
    assert (cav.x_edge, cav.y_edge, cav.lid, cav.forces, cav.inlet, cav.outlet) == \
        (G.EDGE_WALL, G.EDGE_WALL, 1, 0, G.INLET_NONE, 0)
    for c in (G.RIBS, G.PERIODIC_HILLS):
        d = G.desc_for_case(c, 16, 16)
        assert (d.x_edge, d.y_edge, d.forces, d.inlet, d.outlet, d.lid) == \
            (G.EDGE_PERIODIC, G.EDGE_PERIODIC, 1, G.INLET_NONE, 0, 0)
    st = G.desc_for_case(G.STEP, 30, 30)
    assert (st.inlet, st.inlet_y0, st.inlet_h, st.outlet) == (G.INLET_PARABOLIC, 10, 19, 1)
    for c in (G.CYLINDER, G.FLAT_PLATE, G.ROTATING_CYLINDER, G.CYLINDER_NEAR_WALL, G.URBAN_CANYON):
        d = G.desc_for_case(c, 16, 16)
        assert (d.x_edge, d.y_edge, d.inlet, d.outlet, d.forces) == \
            (G.EDGE_OPEN, G.EDGE_PERIODIC, G.INLET_ZOU_HE, 1, 1)


def test_errors_are_reported_not_raised_across_the_abi():
    L = G.lib()
    d = G.LbmDesc()
    assert L.lbm_desc_for_case(99, 8, 8, ctypes.byref(d)) != 0
    assert b"case" in L.lbm_last_error()
    h = ctypes.c_void_p()
    bad = (G.LbmDesc * 1)()
    bad[0].nx, bad[0].ny, bad[0].tau = 2, 2, 1.0
    assert L.lbm_create(bad, 1, 1, 0, None, ctypes.byref(h)) != 0
    assert b"3x3" in L.lbm_last_error()
    # within 2^31 nodes, but the rows padded to 64 nodes (int32 node indices) are not
    bad[0].nx, bad[0].ny = 65, 33038209
    assert 65 * 33038209 <= 2 ** 31 < 128 * 33038209
    assert L.lbm_create(bad, 1, 1, 0, None, ctypes.byref(h)) != 0
    assert b"too large" in L.lbm_last_error()

"""The reference's raw vector files (robtop_amd/formats.py): node order of
enumerate_gs_subset (Grid.cpp:2045-2106) on hand-worked cases, and round trips."""
import os
import tempfile

import numpy as np
import pytest

from robtop_amd import formats


def test_order_all_active_2x2x2():
    slot, n = formats.gs_order(np.ones(8, bool), (2, 2, 2))
    assert n == 8 * 32                       # one node per colour, each block padded to 32
    for z in range(2):
        for y in range(2):
            for x in range(2):
                assert slot[x + 2 * (y + 2 * z)] == 32 * (x + 2 * y + 4 * z)


def test_order_lexicographic_inside_a_colour():
    slot, n = formats.gs_order(np.ones(64, bool), (4, 4, 4))
    assert n == 8 * 32                       # 8 nodes per colour
    lin = lambda x, y, z: x + 4 * (y + 4 * z)
    colour0 = [(0, 0, 0), (2, 0, 0), (0, 2, 0), (2, 2, 0), (0, 0, 2), (2, 0, 2), (0, 2, 2), (2, 2, 2)]
    assert [slot[lin(*p)] for p in colour0] == list(range(8))
    assert slot[lin(1, 0, 0)] == 32 and slot[lin(3, 0, 0)] == 33     # colour 1 starts after block 0
    assert slot[lin(1, 1, 1)] == 7 * 32


def test_order_skips_absent_nodes_and_pads():
    act = np.ones(3 * 2 * 1, bool)
    act[0] = False                           # node (0, 0, 0), colour 0
    slot, n = formats.gs_order(act, (3, 2, 1))
    # colour 0: (2,0,0) only; colour 1: (1,0,0); colour 2: (0,1,0), (2,1,0); colour 3: (1,1,0)
    assert n == 4 * 32
    assert slot[0] == -1 and slot[2] == 0 and slot[1] == 32
    assert slot[3] == 64 and slot[5] == 65 and slot[4] == 96


def test_round_trip():
    rng = np.random.default_rng(0)
    dims = (7, 5, 6)
    act = rng.random(np.prod(dims)) < 0.7
    v = rng.standard_normal((3, act.size))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "u")
        formats.write_vectors(p, v, act, dims)
        _, n = formats.gs_order(act, dims)
        flat = formats.read_vector(p)
        assert flat.size == 3 * n
        back = formats.read_vectors(p, act, dims)
    assert np.array_equal(back[:, act], v[:, act])
    assert np.all(back[:, ~act] == 0)
    # interleaved: the first active colour-0 node's x, y, z are the first three values
    slot, _ = formats.gs_order(act, dims)
    first = int(np.flatnonzero(slot == 0)[0])
    assert np.array_equal(flat[:3], v[:, first])


@pytest.mark.gpu
def test_hierarchy_displacement_to_force_file():
    """writeDisplacement then readForce through the hierarchy returns the field."""
    import robtop_amd as rt
    from robtop_amd.workloads import box_bits, cube_side
    print("[formats] 32x16x16 box", flush=True)
    R = cube_side(32, 16, 16)
    hg = rt.HierarchyGrid.genFromMesh(box_bits(R, 32, 16, 16), R)
    g0 = hg[0]
    _, vact = g0.masks()
    rng = np.random.default_rng(1)
    u = rng.standard_normal((3, g0.n_nodes_box)) * vact
    hg.set_displacement(u)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "disp")
        hg.write_displacement(p)
        hg.read_force(p)
    assert np.array_equal(hg.get_force()[:, vact], u[:, vact])
    hg.close()

"""The reference's raw vector files ("bio" format, utils/binaryIO.h:69-108).

``bio::write_vectors`` writes N FP64 vectors interleaved, v0[0] v1[0] v2[0]
v0[1] ... (binaryIO.h:80-90), with no header; ``bio::read_vector`` reads the
whole file back as one flat array (binaryIO.h:92-108). Node vectors in these
files (``writeDisplacement``, Grid.cpp:886-897; ``readForce``, Grid.cpp:1679-1700)
are in the reference's Gauss-Seidel node order (``enumerate_gs_subset``,
Grid.cpp:2045-2106):
  - colour = (x % 2) + 2 (y % 2) + 4 (z % 2) of the node's lattice position;
  - colour blocks one after another, colour 0 first;
  - inside a block, nodes in lexicographic order (x fastest, then y, then z);
  - every block padded with unused slots to a multiple of 32 (snippet::Round<32>).
The hierarchy's box is a sub-box of the reference's cube whose origin is a
multiple of the coarsest spacing, so box parity is lattice parity and box
order is lexicographic order.
"""
import numpy as np


def gs_order(vact_box, nodes_xyz):
    """(slot of each active node in the reference's node order, total slots).

    vact_box: active-node mask of the finest level's node box, x fastest;
    nodes_xyz: the box size (nx, ny, nz). Returns an int64 array over the box
    (-1 for absent nodes) and the padded length n_gsvertices."""
    nx, ny, nz = (int(v) for v in nodes_xyz)
    act = np.asarray(vact_box).reshape(nz, ny, nx).astype(bool)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    colour = (x % 2) + 2 * (y % 2) + 4 * (z % 2)
    slot = np.full(act.shape, -1, np.int64)
    base = 0
    for c in range(8):
        sel = act & (colour == c)
        cnt = int(sel.sum())
        slot[sel] = base + np.arange(cnt)          # boolean indexing walks x fastest, then y, then z
        base += (cnt + 31) // 32 * 32
    return slot.ravel(), base


def write_vectors(path, v_box, vact_box, nodes_xyz):
    """Write a node vector field (3, nodes in the box) in the reference's layout."""
    slot, n = gs_order(vact_box, nodes_xyz)
    v = np.asarray(v_box, np.float64).reshape(3, -1)
    out = np.zeros((n, 3), np.float64)
    live = slot >= 0
    out[slot[live]] = v[:, live].T
    out.tofile(path)


def read_vectors(path, vact_box, nodes_xyz):
    """Read a node vector field written in the reference's layout into (3, nodes in the box)."""
    slot, n = gs_order(vact_box, nodes_xyz)
    flat = np.fromfile(path, dtype=np.float64)
    if flat.size != 3 * n:
        raise ValueError(f"{path}: {flat.size} values, expected 3 x {n} (the reference's padded node count)")
    a = flat.reshape(n, 3)
    live = slot >= 0
    out = np.zeros((3, slot.size))
    out[:, live] = a[slot[live]].T
    return out


def write_vector(path, values):
    """bio::write_vector: one flat array, raw little-endian."""
    np.ascontiguousarray(values).tofile(path)


def read_vector(path, dtype=np.float64):
    """bio::read_vector: the whole file as one flat array."""
    return np.fromfile(path, dtype=dtype)

"""Seeded test problems shared by the oracle and GPU tests."""
import numpy as np

from robtop_amd.workloads import _pack_lsb


def active_cube(kind, R):
    """[x, y, z] boolean solid sets used by the tests."""
    a = np.zeros((R, R, R), bool)
    if kind == "box3":        # 3 layers (L0, L2 coarsest)
        a[:24, :12, :12] = True
    elif kind == "holes3":    # ragged: holes and a notch
        a[:24, :12, :12] = True
        a[5:9, 3:7, 2:6] = False
        a[15:24, 8:12, 0:4] = False
        a[11, :, 7] = False
    elif kind == "box4":      # 4 layers (L0, L2, L3)
        a[:48, :24, :24] = True
    elif kind == "lshape4":   # 4 layers, L-shaped, offset from the origin
        a[8:56, 4:28, 4:28] = True
        a[32:56, 16:28, 4:28] = False
    elif kind == "cantilever64":   # the 64x32x32 cantilever box
        a[:64, :32, :32] = True
    else:
        raise ValueError(kind)
    return a


def bits_of(a):
    return _pack_lsb(np.ascontiguousarray(a.transpose(2, 1, 0)).reshape(-1))


def problem_inputs(lv0, seed, fixed_plane_x=True, rho_range=(0.05, 1.0)):
    """(fixed node flags, force [3, nn], rho [ne]) on the finest box of an
    oracle level dict (keys n, e, o, vact, eact)."""
    rng = np.random.default_rng(seed)
    n, e = lv0["n"], lv0["e"]
    vact = lv0["vact"].reshape(n[2], n[1], n[0])
    eact = lv0["eact"].reshape(e[2], e[1], e[0])
    fixed = np.zeros_like(vact)
    xs = np.arange(n[0]) + lv0["o"][0]
    first = xs[np.nonzero(vact.any(axis=(0, 1)))[0][0]]
    fixed[:, :, xs == first] = True
    fixed &= vact
    f = rng.standard_normal((3, n[2], n[1], n[0])) * vact[None]
    f[:, fixed] = 0
    rho = rng.uniform(*rho_range, size=(e[2], e[1], e[0])).astype(np.float32) * eact
    return fixed.reshape(-1), f.reshape(3, -1), rho.reshape(-1)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.max(np.abs(a - b))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d

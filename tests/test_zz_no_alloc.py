"""rt_voxelize_mesh_noalloc, the entry point that allocates no device memory
(its scratch is a static per-device array), against rt_voxelize_mesh on the
same meshes: identical bits, cube side and box, on two calls in a row (the
second with a smaller mesh and cube, so stale scratch and stale bits of the
first call must not leak into it), and its size and index errors."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torus(c, R, r, nu, nv):
    u = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    v = np.linspace(0, 2 * np.pi, nv, endpoint=False)
    U, W = np.meshgrid(u, v, indexing="ij")
    V = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1)
    idx = lambda i, j: (i % nu) * nv + (j % nv)
    F = []
    for i in range(nu):
        for j in range(nv):
            F += [[idx(i, j), idx(i + 1, j), idx(i, j + 1)], [idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)]]
    return (V.reshape(-1, 3) + np.asarray(c)).astype(np.float32), np.array(F, np.int32)


def _hip():
    """The HIP runtime librobtop_amd itself uses (torch ships a second copy;
    device memory of one runtime must not be handed to the other)."""
    from robtop_amd._lib import lib
    lib()
    paths = sorted({l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l and "/torch/" not in l})
    assert paths, "librobtop_amd's HIP runtime is not mapped"
    H = ctypes.CDLL(paths[0])
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return H


class _Dev:
    """a device buffer of the caller (allocated here, by the test)"""

    def __init__(self, H, nbytes):
        self.H, self.p = H, ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(nbytes, 4)) == 0

    def put(self, a):
        assert self.H.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0

    def get(self, a):
        assert self.H.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, 2) == 0
        return a

    def free(self):
        self.H.hipFree(self.p)


def _noalloc(V, F, reso, words=None):
    from robtop_amd._lib import lib
    H = _hip()
    R = 1
    while R <= reso:
        R <<= 1
    nw = (R ** 3 + 31) // 32 if words is None else words
    V = np.ascontiguousarray(V, np.float32)
    F = np.ascontiguousarray(F, np.int32)
    dv, df, db = _Dev(H, V.nbytes), _Dev(H, F.nbytes), _Dev(H, 4 * max(nw, 1))
    dv.put(V)
    df.put(F)
    assert H.hipMemset(db.p, 0xFF, 4 * max(nw, 1)) == 0     # garbage the call must clear
    reso3 = np.zeros(3, np.int32)
    box = np.zeros(6, np.float32)
    code = lib().rt_voxelize_mesh_noalloc(dv.p, V.shape[0], df.p, F.shape[0], int(reso), db.p, nw,
                                          reso3.ctypes.data_as(ctypes.c_void_p), box.ctypes.data_as(ctypes.c_void_p))
    bits = db.get(np.zeros(max(nw, 1), np.uint32))[:nw]
    for d in (dv, df, db):
        d.free()
    return code, bits, int(reso3[0]), box.reshape(2, 3)


def test_noalloc_voxelizer_matches_allocating_entry_point():
    import robtop_amd as rt
    cases = [("torus 96x48 reso 127", _torus((60, 64, 70), 40, 12, 96, 48), 127),
             ("torus 40x20 reso 63", _torus((30, 31, 29), 20, 7, 40, 20), 63)]
    for name, (V, F), reso in cases:
        print(f"[noalloc] {name}: {F.shape[0]} faces", flush=True)
        code, bits, R, box = _noalloc(V, F, reso)
        assert code == 0, rt._lib.lib().rt_last_error().decode()
        ref_bits, ref_R, ref_box = rt.voxelize_mesh(V, F, reso)
        assert R == ref_R and np.array_equal(box, ref_box)
        assert np.array_equal(bits, ref_bits), (name, int(np.count_nonzero(bits != ref_bits)))
        assert np.count_nonzero(bits) > 0


def test_noalloc_voxelizer_errors():
    from robtop_amd._lib import RT_EARG, lib
    V, F = _torus((30, 31, 29), 20, 7, 40, 20)
    print("[noalloc] bit buffer too small", flush=True)
    code, _, _, _ = _noalloc(V, F, 63, words=100)
    assert code == RT_EARG and b"too small" in lib().rt_last_error()
    print("[noalloc] face index out of range", flush=True)
    Fb = F.copy()
    Fb[5, 1] = V.shape[0]
    code, _, _, _ = _noalloc(V, Fb, 63)
    assert code == RT_EARG and b"out of range" in lib().rt_last_error()
    print("[noalloc] more faces than the scratch holds", flush=True)
    big = np.zeros(((1 << 19) + 1, 3), np.int32)
    code, _, _, _ = _noalloc(V, big, 63)
    assert code == RT_EARG and b"scratch limit" in lib().rt_last_error()

"""ORACLE — test infrastructure only.

ctypes wrapper over ``oracle/liboracle.so``, the FP64 CPU restatement of the
reference's multigrid FEM / SIMP path (``oracle/oracle.cpp`` cites the
reference file:line of each function). Only ``tests/``, the smoke check and the
CPU-baseline leg of ``bench.py`` import this module; the library under test
never does.

Arrays use the oracle's layout: a dense box per level, x fastest; vectors are
``[3, nnodes]`` float64; stencils ``[27, 9, nnodes]``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        F = ctypes.c_float
        sig = {
            "or_last_error": ([], ctypes.c_char_p),
            "or_create": ([I, P], P),
            "or_destroy": ([P], None),
            "or_num_levels": ([P], I),
            "or_num_layers": ([P], I),
            "or_layer_active": ([P, I], I),
            "or_level_info": ([P, I, P], None),
            "or_get_masks": ([P, I, P, P], None),
            "or_set_material": ([P, D, D, F], None),
            "or_ke": ([D, D, D, P], None),
            "or_set_fixed": ([P, P], None),
            "or_set_rho": ([P, P], None),
            "or_get_rho": ([P, P], None),
            "or_set_rhop": ([P, P], None),
            "or_get_rhop": ([P, P], None),
            "or_set_shell": ([P, P], None),
            "or_set_vec": ([P, I, I, P], None),
            "or_get_vec": ([P, I, I, P], None),
            "or_get_stencil": ([P, I, P], None),
            "or_set_stencil": ([P, I, P], None),
            "or_gs_relax": ([P, I, I, I], None),
            "or_update_residual": ([P, I], None),
            "or_apply": ([P, P, P], None),
            "or_restrict_residual": ([P, I], None),
            "or_prolongate_correction": ([P, I], None),
            "or_update_stencil": ([P], I),
            "or_galerkin4": ([P], I),
            "or_galerkin2": ([P, I], I),
            "or_build_coarsest": ([P], I),
            "or_solve_coarsest": ([P], None),
            "or_vcycle": ([P, I, I, I], D),
            "or_relative_residual": ([P], D),
            "or_compliance": ([P], D),
            "or_sensitivity": ([P, P], None),
            "or_filter": ([P, F, P, P], None),
            "or_oc_update": ([P, F, F, F, F, P, P], F),
            "or_set_load_nodes": ([P, P, P], None),
            "or_force_project": ([P], None),
            "or_modified_pm": ([P, I, P], D),
            "or_mma_create": ([I, I, P, P], P),
            "or_mma_destroy": ([P], None),
            "or_mma_update": ([P, P, P, P], I),
            "or_mma_get_x": ([P, P], None),
            "or_mma_set_x": ([P, P], None),
            "or_mma_get_bounds": ([P, P], None),
            "or_mma_stats": ([P, P, P, P, P], None),
            "or_mma_last_error": ([], ctypes.c_char_p),
            "or_bitsat_rank": ([P, I, ctypes.c_long], I),
            "or_voxelize_mesh": ([P, I, P, I, I, P, I, P, P], I),
            "or_voxelize_triangles": ([P, I, P, P, I, P], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ke(E=1e6, nu=0.4, h=1.0):
    out = np.zeros(576)
    lib().or_ke(h, E, nu, _p(out))
    return out.reshape(24, 24)


def bitsat_rank(words, i):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    return lib().or_bitsat_rank(_p(words), words.size, i)


def voxelize_mesh(coords, faces, prefer_reso):
    coords = np.ascontiguousarray(coords, dtype=np.float32).reshape(-1)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    R = 1
    while R <= prefer_reso:
        R <<= 1
    nw = (R ** 3 + 31) // 32
    bits = np.zeros(nw, dtype=np.uint32)
    reso = np.zeros(3, dtype=np.int32)
    box = np.zeros(6, dtype=np.float32)
    rc = lib().or_voxelize_mesh(_p(coords), coords.size // 3, _p(faces), faces.size // 3, prefer_reso,
                                _p(bits), nw, _p(reso), _p(box))
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())
    return bits, int(reso[0]), box.reshape(2, 3)


def voxelize_triangles(tris, bmin, unit, R):
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
    bmin = np.ascontiguousarray(bmin, dtype=np.float32)
    unit = np.ascontiguousarray(unit, dtype=np.float32)
    bits = np.zeros((R ** 3 + 31) // 32, dtype=np.uint32)
    lib().or_voxelize_triangles(_p(tris), tris.size // 9, _p(bmin), _p(unit), R, _p(bits))
    return bits


class Hierarchy:
    """Oracle hierarchy built from LSB-first solid bits of a cube of side R."""

    def __init__(self, R, bits):
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        self.h = lib().or_create(R, _p(bits))
        if not self.h:
            raise RuntimeError(lib().or_last_error().decode())
        self.R = R
        self.nlevels = lib().or_num_levels(self.h)
        self.levels = []
        for i in range(self.nlevels):
            info = np.zeros(11, dtype=np.int32)
            lib().or_level_info(self.h, i, _p(info))
            d = dict(layer=int(info[0]), scale=int(info[1]), o=info[2:5].copy(), e=info[5:8].copy(), n=info[8:11].copy())
            d["nn"] = int(np.prod(d["n"]))
            d["ne"] = int(np.prod(d["e"]))
            eact = np.zeros(d["ne"], np.uint8)
            vact = np.zeros(d["nn"], np.uint8)
            lib().or_get_masks(self.h, i, _p(eact), _p(vact))
            d["eact"] = eact.astype(bool)
            d["vact"] = vact.astype(bool)
            self.levels.append(d)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    @property
    def nlayers(self):
        return lib().or_num_layers(self.h)

    def layer_active(self, layer):
        return lib().or_layer_active(self.h, layer)

    def set_material(self, E, nu, p):
        lib().or_set_material(self.h, E, nu, p)

    def set_fixed(self, vfix):
        a = np.ascontiguousarray(vfix, dtype=np.uint8)
        lib().or_set_fixed(self.h, _p(a))

    def set_rho(self, rho):
        a = np.ascontiguousarray(rho, dtype=np.float32)
        lib().or_set_rho(self.h, _p(a))

    def get_rho(self):
        a = np.zeros(self.levels[0]["ne"], np.float32)
        lib().or_get_rho(self.h, _p(a))
        return a

    def set_rhop(self, rhop):
        a = np.ascontiguousarray(rhop, dtype=np.float32)
        lib().or_set_rhop(self.h, _p(a))

    def get_rhop(self):
        a = np.zeros(self.levels[0]["ne"], np.float32)
        lib().or_get_rhop(self.h, _p(a))
        return a

    def set_shell(self, s):
        a = np.ascontiguousarray(s, dtype=np.uint8)
        lib().or_set_shell(self.h, _p(a))

    def set_vec(self, li, which, x):
        a = np.ascontiguousarray(x, dtype=np.float64).reshape(3, -1)
        assert a.shape[1] == self.levels[li]["nn"]
        lib().or_set_vec(self.h, li, which, _p(a))

    def get_vec(self, li, which):
        a = np.zeros((3, self.levels[li]["nn"]))
        lib().or_get_vec(self.h, li, which, _p(a))
        return a

    def get_stencil(self, li):
        a = np.zeros((27, 9, self.levels[li]["nn"]))
        lib().or_get_stencil(self.h, li, _p(a))
        return a

    def set_stencil(self, li, st):
        a = np.ascontiguousarray(st, dtype=np.float64)
        lib().or_set_stencil(self.h, li, _p(a))

    def gs_relax(self, li, n=1, reverse=False):
        lib().or_gs_relax(self.h, li, n, int(reverse))

    def update_residual(self, li):
        lib().or_update_residual(self.h, li)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().or_apply(self.h, _p(x), _p(y))
        return y

    def restrict_residual(self, li):
        lib().or_restrict_residual(self.h, li)

    def prolongate_correction(self, li):
        lib().or_prolongate_correction(self.h, li)

    def update_stencil(self):
        if lib().or_update_stencil(self.h) != 0:
            raise RuntimeError(lib().or_last_error().decode())

    def galerkin4(self):
        lib().or_galerkin4(self.h)

    def galerkin2(self, li):
        lib().or_galerkin2(self.h, li)

    def build_coarsest(self):
        if lib().or_build_coarsest(self.h) != 0:
            raise RuntimeError(lib().or_last_error().decode())

    def solve_coarsest(self):
        lib().or_solve_coarsest(self.h)

    def v_cycle(self, pre=1, post=1, symmetric=False):
        return lib().or_vcycle(self.h, pre, post, int(symmetric))

    def relative_residual(self):
        return lib().or_relative_residual(self.h)

    def compliance(self):
        return lib().or_compliance(self.h)

    def sensitivity(self):
        g = np.zeros(self.levels[0]["ne"], np.float32)
        lib().or_sensitivity(self.h, _p(g))
        return g

    def filter(self, r, g):
        g = np.ascontiguousarray(g, dtype=np.float32)
        out = np.zeros_like(g)
        lib().or_filter(self.h, r, _p(g), _p(out))
        return out

    def set_load_nodes(self, load, normals=None):
        a = np.ascontiguousarray(load, dtype=np.uint8)
        if normals is None:
            lib().or_set_load_nodes(self.h, _p(a), None)
        else:
            nb = np.ascontiguousarray(normals, dtype=np.float64)
            lib().or_set_load_nodes(self.h, _p(a), _p(nb))

    def force_project(self):
        lib().or_force_project(self.h)

    def modified_pm(self, max_itn=500):
        """(worst compliance, V-cycles) from the force in F."""
        st = np.zeros(1, np.int32)
        c = lib().or_modified_pm(self.h, max_itn, _p(st))
        if np.isnan(c):
            raise RuntimeError(lib().or_last_error().decode())
        return c, int(st[0])

    def oc_update(self, Vgoal, step, damp, rhomin, g):
        g = np.ascontiguousarray(g, dtype=np.float32)
        it = np.zeros(1, np.int32)
        thr = lib().or_oc_update(self.h, Vgoal, step, damp, rhomin, _p(g), _p(it))
        return thr, int(it[0])


class Mma:
    """The reference's MMA optimiser (optimization/mma_t.h:159-266): n design
    variables in [lo, hi], m constraints g(x) <= 0."""

    def __init__(self, n, m, lo, hi):
        lo = np.broadcast_to(np.asarray(lo, np.float32), (n,)).copy()
        hi = np.broadcast_to(np.asarray(hi, np.float32), (n,)).copy()
        self.n, self.m = n, m
        self.h = lib().or_mma_create(n, m, _p(lo), _p(hi))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mma_destroy(self.h)
            self.h = None

    def x(self):
        a = np.zeros(self.n, np.float32)
        lib().or_mma_get_x(self.h, _p(a))
        return a

    def set_x(self, x):
        a = np.ascontiguousarray(x, dtype=np.float32)
        lib().or_mma_set_x(self.h, _p(a))

    def bounds(self):
        a = np.zeros((4, self.n), np.float32)
        lib().or_mma_get_bounds(self.h, _p(a))
        return a

    def stats(self):
        lam = np.zeros(self.m, np.float32)
        z = np.zeros(1, np.float32)
        it = np.zeros(1, np.int32)
        ov = np.zeros(1, np.int32)
        lib().or_mma_stats(self.h, _p(lam), _p(z), _p(it), _p(ov))
        return dict(lambda_=lam, z=float(z[0]), sub_iters=int(it[0]), overflowed=bool(ov[0]))

    def update(self, df, dg, g):
        """mma_t::update (mma_t.cpp:18-51); dg is [m, n]. True once converged."""
        df = np.ascontiguousarray(df, dtype=np.float32)
        dg = np.ascontiguousarray(dg, dtype=np.float32).reshape(self.m, self.n)
        g = np.ascontiguousarray(g, dtype=np.float32)
        r = lib().or_mma_update(self.h, _p(df), _p(dg), _p(g))
        if r < 0:
            raise RuntimeError(lib().or_mma_last_error().decode())
        return r == 1

"""SIMP loop around the GPU solve, with the reference's driver names
(optimization/optimization.h, utils/snippet.h, utils/test_utils.cpp)."""
import math
import time
from dataclasses import dataclass

from ._lib import RT_SOLVE_MIXED


@dataclass
class Parameter:
    """struct Parameter (optimization.h:11-22); defaults of config_parser.cpp:4-46.
    ``filter_radius`` is an int in the reference (optimization.h:17)."""
    damp_ratio: float = 0.5
    design_step: float = 0.03
    power_penalty: float = 3.0
    volume_ratio: float = 0.3
    volume_decrease: float = 0.02
    filter_radius: int = 2
    min_rho: float = 1e-3
    gridreso: int = 200
    youngs_modulu: float = 1e6
    poisson_ratio: float = 0.4


class converge_criteria:
    """SIMP stop test (snippet.h:156-213): relative change of the sum of the
    last two values against the two before, and of each constraint; stop once
    the test has passed more than ``max_counter`` times net."""

    def __init__(self, n_constrain, max_counter, stop_thres):
        self._n = n_constrain
        self._max = max_counter
        self._thres = stop_thres
        self._stop = 0
        self._itn = 0
        self._old = [math.nan] * 20
        self._oldc = [[math.nan] * 20 for _ in range(n_constrain)]

    def update(self, value, constrain):
        k = self._itn % 20
        self._old[k] = value
        for i in range(self._n):
            self._oldc[i][k] = constrain[i]

        def at(a, d):
            return a[(self._itn - d) % 20]

        def fdiv(a, b):   # IEEE division as in the C++ original
            if b != 0 or math.isnan(b):
                return a / b if not math.isnan(b) else math.nan
            return math.nan if (a == 0 or math.isnan(a)) else math.copysign(math.inf, a)

        fch = fdiv(abs(at(self._old, 0) + at(self._old, 1) - at(self._old, 2) - at(self._old, 3)),
                   at(self._old, 2) + at(self._old, 3))
        converged = fch < self._thres
        for i in range(self._n):
            a = self._oldc[i]
            gch = fdiv(abs(at(a, 0) + at(a, 1) - at(a, 2) - at(a, 3)), at(a, 2) + at(a, 3))
            if gch > 1e-3:
                converged = False
                break
        if converged:
            self._stop += 1
        else:
            self._stop = max(self._stop - 1, 0)
        self._itn += 1
        return self._stop > self._max


def solveFEM(hg, tol=1e-6, maxit=200, mode=RT_SOLVE_MIXED, pre=1, post=1):
    """Solve K(rho) u = f to ||r||/||f|| <= tol with MG-PCG. The reference
    iterates V-cycles (solveFEM, optimization.cpp:85-91)."""
    return hg.solve(tol=tol, maxit=maxit, mode=mode, pre=pre, post=post)


def computeSensitivity(hg, filter_radius):
    """computeSensitivity (optimization.cu:109-133): element sensitivity then
    the sensitivity filter (Grid.cu:2411-2437)."""
    hg.compute_sensitivity()
    if filter_radius > 0:
        hg.filter_sensitivity(filter_radius)


def updateDensities(hg, vgoal, params):
    """updateDensities (optimization.cu:164-226): OC bisection; returns the multiplier."""
    thr, _ = hg.update_densities(vgoal, params.design_step, params.damp_ratio, params.min_rho)
    return thr


def ordinary_topopt(hg, params, n_iter=100, tol=1e-6, mode=RT_SOLVE_MIXED, pre=1, post=1, init_rho=None,
                    log=None, stop=True):
    """The SIMP loop of TestSuit::testOrdinaryTopopt (test_utils.cpp:589-660):
    Vgoal schedule, update_stencil, solve, compliance, stop test, sensitivity,
    filter, OC update. The solve is MG-PCG to ``tol`` instead of V-cycles to 1e-2.
    Returns per-iteration records."""
    if init_rho is not None:
        hg.init_rho(init_rho)
    vgoal = params.volume_ratio
    stop_check = converge_criteria(1, 5, 1e-3)
    recs = []
    for itn in range(1, n_iter + 1):
        t0 = time.perf_counter()
        vgoal *= (1 - params.volume_decrease)
        vc = vgoal - params.volume_ratio
        if vgoal < params.volume_ratio:
            vgoal = params.volume_ratio
        hg.update_stencil()
        st = solveFEM(hg, tol=tol, mode=mode, pre=pre, post=post)
        c = hg.compliance()
        if math.isnan(c) or abs(c) < 1e-11:
            raise FloatingPointError(f"compliance {c} at iteration {itn}")
        done = stop_check.update(c, [vc]) and vgoal <= params.volume_ratio
        rec = dict(iter=itn, compliance=c, vgoal=vgoal, pcg_iters=st.iterations, rel_res=st.rel_res)
        if not (stop and done):
            computeSensitivity(hg, params.filter_radius)
            thr = updateDensities(hg, vgoal, params)
            rec["multiplier"] = thr
        hg.synchronize()
        rec["seconds"] = time.perf_counter() - t0
        recs.append(rec)
        if log:
            log(rec)
        if stop and done:
            break
    return recs


class MMA:
    """MMA::mma_t (optimization/mma_t.h:159-266) on the GPU: n FP32 design
    variables in [lower, upper], m <= 8 constraints g(x) <= 0."""

    def __init__(self, n, m, lower, upper):
        import ctypes
        import numpy as np
        from ._lib import check, lib, ptr
        lo = np.ascontiguousarray(np.broadcast_to(np.asarray(lower, np.float32), (n,)))
        hi = np.ascontiguousarray(np.broadcast_to(np.asarray(upper, np.float32), (n,)))
        h = ctypes.c_void_p()
        check(lib().rt_mma_create(n, m, ptr(lo), ptr(hi), ctypes.byref(h)))
        self.handle, self.n, self.m = h, n, m

    def close(self):
        from ._lib import lib
        if getattr(self, "handle", None):
            lib().rt_mma_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def x(self):
        import numpy as np
        from ._lib import check, lib, ptr
        a = np.zeros(self.n, np.float32)
        check(lib().rt_mma_get_x(self.handle, ptr(a)))
        return a

    def update(self, df, dg, g):
        """mma_t::update (mma_t.cpp:18-51); dg is [m, n]. True once converged."""
        import ctypes
        import numpy as np
        from ._lib import check, lib, ptr
        df = np.ascontiguousarray(df, dtype=np.float32)
        dg = np.ascontiguousarray(dg, dtype=np.float32).reshape(self.m, self.n)
        g = np.ascontiguousarray(g, dtype=np.float32)
        c = ctypes.c_int()
        check(lib().rt_mma_update(self.handle, ptr(df), ptr(dg), ptr(g), ctypes.byref(c)))
        return bool(c.value)

    def stats(self):
        import ctypes
        import numpy as np
        from ._lib import check, lib, ptr
        lam = np.zeros(self.m, np.float32)
        z = ctypes.c_float()
        it = ctypes.c_int()
        ov = ctypes.c_int()
        check(lib().rt_mma_stats(self.handle, ptr(lam), ctypes.byref(z), ctypes.byref(it), ctypes.byref(ov)))
        return dict(lambda_=lam, z=z.value, sub_iters=it.value, overflowed=bool(ov.value))


def worst_case_topopt(hg, params, n_iter=30, max_pm=500, seed=0, init_rho=None, log=None, stop=True,
                      method="pm", lanczos=None):
    """The worst-case SIMP loop of optimization() (optimization.cpp:583-673):
    Vgoal schedule, update_stencil, modified power method for the worst load
    and its compliance, stop test (converge_criteria(1, 2, 5e-3)), sensitivity
    of the worst compliance, filter and OC update. Load nodes and normals must
    be set (HierarchyGrid.set_load_nodes). Returns per-iteration records.

    method="lanczos" finds the worst load by block Lanczos instead
    (HierarchyGrid.block_lanczos, keyword arguments in `lanczos`), which
    converges to the same eigenpair with exact solves."""
    if method not in ("pm", "lanczos"):
        raise ValueError("method must be 'pm' or 'lanczos'")
    hg.init_rho(params.volume_ratio if init_rho is None else init_rho)
    vgoal = params.volume_ratio
    stop_check = converge_criteria(1, 2, 5e-3)
    recs = []
    for itn in range(1, n_iter + 1):
        t0 = time.perf_counter()
        vgoal *= (1 - params.volume_decrease)
        vc = vgoal - params.volume_ratio
        if vgoal < params.volume_ratio:
            vgoal = params.volume_ratio
        hg.update_stencil()
        if method == "pm":
            hg.rand_force(seed + itn)
            c, nv = hg.modified_pm(max_pm)
            extra = dict(vcycles=nv)
        else:
            c, extra = hg.block_lanczos(seed=seed + itn, **(lanczos or {}))
        if math.isnan(c):
            raise FloatingPointError(f"worst compliance NaN at iteration {itn}")
        done = stop_check.update(c, [vc]) and vgoal <= params.volume_ratio + 1e-3
        rec = dict(iter=itn, worst_compliance=c, vgoal=vgoal, **extra)
        if not (stop and done):
            computeSensitivity(hg, params.filter_radius)
            rec["multiplier"] = updateDensities(hg, vgoal, params)
        hg.synchronize()
        rec["seconds"] = time.perf_counter() - t0
        recs.append(rec)
        if log:
            log(rec)
        if stop and done:
            break
    return recs

"""Host-side mirror of the reference's grid interface (grid/Grid.h).

``HierarchyGrid`` and ``Grid`` keep the reference's method names and argument
meaning; every call goes through the C ABI of librobtop_amd.so to the HIP
kernels. Arrays exchanged with the host are numpy arrays in the level's box
layout (x fastest): nodes ``(nz, ny, nx)``-ordered flat vectors, 3-vectors as
``[3, nnodes]``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr, as_c, lib


class Grid:
    """One stored level of the hierarchy (reference class grid::Grid, Grid.h:172-532)."""

    def __init__(self, hier, index):
        self._h = hier
        self.index = index
        info = np.zeros(11, np.int32)
        check(lib().rt_level_info(hier.handle, index, ptr(info)))
        self._layer = int(info[0])
        self.scale = int(info[1])
        self.origin = info[2:5].copy()
        self.elements = info[5:8].copy()
        self.nodes = info[8:11].copy()

    # geometry ---------------------------------------------------------------
    @property
    def n_nodes_box(self):
        return int(np.prod(self.nodes))

    @property
    def n_elements_box(self):
        return int(np.prod(self.elements))

    def masks(self):
        e = np.zeros(self.n_elements_box, np.uint8)
        v = np.zeros(self.n_nodes_box, np.uint8)
        check(lib().rt_level_masks(self._h.handle, self.index, ptr(e), ptr(v)))
        return e.astype(bool), v.astype(bool)

    def surface_nodes(self):
        """Surface-node mask of the node box (mark_surface_nodes_kernel, Grid.cu:2005-2039)."""
        v = np.zeros(self.n_nodes_box, np.uint8)
        check(lib().rt_surface_nodes_box(self._h.handle, self.index, ptr(v)))
        return v.astype(bool)

    def applyK(self, u):
        """Grid::applyK (Grid.cu:2519-2532) on the finest level: K(rho) u, support
        neighbours as zero, support nodes return u."""
        if self.index != 0:
            raise ValueError("applyK is defined on the finest level")
        a = as_c(u, np.float64).reshape(3, self.n_nodes_box)
        f = np.zeros((3, self.n_nodes_box))
        check(lib().rt_apply_k(self._h.handle, ptr(a), ptr(f)))
        return f

    def is_finest(self):
        return self.index == 0

    def is_coarsest(self):
        return self.index == self._h.n_levels - 1

    # per-level vectors (V-cycle precision) --------------------------------------
    def _set(self, which, v):
        a = as_c(v, np.float64).reshape(3, self.n_nodes_box)
        check(lib().rt_set_level_vec(self._h.handle, self.index, which, ptr(a)))

    def _get(self, which):
        a = np.zeros((3, self.n_nodes_box))
        check(lib().rt_get_level_vec(self._h.handle, self.index, which, ptr(a)))
        return a

    def getDisplacement(self):
        return self._get(_lib.RT_VEC_U)

    def setDisplacement(self, u):
        self._set(_lib.RT_VEC_U, u)

    def getForce(self):
        return self._get(_lib.RT_VEC_F)

    def setForce(self, f):
        self._set(_lib.RT_VEC_F, f)

    def getResidual(self):
        return self._get(_lib.RT_VEC_R)

    def setResidual(self, r):
        self._set(_lib.RT_VEC_R, r)

    def getStencil(self):
        st = np.zeros((27, 9, self.n_nodes_box))
        check(lib().rt_get_stencil_box(self._h.handle, self.index, ptr(st)))
        return st

    # operators (Grid.h:357-437) ------------------------------------------------
    def gs_relax(self, n_times=1, reverse=False):
        check(lib().rt_gs_relax(self._h.handle, self.index, n_times, int(reverse)))

    def update_residual(self):
        check(lib().rt_update_residual(self._h.handle, self.index))

    def restrict_residual(self):
        check(lib().rt_restrict_residual(self._h.handle, self.index))

    def prolongate_correction(self):
        check(lib().rt_prolongate_correction(self._h.handle, self.index))

    def reset_displacement(self):
        check(lib().rt_reset_displacement(self._h.handle, self.index))

    def solve_fem_host(self):
        if not self.is_coarsest():
            raise ValueError("solve_fem_host runs on the coarsest level")
        check(lib().rt_solve_coarsest(self._h.handle))

    def relative_residual(self):
        if self.index != 0:
            raise ValueError("relative_residual is reported for the finest level")
        r = ctypes.c_double()
        check(lib().rt_relative_residual(self._h.handle, ctypes.byref(r)))
        return r.value

    # finest-level SIMP quantities ------------------------------------------------
    def compliance(self):
        return self._h.compliance()

    def filterSensitivity(self, radius):
        self._h.filter_sensitivity(radius)


class HierarchyGrid:
    """Reference class grid::HierarchyGrid (Grid.h:534-692) over one GPU."""

    def __init__(self, handle):
        self.handle = handle
        self.n_levels = check(lib().rt_num_levels(handle))
        self._grids = [Grid(self, i) for i in range(self.n_levels)]

    # construction ------------------------------------------------------------------
    @classmethod
    def genFromMesh(cls, solid_bits, reso):
        """HierarchyGrid::genFromMesh(solid_bits, reso) (Grid.cpp:455-712):
        LSB-first solid bits of a cube of side ``reso``."""
        R = int(reso if np.isscalar(reso) else reso[0])
        bits = as_c(solid_bits, np.uint32)
        h = ctypes.c_void_p()
        check(lib().rt_create_from_voxels(R, ptr(bits), bits.size, ctypes.byref(h)))
        return cls(h)

    from_voxels = genFromMesh

    def close(self):
        if self.handle:
            lib().rt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __getitem__(self, i):
        return self._grids[i]

    def __len__(self):
        return self.n_levels

    def n_grid(self):
        return self.n_levels

    @property
    def n_layers(self):
        return check(lib().rt_num_layers(self.handle))

    def layer_active_elements(self, layer):
        return check(lib().rt_layer_active_elements(self.handle, layer))

    def num_active(self, level=0):
        ne, nn = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().rt_num_active(self.handle, level, ctypes.byref(ne), ctypes.byref(nn)))
        return ne.value, nn.value

    # material / boundary ---------------------------------------------------------------
    def set_material(self, youngs=1e6, poisson=0.4, power_penalty=3.0):
        check(lib().rt_set_material(self.handle, youngs, poisson, power_penalty))

    def template_matrix(self):
        ke = np.zeros(576)
        check(lib().rt_get_template_matrix(self.handle, ptr(ke)))
        return ke.reshape(24, 24)

    def set_support(self, fixed_box):
        a = as_c(fixed_box, np.uint8).reshape(-1)
        check(lib().rt_set_support_box(self.handle, ptr(a)))

    def set_force(self, f_box):
        a = as_c(f_box, np.float64).reshape(3, -1)
        check(lib().rt_set_force_box(self.handle, ptr(a)))

    def get_force(self):
        a = np.zeros((3, self[0].n_nodes_box))
        check(lib().rt_get_force_box(self.handle, ptr(a)))
        return a

    def set_shell(self, shell_box):
        a = as_c(shell_box, np.uint8).reshape(-1)
        check(lib().rt_set_shell_box(self.handle, ptr(a)))

    # densities ---------------------------------------------------------------------------
    def init_rho(self, rho):
        check(lib().rt_init_density(self.handle, float(rho)))

    def set_density(self, rho_box):
        a = as_c(rho_box, np.float32).reshape(-1)
        check(lib().rt_set_density_box(self.handle, ptr(a)))

    def get_density(self):
        a = np.zeros(self[0].n_elements_box, np.float32)
        check(lib().rt_get_density_box(self.handle, ptr(a)))
        return a

    def get_rhop(self):
        a = np.zeros(self[0].n_elements_box, np.float32)
        check(lib().rt_get_rhop_box(self.handle, ptr(a)))
        return a

    # multigrid ------------------------------------------------------------------------------
    def update_stencil(self):
        check(lib().rt_update_stencil(self.handle))

    def set_precision(self, bits):
        check(lib().rt_set_vcycle_precision(self.handle, bits))

    def v_cycle(self, pre_relax=1, post_relax=1):
        r = ctypes.c_double()
        check(lib().rt_vcycle(self.handle, pre_relax, post_relax, ctypes.byref(r)))
        return r.value

    def set_island_deflation(self, theta=0.5, min_size=8):
        """Deflate the rigid-body modes of floating stiff clusters (rho >= theta,
        >= min_size elements, touching no support) in the preconditioner;
        theta <= 0 turns it off. Clusters are rebuilt at every update_stencil."""
        check(lib().rt_set_island_deflation(self.handle, float(theta), int(min_size)))

    def num_islands(self):
        n = ctypes.c_int()
        check(lib().rt_num_islands(self.handle, ctypes.byref(n)))
        return n.value

    def set_warm_start(self, m):
        """Start solves from the energy-optimal combination of the last m solutions (0: off)."""
        check(lib().rt_set_warm_start(self.handle, m))

    def set_learned_warm_start(self, net):
        """Start solves from the energy-optimal combination of the plain start
        and the network's prediction (a robtop_amd.learned.WarmStartNet, or
        None to turn it off). The network must stay alive while it is set."""
        self._net = net
        check(lib().rt_set_learned_warm_start(self.handle, net.handle if net is not None else None))

    def learned_warm_start_info(self):
        """{taken, energy of the plain start, energy of the learned start,
        coefficient of the prediction} of the last solve."""
        a = np.zeros(4, np.float64)
        check(lib().rt_learned_warm_start_info(self.handle, ptr(a)))
        return {"taken": bool(a[0]), "e_plain": float(a[1]), "e_learned": float(a[2]), "coef": float(a[3])}

    def cg_ritz_values(self, k=8):
        """(k smallest, k largest) Ritz values of the preconditioned operator
        from the last reliable-update solve's CG coefficients."""
        lo, hi = np.zeros(k), np.zeros(k)
        n = ctypes.c_int()
        check(lib().rt_cg_ritz_values(self.handle, k, ptr(lo), ptr(hi), ctypes.byref(n)))
        return lo[:min(k, n.value)], hi[:min(k, n.value)]

    def cg_coefficients(self, nmax=4096):
        """(alpha, beta) of the last reliable-update solve's CG recurrence."""
        a, b = np.zeros(nmax), np.zeros(nmax)
        n = ctypes.c_int()
        check(lib().rt_cg_coefficients(self.handle, nmax, ptr(a), ptr(b), ctypes.byref(n)))
        m = min(nmax, n.value)
        return a[:m], b[:m]

    def autotune(self, reps=10):
        """Time each level's launch-shape alternatives and keep the fastest
        (same operator; coarse kernels may sum in another order); returns the choices."""
        a = np.zeros(6, np.float64)
        check(lib().rt_autotune(self.handle, reps, ptr(a)))
        return {"fine_zc": int(a[0]), "fine_res_apply_ms": float(a[1]), "fine_smoother": int(a[2]),
                "fine_sweep_ms": float(a[3]), "coarse_sliced_gs_mask": int(a[4]),
                "coarse_sliced_res_mask": int(a[5])}

    def set_fine_smoother(self, kind):
        """Finest-level sweep: 0 = four colour-pair launches (default), 1 = two tiled launches."""
        check(lib().rt_set_fine_smoother(self.handle, kind))

    def set_coarse_stencil_bits(self, bits):
        """Coarse-level stencil storage of FP32 cycles: 32 (FP32) or 16 (bfloat16)."""
        check(lib().rt_set_coarse_stencil_bits(self.handle, bits))

    def set_preconditioner(self, coarse_sweep_mult=1, gamma=1):
        """Cycle shape of the MG preconditioner used by solve()."""
        check(lib().rt_set_preconditioner(self.handle, coarse_sweep_mult, gamma))

    def set_graphs(self, on=True):
        """Replay the preconditioner cycle as a HIP graph (default) or launch eagerly."""
        check(lib().rt_set_graphs(self.handle, int(bool(on))))

    def solve(self, tol=1e-6, maxit=200, mode=_lib.RT_SOLVE_MIXED, pre=1, post=1):
        st = _lib.SolveStats()
        check(lib().rt_solve(self.handle, tol, maxit, mode, pre, post, ctypes.byref(st)))
        return st

    def set_displacement(self, u_box):
        a = as_c(u_box, np.float64).reshape(3, -1)
        check(lib().rt_set_displacement_box(self.handle, ptr(a)))

    def write_displacement(self, path):
        """HierarchyGrid::writeDisplacement (Grid.cpp:886-897): the finest
        level's displacement as interleaved FP64 xyz in the reference's node order."""
        from .formats import write_vectors
        g0 = self[0]
        write_vectors(path, self.get_displacement(), g0.masks()[1], g0.nodes)

    def read_force(self, path):
        """Grid::readForce (Grid.cpp:1679-1700): a force file in the reference's
        node order and interleaved layout becomes the finest level's force."""
        from .formats import read_vectors
        g0 = self[0]
        self.set_force(read_vectors(path, g0.masks()[1], g0.nodes))

    def get_displacement(self):
        a = np.zeros((3, self[0].n_nodes_box))
        check(lib().rt_get_displacement_box(self.handle, ptr(a)))
        return a

    # SIMP step pieces ------------------------------------------------------------------------
    def compliance(self):
        c = ctypes.c_double()
        check(lib().rt_compliance(self.handle, ctypes.byref(c)))
        return c.value

    def compute_sensitivity(self):
        check(lib().rt_compute_sensitivity(self.handle))

    def filter_sensitivity(self, radius):
        check(lib().rt_filter_sensitivity(self.handle, float(radius)))

    def get_sensitivity(self):
        a = np.zeros(self[0].n_elements_box, np.float32)
        check(lib().rt_get_sensitivity_box(self.handle, ptr(a)))
        return a

    def set_sensitivity(self, g_box):
        a = as_c(g_box, np.float32).reshape(-1)
        check(lib().rt_set_sensitivity_box(self.handle, ptr(a)))

    def update_densities(self, vgoal, step, damp, rho_min):
        thr = ctypes.c_float()
        ntr = ctypes.c_int()
        check(lib().rt_update_densities(self.handle, vgoal, step, damp, rho_min, ctypes.byref(thr), ctypes.byref(ntr)))
        return thr.value, ntr.value

    def volume(self):
        v = ctypes.c_double()
        check(lib().rt_volume(self.handle, ctypes.byref(v)))
        return v.value

    # worst-case load search ------------------------------------------------------
    def set_load_nodes(self, load_box, normals_box=None):
        """setLoadNodes (projection.cpp:129-269): load-node flags on the finest
        node box; outward normals [3, n] select the constrained-direction mode."""
        a = as_c(load_box, np.uint8)
        if normals_box is None:
            check(lib().rt_set_load_nodes(self.handle, ptr(a), None))
        else:
            nb = as_c(normals_box, np.float64).reshape(3, -1)
            check(lib().rt_set_load_nodes(self.handle, ptr(a), ptr(nb)))
        n = ctypes.c_int()
        check(lib().rt_num_load_nodes(self.handle, ctypes.byref(n)))
        return n.value

    def force_project(self):
        """forceProject (projection.cu:165-231) of the finest force."""
        check(lib().rt_force_project(self.handle))

    def rand_force(self, seed):
        """randForce (Grid.cu:2242-2245), seeded."""
        check(lib().rt_rand_force(self.handle, ctypes.c_uint64(seed)))

    def modified_pm(self, max_itn=500):
        """modifiedPM (optimization.cpp:123-261) from the current force; returns
        (worst compliance, V-cycles)."""
        c = ctypes.c_double()
        nv = ctypes.c_int()
        check(lib().rt_modified_pm(self.handle, max_itn, ctypes.byref(c), ctypes.byref(nv)))
        return c.value, nv.value

    def solve_batch(self, f_boxes, tol=1e-6, maxit=1000, pre=2, post=2):
        """nb <= 8 load cases [nb, 3, nn] solved together from zero (no
        reference counterpart); returns (u [nb, 3, nn], iterations, rel_res)."""
        f = as_c(f_boxes, np.float64)
        nb = f.shape[0]
        u = np.zeros_like(f)
        it = np.zeros(nb, np.int32)
        rel = np.zeros(nb, np.float64)
        check(lib().rt_solve_batch(self.handle, nb, ptr(f), ptr(u), tol, maxit, pre, post, ptr(it), ptr(rel)))
        return u, it, rel

    def set_coarse_direct_dofs(self, max_dofs=6000):
        """the preconditioner solves the level above the coarsest exactly when
        it has at most max_dofs unknowns (0: the whole hierarchy)"""
        check(lib().rt_set_coarse_direct_dofs(self.handle, int(max_dofs)))

    def set_level_sweeps(self, level, mult):
        """sweep multiplier of coarse level `level` in the preconditioner (0: the
        multiplier of set_preconditioner)"""
        check(lib().rt_set_level_sweeps(self.handle, int(level), int(mult)))

    def preconditioner_levels(self):
        """levels the preconditioner's cycle visits (0 before update_stencil)"""
        n = ctypes.c_int()
        check(lib().rt_preconditioner_levels(self.handle, ctypes.byref(n)))
        return n.value

    def set_block_solves(self, batched=True):
        """block_lanczos solves a block's right-hand sides together (True) or one by one"""
        check(lib().rt_set_block_solves(self.handle, int(bool(batched))))

    def block_lanczos(self, k=8, max_blocks=30, tol=1e-6, solve_tol=1e-8, mode=None, pre=2, post=2, seed=0):
        """Top eigenpair of the compliance operator on the admissible loads by
        block Lanczos (no reference counterpart; same answer as modified_pm).
        Leaves the unit worst load in the force and its displacement in the
        solution; returns (worst compliance, {blocks, solves, pcg_iterations})."""
        if mode is None:
            mode = _lib.RT_SOLVE_RELIABLE
        c = ctypes.c_double()
        st = _lib.LanczosStats()
        check(lib().rt_block_lanczos(self.handle, k, max_blocks, tol, solve_tol, mode, pre, post, seed,
                                     ctypes.byref(c), ctypes.byref(st)))
        return c.value, {"blocks": st.blocks, "solves": st.solves, "pcg_iterations": st.pcg_iterations}

    def synchronize(self):
        check(lib().rt_synchronize(self.handle))

    def time_vcycle(self, bits=32, pre=1, post=1, n=10):
        ms = ctypes.c_double()
        check(lib().rt_time_vcycle(self.handle, bits, pre, post, n, ctypes.byref(ms)))
        return ms.value

    def time_fine_sweep(self, bits=32, n=10):
        ms = ctypes.c_double()
        check(lib().rt_time_fine_sweep(self.handle, bits, n, ctypes.byref(ms)))
        return ms.value

    def time_vcycle_batch(self, nb=8, pre=2, post=2, n=10):
        """ms of one batched preconditioner cycle for nb vectors"""
        ms = ctypes.c_double()
        check(lib().rt_time_vcycle_batch(self.handle, nb, pre, post, n, ctypes.byref(ms)))
        return ms.value

    def time_fine_residual(self, bits=32, n=10):
        ms = ctypes.c_double()
        check(lib().rt_time_fine_residual(self.handle, bits, n, ctypes.byref(ms)))
        return ms.value


def voxelize_mesh(coords, faces, prefer_reso):
    """voxelize_mesh (voxelizer/voxelize_main.cpp:344-404) on the GPU.
    Returns (LSB-first bits, R, box[2, 3])."""
    coords = as_c(coords, np.float32).reshape(-1)
    faces = as_c(faces, np.int32).reshape(-1)
    R = 1
    while R <= prefer_reso:
        R <<= 1
    nw = (R ** 3 + 31) // 32
    bits = np.zeros(nw, np.uint32)
    reso = np.zeros(3, np.int32)
    box = np.zeros(6, np.float32)
    check(lib().rt_voxelize_mesh(ptr(coords), coords.size // 3, ptr(faces), faces.size // 3, int(prefer_reso),
                                 ptr(bits), nw, ptr(reso), ptr(box)))
    return bits, int(reso[0]), box.reshape(2, 3)

"""Learned warm start: the convolutional network of librobtop_amd (rt_net_*).

The network maps (density, load) on the first coarse level's (x4) node box to
a displacement shape there; HierarchyGrid.set_learned_warm_start makes every
solve start from the energy-optimal combination of the plain start and the
prolonged prediction. Training samples are the (density, load, displacement)
triples of solved hierarchies (``add_sample`` after a solve); training runs
Adam with the library's own forward and backward kernels on the GPU.

Parameter layout (``params``/``set_params``): for each of the 11 convolutions
E0A E0B E1A E1B E2A E2B D1A D1B D0A D0B HEAD, weights [27][cout][cin] then
biases [cout], channel counts padded to multiples of 16.
"""
import ctypes

import numpy as np

from ._lib import check, lib, ptr

LAYERS = ("E0A", "E0B", "E1A", "E1B", "E2A", "E2B", "D1A", "D1B", "D0A", "D0B", "HEAD")
CIN = (16, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16)
COUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16)


def param_slices():
    """[(name, weight slice, bias slice, cin, cout)] of the parameter vector."""
    out, off = [], 0
    for name, ci, co in zip(LAYERS, CIN, COUT):
        w = slice(off, off + 27 * ci * co)
        off += 27 * ci * co
        b = slice(off, off + co)
        off += co
        out.append((name, w, b, ci, co))
    return out


class WarmStartNet:
    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        check(lib().rt_net_create(ctypes.c_uint64(seed), ctypes.byref(h)))
        self.handle = h
        n = ctypes.c_size_t()
        check(lib().rt_net_num_params(self.handle, ctypes.byref(n)))
        self.num_params = n.value

    def close(self):
        if getattr(self, "handle", None):
            lib().rt_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def params(self):
        w = np.zeros(self.num_params, np.float32)
        check(lib().rt_net_get_params(self.handle, ptr(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {w.size}")
        check(lib().rt_net_set_params(self.handle, ptr(w)))

    def forward(self, x):
        """x: (nz, ny, nx, 4) float32 -> (nz, ny, nx, 3) float32."""
        x = np.ascontiguousarray(x, np.float32)
        nz, ny, nx, c = x.shape
        assert c == 4
        dims = np.array([nx, ny, nz], np.int32)
        y = np.zeros((nz, ny, nx, 3), np.float32)
        check(lib().rt_net_forward(self.handle, ptr(dims), ptr(x), ptr(y)))
        return y

    def gradient(self, x, target):
        """(loss, gradient over the parameter vector) of the mean squared error."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        nz, ny, nx, _ = x.shape
        assert t.shape == (nz, ny, nx, 3)
        dims = np.array([nx, ny, nz], np.int32)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_double()
        check(lib().rt_net_gradient(self.handle, ptr(dims), ptr(x), ptr(t), ctypes.byref(loss), ptr(g)))
        return loss.value, g

    def add_sample(self, hg):
        """Store the solved hierarchy's (density, load, displacement) triple."""
        check(lib().rt_net_add_sample(self.handle, hg.handle))

    def evaluate(self, hg):
        """Loss of the prediction for the solved hierarchy's triple (not stored)."""
        loss = ctypes.c_double()
        check(lib().rt_net_evaluate(self.handle, hg.handle, ctypes.byref(loss)))
        return loss.value

    def predict_coarse(self, hg):
        """the x4 level's U of hg = the network's prediction from the density
        and that level's current F (FP64 level vectors)"""
        check(lib().rt_net_predict_coarse(self.handle, hg.handle))

    def time_forward(self, dims_xyz, reps=20):
        """Mean ms of one forward pass on an (nx, ny, nz) box."""
        dims = np.array(dims_xyz, np.int32)
        ms = ctypes.c_double()
        check(lib().rt_net_time_forward(self.handle, ptr(dims), reps, ctypes.byref(ms)))
        return ms.value

    def num_samples(self):
        n = ctypes.c_int()
        check(lib().rt_net_num_samples(self.handle, ctypes.byref(n)))
        return n.value

    def clear_samples(self):
        check(lib().rt_net_clear_samples(self.handle))

    def train(self, steps, batch=4, lr=1e-3, seed=0):
        """Adam steps on the stored samples; returns the last step's mean loss."""
        loss = ctypes.c_double()
        check(lib().rt_net_train(self.handle, steps, batch, ctypes.c_float(lr), ctypes.c_uint64(seed),
                                 ctypes.byref(loss)))
        return loss.value

    def adam_step(self, grad, lr=1e-3):
        """One Adam update with the given gradient (layout of params())."""
        g = np.ascontiguousarray(grad, np.float32)
        if g.size != self.num_params:
            raise ValueError("gradient size differs from the parameter count")
        check(lib().rt_net_adam_step(self.handle, ptr(g), ctypes.c_float(lr)))

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))


def forward_flops(dims_xyz, padded=True):
    """Multiply-add flops (2 per MAC) of one forward pass on an (nx, ny, nz)
    box: every 3x3x3 convolution at its resolution, with the channel counts
    the matrix cores process (padded) or the real ones."""
    d0 = list(dims_xyz)
    d1 = [(n + 1) // 2 for n in d0]
    d2 = [(n + 1) // 2 for n in d1]
    P = [int(np.prod(d)) for d in (d0, d1, d2)]
    lev = (0, 0, 1, 1, 2, 2, 1, 1, 0, 0, 0)
    cin = CIN if padded else (4,) + CIN[1:]
    cout = COUT if padded else COUT[:-1] + (3,)
    return float(sum(2 * 27 * ci * co * P[l] for ci, co, l in zip(cin, cout, lev)))


def cantilever_hierarchy(ex, ey, ez, seed):
    """The cantilever of the benchmark: an ex x ey x ez box of solid voxels,
    x = 0 face fixed, -z load spread over the x = max face, seeded
    smoothed-random density; returns (hierarchy, force)."""
    from .grid import HierarchyGrid
    from .workloads import box_bits, cube_side, cantilever_bc, smoothed_random_density
    R = cube_side(ex, ey, ez)
    hg = HierarchyGrid.genFromMesh(box_bits(R, ex, ey, ez), R)
    hg.set_material(1e6, 0.4, 3.0)
    fixed, f, _ = cantilever_bc(hg)
    hg.set_support(fixed)
    hg.set_force(f)
    eact, _ = hg[0].masks()
    hg.set_density(smoothed_random_density(eact, tuple(hg[0].elements), seed))
    return hg, f


def collect_simp_samples(net, size=(256, 128, 128), triples=2000, iters_per_run=20, first_seed=1, tol=1e-6,
                         log=print):
    """Runs SIMP loops of the cantilever (filter r = 2, OC with volume 0.4,
    step 0.03, damping 0.5) from seeded density fields, adding every solved
    step's (density, load, displacement) triple to the network's samples until
    it holds `triples`. Returns the seeds used."""
    from . import RT_SOLVE_RELIABLE
    from .workloads import smoothed_random_density
    hg, f = cantilever_hierarchy(*size, first_seed)
    eact, _ = hg[0].masks()
    seed = first_seed
    try:
        while net.num_samples() < triples:
            hg.set_density(smoothed_random_density(eact, tuple(hg[0].elements), seed))
            hg.set_displacement(np.zeros_like(f))
            for _ in range(iters_per_run):
                if net.num_samples() >= triples:
                    break
                hg.update_stencil()
                hg.solve(tol=tol, maxit=500, mode=RT_SOLVE_RELIABLE, pre=2, post=2)
                net.add_sample(hg)
                hg.compute_sensitivity()
                hg.filter_sensitivity(2)
                hg.update_densities(0.4, 0.03, 0.5, 1e-3)
            log(f"learned: seed {seed} done, {net.num_samples()} samples")
            seed += 1
    finally:
        hg.close()
    return list(range(first_seed, seed))


def heldout_loss(net, size=(256, 128, 128), seed=10_000, steps=5, tol=1e-6):
    """Mean prediction loss over the first `steps` SIMP steps of a seed not used in training."""
    from . import RT_SOLVE_RELIABLE
    hg, f = cantilever_hierarchy(*size, seed)
    losses = []
    try:
        for _ in range(steps):
            hg.update_stencil()
            hg.solve(tol=tol, maxit=500, mode=RT_SOLVE_RELIABLE, pre=2, post=2)
            losses.append(net.evaluate(hg))
            hg.compute_sensitivity()
            hg.filter_sensitivity(2)
            hg.update_densities(0.4, 0.03, 0.5, 1e-3)
    finally:
        hg.close()
    return float(np.mean(losses))

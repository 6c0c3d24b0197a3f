"""Invalid and degenerate inputs are reported as status codes with a message
(the reference prints and exits, config_parser.cpp:49-50, optimization.cpp:80-81).
The validation below happens before any device work, so it runs on the CPU;
the solver-state checks need a hierarchy and run on the GPU."""
import ctypes

import numpy as np
import pytest

import robtop_amd as rt
from robtop_amd import _lib
from robtop_amd.workloads import box_bits, cube_side


def _create(R, bits, nwords=None):
    L = _lib.lib()
    h = ctypes.c_void_p()
    bits = np.ascontiguousarray(bits, np.uint32)
    code = L.rt_create_from_voxels(R, _lib.ptr(bits), bits.size if nwords is None else nwords, ctypes.byref(h))
    return code, h, L.rt_last_error().decode()


def test_empty_solid_is_rejected():
    code, h, msg = _create(32, np.zeros(32 ** 3 // 32, np.uint32))
    assert code == _lib.RT_EARG and "no solid element" in msg and not h.value


@pytest.mark.parametrize("R", [0, 2, 24, 48])
def test_cube_side_must_be_a_power_of_two(R):
    code, h, msg = _create(R, np.ones(max(1, R ** 3 // 32), np.uint32))
    assert code == _lib.RT_EARG and "power of two" in msg


def test_short_bit_array_is_rejected():
    code, h, msg = _create(32, np.ones(10, np.uint32))
    assert code == _lib.RT_EARG and "too short" in msg


def test_level_past_the_32_bit_load_range_is_rejected():
    # a 1024 x 1024 x 512 slab: 8 * 515 * 515 * 259 = 549.6M slots, whose FP32
    # vectors pass the 2 GiB a buffer load can address
    R = 1024
    bits = np.zeros(R ** 3 // 32, np.uint32)
    bits[: R * R * 512 // 32] = 0xFFFFFFFF          # z < 512, lowest-bit-first x order
    code, h, msg = _create(R, bits)
    assert code == _lib.RT_EARG and not h.value
    assert "level 0" in msg and "32-bit vectors" in msg and "2 GiB" in msg


def test_network_null_handle_is_reported():
    L = _lib.lib()
    dims = np.array([4, 4, 4], np.int32)
    x = np.zeros(4 * 64, np.float32)
    y = np.zeros(3 * 64, np.float32)
    assert L.rt_net_forward(None, _lib.ptr(dims), _lib.ptr(x), _lib.ptr(y)) == _lib.RT_EARG
    assert "null rt_net handle" in L.rt_last_error().decode()


@pytest.mark.gpu
def test_single_voxel_solid_has_too_few_layers():
    a = np.zeros((32, 32, 32), bool)
    a[3, 4, 5] = True
    from robtop_amd.workloads import _pack_lsb
    code, h, msg = _create(32, _pack_lsb(a.reshape(-1)))
    assert code == _lib.RT_EARG and "layers" in msg


@pytest.mark.gpu
def test_solver_state_errors():
    print("[edge] 32x16x16 box", flush=True)
    R = cube_side(32, 16, 16)
    hg = rt.HierarchyGrid.genFromMesh(box_bits(R, 32, 16, 16), R)
    L = _lib.lib()
    st = _lib.SolveStats()
    code = L.rt_solve(hg.handle, 1e-6, 100, rt.RT_SOLVE_RELIABLE, 2, 2, ctypes.byref(st))
    assert code == _lib.RT_ESTATE and "update_stencil" in L.rt_last_error().decode()
    info = np.zeros(6)
    assert L.rt_autotune(hg.handle, 3, _lib.ptr(info)) == _lib.RT_ESTATE
    assert L.rt_set_fine_smoother(hg.handle, 7) == _lib.RT_EARG
    assert L.rt_set_coarse_stencil_bits(hg.handle, 12) == _lib.RT_EARG
    from robtop_amd.learned import WarmStartNet
    net = WarmStartNet(0)
    dims = np.array([0, 4, 4], np.int32)
    assert L.rt_net_forward(net.handle, _lib.ptr(dims), _lib.ptr(np.zeros(4, np.float32)),
                            _lib.ptr(np.zeros(3, np.float32))) == _lib.RT_EARG
    assert L.rt_net_train(net.handle, 1, 1, ctypes.c_float(1e-3), ctypes.c_uint64(0), None) == _lib.RT_ESTATE
    net.close()
    hg.close()

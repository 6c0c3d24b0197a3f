"""SIMP trajectory parity on the reference's small configuration: cantilever
box 64 x 32 x 32, uniform density 0.5, p = 3, E = 1e6, nu = 0.4, x = 0 face
fixed, unit -z load spread over the x = max face; each iteration solves to
relative residual 1e-10, then sensitivity, filter (r = 2) and OC (volume 0.4,
step 0.03, damp 0.5, rho_min 1e-3) (test_utils.cpp:589-660).

The HIP path (FP64 MG-PCG) and the oracle (the reference's V(2,2) cycles
repeated to 1e-10) agree on compliance, sensitivity and densities at every
iteration, well inside the 1e-5 relative bound for displacements, compliance
and densities.
"""
import numpy as np
import pytest

import oracle
import robtop_amd as rt
from problems import rel_err

pytestmark = pytest.mark.gpu


def test_cantilever64_simp_iterations_match_oracle():
    from robtop_amd.workloads import box_bits, cantilever_bc
    ex, ey, ez, R = 64, 32, 32, 64
    print("[simp] cantilever 64x32x32", flush=True)
    bits = box_bits(R, ex, ey, ez)
    g = rt.HierarchyGrid.genFromMesh(bits, R)
    o = oracle.Hierarchy(R, bits)
    g.set_material(1e6, 0.4, 3.0)
    o.set_material(1e6, 0.4, 3.0)
    fixed, f, _ = cantilever_bc(g)
    g.set_support(fixed)
    o.set_fixed(fixed.astype(np.uint8))
    g.set_force(f)
    o.set_vec(0, 1, f)
    g.init_rho(0.5)
    o.set_rho(np.where(o.levels[0]["eact"], 0.5, 0.0).astype(np.float32))
    o.set_rhop(g.get_rhop())
    nn = o.levels[0]["nn"]
    g.set_displacement(np.zeros((3, nn)))
    o.set_vec(0, 0, np.zeros((3, nn)))
    for it in range(3):
        g.update_stencil()
        st = g.solve(tol=1e-10, maxit=300, mode=rt.RT_SOLVE_FP64, pre=2, post=2)
        assert st.rel_res < 5e-10, st.rel_res
        o.update_stencil()
        for _ in range(300):
            if o.v_cycle(2, 2) < 1e-10:
                break
        assert rel_err(g.get_displacement(), o.get_vec(0, 0)) < 1e-7, it
        cg, co = g.compliance(), o.compliance()
        assert abs(cg - co) / co < 1e-8, (it, cg, co)
        g.compute_sensitivity()
        sg = g.get_sensitivity()
        so = o.sensitivity()
        assert rel_err(sg, so) < 1e-5, it
        g.filter_sensitivity(2)
        fo = o.filter(2.0, so)
        assert rel_err(g.get_sensitivity(), fo) < 1e-5, it
        g.update_densities(0.4, 0.03, 0.5, 1e-3)
        o.oc_update(0.4, 0.03, 0.5, 1e-3, fo)
        assert rel_err(g.get_density(), o.get_rho()) < 1e-5, it
        o.set_rhop(g.get_rhop())
    g.close()

"""Benchmark: seconds per SIMP iteration at 511x256x256 (cantilever, R = 512)
to relative residual 1e-6, and the V-cycle's algorithmic HBM bandwidth.

One step = update_stencil (Galerkin) + MG-PCG solve to 1e-6 (FP64 Krylov,
FP32 V-cycle) + compliance + sensitivity + filter (r = 2) + OC update, all on
the GPU through the C ABI. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 511,256,256]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", default="511,256,256")
    ap.add_argument("--workload", default="cantilever", choices=["cantilever", "torus"],
                    help="cantilever: the headline s/SIMP-iteration run; torus: the worst-case loop on a "
                         "voxelized torus with block Lanczos (its own JSON line, not the headline)")
    ap.add_argument("--torus-n", type=int, default=384, help="torus voxels per side")
    ap.add_argument("--no-eigen-check", action="store_true",
                    help="torus workload: skip the FP64 recomputation of the worst compliance (for profiling)")
    ap.add_argument("--lanczos-sequential", action="store_true",
                    help="torus: solve a Lanczos block's right-hand sides one after another instead of together")
    ap.add_argument("--lanczos-solve-tol", type=float, default=1e-6,
                    help="relative residual of the PCG solves inside block Lanczos (torus workload)")
    ap.add_argument("--pre", type=int, default=2)
    ap.add_argument("--post", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=5000,
                    help="PCG iteration cap per solve; a timed step that reaches it without the tolerance makes "
                         "the run invalid (exit status 3)")
    ap.add_argument("--coarse-mult", type=int, default=None,
                    help="sweep multiplier on coarse levels (preconditioner); default 2 for the cantilever, 1 for the torus")
    ap.add_argument("--gamma", type=int, default=1, help="preconditioner cycle index: 1 = V, 2 = W")
    ap.add_argument("--level-sweeps", default=None,
                    help="per-level sweep multipliers of the preconditioner, 'level:mult,...' (overrides --coarse-mult "
                         "on those levels); default '1:1' for the cantilever (the x4 level smooths once, the smaller "
                         "levels --coarse-mult times), none for the torus")
    ap.add_argument("--coarse-direct-dofs", type=int, default=6000,
                    help="the preconditioner solves the level above the coarsest exactly (dense FP64 inverse) "
                         "when it has at most this many unknowns (0: cycle over every level)")
    ap.add_argument("--fine-smoother", default="pairs", choices=["tiled", "pairs"],
                    help="finest-level sweep: four colour-pair launches (default) or two tiled launches")
    ap.add_argument("--warm-start", type=int, default=0,
                    help="start each solve from the energy-optimal combination of the last M solutions (0: the "
                         "previous solution, energy-scaled)")
    ap.add_argument("--islands", type=float, default=0.02, metavar="THETA",
                    help="deflate the rigid-body modes of floating clusters of elements with rho >= THETA in the "
                         "preconditioner (0: off)")
    ap.add_argument("--island-min", type=int, default=1, help="smallest floating cluster deflated (elements)")
    ap.add_argument("--coarse-bf16", action="store_true",
                    help="store the coarse-level stencils of the FP32 preconditioner in bfloat16")
    ap.add_argument("--mode", default="reliable", choices=["mixed", "reliable", "fp64"],
                    help="MG-PCG variant: FP64 Krylov vectors (mixed), FP32 vectors with FP64 residual "
                         "replacement (reliable), or all FP64")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the preconditioner kernels one by one instead of replaying a HIP graph "
                         "(use under rocprofv3: its kernel tracing has aborted the queue on graph replays)")
    ap.add_argument("--learned", type=int, default=0, metavar="TRIPLES",
                    help="learned warm start: train the network on TRIPLES (density, load, displacement) triples "
                         "of 256x128x128 SIMP runs, then start every solve from its prediction (0: off)")
    ap.add_argument("--learned-steps", type=int, default=2000, help="Adam steps (batch 4) of that training")
    ap.add_argument("--net-forward", default=None, metavar="NX,NY,NZ",
                    help="only time the learned warm start's network forward pass on this box (e.g. 129,65,65, "
                         "the x4 level of 511x256x256) and print its MFMA roofline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing the alternatives before the run")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--vcycles", type=int, default=20, help="graphless V-cycles timed for the roofline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-only", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="kernels only, for rocprofv3 counter runs: two fine sweeps, one V-cycle, then the "
                         "calibration copy (256 MiB device-to-device)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn the counter_collection.csv files of a --profile run under rocprofv3 --pmc "
                         "FETCH_SIZE / --pmc WRITE_SIZE into HBM bytes per fine-sweep launch (JSON)")
    return ap.parse_args()


def algorithmic_vcycle_bytes(levels, pre, post, coarsest_dofs, P=4, mult=1, gamma=1, Ps=4, level_mult=None):
    """Compulsory bytes of one preconditioner cycle: each array element an
    operation must read or write, counted once per operation; u/f/r in FP32 (P),
    stencils Ps bytes each (FP32 or bfloat16), 1 flag byte per node, rho^p 4 B per element. Coarse
    levels smooth `mult` times pre/post sweeps and level i > 0 is visited
    gamma^(i-1) times (the coarsest once per visit of its parent)."""
    n0, e0 = levels[0]["nodes"], levels[0]["elements"]
    sweep0 = n0 * (9 * P + 1) + 4 * e0
    res0 = sweep0
    total = (pre + post) * sweep0 + res0
    nlev = len(levels)
    visits = 1
    for i in range(1, nlev):
        nf, nc = levels[i - 1]["nodes"], levels[i]["nodes"]
        per_visit = 3 * P * nf + 3 * P * nc            # restrict (read r_f, write f_c)
        per_visit += 3 * P * nc + nf * (6 * P + 1)     # prolong (read u_c, read+write u_f, flag)
        if i < nlev - 1:
            nl = levels[i]["nodes"]
            per = nl * (243 * Ps + 9 * P + 1)
            m = (level_mult or {}).get(i, 0) or mult
            per_visit += m * (pre + post) * per + per   # sweeps + residual
        else:
            per_visit += coarsest_dofs ** 2 * 8
        total += visits * per_visit
        if i < nlev - 2:
            visits *= gamma
    return total


def parse_level_sweeps(spec, n_levels):
    """'1:1,2:4' -> {1: 1, 2: 4}"""
    out = {}
    for item in filter(None, spec.split(",")):
        lv, m = (int(t) for t in item.split(":"))
        if not 1 <= lv < n_levels:
            raise SystemExit(f"--level-sweeps: level {lv} is not a coarse level (1..{n_levels - 1})")
        out[lv] = m
    return out


def build_problem(ex, ey, ez, seed):
    import robtop_amd as rt
    from robtop_amd.workloads import box_bits, cube_side, cantilever_bc, smoothed_random_density
    R = cube_side(ex, ey, ez)
    bits = box_bits(R, ex, ey, ez)
    hg = rt.HierarchyGrid.genFromMesh(bits, R)
    hg.set_material(1e6, 0.4, 3.0)
    fixed, f, nload = cantilever_bc(hg)
    hg.set_support(fixed)
    hg.set_force(f)
    g0 = hg[0]
    eact, _ = g0.masks()
    rho = smoothed_random_density(eact, tuple(g0.elements), seed)
    hg.set_density(rho)
    return hg, bits, R, fixed, f, rho


def cpu_baseline(bits, R, fixed, f, rho, pre, post, n_vcycles, n_applies, reps=2):
    """Oracle (FP64 C++/OpenMP restatement of the reference kernels) on the host
    cores: the finest level's share of `reps` V(pre, post) cycles (pre + post
    Gauss-Seidel sweeps and one residual each) at the full size; the SIMP
    iteration is extrapolated from the GPU run's V-cycle and operator counts.
    Coarse levels, Galerkin and OC are left out, so this is a lower bound of
    the CPU path."""
    import oracle
    o = oracle.Hierarchy(R, bits)
    o.set_material(1e6, 0.4, 3.0)
    o.set_fixed(fixed)
    o.set_rho(rho)
    nn = o.levels[0]["nn"]
    o.set_vec(0, 1, f)
    o.set_vec(0, 0, np.zeros((3, nn)))
    t_sweep = t_res = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        o.gs_relax(0, pre + post)
        t1 = time.perf_counter()
        o.update_residual(0)
        t2 = time.perf_counter()
        t_sweep += (t1 - t0) / (pre + post)
        t_res += t2 - t1
    t_sweep /= reps
    t_res /= reps
    est = n_vcycles * ((pre + post) * t_sweep + t_res) + n_applies * t_res
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=est, unit="s/iter", cores=cores, kind="port",
                sample=(f"oracle FP64 finest-level work of {reps} V({pre},{post}) cycles at full size on {cores} "
                        f"threads ({reps * (pre + post)} GS sweeps, {t_sweep:.2f} s each; {reps} residuals, "
                        f"{t_res:.2f} s each), scaled by the GPU run's {n_vcycles:.1f} V-cycles and "
                        f"{n_applies:.1f} operator applies per SIMP iteration; coarse levels, Galerkin and OC "
                        f"excluded (lower bound)"))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


CALIB_BYTES = 256 * 1024 * 1024


def profile_run(a):
    """Kernel-only run for counter collection: two fine-level sweeps (with the
    chosen smoother), one V-cycle, then a plain copy of CALIB_BYTES between two
    device buffers as the calibration dispatch."""
    import torch
    ex, ey, ez = (int(v) for v in a.size.split(","))
    hg, *_ = build_problem(ex, ey, ez, a.seed)
    hg.set_graphs(False)
    hg.set_fine_smoother(1 if a.fine_smoother == "tiled" else 0)
    hg.update_stencil()
    log(f"profile {ex}x{ey}x{ez}: fine sweeps")
    hg.time_fine_sweep(32, 2)
    log("V-cycle")
    hg.time_vcycle(32, a.pre, a.post, 1)
    hg.synchronize()
    log("calibration copy")
    x = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    log("done")


def pmc_summary(fetch_csv, write_csv, size):
    """HBM bytes per fine-sweep launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch). Each counter is scaled by what it reads for the calibration
    copy: the largest copyBuffer dispatch, which moves CALIB_BYTES each way."""
    import csv
    import re
    from collections import defaultdict

    def per_kernel(path, counter):
        agg, calib = defaultdict(list), 0.0
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = re.sub(r"\(.*", "", r["Kernel_Name"]).strip()
            v = float(r["Counter_Value"]) * 1024.0
            if "copyBuffer" in name:
                calib = max(calib, v)
            agg[name].append(v)
        return agg, calib

    fa, fc = per_kernel(fetch_csv, "FETCH_SIZE")
    wa, wc = per_kernel(write_csv, "WRITE_SIZE")
    kf, kw = fc / CALIB_BYTES, wc / CALIB_BYTES
    kname = "k_fine_gs_tile" if any("k_fine_gs_tile" in k for k in fa) else "k_fine_gs_pair"
    launches = 2 if kname == "k_fine_gs_tile" else 4
    gs = [k for k in fa if kname in k]
    fetch = sum(sum(fa[k]) for k in gs) / sum(len(fa[k]) for k in gs) / kf
    write = sum(sum(wa[k]) for k in gs) / sum(len(wa[k]) for k in gs) / kw
    ex, ey, ez = size
    nodes, elems = (ex + 1) * (ey + 1) * (ez + 1), ex * ey * ez
    alg = (nodes * (9 * 4 + 1) + 4 * elems) / launches
    return {"kernel": kname, "size": list(size), "bytes_per_launch": fetch + write,
            "fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
            "calibration": {"copy_bytes": CALIB_BYTES, "fetch_factor": kf, "write_factor": kw},
            "algorithmic_bytes_per_launch": alg, "traffic_over_algorithmic": (fetch + write) / alg}


def torus_run(a):
    """Worst-case SIMP iterations on an analytic torus (R = 0.35, r = 0.12 of the
    box edge) voxelized at n^3: supports on the bottom 5 % of its height, loads
    along the outward normals of the other surface nodes, the worst load found
    by block Lanczos (k = 8, FP32-preconditioned reliable-update solves). After
    the timed iterations the worst compliance is recomputed with FP64 solves at
    the same densities and the relative difference reported."""
    import robtop_amd as rt
    from robtop_amd.workloads import torus_bits, torus_worst_case_bc
    n = a.torus_n
    log(f"torus {n}^3")
    bits, R, solid = torus_bits(n)
    hg = rt.HierarchyGrid.genFromMesh(bits, R)
    hg.set_material(1e6, 0.4, 3.0)
    fixed, load, normals = torus_worst_case_bc(hg, n)
    hg.set_support(fixed)
    nl = hg.set_load_nodes(load.astype(np.uint8), normals)
    hg.set_precision(32)
    hg.set_preconditioner(a.coarse_mult, a.gamma)
    hg.set_coarse_direct_dofs(a.coarse_direct_dofs)
    log(f"hierarchy: {hg.n_levels} levels, {nl} load nodes")
    hg.set_graphs(not a.no_graphs)
    tuned = None
    eact, _ = hg[0].masks()
    hg.set_density(np.where(eact, 0.3, 0.0).astype(np.float32))
    hg.update_stencil()
    if not a.no_autotune:
        tuned = hg.autotune(reps=5)
        log(f"autotune: {tuned}")
    t_v = hg.time_vcycle(32, a.pre, a.post, a.vcycles)
    t_vb = hg.time_vcycle_batch(8, a.pre, a.post, a.vcycles)
    log(f"V-cycle {t_v:.3f} ms; batched cycle for 8 vectors {t_vb:.3f} ms ({t_vb / 8:.3f} ms per vector)")
    hg.set_block_solves(not a.lanczos_sequential)
    params = rt.Parameter(volume_ratio=0.3, volume_decrease=0.05)
    lz = dict(k=8, max_blocks=20, tol=1e-6, solve_tol=a.lanczos_solve_tol, pre=a.pre, post=a.post)
    recs = rt.worst_case_topopt(hg, params, n_iter=a.warmup + a.steps, stop=False, method="lanczos", lanczos=lz,
                                log=lambda r: log(f"iter {r['iter']}: c {r['worst_compliance']:.6e}, "
                                                  f"{r['blocks']} blocks, {r['solves']} solves, "
                                                  f"{r['pcg_iterations']} PCG iterations, {r['seconds']:.2f} s"))
    timed = recs[a.warmup:]
    s_iter = float(np.mean([r["seconds"] for r in timed]))
    hg.update_stencil()
    eig = None
    if not a.no_eigen_check:
        c32, st32 = hg.block_lanczos(seed=99, **lz)
        lz64 = dict(lz, tol=1e-8, solve_tol=1e-11, mode=rt.RT_SOLVE_FP64)
        c64, st64 = hg.block_lanczos(seed=99, **lz64)
        eig = {"worst_compliance_fp32_solves": c32, "worst_compliance_fp64_solves": c64,
               "rel_diff": abs(c32 - c64) / abs(c64), "fp64_blocks": st64["blocks"], "fp64_solves": st64["solves"]}
    ne, nn = hg.num_active(0)
    out = {
        "metric": f"s/worst-case SIMP iter, torus {n}^3, block Lanczos k=8",
        "value": s_iter, "unit": "s/iter", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": s_iter * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"torus R=0.35 r=0.12 voxelized at {n}^3, supports on the bottom 5 %, normal loads on "
                               f"{nl} surface nodes, block Lanczos k=8 (tol 1e-6, solves to {a.lanczos_solve_tol:g}), "
                               f"filter r=2, OC",
                   "active_elements": ne, "active_nodes": nn, "levels": hg.n_levels,
                   "lanczos_per_iter": [{k: r[k] for k in ("blocks", "solves", "pcg_iterations")} for r in timed],
                   "autotune": tuned, "vcycle_ms": t_v, "batched_vcycle_ms_8_vectors": t_vb,
                   "block_solves": "one after another" if a.lanczos_sequential else "together (batched cycle)",
                   "node_box": [int(v) for v in hg[0].nodes]},
        "eigen_check": eig,
    }
    hg.close()
    print(json.dumps(out), flush=True)


def net_forward_run(a):
    """Network forward pass alone (random weights; the kernels' speed does not
    depend on the values): mean ms over --steps passes after --warmup, HIP events."""
    from robtop_amd.learned import WarmStartNet, forward_flops
    dims = tuple(int(v) for v in a.net_forward.split(","))
    net = WarmStartNet(a.seed)
    net.time_forward(dims, max(1, a.warmup))
    ms = net.time_forward(dims, max(1, a.steps))
    fl = forward_flops(dims)
    out = {"metric": "learned warm start network forward pass", "value": ms, "unit": "ms", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "higher_is_better": False, "dtype": "bf16 (fp32 accumulate)",
           "data": "synthetic", "config": {"workload": f"3-level conv encoder-decoder on a {dims} node box"},
           "roofline": {"bound": "mfma", "achieved": fl / (ms * 1e-3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": fl / (ms * 1e-3) / 1e12 / 2500.0, "flops_per_pass": fl}}
    net.close()
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.coarse_mult is None:
        a.coarse_mult = 1 if a.workload == "torus" else 2
    if a.level_sweeps is None:
        a.level_sweeps = "" if a.workload == "torus" else "1:1"
    if a.pmc_summary:
        print(json.dumps(pmc_summary(*a.pmc_summary, tuple(int(v) for v in a.size.split(",")))))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "robtop_amd benchmarks one GPU; --gpus must be 1"}))
        sys.exit(2)
    if a.profile:
        profile_run(a)
        return
    if a.workload == "torus":
        torus_run(a)
        return
    if a.net_forward:
        net_forward_run(a)
        return
    ex, ey, ez = (int(v) for v in a.size.split(","))
    import robtop_amd as rt
    log(f"building {ex}x{ey}x{ez}")
    hg, bits, R, fixed, f, rho = build_problem(ex, ey, ez, a.seed)
    log(f"hierarchy: {hg.n_levels} levels")
    hg.set_preconditioner(a.coarse_mult, a.gamma)
    hg.set_coarse_direct_dofs(a.coarse_direct_dofs)
    level_mults = parse_level_sweeps(a.level_sweeps, hg.n_levels)
    for lv, m in level_mults.items():
        hg.set_level_sweeps(lv, m)
    if a.coarse_bf16:
        hg.set_coarse_stencil_bits(16)
    hg.set_fine_smoother(1 if a.fine_smoother == "tiled" else 0)
    hg.set_warm_start(a.warm_start)
    if a.islands > 0:
        hg.set_island_deflation(a.islands, a.island_min)
    hg.set_graphs(not a.no_graphs)
    learned = None
    if a.learned > 0:
        from robtop_amd.learned import WarmStartNet, collect_simp_samples, heldout_loss
        net = WarmStartNet(a.seed)
        t_a = time.perf_counter()
        seeds = collect_simp_samples(net, triples=a.learned, log=log)
        t_b = time.perf_counter()
        train_loss = net.train(a.learned_steps, batch=4, lr=1e-3, seed=a.seed)
        t_c = time.perf_counter()
        learned = {"triples": net.num_samples(), "runs": len(seeds), "train_loss": train_loss,
                   "heldout_loss": heldout_loss(net), "sample_seconds": t_b - t_a, "train_seconds": t_c - t_b,
                   "adam_steps": a.learned_steps, "taken_steps": 0}
        from robtop_amd.learned import forward_flops
        dims = tuple(int(v) for v in hg[1].nodes)
        fwd_ms = net.time_forward(dims, 20)
        learned["forward"] = {"box": dims, "ms": fwd_ms, "bound": "mfma",
                              "achieved": forward_flops(dims) / (fwd_ms * 1e-3) / 1e12,
                              "achieved_real_channels": forward_flops(dims, padded=False) / (fwd_ms * 1e-3) / 1e12,
                              "peak": 2500.0, "unit": "TFLOP/s"}
        learned["forward"]["frac"] = learned["forward"]["achieved"] / 2500.0
        log(f"learned: {learned}")
        net.clear_samples()
        hg.set_learned_warm_start(net)
    tuned = None
    if not a.no_autotune:
        hg.update_stencil()
        tuned = hg.autotune(reps=5)
        log(f"autotune: {tuned}")
        a.fine_smoother = "tiled" if tuned["fine_smoother"] == 1 else "pairs"
    g0 = hg[0]
    ne0, nn0 = hg.num_active(0)
    levels = []
    for i in range(hg.n_levels):
        ne, nn = hg.num_active(i)
        levels.append(dict(nodes=nn, elements=ne))

    vgoal, step, damp, rmin, rfilt = 0.4, 0.03, 0.5, 1e-3, 2
    mode = {"mixed": rt.RT_SOLVE_MIXED, "reliable": rt.RT_SOLVE_RELIABLE, "fp64": rt.RT_SOLVE_FP64}[a.mode]

    phase = {"update_stencil": 0.0, "solve": 0.0, "sensitivity_filter_oc": 0.0}

    def simp_step(record=False):
        t0 = time.perf_counter()
        hg.update_stencil()
        hg.synchronize()
        t1 = time.perf_counter()
        st = hg.solve(tol=a.tol, maxit=a.maxit, mode=mode, pre=a.pre, post=a.post)
        c = hg.compliance()
        t2 = time.perf_counter()
        if learned is not None and record:
            learned["taken_steps"] += int(hg.learned_warm_start_info()["taken"])
        hg.compute_sensitivity()
        hg.filter_sensitivity(rfilt)
        hg.update_densities(vgoal, step, damp, rmin)
        hg.synchronize()
        t3 = time.perf_counter()
        if record:
            phase["update_stencil"] += t1 - t0
            phase["solve"] += t2 - t1
            phase["sensitivity_filter_oc"] += t3 - t2
        return st, c

    for w in range(a.warmup):
        st, _ = simp_step()
        log(f"warmup {w}: {st.iterations} PCG iterations, rel_res {st.rel_res:.2e}")
    hg.synchronize()
    iters, rels, comps, conv, secs = [], [], [], [], []
    t0 = time.perf_counter()
    for k in range(a.steps):
        ts = time.perf_counter()
        st, c = simp_step(record=True)
        secs.append(time.perf_counter() - ts)
        log(f"step {a.warmup + k}: {st.iterations} PCG iterations, rel_res {st.rel_res:.2e}, "
            f"{'converged' if st.converged else 'NOT CONVERGED'}, c {c:.6e}, {secs[-1]:.3f} s")
        iters.append(st.iterations)
        rels.append(st.rel_res)
        comps.append(c)
        conv.append(bool(st.converged))
    hg.synchronize()
    t1 = time.perf_counter()
    s_per_iter = (t1 - t0) / max(1, a.steps)
    n_conv = int(sum(conv))

    # V-cycle and dominant-kernel timing (HIP events on the handle's stream)
    hg.update_stencil()
    t_v = hg.time_vcycle(32, a.pre, a.post, a.vcycles)
    t_sweep = hg.time_fine_sweep(32, 10)
    # the preconditioner visits pc_levels levels and solves the last exactly
    pc_levels = hg.preconditioner_levels()
    B_V = algorithmic_vcycle_bytes(levels[:pc_levels], a.pre, a.post, 3 * levels[pc_levels - 1]["nodes"],
                                   mult=a.coarse_mult, gamma=a.gamma, Ps=2 if a.coarse_bf16 else 4,
                                   level_mult=level_mults)
    # Compulsory bytes of a sweep: u (read + write), f and the flag once per node,
    # rho^p once per element. The tiled smoother runs a sweep as two launches
    # (one z-parity class each, half the nodes), the colour-pair smoother as four
    # (one (y, z) parity class each); each launch is charged its share.
    n_launch = 2 if a.fine_smoother == "tiled" else 4
    kname = "k_fine_gs_tile" if a.fine_smoother == "tiled" else "k_fine_gs_pair"
    sweep_bytes = levels[0]["nodes"] * (9 * 4 + 1) + 4 * levels[0]["elements"]
    launch_bytes = sweep_bytes / n_launch
    launch_ms = t_sweep / n_launch
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic_fine_gs.json")
    if os.path.exists(tp):
        try:
            tj = json.load(open(tp))
            traffic = tj.get("bytes_per_launch") if tj.get("kernel") == kname else None
        except Exception:
            traffic = None

    out = {
        "metric": "s/SIMP iter at 511×256×256 to rel-res 1e-6; V-cycle HBM GB/s vs peak",
        "value": s_per_iter,
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": s_per_iter * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"cantilever {ex}x{ey}x{ez} voxels (R={R}), seeded smoothed-random density, "
                        f"MG-PCG {a.mode} V({a.pre},{a.post}) to {a.tol:g}, filter r=2, OC",
            "active_elements": ne0, "active_nodes": nn0, "dofs": 3 * nn0, "levels": hg.n_levels,
            "pcg_iters_per_step": float(np.mean(iters)), "rel_res": float(np.max(rels)),
            "converged_steps": n_conv, "maxit": a.maxit,
            "per_step": [{"simp_step": a.warmup + k, "pcg_iterations": int(iters[k]), "rel_res": float(rels[k]),
                          "converged": conv[k], "seconds": float(secs[k])} for k in range(a.steps)],
            "median_seconds_per_step": float(np.median(secs)),
            "compliance_last": comps[-1],
            "seconds_per_step": {k: v / max(1, a.steps) for k, v in phase.items()},
        },
        "vcycle": {"ms": t_v, "algorithmic_bytes": B_V, "GBps": B_V / (t_v * 1e-3) / 1e9,
                   "frac_of_peak": B_V / (t_v * 1e-3) / 1e9 / HBM_PEAK_GBS, "pre": a.pre, "post": a.post,
                   "coarse_mult": a.coarse_mult, "gamma": a.gamma, "levels_visited": pc_levels,
                   "level_sweeps": {str(k): v for k, v in level_mults.items()},
                   "exact_solve_dofs": 3 * levels[pc_levels - 1]["nodes"],
                   "coarse_stencil_bits": 16 if a.coarse_bf16 else 32},
        "roofline": {"kernel": f"{kname} (one of the {n_launch} launches of a finest-level Gauss-Seidel sweep)",
                     "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": launch_bytes, "launch_ms": launch_ms},
    }
    if tuned is not None:
        out["config"]["autotune"] = tuned
    if learned is not None:
        out["config"]["learned_warm_start"] = learned
    if not a.no_cpu:
        n_v = float(np.mean(iters)) + 1.0
        n_app = float(np.mean(iters)) + 2.0
        hg.close()
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(bits, R, fixed, f, rho, a.pre, a.post, n_v, n_app)
    out["valid"] = n_conv == a.steps
    if not out["valid"]:
        out["invalid_reason"] = (f"{a.steps - n_conv} of {a.steps} timed solves stopped at maxit={a.maxit} above "
                                 f"rel-res {a.tol:g}")
    print(json.dumps(out), flush=True)
    if not out["valid"]:
        sys.exit(3)


if __name__ == "__main__":
    main()

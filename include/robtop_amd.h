/*
 * robtop_amd — C ABI of the MI355X-native multigrid FEM / SIMP hot path.
 *
 * Every entry point returns 0 on success and a negative code on failure
 * (RT_E*); rt_last_error() gives the message of the calling thread's last
 * failure. Handles are not thread-safe; each owns one HIP stream and all of its
 * device memory. Host buffers are owned by the caller.
 *
 * Layouts at the boundary:
 *   - solid bits: uint32 words, LSB-first, bit id x + y*R + z*R^2 (what the
 *     reference holds after wordReverse_g, grid/Grid.cu:2662-2687);
 *   - "box" arrays: a level's dense node box (nx, ny, nz) or element box
 *     (nx-1, ny-1, nz-1), x fastest; rt_level_info gives origin and sizes;
 *   - "active" arrays: only the solid elements / existing nodes, in
 *     lexicographic (z, y, x) order — the reference's lexicographic ids before
 *     its colour permutation (eidmap/vidmap, grid/Grid.h:220-221);
 *   - 3-vectors: three consecutive blocks [x..., y..., z...] (SoA) in FP64.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * lavenklau/robtop).
 */
#ifndef ROBTOP_AMD_H
#define ROBTOP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RT_OK = 0,
    RT_EARG = -1,     /* invalid argument / size */
    RT_EHIP = -2,     /* HIP runtime failure */
    RT_ESTATE = -3,   /* call out of order (e.g. solve before update_stencil) */
    RT_ENUM = -4,     /* numerical failure (non-SPD coarsest system, breakdown) */
    RT_EINTERNAL = -5
};

/* solver modes for rt_solve */
enum {
    RT_SOLVE_FP64 = 0,      /* FP64 PCG, FP64 V-cycle preconditioner */
    RT_SOLVE_FP32 = 1,      /* FP32 PCG vectors (FP64 dots), FP32 V-cycle */
    RT_SOLVE_MIXED = 2,     /* FP64 PCG vectors and dots, FP32 V-cycle preconditioner */
    RT_SOLVE_FP32_IR = 3,   /* FP64 iterative refinement (true residual) around FP32 MG-PCG */
    RT_SOLVE_RELIABLE = 4   /* FP32 Krylov vectors and operator, FP64 solution, FP64 residual
                               replacement whenever the FP32 residual has dropped 10x */
};

/* vector selectors for the per-level accessors */
enum { RT_VEC_U = 0, RT_VEC_F = 1, RT_VEC_R = 2 };

typedef struct rt_hier rt_hier;

typedef struct rt_solve_stats {
    int iterations;         /* PCG iterations */
    double rel_res;         /* ||f - K u|| / ||f|| of the returned u, recomputed in FP64 */
    double rel_res_pcg;     /* the PCG recurrence residual at exit */
    double seconds;         /* wall time of the solve */
    int converged;          /* 1 if rel_res <= tol; 0 if maxit ran out first (the call still returns RT_OK) */
} rt_solve_stats;

const char* rt_last_error(void);
const char* rt_version(void);

/* --- hierarchy --------------------------------------------------------- */

/* Build the multigrid hierarchy from a solid voxel set of a cube of side R
 * (R a power of two). Replaces HierarchyGrid::genFromMesh(solid_bits, reso)
 * (grid/Grid.cpp:455-712): levels coarsen by OR-pooling until a level has at
 * most 400 solid elements (Grid.cpp:229), layer 1 is skipped (x4 first
 * coarsening, Grid.cpp:390-393, optimization.cpp:17). Mode is "with support"
 * (wsff/wscf, grid/Grid.h:146-151). */
int rt_create_from_voxels(int R, const uint32_t* solid_bits, size_t nwords, rt_hier** out);
int rt_destroy(rt_hier* h);

/* number of stored levels (layer 1 excluded) and of reference layers */
int rt_num_levels(const rt_hier* h);
int rt_num_layers(const rt_hier* h);
int rt_layer_active_elements(const rt_hier* h, int layer);
/* info[11] = {layer, scale, ox, oy, oz, ex, ey, ez, nx, ny, nz}; origin in the
 * level's absolute element coordinates */
int rt_level_info(const rt_hier* h, int level, int* info);
/* masks of the level's element box and node box (1 = solid / existing) */
int rt_level_masks(const rt_hier* h, int level, uint8_t* eact_box, uint8_t* vact_box);
int rt_num_active(const rt_hier* h, int level, size_t* n_elements, size_t* n_nodes);
/* Surface-node mask of a level in its node box (1 = existing node on the
 * surface). Replaces Grid::mark_surface_nodes_g (grid/Grid.cu:2005-2051),
 * whose flag gates supports and loads in computeProjectionMatrix
 * (grid/Grid.cpp:1291-1403). */
int rt_surface_nodes_box(const rt_hier* h, int level, uint8_t* surf_box);

/* --- material and boundary conditions ---------------------------------- */

/* Young's modulus, Poisson ratio, SIMP power. Replaces setParameters
 * (optimization/optimization.cpp:32-50) + initTemplateMatrix
 * (element/templateMatrix.cpp:31-124) + uploadTemplateMatrix
 * (optimization/optimization.cu:330-343). Element size 1 (Grid.cpp:457-461). */
int rt_set_material(rt_hier* h, double youngs, double poisson, float power_penalty);
/* element stiffness KE (576 doubles, row-major) as used on the device */
int rt_get_template_matrix(const rt_hier* h, double* ke576);

/* Support (fixed) node flags on the finest node box. Replaces the support
 * flag set in Grid::computeProjectionMatrix (grid/Grid.cpp:1291-1403). */
int rt_set_support_box(rt_hier* h, const uint8_t* fixed_box);
int rt_set_support_active(rt_hier* h, const uint8_t* fixed_active);
/* Nodal force (FP64, 3 x n) on the finest level; zeroed at supports.
 * Replaces setForceSupport (optimization/projection.cu:344-369). */
int rt_set_force_box(rt_hier* h, const double* f_box);
int rt_set_force_active(rt_hier* h, const double* f_active);
int rt_get_force_box(const rt_hier* h, double* f_box);
/* shell elements (density forced to 1 by the OC update, optimization.cu:159) */
int rt_set_shell_box(rt_hier* h, const uint8_t* shell_box);

/* --- densities ----------------------------------------------------------- */

/* Replaces initDensities / Grid::init_rho (optimization.cpp:687-690) and the
 * device density array rho_e (Grid.h:196). */
int rt_init_density(rt_hier* h, float rho);
int rt_set_density_box(rt_hier* h, const float* rho_box);
int rt_get_density_box(const rt_hier* h, float* rho_box);
int rt_set_density_active(rt_hier* h, const float* rho_active);
int rt_get_density_active(const rt_hier* h, float* rho_active);
/* rho^p as evaluated on the device (FP32 powf, grid/Grid.cu:834) */
int rt_get_rhop_box(const rt_hier* h, float* rhop_box);

/* --- multigrid ----------------------------------------------------------- */

/* Galerkin coarse operators for the current densities. Replaces
 * HierarchyGrid::update_stencil (grid/Grid.cpp:2108-2119): x4 restriction of
 * the penalised fine operator (Grid.cu:392-506), x2 restrictions
 * (Grid.cu:120-177) and the coarsest dense factorisation (Grid.cpp:1884-1939). */
int rt_update_stencil(rt_hier* h);
/* stencil [27][9][nodes] of a coarse level on its node box */
int rt_get_stencil_box(const rt_hier* h, int level, double* st_box);

/* Precision of the V-cycle used by rt_vcycle and the per-level operations:
 * 64 (the reference's FP64) or 32. */
int rt_set_vcycle_precision(rt_hier* h, int bits);

/* Per-level vectors (U, F, R) of the V-cycle, box layout, FP64 at the boundary. */
int rt_set_level_vec(rt_hier* h, int level, int which, const double* v_box);
int rt_get_level_vec(const rt_hier* h, int level, int which, double* v_box);

/* Grid::gs_relax (grid/Grid.cu:1068-1111); reverse=1 runs colours 7..0 and
 * components z,y,x (the symmetric counterpart, no reference equivalent). */
int rt_gs_relax(rt_hier* h, int level, int nsweeps, int reverse);
/* Grid::update_residual (grid/Grid.cu:1644-1678) */
int rt_update_residual(rt_hier* h, int level);
/* Grid::applyK (grid/Grid.cu:2519-2532): f = K(rho) u on the finest node box
 * (FP64, 3 x n), support neighbours taken as zero; support nodes return u. */
int rt_apply_k(rt_hier* h, const double* u_box, double* f_box);
/* Grid::restrict_residual (grid/Grid.cu:1796-1816): level = the coarse level */
int rt_restrict_residual(rt_hier* h, int level);
/* Grid::prolongate_correction (grid/Grid.cu:1898-1915): level = the fine level */
int rt_prolongate_correction(rt_hier* h, int level);
/* Grid::reset_displacement (grid/Grid.cu:1917-1922) */
int rt_reset_displacement(rt_hier* h, int level);
/* Grid::solve_fem_host (grid/Grid.cpp:1981-2043) on the coarsest level */
int rt_solve_coarsest(rt_hier* h);
/* Grid::relative_residual (grid/Grid.cu:1946-1951) of the finest level */
int rt_relative_residual(rt_hier* h, double* rel);

/* One V(pre, post) cycle on the finest displacement/force. Replaces
 * HierarchyGrid::v_cycle (grid/Grid.cpp:1068-1113); returns ||r||/||f||. */
int rt_vcycle(rt_hier* h, int pre, int post, double* rel_res);

/* --- solve and SIMP step ------------------------------------------------- */

/* Floating-cluster deflation of the preconditioner (no counterpart in the
 * reference): at every rt_update_stencil, the face-connected clusters of at
 * least min_size elements with rho >= theta that touch no support are found,
 * and the preconditioner adds the exact correction on their six rigid-body
 * modes, z += sum_i Z_i (Z_i^T K Z_i)^+ Z_i^T r, to the V-cycle. At most 4096
 * clusters. theta <= 0 turns it off (default). */
int rt_set_island_deflation(rt_hier* h, float theta, int min_size);
int rt_num_islands(const rt_hier* h, int* n);

/* Warm start of rt_solve from a subspace of previous solutions: with m >= 1
 * every solve starts from the combination of the last m returned solutions
 * that minimises the energy norm of the error under the current stiffness
 * (Galerkin projection, an m x m system), and each solve's result is stored.
 * m = 0 (default) starts from the current displacement scaled by the
 * energy-optimal factor. m <= 8; setting m clears the stored solutions. No
 * counterpart in the reference. */
int rt_set_warm_start(rt_hier* h, int m);

/* Implementation of the finest-level Gauss-Seidel sweep (same result bit for
 * bit): 0 (default) = four colour-pair launches updating in place; 1 = two
 * launches per sweep, each relaxing the four colours of one z-parity class
 * tile by tile from an old-value buffer into a second buffer. No counterpart
 * in the reference. */
int rt_set_fine_smoother(rt_hier* h, int kind);

/* Storage of the coarse-level stencils read by FP32 cycles: 32 (FP32, the
 * default) or 16 (bfloat16, rounded from the FP64 Galerkin stencils; halves
 * the coarse levels' stencil traffic). Used by the preconditioner and every
 * FP32 V-cycle; FP64 cycles and the coarsest factorisation keep FP64. No
 * counterpart in the reference (its stencils are FP64). */
int rt_set_coarse_stencil_bits(rt_hier* h, int bits);

/* Shape of the preconditioner cycle of rt_solve (and rt_time_vcycle): levels
 * below the finest smooth coarse_sweep_mult times pre/post sweeps, and each
 * coarse level is visited gamma times (1 = V-cycle, 2 = W-cycle). Default
 * (1, 1); no reference counterpart. rt_vcycle is always the reference's V. */
int rt_set_preconditioner(rt_hier* h, int coarse_sweep_mult, int gamma);
/* The preconditioner's cycle (rt_solve, rt_solve_batch, the worst-case
 * solves) stops one level above the coarsest and solves that level exactly
 * (FP64 dense inverse) when it has at most max_dofs unknowns (default 6000:
 * at 511 x 256 x 256 this is the 4131-unknown level of 16 x 8 x 8 elements).
 * 0 keeps the whole hierarchy; at most 30000 (a 7.2 GB inverse), else
 * RT_EARG. rt_vcycle always visits every level. */
int rt_set_coarse_direct_dofs(rt_hier* h, int max_dofs);
/* Sweep multiplier of one coarse level (1 .. levels-1) in the preconditioner's
 * cycle, overriding rt_set_preconditioner's coarse_sweep_mult there (0 puts
 * the level back on it). */
int rt_set_level_sweeps(rt_hier* h, int level, int mult);
/* number of levels the preconditioner's cycle visits (the last one solved
 * exactly); 0 before the first rt_update_stencil */
int rt_preconditioner_levels(const rt_hier* h, int* levels);

/* Replay the preconditioner cycle of rt_solve as a captured HIP graph (on,
 * the default) or launch its kernels one by one (off); results are identical. */
int rt_set_graphs(rt_hier* h, int on);

/* MG-preconditioned CG to ||r||/||f|| <= tol from the current displacement
 * (new; the reference iterates V-cycles: solveFEM, optimization.cpp:85-91,
 * and test_utils.cpp:633-637). Preconditioner: V(pre, post) with symmetric
 * (reversed) post-smoothing. */
int rt_solve(rt_hier* h, double tol, int maxit, int mode, int pre, int post, rt_solve_stats* stats);

int rt_set_displacement_box(rt_hier* h, const double* u_box);
int rt_get_displacement_box(const rt_hier* h, double* u_box);
int rt_get_displacement_active(const rt_hier* h, double* u_active);

/* Grid::compliance = F.U (grid/Grid.cpp:1976-1979) */
int rt_compliance(rt_hier* h, double* c);

/* computeSensitivity (optimization/optimization.cu:109-133) without the
 * filter: g_e = -p rho^(p-1) u_e^T KE u_e */
int rt_compute_sensitivity(rt_hier* h);
/* Grid::filterSensitivity (grid/Grid.cu:2411-2437) */
int rt_filter_sensitivity(rt_hier* h, double radius);
int rt_get_sensitivity_box(const rt_hier* h, float* g_box);
int rt_set_sensitivity_box(rt_hier* h, const float* g_box);

/* updateDensities (optimization/optimization.cu:164-226): OC bisection on the
 * multiplier; volume averaged over solid elements. Returns the multiplier and
 * the number of trial evaluations. */
int rt_update_densities(rt_hier* h, float vgoal, float step, float damp, float rho_min, float* multiplier,
                        int* trials);
/* mean density over solid elements */
int rt_volume(rt_hier* h, double* v);

/* --- worst-case load search --------------------------------------------- */

/* Load nodes for the worst-case search. Replaces setLoadNodes
 * (optimization/projection.cpp:129-269): load_box flags load nodes on the
 * finest node box (supports and absent nodes are dropped); normals_box (FP64,
 * 3 x n) gives outward normals for the constrained-direction modes wscf / nscf
 * (only the normal component of a load is admissible), or is NULL for free
 * loads (wsff / nsff). The no-support modes (nsff, nscf) are the ones of a
 * hierarchy without supports: their loads must also be balanced (no net force
 * or moment over the load nodes, projection.cpp:208-257, :322-335). */
int rt_set_load_nodes(rt_hier* h, const uint8_t* load_box, const double* normals_box);
int rt_num_load_nodes(const rt_hier* h, int* n);
/* forceProject (optimization/projection.cu:165-231): the finest force is
 * cleared off the load nodes and, with normals, reduced to its normal part;
 * without supports its rigid part over the load nodes is removed as well. */
int rt_force_project(rt_hier* h);
/* Grid::randForce (grid/Grid.cu:2242-2245): force components uniform in
 * [-1, 1) on existing nodes, from a counter-based generator seeded by `seed`
 * (the reference seeds cuRAND from the clock). */
int rt_rand_force(rt_hier* h, uint64_t seed);
/* modifiedPM (optimization/optimization.cpp:123-261), starting from the
 * current force: project and normalise, u = 0, then per step one V(1,1) cycle
 * (precision of rt_set_vcycle_precision) and f <- P u / |P u|, until the load
 * change |fs_new - fs_old| / |fs_old| is at most 1e-4 and the cycle's
 * relative residual at most 1e-2, or max_itn steps. Without supports (the
 * no-support modes) the coarsest levels are solved for the minimum-norm
 * solution and the rigid motion is removed from u after every cycle
 * (displacementProject, projection.cpp:374-394). The worst load and its
 * displacement stay in the force / displacement; returns the worst
 * compliance f.u and the number of V-cycles. */
int rt_modified_pm(rt_hier* h, int max_itn, double* worst_compliance, int* vcycles);

/* Worst-case load search by block Lanczos (no counterpart in the reference,
 * whose modifiedPM is rt_modified_pm): the top eigenpair of the compliance
 * operator C = P N^T K^-1 N P on the admissible loads of rt_set_load_nodes,
 * with block size k (<= 64) from a seeded random block, full
 * reorthogonalisation, and one rt_solve (mode, V(pre, post), relative residual
 * solve_tol, zero start) per operator application. Stops when the top Ritz
 * pair's residual is below tol * theta or after max_blocks blocks. On return
 * the force is the unit Ritz load, the displacement its solution, and
 * *worst_compliance = f . u. */
typedef struct rt_lanczos_stats {
    int blocks;             /* Lanczos blocks built */
    int solves;             /* operator applications (PCG solves), the final one included */
    int pcg_iterations;     /* No counterpart in the reference (it solves one load at a time): nb <= 8
 * load cases K u_b = f_b on the same hierarchy in one batched solve, from
 * zero. f_box and u_box hold nb finest node-box vectors of 3 nn values each
 * (the layout of rt_set_force_box); support and absent nodes of f are
 * ignored. Preconditioned CG per load case with FP64 vectors and dots, in
 * lockstep, preconditioned by one FP32 V(pre, post) cycle for all nb vectors
 * at once: the coarser levels read each stencil coefficient once per cycle
 * for all of them, and the cycle is replayed from a HIP graph. iterations[b]
 * and rel_res[b] (true |f_b - K u_b| / |f_b|) per load case. Needs the V
 * cycle (gamma 1) and FP32 coarse stencils. */
int rt_solve_batch(rt_hier* h, int nb, const double* f_box, double* u_box, double tol, int maxit, int pre, int post,
                   int* iterations, double* rel_res);
/* whether rt_block_lanczos solves the k right-hand sides of a block together
 * with the batched solver (1, the default, for k <= 8 and the mixed or
 * reliable modes) or one after another (0) */
int rt_set_block_solves(rt_hier* h, int batched);
/* PCG iterations over all solves */
} rt_lanczos_stats;
int rt_block_lanczos(rt_hier* h, int k, int max_blocks, double tol, double solve_tol, int mode, int pre, int post,
                     uint64_t seed, double* worst_compliance, rt_lanczos_stats* stats);

/* --- MMA optimiser ------------------------------------------------------- */

typedef struct rt_mma rt_mma;

/* MMA::mma_t(ndim, nConstrain) + init(lower, upper)
 * (optimization/mma_t.h:202-221): n FP32 design variables with bounds
 * lower[n], upper[n] (host or device arrays), 1 <= m <= 8 constraints
 * g(x) <= 0; a0 = 1, a = 0, c = 10, d = 1 as the reference's defaults. */
int rt_mma_create(int n, int m, const float* lower, const float* upper, rt_mma** out);
int rt_mma_destroy(rt_mma* h);
/* MMA::mma_t::update(dev_df, dev_dg, dev_g) (optimization/mma_t.cpp:18-51):
 * df[n] objective gradient, dg[m][n] constraint gradients (row-major), g[m]
 * constraint values at the current design (host or device arrays). Solves the
 * MMA subproblem by the reference's primal-dual iteration (mma_t.cu:330-592),
 * moves the asymptotes (mma_t.cu:161-226); *converged = 1 once the design
 * change has stayed below 1e-3 for more than 10 updates. */
int rt_mma_update(rt_mma* h, const float* df, const float* dg, const float* g, int* converged);
/* The same update with the constraint gradients as m separate rows of n values
 * (host or device arrays), the Scalar** dev_dg form of
 * MMA::mma_t::update(Scalar*, Scalar**, Scalar*) (optimization/mma_t.h:255). */
int rt_mma_update_rows(rt_mma* h, const float* df, const float* const* dg_rows, const float* g, int* converged);
/* MMA::mma_t::get_x (mma_t.cu:156-159): current design (host or device array) */
int rt_mma_get_x(const rt_mma* h, float* x);
/* last update's multipliers lambda[m], artificial variable z, number of
 * subproblem iterations and whether they hit the reference's limit of 80 */
int rt_mma_stats(const rt_mma* h, float* lambda, float* z, int* sub_iters, int* overflowed);

/* --- learned warm start (no counterpart in the reference) ----------------
 *
 * A 3D convolutional encoder-decoder on the first coarse level's (x4) node
 * box: 3 resolutions with 16 / 32 / 64 channels, two 3x3x3 convolutions per
 * resolution with ReLU, 2x2x2 average pooling down, nearest upsampling with
 * additive skips up, and a 3x3x3 output convolution. Input per node: density
 * (mean over the 8^3 fine elements around it) and the restricted load (3
 * components, scaled by 1 / max |load|); output: 3 displacement components
 * (shape only: the solver fits the scale). Weights and activations are
 * bfloat16, accumulation FP32 on the matrix cores; Adam keeps FP32 master
 * weights. Parameters are one FP32 vector: for each of the 11 convolutions
 * (E0A E0B E1A E1B E2A E2B D1A D1B D0A D0B HEAD) the weights [27][cout][cin]
 * then the biases [cout], with cin / cout padded to multiples of 16 (the
 * first layer's inputs 4..15 and the last layer's outputs 3..15 are zero). A
 * tap t holds the offset (t%3-1, t/3%3-1, t/9-1). Not thread-safe per handle. */
typedef struct rt_net rt_net;
int rt_net_create(uint64_t seed, rt_net** out);
int rt_net_destroy(rt_net* n);
int rt_net_num_params(const rt_net* n, size_t* count);
int rt_net_get_params(const rt_net* n, float* w);
/* sets the parameters and resets the optimiser state */
int rt_net_set_params(rt_net* n, const float* w);
/* y[P][3] = net(x[P][4]) on a dense box dims = {nx, ny, nz}, P = nx ny nz,
 * position x + nx (y + ny z); inputs are rounded to bfloat16 */
int rt_net_forward(rt_net* n, const int* dims, const float* x, float* y);
/* loss = mean over P and 3 channels of (net(x) - target)^2 and its gradient
 * with respect to every parameter (layout of rt_net_get_params) */
int rt_net_gradient(rt_net* n, const int* dims, const float* x, const float* target, double* loss, float* grad);
/* stores the hierarchy's current (density, load, displacement) as a training
 * sample, the displacement normalised to unit RMS; call after a solve */
int rt_net_add_sample(rt_net* n, rt_hier* h);
/* loss of the prediction for the hierarchy's current triple (not stored) */
int rt_net_evaluate(rt_net* n, rt_hier* h, double* loss);
/* The network as a coarse-grid correction: the first coarse (x4) level's U
 * becomes the prediction from the density and that level's current F (a
 * restricted residual, say), shape only (unit scale up to the network's
 * error). Works on the FP64 level vectors (rt_set_vcycle_precision(h, 64)). */
int rt_net_predict_coarse(rt_net* n, rt_hier* h);
int rt_net_num_samples(const rt_net* n, int* count);
int rt_net_clear_samples(rt_net* n);
/* steps Adam updates (beta 0.9 / 0.999), each on the mean loss of `batch`
 * samples drawn with replacement; mean_loss = the last step's loss */
int rt_net_train(rt_net* n, int steps, int batch, float lr, uint64_t seed, double* mean_loss);
/* one Adam update (beta 0.9 / 0.999, eps 1e-8, bias-corrected; the optimiser
 * state that rt_net_train uses) with a caller-given gradient grad[num_params] */
int rt_net_adam_step(rt_net* n, const float* grad, float lr);
/* mean milliseconds of one forward pass on a dims box (HIP events, n passes) */
int rt_net_time_forward(rt_net* n, const int* dims, int reps, double* ms);
/* Starts every rt_solve from the energy-optimal combination of the plain
 * start (the current displacement, or the rt_set_warm_start combination) and
 * the prolonged prediction, kept only when it lowers the energy
 * u.Ku / 2 - f.u (otherwise the plain start stays). net = NULL turns it
 * off. The network is not owned and must outlive its use. */
int rt_set_learned_warm_start(rt_hier* h, rt_net* n);
/* info = {1 if the last solve took the learned start else 0, energy of the
 * plain start, energy of the learned start, coefficient of the prediction} */
int rt_learned_warm_start_info(const rt_hier* h, double* info);

/* --- voxelizer ----------------------------------------------------------- */

/* voxelize_mesh (voxelizer/voxelize_main.cpp:344-404): padded cube of side
 * R = smallest power of two > prefer_reso, parity fill on the GPU
 * (voxelize_solid.cu:78-150). bits must hold R^3/32 words; out_box = {min[3],
 * max[3]}. */
int rt_voxelize_mesh(const float* coords, int nverts, const int* faces, int nfaces, int prefer_reso,
                     uint32_t* bits, size_t maxwords, int* out_reso, float* out_box);
/* No counterpart in the reference: rt_voxelize_mesh for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). d_coords[3 nverts], d_faces[3 nfaces] and
 * d_bits[maxwords] are DEVICE arrays of the current device; out_reso[3] and
 * out_box[6] are host arrays. The result equals rt_voxelize_mesh's. It calls
 * no hipMalloc/hipFree: the staged triangles and bounding-box partials live
 * in a static per-device scratch array (about 19 MB, held by each device the
 * library runs on until the process ends), so nfaces is limited to 2^19
 * (524288; more returns RT_EARG). Work runs on the null stream and the call
 * returns after it has finished. Calls that overlap on one device share the
 * scratch and race: serialise them. */
int rt_voxelize_mesh_noalloc(const float* d_coords, int nverts, const int* d_faces, int nfaces, int prefer_reso,
                             uint32_t* d_bits, size_t maxwords, int* out_reso, float* out_box);

/* Autotuner (no counterpart in the reference): on this hierarchy's own sizes,
 * times each launch-shape alternative `reps` times with HIP events and keeps
 * the fastest: z planes per workgroup of the finest residual / apply (8 .. 48),
 * the finest smoother implementation (rt_set_fine_smoother), and per coarse
 * level the sliced (nine waves per 64 nodes) or thread-per-node smoother and
 * residual kernels. The operator is the same; only the summation order of a
 * coarse node's 27 neighbour blocks differs between the two coarse kernel
 * forms (the finest-level choices are bit-identical). Needs rt_update_stencil first; uses the current force and
 * displacement as data. info[6] = {z planes, ms of residual + apply, smoother
 * kind, ms per sweep, mask of coarse levels with the sliced smoother, mask of
 * coarse levels with the sliced residual} (bit l = level l). */
int rt_autotune(rt_hier* h, int reps, double* info);

/* Diagnostic (no counterpart in the reference): the k smallest (ascending)
 * and k largest (descending) Ritz values of the preconditioned operator from
 * the CG coefficients of the last RT_SOLVE_RELIABLE solve (the Lanczos
 * tridiagonal T_jj = 1/a_j + b_{j-1}/a_{j-1}, T_{j,j+1} = sqrt(b_j)/a_j);
 * n_out = its order (the iteration count). Either array may be NULL. */
int rt_cg_ritz_values(const rt_hier* h, int k, double* smallest, double* largest, int* n_out);
/* the CG coefficients themselves: alpha_j, and beta_j (0 past the last) for
 * j < min(nmax, n_out) */
int rt_cg_coefficients(const rt_hier* h, int nmax, double* alpha, double* beta, int* n_out);

/* --- timing hooks (benchmarking) ----------------------------------------- */

int rt_synchronize(rt_hier* h);
/* Runs n preconditioner cycles (precision `bits`, the shape and graph mode
 * rt_solve uses) on the current force and returns the mean milliseconds per
 * cycle measured with HIP events on the handle's stream. */
int rt_time_vcycle(rt_hier* h, int bits, int pre, int post, int n, double* ms_per_cycle);
/* mean milliseconds of one batched FP32 preconditioner cycle for nb vectors
 * (the cycle of rt_solve_batch and of block Lanczos's block solves) */
int rt_time_vcycle_batch(rt_hier* h, int nb, int pre, int post, int n, double* ms_per_cycle);
/* Times n full fine-level smoothing sweeps (colours 0..7, run as four
 * colour-pair launches) of precision `bits`; returns mean ms per sweep. */
int rt_time_fine_sweep(rt_hier* h, int bits, int n, double* ms_per_sweep);
/* Times n finest-level residual evaluations r = f - K u (precision `bits`,
 * u = the current solution, f = the current force); returns mean ms each. */
int rt_time_fine_residual(rt_hier* h, int bits, int n, double* ms_per_residual);

#ifdef __cplusplus
}
#endif

#endif /* ROBTOP_AMD_H */

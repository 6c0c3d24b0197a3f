/*
 * robtop_amd.hpp — header-only C++ facade over the C ABI (robtop_amd.h) that
 * keeps the reference's host interface for the multigrid/SIMP path:
 *
 *   grid::Grid           one stored level      (grid/Grid.h:172-532)
 *   grid::HierarchyGrid  the level hierarchy   (grid/Grid.h:541-680)
 *   free functions       optimization driver   (optimization/optimization.h:28-82)
 *
 * Same names and argument meaning; device memory stays inside the library, so
 * getters that hand out device pointers in the reference (getForce, getRho,
 * getDisplacement, ...) return host copies in the level's box layout here
 * (x fastest; 3-vectors as [x..., y..., z...]). Errors throw robtop_amd::error
 * (the reference prints and continues, cudaCommon.cuh:6-11).
 *
 * Load fields return std::array<double, 3> where the reference returns
 * Eigen::Matrix<double, 3, 1> (optimization.h:75); the value is the same.
 */
#ifndef ROBTOP_AMD_HPP
#define ROBTOP_AMD_HPP

#include "robtop_amd.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace robtop_amd {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != RT_OK) throw error(rc, rt_last_error());
}

using LoadField = std::function<std::array<double, 3>(double[3])>;
using AreaTest = std::function<bool(double[3])>;

namespace grid {

class HierarchyGrid;

/* One stored level. Level 0 is the finest; the reference's layer 1 is not
 * stored (x4 first coarsening, Grid.cpp:390-393). */
class Grid {
public:
    Grid(HierarchyGrid* h, int level);

    int level() const { return level_; }
    int layer() const { return info_[0]; }
    /* element size in finest elements */
    int scale() const { return info_[1]; }
    std::array<int, 3> origin() const { return {info_[2], info_[3], info_[4]}; }
    std::array<int, 3> elementBox() const { return {info_[5], info_[6], info_[7]}; }
    std::array<int, 3> nodeBox() const { return {info_[8], info_[9], info_[10]}; }
    size_t n_nodes_box() const { return (size_t)info_[8] * info_[9] * info_[10]; }
    size_t n_elements_box() const { return (size_t)info_[5] * info_[6] * info_[7]; }
    size_t n_nodes() const;
    size_t n_elements() const;
    bool is_finest() const { return level_ == 0; }
    bool is_coarsest() const;

    std::vector<uint8_t> elementMask() const;
    std::vector<uint8_t> nodeMask() const;
    std::vector<uint8_t> surfaceNodes() const;

    /* Grid::gs_relax (Grid.cu:1068-1111) */
    void gs_relax(int n_times = 1);
    /* Grid::update_residual (Grid.cu:1644-1678) */
    void update_residual(void);
    /* Grid::restrict_residual (Grid.cu:1796-1816), called on the coarse level */
    void restrict_residual(void);
    /* Grid::prolongate_correction (Grid.cu:1898-1915), called on the fine level */
    void prolongate_correction(void);
    /* Grid::reset_displacement (Grid.cu:1917-1922) */
    void reset_displacement(void);
    /* Grid::solve_fem_host (Grid.cpp:1981-2043), coarsest level */
    void solve_fem_host(void);
    /* Grid::relative_residual (Grid.cu:1946-1951), finest level */
    double relative_residual(void);
    /* Grid::compliance (Grid.cpp:1976-1979), finest level */
    double compliance(void);
    /* Grid::filterSensitivity (Grid.cu:2411-2437), finest level */
    void filterSensitivity(double radii);
    /* Grid::applyK (Grid.h:357): f = K(rho) u on the finest level */
    void applyK(const double* u_box, double* f_box);

    std::vector<double> getDisplacement(void) const { return getv(RT_VEC_U); }
    std::vector<double> getForce(void) const { return getv(RT_VEC_F); }
    std::vector<double> getResidual(void) const { return getv(RT_VEC_R); }
    void setDisplacement(const std::vector<double>& v) { setv(RT_VEC_U, v); }
    void setForce(const std::vector<double>& v) { setv(RT_VEC_F, v); }
    std::vector<float> getRho(void) const;
    std::vector<float> getSens(void) const;
    std::vector<double> getStencil(void) const;

private:
    std::vector<double> getv(int which) const;
    void setv(int which, const std::vector<double>& v);
    rt_hier* H() const;
    HierarchyGrid* h_;
    int level_;
    int info_[11];
};

/* The level hierarchy (grid::HierarchyGrid). Owns one library handle. */
class HierarchyGrid {
public:
    HierarchyGrid() = default;
    ~HierarchyGrid() { reset(); }
    HierarchyGrid(const HierarchyGrid&) = delete;
    HierarchyGrid& operator=(const HierarchyGrid&) = delete;

    /* frees the library handle and its device memory */
    void release() { reset(); }

    void set_prefer_reso(int r) { prefer_reso_ = r; }
    void set_skip_layer(bool skip) {
        if (!skip) throw error(RT_EARG, "only the x4 first coarsening (skiplayer1) is supported");
    }

    /* HierarchyGrid::genFromMesh(solid_bit, out_reso) (Grid.cpp:455-712):
     * LSB-first bits of a cube of side out_reso[0]; box [0, R]^3, element size 1 */
    void genFromMesh(const std::vector<unsigned int>& solid_bit, int out_reso[3]) {
        if (out_reso[0] != out_reso[1] || out_reso[0] != out_reso[2])
            throw error(RT_EARG, "the solid must be given on a cube");
        build(out_reso[0], solid_bit.data(), solid_bit.size());
        for (int k = 0; k < 3; k++) { box_[0][k] = 0.f; box_[1][k] = (float)out_reso[0]; }
    }
    /* HierarchyGrid::genFromMesh(pcoords, facevertices) (Grid.cpp:182-452):
     * voxelize at set_prefer_reso, then build */
    void genFromMesh(const std::vector<float>& pcoords, const std::vector<int>& facevertices) {
        int R = 1;
        while (R <= prefer_reso_) R <<= 1;
        std::vector<uint32_t> bits((size_t)R * R * R / 32);
        int reso[3];
        float box[6];
        check(rt_voxelize_mesh(pcoords.data(), (int)(pcoords.size() / 3), facevertices.data(),
                               (int)(facevertices.size() / 3), prefer_reso_, bits.data(), bits.size(), reso, box));
        bits.resize((size_t)reso[0] * reso[0] * reso[0] / 32);
        build(reso[0], bits.data(), bits.size());
        for (int k = 0; k < 3; k++) { box_[0][k] = box[k]; box_[1][k] = box[3 + k]; }
    }

    int n_grid(void) const { return (int)levels_.size(); }
    Grid& operator[](int i) { return *levels_.at(i); }
    rt_hier* handle() const {
        if (!h_) throw error(RT_ESTATE, "hierarchy not built");
        return h_;
    }
    int resolution() const { return R_; }
    /* the finest level's bounding box, [min, max] per axis */
    const float (&box() const)[2][3] { return box_; }

    /* setParameters's material part + uploadTemplateMatrix (optimization.cu:330-343) */
    void setMaterial(double youngs, double poisson, float power_penalty) {
        check(rt_set_material(handle(), youngs, poisson, power_penalty));
    }

    /* Supports and loads from the area tests and load field, evaluated on the
     * finest level's surface nodes at their positions box_min + id * eh
     * (Grid::computeProjectionMatrix, Grid.cpp:1291-1403): a support node is
     * not a load node; the force at a load node is field(pos). */
    void setBoundaryCondition(const AreaTest& fixarea, const AreaTest& loadarea, const LoadField& field) {
        Grid& g = (*this)[0];
        const auto nb = g.nodeBox();
        const auto o = g.origin();
        const std::vector<uint8_t> surf = g.surfaceNodes();
        const size_t n = g.n_nodes_box();
        std::vector<uint8_t> fixed(n, 0), load(n, 0);
        std::vector<double> f(3 * n, 0.0);
        const double eh = (double)(box_[1][0] - box_[0][0]) / R_;
        nload_ = 0;
        for (int z = 0; z < nb[2]; z++)
            for (int y = 0; y < nb[1]; y++)
                for (int x = 0; x < nb[0]; x++) {
                    const size_t v = ((size_t)z * nb[1] + y) * nb[0] + x;
                    if (!surf[v]) continue;
                    double pos[3] = {box_[0][0] + (o[0] + x) * eh, box_[0][1] + (o[1] + y) * eh,
                                     box_[0][2] + (o[2] + z) * eh};
                    if (fixarea(pos)) { fixed[v] = 1; continue; }
                    if (loadarea(pos)) {
                        load[v] = 1;
                        const auto fv = field(pos);
                        for (int c = 0; c < 3; c++) f[c * n + v] = fv[c];
                        nload_++;
                    }
                }
        check(rt_set_support_box(handle(), fixed.data()));
        check(rt_set_force_box(handle(), f.data()));
        // the load nodes are also the admissible load set of the worst-case
        // search (setLoadNodes, projection.cpp:129-269), free direction (wsff)
        check(rt_set_load_nodes(handle(), load.data(), nullptr));
        fixed_ = std::move(fixed);
        load_ = std::move(load);
    }
    /* constrained-direction worst case (wscf, setWorkMode): the admissible load
     * at each load node is along the given unit normal ([3][node box], the
     * reference takes outward normals from CGAL, Grid.cpp:1405-1423) */
    void setLoadNormals(const std::vector<double>& normals_box) {
        if (normals_box.size() != 3 * load_.size()) throw error(RT_EARG, "normals must be 3 x the finest node box");
        check(rt_set_load_nodes(handle(), load_.data(), normals_box.data()));
    }
    /* load-node flags of the finest node box set by setBoundaryCondition */
    const std::vector<uint8_t>& loadMask() const { return load_; }
    size_t n_load_nodes() const { return nload_; }
    /* support flags of the finest node box set by setBoundaryCondition */
    const std::vector<uint8_t>& supportMask() const { return fixed_; }

    /* Grid::init_rho (initDensities, optimization.cpp:687-690) */
    void init_rho(double rho) { check(rt_init_density(handle(), (float)rho)); }
    /* HierarchyGrid::update_stencil (Grid.cpp:2108-2119) */
    void update_stencil(void) { check(rt_update_stencil(handle())); }
    /* HierarchyGrid::v_cycle (Grid.cpp:1068-1113) */
    double v_cycle(int pre_relax = 1, int post_relax = 1) {
        double r = 0;
        check(rt_vcycle(handle(), pre_relax, post_relax, &r));
        return r;
    }
    /* MG-PCG (no reference counterpart; solveFEM iterates v_cycle) */
    rt_solve_stats solve(double tol = 1e-6, int maxit = 200, int mode = RT_SOLVE_MIXED, int pre = 1, int post = 1) {
        rt_solve_stats st{};
        check(rt_solve(handle(), tol, maxit, mode, pre, post, &st));
        return st;
    }
    void set_vcycle_precision(int bits) { check(rt_set_vcycle_precision(handle(), bits)); }
    void synchronize() { check(rt_synchronize(handle())); }

private:
    void reset() {
        levels_.clear();
        if (h_) { rt_destroy(h_); h_ = nullptr; }
    }
    void build(int R, const uint32_t* bits, size_t nwords) {
        reset();
        check(rt_create_from_voxels(R, bits, nwords, &h_));
        R_ = R;
        const int nl = rt_num_levels(h_);
        if (nl < 0) check(nl);
        for (int i = 0; i < nl; i++) levels_.emplace_back(new Grid(this, i));
    }
    rt_hier* h_ = nullptr;
    int R_ = 0;
    int prefer_reso_ = 200;
    size_t nload_ = 0;
    std::vector<uint8_t> fixed_, load_;
    float box_[2][3] = {{0, 0, 0}, {0, 0, 0}};
    std::vector<std::unique_ptr<Grid>> levels_;
};

inline Grid::Grid(HierarchyGrid* h, int level) : h_(h), level_(level) {
    check(rt_level_info(h->handle(), level, info_));
}
inline rt_hier* Grid::H() const { return h_->handle(); }
inline size_t Grid::n_nodes() const {
    size_t ne = 0, nn = 0;
    check(rt_num_active(H(), level_, &ne, &nn));
    return nn;
}
inline size_t Grid::n_elements() const {
    size_t ne = 0, nn = 0;
    check(rt_num_active(H(), level_, &ne, &nn));
    return ne;
}
inline bool Grid::is_coarsest() const { return level_ == h_->n_grid() - 1; }
inline std::vector<uint8_t> Grid::elementMask() const {
    std::vector<uint8_t> e(n_elements_box());
    check(rt_level_masks(H(), level_, e.data(), nullptr));
    return e;
}
inline std::vector<uint8_t> Grid::nodeMask() const {
    std::vector<uint8_t> v(n_nodes_box());
    check(rt_level_masks(H(), level_, nullptr, v.data()));
    return v;
}
inline std::vector<uint8_t> Grid::surfaceNodes() const {
    std::vector<uint8_t> v(n_nodes_box());
    check(rt_surface_nodes_box(H(), level_, v.data()));
    return v;
}
inline void Grid::gs_relax(int n_times) { check(rt_gs_relax(H(), level_, n_times, 0)); }
inline void Grid::update_residual(void) { check(rt_update_residual(H(), level_)); }
inline void Grid::restrict_residual(void) { check(rt_restrict_residual(H(), level_)); }
inline void Grid::prolongate_correction(void) { check(rt_prolongate_correction(H(), level_)); }
inline void Grid::reset_displacement(void) { check(rt_reset_displacement(H(), level_)); }
inline void Grid::solve_fem_host(void) {
    if (!is_coarsest()) throw error(RT_EARG, "solve_fem_host is defined on the coarsest level");
    check(rt_solve_coarsest(H()));
}
inline double Grid::relative_residual(void) {
    double r = 0;
    check(rt_relative_residual(H(), &r));
    return r;
}
inline double Grid::compliance(void) {
    double c = 0;
    check(rt_compliance(H(), &c));
    return c;
}
inline void Grid::filterSensitivity(double radii) { check(rt_filter_sensitivity(H(), radii)); }
inline void Grid::applyK(const double* u_box, double* f_box) { check(rt_apply_k(H(), u_box, f_box)); }
inline std::vector<double> Grid::getv(int which) const {
    std::vector<double> v(3 * n_nodes_box());
    check(rt_get_level_vec(H(), level_, which, v.data()));
    return v;
}
inline void Grid::setv(int which, const std::vector<double>& v) {
    if (v.size() != 3 * n_nodes_box()) throw error(RT_EARG, "vector size does not match the level's node box");
    check(rt_set_level_vec(H(), level_, which, v.data()));
}
inline std::vector<float> Grid::getRho(void) const {
    std::vector<float> r(n_elements_box());
    check(rt_get_density_box(H(), r.data()));
    return r;
}
inline std::vector<float> Grid::getSens(void) const {
    std::vector<float> r(n_elements_box());
    check(rt_get_sensitivity_box(H(), r.data()));
    return r;
}
inline std::vector<double> Grid::getStencil(void) const {
    std::vector<double> s((size_t)243 * n_nodes_box());
    check(rt_get_stencil_box(H(), level_, s.data()));
    return s;
}

}  // namespace grid

/* struct Parameter (optimization/optimization.h:11-22), defaults of
 * config_parser.cpp:4-46 */
struct Parameter {
    float volume_ratio = 0.3f;
    float volume_decrease = 0.02f;
    float design_step = 0.03f;
    int filter_radius = 2;
    float damp_ratio = 0.5f;
    float power_penalty = 3.f;
    float min_rho = 1e-3f;
    int gridreso = 200;
    float youngs_modulu = 1e6f;
    float poisson_ratio = 0.4f;
};

/* The reference keeps its state in process globals (params, grids;
 * optimization.cpp:8-12); these accessors hold the same. */
inline Parameter& params() {
    static Parameter p;
    return p;
}
/* never destroyed: device memory must not be freed from static destructors,
 * which may run after the HIP runtime has shut down; call release() instead */
inline grid::HierarchyGrid& getGrids(void) {
    static grid::HierarchyGrid* g = new grid::HierarchyGrid;
    return *g;
}
struct BoundaryCondition {
    AreaTest fix, load;
    LoadField field;
};
inline BoundaryCondition& boundaryCondition() {
    static BoundaryCondition bc;
    return bc;
}

/* setParameters (optimization.cpp:32-50); shell width and logging flags are
 * accepted and ignored (shell elements come from rt_set_shell_box) */
inline void setParameters(float volRatio, float volDecrease, float designStep, float filterRadi, float dampRatio,
                          float powerPenal, float min_density, int gridreso, float youngs_modulu,
                          float poisson_ratio, float /*shell_width*/ = 0, bool /*logdensity*/ = false,
                          bool /*logcompliance*/ = false) {
    Parameter& p = params();
    p.volume_ratio = volRatio; p.volume_decrease = volDecrease; p.design_step = designStep;
    p.filter_radius = (int)filterRadi; p.damp_ratio = dampRatio; p.power_penalty = powerPenal;
    p.min_rho = min_density; p.gridreso = gridreso; p.youngs_modulu = youngs_modulu;
    p.poisson_ratio = poisson_ratio;
    getGrids().set_prefer_reso(gridreso);
}

/* setBoundaryCondition (optimization.cpp:680-685): stored, applied by buildGrids */
inline void setBoundaryCondition(AreaTest fixarea, AreaTest loadarea, LoadField loadforce) {
    boundaryCondition() = {std::move(fixarea), std::move(loadarea), std::move(loadforce)};
}

inline void applyBoundaryCondition_() {
    const BoundaryCondition& bc = boundaryCondition();
    if (bc.fix && bc.load && bc.field) getGrids().setBoundaryCondition(bc.fix, bc.load, bc.field);
}

/* buildGrids (optimization.cpp:14-30): hierarchy from a triangle mesh */
inline void buildGrids(const std::vector<float>& coords, const std::vector<int>& trifaces) {
    getGrids().genFromMesh(coords, trifaces);
    applyBoundaryCondition_();
}
/* hierarchy from solid bits (HierarchyGrid::genFromMesh(solid_bit, reso)) */
inline void buildGrids(const std::vector<unsigned int>& solid_bits, int reso[3]) {
    getGrids().genFromMesh(solid_bits, reso);
    applyBoundaryCondition_();
}

/* uploadTemplateMatrix (optimization.cu:330-343) */
inline void uploadTemplateMatrix(void) {
    const Parameter& p = params();
    getGrids().setMaterial(p.youngs_modulu, p.poisson_ratio, p.power_penalty);
}
/* initDensities (optimization.cpp:687-690) */
inline void initDensities(double rho) { getGrids().init_rho(rho); }
/* update_stencil (optimization.cpp:692-695) */
inline void update_stencil(void) { getGrids().update_stencil(); }
/* MG-PCG solve of the current hierarchy to ||f - K u|| / ||f|| <= tol (no
 * reference counterpart); returns the solver statistics and throws if maxit
 * runs out first */
inline rt_solve_stats solveFEM(double tol, int mode = RT_SOLVE_MIXED, int maxit = 5000) {
    rt_solve_stats st = getGrids().solve(tol, maxit, mode, 1, 1);
    if (!st.converged)
        throw error(RT_ENUM, "solveFEM: relative residual " + std::to_string(st.rel_res) + " above " +
                                 std::to_string(tol) + " after " + std::to_string(st.iterations) + " iterations");
    return st;
}
/* void solveFEM(void) (optimization.h:47, optimization.cpp:85-91): the
 * reference repeats V(1,1) cycles until the relative residual is at most 1e-4;
 * this reaches the same tolerance with MG-PCG (V(1,1) preconditioner) */
inline void solveFEM(void) { (void)solveFEM(1e-4, RT_SOLVE_RELIABLE); }

/* double modifiedPM(void) (optimization.h:50, optimization.cpp:123-261): the
 * worst-case load of the load nodes set with rt_set_load_nodes, by the
 * modified power method from a random start (randForce, Grid.cu:2242-2245;
 * seeded from a per-process counter here instead of the clock), at most 500
 * steps of one V(1,1) cycle each. The worst load and its displacement stay in
 * the force and displacement; returns the worst compliance. */
inline double modifiedPM(void) {
    static uint64_t seed = 0;
    rt_hier* h = getGrids().handle();
    check(rt_rand_force(h, ++seed));
    double c = 0;
    int cycles = 0;
    check(rt_modified_pm(h, 500, &c, &cycles));
    if (std::isnan(c)) throw error(RT_ENUM, "modifiedPM: NaN compliance");
    return c;
}
/* computeSensitivity (optimization.cu:109-133): element sensitivity, then the
 * filter of radius params.filter_radius */
inline void computeSensitivity(void) {
    check(rt_compute_sensitivity(getGrids().handle()));
    getGrids()[0].filterSensitivity(params().filter_radius);
}
/* updateDensities (optimization.cu:164-226); returns the multiplier */
inline float updateDensities(float Vgoal) {
    const Parameter& p = params();
    float thr = 0;
    int trials = 0;
    check(rt_update_densities(getGrids().handle(), Vgoal, p.design_step, p.damp_ratio, p.min_rho, &thr, &trials));
    return thr;
}

namespace MMA {
using Scalar = float;
/* MMA::mma_t (optimization/mma_t.h:159-266): the method of moving asymptotes
 * on the GPU for ndim FP32 design variables and nConstrain <= 8 constraints
 * g(x) <= 0. Usage as in the reference: construct, init(bounds), then
 * update(df, dg, g) once per design iteration; get_x() returns the design. */
class mma_t {
public:
    mma_t(int ndim, int nConstrain, Scalar /*err_tol*/ = 1e-7f) : n_(ndim), m_(nConstrain) {}
    ~mma_t() { reset(); }
    mma_t(const mma_t&) = delete;
    mma_t& operator=(const mma_t&) = delete;

    /* mma_t::init (mma_t.h:215-224): bounds per variable or uniform */
    void init(const Scalar* lower_bound, const Scalar* upper_bound) {
        reset();
        check(rt_mma_create(n_, m_, lower_bound, upper_bound, &h_));
    }
    void init(Scalar lower_bound, Scalar upper_bound) {
        const std::vector<Scalar> lo((size_t)n_, lower_bound), hi((size_t)n_, upper_bound);
        init(lo.data(), hi.data());
    }
    /* bool mma_t::update(Scalar* dev_df, Scalar** dev_dg, Scalar* dev_g)
     * (mma_t.h:255, mma_t.cpp:18-51): df[n] objective gradient, dg[i][n] the
     * i-th constraint's gradient, g[m] the constraint values (host or device
     * arrays); true once the design change has stayed below 1e-3 for more than
     * 10 updates */
    bool update(Scalar* dev_df, Scalar** dev_dg, Scalar* dev_g) {
        int conv = 0;
        check(rt_mma_update_rows(H(), dev_df, dev_dg, dev_g, &conv));
        return conv != 0;
    }
    /* mma_t::get_x(Scalar* dst) (mma_t.cu:156-159): host or device array of n */
    void get_x(Scalar* dst) const { check(rt_mma_get_x(H(), dst)); }
    std::vector<Scalar> get_x(void) const {
        std::vector<Scalar> x((size_t)n_);
        get_x(x.data());
        return x;
    }
    int n_dim(void) const { return n_; }
    int n_constrain(void) const { return m_; }

private:
    rt_mma* H() const {
        if (!h_) throw error(RT_ESTATE, "mma_t::init has not been called");
        return h_;
    }
    void reset() {
        if (h_) { rt_mma_destroy(h_); h_ = nullptr; }
    }
    int n_, m_;
    rt_mma* h_ = nullptr;
};
}  // namespace MMA

/* converge_criteria (utils/snippet.h:156-213) */
class converge_criteria {
public:
    converge_criteria(int n_constrain, int max_counter, double stop_thres)
        : n_(n_constrain), max_(max_counter), thres_(stop_thres), oldc_(n_constrain) {
        old_.fill(NAN);
        for (auto& a : oldc_) a.fill(NAN);
    }
    bool update(double value, const double* constrain) {
        const int k = itn_ % 20;
        old_[k] = value;
        for (int i = 0; i < n_; i++) oldc_[i][k] = constrain[i];
        auto at = [&](const std::array<double, 20>& a, int d) { return a[((itn_ - d) % 20 + 20) % 20]; };
        const double fch = std::fabs(at(old_, 0) + at(old_, 1) - at(old_, 2) - at(old_, 3)) /
                           (at(old_, 2) + at(old_, 3));
        bool conv = fch < thres_;
        for (int i = 0; i < n_; i++) {
            const double gch = std::fabs(at(oldc_[i], 0) + at(oldc_[i], 1) - at(oldc_[i], 2) - at(oldc_[i], 3)) /
                               (at(oldc_[i], 2) + at(oldc_[i], 3));
            if (gch > 1e-3) { conv = false; break; }
        }
        stop_ = conv ? stop_ + 1 : std::max(stop_ - 1, 0);
        itn_++;
        return stop_ > max_;
    }

private:
    int n_, max_;
    double thres_;
    int stop_ = 0, itn_ = 0;
    std::array<double, 20> old_;
    std::vector<std::array<double, 20>> oldc_;
};

}  // namespace robtop_amd

#endif /* ROBTOP_AMD_HPP */

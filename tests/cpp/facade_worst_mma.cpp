// The C++ facade's worst-case and MMA entry points, called the way the
// reference's own code calls them:
//
//   mma EX         MMA::mma_t on the reference's 2-D test problem
//                  (mma_t.cpp:208-300): mma_t(2, 4), init(-1, 1), then
//                  update(df, dg, g) with the constraint gradients as separate
//                  rows (Scalar**), until update() reports convergence or 200
//                  updates. Prints one JSON line per update with the design.
//   pm EX EY EZ    the worst case of a cantilever box (optimization.cpp:583-673
//                  setup): supports on x = 0, load area the top face z = EZ,
//                  density 0.5; update_stencil, then modifiedPM() (the
//                  reference's double modifiedPM(void), optimization.cpp:123-261)
//                  and solveFEM() (void, to 1e-4). Prints the worst compliance.
//   vox FILE RESO  buildGrids(coords, trifaces) (optimization.cpp:14-30) on the
//                  triangle mesh in FILE (int32 vertex count, int32 face count,
//                  float32 xyz per vertex, int32 vertex triple per face) with
//                  gridreso RESO. Prints the cube side, the box and the finest
//                  level's element mask (origin, sizes, 64-bit FNV-1a hash).
#include "robtop_amd.hpp"

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace ra = robtop_amd;

static int run_mma() {
    ra::MMA::mma_t mma(2, 4);
    mma.init(-1.f, 1.f);
    std::vector<float> x(2), df(2), g(4);
    std::vector<std::vector<float>> dg(4, std::vector<float>(2));
    float* rows[4] = {dg[0].data(), dg[1].data(), dg[2].data(), dg[3].data()};
    for (int it = 1; it <= 200; it++) {
        mma.get_x(x.data());
        const double a = x[0], b = x[1], e = std::exp(10 * (a - b) * (a - b));
        df[0] = (float)(0.5 + 20 * (a - b) * e);
        df[1] = (float)(std::sqrt(3.0) / 2 + 20 * (b - a) * e);
        g[0] = (float)(a * a + b * b - 1);
        g[1] = (float)(a * a - 0.75);
        g[2] = (float)(b * b - 0.75);
        g[3] = (float)((a + b) * (a + b) - 1);
        dg[0] = {(float)(2 * a), (float)(2 * b)};
        dg[1] = {(float)(2 * a), 0.f};
        dg[2] = {0.f, (float)(2 * b)};
        dg[3] = {(float)(2 * (a + b)), (float)(2 * (a + b))};
        const bool conv = mma.update(df.data(), rows, g.data());
        mma.get_x(x.data());
        std::printf("{\"update\": %d, \"x\": [%.9g, %.9g], \"converged\": %d}\n", it, x[0], x[1], conv ? 1 : 0);
        if (conv) break;
    }
    return 0;
}

static int run_pm(int ex, int ey, int ez) {
    int R = 1;
    while (R < ex || R < ey || R < ez) R <<= 1;
    ra::setParameters(0.4f, 0.02f, 0.03f, 2.f, 0.5f, 3.f, 1e-3f, R, 1e6f, 0.4f);
    ra::setBoundaryCondition([](double p[3]) { return p[0] < 0.5; },
                             [ez](double p[3]) { return p[2] > ez - 0.5; },
                             [](double[3]) { return std::array<double, 3>{0.0, 0.0, -1.0}; });
    std::vector<unsigned int> bits((size_t)R * R * R / 32, 0u);
    for (int z = 0; z < ez; z++)
        for (int y = 0; y < ey; y++)
            for (int x = 0; x < ex; x++) {
                const size_t id = (size_t)x + (size_t)R * (y + (size_t)R * z);
                bits[id >> 5] |= 1u << (id & 31);
            }
    int reso[3] = {R, R, R};
    ra::buildGrids(bits, reso);
    ra::uploadTemplateMatrix();
    ra::initDensities(0.5);
    ra::update_stencil();
    const double c = ra::modifiedPM();
    ra::grid::HierarchyGrid& grids = ra::getGrids();
    size_t nl = 0;
    for (uint8_t v : grids.loadMask()) nl += v;
    ra::solveFEM();   // re-solve the worst load to 1e-4 with the reference's void solveFEM()
    const double c2 = grids[0].compliance();
    std::printf("{\"worst_compliance\": %.17g, \"load_nodes\": %zu, \"compliance_after_solveFEM\": %.17g}\n", c, nl,
                c2);
    grids.release();
    return 0;
}

static int run_vox(const char* path, int reso) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) throw std::runtime_error("cannot open mesh file");
    int32_t nv = 0, nf = 0;
    if (std::fread(&nv, 4, 1, fp) != 1 || std::fread(&nf, 4, 1, fp) != 1 || nv < 0 || nf < 0) {
        std::fclose(fp);
        throw std::runtime_error("bad mesh header");
    }
    std::vector<float> coords((size_t)nv * 3);
    std::vector<int> faces((size_t)nf * 3);
    const bool ok = std::fread(coords.data(), 4, coords.size(), fp) == coords.size() &&
                    std::fread(faces.data(), 4, faces.size(), fp) == faces.size();
    std::fclose(fp);
    if (!ok) throw std::runtime_error("short mesh file");
    ra::setParameters(0.4f, 0.02f, 0.03f, 2.f, 0.5f, 3.f, 1e-3f, reso, 1e6f, 0.4f);
    ra::buildGrids(coords, faces);
    ra::grid::HierarchyGrid& grids = ra::getGrids();
    rt_hier* h = grids.handle();
    int info[11];
    if (rt_level_info(h, 0, info) != RT_OK) throw std::runtime_error(rt_last_error());
    const size_t ne = (size_t)info[5] * info[6] * info[7], nn = (size_t)info[8] * info[9] * info[10];
    std::vector<uint8_t> eact(ne), vact(nn);
    if (rt_level_masks(h, 0, eact.data(), vact.data()) != RT_OK) throw std::runtime_error(rt_last_error());
    uint64_t hash = 1469598103934665603ull;
    size_t nact = 0;
    for (uint8_t v : eact) {
        hash = (hash ^ v) * 1099511628211ull;
        nact += v;
    }
    const auto& box = grids.box();
    std::printf("{\"R\": %d, \"origin\": [%d, %d, %d], \"elements\": [%d, %d, %d], \"active\": %zu, "
                "\"hash\": \"%016llx\", \"box\": [%.9g, %.9g, %.9g, %.9g, %.9g, %.9g]}\n",
                grids.resolution(), info[2], info[3], info[4], info[5], info[6], info[7], nact,
                (unsigned long long)hash, box[0][0], box[0][1], box[0][2], box[1][0], box[1][1], box[1][2]);
    grids.release();
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("usage: %s mma | pm EX EY EZ | vox FILE RESO\n", argv[0]);
        return 0;
    }
    try {
        const std::string mode = argv[1];
        if (mode == "mma") return run_mma();
        if (mode == "pm" && argc == 5) return run_pm(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
        if (mode == "vox" && argc == 4) return run_vox(argv[2], std::atoi(argv[3]));
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

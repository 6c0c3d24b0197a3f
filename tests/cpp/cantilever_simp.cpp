// SIMP loop on a cantilever box through the C++ facade, written the way the
// reference's own driver calls its host API (example.cpp:38-67,
// TestSuit::testOrdinaryTopopt, test_utils.cpp:589-660):
//
//   setParameters -> setBoundaryCondition -> buildGrids -> uploadTemplateMatrix
//   -> initDensities -> loop { update_stencil, solve, compliance, stop test,
//                              computeSensitivity, updateDensities }
//
// Usage: cantilever_simp EX EY EZ ITERS OUTDIR
//   box of EX x EY x EZ voxels at the origin of the smallest power-of-two cube;
//   nodes at x = 0 fixed, unit -z load spread over the nodes at x = EX.
// Prints one JSON line per SIMP iteration. Afterwards OUTDIR holds, for the
// final densities (after one more update_stencil + solve):
//   rho.f32 and rhop.f32 (rho and rho^p on the element box), u.f64 / f.f64 (3 x node box), fixed.u8 (node box),
//   info.txt ("R ex ey ez nx ny nz compliance").
#include "robtop_amd.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace ra = robtop_amd;

template <typename T>
static void dump(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) throw std::runtime_error("cannot write " + path);
}

int main(int argc, char** argv) {
    if (argc != 6) {
        std::printf("usage: %s EX EY EZ ITERS OUTDIR\n", argv[0]);
        return 0;
    }
    const int ex = std::atoi(argv[1]), ey = std::atoi(argv[2]), ez = std::atoi(argv[3]);
    const int iters = std::atoi(argv[4]);
    const std::string out = argv[5];
    int R = 1;
    while (R < ex || R < ey || R < ez) R <<= 1;
    try {
        // config: volume 0.4, filter r = 2, damp 0.5, step 0.03, p = 3, rho_min 1e-3, E 1e6, nu 0.4
        ra::setParameters(0.4f, 0.02f, 0.03f, 2.f, 0.5f, 3.f, 1e-3f, R, 1e6f, 0.4f);
        const double nface = (double)(ey + 1) * (ez + 1);
        ra::setBoundaryCondition([](double p[3]) { return p[0] < 0.5; },
                                 [ex](double p[3]) { return p[0] > ex - 0.5; },
                                 [nface](double[3]) { return std::array<double, 3>{0.0, 0.0, -1.0 / nface}; });
        std::vector<unsigned int> bits((size_t)R * R * R / 32, 0u);
        for (int z = 0; z < ez; z++)
            for (int y = 0; y < ey; y++)
                for (int x = 0; x < ex; x++) {
                    const size_t id = (size_t)x + (size_t)R * (y + (size_t)R * z);
                    bits[id >> 5] |= 1u << (id & 31);
                }
        int reso[3] = {R, R, R};
        ra::buildGrids(bits, reso);
        ra::grid::HierarchyGrid& grids = ra::getGrids();
        ra::uploadTemplateMatrix();
        const ra::Parameter& P = ra::params();
        ra::initDensities(P.volume_ratio);
        float Vgoal = P.volume_ratio;
        ra::converge_criteria stop_check(1, 5, 1e-3);
        for (int itn = 1; itn <= iters; itn++) {
            Vgoal *= (1 - P.volume_decrease);
            double Vc = Vgoal - P.volume_ratio;
            if (Vgoal < P.volume_ratio) Vgoal = P.volume_ratio;
            ra::update_stencil();
            const rt_solve_stats st = ra::solveFEM(1e-6);
            const double c = grids[0].compliance();
            std::printf("{\"iter\": %d, \"compliance\": %.17g, \"pcg_iters\": %d, \"rel_res\": %.3e}\n", itn, c,
                        st.iterations, st.rel_res);
            std::fflush(stdout);
            if (stop_check.update(c, &Vc) && Vgoal <= P.volume_ratio) break;
            ra::computeSensitivity();
            ra::updateDensities(Vgoal);
        }
        ra::update_stencil();
        ra::solveFEM(1e-8, RT_SOLVE_FP64);
        const double c = grids[0].compliance();
        ra::grid::Grid& g0 = grids[0];
        const size_t nn = g0.n_nodes_box();
        std::vector<double> u(3 * nn), f(3 * nn);
        ra::check(rt_get_displacement_box(grids.handle(), u.data()));
        ra::check(rt_get_force_box(grids.handle(), f.data()));
        std::vector<float> rhop(g0.n_elements_box());
        ra::check(rt_get_rhop_box(grids.handle(), rhop.data()));
        dump(out + "/rho.f32", g0.getRho());
        dump(out + "/rhop.f32", rhop);
        dump(out + "/u.f64", u);
        dump(out + "/f.f64", f);
        dump(out + "/fixed.u8", grids.supportMask());
        const auto nb = g0.nodeBox();
        std::ofstream info(out + "/info.txt");
        info.precision(17);
        info << R << " " << ex << " " << ey << " " << ez << " " << nb[0] << " " << nb[1] << " " << nb[2] << " " << c
             << "\n";
        grids.release();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// Uses the C++ mirror (include/vgc_geogram.hpp) the way geogram's callers use
// the reference classes.  Without a GPU it checks that context creation fails
// cleanly (exit 3); with one it runs CVT Lloyd, cells and OT and checks
// invariants (exit 0 on success).
#include <cmath>
#include <cstdio>
#include <vector>

#include "vgc_geogram.hpp"

int main() {
    using namespace vgc::geogram;
    const double box[6] = {0, 0, 0, 1, 1, 1};
    const unsigned n = 5000;
    std::vector<double> x(3 * n);
    vgc_sample_points(0, 42, n, x.data());
    Context* ctx = nullptr;
    try {
        ctx = new Context(0);
    } catch (const std::exception& e) {
        std::printf("no GPU: %s\n", e.what());
        return 3;
    }
    RestrictedVoronoiDiagram rvd(*ctx, box);
    rvd.set_points(n, x.data());
    std::vector<double> mg(3 * n, 0.0), m(n, 0.0);
    rvd.compute_centroids(mg.data(), m.data());
    double s = 0;
    for (double v : m) s += v;
    if (std::fabs(s - 1.0) > 1e-12) return 1;
    CentroidalVoronoiTesselation cvt(*ctx, box);
    cvt.set_points(n, x.data());
    cvt.Lloyd_iterations(5);
    // quasi-Newton from the Lloyd points: the energy decreases
    CentroidalVoronoiTesselation cvt2(*ctx, box);
    cvt2.set_points(n, cvt.embedding(0));
    cvt2.Newton_iterations(5);
    if (!(cvt2.optimizer_stats().f < cvt2.optimizer_stats().f0)) return 5;
    // exact kNN through the NearestNeighborSearch mirror: the nearest point
    // of a stored point is itself, at distance 0
    NearestNeighborSearch nn(*ctx);
    nn.set_points(n, x.data());
    index_t nb[4];
    double d2[4];
    nn.get_nearest_neighbors(4, index_t(17), nb, d2);
    if (nb[0] != 17 || d2[0] != 0.0 || !(d2[1] > 0.0) || !(d2[2] >= d2[1])) return 6;
    OptimalTransportMap3d otm(*ctx, box);
    otm.set_points(n, cvt.embedding(0));
    otm.set_epsilon(1e-4);
    otm.optimize(50);
    if (!otm.stats().converged) return 2;
    // learned start with an untrained (zero) network: weights 0, judged, converges
    std::vector<float> params(VGC_MLP_PARAMS, 0.0f);
    OptimalTransportMap3d otm2(*ctx, box);
    otm2.set_points(n, cvt.embedding(0));
    otm2.set_learned_initial_weights(params.data());
    otm2.optimize(50);
    if (!otm2.stats().converged || otm2.stats().newton_iters != otm.stats().newton_iters) return 4;
    // optimize_levels with convex level starts: two levels (the seeds are in
    // a random order already), each converges
    OptimalTransportMap3d otm3(*ctx, box);
    otm3.set_points(n, cvt.embedding(0));
    otm3.set_epsilon(1e-4);
    otm3.optimize_levels({0u, uint32_t(n / 8), uint32_t(n)}, 50, true);
    if (!otm3.stats().converged) return 7;
    // coarse-plan start (every 16th seed), judged against the zero start
    // (which a centroidal diagram with equal masses nearly solves already):
    // converges in no more Newton steps than the zero start
    OptimalTransportMap3d otm4(*ctx, box);
    otm4.set_points(n, cvt.embedding(0));
    otm4.set_epsilon(1e-4);
    otm4.set_coarse_initial_weights(16);
    otm4.optimize(50);
    if (!otm4.stats().converged || otm4.stats().newton_iters > otm.stats().newton_iters) return 8;
    std::printf("ok: sum m = %.17g, OT newton %u\n", s, otm.stats().newton_iters);
    delete ctx;
    return 0;
}

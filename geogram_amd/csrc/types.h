// Data layout shared by the kernels and their CPU emulation: seeds in
// Morton order, the implicit bounding-volume tree over buckets of 16 seeds,
// per-instance offsets, and the distance helpers of the neighbour search.
#pragma once

#include <stdint.h>

#include "wave.h"

namespace vgc {

VGC_SD int imin(int a, int b) { return a < b ? a : b; }

// Morton-ordered seed in FP64 (x, y, z, weight).
struct alignas(32) pt4 {
    double x, y, z, w;
};

// Morton-ordered seed in FP32, for the neighbour search.
struct alignas(16) pf4 {
    float x, y, z, w;
};

// Node of the implicit bounding-volume tree over buckets of kBucket
// consecutive Morton-ordered seeds (float box, rounded outwards).
struct alignas(32) bnode {
    float lo[3];
    float hi[3];
    int pad[2];
};

constexpr int kBucket = 16;

// float conversions rounded down / up (exact bounds for the tree)
VGC_SD float f_down(double x) {
    float f = float(x);
    if (double(f) > x) f = nextafterf(f, -INFINITY);
    return f;
}
VGC_SD float f_up(double x) {
    float f = float(x);
    if (double(f) < x) f = nextafterf(f, INFINITY);
    return f;
}

VGC_SD float box_dist2(const bnode& b, float px, float py, float pz) {
    float dx = fmaxf(fmaxf(b.lo[0] - px, px - b.hi[0]), 0.0f);
    float dy = fmaxf(fmaxf(b.lo[1] - py, py - b.hi[1]), 0.0f);
    float dz = fmaxf(fmaxf(b.lo[2] - pz, pz - b.hi[2]), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

VGC_SD float dist2f(const pf4& a, float px, float py, float pz) {
    float dx = a.x - px, dy = a.y - py, dz = a.z - pz;
    return dx * dx + dy * dy + dz * dz;
}

// Per-instance layout of the Morton-ordered seeds, buckets and trees.
struct InstLayout {
    const int* inst_of;       // Morton position -> instance (nullptr: one)
    const int* p0;            // first Morton position
    const int* n;             // seed count
    const int* b0;            // first bucket
    const int* nb;            // bucket count
    const long long* t0;      // node offset: node k (1 <= k < 2L) at t0 + k
    const int* L;             // leaves (power of two >= nb)
};

// Squared radius (FP32, rounded up) within which a seed can still cut a
// partial cell of squared radius r2 around a seed of weight wi; wmax is the
// largest weight of the instance (Voronoi: weighted = false).  A plane at
// distance d can cut only if (d^2 + wi - wj) / (2 d) < sqrt(4.1)/2 R, the
// security-radius margin of generic_RVD.h:2319-2322.
VGC_SD float radius_for(double r2, double wi, double wmax, int weighted) {
    double rho2;
    if (!weighted) {
        rho2 = 4.1 * r2;
    } else {
        double a = 1.0124228365658292 * sqrt(r2);  // sqrt(4.1) / 2
        double cw = wi - wmax;                     // <= 0
        double rho = a + sqrt(a * a - cw);
        rho2 = rho * rho;
    }
    return f_up(rho2 * 1.0001);
}

}  // namespace vgc

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
// consecutive Morton-ordered seeds (float box, rounded outwards) and the
// largest weight below it (rounded up; -inf for an empty node).
struct alignas(32) bnode {
    float lo[3];
    float hi[3];
    float wmax;
    int pad;
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

// Power-distance filter of the weighted neighbour search.  A partial
// Laguerre cell of seed p (weight wi) lies in the ball B(ctr, rc).  For x in
// the ball, pow_j(x) - pow_i(x) (pow_k(x) = |x - q_k|^2 - w_k) differs from its
// value at ctr by at most 2 rc |q_j - p|, so seed q (weight wj) can cut the
// cell only if
//     |ctr - q|^2 - wj - (|ctr - p|^2 - wi) < 2 rc |q - p|.
// Unlike the ball around p itself, this stays local when the cell sits away
// from its seed, as Laguerre cells of a transport plan do.  The tests run in
// FP32; the margin covers the rounding of the FP32 coordinates (relative
// 2^-24 of the coordinate scale), weights and arithmetic, so a seed that can
// cut is never dropped.
struct PowerFilter {
    int on;
    float cx, cy, cz;  // ball centre = centre of the cell's bounding box
    float rc;          // ball radius, rounded up
    float hx, hy, hz;  // half-widths of the bounding box, rounded up
    float phi;         // |ctr - p|^2 - wi
    float dcp;         // |ctr - p|
    float s;           // 1e-6 x coordinate scale of the instance
};

// Ball and box of a flagged partial cell, handed from the clipping kernel to
// the list tiers (rc < 0: none).
struct fball {
    float cx, cy, cz, rc, hx, hy, hz, pad;
};

// The cell lies in the ball and in the box, so over the cell
// 2 (x - ctr).(p - q) >= -min(2 rc |q - p|, 2 sum_a h_a |q_a - p_a|).
// dc2: squared distance from the centre to q (or a lower bound over a tree
// node), wj: q's weight (or an upper bound), d: |q - p| and ax, ay, az:
// |q_a - p_a| (or upper bounds).  True when q may cut.
template <class TF>
VGC_HD auto pf_may_cut(const PowerFilter& F, const TF& dc2, const TF& wj, const TF& d,
                       const TF& ax, const TF& ay, const TF& az) {
    const TF rb = (2.0f * F.rc) * d;
    const TF rx = 2.0f * (F.hx * ax + F.hy * ay + F.hz * az);
    const TF rd = vminf(rb, rx);
    const TF lhs = dc2 - wj - F.phi - rd;
    const TF mar = 1e-6f * (dc2 + vabsf(wj) + fabsf(F.phi) + rd) + F.s * (F.dcp + 2.0f * d + F.rc);
    return lhs < mar;
}

// Ball form only (d a lower bound of |q - p| on a monotone branch).
template <class TF>
VGC_HD auto pf_may_cut_ball(const PowerFilter& F, const TF& dc2, const TF& wj, const TF& d) {
    const TF rd = (2.0f * F.rc) * d;
    const TF lhs = dc2 - wj - F.phi - rd;
    const TF mar = 1e-6f * (dc2 + vabsf(wj) + fabsf(F.phi) + rd) + F.s * (F.dcp + 2.0f * d + F.rc);
    return lhs < mar;
}

// Tree node: dc2 = squared distance from the centre to the node's box,
// wm = the node's max weight, dmin / dmax = nearest / farthest distance from
// p to the box, fx, fy, fz = farthest per-axis offsets.  Two lower bounds of
// the left side over the node: with the farthest offsets, and, since
// |ctr - q| >= |q - p| - |ctr - p| and (d - dcp)^2 - 2 rc d grows for
// d >= dcp + rc, with d = dmin there.
template <class TF>
VGC_HD auto pf_node_may_cut(const PowerFilter& F, const TF& dc2, const TF& wm, const TF& dmin,
                            const TF& dmax, const TF& fx, const TF& fy, const TF& fz) {
    auto may = pf_may_cut(F, dc2, wm, dmax, fx, fy, fz);
    const TF t = dmin - F.dcp;
    return may & ((dmin < F.dcp + F.rc) | pf_may_cut_ball(F, t * t, wm, dmin));
}

// Filter of seed p (FP64 position, weight) from a ball; the instance box
// gives the coordinate scale.
VGC_SD PowerFilter filter_from_ball(const fball& b, const pt4& p, const double* box) {
    PowerFilter F;
    F.on = b.rc >= 0.0f;
    F.cx = b.cx;
    F.cy = b.cy;
    F.cz = b.cz;
    F.rc = b.rc;
    F.hx = b.hx;
    F.hy = b.hy;
    F.hz = b.hz;
    const double fx = double(b.cx) - p.x, fy = double(b.cy) - p.y, fz = double(b.cz) - p.z;
    const double off2 = fx * fx + fy * fy + fz * fz;
    F.phi = float(off2 - p.w);
    F.dcp = f_up(sqrt(off2));
    double cs = 0.0;
    for (int k = 0; k < 6; ++k) cs = fmax(cs, fabs(box[k]));
    F.s = float(1e-6 * cs);
    return F;
}

}  // namespace vgc

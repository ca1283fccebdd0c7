// Semi-discrete optimal transport: per-element kernels shared by the gfx950
// backend and the CPU emulation.
//
// The Laguerre cells of seeds p_i with weights w_i (power distance
// |x - p_i|^2 - w_i, the convention of PeriodicDelaunay3d::set_weights,
// delaunay/periodic_delaunay_3d.cpp:3093, 3719-3726) have masses m_i(w) for
// the uniform density on the box.  Newton's method solves m(w) = nu with the
// Jacobian
//     dm_i/dw_j = -A_ij / (2 |p_i - p_j|)   (j a Laguerre neighbour of i),
//     dm_i/dw_i = sum_j A_ij / (2 |p_i - p_j|),
// A_ij the area of the facet shared by the two cells: a weighted graph
// Laplacian, symmetric positive semi-definite, kernel = constants.  It is the
// Newton system the Exploragram solver called by compute_OTM
// (examples/exploragram/compute_OTM/main.cpp:283-297) hands to OpenNL.
#pragma once

#include "types.h"

namespace vgc {

// Row i of the Hessian from the cells' neighbour CSR (caller numbering):
// off-diagonal values into val[ptr[i]..ptr[i+1]) and the diagonal.
// period: the periods of a periodic box (null otherwise); the distance to a
// neighbour is then the one to its nearest image.
VGC_SD double min_image(double e, double L) { return e - L * rint(e / L); }

VGC_SD void ot_hessian_row(int i, const double* __restrict__ xyz, const long long* __restrict__ ptr,
                           const int* __restrict__ col, const double* __restrict__ area,
                           double* __restrict__ val, double* __restrict__ diag,
                           const double* __restrict__ period) {
    const double px = xyz[3 * (long long)i], py = xyz[3 * (long long)i + 1],
                 pz = xyz[3 * (long long)i + 2];
    double d = 0.0;
    for (long long k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int j = col[k];
        double ex = xyz[3 * (long long)j] - px, ey = xyz[3 * (long long)j + 1] - py,
               ez = xyz[3 * (long long)j + 2] - pz;
        if (period) {
            ex = min_image(ex, period[0]);
            ey = min_image(ey, period[1]);
            ez = min_image(ez, period[2]);
        }
        const double h = -area[k] / (2.0 * sqrt(ex * ex + ey * ey + ez * ez));
        val[k] = h;
        d -= h;
    }
    diag[i] = d;
}

// y_i = diag_i x_i + sum_k val_k x_col_k
VGC_SD double ot_spmv_row(int i, const long long* __restrict__ ptr, const int* __restrict__ col,
                          const double* __restrict__ val, const double* __restrict__ diag,
                          const double* __restrict__ x) {
    double s = diag[i] * x[i];
    for (long long k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
    return s;
}

// The same product with kSpmvLanes lanes per row (coalesced reads of the
// row's entries): lane l sums entries ptr_i + l, + kSpmvLanes, ... in order;
// the lanes are combined by a butterfly over offsets 8, 4, 2, 1, then the
// diagonal term is added.  The CPU emulation follows the same order.
constexpr int kSpmvLanes = 16;

// Scalars per instance of the batched PCG (device resident): [0] r.z,
// [1] r.r, [2] p.q, [3] alpha, [4] beta, [5] tol^2 b.b, [6] stop (1
// converged, 2 breakdown), [7] iterations, [8 + q] r.z of the direction of
// the iterations of parity q, [10] beta is current.
constexpr int kCgStride = 16;

VGC_SD double ot_spmv_lane(int i, int l, const long long* __restrict__ ptr,
                           const int* __restrict__ col, const double* __restrict__ val,
                           const double* __restrict__ x) {
    double s = 0.0;
    for (long long k = ptr[i] + l; k < ptr[i + 1]; k += kSpmvLanes) s += val[k] * x[col[k]];
    return s;
}


// ------------------------------------------------- batched (segmented) solves
// Instance k holds the caller-order range [off[k], off[k+1]).  Reductions per
// instance run over a fixed block table: instance k owns cnt[k] blocks
// (first[k] ... first[k] + cnt[k] - 1); thread t of its block b takes the
// elements off[k] + b * 256 + t + j * cnt[k] * 256 in order, the block sums
// by a halving tree, and one block per instance adds its blocks' partials by
// the same tree: fixed order, bit-reproducible.
struct SegTable {
    int n_inst;
    int nblk;
    const long long* off;   // n_inst + 1
    const int* blk_inst;    // nblk
    const int* blk_idx;     // nblk
    const int* cnt;         // n_inst
    const int* first;       // n_inst
    const int* inst_of;     // caller-order element -> instance
};
constexpr int kSegMaxBlocks = 64;  // blocks per instance

// --------------------------------------------- convex start from a coarse plan
// With phi_i = |p_i|^2 - w_i, the cell of seed i is {x : phi_i - 2 x.p_i
// minimal}: it is non-empty exactly when (p_i, phi_i) is a vertex of the
// lower convex hull of the lifted seeds.  Sampling a strictly convex function
// phi therefore gives weights with no empty cell (cell i contains
// grad phi(p_i) / 2).  The start is
//     phi(p) = s log sum_j exp(piece_j(p) / s),
//     piece_j(p) = a_j + g_j.p + mu_j H_Rj(|p - b_j|^2),
//     H_R(r2) = r2 for r <= R, 2 R r - R^2 beyond (convex, linear far away),
// a soft maximum of convex pieces, hence convex; each piece is the tangent
// of a coarse transport plan's potential at a coarse seed plus a curvature
// term that spreads the fine seeds near that coarse seed over its coarse
// Laguerre cell.  kStartPiece doubles per piece: a, g[3], mu, b[3], R.
// Terms more than 40 s below the running maximum are skipped: each would
// add less than half an ulp to a sum that is at least 1, so the result is
// the same.
constexpr int kStartPiece = 9;

VGC_SD double ot_convex_start_row(const double* __restrict__ p, const double* __restrict__ piece,
                                  int m, double s) {
    const double px = p[0], py = p[1], pz = p[2];
    const double inv_s = 1.0 / s;
    double mx = -1.0e300, sum = 0.0;
    for (int j = 0; j < m; ++j) {
        const double* c = piece + (long long)kStartPiece * j;
        const double dx = px - c[5], dy = py - c[6], dz = pz - c[7];
        const double r2 = dx * dx + dy * dy + dz * dz, R = c[8];
        const double h = r2 <= R * R ? r2 : R * (2.0 * sqrt(r2) - R);
        const double v = c[0] + c[1] * px + c[2] * py + c[3] * pz + c[4] * h;
        if (v > mx) {
            sum = sum * exp((mx - v) * inv_s) + 1.0;
            mx = v;
        } else if (v > mx - 40.0 * s) {
            sum += exp((v - mx) * inv_s);
        }
    }
    return px * px + py * py + pz * pz - (mx + s * log(sum));
}

}  // namespace vgc

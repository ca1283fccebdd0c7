// Types and scalar helpers shared by the gfx950 kernels and their CPU
// emulation.
#pragma once

#include <stdint.h>

#if defined(VGC_EMU)
#include <cmath>
#else
#include <hip/hip_runtime.h>
#endif

#include "clip_cell.h"

namespace vgc {

// 16-bit quantisation per axis, bits interleaved: 48-bit Morton code; the
// instance id sits above it so instances sort into contiguous ranges.
VGC_SD uint64_t spread16(uint64_t v) {
    v &= 0xFFFFull;
    v = (v | (v << 16)) & 0x0000FF0000FFull;
    v = (v | (v << 8)) & 0x00F00F00F00Full;
    v = (v | (v << 4)) & 0x0C30C30C30C3ull;
    v = (v | (v << 2)) & 0x249249249249ull;
    return v;
}

VGC_SD uint64_t quant16(double x, double lo, double hi) {
    double t = (x - lo) / (hi - lo);
    if (!(t > 0.0)) t = 0.0;
    if (t >= 1.0) t = 0.99999999;
    return uint64_t(t * 65536.0);
}

VGC_SD uint64_t morton_key(double x, double y, double z, const double* box, uint64_t inst) {
    uint64_t qx = quant16(x, box[0], box[3]);
    uint64_t qy = quant16(y, box[1], box[4]);
    uint64_t qz = quant16(z, box[2], box[5]);
    return (inst << 48) | spread16(qx) | (spread16(qy) << 1) | (spread16(qz) << 2);
}

// --------------------------------------------------------------- kNN
// K nearest other seeds of each query (FP32 ranking), sorted by (d2, id);
// unfilled entries are -1.  Traversal: own bucket and enough Morton-adjacent
// buckets to hold K seeds first (tight radius), then the implicit tree from
// the root, left to right, pruning boxes beyond the current K-th distance.

template <int K>
VGC_SD bool knn_less(float d, int id, float bd, int bi) {
    return d < bd || (d == bd && id < bi);
}

template <int K>
VGC_SD void knn_insert(float (&bd)[K], int (&bi)[K], float d, int id) {
#pragma unroll
    for (int k = K - 1; k > 0; --k) {
        bool lp = knn_less<K>(d, id, bd[k - 1], bi[k - 1]);
        bool lc = knn_less<K>(d, id, bd[k], bi[k]);
        float nd = lp ? bd[k - 1] : (lc ? d : bd[k]);
        int ni = lp ? bi[k - 1] : (lc ? id : bi[k]);
        bd[k] = nd;
        bi[k] = ni;
    }
    bool l0 = knn_less<K>(d, id, bd[0], bi[0]);
    bd[0] = l0 ? d : bd[0];
    bi[0] = l0 ? id : bi[0];
}

template <int K>
VGC_SD void knn_scan_bucket(const pf4* pf, int p0, int n, int b, int self, float px,
                            float py, float pz, float (&bd)[K], int (&bi)[K]) {
    int s = p0 + b * kBucket;
    int e = s + kBucket;
    if (e > p0 + n) e = p0 + n;
    // the bucket's points loaded together, then taken in order
    pf4 a[kBucket];
#pragma unroll
    for (int k = 0; k < kBucket; ++k) a[k] = s + k < e ? pf[s + k] : pf4{};
#pragma unroll
    for (int k = 0; k < kBucket; ++k) {
        const int m = s + k;
        if (m >= e || m == self) continue;
        float d = dist2f(a[k], px, py, pz);
        if (knn_less<K>(d, m, bd[K - 1], bi[K - 1])) knn_insert<K>(bd, bi, d, m);
    }
}

// K nearest seeds of instance `inst` to the FP32 point (px, py, pz), seed
// `self` (a Morton position, or -1) excluded; the walk starts with the
// buckets around Morton position qpos (the query's place in the order).
template <int K>
VGC_SD void knn_query_pt(const pf4* pf, const bnode* nodes, const InstLayout& IL, int inst,
                         int qpos, float px, float py, float pz, int self, float (&bd)[K],
                         int (&bi)[K]) {
    const int p0 = IL.p0[inst], n = IL.n[inst], nb = IL.nb[inst], L = IL.L[inst];
    const bnode* T = nodes + IL.t0[inst];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        bd[k] = INFINITY;
        bi[k] = -1;
    }
    // Morton neighbours first: buckets bq, bq-1, bq+1, bq-2, ... until K
    // other seeds have been seen, so the tree walk starts with a finite radius
    const int bq = (qpos - p0) / kBucket;
    int blo = bq, bhi = bq;
    knn_scan_bucket<K>(pf, p0, n, bq, self, px, py, pz, bd, bi);
    int seen = imin(p0 + n, p0 + (bq + 1) * kBucket) - (p0 + bq * kBucket) - (self >= 0 ? 1 : 0);
    while (seen < K && (blo > 0 || bhi < nb - 1)) {
        if (blo > 0) {
            --blo;
            knn_scan_bucket<K>(pf, p0, n, blo, self, px, py, pz, bd, bi);
            seen += kBucket;
        }
        if (bhi < nb - 1) {
            ++bhi;
            knn_scan_bucket<K>(pf, p0, n, bhi, self, px, py, pz, bd, bi);
            seen += imin(p0 + n, p0 + (bhi + 1) * kBucket) - (p0 + bhi * kBucket);
        }
    }
    int node = 1;
    while (true) {
        float lim = bd[K - 1] * 1.00001f;
        bool inside = box_dist2(T[node], px, py, pz) <= lim;
        if (inside && node < L) {
            node = 2 * node;
            continue;
        }
        if (inside) {
            int b = node - L;
            if (b < nb && (b < blo || b > bhi))
                knn_scan_bucket<K>(pf, p0, n, b, self, px, py, pz, bd, bi);
        }
        while (node & 1) node >>= 1;
        if (node == 0) break;
        node += 1;
    }
}

template <int K>
VGC_SD void knn_query(const pf4* pf, const bnode* nodes, const InstLayout& IL, int q,
                      int* out, float* out_d2) {
    const int inst = IL.inst_of ? IL.inst_of[q] : 0;
    const pf4 pq = pf[q];
    float bd[K];
    int bi[K];
    knn_query_pt<K>(pf, nodes, IL, inst, q, pq.x, pq.y, pq.z, q, bd, bi);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        out[k] = bi[k];
        if (out_d2) out_d2[k] = bd[k];
    }
}

// ------------------------------------------------------- exact kNN
// Exact k nearest neighbours (FP64 distances) by FP32 radius lists: an FP32
// radius that holds at least kk points, every point within it listed, the
// list sorted by (FP64 squared distance, caller's index), and the kk-th
// entry certified nearer than any point left out.
//
// e: bound of the error of an FP32 difference vector against the exact one,
// from the FP32 rounding of coordinates of magnitude <= S:
// per axis 2 u S (two roundings) + u 2 S (the subtraction), u = 2^-24, so
// |v32 - v| <= sqrt(3) 4 u S (rounded up below).
VGC_SD double knn_vec_err(double S) { return 7.0 * 5.9604644775390625e-8 * S * (1.0 + 1e-6); }

// FP32 squared radius that holds every point whose FP64 squared distance is
// at most that of a point found at FP32 squared distance d32.
VGC_SD float knn_list_radius(float d32, double e) {
    const double g = 1.0 + 4.0 * 5.9604644775390625e-8;  // 1 + gamma_3 (FP32), rounded up
    const double a = sqrt(double(d32) * g) * (1.0 + 1e-6) + 2.5 * e;
    return f_up(a * a * g * (1.0 + 1e-6));
}

// Lower bound of the FP64 squared distance (as computed in FP64) of a point
// whose FP32 squared distance exceeds r32.
VGC_SD double knn_outside_bound(float r32, double e) {
    const double g = 1.0 + 4.0 * 5.9604644775390625e-8;
    double a = sqrt(double(r32) / g) - e;
    if (!(a > 0.0)) return 0.0;
    return a * a * (1.0 - 4e-15);
}

// FP32 squared radius of query t of the exact kNN, a first guess checked
// later by the count and the certification (which grow it when short): the
// kk-th FP32 squared distance of the 16 nearest points found by the tree
// walk, or beyond 16 the 16th scaled by twice the volume ratio (the factor
// two covers points spread on a surface, whose count grows as r^2).
// A walk that keeps the W <= 16 nearest (kk <= W) finds the same kk-th
// distance as one that keeps 16: the GPU takes W = 1 or 4 when kk allows.
constexpr int kKnnSeedK = 16;
template <int W = kKnnSeedK>
VGC_SD float knn_seed_radius_row(const pf4* pf, const bnode* nodes, const InstLayout& IL,
                                 const pf4* qf, const int* qpos, int t, int kk, double e) {
    float bd[W];
    int bi[W];
    const pf4 q = qf[t];
    knn_query_pt<W>(pf, nodes, IL, 0, qpos ? qpos[t] : t, q.x, q.y, q.z, -1, bd, bi);
    // entry kk - 1 by constant indices (a variable index would put the
    // arrays in scratch memory)
    const int want = (kk < W ? kk : W) - 1;
    float d = bd[0];
#pragma unroll
    for (int k = 1; k < W; ++k) d = k == want ? bd[k] : d;
    if (kk > W) d *= 2.0f * powf(float(kk) / float(W), 2.0f / 3.0f);
    if (!(d < INFINITY)) return INFINITY;
    return knn_list_radius(d, e);
}

// Listed entries of query t: caller's index and FP64 squared distance (as
// computed by Geom::distance2, a sum of three squared differences).
VGC_SD void knn_d64_row(const pt4* qp, const pt4* pts, const int* orig, int t, long long b,
                        long long e, const int* lm, int* li, double* ld) {
    const pt4 q = qp[t];
    for (long long k = b; k < e; ++k) {
        const int m = lm[k];
        const pt4 p = pts[m];
        const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        li[k] = orig[m];
        ld[k] = dx * dx + dy * dy + dz * dz;
    }
}

// Query t's sorted list [b, b + c): 1 when its kk-th entry is certified (or
// every point is listed) and the row is written, 0 when the radius must grow.
VGC_SD int knn_select_row(int t, long long b, int c, const double* ld, const int* li, float r32,
                          double e, int kk, int n, int self_rule, const int* qrow, int* out_idx,
                          double* out_d2) {
    if (c < kk && c < n) return 0;
    if (c < n && !(ld[b + kk - 1] < knn_outside_bound(r32, e))) return 0;
    const int kout = self_rule ? kk - 1 : kk;
    const long long row = (long long)qrow[t] * kout;
    int m = 0;
    if (!self_rule) {
        for (; m < kk; ++m) {
            out_idx[row + m] = li[b + m];
            out_d2[row + m] = ld[b + m];
        }
    } else {
        const int own = qrow[t];
        bool disconnected = false;
        for (int j = 0; j < kk && m < kout; ++j) {
            const int id = li[b + j];
            if (id == own) continue;
            if (ld[b + j] == 0.0) {
                if (id < own) {
                    disconnected = true;
                    break;
                }
                continue;
            }
            out_idx[row + m] = id;
            out_d2[row + m] = ld[b + j];
            ++m;
        }
        if (disconnected) m = 0;
    }
    for (; m < kout; ++m) {
        out_idx[row + m] = -1;
        out_d2[row + m] = -1.0;
    }
    return 1;
}

// radius of a query whose list must grow
VGC_SD float knn_grow(float r32) {
    const float g = 4.0f * r32;
    return g > 1e-30f ? g : 1e-30f;
}

// Lloyd update of seed i (CentroidalVoronoiTesselation::Lloyd_iterations,
// voronoi/CVT.cpp:154-161): the centroid mg / m when m > 1e-30.  wrap: a
// periodic box; the centroid is brought back into [lo, hi).
VGC_SD void lloyd_update_row(int i, double* xyz, const double* vol, const double* mg,
                             const double* wrap) {
    const double m = vol[i];
    if (!(m > 1e-30)) return;
    const double s = 1.0 / m;
    for (int a = 0; a < 3; ++a) {
        double c = s * mg[3 * (long long)i + a];
        if (wrap) {
            const double lo = wrap[a], hi = wrap[3 + a], L = hi - lo;
            c -= L * floor((c - lo) / L);
            if (c >= hi) c -= L;
            if (c < lo) c = lo;
        }
        xyz[3 * (long long)i + a] = c;
    }
}

// ------------------------------------------------------ periodic images
// Images of seed i for a periodic box: every shift s in {-1, 0, 1}^3, s != 0,
// whose image x + s L lies within delta of the box (per axis
// lo - delta <= x_a + s_a L_a <= hi + delta), in the order sx, sy, sz
// (outer to inner).  Count mode when out == nullptr.
VGC_SD int periodic_images(const double* __restrict__ xyz, const double* __restrict__ w, int i,
                           const double* __restrict__ box, double delta, double* out_xyz,
                           double* out_w, int* out_src) {
    const double x[3] = {xyz[3 * (long long)i], xyz[3 * (long long)i + 1], xyz[3 * (long long)i + 2]};
    int cnt = 0;
    for (int sx = -1; sx <= 1; ++sx)
        for (int sy = -1; sy <= 1; ++sy)
            for (int sz = -1; sz <= 1; ++sz) {
                if (sx == 0 && sy == 0 && sz == 0) continue;
                const int sh[3] = {sx, sy, sz};
                double y[3];
                bool in = true;
                for (int a = 0; a < 3; ++a) {
                    y[a] = x[a] + double(sh[a]) * (box[3 + a] - box[a]);
                    in = in && y[a] >= box[a] - delta && y[a] <= box[3 + a] + delta;
                }
                if (!in) continue;
                if (out_xyz) {
                    for (int a = 0; a < 3; ++a) out_xyz[3 * (long long)cnt + a] = y[a];
                    if (out_w) out_w[cnt] = w[i];
                    out_src[cnt] = i;
                }
                ++cnt;
            }
    return cnt;
}

// ------------------------------------------------- power-nearest queries
// Seed of instance 0 nearest to x (FP64) in power distance |x - p|^2 - w
// (ties: the smaller Morton position); returns its Morton position.  Tree
// nodes are skipped when the FP32 lower bound box_dist2 - wmax exceeds the
// best FP64 value by more than a margin covering the FP32 roundings (cs2 =
// squared coordinate scale); seeds are compared in FP64.
VGC_SD int power_nearest_query(const pt4* pts, const bnode* nodes, const InstLayout& IL, double x,
                               double y, double z, float cs2) {
    const int p0 = IL.p0[0], n = IL.n[0], nb = IL.nb[0], L = IL.L[0];
    const bnode* T = nodes + IL.t0[0];
    const float px = float(x), py = float(y), pz = float(z);
    double best = INFINITY;
    int bi = -1;
    int node = 1;
    while (true) {
        const bnode& B = T[node];
        const float lb = box_dist2(B, px, py, pz) - B.wmax;
        const float bf = float(best);
        const bool inside =
            bi < 0 || lb <= bf + 1e-5f * (fabsf(lb) + fabsf(bf)) + 1e-6f * cs2;
        if (inside && node < L) {
            node = 2 * node;
            continue;
        }
        if (inside) {
            const int b = node - L;
            if (b < nb) {
                const int e = imin(p0 + n, p0 + (b + 1) * kBucket);
                for (int m = p0 + b * kBucket; m < e; ++m) {
                    const pt4 q = pts[m];
                    const double dx = x - q.x, dy = y - q.y, dz = z - q.z;
                    const double d = dx * dx + dy * dy + dz * dz - q.w;
                    if (d < best || (d == best && m < bi)) {
                        best = d;
                        bi = m;
                    }
                }
            }
        }
        while (node & 1) node >>= 1;
        if (node == 0) break;
        node += 1;
    }
    return bi;
}

// ------------------------------------------ cells over the plane limit
// Such a cell is computed in pieces, its parts in boxes that tile its
// instance's box: first the eight octants around the seed, then, for a piece
// still over the limit, the eight halves-of-edges sub-boxes of its box, and
// so on.  Each piece is clipped by the cell's complete neighbour list;
// volumes, first and second moments add up over the pieces and so do the
// facet areas of one neighbour.
constexpr int kPieces = 8;

// the octant boxes of cell t: items 8t .. 8t + 7
VGC_SD void piece_setup_row(int q, const int* cells, const pt4* pts, const int* inst_of,
                            const double* box, double* p_box) {
    const int t = q / kPieces, k = q % kPieces;
    const int c = cells[t];
    const double* b = box + 6 * (inst_of ? inst_of[c] : 0);
    const pt4 p = pts[c];
    const double pc[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
        const double s = fmin(fmax(pc[a], b[a]), b[3 + a]);  // the seed, inside the box
        const bool upper = (k >> a) & 1;
        p_box[6 * (size_t)q + a] = upper ? s : b[a];
        p_box[6 * (size_t)q + 3 + a] = upper ? b[3 + a] : s;
    }
}

// cell t from its finished pieces pidx[pptr[t] .. pptr[t+1]) (in that
// order): sums, and the union of the neighbour lists with the areas of a
// neighbour added, into x_nbr / x_area from x_off[t]
VGC_SD void piece_merge_row(int t, const int* cells, const int* orig, const long long* pptr,
                            const int* pidx, const double* p_vol, const double* p_mg,
                            const double* p_e2, const int* p_cnt, const int* p_nbr,
                            const double* p_area, const uint32_t* p_flags, int stride,
                            const long long* x_off, double* vol, double* mg, double* e2,
                            int* nbr_cnt, uint32_t* flags, int* x_nbr, double* x_area) {
    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, e = 0.0;
    uint32_t any = 0;
    int cnt = 0;
    int* xn = x_nbr + x_off[t];
    double* xa = x_area + x_off[t];
    for (long long u = pptr[t]; u < pptr[t + 1]; ++u) {
        const int q = pidx[u];
        v += p_vol[q];
        gx += p_mg[3 * (size_t)q];
        gy += p_mg[3 * (size_t)q + 1];
        gz += p_mg[3 * (size_t)q + 2];
        if (p_e2) e += p_e2[q];
        any |= p_flags[q];
        for (int m = 0; m < p_cnt[q]; ++m) {
            const int j = p_nbr[(size_t)q * stride + m];
            const double a = p_area[(size_t)q * stride + m];
            int k = 0;
            while (k < cnt && xn[k] != j) ++k;
            if (k == cnt) {
                xn[cnt] = j;
                xa[cnt] = a;
                ++cnt;
            } else {
                xa[k] += a;
            }
        }
    }
    const int c = cells[t], i = orig[c];
    vol[i] = v;
    mg[3 * (size_t)i] = gx;
    mg[3 * (size_t)i + 1] = gy;
    mg[3 * (size_t)i + 2] = gz;
    if (e2) e2[i] = e;
    nbr_cnt[i] = cnt;
    flags[c] = CELL_DONE | (v > 0.0 ? 0u : (uint32_t)CELL_EMPTY) |
               (any & (uint32_t)(CELL_UNCERTAIN | CELL_DISCONNECTED));
}

// the merged neighbour list of piece-computed cell t into the CSR
VGC_SD void piece_copy_row(int t, const int* cells, const int* orig, const long long* x_off,
                           const int* x_nbr, const double* x_area, const int* nbr_cnt,
                           const long long* ptr, int* nbr, double* area) {
    const int i = orig[cells[t]];
    const long long o = ptr[i];
    for (int m = 0; m < nbr_cnt[i]; ++m) {
        nbr[o + m] = x_nbr[x_off[t] + m];
        area[o + m] = x_area[x_off[t] + m];
    }
}

// Admissibility repair of a transport start: an empty cell i gets the
// smallest weight that puts its seed inside its cell with margin delta,
// w_i = max_j (w_j - |p_i - p_j|^2) + delta (j the power-nearest seed of
// p_i, which is not i since p_i lies outside its cell); other weights are
// kept.  pts / nodes: the tree built for the weights w.
VGC_SD double ot_repair_row(int i, const double* xyz, const double* w, const double* vol,
                            const pt4* pts, const bnode* nodes, const InstLayout& IL, float cs2,
                            double delta) {
    if (vol[i] > 0.0) return w[i];
    const double x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
    const int j = power_nearest_query(pts, nodes, IL, x, y, z, cs2);
    const pt4 q = pts[j];
    const double dx = x - q.x, dy = y - q.y, dz = z - q.z;
    const double v = q.w - (dx * dx + dy * dy + dz * dz) + delta;
    return v > w[i] ? v : w[i];
}

// ------------------------------------------------------ radius queries
// Does the subtree of box B hold seeds within r2 of p (FP32 squared
// distance) that may pass the power-distance filter F?
VGC_SD bool radius_node_in(const bnode& B, float px, float py, float pz, float r2,
                           const PowerFilter* F) {
    bool inside = box_dist2(B, px, py, pz) <= r2;
    if (inside && F && F->on) {
        const float fx = fmaxf(px - B.lo[0], B.hi[0] - px);
        const float fy = fmaxf(py - B.lo[1], B.hi[1] - py);
        const float fz = fmaxf(pz - B.lo[2], B.hi[2] - pz);
        const float dfar = sqrtf(fx * fx + fy * fy + fz * fz);
        inside = pf_node_may_cut(*F, box_dist2(B, F->cx, F->cy, F->cz), B.wmax,
                                 sqrtf(box_dist2(B, px, py, pz)), dfar, fx, fy, fz);
    }
    return inside;
}

// Is seed a within r2 of p, and does it pass F?  d: its squared distance.
VGC_SD bool radius_seed_take(const pf4& a, float px, float py, float pz, float r2,
                             const PowerFilter* F, float& d) {
    d = dist2f(a, px, py, pz);
    bool take = d <= r2;
    if (take && F && F->on)
        take = pf_may_cut(*F, dist2f(a, F->cx, F->cy, F->cz), a.w, sqrtf(d), fabsf(a.x - px),
                          fabsf(a.y - py), fabsf(a.z - pz));
    return take;
}

// Seeds (other than q) with FP32 squared distance <= r2 and, with a filter
// F, that pass the power-distance test (types.h), in traversal order
// (depth first, children in order: increasing Morton position).  Count
// mode when out == nullptr.  With budget > 0, a walk that visits more nodes
// returns -1 (the caller hands the query to a wider walk).
VGC_SD int radius_query_pt(const pf4* pf, const bnode* nodes, const InstLayout& IL, int inst,
                           float px, float py, float pz, int self, float r2, const PowerFilter* F,
                           int* out, float* out_d2, int budget = 0) {
    const int p0 = IL.p0[inst], n = IL.n[inst], nb = IL.nb[inst], L = IL.L[inst];
    const bnode* T = nodes + IL.t0[inst];
    int cnt = 0;
    int node = 1;
    int visits = 0;
    while (true) {
        if (budget > 0 && ++visits > budget) return -1;
        const bool inside = radius_node_in(T[node], px, py, pz, r2, F);
        if (inside && node < L) {
            node = 2 * node;
            continue;
        }
        if (inside) {
            int b = node - L;
            if (b < nb) {
                int s = p0 + b * kBucket;
                int e = s + kBucket;
                if (e > p0 + n) e = p0 + n;
                // the bucket's seeds loaded together, then tested in order
                pf4 a[kBucket];
#pragma unroll
                for (int k = 0; k < kBucket; ++k) a[k] = s + k < e ? pf[s + k] : pf4{};
#pragma unroll
                for (int k = 0; k < kBucket; ++k) {
                    const int m = s + k;
                    if (m >= e || m == self) continue;
                    float d;
                    if (radius_seed_take(a[k], px, py, pz, r2, F, d)) {
                        if (out) {
                            out[cnt] = m;
                            out_d2[cnt] = d;
                        }
                        ++cnt;
                    }
                }
            }
        }
        while (node & 1) node >>= 1;
        if (node == 0) break;
        node += 1;
    }
    return cnt;
}

VGC_SD int radius_query(const pf4* pf, const bnode* nodes, const InstLayout& IL, int q,
                        float r2, const PowerFilter* F, int* out, float* out_d2, int budget = 0) {
    const int inst = IL.inst_of ? IL.inst_of[q] : 0;
    const pf4 pq = pf[q];
    return radius_query_pt(pf, nodes, IL, inst, pq.x, pq.y, pq.z, q, r2, F, out, out_d2, budget);
}

// ---------------------------------------------------------- tree build
VGC_SD void bvh_leaf(const pt4* pts, const InstLayout& IL, bnode* nodes, int inst, int leaf) {
    const int p0 = IL.p0[inst], n = IL.n[inst], nb = IL.nb[inst], L = IL.L[inst];
    bnode B;
    B.lo[0] = B.lo[1] = B.lo[2] = INFINITY;
    B.hi[0] = B.hi[1] = B.hi[2] = -INFINITY;
    B.wmax = -INFINITY;
    B.pad = 0;
    if (leaf < nb) {
        int s = p0 + leaf * kBucket;
        int e = s + kBucket;
        if (e > p0 + n) e = p0 + n;
        double lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY;
        double hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY, wm = -INFINITY;
        for (int m = s; m < e; ++m) {
            const pt4 p = pts[m];
            wm = fmax(wm, p.w);
            lo0 = fmin(lo0, p.x);
            lo1 = fmin(lo1, p.y);
            lo2 = fmin(lo2, p.z);
            hi0 = fmax(hi0, p.x);
            hi1 = fmax(hi1, p.y);
            hi2 = fmax(hi2, p.z);
        }
        // the float coordinates used by the search are round-to-nearest of
        // the doubles, so the outward-rounded box contains them
        B.lo[0] = fminf(f_down(lo0), float(lo0));
        B.lo[1] = fminf(f_down(lo1), float(lo1));
        B.lo[2] = fminf(f_down(lo2), float(lo2));
        B.hi[0] = fmaxf(f_up(hi0), float(hi0));
        B.hi[1] = fmaxf(f_up(hi1), float(hi1));
        B.hi[2] = fmaxf(f_up(hi2), float(hi2));
        B.wmax = fmaxf(f_up(wm), float(wm));
    }
    nodes[IL.t0[inst] + L + leaf] = B;
}

VGC_SD void bvh_merge(bnode* nodes, long long t0, int k) {
    const bnode a = nodes[t0 + 2 * k], b = nodes[t0 + 2 * k + 1];
    bnode B;
    for (int c = 0; c < 3; ++c) {
        B.lo[c] = fminf(a.lo[c], b.lo[c]);
        B.hi[c] = fmaxf(a.hi[c], b.hi[c]);
    }
    B.wmax = fmaxf(a.wmax, b.wmax);
    B.pad = 0;
    nodes[t0 + k] = B;
}

}  // namespace vgc

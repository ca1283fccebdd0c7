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
    for (int m = s; m < e; ++m) {
        if (m == self) continue;
        float d = dist2f(pf[m], px, py, pz);
        if (knn_less<K>(d, m, bd[K - 1], bi[K - 1])) knn_insert<K>(bd, bi, d, m);
    }
}

template <int K>
VGC_SD void knn_query(const pf4* pf, const bnode* nodes, const InstLayout& IL, int q,
                      int* out, float* out_d2) {
    const int inst = IL.inst_of ? IL.inst_of[q] : 0;
    const int p0 = IL.p0[inst], n = IL.n[inst], nb = IL.nb[inst], L = IL.L[inst];
    const bnode* T = nodes + IL.t0[inst];
    const pf4 pq = pf[q];
    const float px = pq.x, py = pq.y, pz = pq.z;
    float bd[K];
    int bi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        bd[k] = INFINITY;
        bi[k] = -1;
    }
    // Morton neighbours first: buckets bq, bq-1, bq+1, bq-2, ... until K
    // other seeds have been seen, so the tree walk starts with a finite radius
    const int bq = (q - p0) / kBucket;
    int blo = bq, bhi = bq;
    knn_scan_bucket<K>(pf, p0, n, bq, q, px, py, pz, bd, bi);
    int seen = imin(p0 + n, p0 + (bq + 1) * kBucket) - (p0 + bq * kBucket) - 1;
    while (seen < K && (blo > 0 || bhi < nb - 1)) {
        if (blo > 0) {
            --blo;
            knn_scan_bucket<K>(pf, p0, n, blo, q, px, py, pz, bd, bi);
            seen += kBucket;
        }
        if (bhi < nb - 1) {
            ++bhi;
            knn_scan_bucket<K>(pf, p0, n, bhi, q, px, py, pz, bd, bi);
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
                knn_scan_bucket<K>(pf, p0, n, b, q, px, py, pz, bd, bi);
        }
        while (node & 1) node >>= 1;
        if (node == 0) break;
        node += 1;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        out[k] = bi[k];
        if (out_d2) out_d2[k] = bd[k];
    }
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

// ------------------------------------------------------ radius queries
// Seeds (other than q) with FP32 squared distance <= r2 and, with a filter
// F, that pass the power-distance test (types.h), in traversal order.  Count
// mode when out == nullptr.
VGC_SD int radius_query(const pf4* pf, const bnode* nodes, const InstLayout& IL, int q,
                        float r2, const PowerFilter* F, int* out, float* out_d2) {
    const int inst = IL.inst_of ? IL.inst_of[q] : 0;
    const int p0 = IL.p0[inst], n = IL.n[inst], nb = IL.nb[inst], L = IL.L[inst];
    const bnode* T = nodes + IL.t0[inst];
    const pf4 pq = pf[q];
    const float px = pq.x, py = pq.y, pz = pq.z;
    const bool filt = F && F->on;
    int cnt = 0;
    int node = 1;
    while (true) {
        const bnode B = T[node];
        bool inside = box_dist2(B, px, py, pz) <= r2;
        if (inside && filt) {
            const float fx = fmaxf(px - B.lo[0], B.hi[0] - px);
            const float fy = fmaxf(py - B.lo[1], B.hi[1] - py);
            const float fz = fmaxf(pz - B.lo[2], B.hi[2] - pz);
            const float dfar = sqrtf(fx * fx + fy * fy + fz * fz);
            inside = pf_node_may_cut(*F, box_dist2(B, F->cx, F->cy, F->cz), B.wmax,
                                     sqrtf(box_dist2(B, px, py, pz)), dfar, fx, fy, fz);
        }
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
                for (int m = s; m < e; ++m) {
                    if (m == q) continue;
                    const pf4 a = pf[m];
                    float d = dist2f(a, px, py, pz);
                    bool take = d <= r2;
                    if (take && filt)
                        take = pf_may_cut(*F, dist2f(a, F->cx, F->cy, F->cz), a.w, sqrtf(d),
                                          fabsf(a.x - px), fabsf(a.y - py), fabsf(a.z - pz));
                    if (take) {
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

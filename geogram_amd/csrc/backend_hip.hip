// gfx950 backend of the cell pipeline and the exported C ABI (libvgc.so).
//
// Kernels:
//   k_morton / rocPRIM radix sort / k_gather   Morton order of the seeds
//   k_bvh_leaves, k_bvh_level                  implicit tree over buckets
//   k_knn<K>                                   lane-per-query kNN, FP32 ranking
//   k_clip<S>                                  wavefront-per-cell clipping and
//                                              FP64 integration (clip_cell.h)
//   k_radius_*                                 complete lists for pass B / C
//   k_compact                                  neighbour CSR
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <stdexcept>
#include <string>

#include "api_impl.h"
#include "mlp.h"
#include "ot_kernels.h"

#define VGC_HIP_CHECK(x)                                                              \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                     " at " #x);                                      \
    } while (0)

namespace vgc {

// ----------------------------------------------------------------- kernels

__device__ __forceinline__ int find_inst(const long long* off, int ni, long long t) {
    int lo = 0, hi = ni - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void k_fill_inst(int* inst_orig, int n, const int* p0, int ni) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = ni - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (p0[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    inst_orig[i] = lo;
}

__global__ void k_morton(const double* __restrict__ xyz, const int* __restrict__ inst_orig,
                         const double* __restrict__ box, int n, uint64_t* keys, int* vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int inst = inst_orig ? inst_orig[i] : 0;
    keys[i] = morton_key(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2],
                         box + 6 * inst, (uint64_t)inst);
    vals[i] = i;
}

__global__ void k_gather(const double* __restrict__ xyz, const double* __restrict__ w,
                         const int* __restrict__ inst_orig, const int* __restrict__ perm, int n,
                         pt4* pts, pf4* pf, int* orig, int* inst_of) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int i = perm[s];
    pt4 p;
    p.x = xyz[3 * (size_t)i];
    p.y = xyz[3 * (size_t)i + 1];
    p.z = xyz[3 * (size_t)i + 2];
    p.w = w ? w[i] : 0.0;
    pts[s] = p;
    pf4 f;
    f.x = float(p.x);
    f.y = float(p.y);
    f.z = float(p.z);
    f.w = float(p.w);
    pf[s] = f;
    orig[s] = i;
    if (inst_of) inst_of[s] = inst_orig ? inst_orig[i] : 0;
}

__global__ void k_bvh_leaves(const pt4* __restrict__ pts, InstLayout IL, bnode* nodes,
                             const long long* __restrict__ leaf_off, int ni, long long total) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int inst = find_inst(leaf_off, ni, t);
    bvh_leaf(pts, IL, nodes, inst, (int)(t - leaf_off[inst]));
}

__global__ void k_bvh_level(InstLayout IL, bnode* nodes, const long long* __restrict__ lvl_off,
                            int ni, long long total, int h) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int inst = find_inst(lvl_off, ni, t);
    int k = (IL.L[inst] >> h) + (int)(t - lvl_off[inst]);
    bvh_merge(nodes, IL.t0[inst], k);
}

__global__ void k_wmax(const pt4* __restrict__ pts, InstLayout IL, double* out) {
    __shared__ double red[256];
    int inst = blockIdx.x;
    int p0 = IL.p0[inst], n = IL.n[inst];
    double m = -INFINITY;
    for (int s = p0 + threadIdx.x; s < p0 + n; s += blockDim.x) m = fmax(m, pts[s].w);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[inst] = red[0];
}

template <int K>
__global__ __launch_bounds__(256) void k_knn(const pf4* __restrict__ pf,
                                             const bnode* __restrict__ nodes, InstLayout IL,
                                             const int* __restrict__ queries, int nq, int* out,
                                             float* out_d2) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    int q = queries ? queries[t] : t;
    knn_query<K>(pf, nodes, IL, q, out + (size_t)t * K, out_d2 ? out_d2 + (size_t)t * K : nullptr);
}

// Work-groups go to the 8 XCDs of the chip in turn; renumber them so that
// each XCD takes one contiguous range of work-groups (cells in Morton order:
// neighbouring cells read the same seeds and tree nodes, from that XCD's L2).
// A bijection of [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// S = 1 with FP64 signs (every cell's first pass, and the second gather):
// registers capped at 128 for 4 waves per SIMD (some spill to scratch); with
// one-wave work-groups that measured 7 % faster than 168 registers and 3
// waves — the wait for dependent LDS and memory round trips is hidden by
// occupancy.  The exact re-clip (a few cells) keeps its registers.
// F32: the first pass with FP32 corner state (the cells it cannot certify
// are left to the FP64 kernel).
template <int S, int WAVES, bool EXACT, bool GATHER, bool W, bool CORNER = W, bool F32 = false>
__global__ __launch_bounds__(64 * WAVES, (S == 1 && !EXACT) ? 4 : 1) void k_clip(ClipParams P) {
    __shared__ CellLds<S> lds[WAVES];
    const int wave = threadIdx.x >> 6;
    const int item = xcd_block() * WAVES + wave;
    if (item >= P.n_items) return;  // whole wavefront
    process_cell<S, EXACT, GATHER, W, CORNER, typename std::conditional<F32, float, double>::type>(
        P, item, lds[wave]);
}

__global__ void k_radius_r2(const int* __restrict__ items, int ni, const pt4* __restrict__ pts,
                            const double* __restrict__ r2, const int* __restrict__ inst_of,
                            const double* __restrict__ wmaxv, int weighted,
                            const fball* __restrict__ ball, const double* __restrict__ box,
                            float* rho2, PowerFilter* filt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni) return;
    int c = items[t];
    int inst = inst_of ? inst_of[c] : 0;
    rho2[t] = radius_for(r2[c], pts[c].w, weighted ? wmaxv[inst] : 0.0, weighted);
    if (filt) filt[t] = filter_from_ball(ball[c], pts[c], box + 6 * inst);
}

// Radius lists, one lane per query.  A walk over kRadiusBudget nodes (the
// ball of a cell with a very large weight) stops and flags its query
// (wide[t]) for k_radius_wide.
constexpr int kRadiusBudget = 512;
static_assert(kBucket == 16, "k_radius_wide takes four leaves of 16 seeds per step");
__global__ void k_radius_count(const pf4* __restrict__ pf, const bnode* __restrict__ nodes,
                               InstLayout IL, const int* __restrict__ items, int ni,
                               const float* __restrict__ rho2, const PowerFilter* __restrict__ filt,
                               int* cnt, unsigned char* wide) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni) return;
    const int c = radius_query(pf, nodes, IL, items[t], rho2[t], filt ? filt + t : nullptr,
                               nullptr, nullptr, kRadiusBudget);
    wide[t] = c < 0 ? 1 : 0;
    cnt[t] = c < 0 ? 0 : c;
}

__global__ void k_radius_fill(const pf4* __restrict__ pf, const bnode* __restrict__ nodes,
                              InstLayout IL, const int* __restrict__ items, int ni,
                              const float* __restrict__ rho2, const PowerFilter* __restrict__ filt,
                              const long long* __restrict__ off, long long base, int* list,
                              float* d2, unsigned char* wide) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni) return;
    const int c = radius_query(pf, nodes, IL, items[t], rho2[t], filt ? filt + t : nullptr,
                               list + (off[t] - base), d2 + (off[t] - base), kRadiusBudget);
    wide[t] = c < 0 ? 1 : 0;
}

// The flagged radius queries, one wavefront each (a grid of waves strides
// over the queries): the tree level by level, 64 nodes per step, each
// level's passing nodes kept in order in the wave's scratch (two buffers of
// fstride >= L nodes), then the seeds of the passing leaves, four leaves
// (64 seeds) per step.  A node is expanded exactly when the depth-first
// walk of radius_query_pt expands it, by the same tests, and the seeds come
// out in its order (leaves in order, seeds by Morton position): count mode
// (list == nullptr) gives its count, fill mode its list, entry for entry.
__global__ __launch_bounds__(64) void k_radius_wide(const pf4* __restrict__ pf,
                                                    const bnode* __restrict__ nodes, InstLayout IL,
                                                    const int* __restrict__ items, int ni,
                                                    const float* __restrict__ rho2,
                                                    const PowerFilter* __restrict__ filt,
                                                    const unsigned char* __restrict__ wide,
                                                    int* front, long long fstride, int* cnt,
                                                    const long long* __restrict__ off,
                                                    long long base, int* list, float* d2) {
    const int lane = threadIdx.x;
    int* fa = front + (long long)blockIdx.x * 2 * fstride;
    int* fb = fa + fstride;
    for (int t = blockIdx.x; t < ni; t += gridDim.x) {
        if (!wide[t]) continue;  // uniform over the wave
        const int q = items[t];
        const int inst = IL.inst_of ? IL.inst_of[q] : 0;
        const pf4 pq = pf[q];
        const float r2 = rho2[t];
        const PowerFilter* F = filt ? filt + t : nullptr;
        const int p0 = IL.p0[inst], n = IL.n[inst], nb = IL.nb[inst], L = IL.L[inst];
        const bnode* T = nodes + IL.t0[inst];
        int* out = list ? list + (off[t] - base) : nullptr;
        float* od = list ? d2 + (off[t] - base) : nullptr;
        int* A = fa;
        int* B = fb;
        if (lane == 0) A[0] = 1;
        __syncthreads();
        int na = 1, c = 0;
        for (int first = 1; na > 0; first *= 2) {
            if (first < L) {
                // internal level: the passing nodes' children, in order
                int nn = 0;
                for (int i0 = 0; i0 < na; i0 += 64) {
                    const int i = i0 + lane;
                    const int node = i < na ? A[i] : 0;
                    const bool pass = i < na && radius_node_in(T[node], pq.x, pq.y, pq.z, r2, F);
                    const uint64_t m = __ballot(pass);
                    const int r = nn + 2 * mbcnt(m);
                    if (pass) {
                        B[r] = 2 * node;
                        B[r + 1] = 2 * node + 1;
                    }
                    nn += 2 * __popcll(m);
                }
                __syncthreads();
                int* tmp = A;
                A = B;
                B = tmp;
                na = nn;
                continue;
            }
            // leaf level
            for (int i0 = 0; i0 < na; i0 += 64) {
                const int i = i0 + lane;
                const int node = i < na ? A[i] : 0;
                const bool pass =
                    i < na && radius_node_in(T[node], pq.x, pq.y, pq.z, r2, F) && node - L < nb;
                uint64_t lm = __ballot(pass);
                while (lm) {
                    // the next four passing leaves, 16 lanes each
                    int sel = -1;
                    for (int g = 0; g < 4 && lm; ++g) {
                        const int bit = __builtin_ctzll(lm);
                        lm &= lm - 1;
                        if ((lane >> 4) == g) sel = bit;
                    }
                    const int leaf0 = __shfl(node, sel >= 0 ? sel : 0);  // every lane takes part
                    const int leaf = sel >= 0 ? leaf0 : 0;
                    const int mseed = p0 + (leaf - L) * kBucket + (lane & (kBucket - 1));
                    const bool ok = sel >= 0 && mseed < p0 + n && mseed != q;
                    float d = 0.0f;
                    const bool take = ok && radius_seed_take(pf[mseed], pq.x, pq.y, pq.z, r2, F, d);
                    const uint64_t tm = __ballot(take);
                    if (out && take) {
                        const int r = c + mbcnt(tm);
                        out[r] = mseed;
                        od[r] = d;
                    }
                    c += __popcll(tm);
                }
            }
            break;
        }
        if (!list && lane == 0) cnt[t] = c;
        __syncthreads();  // the scratch is reused by the wave's next query
    }
}

__global__ void k_rebase(const long long* __restrict__ off, int n, long long base, long long* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = off[t] - base;
}

// 16 lanes per cell: each cell's row of the per-cell staging (stride
// entries) is read and its CSR range written in contiguous 64-byte runs
constexpr int kCompactLanes = 16;
__global__ void k_compact(int n, const int* __restrict__ cnt, const long long* __restrict__ ptr,
                          const int* __restrict__ nbr_tmp, const double* __restrict__ area_tmp,
                          int stride, int* nbr, double* area) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long i = t / kCompactLanes;
    if (i >= n) return;
    const long long o = ptr[i];
    const int m = min(cnt[i], stride);  // longer lists (pieces) are copied on their own
    const size_t row = (size_t)i * stride;
    for (int k = int(t % kCompactLanes); k < m; k += kCompactLanes) {
        nbr[o + k] = nbr_tmp[row + k];
        area[o + k] = area_tmp[row + k];
    }
}

__global__ void k_lloyd_update(double* xyz, const double* __restrict__ vol,
                               const double* __restrict__ mg, int n,
                               const double* __restrict__ wrap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lloyd_update_row(i, xyz, vol, mg, wrap);
}

__global__ void k_ghost_count(const double* __restrict__ xyz, int n, const double* __restrict__ box,
                              double delta, int* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = periodic_images(xyz, nullptr, i, box, delta, nullptr, nullptr, nullptr);
}

// seeds of the period first (caller order), then the images of seed i at
// n + off[i]
__global__ void k_ghost_fill(const double* __restrict__ xyz, const double* __restrict__ w, int n,
                             const double* __restrict__ box, double delta,
                             const long long* __restrict__ off, double* xe, double* we, int* src) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < 3; ++a) xe[3 * (long long)i + a] = xyz[3 * (long long)i + a];
    if (we) we[i] = w[i];
    src[i] = i;
    const long long o = n + off[i];
    periodic_images(xyz, w, i, box, delta, xe + 3 * o, we ? we + o : nullptr, src + o);
}

// largest squared radius (FP32, rounded up) within which a seed could still
// cut a cell of the period (radius_for of its squared radius r2)
__global__ void k_periodic_reach(const pt4* __restrict__ pts, const int* __restrict__ orig, int ne,
                                 int n_real, const double* __restrict__ r2,
                                 const double* __restrict__ wmax, float* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    float v = 0.0f;
    if (c < ne && orig[c] < n_real)
        v = radius_for(r2[c], pts[c].w, wmax ? wmax[0] : 0.0, wmax ? 1 : 0);
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m));
    if (__lane_id() == 0 && v > 0.0f) atomicMax((unsigned*)out, __float_as_uint(v));
}

// ------------------------------------------------- surface restricted cells
__global__ void k_surface_items(const double* __restrict__ V, const int* __restrict__ F,
                                const double* __restrict__ vw, int nt, double* ov, double* ow) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) surface_items_row(V, F, vw, t, ov, ow);
}

__global__ void k_surface_v0(const double* __restrict__ tv, int n, double* q) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        for (int a = 0; a < 3; ++a) q[3 * (size_t)t + a] = tv[9 * (size_t)t + a];
}

__global__ __launch_bounds__(256) void k_surface_pass(SurfParams P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < P.n_items) surface_item(P, t);
}

// the split items' children (4 each), then the redone items unchanged
__global__ void k_surface_next(const double* __restrict__ tv, const double* __restrict__ tw,
                               const int* __restrict__ lst, int nsplit, int nredo, double* ov,
                               double* ow) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nsplit) {
        surface_split_row(tv, tw, lst[c], c, ov, ow);
    } else if (c < nsplit + nredo) {
        const int t = lst[c];
        const size_t o = 4 * (size_t)nsplit + (size_t)(c - nsplit);
        for (int a = 0; a < 9; ++a) ov[9 * o + a] = tv[9 * (size_t)t + a];
        if (ow)
            for (int a = 0; a < 3; ++a) ow[3 * o + a] = tw[3 * (size_t)t + a];
    }
}

__global__ void k_surface_reduce(int n, const unsigned long long* __restrict__ keys,
                                 const int* __restrict__ idx, long long nrec,
                                 const double* __restrict__ val, const int* __restrict__ flag,
                                 double* m, double* mg) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) surface_reduce_row(s, keys, idx, nrec, val, flag, m, mg);
}

__global__ void k_knn_to_caller(const int* __restrict__ orig, const int* __restrict__ list,
                                const float* __restrict__ d2, int n, int K, int* out, float* outd) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int i = orig[q];
    for (int k = 0; k < K; ++k) {
        int j = list[(size_t)q * K + k];
        out[(size_t)i * K + k] = j >= 0 ? orig[j] : -1;
        outd[(size_t)i * K + k] = d2[(size_t)q * K + k];
    }
}

// ------------------------------------------------------- exact kNN kernels
__global__ void k_iota(int* d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = i;
}

// first position of each query key in the sorted keys (clamped to n - 1)
__global__ void k_lower_bound_u64(const uint64_t* __restrict__ keys, int n,
                                  const uint64_t* __restrict__ q, int nq, int* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const uint64_t k = q[t];
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[t] = lo < n ? lo : n - 1;
}

template <int W>
__global__ __launch_bounds__(256) void k_knn_seed_radius(const pf4* __restrict__ pf,
                                                         const bnode* __restrict__ nodes,
                                                         InstLayout IL, const pf4* __restrict__ qf,
                                                         const int* __restrict__ qpos, int nq,
                                                         int kk, double e, float* rr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nq) rr[t] = knn_seed_radius_row<W>(pf, nodes, IL, qf, qpos, t, kk, e);
}

__global__ void k_knn_count(const pf4* __restrict__ pf, const bnode* __restrict__ nodes,
                            InstLayout IL, const pf4* __restrict__ qf, const int* __restrict__ items,
                            int ni, const float* __restrict__ rr, int* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int t = items[i];
    const pf4 q = qf[t];
    cnt[i] = radius_query_pt(pf, nodes, IL, 0, q.x, q.y, q.z, -1, rr[t], nullptr, nullptr, nullptr);
}

__global__ void k_knn_fill(const pf4* __restrict__ pf, const bnode* __restrict__ nodes,
                           InstLayout IL, const pf4* __restrict__ qf, const int* __restrict__ items,
                           int ni, const float* __restrict__ rr, const long long* __restrict__ off,
                           int* lm, float* lf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int t = items[i];
    const pf4 q = qf[t];
    radius_query_pt(pf, nodes, IL, 0, q.x, q.y, q.z, -1, rr[t], nullptr, lm + off[i], lf + off[i]);
}

__global__ void k_knn_d64(const pt4* __restrict__ qp, const pt4* __restrict__ pts,
                          const int* __restrict__ orig, const int* __restrict__ items, int ni,
                          const long long* __restrict__ off, const int* __restrict__ lm, int* li,
                          double* ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ni) knn_d64_row(qp, pts, orig, items[i], off[i], off[i + 1], lm, li, ld);
}

__global__ void k_knn_select(const int* __restrict__ items, int ni,
                             const long long* __restrict__ off, const int* __restrict__ cnt,
                             const double* __restrict__ ld, const int* __restrict__ li, float* rr,
                             double e, int kk, int n, int self_rule, const int* __restrict__ qrow,
                             int* out_idx, double* out_d2, int* redo, unsigned* nredo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int t = items[i];
    if (!knn_select_row(t, off[i], cnt[i], ld, li, rr[t], e, kk, n, self_rule, qrow, out_idx,
                        out_d2)) {
        rr[t] = knn_grow(rr[t]);
        redo[atomicAdd(nredo, 1u)] = t;
    }
}

// ------------------------------------------------------------- OT kernels
// Deterministic reductions: kRedBlocks blocks of 256 threads; thread t of
// block b accumulates elements b*256+t + k*kRedBlocks*256 in order, the block
// reduces by halving strides, and one block reduces the partials the same way.
constexpr int kRedBlocks = 512;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(256) void k_dot_partial(int n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) s += a[i] * b[i];
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_sum_partial(int n, const double* __restrict__ a,
                                                     double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) s += a[i];
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// CVT gradient: g_i += 2 (m_i x_i - mg_i), caller order
__global__ void k_cvt_grad(int n, const double* __restrict__ xyz, const double* __restrict__ vol,
                           const double* __restrict__ mg, double* g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) {
        const size_t k = 3 * (size_t)i + c;
        g[k] += 2.0 * (vol[i] * xyz[k] - mg[k]);
    }
}

__global__ __launch_bounds__(256) void k_sum_final(const double* __restrict__ part, double* out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) s += part[i];
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) *out = t;
}

// [0] max |m-nu|/nu  [1] sum (m-nu)^2  [2] min m, per block then final
__global__ __launch_bounds__(256) void k_mass_partial(int n, const double* __restrict__ m,
                                                      const double* __restrict__ nu, double* part) {
    __shared__ double sh[256];
    double mx = 0.0, s2 = 0.0, mn = INFINITY;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) {
        double d = m[i] - nu[i];
        mx = fmax(mx, fabs(d) / nu[i]);
        s2 += d * d;
        mn = fmin(mn, m[i]);
    }
    double t = block_sum(s2, sh);
    __syncthreads();
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    double tmx = sh[0];
    __syncthreads();
    sh[threadIdx.x] = mn;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tmx;
        part[kRedBlocks + blockIdx.x] = t;
        part[2 * kRedBlocks + blockIdx.x] = sh[0];
    }
}

__global__ __launch_bounds__(256) void k_mass_final(const double* __restrict__ part, double* out) {
    __shared__ double sh[256];
    double mx = 0.0, s2 = 0.0, mn = INFINITY;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) {
        mx = fmax(mx, part[i]);
        s2 += part[kRedBlocks + i];
        mn = fmin(mn, part[2 * kRedBlocks + i]);
    }
    double t = block_sum(s2, sh);
    __syncthreads();
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    double tmx = sh[0];
    __syncthreads();
    sh[threadIdx.x] = mn;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = tmx;
        out[1] = t;
        out[2] = sh[0];
    }
}

__global__ void k_ot_hessian(int n, const double* __restrict__ xyz, const long long* __restrict__ ptr,
                             const int* __restrict__ col, const double* __restrict__ area,
                             double* val, double* diag, const double* __restrict__ period) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ot_hessian_row(i, xyz, ptr, col, area, val, diag, period);
}

// one thread per seed; every lane reads the same piece at a time (one
// broadcast load per piece and wave)
__global__ __launch_bounds__(256) void k_ot_convex_start(int n, const double* __restrict__ xyz,
                                                         const double* __restrict__ piece, int m,
                                                         double s, double* w) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = ot_convex_start_row(xyz + 3 * (long long)i, piece, m, s);
}

__global__ void k_ot_spmv(int n, const long long* __restrict__ ptr, const int* __restrict__ col,
                          const double* __restrict__ val, const double* __restrict__ diag,
                          const double* __restrict__ x, double* y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = ot_spmv_row(i, ptr, col, val, diag, x);
}

__global__ void k_ot_rhs(int n, const double* __restrict__ nu, const double* __restrict__ m,
                         double* b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = nu[i] - m[i];
}

// b -= sum(b) / n (sum from device memory): the Newton right-hand side
// projected on the range of the Laplacian (whose kernel is the constants)
__global__ void k_sub_mean(int n, double* b, const double* __restrict__ sum) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] -= *sum / double(n);
}

__global__ void k_ot_axpy(int n, const double* __restrict__ w, double t,
                          const double* __restrict__ d, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = w[i] + t * d[i];
}

__global__ void k_ot_pcg_init(int n, const double* __restrict__ b, const double* __restrict__ diag,
                              double* x, double* r, double* z, double* p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ri = b[i];
    double zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
    x[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
}

__global__ void k_ot_pcg_update(int n, double alpha, const double* __restrict__ p,
                                const double* __restrict__ q, const double* __restrict__ diag,
                                double* x, double* r, double* z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += alpha * p[i];
    double ri = r[i] - alpha * q[i];
    r[i] = ri;
    z[i] = diag[i] > 0.0 ? ri / diag[i] : 0.0;
}

__global__ void k_vscale(int n, double a, const double* __restrict__ x, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i];
}

__global__ void k_vaxpy(int n, double a, const double* __restrict__ x, double* y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

__global__ void k_ot_pcg_dir(int n, double beta, const double* __restrict__ z, double* p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(256) void k_min_partial(int n, const double* __restrict__ a, double* part) {
    __shared__ double sh[256];
    double mn = INFINITY;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) mn = fmin(mn, a[i]);
    sh[threadIdx.x] = mn;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_min_final(const double* __restrict__ part, double* out) {
    __shared__ double sh[256];
    double mn = INFINITY;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) mn = fmin(mn, part[i]);
    sh[threadIdx.x] = mn;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ------------------------------------------------------------- MLP kernels
constexpr int kMlpRowsFwd = 64;   // rows per wavefront, inference
constexpr int kMlpRowsGrad = 32;  // rows per wavefront, gradient

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ P,
                                                     const float* __restrict__ X, int n, float* out) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int r0 = wave * kMlpRowsFwd;
    if (r0 >= n) return;
    const int r1 = min(n, r0 + kMlpRowsFwd);
    mlp_rows<false>(P, X, nullptr, r0, r1, 0.0f, out, nullptr);
}

__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ P,
                                                  const float* __restrict__ X,
                                                  const float* __restrict__ T, int n, float scale,
                                                  float* part, float* lpart) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int r0 = wave * kMlpRowsGrad;
    if (r0 >= n) return;
    const int r1 = min(n, r0 + kMlpRowsGrad);
    float l = mlp_rows<true>(P, X, T, r0, r1, scale, nullptr, part + (long long)wave * kMlpParams);
    if (__lane_id() == 0) lpart[wave] = l;
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int nw, float* grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float s = 0.0f;
    for (int w = 0; w < nw; ++w) s += part[(long long)w * kMlpParams + i];
    grad[i] = s;
}

// loss = (sum of the per-wave loss sums, in wave order, in FP64) / n
__global__ void k_mlp_loss(const float* __restrict__ lpart, int nw, int n, double* loss) {
    if (threadIdx.x != 0) return;
    double l = 0.0;
    for (int w = 0; w < nw; ++w) l += lpart[w];
    *loss = l / n;
}

// Scratch of vgc_mlp_grad_noalloc: the per-wave gradient partials and loss
// sums of up to kNoAllocRows rows, in static device memory (each GPU holds
// its own copy for the life of the process).  Host code reaches it through
// the address hipGetSymbolAddress returns for the device (scratch_ptr), never
// through the array's name.
constexpr int kNoAllocRows = 32768;
constexpr int kNoAllocWaves = kNoAllocRows / kMlpRowsGrad;
static __device__ __align__(256) float scratch[(size_t)kNoAllocWaves * (kMlpParams + 1)];

__global__ void k_mlp_adam(float* p, const float* __restrict__ g, float* m, float* v, float lr,
                           float b1, float b2, float c1, float c2, float eps) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpParams) adam_elem(i, p, g, m, v, lr, b1, b2, c1, c2, eps);
}

__global__ void k_ot_features(int n, const double* __restrict__ xyz, const double* __restrict__ nu,
                              const float* __restrict__ d2, const double* __restrict__ box,
                              double inv_cell, double h, float* f) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ot_features_row(i, xyz, nu, d2, box, inv_cell, h, f + (long long)i * kMlpIn);
}


// Device-resident Jacobi-PCG: the scalars live in cg[8] = {rz, rr, pq, alpha,
// beta, threshold, done, iterations}; no host round trip per iteration.  The
// reductions follow the kRedBlocks x 256 order above, so the iterates equal
// those of a host-driven loop.  done: 1 converged (or breakdown); every
// kernel is a no-op afterwards, so extra launched iterations change nothing.
__device__ __forceinline__ void block_sum2(double a, double b, double* sh, double* sh2, double& ta,
                                           double& tb) {
    sh[threadIdx.x] = a;
    sh2[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sh[threadIdx.x] += sh[threadIdx.x + off];
            sh2[threadIdx.x] += sh2[threadIdx.x + off];
        }
        __syncthreads();
    }
    ta = sh[0];
    tb = sh2[0];
}

__global__ __launch_bounds__(256) void k_cg_start_partial(int n, const double* __restrict__ b,
                                                          const double* __restrict__ diag,
                                                          double* x, double* r, double* z,
                                                          double* p, double* part) {
    __shared__ double sh[256], sh2[256];
    double s1 = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) {
        const double ri = b[i], zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
        x[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        s1 += ri * zi;
        s2 += ri * ri;
    }
    double t1, t2;
    block_sum2(s1, s2, sh, sh2, t1, t2);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = t1;
        part[kRedBlocks + blockIdx.x] = t2;
    }
}

__global__ __launch_bounds__(256) void k_cg_start_final(const double* __restrict__ part, double* cg,
                                                        double tol2) {
    __shared__ double sh[256], sh2[256];
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) {
        s1 += part[i];
        s2 += part[kRedBlocks + i];
    }
    double t1, t2;
    block_sum2(s1, s2, sh, sh2, t1, t2);
    if (threadIdx.x == 0) {
        cg[0] = t1;
        cg[1] = t2;
        cg[4] = 0.0;
        cg[5] = tol2 * t2;
        cg[6] = t2 > 0.0 ? 0.0 : 1.0;
        cg[7] = 0.0;
        cg[8] = t1;    // r.z of the first direction (iteration parity 0)
        cg[10] = 1.0;  // beta (cg[4]) is current
    }
}

// PCG iteration, four kernels without atomics.  Each reduction writes
// kRedBlocks block partials; the kernel that needs the total has every one
// of its blocks add the partials in the fixed order (thread t takes t,
// t + 256, ..., then the halving tree), so every block holds the same value
// and block 0 alone writes the scalars.  Scalars cg[16]: [0] r.z, [1] r.r,
// [2] p.q, [3] alpha, [4] beta, [5] tol^2 b.b, [6] stop (1 converged,
// 2 breakdown), [7] iterations, [8 + q] r.z of the direction of the
// iterations of parity q (double-buffered: a kernel reads one slot while its
// block 0 writes the other), [10] beta is current (set by cg_start and
// k_cg_close).  Partials: [0, R) p.q, [R (1 + 2q), R (3 + 2q)) r.z and r.r of
// the update of parity q.
__device__ __forceinline__ void red_final2(const double* __restrict__ pa,
                                           const double* __restrict__ pb, double* sh, double* sh2,
                                           double& ta, double& tb) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < kRedBlocks; k += 256) {
        a += pa[k];
        b += pb[k];
    }
    block_sum2(a, b, sh, sh2, ta, tb);
}

// Closes the update of parity par: counts it, stops on r.r <= tol^2 b.b,
// else beta = r.z / previous r.z.  Block 0 of the caller writes the
// scalars; the next direction uses parity slot next.
__device__ __forceinline__ bool cg_close_update(double* cg, const double* __restrict__ part,
                                                int par, int next, double* sh, double* sh2,
                                                double& beta) {
    double rz, rr;
    red_final2(part + kRedBlocks * (1 + 2 * par), part + kRedBlocks * (2 + 2 * par), sh, sh2, rz,
               rr);
    const bool stop = rr <= cg[5];
    beta = stop ? 0.0 : rz / cg[8 + par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cg[1] = rr;
        cg[7] += 1.0;
        cg[4] = beta;
        if (stop) {
            cg[6] = 1.0;
        } else {
            cg[0] = rz;
            cg[8 + next] = rz;
        }
    }
    return stop;
}

// First: complete the previous iteration (unless beta is current), then the
// direction p = z + beta p.
__global__ __launch_bounds__(256) void k_cg_dir(int n, const double* __restrict__ z,
                                                double* __restrict__ p, double* cg,
                                                const double* __restrict__ part, int par) {
    __shared__ double sh[256], sh2[256];
    if (cg[6] != 0.0) return;  // uniform over the grid
    double beta = cg[4];
    if (cg[10] == 0.0 && cg_close_update(cg, part, par ^ 1, par, sh, sh2, beta)) return;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256)
        p[i] = z[i] + beta * p[i];
}

// Second: q = A p, one 16-lane group per row; a grid of at most
// kSpmvBlocks resident work-groups strides over the rows (a work-group per
// 16 rows would spend its short life being dispatched).
constexpr int kSpmvBlocks = 2048;
__global__ __launch_bounds__(256) void k_cg_spmv(int n, const long long* __restrict__ ptr,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ diag,
                                                 const double* __restrict__ p, double* q,
                                                 double* cg) {
    if (cg[6] != 0.0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) cg[10] = 0.0;  // read by the next k_cg_dir only
    // two row groups per step, their loads issued together (the first
    // kSpmvLanes entries of a row as straight-line code); each row's sum is
    // formed in the order of ot_spmv_lane
    constexpr int U = 2;  // measured: 66.6 us against 73.6 us at U = 4 (1M rows, 15.3M entries, 2048 blocks)
    const long long stride = (long long)gridDim.x * blockDim.x, N = (long long)n * kSpmvLanes;
    for (long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; t0 < N; t0 += U * stride) {
        double s[U];
        int row[U];
        bool ok[U];
        long long k[U], e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long t = t0 + u * stride;
            ok[u] = t < N;
            row[u] = ok[u] ? int(t / kSpmvLanes) : 0;
            k[u] = ptr[row[u]] + int(t % kSpmvLanes);
            e[u] = ok[u] ? ptr[row[u] + 1] : 0;
        }
        double v[U], x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = k[u] < e[u];
            v[u] = in ? val[k[u]] : 0.0;
            x[u] = in ? p[col[k[u]]] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s[u] = 0.0;
            if (k[u] < e[u]) s[u] += v[u] * x[u];
            for (long long kk = k[u] + kSpmvLanes; kk < e[u]; kk += kSpmvLanes) s[u] += val[kk] * p[col[kk]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int m = kSpmvLanes / 2; m >= 1; m >>= 1) s[u] += __shfl_xor(s[u], m);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u] && (t0 % kSpmvLanes) == 0) q[row[u]] = s[u] + diag[row[u]] * p[row[u]];
    }
}

// Third: the block partials of p.q.
__global__ __launch_bounds__(256) void k_cg_dot(int n, const double* __restrict__ p,
                                                const double* __restrict__ q,
                                                const double* __restrict__ cg, double* part) {
    __shared__ double sh[256];
    if (cg[6] != 0.0) return;
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) s += p[i] * q[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Fourth: alpha = r.z / p.q (or the breakdown stop), x += alpha p,
// r -= alpha q, z = r / diag, and the block partials of r.z and r.r.
__global__ __launch_bounds__(256) void k_cg_update(int n, const double* __restrict__ p,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ diag,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z, double* cg,
                                                   double* part, int par) {
    __shared__ double sh[256], sh2[256];
    if (cg[6] != 0.0) return;
    double a = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += 256) a += part[b];
    const double pq = block_sum(a, sh);
    const bool breakdown = !(pq > 0.0);
    const double alpha = breakdown ? 0.0 : cg[8 + par] / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cg[2] = pq;
        cg[3] = alpha;
        if (breakdown) cg[6] = 2.0;  // stop, the iteration does not count
    }
    if (breakdown) return;
    double s1 = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
        z[i] = zi;
        s1 += ri * zi;
        s2 += ri * ri;
    }
    double t1, t2;
    block_sum2(s1, s2, sh, sh2, t1, t2);
    if (threadIdx.x == 0) {
        part[kRedBlocks * (1 + 2 * par) + blockIdx.x] = t1;
        part[kRedBlocks * (2 + 2 * par) + blockIdx.x] = t2;
    }
}

// After the last iteration of a captured chunk (parity par): its update
// closed by one block, so the host reads final scalars between chunks.
__global__ __launch_bounds__(256) void k_cg_close(double* cg, const double* __restrict__ part,
                                                  int par) {
    __shared__ double sh[256], sh2[256];
    if (cg[6] != 0.0 || cg[10] != 0.0) return;
    double beta;
    cg_close_update(cg, part, par, 0, sh, sh2, beta);
    if (threadIdx.x == 0) cg[10] = 1.0;
}

// Morton order of a Newton system: inv[orig[m]] = m; row lengths, rows and
// vectors gathered / scattered through orig.
__global__ void k_inv_perm(int n, const int* __restrict__ orig, int* inv) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) inv[orig[m]] = m;
}
__global__ void k_perm_row_len(int n, const int* __restrict__ orig,
                               const long long* __restrict__ ptr, int* len) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) len[m] = int(ptr[orig[m] + 1] - ptr[orig[m]]);
}
__global__ void k_perm_rows(int n, const int* __restrict__ orig, const int* __restrict__ inv,
                            const long long* __restrict__ ptr, const int* __restrict__ col,
                            const double* __restrict__ val, const long long* __restrict__ ptr_m,
                            int* col_m, double* val_m) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = int(t / kSpmvLanes), l = int(t % kSpmvLanes);
    if (m >= n) return;
    const int i = orig[m];
    const long long b = ptr[i], e = ptr[i + 1], o = ptr_m[m] - b;
    for (long long k = b + l; k < e; k += kSpmvLanes) {
        col_m[o + k] = inv[col[k]];
        val_m[o + k] = val[k];
    }
}
// the Morton position (= caller position range) of row r of a subset of
// instances, through orig: out[r] = orig[off[k] + r - soff[k]], k = inst[r]
__global__ void k_range_map(int n, const int* __restrict__ inst, const long long* __restrict__ soff,
                            const long long* __restrict__ off, const int* __restrict__ orig,
                            int* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = orig[off[inst[r]] + (r - soff[inst[r]])];
}
__global__ void k_gather_f64(int n, const int* __restrict__ orig, const double* __restrict__ src,
                             double* dst) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) dst[m] = src[orig[m]];
}
__global__ void k_scatter_f64(int n, const int* __restrict__ orig, const double* __restrict__ src,
                              double* dst) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) dst[orig[m]] = src[m];
}


// ------------------------------------------------------ batched OT kernels
__device__ __forceinline__ void seg_range(const SegTable& T, int blk, int& k, long long& beg,
                                          long long& end, long long& stride) {
    k = T.blk_inst[blk];
    beg = T.off[k] + (long long)T.blk_idx[blk] * 256 + threadIdx.x;
    end = T.off[k + 1];
    stride = (long long)T.cnt[k] * 256;
}

// per instance: [0] max |m-nu|/nu, [1] sum (m-nu)^2, [2] min m; one block each
__global__ __launch_bounds__(256) void k_seg_mass(SegTable T, const double* __restrict__ m,
                                                  const double* __restrict__ nu, double* out) {
    __shared__ double sh[256], sh2[256];
    const int k = blockIdx.x;
    double mx = 0.0, s2 = 0.0, mn = INFINITY;
    for (long long i = T.off[k] + threadIdx.x; i < T.off[k + 1]; i += 256) {
        const double d = m[i] - nu[i];
        mx = fmax(mx, fabs(d) / nu[i]);
        s2 += d * d;
        mn = fmin(mn, m[i]);
    }
    double t2, dummy;
    block_sum2(s2, 0.0, sh, sh2, t2, dummy);
    __syncthreads();
    sh[threadIdx.x] = mx;
    sh2[threadIdx.x] = mn;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
            sh2[threadIdx.x] = fmin(sh2[threadIdx.x], sh2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * k + 0] = sh[0];
        out[3 * k + 1] = t2;
        out[3 * k + 2] = sh2[0];
    }
}

// b = nu - m for active instances, 0 for the others
__global__ void k_seg_rhs(int n, const int* __restrict__ inst_of, const int* __restrict__ active,
                          const double* __restrict__ nu, const double* __restrict__ m, double* b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = active[inst_of[i]] ? nu[i] - m[i] : 0.0;
}

// out = w + t[instance] d
__global__ void k_seg_axpy(int n, const int* __restrict__ inst_of, const double* __restrict__ w,
                           const double* __restrict__ t, const double* __restrict__ d,
                           double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = w[i] + t[inst_of[i]] * d[i];
}

// dst = src on the instances with sel != 0
__global__ void k_seg_copy(int n, const int* __restrict__ inst_of, const int* __restrict__ sel,
                           const double* __restrict__ src, double* dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && sel[inst_of[i]]) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_cgb_start_partial(SegTable T, const double* __restrict__ b,
                                                           const double* __restrict__ diag,
                                                           double* x, double* r, double* z,
                                                           double* p, double* part) {
    __shared__ double sh[256], sh2[256];
    int k;
    long long beg, end, stride;
    seg_range(T, blockIdx.x, k, beg, end, stride);
    double s1 = 0.0, s2 = 0.0;
    for (long long i = beg; i < end; i += stride) {
        const double ri = b[i], zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
        x[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        s1 += ri * zi;
        s2 += ri * ri;
    }
    double t1, t2;
    block_sum2(s1, s2, sh, sh2, t1, t2);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = t1;
        part[T.nblk + blockIdx.x] = t2;
    }
}

__device__ __forceinline__ void seg_final2(const SegTable& T, int k, const double* part, double* sh,
                                           double* sh2, double& t1, double& t2) {
    const int c = T.cnt[k], f = T.first[k];
    double a = (int)threadIdx.x < c ? part[f + threadIdx.x] : 0.0;
    double b = (int)threadIdx.x < c ? part[T.nblk + f + threadIdx.x] : 0.0;
    block_sum2(a, b, sh, sh2, t1, t2);
}

__global__ __launch_bounds__(256) void k_cgb_start_final(SegTable T, const double* __restrict__ part,
                                                         double* cg,
                                                         const double* __restrict__ tol2) {
    __shared__ double sh[256], sh2[256];
    const int k = blockIdx.x;
    double t1, t2;
    seg_final2(T, k, part, sh, sh2, t1, t2);
    if (threadIdx.x == 0) {
        double* c = cg + kCgStride * k;
        c[0] = t1;
        c[1] = t2;
        c[4] = 0.0;
        c[5] = tol2[k] * t2;
        c[6] = t2 > 0.0 ? 0.0 : 1.0;
        c[7] = 0.0;
        c[8] = t1;    // r.z of the first direction (iteration parity 0)
        c[10] = 1.0;  // beta (c[4]) is current
    }
}

// Batched PCG iteration: the kernels of the single-instance iteration
// (k_cg_dir ... k_cg_update) per instance over the block table: every block
// of instance k adds the instance's cnt[k] block partials in the fixed order
// (thread t takes partial t, then the halving tree), so the instance's
// blocks agree, and its block 0 (blk_idx 0) writes its scalars
// (kCgStride per instance).  Partials: [0, nblk) p.q, [nblk (1 + 2q),
// nblk (3 + 2q)) r.z and r.r of the update of parity q.
__device__ __forceinline__ bool cgb_close_update(const SegTable& T, int k, double* c,
                                                 const double* __restrict__ part, int par, int next,
                                                 bool writer, double* sh, double* sh2,
                                                 double& beta) {
    const int cnt = T.cnt[k], f = T.first[k];
    const double* pa = part + T.nblk * (1 + 2 * par);
    const double a = (int)threadIdx.x < cnt ? pa[f + threadIdx.x] : 0.0;
    const double b = (int)threadIdx.x < cnt ? pa[T.nblk + f + threadIdx.x] : 0.0;
    double rz, rr;
    block_sum2(a, b, sh, sh2, rz, rr);
    const bool stop = rr <= c[5];
    beta = stop ? 0.0 : rz / c[8 + par];
    if (writer && threadIdx.x == 0) {
        c[1] = rr;
        c[7] += 1.0;
        c[4] = beta;
        if (stop) {
            c[6] = 1.0;
        } else {
            c[0] = rz;
            c[8 + next] = rz;
        }
    }
    return stop;
}

__global__ __launch_bounds__(256) void k_cgb_dir(SegTable T, const double* __restrict__ z,
                                                 double* __restrict__ p, double* cg,
                                                 const double* __restrict__ part, int par) {
    __shared__ double sh[256], sh2[256];
    int k;
    long long beg, end, stride;
    seg_range(T, blockIdx.x, k, beg, end, stride);
    double* c = cg + kCgStride * k;
    if (c[6] != 0.0) return;  // uniform over the block and the instance
    double beta = c[4];
    if (c[10] == 0.0 &&
        cgb_close_update(T, k, c, part, par ^ 1, par, T.blk_idx[blockIdx.x] == 0, sh, sh2, beta))
        return;
    for (long long i = beg; i < end; i += stride) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(256) void k_cgb_spmv(int n, const int* __restrict__ inst_of,
                                                  const long long* __restrict__ ptr,
                                                  const int* __restrict__ col,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ diag,
                                                  const double* __restrict__ p, double* q,
                                                  const double* __restrict__ cg) {
    // as k_cg_spmv (grid-stride), rows of stopped instances skipped; four
    // row groups per step here (two: 5.92 s of CG in the 256-instance
    // transport against 5.72 s; the stop-flag read lengthens each row's
    // chain of dependent loads)
    constexpr int U = 4;
    const long long stride = (long long)gridDim.x * blockDim.x, N = (long long)n * kSpmvLanes;
    for (long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; t0 < N; t0 += U * stride) {
        double s[U];
        int row[U];
        bool ok[U];
        long long k[U], e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long t = t0 + u * stride;
            row[u] = t < N ? int(t / kSpmvLanes) : 0;
            ok[u] = t < N && cg[kCgStride * inst_of[row[u]] + 6] == 0.0;
            k[u] = ptr[row[u]] + int(t % kSpmvLanes);
            e[u] = ok[u] ? ptr[row[u] + 1] : 0;
        }
        double v[U], x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = k[u] < e[u];
            v[u] = in ? val[k[u]] : 0.0;
            x[u] = in ? p[col[k[u]]] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s[u] = 0.0;
            if (k[u] < e[u]) s[u] += v[u] * x[u];
            for (long long kk = k[u] + kSpmvLanes; kk < e[u]; kk += kSpmvLanes) s[u] += val[kk] * p[col[kk]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int m = kSpmvLanes / 2; m >= 1; m >>= 1) s[u] += __shfl_xor(s[u], m);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u] && (t0 % kSpmvLanes) == 0) q[row[u]] = s[u] + diag[row[u]] * p[row[u]];
    }
}

__global__ __launch_bounds__(256) void k_cgb_dot(SegTable T, const double* __restrict__ p,
                                                 const double* __restrict__ q, double* cg,
                                                 double* part) {
    __shared__ double sh[256], sh2[256];
    int k;
    long long beg, end, stride;
    seg_range(T, blockIdx.x, k, beg, end, stride);
    double* c = cg + kCgStride * k;
    if (c[6] != 0.0) return;
    if (T.blk_idx[blockIdx.x] == 0 && threadIdx.x == 0) c[10] = 0.0;  // read by the next dir only
    double s = 0.0;
    for (long long i = beg; i < end; i += stride) s += p[i] * q[i];
    double t1, t2;
    block_sum2(s, 0.0, sh, sh2, t1, t2);
    if (threadIdx.x == 0) part[blockIdx.x] = t1;
}

__global__ __launch_bounds__(256) void k_cgb_update(SegTable T, const double* __restrict__ p,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ diag,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ z, double* cg,
                                                    double* part, int par) {
    __shared__ double sh[256], sh2[256];
    int k;
    long long beg, end, stride;
    seg_range(T, blockIdx.x, k, beg, end, stride);
    double* c = cg + kCgStride * k;
    if (c[6] != 0.0) return;
    const int cnt = T.cnt[k], f = T.first[k];
    const double a = (int)threadIdx.x < cnt ? part[f + threadIdx.x] : 0.0;
    double pq, dummy;
    block_sum2(a, 0.0, sh, sh2, pq, dummy);
    const bool breakdown = !(pq > 0.0);
    const double alpha = breakdown ? 0.0 : c[8 + par] / pq;
    if (T.blk_idx[blockIdx.x] == 0 && threadIdx.x == 0) {
        c[2] = pq;
        c[3] = alpha;
        if (breakdown) c[6] = 2.0;
    }
    if (breakdown) return;
    double s1 = 0.0, s2 = 0.0;
    for (long long i = beg; i < end; i += stride) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
        z[i] = zi;
        s1 += ri * zi;
        s2 += ri * ri;
    }
    double t1, t2;
    block_sum2(s1, s2, sh, sh2, t1, t2);
    if (threadIdx.x == 0) {
        part[T.nblk * (1 + 2 * par) + blockIdx.x] = t1;
        part[T.nblk * (2 + 2 * par) + blockIdx.x] = t2;
    }
}

// one block per instance after a chunk's last iteration (parity par)
__global__ __launch_bounds__(256) void k_cgb_close(SegTable T, double* cg,
                                                   const double* __restrict__ part, int par) {
    __shared__ double sh[256], sh2[256];
    const int k = blockIdx.x;
    double* c = cg + kCgStride * k;
    if (c[6] != 0.0 || c[10] != 0.0) return;
    double beta;
    cgb_close_update(T, k, c, part, par, 0, true, sh, sh2, beta);
    if (threadIdx.x == 0) c[10] = 1.0;
}

// per instance (one block each): b -= mean of b over the instance's rows —
// the Newton right-hand side projected on the range of its Laplacian
__global__ __launch_bounds__(256) void k_seg_center(SegTable T, double* b) {
    __shared__ double sh[256], sh2[256];
    const int k = blockIdx.x;
    const long long beg = T.off[k], end = T.off[k + 1];
    double s = 0.0;
    for (long long i = beg + threadIdx.x; i < end; i += 256) s += b[i];
    double t, dummy;
    block_sum2(s, 0.0, sh, sh2, t, dummy);
    const double mean = t / double(end - beg);
    for (long long i = beg + threadIdx.x; i < end; i += 256) b[i] -= mean;
}

__global__ void k_set0_i64(long long* p) { *p = 0; }

__global__ void k_piece_setup(int nq, const int* __restrict__ cells, const pt4* __restrict__ pts,
                              const int* __restrict__ inst_of, const double* __restrict__ box,
                              double* p_box) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) piece_setup_row(q, cells, pts, inst_of, box, p_box);
}
__global__ void k_piece_merge(int m, const int* __restrict__ cells, const int* __restrict__ orig,
                              const long long* __restrict__ pptr, const int* __restrict__ pidx,
                              const double* __restrict__ p_vol, const double* __restrict__ p_mg,
                              const double* __restrict__ p_e2, const int* __restrict__ p_cnt,
                              const int* __restrict__ p_nbr, const double* __restrict__ p_area,
                              const uint32_t* __restrict__ p_flags, int stride,
                              const long long* __restrict__ x_off, double* vol, double* mg,
                              double* e2, int* nbr_cnt, uint32_t* flags, int* x_nbr,
                              double* x_area) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m)
        piece_merge_row(t, cells, orig, pptr, pidx, p_vol, p_mg, p_e2, p_cnt, p_nbr, p_area,
                        p_flags, stride, x_off, vol, mg, e2, nbr_cnt, flags, x_nbr, x_area);
}
__global__ void k_piece_copy(int m, const int* __restrict__ cells, const int* __restrict__ orig,
                             const long long* __restrict__ x_off, const int* __restrict__ x_nbr,
                             const double* __restrict__ x_area, const int* __restrict__ nbr_cnt,
                             const long long* __restrict__ ptr, int* nbr, double* area) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) piece_copy_row(t, cells, orig, x_off, x_nbr, x_area, nbr_cnt, ptr, nbr, area);
}

__global__ void k_ot_repair(int n, const double* __restrict__ xyz, const double* __restrict__ w,
                            const double* __restrict__ vol, const pt4* __restrict__ pts,
                            const bnode* __restrict__ nodes, InstLayout IL, float cs2, double delta,
                            double* wout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wout[i] = ot_repair_row(i, xyz, w, vol, pts, nodes, IL, cs2, delta);
}

// weight of the power-nearest seed (instance 0 of the tree) of each query
__global__ void k_power_nearest(const pt4* __restrict__ pts, const bnode* __restrict__ nodes,
                                InstLayout IL, const double* __restrict__ q, int nq, float cs2,
                                double* wout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const int j = power_nearest_query(pts, nodes, IL, q[3 * (size_t)t + 0], q[3 * (size_t)t + 1],
                                      q[3 * (size_t)t + 2], cs2);
    wout[t] = pts[j].w;
}

// Cells of `items` (all n_items cells if null) that are not plainly done —
// not CELL_DONE, or carrying a flag of `mask` — appended with their flags
// (unordered; the host sorts them).  Cells of the instances marked in
// `skip` (by inst_of; instance 0 if null) are left out.
__global__ void k_census(const int* __restrict__ items, int n_items,
                         const uint32_t* __restrict__ flags, uint32_t mask,
                         const unsigned* __restrict__ skip, const int* __restrict__ inst_of,
                         int* out_cell, uint32_t* out_flag, unsigned* count) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_items) return;
    const int c = items ? items[t] : t;
    if (skip && skip[inst_of ? inst_of[c] : 0]) return;
    const uint32_t f = flags[c];
    if ((f & mask) || !(f & CELL_DONE)) {
        const unsigned k = atomicAdd(count, 1u);
        out_cell[k] = c;
        out_flag[k] = f;
    }
}

__global__ void k_gather_u32(const int* __restrict__ items, int n, const uint32_t* __restrict__ src,
                             uint32_t* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = src[items[t]];
}

// ----------------------------------------------------------------- backend

static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct HipBackend {
    explicit HipBackend(int device) : dev_(device) {
        VGC_HIP_CHECK(hipSetDevice(dev_));
        VGC_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        VGC_HIP_CHECK(hipEventCreateWithFlags(&enter_ev_, hipEventDisableTiming));
    }
    ~HipBackend() {
        for (auto& e : ev_)
            for (auto& x : e)
                if (x) (void)hipEventDestroy(x);
        if (temp_) (void)hipFree(temp_);
        if (red_) (void)hipFree(red_);
        if (vdot_) (void)hipFree(vdot_);
        if (cg_part_) (void)hipFree(cg_part_);
        if (cg_exec_) (void)hipGraphExecDestroy(cg_exec_);
        if (cgb_exec_) (void)hipGraphExecDestroy(cgb_exec_);
        if (seg_part_) (void)hipFree(seg_part_);
        if (wide_) (void)hipFree(wide_);
        if (wide_front_) (void)hipFree(wide_front_);
        if (enter_ev_) (void)hipEventDestroy(enter_ev_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    hipStream_t stream() const { return stream_; }

    // The caller's stream (the null stream unless set): each call first
    // waits, on the context's own stream, for the work queued there before
    // it — device inputs written by the caller's asynchronous kernels.
    void set_caller_stream(hipStream_t s) { caller_ = s; }
    void enter() {
        VGC_HIP_CHECK(hipEventRecord(enter_ev_, caller_));
        VGC_HIP_CHECK(hipStreamWaitEvent(stream_, enter_ev_, 0));
    }

    void* dalloc(size_t b) {
        void* p = nullptr;
        VGC_HIP_CHECK(hipMalloc(&p, b ? b : 1));
        return p;
    }
    void dfree(void* p) { (void)hipFree(p); }
    void h2d(void* d, const void* s, size_t b) {
        VGC_HIP_CHECK(hipMemcpyAsync(d, s, b, hipMemcpyHostToDevice, stream_));
        VGC_HIP_CHECK(hipStreamSynchronize(stream_));
    }
    void d2h(void* d, const void* s, size_t b) {
        VGC_HIP_CHECK(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToHost, stream_));
        VGC_HIP_CHECK(hipStreamSynchronize(stream_));
    }
    void d2d(void* d, const void* s, size_t b) {
        VGC_HIP_CHECK(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, stream_));
    }
    void dzero(void* d, size_t b) { VGC_HIP_CHECK(hipMemsetAsync(d, 0, b, stream_)); }
    void sync() { VGC_HIP_CHECK(hipStreamSynchronize(stream_)); }
    // HIP events on the launch stream around a phase (slot 0: kNN, 1: pass A)
    void event_begin(int k) {
        if (!ev_[k][0]) {
            VGC_HIP_CHECK(hipEventCreate(&ev_[k][0]));
            VGC_HIP_CHECK(hipEventCreate(&ev_[k][1]));
        }
        VGC_HIP_CHECK(hipEventRecord(ev_[k][0], stream_));
    }
    void event_end(int k) { VGC_HIP_CHECK(hipEventRecord(ev_[k][1], stream_)); }
    double event_ms(int k) {
        float ms = 0.f;
        VGC_HIP_CHECK(hipEventSynchronize(ev_[k][1]));
        VGC_HIP_CHECK(hipEventElapsedTime(&ms, ev_[k][0], ev_[k][1]));
        return ms;
    }
    void check() { VGC_HIP_CHECK(hipGetLastError()); }

    void* temp(size_t b) {
        if (b > temp_bytes_) {
            if (temp_) VGC_HIP_CHECK(hipFree(temp_));
            temp_bytes_ = b + b / 4;
            VGC_HIP_CHECK(hipMalloc(&temp_, temp_bytes_));
        }
        return temp_;
    }

    void fill_inst(int* inst_orig, int n, const int* p0, int ni) {
        k_fill_inst<<<blocks_for(n, 256), 256, 0, stream_>>>(inst_orig, n, p0, ni);
        check();
    }
    void morton(const double* xyz, const int* inst_orig, const double* box, int n, uint64_t* keys,
                int* vals) {
        k_morton<<<blocks_for(n, 256), 256, 0, stream_>>>(xyz, inst_orig, box, n, keys, vals);
        check();
    }
    void sort_pairs(uint64_t* keys, int* vals, uint64_t* keys2, int* vals2, int n, int end_bit) {
        size_t tb = 0;
        VGC_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, keys, keys2, vals, vals2, (unsigned)n,
                                                0, end_bit, stream_));
        void* t = temp(tb);
        VGC_HIP_CHECK(rocprim::radix_sort_pairs(t, tb, keys, keys2, vals, vals2, (unsigned)n, 0,
                                                end_bit, stream_));
    }
    void gather(const double* xyz, const double* w, const int* inst_orig, const int* perm, int n,
                pt4* pts, pf4* pf, int* orig, int* inst_of) {
        k_gather<<<blocks_for(n, 256), 256, 0, stream_>>>(xyz, w, inst_orig, perm, n, pts, pf, orig,
                                                          inst_of);
        check();
    }
    void bvh_leaves(const pt4* pts, const InstLayout& IL, bnode* nodes, const long long* leaf_off,
                    int ni, long long total) {
        k_bvh_leaves<<<blocks_for(total, 256), 256, 0, stream_>>>(pts, IL, nodes, leaf_off, ni, total);
        check();
    }
    void bvh_level(const InstLayout& IL, bnode* nodes, const long long* lvl_off, int ni,
                   long long total, int h) {
        k_bvh_level<<<blocks_for(total, 256), 256, 0, stream_>>>(IL, nodes, lvl_off, ni, total, h);
        check();
    }
    void wmax(const pt4* pts, const InstLayout& IL, int ni, double* out) {
        k_wmax<<<ni, 256, 0, stream_>>>(pts, IL, out);
        check();
    }
    void knn(int K, const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* queries,
             int nq, int* out, float* out_d2) {
        unsigned g = blocks_for(nq, 256);
        switch (K) {
            case 16: k_knn<16><<<g, 256, 0, stream_>>>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 24: k_knn<24><<<g, 256, 0, stream_>>>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 32: k_knn<32><<<g, 256, 0, stream_>>>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 48: k_knn<48><<<g, 256, 0, stream_>>>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 64: k_knn<64><<<g, 256, 0, stream_>>>(pf, nodes, IL, queries, nq, out, out_d2); break;
            default: throw std::runtime_error("unsupported K");
        }
        check();
    }
    template <bool E, bool G, bool W>
    void clip_egw(int S, const ClipParams& P) {
        const unsigned g4 = blocks_for(P.n_items, 4), g1 = blocks_for(P.n_items, 1);
        if (S == 1) {
            if constexpr (G && !W) {
                if (P.corner) {
                    // second gather of Voronoi cells, with the corner test
                    k_clip<1, 1, E, G, W, true><<<g1, 64, 0, stream_>>>(P);
                    return;
                }
            }
            if constexpr (G && !E) {
                if (P.fp32) {
                    if (waves_ == 1)
                        k_clip<1, 1, false, true, W, W, true><<<g1, 64, 0, stream_>>>(P);
                    else
                        k_clip<1, 4, false, true, W, W, true><<<g4, 256, 0, stream_>>>(P);
                    return;
                }
            }
            if (waves_ == 1)
                k_clip<1, 1, E, G, W><<<g1, 64, 0, stream_>>>(P);
            else
                k_clip<1, 4, E, G, W><<<g4, 256, 0, stream_>>>(P);
        } else if (S == 2) {
            k_clip<2, 4, E, G, W><<<g4, 256, 0, stream_>>>(P);
        } else if (S == 4) {
            k_clip<4, 1, E, G, W><<<g1, 64, 0, stream_>>>(P);
        } else {
            // 512 corners: list mode only (the last tier)
            if constexpr (!G)
                k_clip<8, 1, E, false, W><<<g1, 64, 0, stream_>>>(P);
            else
                throw std::runtime_error("clip: 512-corner cells run from lists only");
        }
    }
    template <bool E, bool G>
    void clip_eg(int S, const ClipParams& P) {
        if (P.weighted)
            clip_egw<E, G, true>(S, P);
        else
            clip_egw<E, G, false>(S, P);
    }
    // wavefronts per work-group of the S = 1 kernels (1 or 4); the cells'
    // results do not depend on it
    void set_clip_waves(int w) { waves_ = w == 1 ? 1 : 4; }
    int clip_waves() const { return waves_; }
    void clip(int S, bool exact, bool gather, const ClipParams& P) {
        if (P.n_items <= 0) return;
        if (exact) {
            if (gather)
                clip_eg<true, true>(S, P);
            else
                clip_eg<true, false>(S, P);
        } else {
            if (gather)
                clip_eg<false, true>(S, P);
            else
                clip_eg<false, false>(S, P);
        }
        check();
    }
    void radius_r2(const int* items, int ni, const pt4* pts, const double* r2, const int* inst_of,
                   const double* wmaxv, int weighted, const fball* ball, const double* box,
                   float* rho2, PowerFilter* filt) {
        k_radius_r2<<<blocks_for(ni, 256), 256, 0, stream_>>>(items, ni, pts, r2, inst_of, wmaxv,
                                                              weighted, ball, box, rho2, filt);
        check();
    }
    // largest leaf count of the instances' trees (scratch of k_radius_wide)
    void set_tree_leaves(int lmax) { tree_leaves_ = lmax; }
    void radius_count(const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* items,
                      int ni, const float* rho2, const PowerFilter* filt, int* cnt) {
        if (ni <= 0) return;
        unsigned char* wide = wide_flags(ni);
        k_radius_count<<<blocks_for(ni, 64), 64, 0, stream_>>>(pf, nodes, IL, items, ni, rho2, filt,
                                                               cnt, wide);
        radius_wide(pf, nodes, IL, items, ni, rho2, filt, wide, cnt, nullptr, 0, nullptr, nullptr);
        check();
    }
    void radius_fill(const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* items,
                     int ni, const float* rho2, const PowerFilter* filt, const long long* off,
                     long long base, int* list, float* d2) {
        if (ni <= 0) return;
        unsigned char* wide = wide_flags(ni);
        k_radius_fill<<<blocks_for(ni, 64), 64, 0, stream_>>>(pf, nodes, IL, items, ni, rho2, filt,
                                                              off, base, list, d2, wide);
        radius_wide(pf, nodes, IL, items, ni, rho2, filt, wide, nullptr, off, base, list, d2);
        check();
    }
    // the flagged queries by k_radius_wide: up to 64 waves, each with two
    // frontier buffers of tree_leaves_ nodes (at most 256 MB in all)
    void radius_wide(const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* items,
                     int ni, const float* rho2, const PowerFilter* filt, const unsigned char* wide,
                     int* cnt, const long long* off, long long base, int* list, float* d2) {
        const long long fstride = std::max(1, tree_leaves_);
        long long g = (256ll << 20) / (2 * sizeof(int) * fstride);
        g = std::max(1ll, std::min(g, std::min(64ll, (long long)ni)));
        const long long need = g * 2 * fstride;
        if (wide_front_cap_ < need) {
            if (wide_front_) VGC_HIP_CHECK(hipFree(wide_front_));
            VGC_HIP_CHECK(hipMalloc(&wide_front_, sizeof(int) * (size_t)need));
            wide_front_cap_ = need;
        }
        k_radius_wide<<<(unsigned)g, 64, 0, stream_>>>(pf, nodes, IL, items, ni, rho2, filt, wide,
                                                       wide_front_, fstride, cnt, off, base, list,
                                                       d2);
    }
    unsigned char* wide_flags(int ni) {
        if (wide_cap_ < ni) {
            if (wide_) VGC_HIP_CHECK(hipFree(wide_));
            VGC_HIP_CHECK(hipMalloc(&wide_, (size_t)ni));
            wide_cap_ = ni;
        }
        return wide_;
    }
    void rebase(const long long* off, int n, long long base, long long* out) {
        k_rebase<<<blocks_for(n, 256), 256, 0, stream_>>>(off, n, base, out);
        check();
    }
    // k_census into out_cell / out_flag (n_items capacity); returns the count
    unsigned census(const int* items, int n_items, const uint32_t* flags, uint32_t mask,
                    const unsigned* skip, const int* inst_of, int* out_cell, uint32_t* out_flag,
                    unsigned* d_count) {
        VGC_HIP_CHECK(hipMemsetAsync(d_count, 0, sizeof(unsigned), stream_));
        if (n_items > 0)
            k_census<<<blocks_for(n_items, 256), 256, 0, stream_>>>(
                items, n_items, flags, mask, skip, inst_of, out_cell, out_flag, d_count);
        check();
        unsigned cnt = 0;
        d2h(&cnt, d_count, sizeof(unsigned));
        return cnt;
    }
    void power_nearest(const pt4* pts, const bnode* nodes, const InstLayout& IL, const double* q,
                       int nq, float cs2, double* wout) {
        if (nq > 0)
            k_power_nearest<<<blocks_for(nq, 256), 256, 0, stream_>>>(pts, nodes, IL, q, nq, cs2,
                                                                      wout);
        check();
    }
    void gather_u32(const int* items, int n, const uint32_t* src, uint32_t* out) {
        if (n > 0) k_gather_u32<<<blocks_for(n, 256), 256, 0, stream_>>>(items, n, src, out);
        check();
    }
    void scan_i32(const int* in, long long* out, int n) {
        size_t tb = 0;
        VGC_HIP_CHECK(rocprim::inclusive_scan(nullptr, tb, in, out + 1, (size_t)n,
                                              rocprim::plus<long long>(), stream_));
        void* t = temp(tb);
        VGC_HIP_CHECK(rocprim::inclusive_scan(t, tb, in, out + 1, (size_t)n,
                                              rocprim::plus<long long>(), stream_));
        k_set0_i64<<<1, 1, 0, stream_>>>(out);
        check();
    }
    void seg_sort(const float* kin, const int* vin, float* kout, int* vout, long long total,
                  int nseg, const long long* off) {
        if (total <= 0) return;
        // rocPRIM takes 32-bit sizes and sizes its scratch by them
        if (total > 0x7FFFFFFFll || nseg < 0)
            throw std::runtime_error("seg_sort: list chunk exceeds 2^31 entries");
        size_t tb = 0;
        VGC_HIP_CHECK(rocprim::segmented_radix_sort_pairs(nullptr, tb, kin, kout, vin, vout,
                                                          (unsigned)total, (unsigned)nseg, off,
                                                          off + 1, 0, 32, stream_));
        void* t = temp(tb);
        VGC_HIP_CHECK(rocprim::segmented_radix_sort_pairs(t, tb, kin, kout, vin, vout,
                                                          (unsigned)total, (unsigned)nseg, off,
                                                          off + 1, 0, 32, stream_));
    }
    void compact(int n, const int* cnt, const long long* ptr, const int* nbr_tmp,
                 const double* area_tmp, int stride, int* nbr, double* area) {
        k_compact<<<blocks_for((long long)n * kCompactLanes, 256), 256, 0, stream_>>>(
            n, cnt, ptr, nbr_tmp, area_tmp, stride, nbr, area);
        check();
    }
    void lloyd_update(double* xyz, const double* vol, const double* mg, int n,
                      const double* wrap = nullptr) {
        k_lloyd_update<<<blocks_for(n, 256), 256, 0, stream_>>>(xyz, vol, mg, n, wrap);
        check();
    }
    void ghost_count(const double* xyz, int n, const double* box, double delta, int* cnt) {
        k_ghost_count<<<blocks_for(n, 256), 256, 0, stream_>>>(xyz, n, box, delta, cnt);
        check();
    }
    void ghost_fill(const double* xyz, const double* w, int n, const double* box, double delta,
                    const long long* off, double* xe, double* we, int* src) {
        k_ghost_fill<<<blocks_for(n, 256), 256, 0, stream_>>>(xyz, w, n, box, delta, off, xe, we,
                                                              src);
        check();
    }
    void periodic_reach(const pt4* pts, const int* orig, int ne, int n_real, const double* r2,
                        const double* wmax, float* out) {
        VGC_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(float), stream_));
        k_periodic_reach<<<blocks_for(ne, 256), 256, 0, stream_>>>(pts, orig, ne, n_real, r2, wmax,
                                                                   out);
        check();
    }
    // ----------------------------------------------------------------- OT
    double* red_part() {
        if (!red_) VGC_HIP_CHECK(hipMalloc(&red_, sizeof(double) * 3 * kRedBlocks));
        return red_;
    }
    void ot_dot(int n, const double* a, const double* b, double* out) {
        k_dot_partial<<<kRedBlocks, 256, 0, stream_>>>(n, a, b, red_part());
        k_sum_final<<<1, 256, 0, stream_>>>(red_part(), out);
        check();
    }
    // vector helpers of the host-driven optimisers: a . b (fixed-order
    // reduction, read back), out = a x, y += a x
    double vdot(int n, const double* a, const double* b) {
        if (!vdot_) VGC_HIP_CHECK(hipMalloc(&vdot_, sizeof(double)));
        ot_dot(n, a, b, vdot_);
        double r = 0.0;
        d2h(&r, vdot_, sizeof(double));
        return r;
    }
    void vscale(int n, double a, const double* x, double* out) {
        k_vscale<<<blocks_for(n, 256), 256, 0, stream_>>>(n, a, x, out);
        check();
    }
    void vaxpy(int n, double a, const double* x, double* y) {
        k_vaxpy<<<blocks_for(n, 256), 256, 0, stream_>>>(n, a, x, y);
        check();
    }
    void red_sum(int n, const double* a, double* out) {
        k_sum_partial<<<kRedBlocks, 256, 0, stream_>>>(n, a, red_part());
        k_sum_final<<<1, 256, 0, stream_>>>(red_part(), out);
        check();
    }
    void cvt_grad(int n, const double* xyz, const double* vol, const double* mg, double* g) {
        k_cvt_grad<<<blocks_for(n, 256), 256, 0, stream_>>>(n, xyz, vol, mg, g);
        check();
    }
    void ot_mass_stats(int n, const double* m, const double* nu, double* out) {
        k_mass_partial<<<kRedBlocks, 256, 0, stream_>>>(n, m, nu, red_part());
        k_mass_final<<<1, 256, 0, stream_>>>(red_part(), out);
        check();
    }
    void ot_min(int n, const double* a, double* out) {
        k_min_partial<<<kRedBlocks, 256, 0, stream_>>>(n, a, red_part());
        k_min_final<<<1, 256, 0, stream_>>>(red_part(), out);
        check();
    }
    void ot_hessian(int n, const double* xyz, const long long* ptr, const int* col,
                    const double* area, double* val, double* diag, const double* period = nullptr) {
        k_ot_hessian<<<blocks_for(n, 256), 256, 0, stream_>>>(n, xyz, ptr, col, area, val, diag,
                                                              period);
        check();
    }
    void ot_convex_start(int n, const double* xyz, const double* piece, int m, double s,
                         double* w) {
        k_ot_convex_start<<<blocks_for(n, 256), 256, 0, stream_>>>(n, xyz, piece, m, s, w);
        check();
    }
    void ot_spmv(int n, const long long* ptr, const int* col, const double* val, const double* diag,
                 const double* x, double* y) {
        k_ot_spmv<<<blocks_for(n, 256), 256, 0, stream_>>>(n, ptr, col, val, diag, x, y);
        check();
    }
    void ot_rhs(int n, const double* nu, const double* m, double* b, double* tmp) {
        k_ot_rhs<<<blocks_for(n, 256), 256, 0, stream_>>>(n, nu, m, b);
        red_sum(n, b, tmp);
        k_sub_mean<<<blocks_for(n, 256), 256, 0, stream_>>>(n, b, tmp);
        check();
    }
    void ot_axpy(int n, const double* w, double t, const double* d, double* out) {
        k_ot_axpy<<<blocks_for(n, 256), 256, 0, stream_>>>(n, w, t, d, out);
        check();
    }
    void ot_pcg_init(int n, const double* b, const double* diag, double* x, double* r, double* z,
                     double* p) {
        k_ot_pcg_init<<<blocks_for(n, 256), 256, 0, stream_>>>(n, b, diag, x, r, z, p);
        check();
    }
    void ot_pcg_update(int n, double alpha, const double* p, const double* q, const double* diag,
                       double* x, double* r, double* z) {
        k_ot_pcg_update<<<blocks_for(n, 256), 256, 0, stream_>>>(n, alpha, p, q, diag, x, r, z);
        check();
    }
    void ot_pcg_dir(int n, double beta, const double* z, double* p) {
        k_ot_pcg_dir<<<blocks_for(n, 256), 256, 0, stream_>>>(n, beta, z, p);
        check();
    }
    // device-resident PCG (scalars in cg[8])
    void cg_start(int n, const double* b, const double* diag, double* x, double* r, double* z,
                  double* p, double* cg, double tol2) {
        k_cg_start_partial<<<kRedBlocks, 256, 0, stream_>>>(n, b, diag, x, r, z, p, red_part());
        k_cg_start_final<<<1, 256, 0, stream_>>>(red_part(), cg, tol2);
        check();
    }
    // k iterations as one hipGraph, captured once per argument set and
    // replayed until the solve ends (the kernels read their scalars from cg)
    void cg_iters(int k, int n, const long long* ptr, const int* col, const double* val,
                  const double* diag, double* x, double* r, double* z, double* p, double* q,
                  double* cg) {
        double* part = cg_part(n);  // allocated outside the capture
        const void* key[13] = {(const void*)(long long)k, (const void*)(long long)n, ptr, col, val,
                               diag, x, r, z, p, q, cg, part};
        bool same = cg_exec_ != nullptr;
        for (int i = 0; i < 13 && same; ++i) same = cg_key_[i] == key[i];
        if (!same) {
            if (cg_exec_) VGC_HIP_CHECK(hipGraphExecDestroy(cg_exec_));
            cg_exec_ = nullptr;
            hipGraph_t g = nullptr;
            VGC_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < k; ++i) cg_iter(n, ptr, col, val, diag, x, r, z, p, q, cg, i & 1);
            k_cg_close<<<1, 256, 0, stream_>>>(cg, part, (k - 1) & 1);
            VGC_HIP_CHECK(hipStreamEndCapture(stream_, &g));
            VGC_HIP_CHECK(hipGraphInstantiate(&cg_exec_, g, nullptr, nullptr, 0));
            VGC_HIP_CHECK(hipGraphDestroy(g));
            for (int i = 0; i < 13; ++i) cg_key_[i] = key[i];
        }
        VGC_HIP_CHECK(hipGraphLaunch(cg_exec_, stream_));
    }
    void cg_iter(int n, const long long* ptr, const int* col, const double* val,
                 const double* diag, double* x, double* r, double* z, double* p, double* q,
                 double* cg, int par) {
        double* part = cg_part(n);
        k_cg_dir<<<kRedBlocks, 256, 0, stream_>>>(n, z, p, cg, part, par);
        const unsigned gs = blocks_for((long long)n * kSpmvLanes, 256);
        k_cg_spmv<<<gs < kSpmvBlocks ? gs : kSpmvBlocks, 256, 0, stream_>>>(n, ptr, col, val, diag,
                                                                          p, q, cg);
        k_cg_dot<<<kRedBlocks, 256, 0, stream_>>>(n, p, q, cg, part);
        k_cg_update<<<kRedBlocks, 256, 0, stream_>>>(n, p, q, diag, x, r, z, cg, part, par);
        check();
    }
    // block partials (p.q, and r.z, r.r of the two update parities)
    double* cg_part(int) {
        const long long need = 5ll * kRedBlocks;
        if (cg_part_cap_ < need) {
            if (cg_part_) VGC_HIP_CHECK(hipFree(cg_part_));
            VGC_HIP_CHECK(hipMalloc(&cg_part_, sizeof(double) * (size_t)need));
            cg_part_cap_ = need;
        }
        return cg_part_;
    }
    // Morton order of a Newton system (orig: Morton position -> caller index)
    void inv_perm(int n, const int* orig, int* inv) {
        k_inv_perm<<<blocks_for(n, 256), 256, 0, stream_>>>(n, orig, inv);
        check();
    }
    void perm_row_len(int n, const int* orig, const long long* ptr, int* len) {
        k_perm_row_len<<<blocks_for(n, 256), 256, 0, stream_>>>(n, orig, ptr, len);
        check();
    }
    void perm_rows(int n, const int* orig, const int* inv, const long long* ptr, const int* col,
                   const double* val, const long long* ptr_m, int* col_m, double* val_m) {
        k_perm_rows<<<blocks_for((long long)n * kSpmvLanes, 256), 256, 0, stream_>>>(
            n, orig, inv, ptr, col, val, ptr_m, col_m, val_m);
        check();
    }
    void range_map(int n, const int* inst, const long long* soff, const long long* off,
                   const int* orig, int* out) {
        if (n > 0) k_range_map<<<blocks_for(n, 256), 256, 0, stream_>>>(n, inst, soff, off, orig, out);
        check();
    }
    void gather_f64(int n, const int* orig, const double* src, double* dst) {
        k_gather_f64<<<blocks_for(n, 256), 256, 0, stream_>>>(n, orig, src, dst);
        check();
    }
    void scatter_f64(int n, const int* orig, const double* src, double* dst) {
        k_scatter_f64<<<blocks_for(n, 256), 256, 0, stream_>>>(n, orig, src, dst);
        check();
    }

    // --------------------------------------------------------- batched OT
    // block partials of the batched reductions: count doubles
    double* seg_part(long long count) {
        if (seg_part_cap_ < count) {
            if (seg_part_) VGC_HIP_CHECK(hipFree(seg_part_));
            VGC_HIP_CHECK(hipMalloc(&seg_part_, sizeof(double) * (size_t)count));
            seg_part_cap_ = count;
        }
        return seg_part_;
    }
    void seg_mass_stats(const SegTable& T, const double* m, const double* nu, double* out) {
        k_seg_mass<<<T.n_inst, 256, 0, stream_>>>(T, m, nu, out);
        check();
    }
    void seg_rhs(const SegTable& T, int n, const int* inst_of, const int* active, const double* nu,
                 const double* m, double* b) {
        k_seg_rhs<<<blocks_for(n, 256), 256, 0, stream_>>>(n, inst_of, active, nu, m, b);
        k_seg_center<<<T.n_inst, 256, 0, stream_>>>(T, b);
        check();
    }
    void seg_axpy(int n, const int* inst_of, const double* w, const double* t, const double* d,
                  double* out) {
        k_seg_axpy<<<blocks_for(n, 256), 256, 0, stream_>>>(n, inst_of, w, t, d, out);
        check();
    }
    void seg_copy(int n, const int* inst_of, const int* sel, const double* src, double* dst) {
        k_seg_copy<<<blocks_for(n, 256), 256, 0, stream_>>>(n, inst_of, sel, src, dst);
        check();
    }
    void cgb_start(const SegTable& T, const double* b, const double* diag, double* x, double* r,
                   double* z, double* p, double* cg, const double* tol2) {
        double* part = seg_part(2 * (long long)T.nblk);
        k_cgb_start_partial<<<T.nblk, 256, 0, stream_>>>(T, b, diag, x, r, z, p, part);
        k_cgb_start_final<<<T.n_inst, 256, 0, stream_>>>(T, part, cg, tol2);
        check();
    }
    // k batched CG iterations, captured once as a graph (for these buffers and
    // this table) and replayed
    void cgb_iters(int k, const SegTable& T, int n, const long long* ptr, const int* col,
                   const double* val, const double* diag, double* x, double* r, double* z,
                   double* p, double* q, double* cg) {
        double* part = seg_part(5 * (long long)T.nblk);  // allocated outside the capture
        const void* key[22] = {(const void*)(long long)k, (const void*)(long long)n,
                               (const void*)(long long)T.n_inst, (const void*)(long long)T.nblk,
                               T.off, T.blk_inst, T.blk_idx, T.cnt, T.first, T.inst_of, ptr, col,
                               val, diag, x, r, z, p, q, cg, part, nullptr};
        bool same = cgb_exec_ != nullptr;
        for (int i = 0; i < 22 && same; ++i) same = cgb_key_[i] == key[i];
        if (!same) {
            if (cgb_exec_) VGC_HIP_CHECK(hipGraphExecDestroy(cgb_exec_));
            cgb_exec_ = nullptr;
            hipGraph_t g = nullptr;
            VGC_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            cgb_launch(k, T, n, ptr, col, val, diag, x, r, z, p, q, cg, part);
            VGC_HIP_CHECK(hipStreamEndCapture(stream_, &g));
            VGC_HIP_CHECK(hipGraphInstantiate(&cgb_exec_, g, nullptr, nullptr, 0));
            VGC_HIP_CHECK(hipGraphDestroy(g));
            for (int i = 0; i < 22; ++i) cgb_key_[i] = key[i];
        }
        VGC_HIP_CHECK(hipGraphLaunch(cgb_exec_, stream_));
    }
    void cgb_launch(int k, const SegTable& T, int n, const long long* ptr, const int* col,
                    const double* val, const double* diag, double* x, double* r, double* z,
                    double* p, double* q, double* cg, double* part) {
        for (int it = 0; it < k; ++it) {
            k_cgb_dir<<<T.nblk, 256, 0, stream_>>>(T, z, p, cg, part, it & 1);
            const unsigned gs = blocks_for((long long)n * kSpmvLanes, 256);
            k_cgb_spmv<<<gs < kSpmvBlocks ? gs : kSpmvBlocks, 256, 0, stream_>>>(
                n, T.inst_of, ptr, col, val, diag, p, q, cg);
            k_cgb_dot<<<T.nblk, 256, 0, stream_>>>(T, p, q, cg, part);
            k_cgb_update<<<T.nblk, 256, 0, stream_>>>(T, p, q, diag, x, r, z, cg, part, it & 1);
        }
        k_cgb_close<<<T.n_inst, 256, 0, stream_>>>(T, cg, part, (k - 1) & 1);
        check();
    }
    // ---------------------------------------------------------------- MLP
    void mlp_forward(const float* P, const float* X, int n, float* out) {
        int waves = (n + kMlpRowsFwd - 1) / kMlpRowsFwd;
        k_mlp_forward<<<blocks_for(waves, 4), 256, 0, stream_>>>(P, X, n, out);
        check();
    }
    int mlp_grad_waves(int n) { return (n + kMlpRowsGrad - 1) / kMlpRowsGrad; }
    void mlp_grad(const float* P, const float* X, const float* T, int n, float scale, float* part,
                  float* lpart, float* grad) {
        int waves = mlp_grad_waves(n);
        k_mlp_grad<<<blocks_for(waves, 4), 256, 0, stream_>>>(P, X, T, n, scale, part, lpart);
        k_mlp_reduce<<<blocks_for(kMlpParams, 256), 256, 0, stream_>>>(part, waves, grad);
        check();
    }
    void mlp_adam(float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                  float c1, float c2, float eps) {
        k_mlp_adam<<<blocks_for(kMlpParams, 256), 256, 0, stream_>>>(p, g, m, v, lr, b1, b2, c1, c2,
                                                                      eps);
        check();
    }
    // gradient and loss with caller-provided scratch (no allocation)
    void mlp_grad_scratch(const float* P, const float* X, const float* T, int n, float* part,
                          float* lpart, float* grad, double* loss) {
        mlp_grad(P, X, T, n, 1.0f / float(n), part, lpart, grad);
        if (loss) k_mlp_loss<<<1, 64, 0, stream_>>>(lpart, mlp_grad_waves(n), n, loss);
        check();
    }
    void ot_features(int n, const double* xyz, const double* nu, const float* d2,
                     const double* box, double inv_cell, double h, float* f) {
        k_ot_features<<<blocks_for(n, 256), 256, 0, stream_>>>(n, xyz, nu, d2, box, inv_cell, h, f);
        check();
    }

    // ------------------------------------------------------- exact kNN
    void surface_items(const double* V, const int* F, const double* vw, int nt, double* ov,
                       double* ow) {
        k_surface_items<<<blocks_for(nt, 256), 256, 0, stream_>>>(V, F, vw, nt, ov, ow);
        check();
    }
    void surface_v0(const double* tv, int n, double* q) {
        k_surface_v0<<<blocks_for(n, 256), 256, 0, stream_>>>(tv, n, q);
        check();
    }
    void surface_pass(const SurfParams& P) {
        k_surface_pass<<<blocks_for(P.n_items, 256), 256, 0, stream_>>>(P);
        check();
    }
    void surface_next(const double* tv, const double* tw, const int* lst, int nsplit, int nredo,
                      double* ov, double* ow) {
        k_surface_next<<<blocks_for(nsplit + nredo, 256), 256, 0, stream_>>>(tv, tw, lst, nsplit,
                                                                              nredo, ov, ow);
        check();
    }
    void surface_reduce(int n, const unsigned long long* keys, const int* idx, long long nrec,
                        const double* val, const int* flag, double* m, double* mg) {
        k_surface_reduce<<<blocks_for(n, 256), 256, 0, stream_>>>(n, keys, idx, nrec, val, flag, m,
                                                                   mg);
        check();
    }
    void piece_setup(int nq, const int* cells, const pt4* pts, const int* inst_of,
                     const double* box, double* p_box) {
        k_piece_setup<<<blocks_for(nq, 256), 256, 0, stream_>>>(nq, cells, pts, inst_of, box, p_box);
        check();
    }
    void piece_merge(int m, const int* cells, const int* orig, const long long* pptr,
                     const int* pidx, const double* p_vol, const double* p_mg, const double* p_e2,
                     const int* p_cnt, const int* p_nbr, const double* p_area,
                     const uint32_t* p_flags, int stride, const long long* x_off, double* vol,
                     double* mg, double* e2, int* nbr_cnt, uint32_t* flags, int* x_nbr,
                     double* x_area) {
        k_piece_merge<<<blocks_for(m, 64), 64, 0, stream_>>>(m, cells, orig, pptr, pidx, p_vol, p_mg,
                                                             p_e2, p_cnt, p_nbr, p_area, p_flags,
                                                             stride, x_off, vol, mg, e2, nbr_cnt,
                                                             flags, x_nbr, x_area);
        check();
    }
    void piece_copy(int m, const int* cells, const int* orig, const long long* x_off,
                    const int* x_nbr, const double* x_area, const int* nbr_cnt,
                    const long long* ptr, int* nbr, double* area) {
        k_piece_copy<<<blocks_for(m, 64), 64, 0, stream_>>>(m, cells, orig, x_off, x_nbr, x_area,
                                                            nbr_cnt, ptr, nbr, area);
        check();
    }
    void ot_repair(int n, const double* xyz, const double* w, const double* vol, const pt4* pts,
                   const bnode* nodes, const InstLayout& IL, float cs2, double delta,
                   double* wout) {
        k_ot_repair<<<blocks_for(n, 256), 256, 0, stream_>>>(n, xyz, w, vol, pts, nodes, IL, cs2,
                                                              delta, wout);
        check();
    }
    void iota(int* d, int n) {
        if (n > 0) k_iota<<<blocks_for(n, 256), 256, 0, stream_>>>(d, n);
        check();
    }
    void lower_bound_u64(const uint64_t* keys, int n, const uint64_t* q, int nq, int* out) {
        if (nq > 0) k_lower_bound_u64<<<blocks_for(nq, 256), 256, 0, stream_>>>(keys, n, q, nq, out);
        check();
    }
    void knn_seed_radius(const pf4* pf, const bnode* nodes, const InstLayout& IL, const pf4* qf,
                         const int* qpos, int nq, int kk, double e, float* rr) {
        // kk <= 8: the narrowest walk that holds kk (the radius of the
        // 16-wide walk); kk > 16: the 8th nearest scaled, a first guess like
        // the scaled 16th (the lists are exact either way).  On the surface
        // leg (327,680 triangles, 100k seeds, 33-NN lists) the call took
        // 47 ms with the 16th, 27 ms with the 8th, 31 ms with the 4th.
        const unsigned g = blocks_for(nq, 256);
        if (kk == 1)
            k_knn_seed_radius<1><<<g, 256, 0, stream_>>>(pf, nodes, IL, qf, qpos, nq, kk, e, rr);
        else if (kk <= 4)
            k_knn_seed_radius<4><<<g, 256, 0, stream_>>>(pf, nodes, IL, qf, qpos, nq, kk, e, rr);
        else if (kk <= 8 || kk > kKnnSeedK)
            k_knn_seed_radius<8><<<g, 256, 0, stream_>>>(pf, nodes, IL, qf, qpos, nq, kk, e, rr);
        else
            k_knn_seed_radius<kKnnSeedK><<<g, 256, 0, stream_>>>(pf, nodes, IL, qf, qpos, nq, kk, e,
                                                                 rr);
        check();
    }
    void knn_count(const pf4* pf, const bnode* nodes, const InstLayout& IL, const pf4* qf,
                   const int* items, int ni, const float* rr, int* cnt) {
        if (ni > 0) k_knn_count<<<blocks_for(ni, 64), 64, 0, stream_>>>(pf, nodes, IL, qf, items, ni, rr, cnt);
        check();
    }
    void knn_fill(const pf4* pf, const bnode* nodes, const InstLayout& IL, const pf4* qf,
                  const int* items, int ni, const float* rr, const long long* off, int* lm,
                  float* lf) {
        if (ni > 0)
            k_knn_fill<<<blocks_for(ni, 64), 64, 0, stream_>>>(pf, nodes, IL, qf, items, ni, rr, off, lm, lf);
        check();
    }
    void knn_d64(const pt4* qp, const pt4* pts, const int* orig, const int* items, int ni,
                 const long long* off, const int* lm, int* li, double* ld) {
        if (ni > 0)
            k_knn_d64<<<blocks_for(ni, 64), 64, 0, stream_>>>(qp, pts, orig, items, ni, off, lm, li, ld);
        check();
    }
    template <class K, class V>
    void seg_sort_kv(const K* kin, const V* vin, K* kout, V* vout, long long total, int nseg,
                     const long long* off) {
        if (total <= 0) return;
        if (total > 0x7FFFFFFFll || nseg < 0)
            throw std::runtime_error("seg_sort: list chunk exceeds 2^31 entries");
        size_t tb = 0;
        VGC_HIP_CHECK(rocprim::segmented_radix_sort_pairs(nullptr, tb, kin, kout, vin, vout,
                                                          (unsigned)total, (unsigned)nseg, off,
                                                          off + 1, 0, int(8 * sizeof(K)), stream_));
        void* t = temp(tb);
        VGC_HIP_CHECK(rocprim::segmented_radix_sort_pairs(t, tb, kin, kout, vin, vout,
                                                          (unsigned)total, (unsigned)nseg, off,
                                                          off + 1, 0, int(8 * sizeof(K)), stream_));
    }
    void knn_select(const int* items, int ni, const long long* off, const int* cnt,
                    const double* ld, const int* li, float* rr, double e, int kk, int n,
                    int self_rule, const int* qrow, int* out_idx, double* out_d2, int* redo,
                    unsigned* nredo) {
        if (ni > 0)
            k_knn_select<<<blocks_for(ni, 256), 256, 0, stream_>>>(items, ni, off, cnt, ld, li, rr, e,
                                                                   kk, n, self_rule, qrow, out_idx,
                                                                   out_d2, redo, nredo);
        check();
    }

    void knn_to_caller(const int* orig, const int* list, const float* d2, int n, int K, int* out,
                       float* outd) {
        k_knn_to_caller<<<blocks_for(n, 256), 256, 0, stream_>>>(orig, list, d2, n, K, out, outd);
        check();
    }

private:
    int dev_;
    hipStream_t stream_ = nullptr;
    hipStream_t caller_ = nullptr;
    hipEvent_t enter_ev_ = nullptr;
    hipEvent_t ev_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    void* temp_ = nullptr;
    size_t temp_bytes_ = 0;
    double* red_ = nullptr;
    double* vdot_ = nullptr;
    hipGraphExec_t cg_exec_ = nullptr;
    const void* cg_key_[13] = {};
    double* cg_part_ = nullptr;
    long long cg_part_cap_ = 0;
    hipGraphExec_t cgb_exec_ = nullptr;
    const void* cgb_key_[22] = {};
    double* seg_part_ = nullptr;
    long long seg_part_cap_ = 0;
    int tree_leaves_ = 1;
    unsigned char* wide_ = nullptr;
    int wide_cap_ = 0;
    int* wide_front_ = nullptr;
    long long wide_front_cap_ = 0;
    int waves_ = kDefaultClipWaves;
};

typedef Ctx<HipBackend> HipCtx;

}  // namespace vgc

struct vgc_ctx {
    vgc::HipCtx* impl;
};

namespace {
// address of `scratch` on each device, looked up once per device at context
// creation with that device current
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

void lookup_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(vgc::scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch[device] = static_cast<float*>(p);
}
}  // namespace

// Every call makes the context's device current and orders the context's
// stream after the work queued on the caller's stream before the call.
static int vgc_enter(vgc_ctx* ctx) {
    if (!ctx || !ctx->impl) return VGC_ERR_ARG;
    if (hipSetDevice(ctx->impl->device) != hipSuccess) return VGC_ERR_DEVICE;
    try {
        ctx->impl->be.enter();
    } catch (const std::exception& e) {
        ctx->impl->err = e.what();
        return VGC_ERR_DEVICE;
    }
    return VGC_OK;
}

extern "C" {

int vgc_create(vgc_ctx** ctx, int device) {
    if (!ctx) return VGC_ERR_ARG;
    *ctx = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return VGC_ERR_DEVICE;
    try {
        vgc_ctx* c = new vgc_ctx;
        c->impl = new vgc::HipCtx(device);  // makes the device current
        lookup_scratch(device);
        *ctx = c;
        return VGC_OK;
    } catch (const std::exception&) {
        return VGC_ERR_DEVICE;
    }
}

void vgc_destroy(vgc_ctx* ctx) {
    if (!ctx) return;
    delete ctx->impl;
    delete ctx;
}

int vgc_set_stream(vgc_ctx* ctx, void* stream) {
    if (!ctx || !ctx->impl) return VGC_ERR_ARG;
    ctx->impl->be.set_caller_stream((hipStream_t)stream);
    return VGC_OK;
}

const char* vgc_last_error(const vgc_ctx* ctx) {
    if (!ctx || !ctx->impl) return "null context";
    return ctx->impl->err.c_str();
}

int vgc_knn(vgc_ctx* ctx, const double* xyz, uint32_t n, uint32_t k, uint32_t flags, int32_t* nbr,
            double* d2) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_knn(ctx->impl, xyz, n, k, flags, nbr, d2);
}

int vgc_knn_search(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* q, uint32_t nq,
                   uint32_t k, uint32_t flags, int32_t* nbr, double* d2) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_knn_search(ctx->impl, xyz, n, q, nq, k, flags, nbr, d2);
}

int vgc_cells(vgc_ctx* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
              uint32_t k0, uint32_t flags, double* vol, double* mg, uint64_t* nbr_ptr,
              int32_t* nbr, double* farea, uint64_t cap, vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_cells(ctx->impl, xyz, w, &n, 1u, box, k0, flags, vol, mg, nbr_ptr, nbr, farea,
                          cap, stats);
}

int vgc_cells_batch(vgc_ctx* ctx, const double* xyz, const double* w, const uint32_t* inst_n,
                    uint32_t n_inst, const double* boxes, uint32_t k0, uint32_t flags, double* vol,
                    double* mg, uint64_t* nbr_ptr, int32_t* nbr, double* farea, uint64_t cap,
                    vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_cells(ctx->impl, xyz, w, inst_n, n_inst, boxes, k0, flags, vol, mg, nbr_ptr, nbr,
                          farea, cap, stats);
}

int vgc_cvt_func_grad(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* box,
                      uint32_t flags, double* f, double* g, vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_cvt_func_grad(ctx->impl, xyz, n, box, flags, f, g, stats);
}

int vgc_lloyd(vgc_ctx* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
              uint32_t k0, uint32_t flags, vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_lloyd(ctx->impl, xyz, n, box, iters, k0, flags, stats);
}
int vgc_cell_mesh(vgc_ctx* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
                  double shrink, uint32_t flags, uint64_t* cell_vptr, uint64_t* cell_fptr,
                  double* vertices, uint64_t* facet_ptr, int32_t* facet_vertices,
                  int32_t* facet_nbr, int32_t* tets, const vgc_mesh_sizes* cap,
                  vgc_mesh_sizes* need, vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_cell_mesh(ctx->impl, "vgc_cell_mesh", xyz, w, n, box, shrink, flags, cell_vptr,
                              cell_fptr, vertices, facet_ptr, facet_vertices, facet_nbr, tets, cap,
                              need, stats);
}
int vgc_surface_centroids(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* vtx,
                          uint32_t nv, const int32_t* tri, uint32_t nt, const double* vw,
                          uint32_t flags, double* mg, double* m, vgc_surface_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_surface_centroids(ctx->impl, xyz, n, vtx, nv, tri, nt, vw, flags, mg, m, stats);
}
int vgc_rdt(vgc_ctx* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
            uint32_t flags, int32_t* tets, uint64_t cap, uint64_t* nb_tets, vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_rdt(ctx->impl, xyz, w, n, box, flags, tets, cap, nb_tets, stats);
}
int vgc_cvt_newton(vgc_ctx* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
                   uint32_t m, uint32_t flags, vgc_opt_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_cvt_newton(ctx->impl, xyz, n, box, iters, m, flags, stats);
}
int vgc_lloyd_batch(vgc_ctx* ctx, double* xyz, const uint32_t* inst_n, uint32_t n_inst,
                    const double* boxes, uint32_t iters, uint32_t k0, uint32_t flags,
                    vgc_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_lloyd_batch(ctx->impl, xyz, inst_n, n_inst, boxes, iters, k0, flags, stats);
}

int vgc_ot_solve(vgc_ctx* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                 const double* box, double tol, uint32_t max_newton, uint32_t flags,
                 vgc_ot_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_ot_solve(ctx->impl, xyz, nu, w, n, box, tol, max_newton, flags, stats);
}

int vgc_ot_solve_levels(vgc_ctx* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                        const double* box, const uint32_t* levels, uint32_t n_levels, double tol,
                        uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_ot_solve_levels(ctx->impl, xyz, nu, w, n, box, levels, n_levels, tol,
                                    max_newton, flags, stats);
}

int vgc_ot_solve_batch(vgc_ctx* ctx, const double* xyz, const double* nu, double* w,
                       const uint32_t* inst_n, uint32_t n_inst, const double* boxes, double tol,
                       uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_ot_solve_batch(ctx->impl, xyz, nu, w, inst_n, n_inst, boxes, tol, max_newton,
                                   flags, stats);
}
int vgc_ot_features(vgc_ctx* ctx, const double* xyz, const double* nu, uint32_t n,
                    const double* box, uint32_t flags, float* feats) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_ot_features(ctx->impl, xyz, nu, n, box, flags, feats);
}
int vgc_ot_convex_start(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* pieces,
                        uint32_t m, double s, uint32_t flags, double* w) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_ot_convex_start(ctx->impl, xyz, n, pieces, m, s, flags, w);
}
int vgc_mlp_forward(vgc_ctx* ctx, const float* params, const float* feats, uint32_t n,
                    uint32_t flags, float* out) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_mlp_forward(ctx->impl, params, feats, n, flags, out);
}
int vgc_mlp_grad(vgc_ctx* ctx, const float* params, const float* feats, const float* target,
                 uint32_t n, uint32_t flags, float* grad, double* loss) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_mlp_grad(ctx->impl, params, feats, target, n, flags, grad, loss);
}
int vgc_mlp_grad_noalloc(vgc_ctx* ctx, const float* params, const float* feats,
                         const float* target, uint32_t n, float* grad, double* loss) {
    if (int rc = vgc_enter(ctx)) return rc;
    vgc::HipCtx* c = ctx->impl;
    if (!params || !feats || !target || !grad || n == 0) {
        c->err = "vgc_mlp_grad_noalloc: bad argument";
        return VGC_ERR_ARG;
    }
    if (n > (uint32_t)vgc::kNoAllocRows) {
        c->err = "vgc_mlp_grad_noalloc: " + std::to_string(n) + " rows exceed the static scratch (" +
                 std::to_string(vgc::kNoAllocRows) + " rows at most)";
        return VGC_ERR_CAPACITY;
    }
    const int dev = c->device;
    float* base = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
    if (!base) {
        c->err = std::string("vgc_mlp_grad_noalloc: no scratch address for this device (") +
                 hipGetErrorString(dev >= 0 && dev < kMaxDevices ? g_scratch_err[dev]
                                                                  : hipErrorInvalidDevice) +
                 ")";
        return VGC_ERR_DEVICE;
    }
    try {
        float* part = base;
        float* lpart = base + (size_t)vgc::kNoAllocWaves * vgc::kMlpParams;
        c->be.mlp_grad_scratch(params, feats, target, (int)n, part, lpart, grad, loss);
        c->be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        c->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

int vgc_mlp_adam(vgc_ctx* ctx, float* params, const float* grad, float* m, float* v,
                 uint32_t step, float lr, uint32_t flags) {
    if (int rc = vgc_enter(ctx)) return rc;
    return vgc::api_mlp_adam(ctx->impl, params, grad, m, v, step, lr, flags);
}

int vgc_random_float64(uint64_t seed, double* out, uint64_t count) {
    return vgc::random_float64(seed, out, count);
}

int vgc_sample_points(uint32_t kind, uint64_t seed, uint32_t n, double* xyz) {
    return vgc::sample_points(kind, seed, n, xyz);
}
}

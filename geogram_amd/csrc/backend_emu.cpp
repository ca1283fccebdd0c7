// CPU emulation of the gfx950 pipeline, for tests on machines without a GPU.
// It runs the very kernels' code (clip_cell.h through the emulated wave layer,
// the scalar kNN / tree / radius functions of common.h) and the same host
// driver (pipeline.h).  It is a test library (libvgc_emu.so); the product
// library libvgc.so is the HIP build and has no CPU path.
#define VGC_EMU 1

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

#include "api_impl.h"
#include "mlp.h"
#include "ot_kernels.h"

namespace vgc {

struct EmuBackend {
    explicit EmuBackend(int) {}

    void* dalloc(size_t b) {
        void* p = std::malloc(b ? b : 1);
        if (!p) throw std::bad_alloc();
        return p;
    }
    void dfree(void* p) { std::free(p); }
    void h2d(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
    void d2h(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
    void d2d(void* d, const void* s, size_t b) { std::memmove(d, s, b); }
    void dzero(void* d, size_t b) { std::memset(d, 0, b); }
    void sync() {}
    void event_begin(int) {}
    void event_end(int) {}
    double event_ms(int) { return 0.0; }

    void fill_inst(int* inst_orig, int n, const int* p0, int ni) {
        for (int i = 0; i < n; ++i) {
            int lo = 0, hi = ni - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) / 2;
                if (p0[mid] <= i)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            inst_orig[i] = lo;
        }
    }

    void morton(const double* xyz, const int* inst_orig, const double* box, int n,
                uint64_t* keys, int* vals) {
        for (int i = 0; i < n; ++i) {
            int inst = inst_orig ? inst_orig[i] : 0;
            keys[i] = morton_key(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2],
                                 box + 6 * inst, (uint64_t)inst);
            vals[i] = i;
        }
    }

    // stable sort by key (as a radix sort is)
    void sort_pairs(const uint64_t* keys, const int* vals, uint64_t* keys2, int* vals2, int n,
                    int) {
        std::vector<int> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
        for (int i = 0; i < n; ++i) {
            keys2[i] = keys[idx[i]];
            vals2[i] = vals[idx[i]];
        }
    }

    void gather(const double* xyz, const double* w, const int* inst_orig, const int* perm, int n,
                pt4* pts, pf4* pf, int* orig, int* inst_of) {
        for (int s = 0; s < n; ++s) {
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
    }

    void bvh_leaves(const pt4* pts, const InstLayout& IL, bnode* nodes, const long long* leaf_off,
                    int ni, long long) {
        for (int inst = 0; inst < ni; ++inst)
            for (long long l = 0; l < leaf_off[inst + 1] - leaf_off[inst]; ++l)
                bvh_leaf(pts, IL, nodes, inst, (int)l);
    }

    void bvh_level(const InstLayout& IL, bnode* nodes, const long long*, int ni, long long, int h) {
        for (int inst = 0; inst < ni; ++inst) {
            int L = IL.L[inst];
            for (int k = L >> h; k < (2 * L) >> h; ++k) bvh_merge(nodes, IL.t0[inst], k);
        }
    }

    void wmax(const pt4* pts, const InstLayout& IL, int ni, double* out) {
        for (int inst = 0; inst < ni; ++inst) {
            double m = -INFINITY;
            for (int s = IL.p0[inst]; s < IL.p0[inst] + IL.n[inst]; ++s) m = std::fmax(m, pts[s].w);
            out[inst] = m;
        }
    }

    template <int K>
    void knn_k(const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* queries, int nq,
               int* out, float* out_d2) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int t = 0; t < nq; ++t) {
            int q = queries ? queries[t] : t;
            knn_query<K>(pf, nodes, IL, q, out + (size_t)t * K, out_d2 ? out_d2 + (size_t)t * K : nullptr);
        }
    }

    void knn(int K, const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* queries,
             int nq, int* out, float* out_d2) {
        switch (K) {
            case 16: knn_k<16>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 24: knn_k<24>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 32: knn_k<32>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 48: knn_k<48>(pf, nodes, IL, queries, nq, out, out_d2); break;
            case 64: knn_k<64>(pf, nodes, IL, queries, nq, out, out_d2); break;
            default: throw std::runtime_error("unsupported K");
        }
    }

    template <int S, bool E, bool G>
    void clip_s(const ClipParams& P) {
#pragma omp parallel
        {
            std::unique_ptr<CellLds<S>> L(new CellLds<S>());
            std::memset(L.get(), 0, sizeof(CellLds<S>));
#pragma omp for schedule(dynamic, 16)
            for (int it = 0; it < P.n_items; ++it) {
                if constexpr (S == 1 && G && !E) {
                    if (P.fp32) {
                        // the first pass with FP32 corner state
                        if (P.weighted)
                            process_cell<S, false, G, true, true, float>(P, it, *L);
                        else
                            process_cell<S, false, G, false, false, float>(P, it, *L);
                        continue;
                    }
                }
                if (P.weighted)
                    process_cell<S, E, G, true>(P, it, *L);
                else if (G && P.corner)
                    process_cell<S, E, G, false, true>(P, it, *L);
                else
                    process_cell<S, E, G, false>(P, it, *L);
            }
        }
    }

    template <bool E, bool G>
    void clip_eg(int S, const ClipParams& P) {
        if (S == 1) {
            clip_s<1, E, G>(P);
        } else if (S == 2) {
            clip_s<2, E, G>(P);
        } else if (S == 4) {
            clip_s<4, E, G>(P);
        } else {
            if constexpr (!G)
                clip_s<8, E, false>(P);
            else
                throw std::runtime_error("clip: 512-corner cells run from lists only");
        }
    }

    void set_clip_waves(int w) { waves_ = w == 1 ? 1 : 4; }
    int clip_waves() const { return waves_; }
    int waves_ = kDefaultClipWaves;
    void clip(int S, bool exact, bool gather, const ClipParams& P) {
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
    }

    int det4_sign(const double* r) { return det4_sign_exact(r, r + 4, r + 8, r + 12); }

    void radius_r2(const int* items, int ni, const pt4* pts, const double* r2, const int* inst_of,
                   const double* wmaxv, int weighted, const fball* ball, const double* box,
                   float* rho2, PowerFilter* filt) {
        for (int t = 0; t < ni; ++t) {
            int c = items[t];
            int inst = inst_of ? inst_of[c] : 0;
            rho2[t] = radius_for(r2[c], pts[c].w, weighted ? wmaxv[inst] : 0.0, weighted);
            if (filt) filt[t] = filter_from_ball(ball[c], pts[c], box + 6 * inst);
        }
    }

    void set_tree_leaves(int) {}
    void radius_count(const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* items,
                      int ni, const float* rho2, const PowerFilter* filt, int* cnt) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int t = 0; t < ni; ++t)
            cnt[t] = radius_query(pf, nodes, IL, items[t], rho2[t], filt ? filt + t : nullptr,
                                  nullptr, nullptr);
    }

    void radius_fill(const pf4* pf, const bnode* nodes, const InstLayout& IL, const int* items,
                     int ni, const float* rho2, const PowerFilter* filt, const long long* off,
                     long long base, int* list, float* d2) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int t = 0; t < ni; ++t)
            radius_query(pf, nodes, IL, items[t], rho2[t], filt ? filt + t : nullptr,
                         list + (off[t] - base), d2 + (off[t] - base));
    }

    void rebase(const long long* off, int n, long long base, long long* out) {
        for (int t = 0; t < n; ++t) out[t] = off[t] - base;
    }

    unsigned census(const int* items, int n_items, const uint32_t* flags, uint32_t mask,
                    const unsigned* skip, const int* inst_of, int* out_cell, uint32_t* out_flag,
                    unsigned* d_count) {
        unsigned k = 0;
        for (int t = 0; t < n_items; ++t) {
            const int c = items ? items[t] : t;
            if (skip && skip[inst_of ? inst_of[c] : 0]) continue;
            const uint32_t f = flags[c];
            if ((f & mask) || !(f & CELL_DONE)) {
                out_cell[k] = c;
                out_flag[k] = f;
                ++k;
            }
        }
        *d_count = k;
        return k;
    }
    void piece_setup(int nq, const int* cells, const pt4* pts, const int* inst_of,
                     const double* box, double* p_box) {
        for (int q = 0; q < nq; ++q) piece_setup_row(q, cells, pts, inst_of, box, p_box);
    }
    void piece_merge(int m, const int* cells, const int* orig, const long long* pptr,
                     const int* pidx, const double* p_vol, const double* p_mg, const double* p_e2,
                     const int* p_cnt, const int* p_nbr, const double* p_area,
                     const uint32_t* p_flags, int stride, const long long* x_off, double* vol,
                     double* mg, double* e2, int* nbr_cnt, uint32_t* flags, int* x_nbr,
                     double* x_area) {
        for (int t = 0; t < m; ++t)
            piece_merge_row(t, cells, orig, pptr, pidx, p_vol, p_mg, p_e2, p_cnt, p_nbr, p_area,
                            p_flags, stride, x_off, vol, mg, e2, nbr_cnt, flags, x_nbr, x_area);
    }
    void piece_copy(int m, const int* cells, const int* orig, const long long* x_off,
                    const int* x_nbr, const double* x_area, const int* nbr_cnt,
                    const long long* ptr, int* nbr, double* area) {
        for (int t = 0; t < m; ++t)
            piece_copy_row(t, cells, orig, x_off, x_nbr, x_area, nbr_cnt, ptr, nbr, area);
    }
    void ot_repair(int n, const double* xyz, const double* w, const double* vol, const pt4* pts,
                   const bnode* nodes, const InstLayout& IL, float cs2, double delta,
                   double* wout) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i)
            wout[i] = ot_repair_row(i, xyz, w, vol, pts, nodes, IL, cs2, delta);
    }
    void power_nearest(const pt4* pts, const bnode* nodes, const InstLayout& IL, const double* q,
                       int nq, float cs2, double* wout) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int t = 0; t < nq; ++t) {
            const int j = power_nearest_query(pts, nodes, IL, q[3 * (size_t)t + 0],
                                              q[3 * (size_t)t + 1], q[3 * (size_t)t + 2], cs2);
            wout[t] = pts[j].w;
        }
    }
    void gather_u32(const int* items, int n, const uint32_t* src, uint32_t* out) {
        for (int t = 0; t < n; ++t) out[t] = src[items[t]];
    }
    void scan_i32(const int* in, long long* out, int n) {
        out[0] = 0;
        for (int i = 0; i < n; ++i) out[i + 1] = out[i] + in[i];
    }

    void seg_sort(const float* kin, const int* vin, float* kout, int* vout, long long, int nseg,
                  const long long* off) {
        for (int s = 0; s < nseg; ++s) {
            long long b = off[s], e = off[s + 1];
            std::vector<long long> idx(e - b);
            std::iota(idx.begin(), idx.end(), b);
            std::stable_sort(idx.begin(), idx.end(), [&](long long x, long long y) { return kin[x] < kin[y]; });
            for (long long i = 0; i < e - b; ++i) {
                kout[b + i] = kin[idx[i]];
                vout[b + i] = vin[idx[i]];
            }
        }
    }

    void compact(int n, const int* cnt, const long long* ptr, const int* nbr_tmp,
                 const double* area_tmp, int stride, int* nbr, double* area) {
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < std::min(cnt[i], stride); ++k) {
                nbr[ptr[i] + k] = nbr_tmp[(size_t)i * stride + k];
                area[ptr[i] + k] = area_tmp[(size_t)i * stride + k];
            }
    }

    void lloyd_update(double* xyz, const double* vol, const double* mg, int n,
                      const double* wrap = nullptr) {
        for (int i = 0; i < n; ++i) lloyd_update_row(i, xyz, vol, mg, wrap);
    }
    void ghost_count(const double* xyz, int n, const double* box, double delta, int* cnt) {
        for (int i = 0; i < n; ++i)
            cnt[i] = periodic_images(xyz, nullptr, i, box, delta, nullptr, nullptr, nullptr);
    }
    void ghost_fill(const double* xyz, const double* w, int n, const double* box, double delta,
                    const long long* off, double* xe, double* we, int* src) {
        for (int i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) xe[3 * (long long)i + a] = xyz[3 * (long long)i + a];
            if (we) we[i] = w[i];
            src[i] = i;
            const long long o = n + off[i];
            periodic_images(xyz, w, i, box, delta, xe + 3 * o, we ? we + o : nullptr, src + o);
        }
    }
    void periodic_reach(const pt4* pts, const int* orig, int ne, int n_real, const double* r2,
                        const double* wmax, float* out) {
        float v = 0.0f;
        for (int c = 0; c < ne; ++c)
            if (orig[c] < n_real)
                v = std::max(v, radius_for(r2[c], pts[c].w, wmax ? wmax[0] : 0.0, wmax ? 1 : 0));
        *out = v;
    }

    // ----------------------------------------------------------------- OT
    // reductions in the order of the device kernels (512 blocks x 256
    // threads, grid-stride, halving-stride block trees)
    static constexpr int kRedBlocks = 512;
    static double tree_sum(double* v) {
        for (int off = 128; off > 0; off >>= 1)
            for (int t = 0; t < off; ++t) v[t] += v[t + off];
        return v[0];
    }
    template <class F>
    static double red_sum(int n, F f) {
        std::vector<double> part(kRedBlocks);
        double sh[256];
        for (int b = 0; b < kRedBlocks; ++b) {
            for (int t = 0; t < 256; ++t) {
                double s = 0.0;
                for (long long i = (long long)b * 256 + t; i < n; i += (long long)kRedBlocks * 256)
                    s += f((int)i);
                sh[t] = s;
            }
            part[b] = tree_sum(sh);
        }
        for (int t = 0; t < 256; ++t) {
            double s = 0.0;
            for (int i = t; i < kRedBlocks; i += 256) s += part[i];
            sh[t] = s;
        }
        return tree_sum(sh);
    }
    void red_sum(int n, const double* a, double* out) {
        *out = red_sum(n, [&](int i) { return a[i]; });
    }
    void cvt_grad(int n, const double* xyz, const double* vol, const double* mg, double* g) {
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) {
                const size_t k = 3 * (size_t)i + c;
                g[k] += 2.0 * (vol[i] * xyz[k] - mg[k]);
            }
    }
    double vdot(int n, const double* a, const double* b) {
        return red_sum(n, [&](int i) { return a[i] * b[i]; });
    }
    void vscale(int n, double a, const double* x, double* out) {
        for (int i = 0; i < n; ++i) out[i] = a * x[i];
    }
    void vaxpy(int n, double a, const double* x, double* y) {
        for (int i = 0; i < n; ++i) y[i] += a * x[i];
    }
    void ot_dot(int n, const double* a, const double* b, double* out) {
        *out = red_sum(n, [&](int i) { return a[i] * b[i]; });
    }
    void ot_mass_stats(int n, const double* m, const double* nu, double* out) {
        double mx = 0.0, mn = INFINITY;
        for (int i = 0; i < n; ++i) {
            mx = std::fmax(mx, std::fabs(m[i] - nu[i]) / nu[i]);
            mn = std::fmin(mn, m[i]);
        }
        out[0] = mx;
        out[1] = red_sum(n, [&](int i) {
            double d = m[i] - nu[i];
            return d * d;
        });
        out[2] = mn;
    }
    void ot_min(int n, const double* a, double* out) {
        double mn = INFINITY;
        for (int i = 0; i < n; ++i) mn = std::fmin(mn, a[i]);
        *out = mn;
    }
    void ot_hessian(int n, const double* xyz, const long long* ptr, const int* col,
                    const double* area, double* val, double* diag, const double* period = nullptr) {
        for (int i = 0; i < n; ++i) ot_hessian_row(i, xyz, ptr, col, area, val, diag, period);
    }
    void ot_convex_start(int n, const double* xyz, const double* piece, int m, double s,
                         double* w) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) w[i] = ot_convex_start_row(xyz + 3 * (long long)i, piece, m, s);
    }
    void ot_spmv(int n, const long long* ptr, const int* col, const double* val, const double* diag,
                 const double* x, double* y) {
        for (int i = 0; i < n; ++i) y[i] = ot_spmv_row(i, ptr, col, val, diag, x);
    }
    void ot_rhs(int n, const double* nu, const double* m, double* b, double* tmp) {
        for (int i = 0; i < n; ++i) b[i] = nu[i] - m[i];
        red_sum(n, b, tmp);
        for (int i = 0; i < n; ++i) b[i] -= *tmp / double(n);
    }
    void ot_axpy(int n, const double* w, double t, const double* d, double* out) {
        for (int i = 0; i < n; ++i) out[i] = w[i] + t * d[i];
    }
    void ot_pcg_init(int n, const double* b, const double* diag, double* x, double* r, double* z,
                     double* p) {
        for (int i = 0; i < n; ++i) {
            double ri = b[i], zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
            x[i] = 0.0;
            r[i] = ri;
            z[i] = zi;
            p[i] = zi;
        }
    }
    void ot_pcg_update(int n, double alpha, const double* p, const double* q, const double* diag,
                       double* x, double* r, double* z) {
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            double ri = r[i] - alpha * q[i];
            r[i] = ri;
            z[i] = diag[i] > 0.0 ? ri / diag[i] : 0.0;
        }
    }
    void ot_pcg_dir(int n, double beta, const double* z, double* p) {
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    // device-resident PCG of the gfx950 backend, same operations and orders
    void cg_start(int n, const double* b, const double* diag, double* x, double* r, double* z,
                  double* p, double* cg, double tol2) {
        ot_pcg_init(n, b, diag, x, r, z, p);
        const double rz = red_sum(n, [&](int i) { return r[i] * z[i]; });
        const double bb = red_sum(n, [&](int i) { return r[i] * r[i]; });
        cg[0] = rz;
        cg[1] = bb;
        cg[4] = 0.0;
        cg[5] = tol2 * bb;
        cg[6] = bb > 0.0 ? 0.0 : 1.0;
        cg[7] = 0.0;
    }
    void cg_iters(int k, int n, const long long* ptr, const int* col, const double* val,
                  const double* diag, double* x, double* r, double* z, double* p, double* q,
                  double* cg) {
        for (int i = 0; i < k; ++i) cg_iter(n, ptr, col, val, diag, x, r, z, p, q, cg);
    }
    // the four PCG kernels, reduction orders as on the device
    void cg_iter(int n, const long long* ptr, const int* col, const double* val,
                 const double* diag, double* x, double* r, double* z, double* p, double* q,
                 double* cg) {
        if (cg[6] != 0.0) return;
        const double beta = cg[4];
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double sl[kSpmvLanes];
            for (int l = 0; l < kSpmvLanes; ++l) sl[l] = ot_spmv_lane(i, l, ptr, col, val, p);
            for (int m = kSpmvLanes / 2; m >= 1; m >>= 1) {
                double t[kSpmvLanes];
                for (int l = 0; l < kSpmvLanes; ++l) t[l] = sl[l] + sl[l ^ m];
                for (int l = 0; l < kSpmvLanes; ++l) sl[l] = t[l];
            }
            q[i] = sl[0] + diag[i] * p[i];
        }
        const double pq = red_sum(n, [&](int i) { return p[i] * q[i]; });
        cg[2] = pq;
        if (!(pq > 0.0)) {
            cg[3] = 0.0;
            cg[6] = 2.0;
            return;
        }
        cg[3] = cg[0] / pq;
        ot_pcg_update(n, cg[3], p, q, diag, x, r, z);
        const double rz = red_sum(n, [&](int i) { return r[i] * z[i]; });
        const double rr = red_sum(n, [&](int i) { return r[i] * r[i]; });
        cg[1] = rr;
        cg[7] += 1.0;
        if (rr <= cg[5]) {
            cg[4] = 0.0;
            cg[6] = 1.0;
            return;
        }
        cg[4] = rz / cg[0];
        cg[0] = rz;
    }
    // Morton order of a Newton system
    void inv_perm(int n, const int* orig, int* inv) {
        for (int m = 0; m < n; ++m) inv[orig[m]] = m;
    }
    void perm_row_len(int n, const int* orig, const long long* ptr, int* len) {
        for (int m = 0; m < n; ++m) len[m] = int(ptr[orig[m] + 1] - ptr[orig[m]]);
    }
    void perm_rows(int n, const int* orig, const int* inv, const long long* ptr, const int* col,
                   const double* val, const long long* ptr_m, int* col_m, double* val_m) {
#pragma omp parallel for schedule(static)
        for (int m = 0; m < n; ++m) {
            const int i = orig[m];
            const long long b = ptr[i], e = ptr[i + 1], o = ptr_m[m] - b;
            for (long long k = b; k < e; ++k) {
                col_m[o + k] = inv[col[k]];
                val_m[o + k] = val[k];
            }
        }
    }
    void range_map(int n, const int* inst, const long long* soff, const long long* off,
                   const int* orig, int* out) {
        for (int r = 0; r < n; ++r) out[r] = orig[off[inst[r]] + (r - soff[inst[r]])];
    }
    void gather_f64(int n, const int* orig, const double* src, double* dst) {
        for (int m = 0; m < n; ++m) dst[m] = src[orig[m]];
    }
    void scatter_f64(int n, const int* orig, const double* src, double* dst) {
        for (int m = 0; m < n; ++m) dst[orig[m]] = src[m];
    }

    // --------------------------------------------------------- batched OT
    // same block table and reduction orders as the gfx950 kernels
    std::vector<double> seg_part_;
    template <class F>
    static void seg_block_sums(const SegTable& T, int blk, F f, double& t1, double& t2) {
        const int k = T.blk_inst[blk];
        const long long beg0 = T.off[k] + (long long)T.blk_idx[blk] * 256, end = T.off[k + 1];
        const long long stride = (long long)T.cnt[k] * 256;
        double a[256], b[256];
        for (int t = 0; t < 256; ++t) {
            double s1 = 0.0, s2 = 0.0;
            for (long long i = beg0 + t; i < end; i += stride) f(i, s1, s2);
            a[t] = s1;
            b[t] = s2;
        }
        t1 = tree_sum(a);
        t2 = tree_sum(b);
    }
    static void seg_final(const SegTable& T, int k, const double* part, double& t1, double& t2) {
        double a[256], b[256];
        for (int t = 0; t < 256; ++t) {
            a[t] = t < T.cnt[k] ? part[T.first[k] + t] : 0.0;
            b[t] = t < T.cnt[k] ? part[T.nblk + T.first[k] + t] : 0.0;
        }
        t1 = tree_sum(a);
        t2 = tree_sum(b);
    }
    void seg_mass_stats(const SegTable& T, const double* m, const double* nu, double* out) {
        for (int k = 0; k < T.n_inst; ++k) {
            double a[256], mx = 0.0, mn = INFINITY;
            for (int t = 0; t < 256; ++t) {
                double s2 = 0.0;
                for (long long i = T.off[k] + t; i < T.off[k + 1]; i += 256) {
                    const double d = m[i] - nu[i];
                    mx = std::fmax(mx, std::fabs(d) / nu[i]);
                    s2 += d * d;
                    mn = std::fmin(mn, m[i]);
                }
                a[t] = s2;
            }
            out[3 * k + 0] = mx;
            out[3 * k + 1] = tree_sum(a);
            out[3 * k + 2] = mn;
        }
    }
    void seg_rhs(const SegTable& T, int n, const int* inst_of, const int* active, const double* nu,
                 const double* m, double* b) {
        for (int i = 0; i < n; ++i) b[i] = active[inst_of[i]] ? nu[i] - m[i] : 0.0;
        for (int k = 0; k < T.n_inst; ++k) {
            double a[256];
            for (int t = 0; t < 256; ++t) {
                double s = 0.0;
                for (long long i = T.off[k] + t; i < T.off[k + 1]; i += 256) s += b[i];
                a[t] = s;
            }
            const double mean = tree_sum(a) / double(T.off[k + 1] - T.off[k]);
            for (long long i = T.off[k]; i < T.off[k + 1]; ++i) b[i] -= mean;
        }
    }
    void seg_axpy(int n, const int* inst_of, const double* w, const double* t, const double* d,
                  double* out) {
        for (int i = 0; i < n; ++i) out[i] = w[i] + t[inst_of[i]] * d[i];
    }
    void seg_copy(int n, const int* inst_of, const int* sel, const double* src, double* dst) {
        for (int i = 0; i < n; ++i)
            if (sel[inst_of[i]]) dst[i] = src[i];
    }
    void cgb_start(const SegTable& T, const double* b, const double* diag, double* x, double* r,
                   double* z, double* p, double* cg, const double* tol2) {
        seg_part_.assign(2 * (size_t)T.nblk, 0.0);
        for (int blk = 0; blk < T.nblk; ++blk) {
            double t1, t2;
            seg_block_sums(T, blk, [&](long long i, double& s1, double& s2) {
                const double ri = b[i], zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
                x[i] = 0.0;
                r[i] = ri;
                z[i] = zi;
                p[i] = zi;
                s1 += ri * zi;
                s2 += ri * ri;
            }, t1, t2);
            seg_part_[blk] = t1;
            seg_part_[T.nblk + blk] = t2;
        }
        for (int k = 0; k < T.n_inst; ++k) {
            double t1, t2;
            seg_final(T, k, seg_part_.data(), t1, t2);
            double* c = cg + kCgStride * k;
            c[0] = t1;
            c[1] = t2;
            c[4] = 0.0;
            c[5] = tol2[k] * t2;
            c[6] = t2 > 0.0 ? 0.0 : 1.0;
            c[7] = 0.0;
        }
    }
    // the batched PCG kernels, block table and reduction orders as on the
    // device
    void cgb_iters(int kk, const SegTable& T, int n, const long long* ptr, const int* col,
                   const double* val, const double* diag, double* x, double* r, double* z,
                   double* p, double* q, double* cg) {
        seg_part_.assign(2 * (size_t)T.nblk, 0.0);
        for (int it = 0; it < kk; ++it) {
            for (int i = 0; i < n; ++i) {
                const double* c = cg + kCgStride * T.inst_of[i];
                if (c[6] == 0.0) p[i] = z[i] + c[4] * p[i];
            }
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                if (cg[kCgStride * T.inst_of[i] + 6] != 0.0) continue;
                double sl[kSpmvLanes];
                for (int l = 0; l < kSpmvLanes; ++l) sl[l] = ot_spmv_lane(i, l, ptr, col, val, p);
                for (int m = kSpmvLanes / 2; m >= 1; m >>= 1) {
                    double t[kSpmvLanes];
                    for (int l = 0; l < kSpmvLanes; ++l) t[l] = sl[l] + sl[l ^ m];
                    for (int l = 0; l < kSpmvLanes; ++l) sl[l] = t[l];
                }
                q[i] = sl[0] + diag[i] * p[i];
            }
            for (int blk = 0; blk < T.nblk; ++blk) {
                const int k = T.blk_inst[blk];
                if (cg[kCgStride * k + 6] != 0.0) continue;
                double t1, t2;
                seg_block_sums(T, blk, [&](long long i, double& s1, double&) { s1 += p[i] * q[i]; },
                               t1, t2);
                seg_part_[blk] = t1;
            }
            for (int k = 0; k < T.n_inst; ++k) {
                double* c = cg + kCgStride * k;
                if (c[6] != 0.0) continue;
                double pq, dummy;
                seg_final(T, k, seg_part_.data(), pq, dummy);
                c[2] = pq;
                if (!(pq > 0.0)) {
                    c[3] = 0.0;
                    c[6] = 2.0;
                } else {
                    c[3] = c[0] / pq;
                }
            }
#pragma omp parallel for schedule(dynamic, 4)
            for (int blk = 0; blk < T.nblk; ++blk) {
                const int k = T.blk_inst[blk];
                if (cg[kCgStride * k + 6] != 0.0) continue;
                const double alpha = cg[kCgStride * k + 3];
                double t1, t2;
                seg_block_sums(T, blk, [&](long long i, double& s1, double& s2) {
                    x[i] += alpha * p[i];
                    const double ri = r[i] - alpha * q[i];
                    r[i] = ri;
                    const double zi = diag[i] > 0.0 ? ri / diag[i] : 0.0;
                    z[i] = zi;
                    s1 += ri * zi;
                    s2 += ri * ri;
                }, t1, t2);
                seg_part_[blk] = t1;
                seg_part_[T.nblk + blk] = t2;
            }
            for (int k = 0; k < T.n_inst; ++k) {
                double* c = cg + kCgStride * k;
                if (c[6] != 0.0) continue;
                double rz, rr;
                seg_final(T, k, seg_part_.data(), rz, rr);
                c[1] = rr;
                c[7] += 1.0;
                if (rr <= c[5]) {
                    c[4] = 0.0;
                    c[6] = 1.0;
                } else {
                    c[4] = rz / c[0];
                    c[0] = rz;
                }
            }
        }
    }

    // ---------------------------------------------------------------- MLP
    static constexpr int kRowsFwd = 64, kRowsGrad = 32;
    void mlp_forward(const float* P, const float* X, int n, float* out) {
        for (int r0 = 0; r0 < n; r0 += kRowsFwd)
            mlp_rows<false>(P, X, nullptr, r0, std::min(n, r0 + kRowsFwd), 0.0f, out, nullptr);
    }
    int mlp_grad_waves(int n) { return (n + kRowsGrad - 1) / kRowsGrad; }
    void mlp_grad(const float* P, const float* X, const float* T, int n, float scale, float* part,
                  float* lpart, float* grad) {
        int waves = mlp_grad_waves(n);
#pragma omp parallel for schedule(static)
        for (int w = 0; w < waves; ++w)
            lpart[w] = mlp_rows<true>(P, X, T, w * kRowsGrad, std::min(n, (w + 1) * kRowsGrad), scale,
                                      nullptr, part + (long long)w * kMlpParams);
        for (int i = 0; i < kMlpParams; ++i) {
            float s = 0.0f;
            for (int w = 0; w < waves; ++w) s += part[(long long)w * kMlpParams + i];
            grad[i] = s;
        }
    }
    void mlp_adam(float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                  float c1, float c2, float eps) {
        for (int i = 0; i < kMlpParams; ++i) adam_elem(i, p, g, m, v, lr, b1, b2, c1, c2, eps);
    }
    void ot_features(int n, const double* xyz, const double* nu, const float* d2,
                     const double* box, double inv_cell, double h, float* f) {
        for (int i = 0; i < n; ++i) ot_features_row(i, xyz, nu, d2, box, inv_cell, h, f + (long long)i * kMlpIn);
    }

    // ------------------------------------------ surface restricted cells
    void surface_items(const double* V, const int* F, const double* vw, int nt, double* ov,
                       double* ow) {
        for (int t = 0; t < nt; ++t) surface_items_row(V, F, vw, t, ov, ow);
    }
    void surface_v0(const double* tv, int n, double* q) {
        for (int t = 0; t < n; ++t)
            for (int a = 0; a < 3; ++a) q[3 * (size_t)t + a] = tv[9 * (size_t)t + a];
    }
    void surface_pass(const SurfParams& P) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int t = 0; t < P.n_items; ++t) surface_item(P, t);
    }
    void surface_next(const double* tv, const double* tw, const int* lst, int nsplit, int nredo,
                      double* ov, double* ow) {
        for (int c = 0; c < nsplit; ++c) surface_split_row(tv, tw, lst[c], c, ov, ow);
        for (int c = nsplit; c < nsplit + nredo; ++c) {
            const int t = lst[c];
            const size_t o = 4 * (size_t)nsplit + (size_t)(c - nsplit);
            for (int a = 0; a < 9; ++a) ov[9 * o + a] = tv[9 * (size_t)t + a];
            if (ow)
                for (int a = 0; a < 3; ++a) ow[3 * o + a] = tw[3 * (size_t)t + a];
        }
    }
    void surface_reduce(int n, const unsigned long long* keys, const int* idx, long long nrec,
                        const double* val, const int* flag, double* m, double* mg) {
#pragma omp parallel for schedule(static)
        for (int s = 0; s < n; ++s) surface_reduce_row(s, keys, idx, nrec, val, flag, m, mg);
    }
    // ------------------------------------------------------- exact kNN
    void iota(int* d, int n) {
        for (int i = 0; i < n; ++i) d[i] = i;
    }
    void lower_bound_u64(const uint64_t* keys, int n, const uint64_t* q, int nq, int* out) {
        for (int t = 0; t < nq; ++t) {
            const int lo = int(std::lower_bound(keys, keys + n, q[t]) - keys);
            out[t] = lo < n ? lo : n - 1;
        }
    }
    void knn_seed_radius(const pf4* pf, const bnode* nodes, const InstLayout& IL, const pf4* qf,
                         const int* qpos, int nq, int kk, double e, float* rr) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int t = 0; t < nq; ++t) rr[t] = knn_seed_radius_row(pf, nodes, IL, qf, qpos, t, kk, e);
    }
    void knn_count(const pf4* pf, const bnode* nodes, const InstLayout& IL, const pf4* qf,
                   const int* items, int ni, const float* rr, int* cnt) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < ni; ++i) {
            const pf4 q = qf[items[i]];
            cnt[i] = radius_query_pt(pf, nodes, IL, 0, q.x, q.y, q.z, -1, rr[items[i]], nullptr,
                                     nullptr, nullptr);
        }
    }
    void knn_fill(const pf4* pf, const bnode* nodes, const InstLayout& IL, const pf4* qf,
                  const int* items, int ni, const float* rr, const long long* off, int* lm,
                  float* lf) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < ni; ++i) {
            const pf4 q = qf[items[i]];
            radius_query_pt(pf, nodes, IL, 0, q.x, q.y, q.z, -1, rr[items[i]], nullptr, lm + off[i],
                            lf + off[i]);
        }
    }
    void knn_d64(const pt4* qp, const pt4* pts, const int* orig, const int* items, int ni,
                 const long long* off, const int* lm, int* li, double* ld) {
        for (int i = 0; i < ni; ++i) knn_d64_row(qp, pts, orig, items[i], off[i], off[i + 1], lm, li, ld);
    }
    // stable sort of each segment by key (as a radix sort is)
    template <class K, class V>
    void seg_sort_kv(const K* kin, const V* vin, K* kout, V* vout, long long, int nseg,
                     const long long* off) {
        for (int sg = 0; sg < nseg; ++sg) {
            const long long b = off[sg], e = off[sg + 1];
            std::vector<long long> idx(e - b);
            std::iota(idx.begin(), idx.end(), b);
            std::stable_sort(idx.begin(), idx.end(),
                             [&](long long x, long long y) { return kin[x] < kin[y]; });
            for (long long k = b; k < e; ++k) {
                kout[k] = kin[idx[k - b]];
                vout[k] = vin[idx[k - b]];
            }
        }
    }
    void knn_select(const int* items, int ni, const long long* off, const int* cnt,
                    const double* ld, const int* li, float* rr, double e, int kk, int n,
                    int self_rule, const int* qrow, int* out_idx, double* out_d2, int* redo,
                    unsigned* nredo) {
        for (int i = 0; i < ni; ++i) {
            const int t = items[i];
            if (!knn_select_row(t, off[i], cnt[i], ld, li, rr[t], e, kk, n, self_rule, qrow, out_idx,
                                out_d2)) {
                rr[t] = knn_grow(rr[t]);
                redo[(*nredo)++] = t;
            }
        }
    }

    void knn_to_caller(const int* orig, const int* list, const float* d2, int n, int K, int* out,
                       float* outd) {
        for (int q = 0; q < n; ++q) {
            int i = orig[q];
            for (int k = 0; k < K; ++k) {
                int j = list[(size_t)q * K + k];
                out[(size_t)i * K + k] = j >= 0 ? orig[j] : -1;
                outd[(size_t)i * K + k] = d2[(size_t)q * K + k];
            }
        }
    }
};

typedef Ctx<EmuBackend> EmuCtx;

}  // namespace vgc

using vgc::EmuCtx;

extern "C" {

int vgc_emu_create(void** ctx) {
    if (!ctx) return VGC_ERR_ARG;
    *ctx = new EmuCtx(0);
    return VGC_OK;
}
void vgc_emu_destroy(void* ctx) { delete (EmuCtx*)ctx; }
const char* vgc_emu_last_error(const void* ctx) {
    return ctx ? ((const EmuCtx*)ctx)->err.c_str() : "null context";
}
int vgc_emu_knn(void* ctx, const double* xyz, uint32_t n, uint32_t k, uint32_t flags, int32_t* nbr,
                double* d2) {
    return vgc::api_knn((EmuCtx*)ctx, xyz, n, k, flags, nbr, d2);
}
int vgc_emu_knn_search(void* ctx, const double* xyz, uint32_t n, const double* q, uint32_t nq,
                       uint32_t k, uint32_t flags, int32_t* nbr, double* d2) {
    return vgc::api_knn_search((EmuCtx*)ctx, xyz, n, q, nq, k, flags, nbr, d2);
}
int vgc_emu_cells(void* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
                  uint32_t k0, uint32_t flags, double* vol, double* mg, uint64_t* nbr_ptr,
                  int32_t* nbr, double* farea, uint64_t cap, vgc_stats* st) {
    return vgc::api_cells((EmuCtx*)ctx, xyz, w, &n, 1u, box, k0, flags, vol, mg, nbr_ptr, nbr,
                          farea, cap, st);
}
int vgc_emu_cells_batch(void* ctx, const double* xyz, const double* w, const uint32_t* inst_n,
                        uint32_t n_inst, const double* boxes, uint32_t k0, uint32_t flags,
                        double* vol, double* mg, uint64_t* nbr_ptr, int32_t* nbr, double* farea,
                        uint64_t cap, vgc_stats* st) {
    return vgc::api_cells((EmuCtx*)ctx, xyz, w, inst_n, n_inst, boxes, k0, flags, vol, mg, nbr_ptr,
                          nbr, farea, cap, st);
}
int vgc_emu_ot_solve(void* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                     const double* box, double tol, uint32_t max_newton, uint32_t flags,
                     vgc_ot_stats* stats) {
    return vgc::api_ot_solve((EmuCtx*)ctx, xyz, nu, w, n, box, tol, max_newton, flags, stats);
}
int vgc_emu_ot_solve_levels(void* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                            const double* box, const uint32_t* levels, uint32_t n_levels,
                            double tol, uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats) {
    return vgc::api_ot_solve_levels((EmuCtx*)ctx, xyz, nu, w, n, box, levels, n_levels, tol,
                                    max_newton, flags, stats);
}
// test hook: weights of the power-nearest seeds (xyz, w: m seeds) of nq points
int vgc_emu_power_nearest(void* ctx, const double* xyz, const double* w, uint32_t m,
                          const double* box, const double* q, uint32_t nq, double* wout) {
    EmuCtx* c = (EmuCtx*)ctx;
    vgc::Problem pb;
    pb.n = (int)m;
    pb.n_inst = 1;
    pb.inst_n = {(int)m};
    pb.box.assign(box, box + 6);
    pb.d_xyz = xyz;
    pb.d_w = w;
    vgc::EngineStats es;
    return c->eng.power_nearest_weights(pb, q, (int)nq, wout, es) ? VGC_ERR_INTERNAL : VGC_OK;
}
int vgc_emu_ot_solve_batch(void* ctx, const double* xyz, const double* nu, double* w,
                           const uint32_t* inst_n, uint32_t n_inst, const double* boxes, double tol,
                           uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats) {
    return vgc::api_ot_solve_batch((EmuCtx*)ctx, xyz, nu, w, inst_n, n_inst, boxes, tol, max_newton,
                                   flags, stats);
}
int vgc_emu_ot_features(void* ctx, const double* xyz, const double* nu, uint32_t n,
                        const double* box, uint32_t flags, float* feats) {
    return vgc::api_ot_features((EmuCtx*)ctx, xyz, nu, n, box, flags, feats);
}
int vgc_emu_ot_convex_start(void* ctx, const double* xyz, uint32_t n, const double* pieces,
                            uint32_t m, double s, uint32_t flags, double* w) {
    return vgc::api_ot_convex_start((EmuCtx*)ctx, xyz, n, pieces, m, s, flags, w);
}
int vgc_emu_mlp_forward(void* ctx, const float* params, const float* feats, uint32_t n,
                        uint32_t flags, float* out) {
    return vgc::api_mlp_forward((EmuCtx*)ctx, params, feats, n, flags, out);
}
int vgc_emu_mlp_grad(void* ctx, const float* params, const float* feats, const float* target,
                     uint32_t n, uint32_t flags, float* grad, double* loss) {
    return vgc::api_mlp_grad((EmuCtx*)ctx, params, feats, target, n, flags, grad, loss);
}
int vgc_emu_mlp_adam(void* ctx, float* params, const float* grad, float* m, float* v,
                     uint32_t step, float lr, uint32_t flags) {
    return vgc::api_mlp_adam((EmuCtx*)ctx, params, grad, m, v, step, lr, flags);
}
// test hook: entries per chunk of the list tiers
void vgc_emu_set_list_chunk(void* ctx, long long n) { ((EmuCtx*)ctx)->eng.list_chunk = n; }
int vgc_emu_det4_sign(const double* rows) {
    return vgc::det4_sign_exact(rows, rows + 4, rows + 8, rows + 12);
}
int vgc_emu_lloyd_batch(void* ctx, double* xyz, const uint32_t* inst_n, uint32_t n_inst,
                        const double* boxes, uint32_t iters, uint32_t k0, uint32_t flags,
                        vgc_stats* st) {
    return vgc::api_lloyd_batch((EmuCtx*)ctx, xyz, inst_n, n_inst, boxes, iters, k0, flags, st);
}
int vgc_emu_cvt_func_grad(void* ctx, const double* xyz, uint32_t n, const double* box,
                          uint32_t flags, double* f, double* g, vgc_stats* st) {
    return vgc::api_cvt_func_grad((EmuCtx*)ctx, xyz, n, box, flags, f, g, st);
}
int vgc_emu_cell_mesh(void* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
                  double shrink, uint32_t flags, uint64_t* cell_vptr, uint64_t* cell_fptr,
                  double* vertices, uint64_t* facet_ptr, int32_t* facet_vertices,
                  int32_t* facet_nbr, int32_t* tets, const vgc_mesh_sizes* cap,
                  vgc_mesh_sizes* need, vgc_stats* stats) {
    return vgc::api_cell_mesh((EmuCtx*)ctx, "vgc_cell_mesh", xyz, w, n, box, shrink, flags, cell_vptr,
                              cell_fptr, vertices, facet_ptr, facet_vertices, facet_nbr, tets, cap,
                              need, stats);
}
int vgc_emu_surface_centroids(void* ctx, const double* xyz, uint32_t n, const double* vtx,
                              uint32_t nv, const int32_t* tri, uint32_t nt, const double* vw,
                              uint32_t flags, double* mg, double* m, vgc_surface_stats* stats) {
    return vgc::api_surface_centroids((EmuCtx*)ctx, xyz, n, vtx, nv, tri, nt, vw, flags, mg, m,
                                      stats);
}
int vgc_emu_rdt(void* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
            uint32_t flags, int32_t* tets, uint64_t cap, uint64_t* nb_tets, vgc_stats* stats) {
    return vgc::api_rdt((EmuCtx*)ctx, xyz, w, n, box, flags, tets, cap, nb_tets, stats);
}
int vgc_emu_cvt_newton(void* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
                       uint32_t m, uint32_t flags, vgc_opt_stats* st) {
    return vgc::api_cvt_newton((EmuCtx*)ctx, xyz, n, box, iters, m, flags, st);
}
int vgc_emu_lloyd(void* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
                  uint32_t k0, uint32_t flags, vgc_stats* st) {
    return vgc::api_lloyd((EmuCtx*)ctx, xyz, n, box, iters, k0, flags, st);
}
}

// C-ABI implementation, written once against a backend: the gfx950 library
// (libvgc.so) instantiates it with the HIP backend; the CPU emulation library
// used by the tests instantiates it with the emulation backend.
#pragma once

#include <chrono>
#include <exception>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vgc.h"
#include "lbfgs.h"
#include "mlp.h"
#include "ot_solver.h"
#include "pipeline.h"
#include "surface.h"

namespace vgc {

template <class B>
struct Ctx {
    explicit Ctx(int dev) : be(dev), eng(be), ot(be, eng), device(dev) {}
    ~Ctx() {
        for (auto& s : stage)
            if (s.p) be.dfree(s.p);
    }
    B be;
    Engine<B> eng;
    OtSolver<B> ot;
    int device;
    std::string err;
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    std::vector<Buf> stage;
    void* staging(int id, size_t bytes) {
        if ((int)stage.size() <= id) stage.resize(id + 1);
        Buf& b = stage[id];
        if (b.bytes < bytes) {
            if (b.p) be.dfree(b.p);
            b.p = be.dalloc(bytes);
            b.bytes = bytes;
        }
        return b.p;
    }
};

inline void fill_stats(const EngineStats& e, vgc_stats* s) {
    if (!s) return;
    s->nb_cells = e.nb_cells;
    s->nb_neighbors = e.nb_neighbors;
    s->nb_pass_b = e.nb_pass_b;
    s->nb_pass_c = e.nb_pass_c;
    s->nb_pass_e = e.nb_pass_e;
    s->nb_uncertain = e.nb_uncertain;
    s->nb_failed = e.nb_failed;
    s->nb_empty = e.nb_empty;
    s->nb_disconnected = e.nb_disconnected;
    s->nb_clips = e.nb_clips;
    s->nb_tests = e.nb_tests;
    s->radius_entries = e.radius_entries;
    s->t_build_ms = e.t_build_ms;
    s->t_knn_ms = e.t_knn_ms;
    s->t_clip_ms = e.t_clip_ms;
    s->t_pass_bc_ms = e.t_pass_bc_ms;
    s->t_compact_ms = e.t_compact_ms;
    s->t_total_ms = e.t_total_ms;
    s->t_clip_a_gpu_ms = e.t_clip_a_gpu_ms;
    s->t_knn_gpu_ms = e.t_knn_gpu_ms;
    s->a_cells = e.a_cells;
    s->a_tests = e.a_tests;
    s->a_live = e.a_live;
    s->a_new = e.a_new;
    s->a_tets = e.a_tets;
    s->a_facets = e.a_facets;
    s->nb_fail_capacity = e.nb_fail_capacity;
    s->why_rings = e.why_rings;
    s->why_walk = e.why_walk;
    s->why_narrow = e.why_narrow;
    s->nb_regather = e.nb_regather;
    s->tuned_waves = (uint64_t)e.tuned_waves;
    s->nb_periodic_retries = e.nb_periodic_retries;
    s->nb_pieces = e.nb_pieces;
    s->nb_fp64_redo = e.nb_fp64_redo;
    s->t_redo_gpu_ms = e.t_redo_gpu_ms;
}

inline bool valid_k0(uint32_t k) { return k == 16 || k == 24 || k == 32 || k == 48 || k == 64; }

template <class B>
int api_cells(Ctx<B>* ctx, const double* xyz, const double* w, const uint32_t* inst_n,
              uint32_t n_inst, const double* boxes, uint32_t k0, uint32_t flags, double* vol,
              double* mg, uint64_t* nbr_ptr, int32_t* nbr, double* farea, uint64_t cap,
              vgc_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (n_inst == 0 || !inst_n || !boxes || !xyz) {
            ctx->err = "vgc_cells: null argument";
            return VGC_ERR_ARG;
        }
        if (k0 != 0 && !valid_k0(k0)) {
            ctx->err = "vgc_cells: k0 must be 0, 16, 24, 32, 48 or 64";
            return VGC_ERR_ARG;
        }
        long long n = 0;
        Problem pb;
        pb.n_inst = (int)n_inst;
        pb.inst_n.resize(n_inst);
        pb.box.assign(boxes, boxes + 6 * (size_t)n_inst);
        for (uint32_t i = 0; i < n_inst; ++i) {
            if (inst_n[i] == 0) {
                ctx->err = "vgc_cells: empty instance";
                return VGC_ERR_ARG;
            }
            pb.inst_n[i] = (int)inst_n[i];
            n += inst_n[i];
            const double* b = boxes + 6 * (size_t)i;
            if (!(b[3] > b[0] && b[4] > b[1] && b[5] > b[2])) {
                ctx->err = "vgc_cells: degenerate box";
                return VGC_ERR_ARG;
            }
        }
        if (n > (1LL << 30)) {
            ctx->err = "vgc_cells: too many seeds";
            return VGC_ERR_ARG;
        }
        pb.n = (int)n;
        pb.k0 = (int)k0;
        pb.periodic = (flags & VGC_PERIODIC) != 0;
        if (pb.periodic && n_inst != 1) {
            ctx->err = "vgc_cells: a periodic domain takes one instance";
            return VGC_ERR_ARG;
        }
        pb.want_neighbors = !(flags & VGC_NO_NEIGHBORS) && (nbr || farea);
        pb.count_work = (flags & VGC_COUNT_WORK) != 0;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        B& be = ctx->be;
        if (dev) {
            pb.d_xyz = xyz;
            pb.d_w = w;
        } else {
            double* dx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            be.h2d(dx, xyz, sizeof(double) * 3 * (size_t)n);
            pb.d_xyz = dx;
            if (w) {
                double* dw = (double*)ctx->staging(1, sizeof(double) * (size_t)n);
                be.h2d(dw, w, sizeof(double) * (size_t)n);
                pb.d_w = dw;
            }
        }
        EngineStats es;
        ctx->eng.autotune = (flags & VGC_AUTOTUNE) != 0;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        if (ctx->eng.cells(pb, es)) {
            ctx->err = "vgc_cells: " + ctx->eng.error;
            return VGC_ERR_INTERNAL;
        }
        Engine<B>& E = ctx->eng;
        auto out = [&](void* dst, const void* src, size_t bytes) {
            if (!dst) return;
            if (dev)
                be.d2d(dst, src, bytes);
            else
                be.d2h(dst, src, bytes);
        };
        out(vol, E.d_vol, sizeof(double) * (size_t)n);
        out(mg, E.d_mg, sizeof(double) * 3 * (size_t)n);
        out(nbr_ptr, E.d_ptr, sizeof(long long) * ((size_t)n + 1));
        fill_stats(es, stats);
        if (nbr || farea) {
            if ((uint64_t)E.total_neighbors > cap) {
                ctx->err = "vgc_cells: neighbour capacity too small";
                return VGC_ERR_CAPACITY;
            }
            out(nbr, E.d_nbr, sizeof(int) * (size_t)E.total_neighbors);
            out(farea, E.d_area, sizeof(double) * (size_t)E.total_neighbors);
        }
        be.sync();
        if (es.nb_failed || es.nb_uncertain) {
            ctx->err = "vgc_cells: " + std::to_string(es.nb_failed) + " cells failed, " +
                       std::to_string(es.nb_uncertain) + " uncertain";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

// CVT energy and gradient of the Voronoi cells of xyz in the box:
// f += sum_i integral over cell i of |x - p_i|^2, g_i += 2 (m_i p_i - mg_i).
template <class B>
int api_cvt_func_grad(Ctx<B>* ctx, const double* xyz, uint32_t n, const double* box,
                      uint32_t flags, double* f, double* g, vgc_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !box || !f || !g || n == 0) {
            ctx->err = "vgc_cvt_func_grad: null argument";
            return VGC_ERR_ARG;
        }
        if (!(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) {
            ctx->err = "vgc_cvt_func_grad: degenerate box";
            return VGC_ERR_ARG;
        }
        if (n > (1u << 30)) {
            ctx->err = "vgc_cvt_func_grad: too many seeds";
            return VGC_ERR_ARG;
        }
        Problem pb;
        pb.n_inst = 1;
        pb.inst_n.assign(1, (int)n);
        pb.box.assign(box, box + 6);
        pb.n = (int)n;
        pb.k0 = 0;
        pb.want_neighbors = false;
        pb.want_e2 = true;
        pb.periodic = (flags & VGC_PERIODIC) != 0;
        pb.count_work = (flags & VGC_COUNT_WORK) != 0;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        B& be = ctx->be;
        double* dg = g;
        if (dev) {
            pb.d_xyz = xyz;
        } else {
            double* dx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            be.h2d(dx, xyz, sizeof(double) * 3 * (size_t)n);
            pb.d_xyz = dx;
            dg = (double*)ctx->staging(1, sizeof(double) * 3 * (size_t)n);
            be.h2d(dg, g, sizeof(double) * 3 * (size_t)n);
        }
        EngineStats es;
        ctx->eng.autotune = (flags & VGC_AUTOTUNE) != 0;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        Engine<B>& E = ctx->eng;
        if (E.cells(pb, es)) {
            ctx->err = "vgc_cvt_func_grad: " + E.error;
            return VGC_ERR_INTERNAL;
        }
        double* df = (double*)ctx->staging(2, sizeof(double));
        be.red_sum((int)n, E.d_e2, df);
        double fs = 0.0;
        be.d2h(&fs, df, sizeof(double));
        *f += fs;
        be.cvt_grad((int)n, pb.d_xyz, E.d_vol, E.d_mg, dg);
        if (!dev) be.d2h(g, dg, sizeof(double) * 3 * (size_t)n);
        be.sync();
        fill_stats(es, stats);
        if (es.nb_failed || es.nb_uncertain) {
            ctx->err = "vgc_cvt_func_grad: " + std::to_string(es.nb_failed) + " cells failed, " +
                       std::to_string(es.nb_uncertain) + " uncertain";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

// Bounding box of n FP64 points (device memory copied to the host when
// dev) for the Morton quantisation, and their largest coordinate magnitude.
template <class B>
void host_extent(B& be, const double* xyz, uint32_t n, bool dev, double lo[3], double hi[3],
                 double& S) {
    std::vector<double> hx;
    const double* hp = xyz;
    if (dev) {
        hx.resize(3 * (size_t)n);
        be.d2h(hx.data(), xyz, sizeof(double) * 3 * (size_t)n);
        hp = hx.data();
    }
    for (int c = 0; c < 3; ++c) lo[c] = hi[c] = hp[c];
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], hp[3 * (size_t)i + c]);
            hi[c] = std::max(hi[c], hp[3 * (size_t)i + c]);
        }
    S = 0.0;
    for (int c = 0; c < 3; ++c) S = std::max(S, std::max(std::fabs(lo[c]), std::fabs(hi[c])));
}

// Exact kNN (FP64 squared distances, ties by caller's index) of the n points
// xyz: of every point itself (q == nullptr: k nearest others, with the
// coincident-point rule) or of nq query points q (k nearest points, none
// excluded).
template <class B>
int knn_common(Ctx<B>* ctx, const char* name, const double* xyz, uint32_t n, const double* q,
               uint32_t nq, uint32_t k, uint32_t flags, int32_t* nbr, double* d2) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        const bool self = q == nullptr;
        if (!xyz || !nbr || n < 1 || k < 1 || (!self && nq < 1)) {
            ctx->err = std::string(name) + ": bad argument";
            return VGC_ERR_ARG;
        }
        if ((self && k > n - 1) || (!self && k > n)) {
            ctx->err = std::string(name) + (self ? ": k exceeds n - 1" : ": k exceeds n");
            return VGC_ERR_ARG;
        }
        if (n > (1u << 30) || (uint64_t)(self ? n : nq) * k > (1ull << 34)) {
            ctx->err = std::string(name) + ": problem too large";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        Problem pb;
        pb.n = (int)n;
        pb.n_inst = 1;
        pb.inst_n = {(int)n};
        double lo[3], hi[3], S;
        host_extent(be, xyz, n, dev, lo, hi, S);
        if (!self) {
            double qlo[3], qhi[3], qS;
            host_extent(be, q, nq, dev, qlo, qhi, qS);
            S = std::max(S, qS);
        }
        if (!(S < 1e300)) {
            ctx->err = std::string(name) + ": coordinates must be finite";
            return VGC_ERR_ARG;
        }
        // Morton box of the stored points (queries outside it are clamped)
        pb.box = {lo[0], lo[1], lo[2], hi[0] + 1e-300 + (hi[0] - lo[0]) * 1e-12,
                  hi[1] + 1e-300 + (hi[1] - lo[1]) * 1e-12, hi[2] + 1e-300 + (hi[2] - lo[2]) * 1e-12};
        for (int c = 0; c < 3; ++c)
            if (!(pb.box[3 + c] > pb.box[c])) pb.box[3 + c] = pb.box[c] + 1.0;
        const double* dq = q;
        if (dev) {
            pb.d_xyz = xyz;
        } else {
            double* dx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            be.h2d(dx, xyz, sizeof(double) * 3 * (size_t)n);
            pb.d_xyz = dx;
            if (!self) {
                double* sq = (double*)ctx->staging(5, sizeof(double) * 3 * (size_t)nq);
                be.h2d(sq, q, sizeof(double) * 3 * (size_t)nq);
                dq = sq;
            }
        }
        EngineStats es;
        Engine<B>& E = ctx->eng;
        if (E.build(pb, es)) {
            ctx->err = std::string(name) + ": " + E.error;
            return VGC_ERR_INTERNAL;
        }
        const size_t rows = self ? n : nq;
        int* d_out = dev ? nbr : (int*)ctx->staging(3, sizeof(int) * rows * k);
        double* d_outd = (dev && d2) ? d2 : (double*)ctx->staging(4, sizeof(double) * rows * k);
        if (E.knn_exact(self ? nullptr : dq, (int)rows, (int)(self ? k + 1 : k), self, S, d_out,
                        d_outd, es)) {
            ctx->err = std::string(name) + ": " + E.error;
            return VGC_ERR_INTERNAL;
        }
        if (!dev) {
            be.d2h(nbr, d_out, sizeof(int) * rows * k);
            if (d2) be.d2h(d2, d_outd, sizeof(double) * rows * k);
        }
        be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_knn(Ctx<B>* ctx, const double* xyz, uint32_t n, uint32_t k, uint32_t flags, int32_t* nbr,
            double* d2) {
    return knn_common(ctx, "vgc_knn", xyz, n, nullptr, 0u, k, flags, nbr, d2);
}

template <class B>
int api_knn_search(Ctx<B>* ctx, const double* xyz, uint32_t n, const double* q, uint32_t nq,
                   uint32_t k, uint32_t flags, int32_t* nbr, double* d2) {
    if (!q) {
        if (ctx) ctx->err = "vgc_knn_search: null query points";
        return VGC_ERR_ARG;
    }
    return knn_common(ctx, "vgc_knn_search", xyz, n, q, nq, k, flags, nbr, d2);
}

// Cell meshes (ConvexCell::append_to_mesh / save) and primal tetrahedra
// (RestrictedVoronoiDiagram::compute_RDT, volumetric) of the cells of xyz
// (weights w or NULL) in the box.
template <class B>
int api_cell_mesh(Ctx<B>* ctx, const char* name, const double* xyz, const double* w, uint32_t n,
                  const double* box, double shrink, uint32_t flags, uint64_t* cell_vptr,
                  uint64_t* cell_fptr, double* vertices, uint64_t* facet_ptr,
                  int32_t* facet_vertices, int32_t* facet_nbr, int32_t* tets,
                  const vgc_mesh_sizes* cap, vgc_mesh_sizes* need, vgc_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !box || n == 0 || !(shrink >= 0.0 && shrink <= 1.0)) {
            ctx->err = std::string(name) + ": bad argument";
            return VGC_ERR_ARG;
        }
        if (!(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) {
            ctx->err = std::string(name) + ": degenerate box";
            return VGC_ERR_ARG;
        }
        if (n > (1u << 28)) {
            ctx->err = std::string(name) + ": too many seeds";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        Problem pb;
        pb.n = (int)n;
        pb.n_inst = 1;
        pb.inst_n = {(int)n};
        pb.box.assign(box, box + 6);
        pb.k0 = 0;
        if (dev) {
            pb.d_xyz = xyz;
            pb.d_w = w;
        } else {
            double* dx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            be.h2d(dx, xyz, sizeof(double) * 3 * (size_t)n);
            pb.d_xyz = dx;
            if (w) {
                double* dw = (double*)ctx->staging(1, sizeof(double) * (size_t)n);
                be.h2d(dw, w, sizeof(double) * (size_t)n);
                pb.d_w = dw;
            }
        }
        ctx->eng.autotune = false;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        const bool want_mesh = vertices && facet_ptr && facet_vertices && facet_nbr;
        long long capv[4] = {0, 0, 0, 0};
        if (cap) {
            capv[0] = want_mesh ? (long long)cap->nb_vertices : 0;
            capv[1] = want_mesh ? (long long)cap->nb_facets : 0;
            capv[2] = want_mesh ? (long long)cap->nb_facet_vertices : 0;
            capv[3] = tets ? (long long)cap->nb_tets : 0;
        }
        // the count pass decides; staging for host outputs is sized after it
        MeshOut mo;
        long long* d_coff = (long long*)ctx->staging(20, sizeof(long long) * 4 * ((size_t)n + 1));
        mo.cell_off = d_coff;
        long long tot[4] = {0, 0, 0, 0};
        EngineStats es;
        // first the counts only (capacities 0 unless device outputs are
        // given and can hold any size): the totals size the host staging
        long long zero[4] = {-1, -1, -1, -1};
        int rc = ctx->eng.mesh(pb, shrink, zero, mo, tot, es);
        if (rc < 0) {
            ctx->err = std::string(name) + ": " + ctx->eng.error;
            return VGC_ERR_INTERNAL;
        }
        if (need) {
            need->nb_vertices = (uint64_t)tot[0];
            need->nb_facets = (uint64_t)tot[1];
            need->nb_facet_vertices = (uint64_t)tot[2];
            need->nb_tets = (uint64_t)tot[3];
        }
        auto out = [&](void* dst, const void* src, size_t bytes) {
            if (!dst || !bytes) return;
            if (dev)
                be.d2d(dst, src, bytes);
            else
                be.d2h(dst, src, bytes);
        };
        // per-cell offsets (uint64 CSR: the offsets are non-negative)
        out(cell_vptr, d_coff, sizeof(long long) * ((size_t)n + 1));
        out(cell_fptr, d_coff + (n + 1), sizeof(long long) * ((size_t)n + 1));
        bool fits = true;
        for (int k = 0; k < 4; ++k) {
            const bool wanted = k < 3 ? want_mesh : tets != nullptr;
            if (wanted && tot[k] > capv[k]) fits = false;
        }
        if (!want_mesh && !tets) {
            be.sync();
            return VGC_OK;
        }
        if (!fits) {
            be.sync();
            ctx->err = std::string(name) + ": output capacity too small (sizes in `need`)";
            return VGC_ERR_CAPACITY;
        }
        if (dev) {
            mo.v = want_mesh ? vertices : nullptr;
            mo.fptr = want_mesh ? (long long*)facet_ptr : nullptr;
            mo.fv = want_mesh ? facet_vertices : nullptr;
            mo.fnbr = want_mesh ? facet_nbr : nullptr;
            mo.tet = tets;
        } else {
            auto st = [&](int id, size_t b) { return ctx->staging(id, b ? b : 1); };
            if (want_mesh) {
                mo.v = (double*)st(21, sizeof(double) * 3 * (size_t)tot[0]);
                mo.fptr = (long long*)st(22, sizeof(long long) * ((size_t)tot[1] + 1));
                mo.fv = (int*)st(23, sizeof(int) * (size_t)tot[2]);
                mo.fnbr = (int*)st(24, sizeof(int) * (size_t)tot[1]);
            }
            if (tets) mo.tet = (int*)st(25, sizeof(int) * 4 * (size_t)tot[3]);
        }
        mo.cell_off = nullptr;
        long long tot2[4];
        long long big[4] = {tot[0], tot[1], tot[2], tot[3]};
        rc = ctx->eng.mesh(pb, shrink, big, mo, tot2, es);
        if (rc != 0) {
            ctx->err = std::string(name) + ": " +
                       (rc < 0 ? ctx->eng.error : std::string("sizes changed between passes"));
            return VGC_ERR_INTERNAL;
        }
        if (!dev) {
            if (want_mesh) {
                be.d2h(vertices, mo.v, sizeof(double) * 3 * (size_t)tot[0]);
                be.d2h(facet_ptr, mo.fptr, sizeof(long long) * ((size_t)tot[1] + 1));
                be.d2h(facet_vertices, mo.fv, sizeof(int) * (size_t)tot[2]);
                be.d2h(facet_nbr, mo.fnbr, sizeof(int) * (size_t)tot[1]);
            }
            if (tets) be.d2h(tets, mo.tet, sizeof(int) * 4 * (size_t)tot[3]);
        }
        be.sync();
        fill_stats(es, stats);
        if (es.nb_failed || es.nb_uncertain) {
            ctx->err = std::string(name) + ": flagged cells";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_rdt(Ctx<B>* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
            uint32_t flags, int32_t* tets, uint64_t cap, uint64_t* nb_tets, vgc_stats* stats) {
    vgc_mesh_sizes c{0, 0, 0, cap}, need{0, 0, 0, 0};
    const int rc = api_cell_mesh(ctx, "vgc_rdt", xyz, w, n, box, 0.0, flags, nullptr, nullptr,
                                 nullptr, nullptr, nullptr, nullptr, tets, &c, &need, stats);
    if (nb_tets) *nb_tets = need.nb_tets;
    if (!tets && rc == VGC_OK) return need.nb_tets ? VGC_ERR_CAPACITY : VGC_OK;
    return rc;
}

// Quasi-Newton CVT (CentroidalVoronoiTesselation::Newton_iterations,
// voronoi/CVT.cpp:272-302): L-BFGS with m pairs on the CVT energy of the
// Voronoi cells of xyz in the box, security radius on, xyz updated in place.
template <class B>
int api_cvt_newton(Ctx<B>* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
                   uint32_t m, uint32_t flags, vgc_opt_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !box || n == 0 || m == 0 || m > 64) {
            ctx->err = "vgc_cvt_newton: bad argument (1 <= m <= 64)";
            return VGC_ERR_ARG;
        }
        if (!(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) {
            ctx->err = "vgc_cvt_newton: degenerate box";
            return VGC_ERR_ARG;
        }
        if (n > (1u << 28)) {
            ctx->err = "vgc_cvt_newton: too many seeds";
            return VGC_ERR_ARG;
        }
        auto t0 = std::chrono::steady_clock::now();
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        double* dx = xyz;
        if (!dev) {
            dx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            be.h2d(dx, xyz, sizeof(double) * 3 * (size_t)n);
        }
        Engine<B>& E = ctx->eng;
        E.autotune = (flags & VGC_AUTOTUNE) != 0;
        E.waves4 = (flags & VGC_WAVES4) != 0;
        E.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        E.begin_call();
        double* d_f = (double*)ctx->staging(2, sizeof(double));
        uint64_t failed = 0;
        // energy and gradient: the caller's funcgrad of CVT.cpp:324-339
        auto fg = [&](const double* x, double* g) -> double {
            Problem pb;
            pb.n_inst = 1;
            pb.inst_n.assign(1, (int)n);
            pb.box.assign(box, box + 6);
            pb.n = (int)n;
            pb.k0 = 0;
            pb.want_neighbors = false;
            pb.want_e2 = true;
            pb.periodic = (flags & VGC_PERIODIC) != 0;
            pb.d_xyz = x;
            EngineStats es;
            if (E.cells(pb, es)) throw std::runtime_error("vgc_cvt_newton: " + E.error);
            failed += es.nb_failed + es.nb_uncertain;
            be.red_sum((int)n, E.d_e2, d_f);
            double f = 0.0;
            be.d2h(&f, d_f, sizeof(double));
            be.dzero(g, sizeof(double) * 3 * (size_t)n);
            be.cvt_grad((int)n, x, E.d_vol, E.d_mg, g);
            return f;
        };
        LbfgsStats ls;
        lbfgs_minimize(be, dx, 3ll * n, (int)m, (int)iters, fg, ls);
        if (!dev) be.d2h(xyz, dx, sizeof(double) * 3 * (size_t)n);
        be.sync();
        if (stats) {
            stats->iterations = (uint32_t)ls.iterations;
            stats->evaluations = (uint32_t)ls.evaluations;
            stats->line_search_info = ls.line_search_info;
            stats->failed_cells = (uint32_t)std::min<uint64_t>(failed, 0xFFFFFFFFu);
            stats->f0 = ls.f0;
            stats->f = ls.f;
            stats->gnorm0 = ls.gnorm0;
            stats->gnorm = ls.gnorm;
            stats->t_total_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
        }
        if (failed) {
            ctx->err = "vgc_cvt_newton: flagged cells";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_lloyd_batch(Ctx<B>* ctx, double* xyz, const uint32_t* inst_n, uint32_t n_inst,
                    const double* boxes, uint32_t iters, uint32_t k0, uint32_t flags,
                    vgc_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !boxes || !inst_n || n_inst < 1) {
            ctx->err = "vgc_lloyd: bad argument";
            return VGC_ERR_ARG;
        }
        if (k0 != 0 && !valid_k0(k0)) {
            ctx->err = "vgc_lloyd: k0 must be 0, 16, 24, 32, 48 or 64";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        Problem pb;
        pb.n_inst = (int)n_inst;
        pb.inst_n.resize(n_inst);
        long long n = 0;
        for (uint32_t i = 0; i < n_inst; ++i) {
            if (inst_n[i] == 0) {
                ctx->err = "vgc_lloyd: empty instance";
                return VGC_ERR_ARG;
            }
            pb.inst_n[i] = (int)inst_n[i];
            n += inst_n[i];
        }
        if (n > 0x7FFFFFFFll) {
            ctx->err = "vgc_lloyd: more than 2^31 - 1 seeds";
            return VGC_ERR_ARG;
        }
        pb.n = (int)n;
        pb.box.assign(boxes, boxes + 6 * (size_t)n_inst);
        pb.k0 = (int)k0;
        pb.periodic = (flags & VGC_PERIODIC) != 0;
        if (pb.periodic && n_inst != 1) {
            ctx->err = "vgc_lloyd: a periodic domain takes one instance";
            return VGC_ERR_ARG;
        }
        double* dx = xyz;
        if (!dev) {
            dx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            be.h2d(dx, xyz, sizeof(double) * 3 * (size_t)n);
        }
        EngineStats es;
        ctx->eng.autotune = (flags & VGC_AUTOTUNE) != 0;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        if (ctx->eng.lloyd(pb, dx, (int)iters, es)) {
            ctx->err = "vgc_lloyd: " + ctx->eng.error;
            return VGC_ERR_INTERNAL;
        }
        if (!dev) be.d2h(xyz, dx, sizeof(double) * 3 * (size_t)n);
        be.sync();
        fill_stats(es, stats);
        if (es.nb_failed || es.nb_uncertain) {
            ctx->err = "vgc_lloyd: flagged cells";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_lloyd(Ctx<B>* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
              uint32_t k0, uint32_t flags, vgc_stats* stats) {
    return api_lloyd_batch(ctx, xyz, &n, 1u, box, iters, k0, flags, stats);
}

template <class B>
int api_ot_solve(Ctx<B>* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                 const double* box, double tol, uint32_t max_newton, uint32_t flags,
                 vgc_ot_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !nu || !w || !box || n < 1 || !(tol > 0.0)) {
            ctx->err = "vgc_ot_solve: bad argument";
            return VGC_ERR_ARG;
        }
        if (!(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) {
            ctx->err = "vgc_ot_solve: degenerate box";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        Problem pb;
        pb.n = (int)n;
        pb.n_inst = 1;
        pb.inst_n = {(int)n};
        pb.box.assign(box, box + 6);
        pb.k0 = 0;
        pb.periodic = (flags & VGC_PERIODIC) != 0;
        const double *dx = xyz, *dnu = nu;
        double* dw = w;
        if (!dev) {
            double* sx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            double* sn = (double*)ctx->staging(5, sizeof(double) * (size_t)n);
            double* sw = (double*)ctx->staging(6, sizeof(double) * (size_t)n);
            be.h2d(sx, xyz, sizeof(double) * 3 * (size_t)n);
            be.h2d(sn, nu, sizeof(double) * (size_t)n);
            be.h2d(sw, w, sizeof(double) * (size_t)n);
            dx = sx;
            dnu = sn;
            dw = sw;
        }
        // the target masses must add up to the domain volume
        std::vector<double> hnu(n);
        be.d2h(hnu.data(), dnu, sizeof(double) * (size_t)n);
        double tot = 0.0, mn = 1e300;
        for (double v : hnu) {
            tot += v;
            mn = std::min(mn, v);
        }
        const double V = (box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2]);
        if (!(mn > 0.0) || std::fabs(tot - V) > 1e-9 * V) {
            ctx->err = "vgc_ot_solve: target masses must be positive and sum to the box volume";
            return VGC_ERR_ARG;
        }
        pb.d_xyz = dx;
        OtStats os;
        const uint32_t maxn = max_newton ? max_newton : 100;
        ctx->eng.autotune = (flags & VGC_AUTOTUNE) != 0;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        if (ctx->ot.solve(pb, dnu, dw, tol, (int)maxn, 1e-7, 4000, os,
                          (flags & VGC_OT_JUDGE_START) != 0, (flags & VGC_OT_VERBOSE) != 0,
                          (flags & VGC_OT_INEXACT) != 0, (flags & VGC_OT_REPAIR) != 0)) {
            ctx->err = "vgc_ot_solve: " + ctx->eng.error;
            return VGC_ERR_INTERNAL;
        }
        if (!dev) be.d2h(w, dw, sizeof(double) * (size_t)n);
        be.sync();
        if (stats) {
            stats->newton_iters = (uint32_t)os.newton_iters;
            stats->cg_iters = (uint32_t)os.cg_iters;
            stats->cell_passes = (uint32_t)os.cell_passes;
            stats->halvings = (uint32_t)os.halvings;
            stats->used_fallback_start = (uint32_t)os.used_fallback_start;
            stats->converged = (uint32_t)os.converged;
            stats->repair_rounds = (uint32_t)os.repair_rounds;
            stats->pad = 0;
            stats->err0 = os.err0;
            stats->final_err = os.final_err;
            stats->t_total_ms = os.t_total_ms;
            stats->t_cells_ms = os.t_cells_ms;
            stats->t_cg_ms = os.t_cg_ms;
        }
        if (!os.converged) {
            ctx->err = "vgc_ot_solve: not converged";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

inline void copy_ot_stats(const OtStats& os, vgc_ot_stats* s) {
    s->newton_iters = (uint32_t)os.newton_iters;
    s->cg_iters = (uint32_t)os.cg_iters;
    s->cell_passes = (uint32_t)os.cell_passes;
    s->halvings = (uint32_t)os.halvings;
    s->used_fallback_start = (uint32_t)os.used_fallback_start;
    s->converged = (uint32_t)os.converged;
    s->repair_rounds = (uint32_t)os.repair_rounds;
    s->pad = 0;
    s->err0 = os.err0;
    s->final_err = os.final_err;
    s->t_total_ms = os.t_total_ms;
    s->t_cells_ms = os.t_cells_ms;
    s->t_cg_ms = os.t_cg_ms;
}

template <class B>
int api_ot_solve_batch(Ctx<B>* ctx, const double* xyz, const double* nu, double* w,
                       const uint32_t* inst_n, uint32_t n_inst, const double* boxes, double tol,
                       uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !nu || !w || !inst_n || !boxes || n_inst < 1 || !(tol > 0.0)) {
            ctx->err = "vgc_ot_solve_batch: bad argument";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        Problem pb;
        pb.n_inst = (int)n_inst;
        pb.inst_n.resize(n_inst);
        long long n = 0;
        for (uint32_t k = 0; k < n_inst; ++k) {
            const double* b = boxes + 6 * (size_t)k;
            if (inst_n[k] == 0 || !(b[3] > b[0] && b[4] > b[1] && b[5] > b[2])) {
                ctx->err = "vgc_ot_solve_batch: empty instance or degenerate box";
                return VGC_ERR_ARG;
            }
            pb.inst_n[k] = (int)inst_n[k];
            n += inst_n[k];
        }
        if (n > 0x7FFFFFFFll) {
            ctx->err = "vgc_ot_solve_batch: more than 2^31 - 1 seeds";
            return VGC_ERR_ARG;
        }
        pb.n = (int)n;
        pb.box.assign(boxes, boxes + 6 * (size_t)n_inst);
        pb.k0 = 0;
        const double *dx = xyz, *dnu = nu;
        double* dw = w;
        if (!dev) {
            double* sx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            double* sn = (double*)ctx->staging(5, sizeof(double) * (size_t)n);
            double* sw = (double*)ctx->staging(6, sizeof(double) * (size_t)n);
            be.h2d(sx, xyz, sizeof(double) * 3 * (size_t)n);
            be.h2d(sn, nu, sizeof(double) * (size_t)n);
            be.h2d(sw, w, sizeof(double) * (size_t)n);
            dx = sx;
            dnu = sn;
            dw = sw;
        }
        // each instance's masses must add up to its box volume
        std::vector<double> hnu(n);
        be.d2h(hnu.data(), dnu, sizeof(double) * (size_t)n);
        for (uint32_t k = 0, o = 0; k < n_inst; o += inst_n[k], ++k) {
            const double* b = boxes + 6 * (size_t)k;
            const double V = (b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
            double tot = 0.0, mn = 1e300;
            for (uint32_t i = o; i < o + inst_n[k]; ++i) {
                tot += hnu[i];
                mn = std::min(mn, hnu[i]);
            }
            if (!(mn > 0.0) || std::fabs(tot - V) > 1e-9 * V) {
                ctx->err = "vgc_ot_solve_batch: target masses of instance " + std::to_string(k) +
                           " must be positive and sum to its box volume";
                return VGC_ERR_ARG;
            }
        }
        pb.d_xyz = dx;
        std::vector<OtStats> os;
        const uint32_t maxn = max_newton ? max_newton : 100;
        ctx->eng.autotune = (flags & VGC_AUTOTUNE) != 0;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        if (ctx->ot.solve_batch(pb, dnu, dw, tol, (int)maxn, 1e-7, 4000, os,
                                (flags & VGC_OT_VERBOSE) != 0, (flags & VGC_OT_INEXACT) != 0)) {
            ctx->err = "vgc_ot_solve_batch: " + ctx->eng.error;
            return VGC_ERR_INTERNAL;
        }
        if (!dev) be.d2h(w, dw, sizeof(double) * (size_t)n);
        be.sync();
        bool all = true;
        for (uint32_t k = 0; k < n_inst; ++k) {
            if (stats) copy_ot_stats(os[k], stats + k);
            all = all && os[k].converged;
        }
        if (!all) {
            ctx->err = "vgc_ot_solve_batch: not every instance converged";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

// ------------------------------------------------------- OT warm start
template <class B>
int api_ot_features(Ctx<B>* ctx, const double* xyz, const double* nu, uint32_t n,
                    const double* box, uint32_t flags, float* feats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !nu || !box || !feats || n < 17) {
            ctx->err = "vgc_ot_features: bad argument (n must exceed 16)";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        const double *dx = xyz, *dnu = nu;
        if (!dev) {
            double* sx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            double* sn = (double*)ctx->staging(5, sizeof(double) * (size_t)n);
            be.h2d(sx, xyz, sizeof(double) * 3 * (size_t)n);
            be.h2d(sn, nu, sizeof(double) * (size_t)n);
            dx = sx;
            dnu = sn;
        }
        Problem pb;
        pb.n = (int)n;
        pb.n_inst = 1;
        pb.inst_n = {(int)n};
        pb.box.assign(box, box + 6);
        pb.d_xyz = dx;
        EngineStats es;
        Engine<B>& E = ctx->eng;
        if (E.build(pb, es)) {
            ctx->err = "vgc_ot_features: " + E.error;
            return VGC_ERR_INTERNAL;
        }
        int* d_list = nullptr;
        float* d_d2 = (float*)ctx->staging(2, sizeof(float) * (size_t)n * 16);
        E.knn_lists(16, d_list, d_d2);
        int* d_out = (int*)ctx->staging(3, sizeof(int) * (size_t)n * 16);
        float* d_outd = (float*)ctx->staging(4, sizeof(float) * (size_t)n * 16);
        be.knn_to_caller(E.d_orig, d_list, d_d2, (int)n, 16, d_out, d_outd);
        double* d_box = (double*)ctx->staging(7, sizeof(double) * 6);
        be.h2d(d_box, box, sizeof(double) * 6);
        const double V = (box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2]);
        const double h = std::cbrt(V / n);
        float* d_f = dev ? feats : (float*)ctx->staging(8, sizeof(float) * (size_t)n * kMlpIn);
        be.ot_features((int)n, dx, dnu, d_outd, d_box, n / V, h, d_f);
        if (!dev) be.d2h(feats, d_f, sizeof(float) * (size_t)n * kMlpIn);
        be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_ot_convex_start(Ctx<B>* ctx, const double* xyz, uint32_t n, const double* pieces,
                        uint32_t m, double s, uint32_t flags, double* w) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !pieces || !w || n == 0 || m == 0 || !(s > 0.0) || !std::isfinite(s) ||
            n > (uint32_t)INT32_MAX || m > (uint32_t)INT32_MAX) {
            ctx->err = "vgc_ot_convex_start: bad argument";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        const double *dx = xyz, *dp = pieces;
        double* dw = w;
        if (!dev) {
            double* sx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            double* sp = (double*)ctx->staging(1, sizeof(double) * kStartPiece * (size_t)m);
            dw = (double*)ctx->staging(2, sizeof(double) * (size_t)n);
            be.h2d(sx, xyz, sizeof(double) * 3 * (size_t)n);
            be.h2d(sp, pieces, sizeof(double) * kStartPiece * (size_t)m);
            dx = sx;
            dp = sp;
        }
        be.ot_convex_start((int)n, dx, dp, (int)m, s, dw);
        if (!dev) be.d2h(w, dw, sizeof(double) * (size_t)n);
        be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_mlp_forward(Ctx<B>* ctx, const float* params, const float* feats, uint32_t n,
                    uint32_t flags, float* out) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!params || !feats || !out || n == 0) {
            ctx->err = "vgc_mlp_forward: bad argument";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        const float *dp = params, *df = feats;
        float* dout = out;
        if (!dev) {
            float* sp = (float*)ctx->staging(9, sizeof(float) * kMlpParams);
            float* sf = (float*)ctx->staging(8, sizeof(float) * (size_t)n * kMlpIn);
            dout = (float*)ctx->staging(10, sizeof(float) * (size_t)n);
            be.h2d(sp, params, sizeof(float) * kMlpParams);
            be.h2d(sf, feats, sizeof(float) * (size_t)n * kMlpIn);
            dp = sp;
            df = sf;
        }
        be.mlp_forward(dp, df, (int)n, dout);
        if (!dev) be.d2h(out, dout, sizeof(float) * (size_t)n);
        be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_mlp_grad(Ctx<B>* ctx, const float* params, const float* feats, const float* target,
                 uint32_t n, uint32_t flags, float* grad, double* loss) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!params || !feats || !target || !grad || n == 0) {
            ctx->err = "vgc_mlp_grad: bad argument";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        const float *dp = params, *df = feats, *dt = target;
        float* dg = grad;
        if (!dev) {
            float* sp = (float*)ctx->staging(9, sizeof(float) * kMlpParams);
            float* sf = (float*)ctx->staging(8, sizeof(float) * (size_t)n * kMlpIn);
            float* st = (float*)ctx->staging(10, sizeof(float) * (size_t)n);
            dg = (float*)ctx->staging(11, sizeof(float) * kMlpParams);
            be.h2d(sp, params, sizeof(float) * kMlpParams);
            be.h2d(sf, feats, sizeof(float) * (size_t)n * kMlpIn);
            be.h2d(st, target, sizeof(float) * (size_t)n);
            dp = sp;
            df = sf;
            dt = st;
        }
        const int waves = be.mlp_grad_waves((int)n);
        float* part = (float*)ctx->staging(12, sizeof(float) * (size_t)waves * kMlpParams);
        float* lpart = (float*)ctx->staging(13, sizeof(float) * (size_t)waves);
        be.mlp_grad(dp, df, dt, (int)n, 1.0f / float(n), part, lpart, dg);
        if (!dev) be.d2h(grad, dg, sizeof(float) * kMlpParams);
        if (loss) {
            std::vector<float> lp(waves);
            be.d2h(lp.data(), lpart, sizeof(float) * waves);
            double l = 0.0;
            for (float v : lp) l += v;
            *loss = l / n;
        }
        be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

template <class B>
int api_mlp_adam(Ctx<B>* ctx, float* params, const float* grad, float* m, float* v, uint32_t step,
                 float lr, uint32_t flags) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!params || !grad || !m || !v || step == 0) {
            ctx->err = "vgc_mlp_adam: bad argument";
            return VGC_ERR_ARG;
        }
        if (!(flags & VGC_DEVICE_PTRS)) {
            ctx->err = "vgc_mlp_adam: device pointers required";
            return VGC_ERR_ARG;
        }
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
        ctx->be.mlp_adam(params, grad, m, v, lr, b1, b2, c1, c2, 1e-8f);
        ctx->be.sync();
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

// ----------------------------------------------------- synthetic inputs
inline int random_float64(uint64_t seed, double* out, uint64_t count) {
    if (!out && count) return VGC_ERR_ARG;
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (uint64_t i = 0; i < count; ++i) out[i] = U(eng);
    return VGC_OK;
}

inline int sample_points(uint32_t kind, uint64_t seed, uint32_t n, double* xyz) {
    if (!xyz || kind > 2) return VGC_ERR_ARG;
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    uint32_t nu = kind == 0 ? n : (kind == 1 ? n / 2 : 0);
    for (uint32_t i = 0; i < nu; ++i) {
        xyz[3 * (size_t)i + 0] = U(eng);
        xyz[3 * (size_t)i + 1] = U(eng);
        xyz[3 * (size_t)i + 2] = U(eng);
    }
    if (nu == n) return VGC_OK;
    const int nc = kind == 1 ? 8 : 16;
    const double smin = kind == 1 ? 0.01 : 0.02, smax = kind == 1 ? 0.1 : 0.15;
    std::vector<double> mean(3 * nc), sig(nc);
    for (int c = 0; c < nc; ++c) {
        for (int a = 0; a < 3; ++a) mean[3 * c + a] = 0.2 + 0.6 * U(eng);
        sig[c] = smin + (smax - smin) * U(eng);
    }
    const double two_pi = 6.283185307179586476925286766559;
    for (uint32_t i = nu; i < n; ++i) {
        int c = (int)(U(eng) * nc);
        if (c >= nc) c = nc - 1;
        while (true) {
            // Box-Muller: two pairs, the fourth normal is discarded
            double u1 = 1.0 - U(eng), u2 = U(eng), u3 = 1.0 - U(eng), u4 = U(eng);
            double r1 = std::sqrt(-2.0 * std::log(u1)), r2 = std::sqrt(-2.0 * std::log(u3));
            double z0 = r1 * std::cos(two_pi * u2), z1 = r1 * std::sin(two_pi * u2);
            double z2 = r2 * std::cos(two_pi * u4);
            double x = mean[3 * c + 0] + sig[c] * z0;
            double y = mean[3 * c + 1] + sig[c] * z1;
            double z = mean[3 * c + 2] + sig[c] * z2;
            if (x > 0.0 && x < 1.0 && y > 0.0 && y < 1.0 && z > 0.0 && z < 1.0) {
                xyz[3 * (size_t)i + 0] = x;
                xyz[3 * (size_t)i + 1] = y;
                xyz[3 * (size_t)i + 2] = z;
                break;
            }
        }
    }
    return VGC_OK;
}

// Multilevel transport (OptimalTransportMap3d::optimize_levels): the seeds
// come in levels, levels[0] = 0 < levels[1] < ... < levels[n_levels] = n;
// level k solves the transport of the first levels[k+1] seeds, their masses
// rescaled to the box volume, starting from the previous level's weights and,
// for the new seeds, the weight of their power-nearest seed of the previous
// level.  stats: one entry per level (may be null).
// The start of a level from the previous level's plan (m0 seeds, weights
// dw[0, m0)): one convex piece per previous seed j, the tangent of the
// potential phi = |q|^2 - w at q_j with slope 2 c_j (c_j the centroid of its
// Laguerre cell) plus mu_j H(|p - b_j|^2), mu_j = (Laguerre volume / Voronoi
// volume)^(1/3), b_j the centroid of its Voronoi cell among the m0 seeds, H
// quadratic up to that cell's radius (capped so that c_j + mu_j (p - b_j)
// stays in the box) and linear beyond; the weights of all m seeds come from
// the soft maximum of the pieces (ot_convex_start_row), s the median of
// 4 x (Voronoi radius) x (Laguerre radius).  A convex function of the
// seeds: no cell of the start is empty.
template <class B>
int convex_level_start(Ctx<B>* ctx, const double* dx, double* dw, uint32_t m0, uint32_t m,
                       const double* box) {
    B& be = ctx->be;
    Engine<B>& E = ctx->eng;
    Problem pc;
    pc.n = (int)m0;
    pc.n_inst = 1;
    pc.inst_n = {(int)m0};
    pc.box.assign(box, box + 6);
    pc.d_xyz = dx;
    pc.want_neighbors = false;
    EngineStats es;
    std::vector<double> vl(m0), ml(3 * (size_t)m0), vv(m0), mv(3 * (size_t)m0), q(3 * (size_t)m0),
        w(m0);
    pc.d_w = dw;
    if (E.cells(pc, es)) return 1;
    be.d2h(vl.data(), E.d_vol, sizeof(double) * m0);
    be.d2h(ml.data(), E.d_mg, sizeof(double) * 3 * (size_t)m0);
    pc.d_w = nullptr;
    if (E.cells(pc, es)) return 1;
    be.d2h(vv.data(), E.d_vol, sizeof(double) * m0);
    be.d2h(mv.data(), E.d_mg, sizeof(double) * 3 * (size_t)m0);
    be.d2h(q.data(), dx, sizeof(double) * 3 * (size_t)m0);
    be.d2h(w.data(), dw, sizeof(double) * m0);
    be.sync();
    const double kPi = 3.14159265358979323846;
    std::vector<double> pieces((size_t)kStartPiece * m0), prod(m0);
    for (uint32_t j = 0; j < m0; ++j) {
        if (!(vl[j] > 0.0) || !(vv[j] > 0.0)) {
            E.error = "convex level start: an empty cell in the previous level's plan";
            return 1;
        }
        prod[j] = 4.0 * std::cbrt(3.0 * vv[j] / (4.0 * kPi)) * std::cbrt(3.0 * vl[j] / (4.0 * kPi));
    }
    std::vector<double> srt(prod);
    std::sort(srt.begin(), srt.end());
    const double s =
        (m0 % 2) ? srt[m0 / 2] : 0.5 * (srt[m0 / 2 - 1] + srt[m0 / 2]);
    for (uint32_t j = 0; j < m0; ++j) {
        double c[3], b[3], dwall = 1e300, r2 = 0.0, gq = 0.0, qq = 0.0;
        for (int a = 0; a < 3; ++a) {
            c[a] = ml[3 * (size_t)j + a] / vl[j];
            b[a] = mv[3 * (size_t)j + a] / vv[j];
            dwall = std::min(dwall, std::min(c[a] - box[a], box[3 + a] - c[a]));
            const double qa = q[3 * (size_t)j + a];
            r2 += (qa - b[a]) * (qa - b[a]);
            gq += 2.0 * c[a] * qa;
            qq += qa * qa;
        }
        const double mu = std::cbrt(vl[j] / vv[j]);
        const double R = std::min(std::cbrt(3.0 * vv[j] / (4.0 * kPi)), 0.98 * dwall / mu);
        const double h = r2 <= R * R ? r2 : R * (2.0 * std::sqrt(r2) - R);
        double* pj = pieces.data() + (size_t)kStartPiece * j;
        pj[0] = qq - w[j] - gq - mu * h;
        for (int a = 0; a < 3; ++a) {
            pj[1 + a] = 2.0 * c[a];
            pj[5 + a] = b[a];
        }
        pj[4] = mu;
        pj[8] = R;
    }
    double* dp = (double*)ctx->staging(8, sizeof(double) * pieces.size());
    be.h2d(dp, pieces.data(), sizeof(double) * pieces.size());
    be.ot_convex_start((int)m, dx, dp, (int)m0, s, dw);
    return 0;
}

// Masses of a coarse level with VGC_OT_CONVEX_LEVELS: each of the first m
// seeds carries the masses of all seeds whose nearest seed among the first m
// it is (exact FP64 nearest, ties by index), summed on the host in index
// order.  The coarse plan then transports the fine plan's own measure, so its
// convex extension starts the finer level close to its masses.
template <class B>
int nearest_masses(Ctx<B>* ctx, const double* dx, const std::vector<double>& hnu, uint32_t n,
                   uint32_t m, std::vector<double>& out) {
    B& be = ctx->be;
    Engine<B>& E = ctx->eng;
    double lo[3], hi[3], S;
    host_extent(be, dx, n, true, lo, hi, S);
    if (!(S < 1e300)) {
        E.error = "coordinates must be finite";
        return 1;
    }
    Problem pb;
    pb.n = (int)m;
    pb.n_inst = 1;
    pb.inst_n = {(int)m};
    pb.box = {lo[0], lo[1], lo[2], hi[0] + 1e-300 + (hi[0] - lo[0]) * 1e-12,
              hi[1] + 1e-300 + (hi[1] - lo[1]) * 1e-12, hi[2] + 1e-300 + (hi[2] - lo[2]) * 1e-12};
    for (int c = 0; c < 3; ++c)
        if (!(pb.box[3 + c] > pb.box[c])) pb.box[3 + c] = pb.box[c] + 1.0;
    pb.d_xyz = dx;
    EngineStats es;
    if (E.build(pb, es)) return 1;
    int* d_idx = (int*)ctx->staging(9, sizeof(int) * (size_t)n);
    double* d_d2 = (double*)ctx->staging(10, sizeof(double) * (size_t)n);
    if (E.knn_exact(dx, (int)n, 1, false, S, d_idx, d_d2, es)) return 1;
    std::vector<int> idx(n);
    be.d2h(idx.data(), d_idx, sizeof(int) * (size_t)n);
    be.sync();
    out.assign(m, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= (int)m) {
            E.error = "nearest coarse seed missing";
            return 1;
        }
        out[idx[i]] += hnu[i];
    }
    return 0;
}

template <class B>
int api_ot_solve_levels(Ctx<B>* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                        const double* box, const uint32_t* levels, uint32_t n_levels, double tol,
                        uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    try {
        if (!xyz || !nu || !w || !box || !levels || n < 1 || n_levels < 1 || !(tol > 0.0)) {
            ctx->err = "vgc_ot_solve_levels: bad argument";
            return VGC_ERR_ARG;
        }
        if (!(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) {
            ctx->err = "vgc_ot_solve_levels: degenerate box";
            return VGC_ERR_ARG;
        }
        if (levels[0] != 0 || levels[n_levels] != n) {
            ctx->err = "vgc_ot_solve_levels: levels must start at 0 and end at n";
            return VGC_ERR_ARG;
        }
        for (uint32_t k = 0; k < n_levels; ++k)
            if (!(levels[k + 1] > levels[k])) {
                ctx->err = "vgc_ot_solve_levels: levels must increase";
                return VGC_ERR_ARG;
            }
        B& be = ctx->be;
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        const double *dx = xyz, *dnu = nu;
        double* dw = w;
        if (!dev) {
            double* sx = (double*)ctx->staging(0, sizeof(double) * 3 * (size_t)n);
            double* sn = (double*)ctx->staging(5, sizeof(double) * (size_t)n);
            double* sw = (double*)ctx->staging(6, sizeof(double) * (size_t)n);
            be.h2d(sx, xyz, sizeof(double) * 3 * (size_t)n);
            be.h2d(sn, nu, sizeof(double) * (size_t)n);
            dx = sx;
            dnu = sn;
            dw = sw;
        }
        std::vector<double> hnu(n);
        be.d2h(hnu.data(), dnu, sizeof(double) * (size_t)n);
        const double V = (box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2]);
        double tot = 0.0, mn = 1e300;
        for (double v : hnu) {
            tot += v;
            mn = std::min(mn, v);
        }
        if (!(mn > 0.0) || std::fabs(tot - V) > 1e-9 * V) {
            ctx->err =
                "vgc_ot_solve_levels: target masses must be positive and sum to the box volume";
            return VGC_ERR_ARG;
        }
        double* dnl = (double*)ctx->staging(7, sizeof(double) * (size_t)n);
        const uint32_t maxn = max_newton ? max_newton : 100;
        ctx->eng.autotune = (flags & VGC_AUTOTUNE) != 0;
        ctx->eng.waves4 = (flags & VGC_WAVES4) != 0;
        ctx->eng.fp64_first = (flags & VGC_FP64_FIRST) != 0;
        ctx->eng.begin_call();
        bool all_conv = true;
        std::vector<double> hl;
        for (uint32_t k = 0; k < n_levels; ++k) {
            const uint32_t m = levels[k + 1], m0 = levels[k];
            // this level's masses: the first m, rescaled to the box volume
            // (VGC_OT_CONVEX_LEVELS: the masses of their nearest seeds)
            if ((flags & VGC_OT_CONVEX_LEVELS) && k + 1 < n_levels) {
                if (nearest_masses(ctx, dx, hnu, n, m, hl)) {
                    ctx->err = "vgc_ot_solve_levels: " + ctx->eng.error;
                    return VGC_ERR_INTERNAL;
                }
            } else {
                double sm = 0.0;
                for (uint32_t i = 0; i < m; ++i) sm += hnu[i];
                hl.resize(m);
                for (uint32_t i = 0; i < m; ++i) hl[i] = hnu[i] * (V / sm);
            }
            be.h2d(dnl, hl.data(), sizeof(double) * (size_t)m);
            // start: zero on the first level; then the previous solution and,
            // for the new seeds, their power-nearest previous seed's weight
            if (k == 0) {
                be.dzero(dw, sizeof(double) * (size_t)m);
            } else if (flags & VGC_OT_CONVEX_LEVELS) {
                if (convex_level_start(ctx, dx, dw, m0, m, box)) {
                    ctx->err = "vgc_ot_solve_levels: " + ctx->eng.error;
                    return VGC_ERR_INTERNAL;
                }
            } else {
                Problem pc;
                pc.n = (int)m0;
                pc.n_inst = 1;
                pc.inst_n = {(int)m0};
                pc.box.assign(box, box + 6);
                pc.d_xyz = dx;
                pc.d_w = dw;
                EngineStats es;
                if (ctx->eng.power_nearest_weights(pc, dx + 3 * (size_t)m0, (int)(m - m0), dw + m0,
                                                   es)) {
                    ctx->err = "vgc_ot_solve_levels: " + ctx->eng.error;
                    return VGC_ERR_INTERNAL;
                }
            }
            Problem pb;
            pb.n = (int)m;
            pb.n_inst = 1;
            pb.inst_n = {(int)m};
            pb.box.assign(box, box + 6);
            pb.k0 = 0;
            pb.d_xyz = dx;
            OtStats os;
            if (ctx->ot.solve(pb, dnl, dw, tol, (int)maxn, 1e-7, 4000, os, false,
                              (flags & VGC_OT_VERBOSE) != 0, (flags & VGC_OT_INEXACT) != 0,
                              (flags & VGC_OT_REPAIR) != 0)) {
                ctx->err = "vgc_ot_solve_levels: " + ctx->eng.error;
                return VGC_ERR_INTERNAL;
            }
            all_conv = all_conv && os.converged;
            if (stats) copy_ot_stats(os, stats + k);
        }
        if (!dev) be.d2h(w, dw, sizeof(double) * (size_t)n);
        be.sync();
        if (!all_conv) {
            ctx->err = "vgc_ot_solve_levels: not converged";
            return VGC_PARTIAL;
        }
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

// Voronoi cells restricted to a triangle mesh: per seed, the mass and first
// moment of its part of the surface (compute_centroids_on_surface,
// voronoi/RVD.cpp:371-415).  Passes over items (the triangles, then the
// items split or redone): nearest seed of each item's first vertex, the
// per-item walk, then the flags decide the next pass.
template <class B>
int api_surface_centroids(Ctx<B>* ctx, const double* xyz, uint32_t n, const double* vtx,
                          uint32_t nv, const int32_t* tri, uint32_t nt, const double* vw,
                          uint32_t flags, double* mg, double* m, vgc_surface_stats* stats) {
    if (!ctx) return VGC_ERR_ARG;
    const char* name = "vgc_surface_centroids";
    struct Owned {
        B& be;
        std::vector<void*> p;
        ~Owned() {
            for (void* q : p)
                if (q) be.dfree(q);
        }
    };
    try {
        if (!xyz || !vtx || !tri || !mg || !m || n < 1 || nv < 3 || nt < 1) {
            ctx->err = std::string(name) + ": bad argument";
            return VGC_ERR_ARG;
        }
        if (n > (1u << 28) || nt > (1u << 26) || nv > (1u << 30)) {
            ctx->err = std::string(name) + ": problem too large";
            return VGC_ERR_ARG;
        }
        B& be = ctx->be;
        auto t0 = std::chrono::steady_clock::now();
        const bool dev = (flags & VGC_DEVICE_PTRS) != 0;
        // triangle indices in range (read back when on the device)
        {
            std::vector<int32_t> ht;
            const int32_t* hp = tri;
            if (dev) {
                ht.resize(3 * (size_t)nt);
                be.d2h(ht.data(), tri, sizeof(int32_t) * 3 * (size_t)nt);
                be.sync();
                hp = ht.data();
            }
            for (size_t i = 0; i < 3 * (size_t)nt; ++i)
                if (hp[i] < 0 || (uint32_t)hp[i] >= nv) {
                    ctx->err = std::string(name) + ": triangle vertex index out of range";
                    return VGC_ERR_ARG;
                }
        }
        Problem pb;
        pb.n = (int)n;
        pb.n_inst = 1;
        pb.inst_n = {(int)n};
        double lo[3], hi[3], S, vlo[3], vhi[3], vS;
        host_extent(be, xyz, n, dev, lo, hi, S);
        host_extent(be, vtx, nv, dev, vlo, vhi, vS);
        S = std::max(S, vS);
        if (!(S < 1e300)) {
            ctx->err = std::string(name) + ": coordinates must be finite";
            return VGC_ERR_ARG;
        }
        pb.box = {lo[0], lo[1], lo[2], hi[0] + 1e-300 + (hi[0] - lo[0]) * 1e-12,
                  hi[1] + 1e-300 + (hi[1] - lo[1]) * 1e-12, hi[2] + 1e-300 + (hi[2] - lo[2]) * 1e-12};
        for (int c = 0; c < 3; ++c)
            if (!(pb.box[3 + c] > pb.box[c])) pb.box[3 + c] = pb.box[c] + 1.0;
        const double* dx = xyz;
        const double* dV = vtx;
        const int32_t* dF = tri;
        const double* dW = vw;
        if (!dev) {
            double* a = (double*)ctx->staging(30, sizeof(double) * 3 * (size_t)n);
            be.h2d(a, xyz, sizeof(double) * 3 * (size_t)n);
            dx = a;
            double* b = (double*)ctx->staging(31, sizeof(double) * 3 * (size_t)nv);
            be.h2d(b, vtx, sizeof(double) * 3 * (size_t)nv);
            dV = b;
            int32_t* c = (int32_t*)ctx->staging(32, sizeof(int32_t) * 3 * (size_t)nt);
            be.h2d(c, tri, sizeof(int32_t) * 3 * (size_t)nt);
            dF = c;
            if (vw) {
                double* d = (double*)ctx->staging(33, sizeof(double) * (size_t)nv);
                be.h2d(d, vw, sizeof(double) * (size_t)nv);
                dW = d;
            }
        }
        pb.d_xyz = dx;
        EngineStats es;
        Engine<B>& E = ctx->eng;
        if (E.build(pb, es)) {
            ctx->err = std::string(name) + ": " + E.error;
            return VGC_ERR_INTERNAL;
        }
        Owned own{be, {}};
        auto alloc = [&](size_t bytes) {
            void* p = be.dalloc(bytes < 256 ? 256 : bytes);
            own.p.push_back(p);
            return p;
        };
        auto release = [&](void* p) {
            for (auto& q : own.p)
                if (q == p) {
                    be.dfree(q);
                    q = nullptr;
                }
        };
        // neighbour lists: K nearest other seeds, doubled while items need more
        int K = (int)std::min<uint32_t>(32u, n - 1);
        int* d_knn = nullptr;
        double* d_knd = nullptr;
        auto build_lists = [&]() -> bool {
            if (K == 0) return true;
            if (d_knn) release(d_knn);
            if (d_knd) release(d_knd);
            d_knn = (int*)alloc(sizeof(int) * (size_t)n * K);
            d_knd = (double*)alloc(sizeof(double) * (size_t)n * K);
            return E.knn_exact(nullptr, (int)n, K + 1, true, S, d_knn, d_knd, es) == 0;
        };
        if (!build_lists()) {
            ctx->err = std::string(name) + ": " + E.error;
            return VGC_ERR_INTERNAL;
        }
        // items of the current pass
        long long n_cur = nt;
        double* tv = (double*)alloc(sizeof(double) * 9 * (size_t)n_cur);
        double* tw = dW ? (double*)alloc(sizeof(double) * 3 * (size_t)n_cur) : nullptr;
        be.surface_items(dV, dF, dW, (int)nt, tv, tw);
        // records and the flags of every item ever processed
        unsigned long long rec_cap = 4ull * nt + 2ull * n + 1024ull;
        unsigned long long* rkey = (unsigned long long*)alloc(sizeof(unsigned long long) * rec_cap);
        double* rval = (double*)alloc(sizeof(double) * 4 * rec_cap);
        unsigned long long* d_count = (unsigned long long*)alloc(sizeof(unsigned long long));
        be.dzero(d_count, sizeof(unsigned long long));
        long long flag_cap = 2ll * nt;
        int* gflag = (int*)alloc(sizeof(int) * (size_t)flag_cap);
        long long base = 0;
        unsigned long long count = 0;
        vgc_surface_stats sst{};
        std::vector<int> hflag;
        for (int pass = 0; n_cur > 0; ++pass) {
            if (pass >= 40 || base + n_cur > (1ll << 31)) {
                ctx->err = std::string(name) + ": items keep splitting (degenerate input?)";
                return VGC_ERR_INTERNAL;
            }
            if (base + n_cur > flag_cap) {
                const long long cap2 = std::max(2 * flag_cap, base + n_cur);
                int* g2 = (int*)alloc(sizeof(int) * (size_t)cap2);
                be.d2d(g2, gflag, sizeof(int) * (size_t)base);
                release(gflag);
                gflag = g2;
                flag_cap = cap2;
            }
            // nearest seed of each item's first vertex
            double* q = (double*)alloc(sizeof(double) * 3 * (size_t)n_cur);
            int* seed0 = (int*)alloc(sizeof(int) * (size_t)n_cur);
            double* sd2 = (double*)alloc(sizeof(double) * (size_t)n_cur);
            be.surface_v0(tv, (int)n_cur, q);
            if (E.knn_exact(q, (int)n_cur, 1, false, S, seed0, sd2, es)) {
                ctx->err = std::string(name) + ": " + E.error;
                return VGC_ERR_INTERNAL;
            }
            release(q);
            release(sd2);
            SurfParams P;
            P.xyz = dx;
            P.n = (int)n;
            P.knn = d_knn;
            P.knn_d2 = d_knd;
            P.K = K;
            P.k_all = K >= (int)n - 1;
            P.tv = tv;
            P.tw = tw;
            P.seed0 = seed0;
            P.n_items = (int)n_cur;
            P.item_base = (uint32_t)base;
            P.rec_count = d_count;
            P.flags = gflag + base;
            for (;;) {
                P.rec_key = rkey;
                P.rec_val = rval;
                P.rec_cap = rec_cap;
                be.surface_pass(P);
                unsigned long long c2 = 0;
                be.d2h(&c2, d_count, sizeof(c2));
                be.sync();
                if (c2 <= rec_cap) {
                    count = c2;
                    break;
                }
                // records overflowed: a larger buffer, the pass again
                const unsigned long long cap2 = c2 + c2 / 2 + 1024;
                unsigned long long* k2 = (unsigned long long*)alloc(sizeof(unsigned long long) * cap2);
                double* v2 = (double*)alloc(sizeof(double) * 4 * cap2);
                be.d2d(k2, rkey, sizeof(unsigned long long) * count);
                be.d2d(v2, rval, sizeof(double) * 4 * count);
                release(rkey);
                release(rval);
                rkey = k2;
                rval = v2;
                rec_cap = cap2;
                be.h2d(d_count, &count, sizeof(count));
            }
            release(seed0);
            hflag.resize((size_t)n_cur);
            be.d2h(hflag.data(), gflag + base, sizeof(int) * (size_t)n_cur);
            be.sync();
            std::vector<int> split, redo;
            for (long long t = 0; t < n_cur; ++t) {
                if (hflag[t] == SURF_OVERFLOW) split.push_back((int)t);
                if (hflag[t] == SURF_NEED_K) redo.push_back((int)t);
            }
            sst.nb_items += (uint64_t)n_cur;
            sst.nb_split += split.size();
            sst.nb_redo += redo.size();
            base += n_cur;
            const long long n_next = 4ll * (long long)split.size() + (long long)redo.size();
            if (n_next == 0) break;
            double* tv2 = (double*)alloc(sizeof(double) * 9 * (size_t)n_next);
            double* tw2 = tw ? (double*)alloc(sizeof(double) * 3 * (size_t)n_next) : nullptr;
            std::vector<int> lst(split);
            lst.insert(lst.end(), redo.begin(), redo.end());
            int* d_lst = (int*)alloc(sizeof(int) * lst.size());
            be.h2d(d_lst, lst.data(), sizeof(int) * lst.size());
            be.surface_next(tv, tw, d_lst, (int)split.size(), (int)redo.size(), tv2, tw2);
            be.sync();
            release(d_lst);
            release(tv);
            if (tw) release(tw);
            tv = tv2;
            tw = tw2;
            n_cur = n_next;
            if (!redo.empty() && K < (int)n - 1) {
                K = (int)std::min<long long>(2ll * K, (long long)n - 1);
                if (!build_lists()) {
                    ctx->err = std::string(name) + ": " + E.error;
                    return VGC_ERR_INTERNAL;
                }
            }
        }
        // records in (seed, item) order, summed per seed
        const int nrec = (int)count;
        uint64_t* k2 = (uint64_t*)alloc(sizeof(uint64_t) * (size_t)std::max(nrec, 1));
        int* i1 = (int*)alloc(sizeof(int) * (size_t)std::max(nrec, 1));
        int* i2 = (int*)alloc(sizeof(int) * (size_t)std::max(nrec, 1));
        if (nrec > 0) {
            be.iota(i1, nrec);
            be.sort_pairs((uint64_t*)rkey, i1, k2, i2, nrec, 64);
        }
        double* om = dev ? m : (double*)ctx->staging(34, sizeof(double) * (size_t)n);
        double* omg = dev ? mg : (double*)ctx->staging(35, sizeof(double) * 3 * (size_t)n);
        be.surface_reduce((int)n, (const unsigned long long*)k2, i2, nrec, rval, gflag, om, omg);
        if (!dev) {
            be.d2h(m, om, sizeof(double) * (size_t)n);
            be.d2h(mg, omg, sizeof(double) * 3 * (size_t)n);
        }
        be.sync();
        sst.nb_records = count;
        sst.k_final = (uint32_t)K;
        sst.t_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                       .count();
        if (stats) *stats = sst;
        return VGC_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VGC_ERR_DEVICE;
    }
}

}  // namespace vgc

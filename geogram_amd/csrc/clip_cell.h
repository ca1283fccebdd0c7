// One Voronoi / Laguerre cell per 64-lane wavefront: clip the domain box by
// bisector (power) planes in nearest-first order until the security radius is
// reached, then integrate volume, first moment and facet areas.
//
// Semantics follow geogram's VBW::ConvexCell (voronoi/convex_cell.cpp) driven
// by the security-radius loop of RestrictedVoronoiDiagram::clip_by_cell_SR
// (voronoi/generic_RVD.h:2278-2347):
//   * the cell is kept in dual form: planes are the dual vertices, corners are
//     the dual triangles (three plane ids, counter-clockwise seen from
//     outside), with the three corners across their edges;
//   * a corner conflicts with a new plane when det4(Pa,Pb,Pc,P) >= 0
//     (convex_cell.cpp:905-911): ties remove the corner;
//   * the hole left by the conflicting corners is re-triangulated by one new
//     corner per border edge (convex_cell.cpp:761-865);
//   * clipping stops at the first neighbour j with |pi-pj|^2 > 4.1 R^2,
//     R = largest corner distance to pi (generic_RVD.h:2303-2323);
//   * a coincident neighbour with a smaller original index leaves the cell
//     unclipped, a coincident one with a larger index is skipped
//     (delaunay/delaunay_nn.cpp:125-133);
//   * volume and first moment are fans of tetrahedra from one corner, facet
//     areas fans of triangles around each face (convex_cell.cpp:1160-1307).
//
// gfx950 layout.  Coordinates are relative to the seed, so a plane is
// e.y <= h with e = pj - pi and h = (|e|^2 + wi - wj) / 2.  Lane l of the
// wave owns corners l, 64+l, ... (S slots): homogeneous point (N, D) with
// D = det(ea, eb, ec) > 0 and N / D the corner, so the conflict
// determinant is the four-FMA dot product t = e.N - h D.  Each corner also
// keeps the magnitudes that bound the rounding error of t: a sign is
// certified when |t| > gamma_9 * (|e|_1 max|N| + |h| |D|); an uncertified one
// follows the tie rule above and flags the cell.  Plane equations, new-corner
// jobs and the two border-walk tables live in the wave's slice of LDS.
//
// Corner state in FP32 (the first pass over every cell).  Plane rows stay
// FP64 in LDS; the corners' N, D and error magnitudes are FP32, built from
// the rows rounded to FP32, and a sign is certified against the exact
// determinant of the FP64 rows (the bound of Prec<float>).  A cell with a sign
// or corner it cannot certify stops at once and is left to the FP64 pass.
// Every other cell took the decisions the FP64 pass takes, in the same order
// and into the same corner slots; its corners are rebuilt in FP64 from their
// plane triples for the integration, exactly as the FP64 pass computed them,
// so its outputs equal the FP64 pass's bit for bit.
#pragma once

#include <type_traits>

#include "wave.h"
#include "types.h"
#include "exact_det.h"

namespace vgc {

enum CellFlag : uint32_t {
    CELL_DONE = 1u,          // outputs written
    CELL_NEED_K = 2u,        // neighbour list exhausted before the security radius
    CELL_OVERFLOW = 4u,      // corner / plane / face capacity exceeded
    CELL_UNCERTAIN = 8u,     // a sign fell inside its error bound (tie rule applied)
    CELL_EMPTY = 16u,        // empty Laguerre cell
    CELL_DISCONNECTED = 32u, // coincident seed with a smaller index
    CELL_BAD_TOPO = 64u,     // conflict zone was not a disk (only after uncertain signs)
    // why a gathering cell went to the list tiers (statistics only)
    CELL_WHY_RINGS = 128u,   // filtered-ring limit
    CELL_WHY_WALK = 256u,    // a tree walk outgrew the wave's buffers
    CELL_WHY_NARROW = 512u,  // a ring could not be narrowed further
};
// a cell that is done and carries none of these needs no further pass
constexpr uint32_t kCensusMask = CELL_NEED_K | CELL_OVERFLOW | CELL_UNCERTAIN | CELL_EMPTY |
                                 CELL_DISCONNECTED | CELL_BAD_TOPO;

// plane slots per cell (8-bit ids): 128 with 64 or 128 corner slots, 256 with
// 256 or 512 corner slots
template <int S>
struct PMax {
    static constexpr int value = S >= 4 ? 256 : 128;
};

// Corner ids in the adjacency words: three fields of 8 bits up to 256
// corners, of 10 bits for 512; the per-plane corner tables widen with them.
template <int S>
struct CornerBits {
    static constexpr int value = S <= 4 ? 8 : 10;
    static constexpr uint32_t mask = (1u << value) - 1u;
    using id_t = typename std::conditional<(S <= 4), uint8_t, uint16_t>::type;
};

constexpr int kCandCap = 128;   // candidates per distance ring
constexpr int kFrontCap = 320;  // tree nodes per level of the walk (two levels: the depth-first stack)
constexpr int kBucketCap = 128; // buckets per ring
// wavefronts per work-group of the first pass (1 or 4; same results): a
// work-group's LDS is released when its last wave ends, so one long cell
// holds its group's slots; 1 measured 13 % faster than 4 on the 2M-seed
// mixture (2: 9 %)
constexpr int kDefaultClipWaves = 1;
constexpr int kMaxFilteredRings = 48;  // power-filtered rings of the first pass
// Voronoi cells: filtered rings of the first pass; cells that need more go to
// a second gather with the corner test (clustered seeds: large cells beside
// dense clusters, whose balls hold thousands of seeds that cannot cut)
constexpr int kMaxVoronoiRings = 16;
constexpr double kGamma9 = 1.2e-15;  // > gamma_9 for the conflict determinant
constexpr double kGamma5 = 0.8e-15;  // > gamma_5 for the corner determinant D

// Error bounds of the corner state of scalar type T.  gamma9: relative bound
// on the conflict determinant t = e.N - h D against the exact determinant of
// the FP64 plane rows; gamma5: corners with D <= gamma5 |D|_abs are
// degenerate (flagged).
template <class T>
struct Prec;
template <>
struct Prec<double> {
    static constexpr double gamma9 = kGamma9;
    static constexpr double gamma5 = kGamma5;
    static constexpr bool fp32 = false;
};
// FP32: each of the four factors of a term of t is an FP64 row entry rounded
// to FP32 (relative 2^-24) and the evaluation adds at most 9 roundings, so a
// term is off by at most gamma_13 < 7.8e-7 of its magnitude; 1.2e-6 also
// covers the rounding of the magnitude sum.  N and D are off by at most
// gamma9 of their magnitudes, so with D > 1e-5 |D|_abs the position N / D is
// known to 12 % of D.  Magnitudes outside [1e-30, 1e30] (FP32 underflow or
// overflow) are uncertain.
template <>
struct Prec<float> {
    static constexpr float gamma9 = 1.2e-6f;
    static constexpr float gamma5 = 1.0e-5f;
    static constexpr bool fp32 = true;
};

struct ClipParams {
    const pt4* pts;        // seeds in Morton order (x, y, z, weight)
    const int* orig;       // Morton position -> caller's index
    const int* inst_of;    // Morton position -> instance (nullptr: one instance)
    const double* box;     // 6 per instance: xmin ymin zmin xmax ymax zmax
    const double* wmax;    // per instance max weight (weighted only)
    const int* inst_n;     // per instance seed count
    const int* items;      // work items (Morton positions); nullptr = 0..n_items-1
    int n_items;
    // neighbour search inside the kernel (gather mode)
    const pf4* pf;          // seeds in Morton order, FP32
    const bnode* nodes;     // implicit trees
    InstLayout IL;
    const int* list;        // neighbour lists (Morton positions), -1 terminated
    const long long* list_ptr;  // per item start (nullptr: cell * list_stride)
    const int* list_len;        // per item length (nullptr: list_stride)
    int list_stride;
    int list_complete;      // lists hold every seed the cell can need
    int ring_limit;         // gather mode: filtered rings before the cell is handed on
    int weighted;
    int corner;             // gather mode, Voronoi cells: the corner-test kernel
    unsigned* abort_flag;   // if set, per instance: raised by an empty cell; the
                            // instance's later cells skip
    // outputs, indexed by the caller's seed index
    double* vol;
    double* mg;             // 3 per seed: integral of x over the cell
    double* e2;             // if set, per seed: integral of |x - p|^2 over the cell
    int* nbr_cnt;
    int* nbr_tmp;           // face_stride per seed: caller's index of neighbour
    double* area_tmp;       // face_stride per seed
    int face_stride;
    // per item (Morton position) bookkeeping
    uint32_t* flags;        // indexed by Morton position
    double* r2;             // squared radius of the (partial) cell, by Morton position
    fball* ball;            // weighted: ball of a flagged partial cell (list tiers)
    // [0] planes that cut, [1] planes tested, [2] live corners at each test,
    // [3] corners created, [4] integration tetrahedra, [5] facets
    unsigned long long* counters;
    // cell meshes and primal tetrahedra (ConvexCell::append_to_mesh,
    // RestrictedVoronoiDiagram::compute_RDT): 0 off, 1 count, 2 write.
    // Per cell (caller's index i), four counts / offsets at stride n (counts)
    // and n + 1 (exclusive offsets): [0] vertices, [1] facets, [2] facet
    // vertex entries, [3] tetrahedra.
    int mesh_mode;
    int mesh_n;
    int* mesh_cnt;
    const long long* mesh_off;
    double mesh_shrink;     // vertices pulled towards the cell's centroid
    double* mesh_v;         // 3 per vertex (absolute coordinates)
    long long* mesh_fptr;   // per facet: first entry in mesh_fv
    int* mesh_fv;           // global vertex indices, facets counter-clockwise
                            // seen from outside
    int* mesh_fnbr;         // per facet: neighbour (caller's index), -1 domain
    int* mesh_tet;          // 4 per tetrahedron (caller's indices)
    // periodic domains: seeds with a caller index >= n_real are images
    // (ghosts) of seeds of the period, candidates only (no cell of their
    // own); src maps every caller index to its seed (null: identity)
    int n_real;
    const int* src;
    // pieces of cells over the plane limit: item k clips its cell inside
    // item_box[6k .. 6k+5] (one octant of the domain around the seed) and,
    // with piece set, writes its outputs and flags by item, not by seed
    const double* item_box;
    int piece;
    // host dispatch: run the first pass (S = 1, gather, FP64 signs) with FP32
    // corner state
    int fp32;
};

template <int S>
struct CellLds {
    static constexpr int kPMax = PMax<S>::value;
    double pe[kPMax][4];   // e.x e.y e.z h
    int pg[kPMax];         // Morton position of the neighbour, -1 for the box
    int v2t[kPMax];        // first corner of each face (integration)
    uint32_t job[64 * S];  // new corner jobs: a | b << 8 | across << 16
    uint32_t cpl[64 * S];  // corner planes (integration walk)
    uint32_t cadj[64 * S]; // corner adjacency (integration walk)
    typename CornerBits<S>::id_t st_tab[kPMax];  // new corner whose first plane is p
    typename CornerBits<S>::id_t en_tab[kPMax];  // new corner whose second plane is p
    union {
        // corner points (integration, after the last ring)
        double q[64 * S][3];
        // gather mode: candidates of the current distance ring, sorted by
        // (FP32 squared distance, Morton position), and the tree walk buffers
        struct {
            uint64_t cand[kCandCap];
            union {
                struct {
                    int front[2][kFrontCap];
                    int bucket[kBucketCap];
                };
                // positions of the next 64 candidates, loaded together
                // before they are clipped one by one
                pt4 stage[kWave];
            };
        };
    };
};

// ---------------------------------------------------------------- helpers

template <int S>
VGC_HD LB conf_bit(const uint64_t (&m)[S], const LI& cid) {
    LU64 mm = LU64(m[0]);
#pragma unroll
    for (int s = 1; s < S; ++s) mm = select(cid >= 64 * s, LU64(m[s]), mm);
    LU64 sh = cvt<uint64_t>(cid & 63);
    return ((mm >> sh) & LU64(1)) != LU64(0);
}

VGC_HD LU byte_of(const LU& w, int k) { return (w >> LU(8 * k)) & LU(0xFF); }

// field k of an adjacency word
template <int S>
VGC_HD LU adj_of(const LU& w, int k) {
    return (w >> LU(CornerBits<S>::value * k)) & LU(CornerBits<S>::mask);
}

template <class T>
struct CornerT {
    Lane<T> nx, ny, nz, d, nmax, dabs, r2;
    LB degenerate;
};
typedef CornerT<double> Corner;

// Corner of planes a, b, c (outward normals e, offsets h), oriented so that
// D = det(ea, eb, ec) > 0.  N = ha (eb x ec) + hb (ec x ea) + hc (ea x eb).
// In FP32, r2 is an upper bound of the squared distance of the exact corner
// of the FP64 rows (the security radius may then only stop later).
template <class T>
VGC_HD CornerT<T> make_corner(const Lane<T>& eax, const Lane<T>& eay, const Lane<T>& eaz,
                              const Lane<T>& ha, const Lane<T>& ebx, const Lane<T>& eby,
                              const Lane<T>& ebz, const Lane<T>& hb, const Lane<T>& ecx,
                              const Lane<T>& ecy, const Lane<T>& ecz, const Lane<T>& hc) {
    typedef Lane<T> LD;  // the lane type of the corner state
    CornerT<T> C;
    // cross products and their magnitude sums
    LD bcx = vfma(eby, ecz, -(ebz * ecy)), bcy = vfma(ebz, ecx, -(ebx * ecz)),
       bcz = vfma(ebx, ecy, -(eby * ecx));
    LD cax = vfma(ecy, eaz, -(ecz * eay)), cay = vfma(ecz, eax, -(ecx * eaz)),
       caz = vfma(ecx, eay, -(ecy * eax));
    LD abx = vfma(eay, ebz, -(eaz * eby)), aby = vfma(eaz, ebx, -(eax * ebz)),
       abz = vfma(eax, eby, -(eay * ebx));
    LD Aeax = vabs(eax), Aeay = vabs(eay), Aeaz = vabs(eaz);
    LD Aebx = vabs(ebx), Aeby = vabs(eby), Aebz = vabs(ebz);
    LD Aecx = vabs(ecx), Aecy = vabs(ecy), Aecz = vabs(ecz);
    LD abcx = vfma(Aeby, Aecz, Aebz * Aecy), abcy = vfma(Aebz, Aecx, Aebx * Aecz),
       abcz = vfma(Aebx, Aecy, Aeby * Aecx);
    LD acax = vfma(Aecy, Aeaz, Aecz * Aeay), acay = vfma(Aecz, Aeax, Aecx * Aeaz),
       acaz = vfma(Aecx, Aeay, Aecy * Aeax);
    LD aabx = vfma(Aeay, Aebz, Aeaz * Aeby), aaby = vfma(Aeaz, Aebx, Aeax * Aebz),
       aabz = vfma(Aeax, Aeby, Aeay * Aebx);
    LD Aha = vabs(ha), Ahb = vabs(hb), Ahc = vabs(hc);
    C.nx = vfma(ha, bcx, vfma(hb, cax, hc * abx));
    C.ny = vfma(ha, bcy, vfma(hb, cay, hc * aby));
    C.nz = vfma(ha, bcz, vfma(hb, caz, hc * abz));
    C.d = vfma(eax, bcx, vfma(eay, bcy, eaz * bcz));
    LD nax = vfma(Aha, abcx, vfma(Ahb, acax, Ahc * aabx));
    LD nay = vfma(Aha, abcy, vfma(Ahb, acay, Ahc * aaby));
    LD naz = vfma(Aha, abcz, vfma(Ahb, acaz, Ahc * aabz));
    C.nmax = vmax(nax, vmax(nay, naz));
    C.dabs = vfma(Aeax, abcx, vfma(Aeay, abcy, Aeaz * abcz));
    C.degenerate = C.d <= T(Prec<T>::gamma5) * C.dabs;
    if constexpr (Prec<T>::fp32)
        C.degenerate = C.degenerate | !(C.dabs >= T(1e-30f)) | !(C.dabs < T(1e30f)) |
                       !(C.nmax < T(1e30f));
    LD dd = select(C.degenerate, LD(T(1)), C.d);
    LD inv = T(1) / dd;
    LD qx = C.nx * inv, qy = C.ny * inv, qz = C.nz * inv;
    C.r2 = vfma(qx, qx, vfma(qy, qy, qz * qz));
    if constexpr (Prec<T>::fp32) {
        // (|N| + a)^2 / (D - gamma9 |D|_abs)^2 with a = 1.8 gamma9 max|N|abs
        // (the length of N's error), plus 1e-5 for the rounding of this
        // evaluation; |N| <= sqrt(3) max_i |N_i| in the cross term
        const LD a = T(1.8f * Prec<T>::gamma9) * C.nmax;
        const LD m = vmax(vabs(C.nx), vmax(vabs(C.ny), vabs(C.nz)));
        const LD num = vfma(C.nx, C.nx, vfma(C.ny, C.ny, C.nz * C.nz)) + a * vfma(T(3.47f), m, a);
        const LD den = select(C.degenerate, LD(T(1)), C.d - T(Prec<T>::gamma9) * C.dabs);
        C.r2 = (num / (den * den)) * T(1.00001f);
    }
    return C;
}

// ------------------------------------------------------------- the cell

template <int S, class T = double>
struct CellState {
    typedef T scalar;
    Lane<T> nx[S], ny[S], nz[S], d[S], nmax[S], dabs[S], r2[S];
    LU pl[S], adj[S];
    uint64_t valid[S];
};

// Plane row p of the wave's table (FP64 in LDS) in the corner state's type.
template <class T, int S>
VGC_HD void load_row(const CellLds<S>& L, const LI& p, Lane<T>& ex, Lane<T>& ey, Lane<T>& ez,
                     Lane<T>& h) {
    const double* pe = &L.pe[0][0];
    ex = cvt<T>(lds_ld(pe, p * 4 + 0));
    ey = cvt<T>(lds_ld(pe, p * 4 + 1));
    ez = cvt<T>(lds_ld(pe, p * 4 + 2));
    h = cvt<T>(lds_ld(pe, p * 4 + 3));
}

// Box planes 0..5 = +x -x +y -y +z -z (outward), corners in lanes 0..7.
template <int S, class T>
VGC_HD void init_box(CellState<S, T>& st, CellLds<S>& L, const double* box, const pt4& pc) {
    const LI lane = lane_id();
    {
        LB m = lane < 6;
        LI axis = lane >> 1;
        LB neg = (lane & 1) != 0;
        LD sg = select(neg, LD(-1.0), LD(1.0));
        LD pcoord = select(axis == 0, LD(pc.x), select(axis == 1, LD(pc.y), LD(pc.z)));
        LD bmin = select(axis == 0, LD(box[0]), select(axis == 1, LD(box[1]), LD(box[2])));
        LD bmax = select(axis == 0, LD(box[3]), select(axis == 1, LD(box[4]), LD(box[5])));
        LD h = select(neg, pcoord - bmin, bmax - pcoord);
        LI li = select(m, lane, LI(0));
        lds_st(&L.pe[0][0], li * 4 + 0, select(axis == 0, sg, LD(0.0)), m);
        lds_st(&L.pe[0][0], li * 4 + 1, select(axis == 1, sg, LD(0.0)), m);
        lds_st(&L.pe[0][0], li * 4 + 2, select(axis == 2, sg, LD(0.0)), m);
        lds_st(&L.pe[0][0], li * 4 + 3, h, m);
        lds_st(L.pg, li, LI(-1), m);
    }
    wave_sync();
    {
        LB m = lane < 8;
        LI sx = lane & 1, sy = (lane >> 1) & 1, sz = (lane >> 2) & 1;
        // max faces are planes 0,2,4 (normal +), min faces 1,3,5 (normal -)
        LI X = select(sx != 0, LI(0), LI(1));
        LI Y = select(sy != 0, LI(2), LI(3));
        LI Z = select(sz != 0, LI(4), LI(5));
        // product of the normal signs is +1 when the number of min faces is even
        LI nmin = (1 - sx) + (1 - sy) + (1 - sz);
        LB pos = (nmin & 1) == 0;
        LI a = select(pos, X, Y), b = select(pos, Y, X), c = Z;
        LI l8 = lane & 7;
        LI a0 = l8 ^ 4;
        LI a1 = select(pos, l8 ^ 1, l8 ^ 2);
        LI a2 = select(pos, l8 ^ 2, l8 ^ 1);
        Lane<T> eax, eay, eaz, ha, ebx, eby, ebz, hb, ecx, ecy, ecz, hc;
        load_row<T>(L, a, eax, eay, eaz, ha);
        load_row<T>(L, b, ebx, eby, ebz, hb);
        load_row<T>(L, c, ecx, ecy, ecz, hc);
        CornerT<T> C = make_corner<T>(eax, eay, eaz, ha, ebx, eby, ebz, hb, ecx, ecy, ecz, hc);
        st.nx[0] = C.nx;
        st.ny[0] = C.ny;
        st.nz[0] = C.nz;
        st.d[0] = C.d;
        st.nmax[0] = C.nmax;
        st.dabs[0] = C.dabs;
        st.r2[0] = select(m, C.r2, Lane<T>(T(0)));
        st.pl[0] = cvt<uint32_t>(a) | (cvt<uint32_t>(b) << LU(8)) | (cvt<uint32_t>(c) << LU(16));
        st.adj[0] = cvt<uint32_t>(a0) | (cvt<uint32_t>(a1) << LU(CornerBits<S>::value)) |
                    (cvt<uint32_t>(a2) << LU(2 * CornerBits<S>::value));
        st.valid[0] = 0xFFull;
#pragma unroll
        for (int s = 1; s < S; ++s) {
            st.valid[s] = 0;
            st.nx[s] = st.ny[s] = st.nz[s] = T(0);
            st.d[s] = T(1);
            st.nmax[s] = st.dabs[s] = st.r2[s] = T(0);
            st.pl[s] = 0;
            st.adj[s] = 0;
        }
    }
}

// Exact sign of det4(Pa, Pb, Pc, Pn) for the lanes in m, with the rows
// (-e, h) of the corner planes pl and of the new plane: >= 0 means conflict
// (voronoi/convex_cell.cpp:905-911).  2 when not decidable.
template <int S>
VGC_HD LI exact_sign_lanes(const CellLds<S>& L, const LU& pl, const LB& m, double ex, double ey,
                           double ez, double h) {
    constexpr int kPMax = PMax<S>::value;
    LI a = cvt<int>(byte_of(pl, 0)) & (kPMax - 1);
    LI b = cvt<int>(byte_of(pl, 1)) & (kPMax - 1);
    LI c = cvt<int>(byte_of(pl, 2)) & (kPMax - 1);
    const double* pe = &L.pe[0][0];
    LD a0 = lds_ld(pe, a * 4 + 0), a1 = lds_ld(pe, a * 4 + 1), a2 = lds_ld(pe, a * 4 + 2),
       a3 = lds_ld(pe, a * 4 + 3);
    LD b0 = lds_ld(pe, b * 4 + 0), b1 = lds_ld(pe, b * 4 + 1), b2 = lds_ld(pe, b * 4 + 2),
       b3 = lds_ld(pe, b * 4 + 3);
    LD c0 = lds_ld(pe, c * 4 + 0), c1 = lds_ld(pe, c * 4 + 1), c2 = lds_ld(pe, c * 4 + 2),
       c3 = lds_ld(pe, c * 4 + 3);
#if defined(VGC_EMU)
    LI r = LI(0);
    for (int l = 0; l < kWave; ++l) {
        if (!m[l]) continue;
        double ra[4] = {-a0[l], -a1[l], -a2[l], a3[l]};
        double rb[4] = {-b0[l], -b1[l], -b2[l], b3[l]};
        double rc[4] = {-c0[l], -c1[l], -c2[l], c3[l]};
        double rn[4] = {-ex, -ey, -ez, h};
        r[l] = det4_sign_exact(ra, rb, rc, rn);
    }
    return r;
#else
    int r = 0;
    if (m) {
        double ra[4] = {-a0, -a1, -a2, a3};
        double rb[4] = {-b0, -b1, -b2, b3};
        double rc[4] = {-c0, -c1, -c2, c3};
        double rn[4] = {-ex, -ey, -ez, h};
        r = det4_sign_exact(ra, rb, rc, rn);
    }
    return r;
#endif
}

// Clip by plane (e, h) stored at slot np with neighbour id j.  Returns the
// updated flags; *cut tells whether the plane removed a corner.
template <int S, bool EXACT, class T>
VGC_HD uint32_t clip_plane(CellState<S, T>& st, CellLds<S>& L, int& np, double ex, double ey,
                           double ez, double h, int j, uint32_t flags, bool& cut,
                           unsigned long long& nlive, unsigned long long& nnew) {
    static_assert(!(EXACT && Prec<T>::fp32), "exact signs run with FP64 corners");
    typedef Lane<T> LT;
    constexpr int kPMax = PMax<S>::value;
    const LI lane = lane_id();
    // the plane in the corner state's type (FP64: itself)
    const T ext = T(ex), eyt = T(ey), ezt = T(ez), ht = T(h);
    const T e1 = vabs(ext) + vabs(eyt) + vabs(ezt);
    const T ha = vabs(ht);
    uint64_t cm[S];
    bool anyc = false, allc = true, anyu = false;
#pragma unroll
    for (int s = 0; s < S; ++s) nlive += (unsigned long long)popc64(st.valid[s]);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LT t = vfma(LT(ext), st.nx[s], vfma(LT(eyt), st.ny[s], vfma(LT(ezt), st.nz[s], -(ht * st.d[s]))));
        LT bnd = T(Prec<T>::gamma9) * vfma(LT(e1), st.nmax[s], ha * st.dabs[s]);
        LB vs = mask_bit(st.valid[s]);
        LB conf = t >= T(0);
        LB unc;
        if constexpr (Prec<T>::fp32)
            unc = vs & !((vabs(t) > bnd) & (bnd >= T(1e-30f)) & (bnd < T(1e30f)));
        else
            unc = vs & (vabs(t) <= bnd);
        if constexpr (EXACT) {
            if (any(unc)) {
                LI sg = exact_sign_lanes<S>(L, st.pl[s], unc, ex, ey, ez, h);
                LB known = unc & (sg != 2);
                conf = select(known, sg >= 0, conf);
                unc = unc & (sg == 2);
            }
        }
        cm[s] = ballot(vs & conf);
        uint64_t um = ballot(unc);
        anyc = anyc || (cm[s] != 0);
        allc = allc && (cm[s] == st.valid[s]);
        anyu = anyu || (um != 0);
    }
    cut = false;
    if (!anyc) {
        if (anyu) flags |= CELL_UNCERTAIN;
        return flags;
    }
    cut = true;
    if (anyu) flags |= CELL_UNCERTAIN;
    if (allc) {
#pragma unroll
        for (int s = 0; s < S; ++s) st.valid[s] = 0;
        return flags | CELL_EMPTY;
    }
    if (np >= kPMax) return flags | CELL_OVERFLOW;
    // new plane
    {
        LB m = lane < 4;
        LD v = select(lane == 0, LD(ex), select(lane == 1, LD(ey), select(lane == 2, LD(ez), LD(h))));
        lds_st(&L.pe[0][0], np * 4 + (lane & 3), v, m);
        lds_st(L.pg, LI(np), LI(j), lane == 0);
    }
    // border edges: conflict corner, neighbour across edge k kept
    uint64_t bk[3][S];
    int total = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LB cs = mask_bit(cm[s]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            LI nb = cvt<int>(adj_of<S>(st.adj[s], k));
            LB nbc = conf_bit<S>(cm, nb);
            bk[k][s] = ballot(cs & !nbc);
            total += popc64(bk[k][s]);
        }
    }
    uint64_t fr[S];
    int nfree = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        fr[s] = ~(st.valid[s] & ~cm[s]);
        nfree += popc64(fr[s]);
    }
    if (total > nfree) return flags | CELL_OVERFLOW;
    nnew += (unsigned long long)total;
    // jobs in (edge k, slot, lane) order
    {
        int base = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                LB m = mask_bit(bk[k][s]);
                LI r = base + mbcnt(bk[k][s]);
                LU pa = byte_of(st.pl[s], k), pb = byte_of(st.pl[s], (k + 1) % 3);
                LU y = adj_of<S>(st.adj[s], k);
                lds_st(L.job, select(m, r, LI(0)), pa | (pb << LU(8)) | (y << LU(16)), m);
                base += popc64(bk[k][s]);
            }
        }
    }
    wave_sync();
    // targets: the first `total` free corner slots compute the new corners
    uint64_t tm[S];
    {
        int fbase = 0;
        const double hn = h;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            LI fr_rank = fbase + mbcnt(fr[s]);
            LB is_t = mask_bit(fr[s]) & (fr_rank < total);
            tm[s] = ballot(is_t);
            fbase += popc64(fr[s]);
            if (tm[s] != 0) {
                LU jb = lds_ld(L.job, select(is_t, fr_rank, LI(0)));
                LI a = cvt<int>(jb & LU(0xFF)) & (kPMax - 1);
                LI b = cvt<int>((jb >> LU(8)) & LU(0xFF)) & (kPMax - 1);
                LU y = (jb >> LU(16)) & LU(CornerBits<S>::mask);
                LT eax, eay, eaz, hav, ebx, eby, ebz, hbv;
                load_row<T>(L, a, eax, eay, eaz, hav);
                load_row<T>(L, b, ebx, eby, ebz, hbv);
                CornerT<T> C = make_corner<T>(eax, eay, eaz, hav, ebx, eby, ebz, hbv, LT(ext),
                                              LT(eyt), LT(ezt), LT(T(hn)));
                if (any(is_t & C.degenerate)) flags |= CELL_UNCERTAIN;
                st.nx[s] = select(is_t, C.nx, st.nx[s]);
                st.ny[s] = select(is_t, C.ny, st.ny[s]);
                st.nz[s] = select(is_t, C.nz, st.nz[s]);
                st.d[s] = select(is_t, C.d, st.d[s]);
                st.nmax[s] = select(is_t, C.nmax, st.nmax[s]);
                st.dabs[s] = select(is_t, C.dabs, st.dabs[s]);
                st.r2[s] = select(is_t, C.r2, st.r2[s]);
                st.pl[s] = select(is_t, cvt<uint32_t>(a) | (cvt<uint32_t>(b) << LU(8)) |
                                            (LU(uint32_t(np)) << LU(16)),
                                  st.pl[s]);
                st.adj[s] = select(is_t, y, st.adj[s]);
                LI cid = lane + 64 * s;
                lds_st(L.st_tab, a, cvt<typename CornerBits<S>::id_t>(cid), is_t);
                lds_st(L.en_tab, b, cvt<typename CornerBits<S>::id_t>(cid), is_t);
            }
        }
    }
    wave_sync();
    // link the new corners around the new face; kept corners re-point to them
    bool bad = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LI cid = lane + 64 * s;
        if (tm[s] != 0) {
            LB is_t = mask_bit(tm[s]);
            LI a = cvt<int>(byte_of(st.pl[s], 0)) & (kPMax - 1);
            LI b = cvt<int>(byte_of(st.pl[s], 1)) & (kPMax - 1);
            LI n1 = cvt<int>(lds_ld(L.st_tab, b)) & (64 * S - 1);
            LI n2 = cvt<int>(lds_ld(L.en_tab, a)) & (64 * S - 1);
            LI own_a = cvt<int>(lds_ld(L.st_tab, a));
            LI own_b = cvt<int>(lds_ld(L.en_tab, b));
            LB wrong = is_t & ((own_a != cid) | (own_b != cid) | !conf_bit<S>(tm, n1) |
                               !conf_bit<S>(tm, n2));
            bad = bad || any(wrong);
            st.adj[s] = select(is_t, (st.adj[s] & LU(CornerBits<S>::mask)) |
                                         (cvt<uint32_t>(n1) << LU(CornerBits<S>::value)) |
                                         (cvt<uint32_t>(n2) << LU(2 * CornerBits<S>::value)),
                               st.adj[s]);
        }
        LB keep = mask_bit(st.valid[s] & ~cm[s]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            LI nb = cvt<int>(adj_of<S>(st.adj[s], k));
            LB hit = keep & conf_bit<S>(cm, nb);
            if (any(hit)) {
                LI v = cvt<int>(byte_of(st.pl[s], (k + 1) % 3)) & (kPMax - 1);
                LI z = cvt<int>(lds_ld(L.st_tab, v)) & (64 * S - 1);
                bad = bad || any(hit & !conf_bit<S>(tm, z));
                LU cleared = st.adj[s] & ~LU(CornerBits<S>::mask << (CornerBits<S>::value * k));
                st.adj[s] = select(hit, cleared | (cvt<uint32_t>(z) << LU(CornerBits<S>::value * k)),
                                   st.adj[s]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) st.valid[s] = (st.valid[s] & ~cm[s]) | tm[s];
    ++np;
    wave_sync();
    if (bad) flags |= CELL_BAD_TOPO;
    return flags;
}

// State of one cell's security-radius loop.
template <int S, class T = double>
struct CellRun {
    CellState<S, T> st;
    int np;
    uint32_t flags;
    bool stopped;
    pt4 pc, prev;
    double wi, cw;
    int c, orig_c, prev_slot;
    const double* box;
    unsigned long long nclip, ntest, nlive, nnew;
};

// One candidate neighbour j (Morton position) in nearest-first order: the
// coincident-seed rule, the security-radius stop, then the clip.  Returns
// true when the cell is finished (or failed) and the candidate loop must end.
template <int S, bool EXACT, bool W, class T>
VGC_HD bool consider(const ClipParams& P, CellLds<S>& L, CellRun<S, T>& R, int j, const pt4& pj) {
    if (j == R.c) return false;
    const double ex = pj.x - R.pc.x, ey = pj.y - R.pc.y, ez = pj.z - R.pc.z;
    const double d2 = ex * ex + ey * ey + ez * ez;
    if (d2 == 0.0) {
        if (P.orig[j] < R.orig_c) {
            wave_sync();
            init_box<S, T>(R.st, L, R.box, R.pc);
            R.np = 6;
            R.flags = CELL_DISCONNECTED;
            R.stopped = true;
            return true;
        }
        return false;
    }
    // security radius (generic_RVD.h:2303-2323)
    bool reach;
    if (!W) {
        bool anyin = false;
#pragma unroll
        for (int s = 0; s < S; ++s)
            anyin = anyin || any(mask_bit(R.st.valid[s]) & (T(4.1) * R.st.r2[s] >= T(d2)));
        reach = !anyin;
    } else {
        const double sd = d2 + R.cw;
        bool anyin = false;
#pragma unroll
        for (int s = 0; s < S; ++s)
            anyin = anyin || any(mask_bit(R.st.valid[s]) & (T(4.1 * d2) * R.st.r2[s] >= T(sd * sd)));
        reach = (sd > 0.0) && !anyin;
    }
    if (reach) {
        R.stopped = true;
        return true;
    }
    // previous candidate: a coincident copy of a seed gives the same plane;
    // under the exact tie rule the copy takes over the face of the first one
    if (pj.x == R.prev.x && pj.y == R.prev.y && pj.z == R.prev.z && pj.w == R.prev.w) {
        if (R.prev_slot >= 0) {
            lds_st(L.pg, LI(R.prev_slot), LI(j), lane_id() == 0);
            wave_sync();
        }
        return false;
    }
    R.prev = pj;
    const double h = W ? 0.5 * (d2 + (R.wi - pj.w)) : 0.5 * d2;
    bool cut = false;
    ++R.ntest;
    R.flags = clip_plane<S, EXACT, T>(R.st, L, R.np, ex, ey, ez, h, j, R.flags, cut, R.nlive, R.nnew);
    R.prev_slot = (cut && !(R.flags & (CELL_EMPTY | CELL_OVERFLOW))) ? R.np - 1 : -1;
    if (cut) ++R.nclip;
    // FP32 corners: the first sign or corner that is not certified ends the
    // cell (the FP64 pass takes it over)
    constexpr uint32_t kEnd =
        CELL_EMPTY | CELL_OVERFLOW | CELL_BAD_TOPO | (Prec<T>::fp32 ? uint32_t(CELL_UNCERTAIN) : 0u);
    if (R.flags & kEnd) {
        R.stopped = (R.flags & CELL_EMPTY) != 0;
        return true;
    }
    return false;
}

// Tree node / seed loads per lane (masked lanes get zeros).
struct BoxLanes {
    LF lx, ly, lz, hx, hy, hz, wm;
};
VGC_HD BoxLanes load_box(const bnode* T, const LI& node, const LB& m) {
    BoxLanes B;
#if defined(VGC_EMU)
    for (int l = 0; l < kWave; ++l) {
        bnode b{};
        if (m[l]) b = T[node[l]];
        B.lx[l] = b.lo[0];
        B.ly[l] = b.lo[1];
        B.lz[l] = b.lo[2];
        B.hx[l] = b.hi[0];
        B.hy[l] = b.hi[1];
        B.hz[l] = b.hi[2];
        B.wm[l] = b.wmax;
    }
#else
    bnode b{};
    if (m) b = T[node];
    B.lx = b.lo[0];
    B.ly = b.lo[1];
    B.lz = b.lo[2];
    B.hx = b.hi[0];
    B.hy = b.hi[1];
    B.hz = b.hi[2];
    B.wm = b.wmax;
#endif
    return B;
}

// Seed (FP32 x, y, z, w) per lane (masked lanes get zeros).
struct SeedLanes {
    LF x, y, z, w;
};
VGC_HD SeedLanes lane_seed(const pf4* pf, const LI& idx, const LB& m) {
    SeedLanes q;
#if defined(VGC_EMU)
    for (int l = 0; l < kWave; ++l) {
        pf4 a{};
        if (m[l]) a = pf[idx[l]];
        q.x[l] = a.x;
        q.y[l] = a.y;
        q.z[l] = a.z;
        q.w[l] = a.w;
    }
#else
    pf4 a{};
    if (m) a = pf[idx];
    q.x = a.x;
    q.y = a.y;
    q.z = a.z;
    q.w = a.w;
#endif
    return q;
}

VGC_HD LF lane_dist2(const pf4* pf, const LI& idx, const LB& m, float px, float py, float pz) {
#if defined(VGC_EMU)
    LF r = LF(0.0f);
    for (int l = 0; l < kWave; ++l)
        if (m[l]) r[l] = dist2f(pf[idx[l]], px, py, pz);
    return r;
#else
    return m ? dist2f(pf[idx], px, py, pz) : 0.0f;
#endif
}

// Seeds q != c of the instance with rlo2 < d2 <= rhi2 (FP32 squared
// distance) into L.cand as (d2 bits << 32 | q), unsorted.  Returns their
// number (only the first kCandCap are stored), or -1 when the tree walk
// overflows its buffers.  Breadth-first walk of the implicit tree, one node
// per lane; boxes entirely inside the inner ball are skipped.  With F.on,
// only seeds (and subtrees) that pass the power-distance filter are taken.
template <int S>
VGC_HD int gather_ring(const ClipParams& P, CellLds<S>& L, int c, int inst, float px, float py,
                       float pz, float rlo2, float rhi2, const PowerFilter& F) {
    const LI lane = lane_id();
    const int p0 = P.IL.p0[inst], n = P.IL.n[inst], nb = P.IL.nb[inst], Lv = P.IL.L[inst];
    const bnode* T = P.nodes + P.IL.t0[inst];
    lds_st(&L.front[0][0], LI(0), LI(1), lane == 0);
    wave_sync();
    int nf = 1, cur = 0, nbk = 0;
    bool overflow = false;
    while (nf > 0) {
        int nn = 0;
        for (int base = 0; base < nf; base += kWave) {
            LI idx = base + lane;
            LB in = idx < nf;
            LI node = lds_ld(&L.front[cur][0], select(in, idx, LI(0)));
            BoxLanes B = load_box(T, node, in);
            LF dx = vmaxf(vmaxf(B.lx - px, px - B.hx), 0.0f);
            LF dy = vmaxf(vmaxf(B.ly - py, py - B.hy), 0.0f);
            LF dz = vmaxf(vmaxf(B.lz - pz, pz - B.hz), 0.0f);
            LF dmin = dx * dx + dy * dy + dz * dz;
            LF fx = vmaxf(px - B.lx, B.hx - px), fy = vmaxf(py - B.ly, B.hy - py),
               fz = vmaxf(pz - B.lz, B.hz - pz);
            LF dmax = fx * fx + fy * fy + fz * fz;
            LB hit = in & (dmin <= rhi2) & (dmax * 1.00001f > rlo2);
            if (F.on) {
                // lower bound of the filter's left side over the node
                LF cx = vmaxf(vmaxf(B.lx - F.cx, F.cx - B.hx), 0.0f);
                LF cy = vmaxf(vmaxf(B.ly - F.cy, F.cy - B.hy), 0.0f);
                LF cz = vmaxf(vmaxf(B.lz - F.cz, F.cz - B.hz), 0.0f);
                LF dc2 = cx * cx + cy * cy + cz * cz;
                hit = hit & pf_node_may_cut(F, dc2, B.wm, vsqrtf(dmin), vsqrtf(dmax), fx, fy, fz);
            }
            LB internal = node < Lv;
            LB push = hit & internal;
            LB leaf = hit & !internal;
            // a node whose children are internal too pushes its four
            // grandchildren (the children's boxes would only prune): half the
            // dependent rounds down the tree
            LB deep = push & (2 * node < Lv);
            uint64_t mp = ballot(push), md = ballot(deep), ml = ballot(leaf);
            LI rp = nn + 2 * mbcnt(mp) + 2 * mbcnt(md);
            LI first = select(deep, 4 * node, 2 * node);
            LB okp = push & (rp + select(deep, LI(3), LI(1)) < kFrontCap);
            lds_st(&L.front[cur ^ 1][0], select(okp, rp, LI(0)), first, okp);
            lds_st(&L.front[cur ^ 1][0], select(okp, rp + 1, LI(0)), first + 1, okp);
            LB okd = okp & deep;
            lds_st(&L.front[cur ^ 1][0], select(okd, rp + 2, LI(0)), first + 2, okd);
            lds_st(&L.front[cur ^ 1][0], select(okd, rp + 3, LI(0)), first + 3, okd);
            LI rl = nbk + mbcnt(ml);
            LB okl = leaf & (rl < kBucketCap);
            lds_st(L.bucket, select(okl, rl, LI(0)), node - Lv, okl);
            nn += 2 * popc64(mp) + 2 * popc64(md);
            nbk += popc64(ml);
        }
        wave_sync();
        if (nn > kFrontCap || nbk > kBucketCap) overflow = true;
        if (overflow) break;
        nf = nn;
        cur ^= 1;
    }
    if (overflow) return -1;
    int cnt = 0;
    for (int base = 0; base < nbk; base += kWave / kBucket) {
        LI bi = base + (lane >> 4);
        LI k = lane & (kBucket - 1);
        LB inb = bi < nbk;
        LI b = lds_ld(L.bucket, select(inb, bi, LI(0)));
        LI m = p0 + b * kBucket + k;
        LB ok = inb & (b < nb) & (m < p0 + n) & (m != c);
        LB take;
        LF d2;
        if (!F.on) {
            d2 = lane_dist2(P.pf, select(ok, m, LI(0)), ok, px, py, pz);
            take = ok & (d2 > rlo2) & (d2 <= rhi2);
        } else {
            SeedLanes q = lane_seed(P.pf, select(ok, m, LI(0)), ok);
            LF ex = q.x - px, ey = q.y - py, ez = q.z - pz;
            d2 = ex * ex + ey * ey + ez * ez;
            LF gx = q.x - F.cx, gy = q.y - F.cy, gz = q.z - F.cz;
            LF dc2 = gx * gx + gy * gy + gz * gz;
            take = ok & (d2 > rlo2) & (d2 <= rhi2) &
                   pf_may_cut(F, dc2, q.w, vsqrtf(d2), vabsf(ex), vabsf(ey), vabsf(ez));
        }
        uint64_t mt = ballot(take);
        LI r = cnt + mbcnt(mt);
        LB st = take & (r < kCandCap);
        LU64 key = (cvt<uint64_t>(fbits(d2)) << LU64(32)) | cvt<uint64_t>(cvt<uint32_t>(m));
        lds_st(L.cand, select(st, r, LI(0)), key, st);
        cnt += popc64(mt);
    }
    wave_sync();
    return cnt;
}

// Bitonic sort of L.cand[0..cnt) (cnt <= kCandCap), padded to 64 or 128.
template <int S>
VGC_HD void sort_cand(CellLds<S>& L, int cnt) {
    const LI lane = lane_id();
    const int P2 = cnt <= 64 ? 64 : 128;
    for (int i0 = 0; i0 < P2; i0 += kWave) {
        LI i = i0 + lane;
        lds_st(L.cand, i, LU64(~0ull), i >= cnt);
    }
    wave_sync();
    for (int k = 2; k <= P2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            LI p = lane;  // pair index, P2 / 2 <= 64 pairs
            LB act = p < P2 / 2;
            LI i = (p / j) * (2 * j) + (p % j);
            LI ip = i + j;
            LI is = select(act, i, LI(0)), ips = select(act, ip, LI(0));
            LU64 a = lds_ld(L.cand, is), b = lds_ld(L.cand, ips);
            LB asc = (i & k) == 0;
            LB sw = act & ((a > b) == asc);
            wave_sync();
            lds_st(L.cand, is, b, sw);
            lds_st(L.cand, ips, a, sw);
            wave_sync();
        }
    }
}

// FP64 lanes to FP32, rounded up.
VGC_HD LF f_up_lanes(const LD& x) {
#if defined(VGC_EMU)
    LF r;
    for (int l = 0; l < kWave; ++l) r[l] = f_up(x[l]);
    return r;
#else
    return f_up(x);
#endif
}

// Corners of the partial cell for the seed test of the depth-first rings
// (S = 1: one corner per lane), relative to the seed in FP32: position and
// r = |v|^2 - wi plus its share of the error margin, rounded up; delta bounds
// the position error of a corner (FP64 rounding) and of a seed's FP32 copy.
struct CornerTab {
    LF x, y, z, r;
    uint64_t live;
    float delta;
};

// Can seed q (FP32, relative to the seed: e, squared length d2, weight wq)
// cut the partial cell?  It can iff some corner v has
// |v - q|^2 - wq < |v|^2 - wi; evaluated in FP32 against every live corner
// with the margin 1e-5 (|v - q|^2 + |v|^2 + |q|^2 + |r| + |wq|)
// + 4 (|v| + |q|) delta + 4 delta^2, which covers the rounding of the copies
// and of this evaluation, so no seed that can cut is dropped.  Lanes outside
// `want` may come back either way.
VGC_HD LB corner_may_cut(const CornerTab& CT, const LF& ex, const LF& ey, const LF& ez,
                         const LF& d2, const LF& wq, const LB& want) {
    const LF qs = wq + 1e-5f * (vabsf(wq) + d2) + (4.0f * CT.delta) * vsqrtf(d2);
    LB hit = LB(false);
    uint64_t m = CT.live;
    int k = 0;
    while (m) {
        const int v = ffs64(m);
        m &= m - 1;
        const float vx = readlane(CT.x, v), vy = readlane(CT.y, v), vz = readlane(CT.z, v),
                    rv = readlane(CT.r, v);
        const LF dx = ex - vx, dy = ey - vy, dz = ez - vz;
        const LF dd = dx * dx + dy * dy + dz * dz;
        hit = hit | (dd * (1.0f - 1e-5f) < qs + rv);
        if ((++k & 7) == 0 && ballot(want & !hit) == 0) break;
    }
    return hit;
}

// The candidates L.cand[0..cnt) that pass the corner test, compacted in
// place (order kept); returns their count.
template <int S>
VGC_HD int corner_filter_cand(const ClipParams& P, CellLds<S>& L, int cnt, float px, float py,
                              float pz, const CornerTab& CT) {
    static_assert(kCandCap == 2 * kWave, "two candidates per lane");
    const LI lane = lane_id();
    const LB in0 = lane < cnt, in1 = lane + kWave < cnt;
    const LU64 k0 = lds_ld(L.cand, select(in0, lane, LI(0)));
    const LU64 k1 = lds_ld(L.cand, select(in1, lane + kWave, LI(0)));
    LB keep[2] = {in0, in1};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!any(keep[h])) continue;
        const LI m = cvt<int>(cvt<uint32_t>(h ? k1 : k0));
        const SeedLanes q = lane_seed(P.pf, select(keep[h], m, LI(0)), keep[h]);
        const LF ex = q.x - px, ey = q.y - py, ez = q.z - pz;
        const LF d2 = ex * ex + ey * ey + ez * ez;
        keep[h] = keep[h] & corner_may_cut(CT, ex, ey, ez, d2, q.w, keep[h]);
    }
    wave_sync();  // every key read before the first write
    const uint64_t b0 = ballot(keep[0]), b1 = ballot(keep[1]);
    lds_st(L.cand, select(keep[0], mbcnt(b0), LI(0)), k0, keep[0]);
    lds_st(L.cand, select(keep[1], popc64(b0) + mbcnt(b1), LI(0)), k1, keep[1]);
    wave_sync();
    return popc64(b0) + popc64(b1);
}

// Seeds q != c of the instance with rlo2 < d2 <= rhi2 (FP32 squared
// distance from the seed) that pass the power filter F (if on) and, once a
// ring of the cell has overflowed (ct_on), the corner test CT, into L.cand as
// (d2 bits << 32 | q), unsorted.  Depth-first walk of the tree: a stack in
// the wave's LDS, up to 64 nodes popped per step (one per lane), passing
// leaves scanned at once, so the stack stays short however many nodes pass.
// When the candidates would overflow, the buffer is sorted and only the
// seeds strictly nearer than its 64th are kept; rhi2 is lowered to just
// below that distance, so the ring (rlo2, rhi2] it returns is complete.
// Returns the count (<= kCandCap), or -1 when the stack overflows or the
// nearest 64 are all at one distance.
template <int S, bool CORNER>
VGC_HD int gather_dfs(const ClipParams& P, CellLds<S>& L, int c, int inst, float px, float py,
                      float pz, float rlo2, float& rhi2, const PowerFilter& F,
                      const CornerTab& CT, bool& ct_on) {
    constexpr int kStack = 2 * kFrontCap;
    const LI lane = lane_id();
    const int p0 = P.IL.p0[inst], n = P.IL.n[inst], nb = P.IL.nb[inst], Lv = P.IL.L[inst];
    const bnode* T = P.nodes + P.IL.t0[inst];
    int* stk = &L.front[0][0];
    lds_st(stk, LI(0), LI(1), lane == 0);
    wave_sync();
    int top = 1, cnt = 0;
    float lim = rhi2;
    while (top > 0) {
        const int k = top < kWave ? top : kWave;
        top -= k;
        LB in = lane < k;
        LI node = lds_ld(stk, select(in, top + lane, LI(0)));
        wave_sync();  // popped before the pushes below reuse the slots
        BoxLanes B = load_box(T, node, in);
        LF dx = vmaxf(vmaxf(B.lx - px, px - B.hx), 0.0f);
        LF dy = vmaxf(vmaxf(B.ly - py, py - B.hy), 0.0f);
        LF dz = vmaxf(vmaxf(B.lz - pz, pz - B.hz), 0.0f);
        LF dmin = dx * dx + dy * dy + dz * dz;
        LF fx = vmaxf(px - B.lx, B.hx - px), fy = vmaxf(py - B.ly, B.hy - py),
           fz = vmaxf(pz - B.lz, B.hz - pz);
        LF dmax = fx * fx + fy * fy + fz * fz;
        LF cx = vmaxf(vmaxf(B.lx - F.cx, F.cx - B.hx), 0.0f);
        LF cy = vmaxf(vmaxf(B.ly - F.cy, F.cy - B.hy), 0.0f);
        LF cz = vmaxf(vmaxf(B.lz - F.cz, F.cz - B.hz), 0.0f);
        LF dc2 = cx * cx + cy * cy + cz * cz;
        LB hit = in & (dmin <= lim) & (dmax * 1.00001f > rlo2);
        if (F.on) hit = hit & pf_node_may_cut(F, dc2, B.wm, vsqrtf(dmin), vsqrtf(dmax), fx, fy, fz);
        LB internal = node < Lv;
        LB push = hit & internal;
        // nodes whose children are internal too push their four grandchildren
        LB deep = push & (2 * node < Lv);
        uint64_t mp = ballot(push), md = ballot(deep);
        // no room for every child: drop the farthest subtrees, and bound the
        // ring below the nearest distance of what was dropped
        while (top + 2 * popc64(mp) + 2 * popc64(md) > kStack) {
            const float far = wave_maxf(select(push, dmin, LF(-1.0f)));
            LB drop = push & (dmin >= far);
            push = push & !drop;
            deep = deep & !drop;
            hit = hit & !drop;
            mp = ballot(push);
            md = ballot(deep);
            const float nl = nextafterf(far, 0.0f);
            if (!(nl > rlo2)) return -1;
            lim = nl < lim ? nl : lim;
        }
        LB leaf = hit & !internal & (dmin <= lim);
        uint64_t ml = ballot(leaf);
        LI rp = top + 2 * mbcnt(mp) + 2 * mbcnt(md);
        LI first = select(deep, 4 * node, 2 * node);
        lds_st(stk, select(push, rp, LI(0)), first, push);
        lds_st(stk, select(push, rp + 1, LI(0)), first + 1, push);
        lds_st(stk, select(deep, rp + 2, LI(0)), first + 2, deep);
        lds_st(stk, select(deep, rp + 3, LI(0)), first + 3, deep);
        top += 2 * popc64(mp) + 2 * popc64(md);
        lds_st(L.bucket, select(leaf, mbcnt(ml), LI(0)), node - Lv, leaf);
        const int nl = popc64(ml);
        wave_sync();
        for (int base = 0; base < nl; base += kWave / kBucket) {
            LI bi = base + (lane >> 4);
            LI kk = lane & (kBucket - 1);
            LB inb = bi < nl;
            LI b = lds_ld(L.bucket, select(inb, bi, LI(0)));
            LI m = p0 + b * kBucket + kk;
            LB ok = inb & (b < nb) & (m < p0 + n) & (m != c);
            SeedLanes q = lane_seed(P.pf, select(ok, m, LI(0)), ok);
            LF ex = q.x - px, ey = q.y - py, ez = q.z - pz;
            LF d2 = ex * ex + ey * ey + ez * ez;
            LF gx = q.x - F.cx, gy = q.y - F.cy, gz = q.z - F.cz;
            LF gc2 = gx * gx + gy * gy + gz * gz;
            LB take = ok & (d2 > rlo2) & (d2 <= lim);
            if (F.on) take = take & pf_may_cut(F, gc2, q.w, vsqrtf(d2), vabsf(ex), vabsf(ey), vabsf(ez));
            if constexpr (CORNER) {
                if (ct_on && CT.live && any(take))
                    take = take & corner_may_cut(CT, ex, ey, ez, d2, q.w, take);
                if (cnt + popc64(ballot(take)) > kCandCap && !ct_on && CT.live) {
                    // the ring overflows: from here on, and in the cell's
                    // later rings, seeds must also pass the corner test (the
                    // buffer and this batch first)
                    ct_on = true;
                    cnt = corner_filter_cand<S>(P, L, cnt, px, py, pz, CT);
                    if (any(take)) take = take & corner_may_cut(CT, ex, ey, ez, d2, q.w, take);
                }
            }
            if (cnt + popc64(ballot(take)) > kCandCap) {
                // keep the seeds strictly nearer than the 64th nearest so far
                sort_cand<S>(L, cnt);
                const float d64 = __builtin_bit_cast(float, uint32_t(L.cand[kWave - 1] >> 32));
                const float nl2 = nextafterf(d64, 0.0f);
                if (!(nl2 > rlo2)) return -1;
                const uint64_t kmax = uint64_t(__builtin_bit_cast(uint32_t, nl2)) << 32 |
                                      0xFFFFFFFFull;
                const int kept = popc64(ballot(lds_ld(L.cand, lane) <= LU64(kmax)));
                if (kept == 0) return -1;
                cnt = kept;
                lim = nl2;
                take = take & (d2 <= lim);
                wave_sync();
            }
            uint64_t mt = ballot(take);
            LI r = cnt + mbcnt(mt);
            LB st = take & (r < kCandCap);
            LU64 key = (cvt<uint64_t>(fbits(d2)) << LU64(32)) | cvt<uint64_t>(cvt<uint32_t>(m));
            lds_st(L.cand, select(st, r, LI(0)), key, st);
            cnt += popc64(mt);
        }
        wave_sync();
    }
    if (cnt > 0 && lim < rhi2) {
        // candidates taken before the bound came down
        sort_cand<S>(L, cnt);
        const uint64_t kmax = uint64_t(__builtin_bit_cast(uint32_t, lim)) << 32 | 0xFFFFFFFFull;
        int kept = 0;
        for (int i0 = 0; i0 < cnt; i0 += kWave) {
            LI i = i0 + lane;
            kept += popc64(ballot((i < cnt) & (lds_ld(L.cand, select(i < cnt, i, LI(0))) <= LU64(kmax))));
        }
        cnt = kept;
    }
    rhi2 = lim;
    return cnt;
}

// Ball around the partial cell of a seed for the power-distance filter
// (types.h; all weights zero for Voronoi cells): centre = middle of the corners' bounding box, radius =
// farthest corner plus the corner's certified position error
// (|dN| <= g nmax, |dD| <= g dabs).  Returns false (no filter) when a corner
// is too close to degenerate for that bound.
template <int S, bool CORNER, class T>
VGC_HD bool make_filter(const CellRun<S, T>& R, PowerFilter& F, CornerTab& CT) {
    const CellState<S, T>& st = R.st;
    // position error of a corner: (G max|N| + |q| G |D|_abs) / (D - G |D|_abs)
    // per coordinate, G = gamma9 of the corner state; in FP32 also a bound of
    // the error's length (sqrt(3) per-coordinate bounds)
    const double G = Prec<T>::fp32 ? 1.8 * double(Prec<T>::gamma9) : kGamma9;
    CT.live = 0;
    LD lox = LD(INFINITY), loy = LD(INFINITY), loz = LD(INFINITY);
    LD hix = LD(-INFINITY), hiy = LD(-INFINITY), hiz = LD(-INFINITY);
    bool bad = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LB vs = mask_bit(st.valid[s]);
        const LD sd = cvt<double>(st.d[s]), sdabs = cvt<double>(st.dabs[s]);
        LD den = sd - G * sdabs;
        bad = bad || any(vs & !(den > 0.0));
        LD inv = 1.0 / select(vs, sd, LD(1.0));
        LD qx = cvt<double>(st.nx[s]) * inv, qy = cvt<double>(st.ny[s]) * inv,
           qz = cvt<double>(st.nz[s]) * inv;
        lox = select(vs, vmin(lox, qx), lox);
        loy = select(vs, vmin(loy, qy), loy);
        loz = select(vs, vmin(loz, qz), loz);
        hix = select(vs, vmax(hix, qx), hix);
        hiy = select(vs, vmax(hiy, qy), hiy);
        hiz = select(vs, vmax(hiz, qz), hiz);
    }
    if (bad) return false;
    const double ox = 0.5 * (-wave_max(-lox) + wave_max(hix));
    const double oy = 0.5 * (-wave_max(-loy) + wave_max(hiy));
    const double oz = 0.5 * (-wave_max(-loz) + wave_max(hiz));
    F.cx = float(R.pc.x + ox);
    F.cy = float(R.pc.y + oy);
    F.cz = float(R.pc.z + oz);
    // the FP32 centre, relative to the seed
    const double fx = double(F.cx) - R.pc.x, fy = double(F.cy) - R.pc.y,
                 fz = double(F.cz) - R.pc.z;
    LD rmax = LD(0.0), hmx = LD(0.0), hmy = LD(0.0), hmz = LD(0.0), emax = LD(0.0);
    LD cvx = LD(0.0), cvy = LD(0.0), cvz = LD(0.0), cvn = LD(0.0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LB vs = mask_bit(st.valid[s]);
        const LD sdabs = cvt<double>(st.dabs[s]);
        LD dd = select(vs, cvt<double>(st.d[s]), LD(1.0));
        LD inv = 1.0 / dd;
        LD qx = cvt<double>(st.nx[s]) * inv, qy = cvt<double>(st.ny[s]) * inv,
           qz = cvt<double>(st.nz[s]) * inv;
        LD gx = qx - fx, gy = qy - fy, gz = qz - fz;
        LD dist = vsqrt(vfma(gx, gx, vfma(gy, gy, gz * gz)));
        LD qn = vsqrt(vfma(qx, qx, vfma(qy, qy, qz * qz)));
        LD err = G * (cvt<double>(st.nmax[s]) + qn * sdabs) / select(vs, dd - G * sdabs, LD(1.0));
        rmax = vmax(rmax, select(vs, dist + err, LD(0.0)));
        hmx = vmax(hmx, select(vs, vabs(gx) + err, LD(0.0)));
        hmy = vmax(hmy, select(vs, vabs(gy) + err, LD(0.0)));
        hmz = vmax(hmz, select(vs, vabs(gz) + err, LD(0.0)));
        if (CORNER && S == 1) {
            emax = vmax(emax, select(vs, err, LD(0.0)));
            cvx = qx;
            cvy = qy;
            cvz = qz;
            cvn = qn;
        }
    }
    const double rc = wave_max(rmax) * (1.0 + 1e-9);
    F.hx = f_up(wave_max(hmx) * (1.0 + 1e-9));
    F.hy = f_up(wave_max(hmy) * (1.0 + 1e-9));
    F.hz = f_up(wave_max(hmz) * (1.0 + 1e-9));
    const double off2 = fx * fx + fy * fy + fz * fz;
    if (!(rc < 1e30) || !(off2 < 1e30)) return false;
    F.rc = f_up(rc);
    F.phi = float(off2 - R.wi);
    F.dcp = f_up(sqrt(off2));
    double cs = 0.0;
    for (int k = 0; k < 6; ++k) cs = fmax(cs, fabs(R.box[k]));
    F.s = float(1e-6 * cs);
    F.on = 1;
    if (CORNER && S == 1) {
        // corner table of the seed test (corner_may_cut)
        const double delta = (wave_max(emax) + double(F.s)) * (1.0 + 1e-9);
        const LD vv = vfma(cvx, cvx, vfma(cvy, cvy, cvz * cvz));
        const LD r = vv - R.wi;
        const LD rv = r + 1e-5 * (vabs(r) + vv) + 4.0 * delta * cvn + 4.0 * delta * delta;
        CT.x = cvt<float>(cvx);
        CT.y = cvt<float>(cvy);
        CT.z = cvt<float>(cvz);
        CT.r = f_up_lanes(rv);
        CT.delta = f_up(delta);
        CT.live = st.valid[0];
    }
    return true;
}

// Candidates base .. base+63 of the sorted ring (those below cnt): one load
// per lane of their FP64 positions into L.stage, so that the clipping loop
// reads them from the LDS instead of waiting on a global load per candidate.
template <int S>
VGC_HD void stage_cand(const ClipParams& P, CellLds<S>& L, int base, int cnt) {
    const LI k = base + lane_id();
    const LB ok = k < cnt;
    const LU64 key = lds_ld(L.cand, select(ok, k, LI(0)));
    const LI j = cvt<int>(cvt<uint32_t>(key));
    const Lane<pt4> q = gld(P.pts, select(ok, j, LI(0)), ok);
    wave_sync();  // the previous batch is consumed
    lds_st(L.stage, lane_id(), q, ok);
    wave_sync();
}

// List mode: entries lbeg+base .. +63 of the cell's list (those below len),
// their ids into L.cand and FP64 positions into L.stage, one load per lane,
// so that the clipping loop does not wait on two dependent loads per
// candidate.  An id of -1 (end of a short list) is kept as -1.
template <int S>
VGC_HD void stage_list(const ClipParams& P, CellLds<S>& L, long long lbeg, int base, int len) {
    const LI k = base + lane_id();
    const LB ok = k < len;
    const LI j = gld(P.list + lbeg, select(ok, k, LI(0)), ok);
    const LB okj = ok & (j >= 0);
    const Lane<pt4> q = gld(P.pts, select(okj, j, LI(0)), okj);
    wave_sync();  // the previous batch is consumed
    lds_st(L.cand, lane_id(), cvt<uint64_t>(cvt<uint32_t>(select(ok, j, LI(-1)))), LB(true));
    lds_st(L.stage, lane_id(), q, okj);
    wave_sync();
}

// Successor of corner t around face f (corner t has plane f): the corner
// across t's edge (previous plane, f); counter-clockwise seen from outside,
// since each corner's planes a, b, c have det(ea, eb, ec) > 0.
template <int S>
VGC_HD LI face_next(const CellLds<S>& L, const LI& t, const LI& f) {
    const LU plt = lds_ld(L.cpl, t);
    const LU adt = lds_ld(L.cadj, t);
    const LU fu = cvt<uint32_t>(f);
    constexpr int B = CornerBits<S>::value;
    const LI sh = select(byte_of(plt, 0) == fu, LI(2 * B), select(byte_of(plt, 1) == fu, LI(0), LI(B)));
    return cvt<int>((adt >> cvt<uint32_t>(sh)) & LU(CornerBits<S>::mask)) & (64 * S - 1);
}

// Wave-wide exclusive prefix sum of non-negative lane values.
VGC_HD LI wave_exscan(const LI& v) {
    const LI lane = lane_id();
    LI s = v;
    for (int d = 1; d < kWave; d <<= 1) {
        const LI o = shfl(s, select(lane >= d, lane - d, LI(0)));
        s = select(lane >= d, s + o, s);
    }
    return s - v;
}

// Sign of orient_3d(p0, p1, p2, p3) = sign det(p1 - p0, p2 - p0, p3 - p0)
// (numerics/predicates.cpp:1889-1899): FP64 with a static error bound, else
// exact as -det4 of the rows (p, 1).
VGC_SD int orient3d_sign(const pt4& p0, const pt4& p1, const pt4& p2, const pt4& p3) {
    const double a11 = p1.x - p0.x, a12 = p1.y - p0.y, a13 = p1.z - p0.z;
    const double a21 = p2.x - p0.x, a22 = p2.y - p0.y, a23 = p2.z - p0.z;
    const double a31 = p3.x - p0.x, a32 = p3.y - p0.y, a33 = p3.z - p0.z;
    const double m1 = a22 * a33 - a23 * a32, m2 = a12 * a33 - a13 * a32, m3 = a12 * a23 - a13 * a22;
    const double det = a11 * m1 - a21 * m2 + a31 * m3;
    const double mag = fabs(a11) * (fabs(a22 * a33) + fabs(a23 * a32)) +
                       fabs(a21) * (fabs(a12 * a33) + fabs(a13 * a32)) +
                       fabs(a31) * (fabs(a12 * a23) + fabs(a13 * a22));
    // differences exact to 1 ulp, 3 products and 2 sums per term: gamma_8
    // of the magnitude covers the evaluation
    if (fabs(det) > 2e-15 * mag) return det > 0.0 ? 1 : -1;
    const double r0[4] = {p0.x, p0.y, p0.z, 1.0}, r1[4] = {p1.x, p1.y, p1.z, 1.0};
    const double r2[4] = {p2.x, p2.y, p2.z, 1.0}, r3[4] = {p3.x, p3.y, p3.z, 1.0};
    return -det4_sign_exact(r0, r1, r2, r3);
}

// Cell mesh and primal tetrahedra of a finished cell (mesh_mode 1: counts,
// 2: write at the cell's offsets).  Vertices: the live corners in slot
// order; facets: the non-empty faces in plane order, each a cycle of
// corners (ConvexCell::append_to_mesh, voronoi/convex_cell.cpp:371-444, and
// ConvexCell::save, :282-358); tetrahedra: (i, j, k, l) for every corner
// whose three planes are bisectors with seeds j, k, l of larger caller
// index than i, oriented so that orient_3d > 0 or = 0
// (RestrictedVoronoiDiagram::compute_RDT, voronoi/RVD.cpp:2302-2327, and
// PrimalTetrahedronAction, generic_RVD.h:1035-1055).  (gx, gy, gz): the
// cell's centroid relative to the seed.
template <int S, class CS>
VGC_HD void mesh_emit(const ClipParams& P, CellLds<S>& L, const CS& st, int np,
                      int orig_c, const pt4& pc, double gx, double gy, double gz) {
    constexpr int kPMax = PMax<S>::value;
    const LI lane = lane_id();
    const long long n = P.mesh_n, n1 = n + 1;
    const bool write = P.mesh_mode == 2;
    long long voff = 0, foff = 0, fvoff = 0, toff = 0;
    if (write) {
        voff = P.mesh_off[orig_c];
        foff = P.mesh_off[n1 + orig_c];
        fvoff = P.mesh_off[2 * n1 + orig_c];
        toff = P.mesh_off[3 * n1 + orig_c];
    }
    // vertices: rank of each live corner (kept in L.job for the facet walk)
    int nv = 0;
    const double sh = P.mesh_shrink;
    wave_sync();
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const LB vs = mask_bit(st.valid[s]);
        const LI cid = lane + 64 * s;
        const LI r = nv + mbcnt(st.valid[s]);
        lds_st(L.job, cid, cvt<uint32_t>(r), vs);
        if (write && P.mesh_v) {
            const LD qx = lds_ld(&L.q[0][0], cid * 3 + 0), qy = lds_ld(&L.q[0][0], cid * 3 + 1),
                     qz = lds_ld(&L.q[0][0], cid * 3 + 2);
            const LD vx = pc.x + (sh * gx + (1.0 - sh) * qx);
            const LD vy = pc.y + (sh * gy + (1.0 - sh) * qy);
            const LD vz = pc.z + (sh * gz + (1.0 - sh) * qz);
            const LI gi = cvt<int>(r) * 3;
#if defined(VGC_EMU)
            for (int l = 0; l < kWave; ++l)
                if (vs[l]) {
                    P.mesh_v[3 * voff + gi[l] + 0] = vx[l];
                    P.mesh_v[3 * voff + gi[l] + 1] = vy[l];
                    P.mesh_v[3 * voff + gi[l] + 2] = vz[l];
                }
#else
            if (vs) {
                P.mesh_v[3 * voff + gi + 0] = vx;
                P.mesh_v[3 * voff + gi + 1] = vy;
                P.mesh_v[3 * voff + gi + 2] = vz;
            }
#endif
        }
        nv += popc64(st.valid[s]);
    }
    wave_sync();
    // facets, lane f walks face f
    int nf = 0, nfv = 0;
    for (int f0 = 0; f0 < np; f0 += 64) {
        const LI f = f0 + lane;
        const LB inr = f < np;
        const LI fs = select(inr, f, LI(0));
        LI t0 = lds_ld(L.v2t, fs);
        const LB act = inr & (t0 != LI(0x7FFFFFFF));
        t0 = select(act, t0, LI(0)) & (64 * S - 1);
        LI size = select(act, LI(1), LI(0));
        {
            LI t = face_next<S>(L, t0, fs);
            LB run = act & (t != t0);
            for (int guard = 0; guard < 64 * S && any(run); ++guard) {
                size = select(run, size + 1, size);
                t = select(run, face_next<S>(L, t, fs), t);
                run = run & (t != t0);
            }
        }
        const LB keep = act & (size >= 3);
        const uint64_t km = ballot(keep);
        const LI frank = nf + mbcnt(km);
        const LI start = nfv + wave_exscan(select(keep, size, LI(0)));
        if (write && P.mesh_fptr && km) {
            const LI pgf = lds_ld(L.pg, fs);
            LI nb = LI(-1);
#if defined(VGC_EMU)
            for (int l = 0; l < kWave; ++l)
                if (keep[l]) {
                    nb[l] = pgf[l] >= 0 ? P.orig[pgf[l]] : -1;
                    P.mesh_fptr[foff + frank[l]] = fvoff + start[l];
                    P.mesh_fnbr[foff + frank[l]] = nb[l];
                }
#else
            if (keep) {
                nb = pgf >= 0 ? P.orig[pgf] : -1;
                P.mesh_fptr[foff + frank] = fvoff + start;
                P.mesh_fnbr[foff + frank] = nb;
            }
#endif
            LI t = t0;
            LI k = LI(0);
            LB run = keep;
            for (int guard = 0; guard < 64 * S && any(run); ++guard) {
                const LI vr = cvt<int>(lds_ld(L.job, t));
#if defined(VGC_EMU)
                for (int l = 0; l < kWave; ++l)
                    if (run[l]) P.mesh_fv[fvoff + start[l] + k[l]] = int(voff) + vr[l];
#else
                if (run) P.mesh_fv[fvoff + start + k] = int(voff) + vr;
#endif
                k = k + 1;
                t = select(run, face_next<S>(L, t, fs), t);
                run = run & (t != t0);
            }
        }
        nf += popc64(km);
        nfv += popc64(km) ? (int)wave_sum(cvt<double>(select(keep, size, LI(0)))) : 0;
    }
    // primal tetrahedra
    int nt = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const LB vs = mask_bit(st.valid[s]);
        const LI a = cvt<int>(byte_of(st.pl[s], 0)) & (kPMax - 1);
        const LI b = cvt<int>(byte_of(st.pl[s], 1)) & (kPMax - 1);
        const LI c = cvt<int>(byte_of(st.pl[s], 2)) & (kPMax - 1);
        const LI ga = lds_ld(L.pg, a), gb = lds_ld(L.pg, b), gc = lds_ld(L.pg, c);
        LB cand = vs & (ga >= 0) & (gb >= 0) & (gc >= 0);
        LI oa = LI(-1), ob = LI(-1), oc = LI(-1);
#if defined(VGC_EMU)
        for (int l = 0; l < kWave; ++l)
            if (cand[l]) {
                oa[l] = P.orig[ga[l]];
                ob[l] = P.orig[gb[l]];
                oc[l] = P.orig[gc[l]];
            }
#else
        if (cand) {
            oa = P.orig[ga];
            ob = P.orig[gb];
            oc = P.orig[gc];
        }
#endif
        const LB mine = cand & (oa > orig_c) & (ob > orig_c) & (oc > orig_c);
        const uint64_t mm = ballot(mine);
        if (write && P.mesh_tet && mm) {
            const LI tr = nt + mbcnt(mm);
#if defined(VGC_EMU)
            for (int l = 0; l < kWave; ++l)
                if (mine[l]) {
                    int v0 = orig_c, v1 = oa[l];
                    if (orient3d_sign(pc, P.pts[ga[l]], P.pts[gb[l]], P.pts[gc[l]]) < 0) {
                        v0 = oa[l];
                        v1 = orig_c;
                    }
                    int* T = P.mesh_tet + 4 * (toff + tr[l]);
                    T[0] = v0;
                    T[1] = v1;
                    T[2] = ob[l];
                    T[3] = oc[l];
                }
#else
            if (mine) {
                int v0 = orig_c, v1 = oa;
                if (orient3d_sign(pc, P.pts[ga], P.pts[gb], P.pts[gc]) < 0) {
                    v0 = oa;
                    v1 = orig_c;
                }
                int* T = P.mesh_tet + 4 * (toff + tr);
                T[0] = v0;
                T[1] = v1;
                T[2] = ob;
                T[3] = oc;
            }
#endif
        }
        nt += popc64(mm);
    }
    if (!write && lane0()) {
        P.mesh_cnt[orig_c] = nv;
        P.mesh_cnt[n + orig_c] = nf;
        P.mesh_cnt[2 * n + orig_c] = nfv;
        P.mesh_cnt[3 * n + orig_c] = nt;
    }
}

// Security-radius loop + integration for work item `item`.  GATHER: the
// wave finds the neighbours itself, ring after ring of distance; otherwise
// it reads them from P.list.
template <int S, bool EXACT, bool GATHER, bool W, bool CORNER = W, class T = double>
VGC_HD void process_cell(const ClipParams& P, int item, CellLds<S>& L) {
    constexpr int kPMax = PMax<S>::value;
    const LI lane = lane_id();
    wave_sync();  // the previous cell of this wave is done with the LDS
    const int c = P.items ? P.items[item] : item;
    const int inst = P.inst_of ? P.inst_of[c] : 0;
    // a pass that only asks whether an instance has an empty cell stops that
    // instance at its first one
    if (P.abort_flag && flag_load(P.abort_flag + inst)) return;
    if (P.n_real > 0 && P.orig[c] >= P.n_real) {
        // an image of a seed across the period: a candidate only
        if (lane0()) P.flags[c] = CELL_DONE;
        return;
    }
    const pt4 pc = P.pts[c];
    const double* box = P.item_box ? P.item_box + 6 * (long long)item : P.box + 6 * inst;
    // outputs by item for pieces, else by seed (caller's index) and Morton
    // position
    const int oc = P.piece ? item : P.orig[c];
    const int fc = P.piece ? item : c;
    if (P.item_box && !(box[3] > box[0] && box[4] > box[1] && box[5] > box[2])) {
        // a flat octant (the seed on the domain's boundary): nothing of the
        // cell in it
        if (lane0()) {
            P.vol[oc] = 0.0;
            P.mg[3 * (long long)oc + 0] = P.mg[3 * (long long)oc + 1] = 0.0;
            P.mg[3 * (long long)oc + 2] = 0.0;
            P.nbr_cnt[oc] = 0;
            if (P.e2) P.e2[oc] = 0.0;
            P.flags[fc] = CELL_DONE | CELL_EMPTY;
            P.r2[fc] = 0.0;
        }
        return;
    }

    CellRun<S, T> R;
    R.pc = pc;
    R.c = c;
    R.orig_c = P.orig[c];
    R.wi = pc.w;
    R.cw = W ? (pc.w - P.wmax[inst]) : 0.0;
    R.box = box;
    R.np = 6;
    R.flags = 0;
    R.stopped = false;
    R.prev.x = R.prev.y = R.prev.z = R.prev.w = NAN;
    R.prev_slot = -1;
    R.nclip = R.ntest = R.nlive = R.nnew = 0;
    init_box<S, T>(R.st, L, box, pc);
    bool complete = false;

    if (!GATHER) {
        long long lbeg;
        int llen;
        if (P.list_ptr) {
            lbeg = P.list_ptr[item];
            llen = P.list_len[item];
        } else {
            lbeg = (long long)c * P.list_stride;  // n x K lists by Morton position
            llen = P.list_stride;
        }
        complete = P.list_complete != 0 || llen >= P.inst_n[inst] - 1;
        for (int jj = 0; jj < llen; ++jj) {
            if ((jj & (kWave - 1)) == 0) stage_list<S>(P, L, lbeg, jj, llen);
            const int j = int(uint32_t(L.cand[jj & (kWave - 1)]));
            if (j < 0) {
                complete = true;  // short list: the instance has fewer seeds
                break;
            }
            if (consider<S, EXACT, W, T>(P, L, R, j, L.stage[jj & (kWave - 1)])) break;
        }
    } else {
        const pf4 pq = P.pf[c];
        const int p0i = P.IL.p0[inst], ni = P.IL.n[inst];
        // first ring, about 56 seeds: 2.4 times the distance to the 4th
        // nearest seed of the own bucket (Morton neighbours, an upper bound of
        // the 4th nearest distance)
        const int bq = (c - p0i) / kBucket;
        float rhi2;
        {
            LI m = p0i + bq * kBucket + (lane & (kBucket - 1));
            LB ok = (lane < kBucket) & (m < p0i + ni) & (m != c);
            LF d2 = lane_dist2(P.pf, select(ok, m, LI(0)), ok, pq.x, pq.y, pq.z);
            d2 = select(ok, d2, LF(INFINITY));
            float d4 = INFINITY;
            for (int k = 0; k < kBucket; ++k) {
                const float t = readlane(d2, k);
                if (popc64(ballot(d2 < t)) == 3 && t < d4) d4 = t;
            }
            rhi2 = 5.8f * d4;
            if (!(rhi2 > 0.0f) || !(rhi2 < INFINITY)) {
                const bnode ob = P.nodes[P.IL.t0[inst] + P.IL.L[inst] + bq];
                const float gx = ob.hi[0] - ob.lo[0], gy = ob.hi[1] - ob.lo[1],
                            gz = ob.hi[2] - ob.lo[2];
                rhi2 = 0.3f * (gx * gx + gy * gy + gz * gz);
            }
        }
        const float bx = float(box[3] - box[0]), by = float(box[4] - box[1]),
                    bz = float(box[5] - box[2]);
        const float all2 = f_up(1.0001 * double(bx * bx + by * by + bz * bz));
        if (!(rhi2 > 0.0f)) rhi2 = all2 * 1e-6f;
        float rlo2 = -1.0f;
        PowerFilter F{};
        CornerTab CT;
        CT.live = 0;
        bool ct_on = false;
        int nfilt = 0;
        while (true) {
            // first ring: level-synchronous walk; later rings: depth first,
            // bounded to the nearest seeds when the buffer fills
            int cnt = rlo2 >= 0.0f
                          ? gather_dfs<S, CORNER>(P, L, c, inst, pq.x, pq.y, pq.z, rlo2, rhi2, F, CT, ct_on)
                          : gather_ring<S>(P, L, c, inst, pq.x, pq.y, pq.z, rlo2, rhi2, F);
            if (F.on && ++nfilt > P.ring_limit) {
                // a second gather (then the list tiers) takes the cell
                R.flags |= CELL_NEED_K | CELL_WHY_RINGS;
                break;
            }
            if (cnt < 0) R.flags |= CELL_WHY_WALK;
            if (cnt < 0 || cnt > kCandCap) {
                // too many seeds (or tree nodes) in the ring: narrow it, same
                // shell-volume fraction in r^3, and gather again
                float lo = rlo2 > 0.0f ? rlo2 : 0.0f;
                float r3lo = lo * sqrtf(lo), r3hi = rhi2 * sqrtf(rhi2);
                float frac = cnt < 0 ? 0.125f : 0.75f * kCandCap / float(cnt);
                float r3 = r3lo + (r3hi - r3lo) * frac;
                float nh = cbrtf(r3);
                nh = nh * nh;
                if (!(nh > lo * 1.000001f) || !(nh < rhi2)) {
                    R.flags |= CELL_NEED_K | CELL_WHY_NARROW;  // seeds at one distance
                    break;
                }
                rhi2 = nh;
                continue;
            }
            sort_cand<S>(L, cnt);
            bool done = false;
            for (int jj = 0; jj < cnt; ++jj) {
                if ((jj & (kWave - 1)) == 0) stage_cand<S>(P, L, jj, cnt);
                const int j = int(uint32_t(L.cand[jj]));
                if (consider<S, EXACT, W, T>(P, L, R, j, L.stage[jj & (kWave - 1)])) {
                    done = true;
                    break;
                }
            }
            if (done) break;
            // ring exhausted: can a seed beyond rhi2 still cut the cell?
            LD r2m = LD(0.0);
#pragma unroll
            for (int s = 0; s < S; ++s)
                r2m = vmax(r2m, select(mask_bit(R.st.valid[s]), cvt<double>(R.st.r2[s]), LD(0.0)));
            const double r2now = wave_max(r2m);
            const float need2 = radius_for(r2now, R.wi, W ? P.wmax[inst] : 0.0, W);
            if (need2 <= rhi2 || rhi2 >= all2) {
                complete = true;
                break;
            }
            rlo2 = rhi2;
            if (make_filter<S, CORNER>(R, F, CT)) {
                // every seed beyond the ring, within the max-weight radius,
                // that passes the power filter of the cell as it is now
                // (narrowed above if too many)
                rhi2 = need2 < all2 ? need2 : all2;
            } else {
                F.on = 0;
                CT.live = 0;
                rhi2 = need2 < all2 ? need2 : all2;
            }
            wave_sync();
        }
    }
    const bool stopped = R.stopped;
    uint32_t flags = R.flags;
    CellState<S, T>& st = R.st;
    int np = R.np;
    const unsigned long long nclip = R.nclip, ntest = R.ntest, nlive = R.nlive, nnew = R.nnew;
    const int orig_c = R.orig_c;
    if (!stopped && !complete && !(flags & (CELL_OVERFLOW | CELL_BAD_TOPO))) flags |= CELL_NEED_K;

    // squared radius of the (possibly partial) cell
    LD r2m = LD(0.0);
#pragma unroll
    for (int s = 0; s < S; ++s)
        r2m = vmax(r2m, select(mask_bit(st.valid[s]), cvt<double>(st.r2[s]), LD(0.0)));
    double r2max = wave_max(r2m);

    if (P.counters && lane0()) {
        atomic_add_u64(P.counters + 0, nclip);
        atomic_add_u64(P.counters + 1, ntest);
        atomic_add_u64(P.counters + 2, nlive);
        atomic_add_u64(P.counters + 3, nnew);
    }
    if (Prec<T>::fp32 && (flags & CELL_UNCERTAIN)) {
        // FP32 corners that could not certify a decision: only the flag (the
        // FP64 pass redoes the cell and writes every output)
        if (lane0()) {
            P.flags[fc] = CELL_UNCERTAIN;
            P.r2[fc] = 0.0;
        }
        return;
    }
    if (flags & (CELL_NEED_K | CELL_OVERFLOW | CELL_BAD_TOPO)) {
        // the partial cell contains the cell: its ball filters the list tiers
        PowerFilter G{};
        CornerTab GT;
        const bool okf = P.ball && !(flags & CELL_BAD_TOPO) && make_filter<S, false>(R, G, GT);
        if (lane0()) {
            P.flags[fc] = flags;
            P.r2[fc] = r2max;
            if (P.ball) {
                fball b;
                b.cx = G.cx;
                b.cy = G.cy;
                b.cz = G.cz;
                b.rc = okf ? G.rc : -1.0f;
                b.hx = G.hx;
                b.hy = G.hy;
                b.hz = G.hz;
                b.pad = 0.0f;
                P.ball[fc] = b;
            }
        }
        return;
    }

    // ------------------------------------------------------ integration
    if (flags & CELL_EMPTY) {
        if (P.abort_flag && lane0()) flag_raise(P.abort_flag + inst);
        if (lane0()) {
            P.vol[oc] = 0.0;
            P.mg[3 * (long long)oc + 0] = 0.0;
            P.mg[3 * (long long)oc + 1] = 0.0;
            P.mg[3 * (long long)oc + 2] = 0.0;
            P.nbr_cnt[oc] = 0;
            if (P.e2) P.e2[oc] = 0.0;
            if (P.mesh_mode == 1)
                for (int k = 0; k < 4; ++k) P.mesh_cnt[(long long)k * P.mesh_n + orig_c] = 0;
            P.flags[fc] = flags | CELL_DONE;
            P.r2[fc] = 0.0;
        }
        return;
    }
    // corner points and topology to LDS
    LD r2f = LD(0.0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LB vs = mask_bit(st.valid[s]);
        LI cid = lane + 64 * s;
        LD nx, ny, nz, dd;
        if constexpr (Prec<T>::fp32) {
            // the FP64 corner of the plane triple, as the FP64 pass builds it
            LD eax, eay, eaz, ha, ebx, eby, ebz, hb, ecx, ecy, ecz, hc;
            load_row<double>(L, cvt<int>(byte_of(st.pl[s], 0)) & (kPMax - 1), eax, eay, eaz, ha);
            load_row<double>(L, cvt<int>(byte_of(st.pl[s], 1)) & (kPMax - 1), ebx, eby, ebz, hb);
            load_row<double>(L, cvt<int>(byte_of(st.pl[s], 2)) & (kPMax - 1), ecx, ecy, ecz, hc);
            const Corner C = make_corner<double>(eax, eay, eaz, ha, ebx, eby, ebz, hb, ecx, ecy, ecz, hc);
            nx = C.nx;
            ny = C.ny;
            nz = C.nz;
            dd = C.d;
            r2f = vmax(r2f, select(vs, C.r2, LD(0.0)));
        } else {
            nx = st.nx[s];
            ny = st.ny[s];
            nz = st.nz[s];
            dd = st.d[s];
        }
        LD inv = 1.0 / select(vs, dd, LD(1.0));
        lds_st(&L.q[0][0], cid * 3 + 0, nx * inv, vs);
        lds_st(&L.q[0][0], cid * 3 + 1, ny * inv, vs);
        lds_st(&L.q[0][0], cid * 3 + 2, nz * inv, vs);
        lds_st(L.cpl, cid, st.pl[s], vs);
        lds_st(L.cadj, cid, st.adj[s], vs);
    }
    // the FP64 cell's radius (the FP32 one bounds it from above)
    if constexpr (Prec<T>::fp32) r2max = wave_max(r2f);
    for (int f0 = 0; f0 < np; f0 += 64) {
        LI f = f0 + lane;
        lds_st(L.v2t, select(f < np, f, LI(0)), LI(0x7FFFFFFF), f < np);
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < S; ++s) {
        LB vs = mask_bit(st.valid[s]);
        LI cid = lane + 64 * s;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            LI p = cvt<int>(byte_of(st.pl[s], k)) & (kPMax - 1);
            lds_min(L.v2t, p, cid, vs);
        }
    }
    wave_sync();
    int cid0 = 0;
#pragma unroll
    for (int s = S - 1; s >= 0; --s)
        if (st.valid[s]) cid0 = 64 * s + ffs64(st.valid[s]);
    const double ox = L.q[cid0][0], oy = L.q[cid0][1], oz = L.q[cid0][2];

    LD vacc = LD(0.0), mxa = LD(0.0), mya = LD(0.0), mza = LD(0.0), e2a = LD(0.0);
    LD ntet = LD(0.0);
    int nfacets = 0;
    int out_base = 0;
    bool face_overflow = false;
    const long long obase = (long long)oc * P.face_stride;
    for (int f0 = 0; f0 < np; f0 += 64) {
        LI f = f0 + lane;
        LB inr = f < np;
        LI fs = select(inr, f, LI(0));
        LI t0 = lds_ld(L.v2t, fs);
        LB act = inr & (t0 != LI(0x7FFFFFFF));
        t0 = select(act, t0, LI(0)) & (64 * S - 1);
        // successor around face f: across edge (p[lv+2] -> p[lv])
        auto nxt = [&](const LI& t) -> LI { return face_next<S>(L, t, fs); };
        LD q0x = lds_ld(&L.q[0][0], t0 * 3 + 0), q0y = lds_ld(&L.q[0][0], t0 * 3 + 1),
           q0z = lds_ld(&L.q[0][0], t0 * 3 + 2);
        LI t1 = nxt(t0);
        LI t = nxt(t1);
        LD q1x = lds_ld(&L.q[0][0], t1 * 3 + 0), q1y = lds_ld(&L.q[0][0], t1 * 3 + 1),
           q1z = lds_ld(&L.q[0][0], t1 * 3 + 2);
        LD area = LD(0.0), fv = LD(0.0), fmx = LD(0.0), fmy = LD(0.0), fmz = LD(0.0),
           fe = LD(0.0);
        // |o|^2 + |q0|^2, constant along the face (second moments)
        const LD oq0 = (ox * ox + oy * oy + oz * oz) + (q0x * q0x + q0y * q0y + q0z * q0z);
        LB run = act & (t != t0);
        for (int guard = 0; guard < 64 * S && any(run); ++guard) {
            LD qx = lds_ld(&L.q[0][0], t * 3 + 0), qy = lds_ld(&L.q[0][0], t * 3 + 1),
               qz = lds_ld(&L.q[0][0], t * 3 + 2);
            // tetrahedron (apex, q0, q1, q) as in tet_volume (convex_cell.cpp:1193-1213)
            LD ux = q0x - ox, uy = q0y - oy, uz = q0z - oz;
            LD vx = q1x - ox, vy = q1y - oy, vz = q1z - oz;
            LD wx = qx - ox, wy = qy - oy, wz = qz - oz;
            LD cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            LD tv = vabs(cx * wx + cy * wy + cz * wz) * (1.0 / 6.0);
            // triangle (q0, q1, q) as in triangle_area (convex_cell.cpp:1147-1158)
            LD ax = q1x - q0x, ay = q1y - q0y, az = q1z - q0z;
            LD bx = qx - q0x, by = qy - q0y, bz = qz - q0z;
            LD wx2 = ay * bz - az * by, wy2 = az * bx - ax * bz, wz2 = ax * by - ay * bx;
            LD ta = 0.5 * vsqrt(wx2 * wx2 + wy2 * wy2 + wz2 * wz2);
            ntet = select(run, ntet + 1.0, ntet);
            fv = select(run, fv + tv, fv);
            const LD sx = ox + q0x + q1x + qx, sy = oy + q0y + q1y + qy, sz = oz + q0z + q1z + qz;
            fmx = select(run, fmx + tv * sx / 4.0, fmx);
            fmy = select(run, fmy + tv * sy / 4.0, fmy);
            fmz = select(run, fmz + tv * sz / 4.0, fmz);
            if (P.e2) {
                // second moment about the seed (seed-relative corners a..d):
                // V/20 (sum |a|^2 + |a + b + c + d|^2)
                const LD nn = oq0 + (q1x * q1x + q1y * q1y + q1z * q1z) + (qx * qx + qy * qy + qz * qz) +
                              (sx * sx + sy * sy + sz * sz);
                fe = select(run, fe + tv * nn * (1.0 / 20.0), fe);
            }
            area = select(run, area + ta, area);
            q1x = select(run, qx, q1x);
            q1y = select(run, qy, q1y);
            q1z = select(run, qz, q1z);
            LI tn = nxt(t);
            t = select(run, tn, t);
            run = run & (t != t0);
        }
        vacc = vacc + fv;
        e2a = e2a + fe;
        mxa = mxa + fmx;
        mya = mya + fmy;
        mza = mza + fmz;
        nfacets += popc64(ballot(act));
        LI pgf = lds_ld(L.pg, fs);
        LB outm = act & (pgf >= 0);
        uint64_t om = ballot(outm);
        LI rank = out_base + mbcnt(om);
        if (out_base + popc64(om) > P.face_stride) {
            face_overflow = true;
        } else {
            // neighbour ids and areas straight to HBM (caller's numbering)
            LI gj = select(outm, pgf, LI(0));
            LI oj = LI(0);
#if defined(VGC_EMU)
            for (int l = 0; l < kWave; ++l)
                if (outm[l]) oj[l] = P.src ? P.src[P.orig[gj[l]]] : P.orig[gj[l]];
            for (int l = 0; l < kWave; ++l)
                if (outm[l]) {
                    P.nbr_tmp[obase + rank[l]] = oj[l];
                    P.area_tmp[obase + rank[l]] = area[l];
                }
#else
            if (outm) {
                oj = P.src ? P.src[P.orig[gj]] : P.orig[gj];
                P.nbr_tmp[obase + rank] = oj;
                P.area_tmp[obase + rank] = area;
            }
#endif
        }
        out_base += popc64(om);
    }
    const double V = wave_sum(vacc);
    const double ntets = wave_sum(ntet);
    // (the FP32 pass runs without cell meshes: the host sends mesh exports to
    // the FP64 pass)
    if (!Prec<T>::fp32 && P.mesh_mode && !face_overflow) {
        const double iv = V > 0.0 ? 1.0 / V : 0.0;
        mesh_emit<S>(P, L, st, np, orig_c, pc, wave_sum(mxa) * iv, wave_sum(mya) * iv,
                     wave_sum(mza) * iv);
    }
    if (P.counters && lane0()) {
        atomic_add_u64(P.counters + 4, (unsigned long long)ntets);
        atomic_add_u64(P.counters + 5, (unsigned long long)nfacets);
    }
    const double Mx = wave_sum(mxa), My = wave_sum(mya), Mz = wave_sum(mza);
    if (face_overflow) flags |= CELL_OVERFLOW;
    if (lane0()) {
        if (!face_overflow) {
            P.vol[oc] = V;
            P.mg[3 * (long long)oc + 0] = Mx + V * pc.x;
            P.mg[3 * (long long)oc + 1] = My + V * pc.y;
            P.mg[3 * (long long)oc + 2] = Mz + V * pc.z;
            P.nbr_cnt[oc] = out_base;
            P.flags[fc] = flags | CELL_DONE;
        } else {
            P.flags[fc] = flags;
            if (P.ball) {
                fball b;
                b.cx = b.cy = b.cz = b.hx = b.hy = b.hz = b.pad = 0.0f;
                b.rc = -1.0f;
                P.ball[fc] = b;
            }
        }
        P.r2[fc] = r2max;
    }
    if (P.e2) {
        const double E2 = wave_sum(e2a);
        if (lane0() && !face_overflow) P.e2[oc] = E2;
    }
}

}  // namespace vgc

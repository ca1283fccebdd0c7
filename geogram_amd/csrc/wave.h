// Wavefront programming layer shared by the gfx950 kernels and their CPU
// emulation.
//
// The clipping kernel runs one 64-lane wavefront per Voronoi / Laguerre cell:
// every lane owns cell corners in registers and the lanes cooperate through
// ballots, lane-rank prefix counts and the wave's slice of LDS.  The same
// source is compiled twice:
//   * for gfx950 (hipcc), where Lane<T> is T, a ballot is __ballot and an LDS
//     gather is an ordinary indexed load;
//   * for the host with VGC_EMU defined (g++), where Lane<T> holds the 64 lane
//     values and every operation loops over the lanes in lane order.
// The emulation reproduces the kernel bit for bit (same operations, same fused
// multiply-adds, same reduction trees), so the kernel's logic can be tested on
// a machine without a GPU.
//
// Rules for code written on top of this layer: control flow that encloses a
// wave_sync() or a ballot must be wave-uniform; per-lane choices are made with
// select() and masked stores, never with if/else on a Lane<bool>.
#pragma once

#include <stdint.h>

#if defined(VGC_EMU)

#include <cmath>
#include <cstring>

#define VGC_HD inline
#define VGC_SD inline
#define VGC_DEVICE_ONLY inline

namespace vgc {

constexpr int kWave = 64;

template <class T>
struct ident {
    typedef T type;
};

template <class T>
struct Lane {
    T v[kWave];
    Lane() {}
    Lane(T s) {  // broadcast
        for (int i = 0; i < kWave; ++i) v[i] = s;
    }
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
};

#define VGC_LANE_BINOP(OP)                                                   \
    template <class T>                                                       \
    inline Lane<T> operator OP(const Lane<T>& a, const Lane<T>& b) {         \
        Lane<T> r;                                                           \
        for (int i = 0; i < kWave; ++i) r.v[i] = T(a.v[i] OP b.v[i]);        \
        return r;                                                            \
    }                                                                        \
    template <class T>                                                       \
    inline Lane<T> operator OP(const Lane<T>& a, typename ident<T>::type b) { \
        Lane<T> r;                                                           \
        for (int i = 0; i < kWave; ++i) r.v[i] = T(a.v[i] OP b);             \
        return r;                                                            \
    }                                                                        \
    template <class T>                                                       \
    inline Lane<T> operator OP(typename ident<T>::type a, const Lane<T>& b) { \
        Lane<T> r;                                                           \
        for (int i = 0; i < kWave; ++i) r.v[i] = T(a OP b.v[i]);             \
        return r;                                                            \
    }                                                                        \
    template <class T>                                                       \
    inline Lane<T>& operator OP##=(Lane<T>& a, const Lane<T>& b) {           \
        a = a OP b;                                                          \
        return a;                                                            \
    }                                                                        \
    template <class T>                                                       \
    inline Lane<T>& operator OP##=(Lane<T>& a, typename ident<T>::type b) {  \
        a = a OP b;                                                          \
        return a;                                                            \
    }

VGC_LANE_BINOP(+)
VGC_LANE_BINOP(-)
VGC_LANE_BINOP(*)
VGC_LANE_BINOP(/)
VGC_LANE_BINOP(%)
VGC_LANE_BINOP(&)
VGC_LANE_BINOP(|)
VGC_LANE_BINOP(^)
VGC_LANE_BINOP(<<)
VGC_LANE_BINOP(>>)
#undef VGC_LANE_BINOP

#define VGC_LANE_CMP(OP)                                                         \
    template <class T>                                                           \
    inline Lane<bool> operator OP(const Lane<T>& a, const Lane<T>& b) {          \
        Lane<bool> r;                                                            \
        for (int i = 0; i < kWave; ++i) r.v[i] = (a.v[i] OP b.v[i]);             \
        return r;                                                                \
    }                                                                            \
    template <class T>                                                           \
    inline Lane<bool> operator OP(const Lane<T>& a, typename ident<T>::type b) { \
        Lane<bool> r;                                                            \
        for (int i = 0; i < kWave; ++i) r.v[i] = (a.v[i] OP b);                  \
        return r;                                                                \
    }                                                                            \
    template <class T>                                                           \
    inline Lane<bool> operator OP(typename ident<T>::type a, const Lane<T>& b) { \
        Lane<bool> r;                                                            \
        for (int i = 0; i < kWave; ++i) r.v[i] = (a OP b.v[i]);                  \
        return r;                                                                \
    }

VGC_LANE_CMP(<)
VGC_LANE_CMP(<=)
VGC_LANE_CMP(>)
VGC_LANE_CMP(>=)
VGC_LANE_CMP(==)
VGC_LANE_CMP(!=)
#undef VGC_LANE_CMP

template <class T>
inline Lane<T> operator-(const Lane<T>& a) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = -a.v[i];
    return r;
}
inline Lane<bool> operator!(const Lane<bool>& a) {
    Lane<bool> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = !a.v[i];
    return r;
}
template <class T>
inline Lane<T> operator~(const Lane<T>& a) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = ~a.v[i];
    return r;
}

typedef Lane<double> LD;
typedef Lane<float> LF;
typedef Lane<int> LI;
typedef Lane<uint32_t> LU;
typedef Lane<uint64_t> LU64;
typedef Lane<bool> LB;

template <class T, class U>
inline Lane<T> cvt(const Lane<U>& a) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = T(a.v[i]);
    return r;
}
template <class T, class U>
inline T cvt(U a) {
    return T(a);
}

template <class T>
inline Lane<T> select(const LB& m, const Lane<T>& a, const Lane<T>& b) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = m.v[i] ? a.v[i] : b.v[i];
    return r;
}
template <class T>
inline Lane<T> select(const LB& m, const Lane<T>& a, typename ident<T>::type b) {
    return select(m, a, Lane<T>(b));
}
template <class T>
inline Lane<T> select(const LB& m, typename ident<T>::type a, const Lane<T>& b) {
    return select(m, Lane<T>(a), b);
}
template <class T>
inline Lane<T> select_s(const LB& m, T a, T b) {
    return select(m, Lane<T>(a), Lane<T>(b));
}

inline LD vfma(const LD& a, const LD& b, const LD& c) {
    LD r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fma(a.v[i], b.v[i], c.v[i]);
    return r;
}
inline LD vabs(const LD& a) {
    LD r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fabs(a.v[i]);
    return r;
}
inline LD vsqrt(const LD& a) {
    LD r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::sqrt(a.v[i]);
    return r;
}
inline LD vmax(const LD& a, const LD& b) {
    LD r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fmax(a.v[i], b.v[i]);
    return r;
}
inline LD vmin(const LD& a, const LD& b) {
    LD r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fmin(a.v[i], b.v[i]);
    return r;
}
inline double vfma(double a, double b, double c) { return std::fma(a, b, c); }
inline double vabs(double a) { return std::fabs(a); }
inline double vsqrt(double a) { return std::sqrt(a); }
inline double vmax(double a, double b) { return std::fmax(a, b); }
inline double vmin(double a, double b) { return std::fmin(a, b); }
// FP32 lanes under the same names (corner state of the FP32 clipping pass)
inline LF vfma(const LF& a, const LF& b, const LF& c) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fma(a.v[i], b.v[i], c.v[i]);
    return r;
}
inline LF vabs(const LF& a) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fabs(a.v[i]);
    return r;
}
inline LF vsqrt(const LF& a) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::sqrt(a.v[i]);
    return r;
}
inline LF vmax(const LF& a, const LF& b) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fmax(a.v[i], b.v[i]);
    return r;
}
inline LF vmin(const LF& a, const LF& b) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fmin(a.v[i], b.v[i]);
    return r;
}
inline float vfma(float a, float b, float c) { return std::fma(a, b, c); }
inline float vabs(float a) { return std::fabs(a); }
inline float vsqrt(float a) { return std::sqrt(a); }
inline float vmax(float a, float b) { return std::fmax(a, b); }
inline float vmin(float a, float b) { return std::fmin(a, b); }

inline LF vmaxf(const LF& a, const LF& b) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fmax(a.v[i], b.v[i]);
    return r;
}
inline LF vmaxf(const LF& a, float b) { return vmaxf(a, LF(b)); }
inline LF vsqrtf(const LF& a) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::sqrt(a.v[i]);
    return r;
}
inline LF vabsf(const LF& a) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fabs(a.v[i]);
    return r;
}
inline float vabsf(float a) { return std::fabs(a); }
inline LF vminf(const LF& a, const LF& b) {
    LF r;
    for (int i = 0; i < kWave; ++i) r.v[i] = std::fmin(a.v[i], b.v[i]);
    return r;
}
inline float vminf(float a, float b) { return std::fmin(a, b); }
inline LU fbits(const LF& a) {
    LU r;
    for (int i = 0; i < kWave; ++i) std::memcpy(&r.v[i], &a.v[i], 4);
    return r;
}
// global-memory gather: base[idx] on the lanes of m, T() elsewhere
template <class T>
inline Lane<T> gld(const T* base, const LI& idx, const LB& m) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = m.v[i] ? base[idx.v[i]] : T();
    return r;
}

inline LI lane_id() {
    LI r;
    for (int i = 0; i < kWave; ++i) r.v[i] = i;
    return r;
}

inline uint64_t ballot(const LB& b) {
    uint64_t m = 0;
    for (int i = 0; i < kWave; ++i)
        if (b.v[i]) m |= (uint64_t(1) << i);
    return m;
}

inline int popc64(uint64_t m) { return __builtin_popcountll(m); }
inline int ffs64(uint64_t m) { return m ? __builtin_ctzll(m) : -1; }

// number of set bits of m below each lane
inline LI mbcnt(uint64_t m) {
    LI r;
    for (int i = 0; i < kWave; ++i)
        r.v[i] = __builtin_popcountll(m & ((uint64_t(1) << i) - 1));
    return r;
}

// bit (lane-th) of a uniform mask, per lane
inline LB mask_bit(uint64_t m) {
    LB r;
    for (int i = 0; i < kWave; ++i) r.v[i] = (m >> i) & 1;
    return r;
}

// test bit idx of a uniform mask at a per-lane index (0..63)
inline LB mask_test(uint64_t m, const LI& idx) {
    LB r;
    for (int i = 0; i < kWave; ++i) r.v[i] = (m >> (idx.v[i] & 63)) & 1;
    return r;
}

template <class T>
inline T readlane(const Lane<T>& a, int lane) {
    return a.v[lane];
}
template <class T>
inline Lane<T> shfl(const Lane<T>& a, const LI& src) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = a.v[src.v[i] & 63];
    return r;
}
template <class T>
inline Lane<T> shfl_xor(const Lane<T>& a, int m) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = a.v[i ^ m];
    return r;
}

// Deterministic butterfly reductions: every lane ends with the same value.
inline double wave_sum(const LD& a) {
    LD x = a;
    for (int m = 32; m >= 1; m >>= 1) x = x + shfl_xor(x, m);
    return x.v[0];
}
inline double wave_max(const LD& a) {
    LD x = a;
    for (int m = 32; m >= 1; m >>= 1) x = vmax(x, shfl_xor(x, m));
    return x.v[0];
}
inline float wave_maxf(const LF& a) {
    LF x = a;
    for (int m = 32; m >= 1; m >>= 1) x = vmaxf(x, shfl_xor(x, m));
    return x.v[0];
}

// LDS access: gather, masked scatter (lane order: higher lane wins on a
// collision, like the hardware's unspecified order; callers never rely on it)
template <class T>
inline Lane<T> lds_ld(const T* p, const LI& idx) {
    Lane<T> r;
    for (int i = 0; i < kWave; ++i) r.v[i] = p[idx.v[i]];
    return r;
}
template <class T>
inline void lds_st(T* p, const LI& idx, const Lane<T>& val, const LB& m) {
    for (int i = 0; i < kWave; ++i)
        if (m.v[i]) p[idx.v[i]] = val.v[i];
}
inline void lds_min(int* p, const LI& idx, const LI& val, const LB& m) {
    for (int i = 0; i < kWave; ++i)
        if (m.v[i] && val.v[i] < p[idx.v[i]]) p[idx.v[i]] = val.v[i];
}
inline void wave_sync() {}

inline bool any(const LB& b) { return ballot(b) != 0; }

// a store of a wave-uniform value, done once
inline bool lane0() { return true; }

inline void atomic_add_u64(unsigned long long* p, unsigned long long v) {
    __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
}
// a flag shared by the cells of a pass: read, and raised once set
inline unsigned flag_load(const unsigned* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void flag_raise(unsigned* p) { __atomic_fetch_or(p, 1u, __ATOMIC_RELAXED); }

}  // namespace vgc

#else  // ---------------------------------------------------------- gfx950

#include <hip/hip_runtime.h>

#define VGC_HD __device__ __forceinline__
#define VGC_SD __host__ __device__ __forceinline__
#define VGC_DEVICE_ONLY __device__ __forceinline__

namespace vgc {

constexpr int kWave = 64;

template <class T>
using Lane = T;
typedef double LD;
typedef float LF;
typedef int LI;
typedef uint32_t LU;
typedef uint64_t LU64;
typedef bool LB;

template <class T, class U>
VGC_HD T cvt(U a) {
    return T(a);
}

template <class T>
VGC_HD T select(bool m, T a, T b) {
    return m ? a : b;
}
template <class T>
VGC_HD T select_s(bool m, T a, T b) {
    return m ? a : b;
}

VGC_HD double vfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
VGC_HD double vabs(double a) { return __builtin_fabs(a); }
VGC_HD double vsqrt(double a) { return __builtin_sqrt(a); }
VGC_HD double vmax(double a, double b) { return __builtin_fmax(a, b); }
VGC_HD double vmin(double a, double b) { return __builtin_fmin(a, b); }
VGC_HD float vfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
VGC_HD float vabs(float a) { return __builtin_fabsf(a); }
VGC_HD float vsqrt(float a) { return __builtin_sqrtf(a); }
VGC_HD float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
VGC_HD float vmin(float a, float b) { return __builtin_fminf(a, b); }

VGC_HD float vmaxf(float a, float b) { return __builtin_fmaxf(a, b); }
VGC_HD float vsqrtf(float a) { return __builtin_sqrtf(a); }
VGC_HD float vabsf(float a) { return __builtin_fabsf(a); }
VGC_HD float vminf(float a, float b) { return __builtin_fminf(a, b); }
VGC_HD uint32_t fbits(float a) { return __float_as_uint(a); }
template <class T>
VGC_HD T gld(const T* base, int idx, bool m) {
    return m ? base[idx] : T();
}

VGC_HD int lane_id() { return __lane_id(); }

VGC_HD uint64_t ballot(bool b) { return __ballot(b); }

VGC_HD int popc64(uint64_t m) { return __popcll(m); }
VGC_HD int ffs64(uint64_t m) { return m ? __builtin_ctzll(m) : -1; }

VGC_HD int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

VGC_HD bool mask_bit(uint64_t m) { return (m >> __lane_id()) & 1; }
VGC_HD bool mask_test(uint64_t m, int idx) { return (m >> (idx & 63)) & 1; }

template <class T>
VGC_HD T readlane(T a, int lane) {
    // lane is wave-uniform at every call: v_readlane into a scalar register
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "readlane of 4- or 8-byte values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a), lane));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, a);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
        return __builtin_bit_cast(T, (uint64_t(hi) << 32) | lo);
    }
}
template <class T>
VGC_HD T shfl(T a, int src) {
    return __shfl(a, src & 63);
}
template <class T>
VGC_HD T shfl_xor(T a, int m) {
    return __shfl_xor(a, m);
}

VGC_HD double wave_sum(double a) {
    double x = a;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = x + __shfl_xor(x, m);
    return x;
}
// Maxima over the wave through DPP lane moves (no LDS): quad swaps, the
// half-row and row mirrors (max over 16 lanes), then lane 15 of rows 0 / 2
// into rows 1 / 3 and lane 31 into rows 2 / 3, so lane 63 holds the maximum.
// A DPP move leaves the rows outside its row mask at their own value.  The
// maximum does not depend on the order of the reduction; every call has
// the whole wave active.
template <int CTRL, int ROWS>
VGC_HD uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
VGC_HD double dpp_f64(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = dpp_u32<CTRL, ROWS>((uint32_t)u);
    const uint32_t hi = dpp_u32<CTRL, ROWS>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
template <int CTRL, int ROWS>
VGC_HD float dpp_f32(float x) {
    return __builtin_bit_cast(float, dpp_u32<CTRL, ROWS>(__builtin_bit_cast(uint32_t, x)));
}
VGC_HD double wave_max(double a) {
    double x = a;
    x = __builtin_fmax(x, dpp_f64<0xB1, 0xF>(x));   // quad_perm [1, 0, 3, 2]
    x = __builtin_fmax(x, dpp_f64<0x4E, 0xF>(x));   // quad_perm [2, 3, 0, 1]
    x = __builtin_fmax(x, dpp_f64<0x141, 0xF>(x));  // row_half_mirror
    x = __builtin_fmax(x, dpp_f64<0x140, 0xF>(x));  // row_mirror
    x = __builtin_fmax(x, dpp_f64<0x142, 0xA>(x));  // row_bcast:15 into rows 1, 3
    x = __builtin_fmax(x, dpp_f64<0x143, 0xC>(x));  // row_bcast:31 into rows 2, 3
    return readlane(x, 63);
}
VGC_HD float wave_maxf(float a) {
    float x = a;
    x = __builtin_fmaxf(x, dpp_f32<0xB1, 0xF>(x));
    x = __builtin_fmaxf(x, dpp_f32<0x4E, 0xF>(x));
    x = __builtin_fmaxf(x, dpp_f32<0x141, 0xF>(x));
    x = __builtin_fmaxf(x, dpp_f32<0x140, 0xF>(x));
    x = __builtin_fmaxf(x, dpp_f32<0x142, 0xA>(x));
    x = __builtin_fmaxf(x, dpp_f32<0x143, 0xC>(x));
    return readlane(x, 63);
}

template <class T>
VGC_HD T lds_ld(const T* p, int idx) {
    return p[idx];
}
template <class T>
VGC_HD void lds_st(T* p, int idx, T val, bool m) {
    if (m) p[idx] = val;
}
VGC_HD void lds_min(int* p, int idx, int val, bool m) {
    if (m) atomicMin(p + idx, val);
}

// Orders the wave's LDS traffic across lanes: the compiler may not move
// memory operations across it, and every lane's earlier LDS accesses complete
// before its later ones.
VGC_HD void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

VGC_HD bool any(bool b) { return __ballot(b) != 0; }

// a store of a wave-uniform value, done by lane 0 only
VGC_HD bool lane0() { return __lane_id() == 0; }

VGC_HD void atomic_add_u64(unsigned long long* p, unsigned long long v) { atomicAdd(p, v); }
// a flag shared by the cells of a pass: read (device-coherent), and raised
// with a vector atomic
VGC_HD unsigned flag_load(const unsigned* p) {
    return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
VGC_HD void flag_raise(unsigned* p) { atomicOr(p, 1u); }

}  // namespace vgc

#endif

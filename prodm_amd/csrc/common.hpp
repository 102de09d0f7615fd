// common.hpp -- shared definitions for the gfx950 MDR kernels and the C ABI.
//
// Everything here is MI355X-specific: wave64 ballots, 256-thread workgroups,
// and FP64 correction arithmetic. Host/device inline helpers mirror the level
// geometry of ProDM's RefactorUtils.hpp:21-59 and the coarse-first layout of
// the MGARD decomposition (coarse nodes of each dim first, then the odd ones).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace prodm {

constexpr int kWave = 64;

// ---- error reporting (thread-local last error, status codes of the C ABI) --
enum Status : int {
    kOk = 0,
    kErrArg = -1,
    kErrGeometry = -2,
    kErrHip = -3,
    kErrNoMem = -4,
    kErrState = -5,
    kErrUnsupported = -6,
};

void set_error(const std::string &msg);
const std::string &last_error();

#define PRODM_HIP(call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::prodm::set_error(std::string(#call " failed: ") + hipGetErrorString(e_));   \
            return ::prodm::kErrHip;                                                      \
        }                                                                                 \
    } while (0)

#define PRODM_CHECK_LAUNCH()                                                              \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess) {                                                           \
            ::prodm::set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_)); \
            return ::prodm::kErrHip;                                                      \
        }                                                                                 \
    } while (0)

// ---- type traits ----------------------------------------------------------
template <class T> struct Traits;
template <> struct Traits<float> {
    using Stream = uint32_t;   // plane word
    using Fixed = uint32_t;    // unsigned fixed point (GroupedBP magnitude)
    using SFixed = int32_t;    // signed fixed point (NegaBinary)
    using Bits = uint32_t;
    static constexpr int kBlock = 32;
    static constexpr int kPrec = 23;
    static constexpr uint32_t kNbMask = 0xaaaaaaaau;
};
template <> struct Traits<double> {
    using Stream = uint64_t;
    using Fixed = uint64_t;
    using SFixed = int64_t;
    using Bits = uint64_t;
    static constexpr int kBlock = 64;
    static constexpr int kPrec = 52;
    static constexpr uint64_t kNbMask = 0xaaaaaaaaaaaaaaaaull;
};

// ---- geometry ---------------------------------------------------------------
// A level step works on a fine box n[3] whose coarse box is c[3]; dims are
// padded to 3 with leading extent-1 ("inactive") dims. Active dims are odd.
struct Box3 {
    uint32_t n[3];
    uint32_t c[3];
};

__host__ __device__ inline uint32_t coarse_of(uint32_t n) { return n == 1 ? 1u : (n >> 1) + 1u; }

// coarse-first position of fine index i along one dim with c coarse nodes
__host__ __device__ inline uint32_t rpos(uint32_t i, uint32_t c) { return (i & 1u) ? c + (i >> 1) : (i >> 1); }

// Index of reordered position (p0,p1,p2) inside the level buffer that
// DirectInterleaver::interleave (DirectInterleaver.hpp:12-37) produces for
// box n minus box c, in row-major order. c = {1,1,1}-padded coarse dims; for
// level 0 pass c = {0,0,0}.
__host__ __device__ inline uint64_t ilv_index(uint32_t p0, uint32_t p1, uint32_t p2, const uint32_t *n,
                                              const uint32_t *c) {
    const uint64_t n12 = (uint64_t)n[1] * n[2];
    const uint64_t plane_in = n12 - (uint64_t)c[1] * c[2];
    if (p0 < c[0]) {
        uint64_t base = (uint64_t)p0 * plane_in;
        if (p1 < c[1]) return base + (uint64_t)p1 * (n[2] - c[2]) + (p2 - c[2]);
        return base + (uint64_t)c[1] * (n[2] - c[2]) + (uint64_t)(p1 - c[1]) * n[2] + p2;
    }
    return (uint64_t)c[0] * plane_in + (uint64_t)(p0 - c[0]) * n12 + (uint64_t)p1 * n[2] + p2;
}

// host-side level geometry (RefactorUtils.hpp:21-59), padded to 3 dims
struct LevelGeom {
    int nd = 0;
    int L = 0;
    uint32_t dims[3] = {1, 1, 1};
    uint32_t ld[33][3];      // level dims, padded
    uint64_t elems[33];      // level elements
    uint64_t points[33];     // points of the level grid
    bool ok = false;
};
LevelGeom make_geom(const uint32_t *dims, int nd, int L);

// ---- device helpers ----------------------------------------------------------
__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }
__device__ inline uint32_t lane_id() { return __lane_id(); }
__device__ inline uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ inline uint32_t readlane_u32(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ inline uint64_t readlane_u64(uint64_t v, uint32_t l) {
    uint32_t lo = readlane_u32((uint32_t)v, l), hi = readlane_u32((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
// Orders one wave's LDS stores before its later LDS loads for the compiler;
// the hardware completes a wave's LDS operations in order, so lanes then read
// what other lanes of the wave wrote (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ inline double readlane_f64(double v, uint32_t l) {
    return __builtin_bit_cast(double, readlane_u64(__builtin_bit_cast(uint64_t, v), l));
}

__device__ inline int dev_frexp_exp(float x) { int e = 0; frexpf(x, &e); return e; }
__device__ inline int dev_frexp_exp(double x) { int e = 0; frexp(x, &e); return e; }
__device__ inline float dev_ldexp(float x, int e) { return ldexpf(x, e); }
__device__ inline double dev_ldexp(double x, int e) { return ldexp(x, e); }

// |x| as an order-preserving unsigned key for atomicMax (NaN excluded by caller)
__device__ inline uint32_t abs_key(float x) { return __float_as_uint(x) & 0x7fffffffu; }
__device__ inline unsigned long long abs_key(double x) {
    return (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
}
template <class T> __host__ __device__ inline T key_to_value(uint64_t k);
template <> __host__ __device__ inline float key_to_value<float>(uint64_t k) {
    uint32_t b = (uint32_t)k; float f; __builtin_memcpy(&f, &b, 4); return f;
}
template <> __host__ __device__ inline double key_to_value<double>(uint64_t k) {
    double d; __builtin_memcpy(&d, &k, 8); return d;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Global address space qualifier: accesses through a pointer read from
// memory (not a kernel argument) are otherwise generic (flat_*); casting to
// a PRODM_GLOBAL pointer makes them global_* (scalar base + 32-bit offset).
#define PRODM_GLOBAL __attribute__((address_space(1)))
// Constant address space: global memory that no wave writes while the
// kernel runs (wave-uniform loads from it are scalar loads).
#define PRODM_CONST __attribute__((address_space(4)))

// Raw buffer accesses: a wave-uniform base in a buffer descriptor plus a
// 32-bit lane byte offset. A lane whose offset is kOffNone lies outside the
// descriptor's range, so the hardware returns 0 for its load and drops its
// store: masked lanes need no exec-mask branches. A descriptor made with
// on = false has range 0 and masks the whole wave. (Descriptor word 3 =
// 32-bit data format, the gfx9 setting for untyped raw accesses.)
constexpr uint32_t kOffNone = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, bool on) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, on ? 0x7FFFFFFF : 0, 0x00020000);
}
template <class T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    }
}
template <class T> __device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, T v) {
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0, 0);
    } else {
        using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, (int)off, 0, 0);
    }
}

// A parameter struct at a wave-uniform address, copied into scalar
// registers word by word (every lane reads the same words).
template <class S> __device__ __forceinline__ S uniform_load(const S *p) {
    static_assert(sizeof(S) % 4 == 0, "parameter structs are whole words");
    constexpr int kW = sizeof(S) / 4;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(p);
    uint32_t w[kW];
#pragma unroll
    for (int i = 0; i < kW; i++) w[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)src[i]);
    S out;
    __builtin_memcpy(&out, w, sizeof(S));
    return out;
}

// Entry of a grouped launch that owns virtual workgroup vb: entry e owns
// [first[e], first[e + 1]), first[0] = 0, first[n] = total > vb.
__device__ __forceinline__ uint32_t group_entry(const uint32_t *__restrict__ first, uint32_t n, uint32_t vb) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= vb) lo = mid;
        else hi = mid;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

}  // namespace prodm

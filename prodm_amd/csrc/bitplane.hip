// bitplane.hip -- bitplane encoders/decoders and max-abs for gfx950.
//
// Layout: one lane per bitplane block (32 elements for u32 words, 64 for
// u64). A workgroup stages its blocks through LDS with coalesced global
// accesses (rows padded by one word so that each lane's block is read or
// written without bank conflicts), and each lane turns its block into plane
// words with an in-register bit-matrix transpose: after transpose32, word k
// holds bit k of every element, LSB = first element of the block
// (GroupedBPEncoder.hpp:345, NegaBinaryBPEncoder.hpp:246). Plane stores and
// loads are then one coalesced word per lane per plane.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "internal.hpp"

namespace prodm {

constexpr int kThreads = 256;

// ------------------------------------------------------------------------------
// fills
// ------------------------------------------------------------------------------
template <class U>
__global__ __launch_bounds__(kThreads) void k_fill(U *dst, uint64_t count, U value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

int launch_fill_u32(uint32_t *dst, uint64_t count, uint32_t value, hipStream_t s) {
    if (!count) return kOk;
    uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(count, kThreads), 8192);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid), dim3(kThreads), 0, s, dst, count, value);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
int launch_fill_u64(uint64_t *dst, uint64_t count, uint64_t value, hipStream_t s) {
    if (!count) return kOk;
    uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(count, kThreads), 8192);
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid), dim3(kThreads), 0, s, dst, count, value);
    PRODM_CHECK_LAUNCH();
    return kOk;
}

// ------------------------------------------------------------------------------
// max-abs: NaN never compares greater (RefactorUtils.hpp:71-73)
// ------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kThreads) void k_maxabs(const T *__restrict__ x, uint64_t n, KeyOf<T> *key) {
    using K = KeyOf<T>;
    K m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        T v = x[i];
        if (v == v) { K k = abs_key(v); m = k > m ? k : m; }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        K o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if (lane_id() == 0 && m) atomicMax(key, m);
}

template <class T> int launch_maxabs(const T *x, uint64_t n, KeyOf<T> *key, hipStream_t s) {
    if (!n) return kOk;
    uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(n, kThreads * 8ull), 4096);
    hipLaunchKernelGGL(k_maxabs<T>, dim3(grid), dim3(kThreads), 0, s, x, n, key);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_maxabs<float>(const float *, uint64_t, uint32_t *, hipStream_t);
template int launch_maxabs<double>(const double *, uint64_t, unsigned long long *, hipStream_t);

template <class T> __device__ inline int level_exp(int32_t exp, const KeyOf<T> *key) {
    return key ? dev_frexp_exp(key_to_value<T>((uint64_t)*key)) : exp;
}

// ------------------------------------------------------------------------------
// bit-matrix transposes: A[r] bit c  ->  A[c] bit r
// ------------------------------------------------------------------------------
__device__ __forceinline__ void transpose32(uint32_t (&A)[32]) {
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int j = 16 >> s;
        const uint32_t m = s == 0 ? 0x0000FFFFu : s == 1 ? 0x00FF00FFu : s == 2 ? 0x0F0F0F0Fu : s == 3 ? 0x33333333u : 0x55555555u;
#pragma unroll
        for (int k = 0; k < 32; k++) {
            if ((k & j) == 0) {
                const uint32_t t = ((A[k] >> j) ^ A[k + j]) & m;
                A[k] ^= t << j;
                A[k + j] ^= t;
            }
        }
    }
}

template <int OFF, int N>
__device__ __forceinline__ void transpose32_at(uint32_t (&A)[N]) {
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int j = 16 >> s;
        const uint32_t m = s == 0 ? 0x0000FFFFu : s == 1 ? 0x00FF00FFu : s == 2 ? 0x0F0F0F0Fu : s == 3 ? 0x33333333u : 0x55555555u;
#pragma unroll
        for (int k = 0; k < 32; k++) {
            if ((k & j) == 0) {
                const uint32_t t = ((A[OFF + k] >> j) ^ A[OFF + k + j]) & m;
                A[OFF + k] ^= t << j;
                A[OFF + k + j] ^= t;
            }
        }
    }
}

// 64 x 64 as four in-place 32 x 32 blocks. Element-major layout: element e
// is (hi[e] << 32) | lo[e]. Word-major layout: word w (bit w of the 64
// elements, element 0 in bit 0) is (w < 32) ? lo[w] | lo[32 + w] << 32
// : hi[w - 32] | hi[w] << 32. The transform is its own inverse.
__device__ __forceinline__ void transpose64(uint32_t (&lo)[64], uint32_t (&hi)[64]) {
    transpose32_at<0>(lo);
    transpose32_at<32>(lo);
    transpose32_at<0>(hi);
    transpose32_at<32>(hi);
}
template <int W> __device__ __forceinline__ uint64_t word64_get(const uint32_t (&lo)[64], const uint32_t (&hi)[64]) {
    if constexpr (W < 32) return ((uint64_t)lo[32 + W] << 32) | lo[W];
    else return ((uint64_t)hi[W] << 32) | hi[W - 32];
}
template <int W> __device__ __forceinline__ void word64_set(uint32_t (&lo)[64], uint32_t (&hi)[64], uint64_t v) {
    if constexpr (W < 32) { lo[W] = (uint32_t)v; lo[32 + W] = (uint32_t)(v >> 32); }
    else { hi[W - 32] = (uint32_t)v; hi[W] = (uint32_t)(v >> 32); }
}
template <int I, int N, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// ------------------------------------------------------------------------------
// block staging geometry
// ------------------------------------------------------------------------------
template <class T> struct BlkStage;
template <> struct BlkStage<float> {
    static constexpr int kTPB = 256;                // threads (= blocks) per workgroup
    static constexpr int kB = 32;                   // elements per block
    static constexpr int kPitch = kB + 1;           // LDS row pitch (words)
};
template <> struct BlkStage<double> {
    static constexpr int kTPB = 64;
    static constexpr int kB = 64;
    static constexpr int kPitch = kB + 1;
};

// coalesced load of the workgroup's TPB*B elements into LDS rows (0 past n)
template <class T, class XP = const T *>
__device__ inline void stage_in(XP __restrict__ x, uint64_t n, uint64_t e0, T *sx) {
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
    // all B loads in flight before the first LDS store
    T v[B];
#pragma unroll
    for (int i = 0; i < B; i++) {
        const uint64_t e = e0 + (uint32_t)(i * TPB) + threadIdx.x;
        v[i] = e < n ? x[e] : (T)0;
    }
#pragma unroll
    for (int i = 0; i < B; i++) {
        const uint32_t le = i * TPB + threadIdx.x;
        sx[(le / B) * PI + (le % B)] = v[i];
    }
}
template <class T>
__device__ inline void stage_out(T *__restrict__ out, uint64_t n, uint64_t e0, const T *sx) {
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
#pragma unroll 8
    for (int i = 0; i < B; i++) {
        const uint32_t le = i * TPB + threadIdx.x;
        const uint64_t e = e0 + le;
        if (e < n) out[e] = sx[(le / B) * PI + (le % B)];
    }
}

// ------------------------------------------------------------------------------
// NegaBinary
// ------------------------------------------------------------------------------
// binary2negabinary of (T_fps) v (NegaBinaryBPEncoder.hpp:43,219-224)
template <class T> __device__ inline typename Traits<T>::Fixed to_negabinary(T v) {
    using F = typename Traits<T>::Fixed;
    using SF = typename Traits<T>::SFixed;
    SF q = 0;
    if (v == v && fabs(v) < (T)9.2233720368547758e18) q = (SF)(int64_t)v;
    return ((F)q + (F)Traits<T>::kNbMask) ^ (F)Traits<T>::kNbMask;
}

// workgroup bid encodes bitplane blocks [bid TPB, (bid + 1) TPB)
template <class T>
__device__ __forceinline__ void nb_encode_blocks(const T *__restrict__ x, uint32_t n, int32_t exp, const KeyOf<T> *key,
                                                 int nb, typename Traits<T>::Stream *__restrict__ planes,
                                                 uint64_t stride, uint32_t bid) {
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
    __shared__ T sx[TPB * PI];
    const uint64_t b0 = (uint64_t)bid * TPB;
    // global-space pointers: the grouped form reads them from memory
    using GS = PRODM_GLOBAL typename Traits<T>::Stream;
    GS *gplanes = (GS *)planes;
    stage_in<T, const PRODM_GLOBAL T *>((const PRODM_GLOBAL T *)x, n, b0 * B, sx);
    __syncthreads();
    const int shift = nb - (level_exp<T>(exp, key) + 2);  // exp += 2 (:21)
    const uint64_t b = b0 + threadIdx.x;
    const uint64_t nw = ((uint64_t)n + B - 1) / B;
    const T *row = sx + threadIdx.x * PI;
    if constexpr (sizeof(T) == 4) {
        uint32_t q[32];
#pragma unroll
        for (int j = 0; j < 32; j++) q[j] = to_negabinary<T>(dev_ldexp(row[j], shift));
        transpose32(q);
        if (b < nw) {
#pragma unroll
            for (int k = 0; k < 32; k++) {
                const int p = nb - 1 - k;  // plane p holds bit nb - 1 - p
                if (p >= 0) gplanes[(uint64_t)p * stride + b] = q[k];
            }
        }
    } else {
        uint32_t lo[64], hi[64];
#pragma unroll
        for (int j = 0; j < 64; j++) {
            const uint64_t q = to_negabinary<T>(dev_ldexp(row[j], shift));
            lo[j] = (uint32_t)q;
            hi[j] = (uint32_t)(q >> 32);
        }
        transpose64(lo, hi);
        if (b < nw) {
            static_for<0, 64>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const int p = nb - 1 - k;
                if (p >= 0) gplanes[(uint64_t)p * stride + b] = word64_get<k>(lo, hi);
            });
        }
    }
}

template <class T>
__global__ __launch_bounds__(BlkStage<T>::kTPB) void k_nb_encode(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                              const KeyOf<T> *key, int nb,
                                                              typename Traits<T>::Stream *__restrict__ planes,
                                                              uint64_t stride) {
    nb_encode_blocks<T>(x, n, exp, key, nb, planes, stride, blockIdx.x);
}

// grouped form: one entry per level buffer of the batch (exponent from the
// level's max key, as the plan's refactor encodes)
template <class T> struct NbEntry {
    const T *x;
    const KeyOf<T> *key;
    typename Traits<T>::Stream *planes;
    uint64_t stride;
    uint32_t n, pad;
};
template <class T>
__global__ __launch_bounds__(BlkStage<T>::kTPB) void k_nb_encode_grouped(const NbEntry<T> *__restrict__ ps,
                                                                      const uint32_t *__restrict__ first,
                                                                      uint32_t nent, uint32_t total, int nb) {
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t e = group_entry(first, nent, vb);
        const NbEntry<T> E = uniform_load(ps + e);
        const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)first[e]);
        nb_encode_blocks<T>(E.x, E.n, 0, E.key, nb, E.planes, E.stride, vb - f0);
        __syncthreads();  // the staging rows are reused by the next workgroup's blocks
    }
}

template <class T>
int nb_batched_launch(const std::vector<EncodeRef> &levels, int nb, const GroupShape &shape, GroupLaunch &L) {
    constexpr int B = Traits<T>::kBlock, TPB = BlkStage<T>::kTPB;
    if (nb < 1 || nb > B) { set_error("batched negabinary encode: bad num_bitplanes"); return kErrArg; }
    L = GroupLaunch{};
    L.kind = kGrpNbEncode;
    L.elem = sizeof(T);
    L.nb = nb;
    L.block = TPB;
    L.shape = shape;
    std::vector<std::pair<uint32_t, NbEntry<T>>> v;
    for (const EncodeRef &r : levels) {
        if (!r.n) continue;
        const uint64_t nw = ((uint64_t)r.n + B - 1) / B;
        if (r.stride < nw || !r.x || !r.planes || !r.key) { set_error("batched negabinary encode: bad level"); return kErrArg; }
        NbEntry<T> E{};
        E.x = static_cast<const T *>(r.x);
        E.key = static_cast<const KeyOf<T> *>(r.key);
        E.planes = static_cast<typename Traits<T>::Stream *>(r.planes);
        E.stride = r.stride;
        E.n = r.n;
        v.push_back({div_up(nw, TPB), E});
    }
    std::stable_sort(v.begin(), v.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    L.first.push_back(0);
    for (auto &x : v) {
        const size_t o = L.params.size();
        L.params.resize(o + sizeof(NbEntry<T>));
        std::memcpy(L.params.data() + o, &x.second, sizeof(NbEntry<T>));
        L.first.push_back(L.first.back() + x.first);
        L.nent++;
    }
    L.total = L.first.back();
    return kOk;
}
template int nb_batched_launch<float>(const std::vector<EncodeRef> &, int, const GroupShape &, GroupLaunch &);
template int nb_batched_launch<double>(const std::vector<EncodeRef> &, int, const GroupShape &, GroupLaunch &);

struct ZeroEntry {
    uint32_t *p;
    uint32_t words, pad;
};
__global__ __launch_bounds__(kThreads) void k_zero_grouped(const ZeroEntry *__restrict__ ps,
                                                          const uint32_t *__restrict__ first, uint32_t nent,
                                                          uint32_t total) {
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t e = group_entry(first, nent, vb);
        const ZeroEntry E = uniform_load(ps + e);
        const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)first[e]);
        const uint32_t i = (vb - f0) * kThreads + threadIdx.x;
        if (i < E.words) E.p[i] = 0u;
    }
}

int zero_batched_launch(const std::vector<std::pair<void *, uint32_t>> &buffers, GroupLaunch &L) {
    L = GroupLaunch{};
    L.kind = kGrpZero;
    L.block = kThreads;
    L.first.push_back(0);
    for (const auto &b : buffers) {
        if (!b.second) continue;
        if (!b.first) { set_error("batched zero: null buffer"); return kErrArg; }
        ZeroEntry E{static_cast<uint32_t *>(b.first), b.second, 0};
        const size_t o = L.params.size();
        L.params.resize(o + sizeof(E));
        std::memcpy(L.params.data() + o, &E, sizeof(E));
        L.first.push_back(L.first.back() + div_up(b.second, kThreads));
        L.nent++;
    }
    L.total = L.first.back();
    return kOk;
}

int bitplane_group_run(const GroupLaunch &L, hipStream_t s) {
    if (L.kind != kGrpNbEncode && L.kind != kGrpZero) { set_error("grouped launch: not a bitplane kind"); return kErrArg; }
    if (!L.d_params || !L.d_first) { set_error("grouped launch: tables not uploaded"); return kErrState; }
    const uint32_t grid = L.grid();
    if (!grid || !L.nent) return kOk;
    if (L.kind == kGrpZero)
        hipLaunchKernelGGL(k_zero_grouped, dim3(grid), dim3(L.block), 0, s, static_cast<const ZeroEntry *>(L.d_params),
                           L.d_first, L.nent, L.total);
    else if (L.elem == 8)
        hipLaunchKernelGGL(k_nb_encode_grouped<double>, dim3(grid), dim3(L.block), 0, s,
                           static_cast<const NbEntry<double> *>(L.d_params), L.d_first, L.nent, L.total, L.nb);
    else
        hipLaunchKernelGGL(k_nb_encode_grouped<float>, dim3(grid), dim3(L.block), 0, s,
                           static_cast<const NbEntry<float> *>(L.d_params), L.d_first, L.nent, L.total, L.nb);
    PRODM_CHECK_LAUNCH();
    return kOk;
}

template <class T>
int launch_nb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb,
                     typename Traits<T>::Stream *planes, uint64_t stride, hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (n == 0 || nb < 1 || nb > B) { set_error("negabinary encode: bad n or num_bitplanes"); return kErrArg; }
    const uint64_t nw = ((uint64_t)n + B - 1) / B;
    if (stride < nw) { set_error("negabinary encode: plane stride below the words per plane"); return kErrArg; }
    const uint32_t grid = div_up(nw, BlkStage<T>::kTPB);
    hipLaunchKernelGGL(k_nb_encode<T>, dim3(grid), dim3(BlkStage<T>::kTPB), 0, s, x, n, exp, key, nb, planes, stride);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_nb_encode<float>(const float *, uint32_t, int32_t, const uint32_t *, int, uint32_t *, uint64_t,
                                     hipStream_t);
template int launch_nb_encode<double>(const double *, uint32_t, int32_t, const unsigned long long *, int, uint64_t *,
                                      uint64_t, hipStream_t);

template <class T>
__global__ __launch_bounds__(BlkStage<T>::kTPB) void k_nb_decode(PlanePtrs pl, uint32_t n, int32_t exp,
                                                              const KeyOf<T> *key, int start, int k,
                                                              T *__restrict__ out) {
    using S = typename Traits<T>::Stream;
    using SF = typename Traits<T>::SFixed;
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
    __shared__ T sx[TPB * PI];
    const uint64_t b0 = (uint64_t)blockIdx.x * TPB;
    const uint64_t b = b0 + threadIdx.x;
    const uint64_t nw = ((uint64_t)n + B - 1) / B;
    const int e2 = level_exp<T>(exp, key) + 2;
    const int end = start + k;
    T *row = sx + threadIdx.x * PI;
    // acc bit (k - 1 - i) comes from plane start + i
    if constexpr (sizeof(T) == 4) {
        uint32_t a[32];
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const int i = k - 1 - r;
            a[r] = (i >= 0 && b < nw) ? static_cast<const S *>(pl.p[i])[b] : 0u;
        }
        transpose32(a);
#pragma unroll
        for (int j = 0; j < 32; j++) {
            const SF val = (SF)((a[j] ^ (uint32_t)Traits<T>::kNbMask) - (uint32_t)Traits<T>::kNbMask);
            const T y = dev_ldexp((T)val, e2 - end);
            row[j] = (end & 1) ? -y : y;  // odd ending plane flips the sign (:153,172-189)
        }
    } else {
        uint32_t lo[64], hi[64];
        static_for<0, 64>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            const int i = k - 1 - r;
            const uint64_t w = (i >= 0 && b < nw) ? static_cast<const S *>(pl.p[i])[b] : 0ull;
            word64_set<r>(lo, hi, w);
        });
        transpose64(lo, hi);
#pragma unroll
        for (int j = 0; j < 64; j++) {
            const uint64_t acc = ((uint64_t)hi[j] << 32) | lo[j];
            const SF val = (SF)((acc ^ (uint64_t)Traits<T>::kNbMask) - (uint64_t)Traits<T>::kNbMask);
            const T y = dev_ldexp((T)val, e2 - end);
            row[j] = (end & 1) ? -y : y;
        }
    }
    __syncthreads();
    stage_out<T>(out, n, b0 * B, sx);
}

template <class T>
int launch_nb_decode(const PlanePtrs &planes, uint32_t n, int32_t exp, const KeyOf<T> *key, int start, int k, T *out,
                     hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (n == 0) return kOk;
    if (k < 0 || start < 0 || start + k > B) { set_error("negabinary decode: bad plane range"); return kErrArg; }
    if (k == 0) {
        // k == 0: the reference returns zeros (:135-138)
        if constexpr (sizeof(T) == 4) return launch_fill_u32(reinterpret_cast<uint32_t *>(out), n, 0u, s);
        else return launch_fill_u64(reinterpret_cast<uint64_t *>(out), n, 0ull, s);
    }
    const uint64_t nw = ((uint64_t)n + B - 1) / B;
    const uint32_t grid = div_up(nw, BlkStage<T>::kTPB);
    hipLaunchKernelGGL(k_nb_decode<T>, dim3(grid), dim3(BlkStage<T>::kTPB), 0, s, planes, n, exp, key, start, k, out);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_nb_decode<float>(const PlanePtrs &, uint32_t, int32_t, const uint32_t *, int, int, float *,
                                     hipStream_t);
template int launch_nb_decode<double>(const PlanePtrs &, uint32_t, int32_t, const unsigned long long *, int, int,
                                      double *, hipStream_t);

// ------------------------------------------------------------------------------
// GroupedBP
// ------------------------------------------------------------------------------
// magnitude of (int64_t) ldexp(x, nb - exp), sign folded (GroupedBPEncoder.hpp:40-44);
// |v| >= 2^63 is undefined in C++: compat reproduces the x86 result 2^63,
// otherwise the truncated magnitude is used (hazard B1).
template <class T>
__device__ inline typename Traits<T>::Fixed gb_mag(T x, int shift, int compat) {
    using F = typename Traits<T>::Fixed;
    const T v = dev_ldexp(x, shift);
    if constexpr (sizeof(T) == 4) {  // one conversion where the magnitude fits the word
        const float a = fabsf(v);
        if (a < 4294967296.0f) return (F)a;
    }
    if (v != v) return 0;
    const T av = fabs(v);
    if (av < (T)9.2233720368547758e18) {
        const int64_t q = (int64_t)v;
        const uint64_t m = (x < 0) ? (uint64_t)0 - (uint64_t)q : (uint64_t)q;
        return (F)m;
    }
    if (compat) return (F)0x8000000000000000ull;
    if (av < (T)1.8446744073709552e19) return (F)(uint64_t)av;
    return (F)~0ull;
}

// recording plane of a block from the OR of its magnitudes
template <class F> __device__ inline int rb_of(F orv, int nb) {
    const F mask = (nb >= (int)(8 * sizeof(F))) ? ~(F)0 : (((F)1 << nb) - 1);
    orv &= mask;
    if (!orv) return nb;
    const int msb = (int)(8 * sizeof(F)) - 1 - (sizeof(F) == 4 ? __clz((uint32_t)orv) : __clzll((uint64_t)orv));
    return nb - 1 - msb;
}

// Pass 1: rb per block and an rb histogram per workgroup (hist[wg][0..nb]).
template <class T>
__global__ __launch_bounds__(BlkStage<T>::kTPB) void k_gb_rb(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                          const KeyOf<T> *key, int nb, int compat,
                                                          uint8_t *__restrict__ rb, uint32_t *__restrict__ hist) {
    using F = typename Traits<T>::Fixed;
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
    __shared__ T sx[TPB * PI];
    __shared__ uint32_t h[65];
    for (int i = threadIdx.x; i <= nb; i += TPB) h[i] = 0;
    const uint64_t b0 = (uint64_t)blockIdx.x * TPB;
    stage_in<T>(x, n, b0 * B, sx);
    __syncthreads();
    const int shift = nb - level_exp<T>(exp, key);
    const uint64_t b = b0 + threadIdx.x;
    const uint32_t nblk = (n - 1) / B + 1;
    const T *row = sx + threadIdx.x * PI;
    F orv = 0;
#pragma unroll 8
    for (int j = 0; j < B; j++) orv |= gb_mag<T>(row[j], shift, compat);
    if (b < nblk) {
        const int r = rb_of<F>(orv, nb);
        rb[b] = (uint8_t)r;
        atomicAdd(&h[r], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= nb; i += TPB) hist[(uint64_t)blockIdx.x * (nb + 1) + i] = h[i];
}

// words per plane p of each workgroup: #(rb <= p) + #(rb == p). One
// workgroup per plane scans these over workgroups: off[wg][i] exclusive.
__global__ __launch_bounds__(1024) void k_gb_scan(const uint32_t *__restrict__ hist, uint32_t nwg, int nbins,
                                                  int pstart, int k, uint32_t *__restrict__ off,
                                                  uint32_t *__restrict__ totals) {
    __shared__ uint32_t sh[1024];
    const int i = blockIdx.x;  // plane index within [pstart, pstart + k)
    const int p = pstart + i;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nwg; c0 += 1024) {
        const uint32_t wg = c0 + threadIdx.x;
        uint32_t v = 0;
        if (wg < nwg) {
            const uint32_t *hh = hist + (uint64_t)wg * nbins;
            uint32_t le = 0;
            for (int r = 0; r <= p && r < nbins; r++) le += hh[r];
            v = le + (p < nbins ? hh[p] : 0);
        }
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            uint32_t t = threadIdx.x >= (uint32_t)o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (wg < nwg) off[(uint64_t)wg * k + i] = carry + sh[threadIdx.x] - v;
        carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[i] = carry;
}

// in-workgroup exclusive position of this lane's words for plane p, given
// its counts (c1 = rb <= p, c2 = rb == p): wave prefix by ballot + mbcnt,
// cross-wave prefix through LDS. All threads must call it (barriers).
template <int TPB>
__device__ inline uint32_t wg_prefix(bool c1, bool c2, uint32_t *wtot) {
    const uint32_t wv = threadIdx.x / 64;
    const uint64_t m1 = ballot64(c1), m2 = ballot64(c2);
    if (lane_id() == 0) wtot[wv] = __popcll(m1) + __popcll(m2);
    __syncthreads();
    uint32_t pre = lanes_below(m1) + lanes_below(m2);
    for (uint32_t q = 0; q < wv; q++) pre += wtot[q];
    __syncthreads();
    return pre;
}

// Pass 2: plane words, compacted per plane.
template <class T>
__global__ __launch_bounds__(BlkStage<T>::kTPB) void k_gb_write(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                             const KeyOf<T> *key, int nb, int compat,
                                                             const uint32_t *__restrict__ off,
                                                             typename Traits<T>::Stream *__restrict__ planes,
                                                             uint64_t stride) {
    using F = typename Traits<T>::Fixed;
    using S = typename Traits<T>::Stream;
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
    __shared__ T sx[TPB * PI];
    __shared__ uint32_t wtot[TPB / 64];
    const uint64_t b0 = (uint64_t)blockIdx.x * TPB;
    stage_in<T>(x, n, b0 * B, sx);
    __syncthreads();
    const int shift = nb - level_exp<T>(exp, key);
    const uint64_t b = b0 + threadIdx.x;
    const uint32_t nblk = (n - 1) / B + 1;
    const T *row = sx + threadIdx.x * PI;
    S sg = 0;
    F orv = 0;
    int r;
    if constexpr (sizeof(T) == 4) {
        uint32_t m[32];
#pragma unroll
        for (int j = 0; j < 32; j++) {
            m[j] = gb_mag<T>(row[j], shift, compat);
            orv |= m[j];
            sg |= (S)(row[j] < 0) << j;
        }
        r = b < nblk ? rb_of<F>(orv, nb) : 255;
        transpose32(m);
#pragma unroll
        for (int k = 31; k >= 0; k--) {
            const int p = nb - 1 - k;  // plane p holds bit nb - 1 - p
            if (p >= 0) {
                const uint32_t pre = wg_prefix<TPB>(r <= p, r == p, wtot);
                uint64_t pos = off[(uint64_t)blockIdx.x * nb + p] + pre;
                S *dst = planes + (uint64_t)p * stride;
                if (r == p) dst[pos++] = sg;
                if (r <= p) dst[pos] = m[k];
            }
        }
    } else {
        uint32_t lo[64], hi[64];
#pragma unroll
        for (int j = 0; j < 64; j++) {
            const uint64_t q = gb_mag<T>(row[j], shift, compat);
            orv |= q;
            lo[j] = (uint32_t)q;
            hi[j] = (uint32_t)(q >> 32);
            sg |= (S)(row[j] < 0) << j;
        }
        r = b < nblk ? rb_of<F>(orv, nb) : 255;
        transpose64(lo, hi);
        static_for<0, 64>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int p = nb - 1 - k;
            if (p >= 0) {
                const uint32_t pre = wg_prefix<TPB>(r <= p, r == p, wtot);
                uint64_t pos = off[(uint64_t)blockIdx.x * nb + p] + pre;
                S *dst = planes + (uint64_t)p * stride;
                if (r == p) dst[pos++] = sg;
                if (r <= p) dst[pos] = word64_get<k>(lo, hi);
            }
        });
    }
}

template <class T> static uint32_t gb_nwg(uint32_t n) {
    const uint32_t nblk = (n - 1) / Traits<T>::kBlock + 1;
    return div_up(nblk, BlkStage<T>::kTPB);
}

template <class T> size_t gb_workspace_bytes(uint32_t n, int nb) {
    const uint32_t nwg = gb_nwg<T>(n);
    return (size_t)nwg * (nb + 1) * 4 + (size_t)nwg * nb * 4 + 256;
}
template size_t gb_workspace_bytes<float>(uint32_t, int);
template size_t gb_workspace_bytes<double>(uint32_t, int);

template <class T>
int launch_gb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb, int compat, uint8_t *rb,
                     typename Traits<T>::Stream *planes, uint64_t stride, uint32_t *words_per_plane, void *ws,
                     size_t ws_bytes, hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (n == 0 || nb < 1 || nb > B) { set_error("grouped encode: bad n or num_bitplanes"); return kErrArg; }
    const uint32_t nblk = (n - 1) / B + 1;
    if (stride < 2ull * nblk) { set_error("grouped encode: plane stride below 2 * nblocks"); return kErrArg; }
    if (ws_bytes < gb_workspace_bytes<T>(n, nb)) { set_error("grouped encode: workspace too small"); return kErrArg; }
    const uint32_t nwg = gb_nwg<T>(n);
    uint32_t *hist = static_cast<uint32_t *>(ws);
    uint32_t *off = hist + (size_t)nwg * (nb + 1);
    hipLaunchKernelGGL(k_gb_rb<T>, dim3(nwg), dim3(BlkStage<T>::kTPB), 0, s, x, n, exp, key, nb, compat, rb, hist);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gb_scan, dim3(nb), dim3(1024), 0, s, hist, nwg, nb + 1, 0, nb, off, words_per_plane);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gb_write<T>, dim3(nwg), dim3(BlkStage<T>::kTPB), 0, s, x, n, exp, key, nb, compat, off, planes,
                       stride);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_gb_encode<float>(const float *, uint32_t, int32_t, const uint32_t *, int, int, uint8_t *,
                                     uint32_t *, uint64_t, uint32_t *, void *, size_t, hipStream_t);
template int launch_gb_encode<double>(const double *, uint32_t, int32_t, const unsigned long long *, int, int,
                                      uint8_t *, uint64_t *, uint64_t, uint32_t *, void *, size_t, hipStream_t);

// ---- GroupedBP progressive decode ----------------------------------------------
// rb histogram per workgroup (same block ranges as the encoder)
template <class T>
__global__ __launch_bounds__(256) void k_gb_rbhist(const uint8_t *__restrict__ rb, uint32_t nblk, int nbins,
                                                   uint32_t *__restrict__ hist) {
    constexpr int TPB = BlkStage<T>::kTPB;
    __shared__ uint32_t h[256];
    for (int i = threadIdx.x; i < nbins; i += 256) h[i] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * TPB;
    for (uint32_t i = threadIdx.x; i < (uint32_t)TPB; i += 256) {
        const uint64_t b = b0 + i;
        if (b < nblk) {
            const int r = rb[b];
            if (r < nbins) atomicAdd(&h[r], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += 256) hist[(uint64_t)blockIdx.x * nbins + i] = h[i];
}

template <class S> __device__ inline S load_word(const uint8_t *base, uint64_t byte_off, bool aligned) {
    if (aligned) return *reinterpret_cast<const S *>(base + byte_off);
    S v = 0;
#pragma unroll
    for (int i = 0; i < (int)sizeof(S); i++) v |= (S)base[byte_off + i] << (8 * i);
    return v;
}

template <class T>
__global__ __launch_bounds__(BlkStage<T>::kTPB) void k_gb_decode(PlanePtrs pl, uint32_t n, int32_t exp,
                                                              const KeyOf<T> *key, int start, int k,
                                                              const uint8_t *__restrict__ rb,
                                                              typename Traits<T>::Stream *__restrict__ signs,
                                                              const uint32_t *__restrict__ off, T *__restrict__ out,
                                                              uint64_t hdr0) {
    using S = typename Traits<T>::Stream;
    constexpr int TPB = BlkStage<T>::kTPB, B = BlkStage<T>::kB, PI = BlkStage<T>::kPitch;
    __shared__ T sx[TPB * PI];
    __shared__ uint32_t wtot[TPB / 64];
    const uint64_t b0 = (uint64_t)blockIdx.x * TPB;
    const uint64_t b = b0 + threadIdx.x;
    const uint32_t nblk = (n - 1) / B + 1;
    const int e0 = level_exp<T>(exp, key);
    const int end = start + k;
    const bool valid = b < nblk;
    const int r = valid ? rb[b] : 255;
    S sg = valid ? signs[b] : 0;
    // acc bit (end - 1 - p) comes from plane p; planes below rb are zero
    constexpr int W = 8 * sizeof(S);
    uint32_t lo[W], hi[sizeof(T) == 8 ? W : 1];
#pragma unroll
    for (int q = 0; q < W; q++) {
        lo[q] = 0;
        if constexpr (sizeof(T) == 8) hi[q] = 0;
    }
    static_for<0, W>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int p = end - 1 - q;  // acc bit q <- plane p
        if (p >= start) {
            const int i = p - start;
            const uint32_t pre = wg_prefix<TPB>(r <= p, r == p, wtot);
            uint64_t pos = off[(uint64_t)blockIdx.x * k + i] + pre;
            const uint8_t *sb = static_cast<const uint8_t *>(pl.p[i]);
            const uint64_t hb = (i == 0) ? hdr0 : 0;
            const bool al = (hb % sizeof(S)) == 0;
            if (r == p) { sg = load_word<S>(sb, hb + pos * sizeof(S), al); pos++; }
            if (r <= p) {
                const S w = load_word<S>(sb, hb + pos * sizeof(S), al);
                if constexpr (sizeof(T) == 4) lo[q] = (uint32_t)w;
                else word64_set<q>(lo, hi, (uint64_t)w);
            }
        }
    });
    // persist the signs of blocks recorded in this window
    if (valid && r >= start && r < end) signs[b] = sg;
    T *row = sx + threadIdx.x * PI;
    if constexpr (sizeof(T) == 4) {
        transpose32(lo);
#pragma unroll
        for (int j = 0; j < 32; j++) {
            T y = dev_ldexp((T)lo[j], e0 - end);
            if ((sg >> j) & 1u) y = -y;
            row[j] = (r >= end) ? (T)0 : y;
        }
    } else {
        transpose64(lo, hi);
#pragma unroll
        for (int j = 0; j < 64; j++) {
            T y = dev_ldexp((T)(((uint64_t)hi[j] << 32) | lo[j]), e0 - end);
            if ((sg >> j) & 1u) y = -y;
            row[j] = (r >= end) ? (T)0 : y;
        }
    }
    __syncthreads();
    stage_out<T>(out, n, b0 * B, sx);
}

template <class T>
int launch_gb_decode(const PlanePtrs &streams, uint32_t n, int32_t exp, const KeyOf<T> *key, int start, int k,
                     const uint8_t *rb, typename Traits<T>::Stream *signs, T *out, void *ws, size_t ws_bytes,
                     uint64_t hdr0, hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (n == 0) return kOk;
    if (k < 0 || start < 0 || start + k > B) { set_error("grouped decode: bad plane range"); return kErrArg; }
    if (k == 0) {
        if constexpr (sizeof(T) == 4) return launch_fill_u32(reinterpret_cast<uint32_t *>(out), n, 0u, s);
        else return launch_fill_u64(reinterpret_cast<uint64_t *>(out), n, 0ull, s);
    }
    const uint32_t nblk = (n - 1) / B + 1;
    const uint32_t nwg = gb_nwg<T>(n);
    const int nbins = start + k;  // rb values >= end never contribute
    if (ws_bytes < (size_t)nwg * (nbins + k) * 4 + (size_t)k * 4) {
        set_error("grouped decode: workspace too small");
        return kErrArg;
    }
    uint32_t *hist = static_cast<uint32_t *>(ws);
    uint32_t *off = hist + (size_t)nwg * nbins;
    uint32_t *totals = off + (size_t)nwg * k;
    hipLaunchKernelGGL(k_gb_rbhist<T>, dim3(nwg), dim3(256), 0, s, rb, nblk, nbins, hist);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gb_scan, dim3(k), dim3(1024), 0, s, hist, nwg, nbins, start, k, off, totals);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gb_decode<T>, dim3(nwg), dim3(BlkStage<T>::kTPB), 0, s, streams, n, exp, key, start, k, rb, signs,
                       off, out, hdr0);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_gb_decode<float>(const PlanePtrs &, uint32_t, int32_t, const uint32_t *, int, int, const uint8_t *,
                                     uint32_t *, float *, void *, size_t, uint64_t, hipStream_t);
template int launch_gb_decode<double>(const PlanePtrs &, uint32_t, int32_t, const unsigned long long *, int, int,
                                      const uint8_t *, uint64_t *, double *, void *, size_t, uint64_t, hipStream_t);

template <class T> size_t gb_decode_workspace_bytes(uint32_t n, int start, int k) {
    const uint32_t nwg = gb_nwg<T>(n);
    return (size_t)nwg * (start + 2 * k) * 4 + 64 * 4 + 256;
}
template size_t gb_decode_workspace_bytes<float>(uint32_t, int, int);
template size_t gb_decode_workspace_bytes<double>(uint32_t, int, int);

// ------------------------------------------------------------------------------
// PerBit
// ------------------------------------------------------------------------------
// Stream p carries bit nb - 1 - p of every element's magnitude, element by
// element, and right after an element's first 1 bit (over planes 0..nb-1)
// its sign; bits fill 64-bit words LSB first (PerBitBPEncoder.hpp:9-39,
// 96-143). Encode runs in three passes over runs of 64 lanes x B
// consecutive elements (one wave): count (first-1 flags per plane), scan
// (bit offset of every wave's run in every stream) and write (each lane
// expands its plane word with its sign bits; the wave assembles its run
// in LDS and stores whole words).

// T_fp of the reference, sign ? -(int64_t) v : (int64_t) v with
// v = ldexp(x, nb - exp) (PerBitBPEncoder.hpp:101-104). The conversion of
// |v| >= 2^63 or NaN is undefined in C++: compat gives the x86 result
// (INT64_MIN, which is its own negation), otherwise the truncated
// magnitude is used, as for GroupedBP (hazard B1).
template <class T>
__device__ inline typename Traits<T>::Fixed pb_mag(T x, int shift, int compat) {
    using F = typename Traits<T>::Fixed;
    if (!compat) return gb_mag<T>(x, shift, 0);
    const T v = dev_ldexp(x, shift);
    if constexpr (sizeof(T) == 4) {  // trunc(|v|): the low word of either sign's conversion
        const float a = fabsf(v);
        if (a < 4294967296.0f) return (F)a;
    }
    const int64_t q = fabs(v) < (T)9.2233720368547758e18 ? (int64_t)v : INT64_MIN;
    return (F)(x < 0 ? (uint64_t)0 - (uint64_t)q : (uint64_t)q);
}

// wave-wide exclusive prefix sum and total (DPP reductions of the device library)
extern "C" __device__ uint32_t __ockl_wfscan_add_u32(uint32_t, bool);
extern "C" __device__ uint32_t __ockl_wfred_add_u32(uint32_t);

// Every lane takes 32 consecutive elements (f32 and f64), a wave 2048, a
// workgroup of 4 waves 8192; a lane's run in a stream is at most 64 bits.
constexpr int kPbTPB = 256, kPbB = 32;

// The lane's 32 elements straight from global memory (16-byte loads when
// the row is whole and aligned) as plane words kept in registers: w[k] bit
// j = bit k of element j's magnitude (k < 32 for f32, < 64 for f64); sg
// bit j = element j < 0.
template <class T> struct PbLane {
    static constexpr int kWords = 8 * sizeof(T);
    uint32_t w[kWords];
    uint32_t sg;
    __device__ __forceinline__ void load(const T *__restrict__ x, uint64_t e, uint32_t n, int shift, int compat, bool vec) {
        T v[kPbB];
        if (vec && e + kPbB <= n) {
            if constexpr (sizeof(T) == 4) {
                const float4 *q = reinterpret_cast<const float4 *>(x + e);
#pragma unroll
                for (int i = 0; i < kPbB / 4; i++) {
                    const float4 t = q[i];
                    v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
                }
            } else {
                const double2 *q = reinterpret_cast<const double2 *>(x + e);
#pragma unroll
                for (int i = 0; i < kPbB / 2; i++) {
                    const double2 t = q[i];
                    v[2 * i] = t.x; v[2 * i + 1] = t.y;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPbB; j++) v[j] = e + j < n ? x[e + j] : (T)0;
        }
        sg = 0;
#pragma unroll
        for (int j = 0; j < kPbB; j++) {
            const uint64_t q = pb_mag<T>(v[j], shift, compat);
            w[j] = (uint32_t)q;
            if constexpr (sizeof(T) == 8) w[32 + j] = (uint32_t)(q >> 32);
            sg |= (uint32_t)(v[j] < 0) << j;
        }
        transpose32_at<0>(w);
        if constexpr (sizeof(T) == 8) transpose32_at<32>(w);
    }
};
// f(p, D, F) for the planes p = 0..nb-1 in order: D = the plane word (bit
// nb - 1 - p of the lane's elements), F = the elements whose first 1 bit is
// in plane p. Unrolled, so that the words stay in registers.
template <class T, class Fn> __device__ __forceinline__ void pb_for_planes(const PbLane<T> &P, int nb, Fn &&f) {
    constexpr int W = PbLane<T>::kWords;
    uint32_t H = 0;
    static_for<0, W>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = W - 1 - decltype(kc)::value;
        const int p = nb - 1 - k;
        if (p >= 0) {
            const uint32_t D = P.w[k];
            const uint32_t F = D & ~H;
            H |= D;
            f(p, D, F);
        }
    });
}

// The first-1 counts of planes 2i and 2i + 1 packed in one word (low and
// high 16 bits: a lane counts at most 32 per plane, a wave at most 2048), so
// that one wave-wide scan or sum serves two planes. f(i, packed count) for
// the pairs i in plane order; planes p < 0 (nb below the word count) count 0.
template <class T, class Fn> __device__ __forceinline__ void pb_for_plane_pairs(const PbLane<T> &P, int nb, Fn &&f) {
    constexpr int W = PbLane<T>::kWords;
    uint32_t H = 0, acc = 0;
    static_for<0, W>([&](auto kc) __attribute__((always_inline)) {
        constexpr int i = decltype(kc)::value;
        constexpr int k = W - 1 - i;
        const bool on = nb - 1 - k >= 0;
        const uint32_t D = on ? P.w[k] : 0u;
        const uint32_t c = (uint32_t)__popc(D & ~H);
        H |= D;
        if constexpr (i % 2 == 0) acc = c;
        else f(i / 2, acc | (c << 16));
    });
}

// Runs are cut per wave (64 lanes x 32 consecutive elements).
template <class T> static uint32_t pb_nwg(uint32_t n) { return div_up(n, (uint64_t)kPbTPB * kPbB); }
template <class T> static uint32_t pb_nwaves(uint32_t n) { return pb_nwg<T>(n) * (kPbTPB / 64); }

// valid elements of wave gw
template <class T> __device__ inline uint32_t pb_wave_valid(uint32_t n, uint32_t gw) {
    constexpr uint64_t E = 64ull * kPbB;
    const uint64_t e = (uint64_t)gw * E;
    return e < n ? (uint32_t)min(E, (uint64_t)n - e) : 0u;
}

// Pass 1: bits[p][gw] = bits of wave gw's run in plane p: its valid
// elements plus those whose first 1 bit is in plane p.
template <class T>
__global__ __launch_bounds__(kPbTPB) void k_pb_count(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                  const KeyOf<T> *key, int nb, int compat,
                                                  uint32_t *__restrict__ bits, uint32_t nwv) {
    constexpr int NW = kPbTPB / 64;
    const int lane = (int)lane_id();
    const uint32_t gw = blockIdx.x * NW + threadIdx.x / 64;
    PbLane<T> P;
    P.load(x, ((uint64_t)blockIdx.x * kPbTPB + threadIdx.x) * kPbB, n, nb - level_exp<T>(exp, key), compat,
           ((uintptr_t)x & 15) == 0);
    uint32_t mine = 0;  // lane p: plane p's first-1 flags of the wave
    constexpr int W = PbLane<T>::kWords;
    pb_for_plane_pairs(P, nb, [&](int i, uint32_t pc) {
        const uint32_t s = __ockl_wfred_add_u32(pc);
        // pair i holds the planes with word indices W - 1 - 2i, W - 2 - 2i
        const int pa = nb - W + 2 * i, pb = pa + 1;
        if (lane == pa) mine = s & 0xFFFFu;
        if (lane == pb) mine = s >> 16;
    });
    if (lane < nb) bits[(uint64_t)lane * nwv + gw] = pb_wave_valid<T>(n, gw) + mine;
}

// Pass 2: exclusive bit offsets off[p][gw] of the wave runs and the stream
// sizes in bytes (whole words). Blocks of kPbScanBlk waves per plane: their
// sums, then every block's waves with the sum of the blocks below.
constexpr int kPbScanT = 1024, kPbScanPer = 4, kPbScanBlk = kPbScanT * kPbScanPer;

// sum over the workgroup (all threads call it); result on thread 0
__device__ inline uint64_t pb_block_sum(uint64_t v, uint64_t *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane_id() == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    uint64_t t = 0;
    if (threadIdx.x == 0)
        for (uint32_t i = 0; i < blockDim.x / 64; i++) t += red[i];
    return t;
}

__global__ __launch_bounds__(kPbScanT) void k_pb_scan_sum(const uint32_t *__restrict__ bits, uint32_t nwv,
                                                         uint64_t *__restrict__ sums) {
    __shared__ uint64_t red[kPbScanT / 64];
    const int p = blockIdx.y;
    const uint32_t *row = bits + (uint64_t)p * nwv;
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < kPbScanPer; i++) {
        const uint32_t w = blockIdx.x * kPbScanBlk + i * kPbScanT + threadIdx.x;
        if (w < nwv) v += row[w];
    }
    const uint64_t t = pb_block_sum(v, red);
    if (threadIdx.x == 0) sums[(uint64_t)p * gridDim.x + blockIdx.x] = t;
}

__global__ __launch_bounds__(kPbScanT) void k_pb_scan_off(const uint32_t *__restrict__ bits, uint32_t nwv,
                                                         const uint64_t *__restrict__ sums,
                                                         uint64_t *__restrict__ off, uint32_t *__restrict__ sizes) {
    __shared__ uint64_t red[kPbScanT / 64];
    __shared__ uint32_t wtot[kPbScanT / 64];
    __shared__ uint64_t base_s;
    const int p = blockIdx.y;
    const uint32_t blk = blockIdx.x, lane = lane_id(), wv = threadIdx.x / 64;
    const uint32_t *row = bits + (uint64_t)p * nwv;
    uint64_t *orow = off + (uint64_t)p * nwv;
    // bits below this block
    const uint64_t b = pb_block_sum(threadIdx.x < blk ? sums[(uint64_t)p * gridDim.x + threadIdx.x] : 0, red);
    if (threadIdx.x == 0) base_s = b;
    // this thread's waves [w0, w0 + kPbScanPer)
    const uint32_t w0 = blk * kPbScanBlk + threadIdx.x * kPbScanPer;
    uint32_t v[kPbScanPer], local = 0;
#pragma unroll
    for (int i = 0; i < kPbScanPer; i++) {
        v[i] = w0 + i < nwv ? row[w0 + i] : 0;
        local += v[i];
    }
    const uint32_t pre = __ockl_wfscan_add_u32(local, false);
    if (lane == 63) wtot[wv] = pre + local;
    __syncthreads();
    uint64_t o = base_s + pre;
    for (uint32_t i = 0; i < wv; i++) o += wtot[i];
#pragma unroll
    for (int i = 0; i < kPbScanPer; i++) {
        if (w0 + i < nwv) {
            orow[w0 + i] = o;
            o += v[i];
            if (w0 + i == nwv - 1) sizes[p] = (uint32_t)(((o + 63) >> 6) * 8);  // the stream's end
        }
    }
}

// Pass 3: the streams. Lane t of wave gw starts its run in plane p at bit
// off[p][gw] + min(t B, valid) + (first-1 flags of the lanes below). The
// wave assembles its run in its own LDS words (in wave order, no
// barriers) and stores them.
template <class T>
__global__ __launch_bounds__(kPbTPB) void k_pb_write(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                  const KeyOf<T> *key, int nb, int compat,
                                                  const uint64_t *__restrict__ off, uint32_t nwv,
                                                  uint64_t *__restrict__ streams, uint64_t stride,
                                                  uint64_t *__restrict__ side) {
    constexpr int B = kPbB, NW = kPbTPB / 64;
    constexpr int kRunWords = (2 * 64 * B + 63) / 64 + 2;  // a wave's run plus its start offset
    __shared__ unsigned long long run[NW][kRunWords];
    const int lane = (int)lane_id();
    const uint32_t wv = threadIdx.x / 64, gw = blockIdx.x * NW + wv;
    unsigned long long *buf = run[wv];
    for (int i = lane; i < kRunWords; i += 64) buf[i] = 0;
    const uint64_t my_off = lane < nb ? off[(uint64_t)lane * nwv + gw] : 0;  // lane p: plane p's offset
    const uint32_t nv = pb_wave_valid<T>(n, gw);
    const uint32_t below = min((uint32_t)lane * B, nv);  // valid elements of the lanes below
    PbLane<T> P;
    P.load(x, ((uint64_t)blockIdx.x * kPbTPB + threadIdx.x) * B, n, nb - level_exp<T>(exp, key), compat,
           ((uintptr_t)x & 15) == 0);
    const uint32_t S = P.sg;  // sign bits of the lane's elements
    // first-1 flags of the lanes below and of the wave: for f32 two planes
    // per scan, computed before the plane loop (16 words); f64's 64 planes
    // would hold 32 of them, so it scans per plane
    constexpr int W = PbLane<T>::kWords;
    constexpr bool kPairs = sizeof(T) == 4;
    uint32_t pre2[kPairs ? W / 2 : 1], tot2[kPairs ? W / 2 : 1];
    if constexpr (kPairs)
        pb_for_plane_pairs(P, nb, [&](int i, uint32_t pc) __attribute__((always_inline)) {
            pre2[i] = __ockl_wfscan_add_u32(pc, false);
            tot2[i] = readlane_u32(pre2[i] + pc, 63);
        });
    pb_for_planes(P, nb, [&](int p, uint32_t D, uint32_t F) __attribute__((always_inline)) {
        const uint64_t o = ((uint64_t)readlane_u32((uint32_t)(my_off >> 32), p) << 32) |
                           readlane_u32((uint32_t)my_off, p);
        uint32_t pre, flags;
        if constexpr (kPairs) {
            // p = nb - W + i for the plane order index i, in pair i / 2
            const int i = p - nb + W;
            const uint32_t sh16 = (i & 1) ? 16u : 0u;
            pre = (pre2[i >> 1] >> sh16) & 0xFFFFu;
            flags = (tot2[i >> 1] >> sh16) & 0xFFFFu;  // the wave's total
        } else {
            const uint32_t c = (uint32_t)__popc(F);
            pre = __ockl_wfscan_add_u32(c, false);
            flags = readlane_u32(pre + c, 63);
        }
        const uint32_t q = (uint32_t)(o & 63) + below + pre;
        // insert each flagged element's sign after its bit, highest first
        uint64_t r = D;
        uint32_t ff = F;
        while (ff) {
            const int j = 31 - __clz(ff);
            ff ^= 1u << j;
            const uint64_t low = (2ull << j) - 1;
            r = (r & low) | ((uint64_t)((S >> j) & 1u) << (j + 1)) | ((r & ~low) << 1);
        }
        const uint32_t wi = q >> 6, sh = q & 63;
        const uint64_t w0 = r << sh, w1 = sh ? r >> (64 - sh) : 0;
        atomicOr(&buf[wi], w0);  // unconditional: OR-ing zero is harmless and needs no branch
        atomicOr(&buf[wi + 1], w1);
        // whole words of this run are stored; a first word shared with the
        // run below (head) and a last word not wholly this run's (tail) go to
        // side[p][gw] for k_pb_fixup
        const uint64_t end = o + nv + flags;
        const uint32_t nwords = (uint32_t)(((o & 63) + nv + flags + 63) >> 6);
        uint64_t *st = streams + (uint64_t)p * stride + (o >> 6);
        uint64_t *sd = side + 2 * ((uint64_t)p * nwv + gw);
        for (uint32_t w = lane; w < nwords; w += 64) {
            const uint64_t v = buf[w];
            buf[w] = 0;
            const bool head = w == 0 && (o & 63);
            const bool tail = !head && w == nwords - 1 && (end & 63);
            // one store to the selected word (no branches)
            *(head ? sd : tail ? sd + 1 : st + w) = v;
        }
    });
}

// Pass 4: the words at run boundaries. Only the last non-empty run can be
// shorter than a word (the others hold 2048 elements), so a word is shared
// by at most two runs: run g's head word is its head ORed with run g - 1's
// tail; the last run's tail word is its own.
__global__ __launch_bounds__(256) void k_pb_fixup(const uint32_t *__restrict__ bits, const uint64_t *__restrict__ off,
                                                  const uint64_t *__restrict__ side, uint32_t nwv, uint32_t glast,
                                                  uint64_t *__restrict__ streams, uint64_t stride) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (g > glast) return;
    const uint64_t i = (uint64_t)p * nwv + g;
    const uint64_t o = off[i], end = o + bits[i];
    uint64_t *st = streams + (uint64_t)p * stride;
    if (o & 63) st[o >> 6] = side[2 * i] | side[2 * (i - 1) + 1];
    const bool one_word = (o & 63) && ((o & 63) + bits[i] <= 64);  // the head word is also the last
    if (g == glast && (end & 63) && !one_word) st[(end - 1) >> 6] = side[2 * i + 1];
}

uint64_t pb_stream_words(uint32_t n) { return (2ull * n + 63) / 64; }

template <class T> size_t pb_workspace_bytes(uint32_t n, int nb) {
    const uint64_t nwv = pb_nwaves<T>(n), nblk = div_up(nwv, kPbScanBlk);
    return nwv * nb * (8 + 16 + 4) + nblk * nb * 8 + 256;
}
template size_t pb_workspace_bytes<float>(uint32_t, int);
template size_t pb_workspace_bytes<double>(uint32_t, int);

template <class T>
int launch_pb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb, int compat,
                     uint64_t *streams, uint64_t stride, uint32_t *sizes, void *ws, size_t ws_bytes, hipStream_t s) {
    if (n == 0 || nb < 1 || nb > (int)(8 * sizeof(T))) { set_error("perbit encode: bad n or num_bitplanes"); return kErrArg; }
    if (stride < pb_stream_words(n)) { set_error("perbit encode: stream stride below ceil(2 n / 64) words"); return kErrArg; }
    if (ws_bytes < pb_workspace_bytes<T>(n, nb)) { set_error("perbit encode: workspace too small"); return kErrArg; }
    const uint32_t nwg = pb_nwg<T>(n), nwv = pb_nwaves<T>(n);
    const uint32_t nblk = div_up(nwv, kPbScanBlk);
    if (nblk > kPbScanT) { set_error("perbit encode: too many elements"); return kErrArg; }
    uint64_t *off = static_cast<uint64_t *>(ws);
    uint64_t *side = off + (size_t)nwv * nb;
    uint64_t *sums = side + 2 * (size_t)nwv * nb;
    uint32_t *bits = reinterpret_cast<uint32_t *>(sums + (size_t)nblk * nb);
    hipLaunchKernelGGL(k_pb_count<T>, dim3(nwg), dim3(kPbTPB), 0, s, x, n, exp, key, nb, compat, bits, nwv);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pb_scan_sum, dim3(nblk, nb), dim3(kPbScanT), 0, s, bits, nwv, sums);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pb_scan_off, dim3(nblk, nb), dim3(kPbScanT), 0, s, bits, nwv, sums, off, sizes);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pb_write<T>, dim3(nwg), dim3(kPbTPB), 0, s, x, n, exp, key, nb, compat, off, nwv,
                       streams, stride, side);
    PRODM_CHECK_LAUNCH();
    const uint32_t glast = (n - 1) / (64u * kPbB);  // the last wave with elements
    hipLaunchKernelGGL(k_pb_fixup, dim3(div_up(glast + 1, 256), nb), dim3(256), 0, s, bits, off, side, nwv, glast,
                       streams, stride);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_pb_encode<float>(const float *, uint32_t, int32_t, const uint32_t *, int, int, uint64_t *,
                                     uint64_t, uint32_t *, void *, size_t, hipStream_t);
template int launch_pb_encode<double>(const double *, uint32_t, int32_t, const unsigned long long *, int, int,
                                      uint64_t *, uint64_t, uint32_t *, void *, size_t, hipStream_t);

}  // namespace prodm

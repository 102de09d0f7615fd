// bitplane.hip -- bitplane encoders/decoders and max-abs for gfx950.
//
// Layout: a wavefront owns a chunk of 64 "groups" of 64 consecutive
// elements. For each group and each plane one __ballot yields the 64-bit
// plane word of those 64 elements (two 32-element blocks for u32 streams,
// one 64-element block for u64 streams); lane c keeps the words of group c,
// so the final stores are coalesced 64-bit per lane. Bit i of a word is
// element i of its block, LSB first (GroupedBPEncoder.hpp:345,
// NegaBinaryBPEncoder.hpp:246).
#include "internal.hpp"

namespace prodm {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr uint32_t kChunk = 64 * 64;  // elements per wave

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

template <class T, int MAXNB>
__global__ __launch_bounds__(kThreads) void k_nb_encode(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                        const KeyOf<T> *key, int nb,
                                                        typename Traits<T>::Stream *__restrict__ planes,
                                                        uint64_t stride) {
    using F = typename Traits<T>::Fixed;
    const uint32_t lane = lane_id();
    const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint64_t base = wave * kChunk;
    if (base >= n) return;
    const int shift = nb - (level_exp<T>(exp, key) + 2);  // exp += 2 (:21)
    uint64_t w[MAXNB];
#pragma unroll
    for (int p = 0; p < MAXNB; p++) w[p] = 0;
    for (int c0 = 0; c0 < 64; c0 += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            uint64_t e = base + (uint64_t)(c0 + u) * 64 + lane;
            v[u] = e < n ? x[e] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            F q = to_negabinary<T>(dev_ldexp(v[u], shift));
#pragma unroll
            for (int p = 0; p < MAXNB; p++) {
                if (p < nb) {
                    uint64_t m = ballot64((q >> (nb - 1 - p)) & 1u);
                    w[p] = (lane == (uint32_t)(c0 + u)) ? m : w[p];
                }
            }
        }
    }
    const uint64_t g = base / 64 + lane;  // group index
    if constexpr (sizeof(T) == 4) {
        const uint64_t nw = ((uint64_t)n + 31) / 32;
        if (2 * g < nw) {
#pragma unroll
            for (int p = 0; p < MAXNB; p++)
                if (p < nb) reinterpret_cast<uint64_t *>(planes + (uint64_t)p * stride)[g] = w[p];
        }
    } else {
        const uint64_t nw = ((uint64_t)n + 63) / 64;
        if (g < nw) {
#pragma unroll
            for (int p = 0; p < MAXNB; p++)
                if (p < nb) planes[(uint64_t)p * stride + g] = w[p];
        }
    }
}

template <class T>
int launch_nb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb,
                     typename Traits<T>::Stream *planes, uint64_t stride, hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (n == 0 || nb < 1 || nb > B) { set_error("negabinary encode: bad n or num_bitplanes"); return kErrArg; }
    const uint64_t nw = ((uint64_t)n + B - 1) / B;
    if (sizeof(T) == 4 && ((stride & 1) || stride < ((nw + 1) & ~1ull))) {
        set_error("negabinary encode: plane stride must be even and >= words rounded up to even");
        return kErrArg;
    }
    if (sizeof(T) == 8 && stride < nw) { set_error("negabinary encode: plane stride too small"); return kErrArg; }
    const uint32_t waves = div_up(n, kChunk);
    const uint32_t grid = div_up(waves, kWavesPerBlock);
    if (nb <= 32)
        hipLaunchKernelGGL((k_nb_encode<T, 32>), dim3(grid), dim3(kThreads), 0, s, x, n, exp, key, nb, planes, stride);
    else
        hipLaunchKernelGGL((k_nb_encode<T, 64>), dim3(grid), dim3(kThreads), 0, s, x, n, exp, key, nb, planes, stride);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_nb_encode<float>(const float *, uint32_t, int32_t, const uint32_t *, int, uint32_t *, uint64_t,
                                     hipStream_t);
template int launch_nb_encode<double>(const double *, uint32_t, int32_t, const unsigned long long *, int, uint64_t *,
                                      uint64_t, hipStream_t);

// plane word (64 bits covering the 64 elements of group g) of plane i
template <class T>
__device__ inline uint64_t group_word_lane_load(const PlanePtrs &pl, int k, uint64_t g, uint64_t nw, uint32_t lane) {
    // u32 streams: lane 2i / 2i+1 load the low / high word of plane i
    // u64 streams: lane i loads the word of plane i
    if constexpr (sizeof(T) == 4) {
        const uint32_t i = lane >> 1;
        const uint64_t wi = 2 * g + (lane & 1u);
        if ((int)i < k && wi < nw) return static_cast<const uint32_t *>(pl.p[i])[wi];
        return 0;
    } else {
        if ((int)lane < k && g < nw) return static_cast<const uint64_t *>(pl.p[lane])[g];
        return 0;
    }
}

template <class T>
__device__ inline uint64_t group_word(uint64_t v, int i) {
    if constexpr (sizeof(T) == 4) {
        uint32_t lo = readlane_u32((uint32_t)v, 2 * i), hi = readlane_u32((uint32_t)v, 2 * i + 1);
        return ((uint64_t)hi << 32) | lo;
    } else {
        return readlane_u64(v, i);
    }
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_nb_decode(PlanePtrs pl, uint32_t n, int32_t exp, const KeyOf<T> *key,
                                                        int start, int k, T *__restrict__ out) {
    using F = typename Traits<T>::Fixed;
    using SF = typename Traits<T>::SFixed;
    constexpr int B = Traits<T>::kBlock;
    const uint32_t lane = lane_id();
    const uint64_t ngroups = ((uint64_t)n + 63) / 64;
    const uint64_t nw = ((uint64_t)n + B - 1) / B;
    const int e2 = level_exp<T>(exp, key) + 2;
    const int end = start + k;
    const uint64_t wave0 = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t g = wave0; g < ngroups; g += nwaves) {
        const uint64_t v = group_word_lane_load<T>(pl, k, g, nw, lane);
        F acc = 0;
        for (int i = 0; i < k; i++) {
            const uint64_t word = group_word<T>(v, i);
            acc |= (F)((word >> lane) & 1u) << (k - 1 - i);
        }
        const SF val = (SF)((acc ^ (F)Traits<T>::kNbMask) - (F)Traits<T>::kNbMask);
        const T y = dev_ldexp((T)val, e2 - end);
        const uint64_t e = g * 64 + lane;
        if (e < n) out[e] = (end & 1) ? -y : y;
    }
}

template <class T>
int launch_nb_decode(const PlanePtrs &planes, uint32_t n, int32_t exp, const KeyOf<T> *key, int start, int k, T *out,
                     hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (n == 0) return kOk;
    if (k < 0 || start < 0 || start + k > B) { set_error("negabinary decode: bad plane range"); return kErrArg; }
    const uint64_t ngroups = ((uint64_t)n + 63) / 64;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(ngroups, kWavesPerBlock), 16384);
    if (k == 0) {
        // k == 0: the reference returns zeros (:135-138)
        if constexpr (sizeof(T) == 4) return launch_fill_u32(reinterpret_cast<uint32_t *>(out), n, 0u, s);
        else return launch_fill_u64(reinterpret_cast<uint64_t *>(out), n, 0ull, s);
    }
    hipLaunchKernelGGL(k_nb_decode<T>, dim3(grid), dim3(kThreads), 0, s, planes, n, exp, key, start, k, out);
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

// blocks per wave chunk: 128 (u32) or 64 (u64)
template <class T> constexpr uint32_t blocks_per_wave() { return kChunk / Traits<T>::kBlock; }
template <class T> constexpr uint32_t blocks_per_wg() { return blocks_per_wave<T>() * kWavesPerBlock; }

// Pass 1: rb per block and an rb histogram per workgroup (hist[wg][0..nb]).
template <class T>
__global__ __launch_bounds__(kThreads) void k_gb_rb(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                    const KeyOf<T> *key, int nb, int compat, uint8_t *__restrict__ rb,
                                                    uint32_t *__restrict__ hist) {
    using F = typename Traits<T>::Fixed;
    constexpr int B = Traits<T>::kBlock;
    __shared__ uint32_t h[65];
    for (int i = threadIdx.x; i <= nb; i += kThreads) h[i] = 0;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint64_t base = wave * kChunk;
    const uint32_t nblk = (n - 1) / B + 1;
    const int shift = nb - level_exp<T>(exp, key);
    if (base < n) {
        for (int c = 0; c < 64; c++) {
            const uint64_t e = base + (uint64_t)c * 64 + lane;
            F m = e < n ? gb_mag<T>(x[e], shift, compat) : (F)0;
            // OR over the block (32 or 64 lanes)
            for (int off = B / 2; off >= 1; off >>= 1) m |= __shfl_xor(m, off, 64);
            if ((lane % B) == 0) {
                const uint64_t b = (base + (uint64_t)c * 64 + lane) / B;
                if (b < nblk) {
                    const int r = rb_of<F>(m, nb);
                    rb[b] = (uint8_t)r;
                    atomicAdd(&h[r], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= nb; i += kThreads) hist[(uint64_t)blockIdx.x * (nb + 1) + i] = h[i];
}

// words of block-group counts per plane p: #(rb <= p) + #(rb == p)
// One workgroup per plane scans cnt over workgroups: off[wg][i] exclusive.
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

// Pass 2: plane words, compacted per plane.
template <class T, int MAXNB>
__global__ __launch_bounds__(kThreads) void k_gb_write(const T *__restrict__ x, uint32_t n, int32_t exp,
                                                       const KeyOf<T> *key, int nb, int compat,
                                                       const uint32_t *__restrict__ off,
                                                       typename Traits<T>::Stream *__restrict__ planes,
                                                       uint64_t stride) {
    using F = typename Traits<T>::Fixed;
    using S = typename Traits<T>::Stream;
    constexpr int B = Traits<T>::kBlock;
    __shared__ uint32_t tot[kWavesPerBlock][64];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / kWave;
    const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + wv;
    const uint64_t base = wave * kChunk;
    const uint32_t nblk = (n - 1) / B + 1;
    const int shift = nb - level_exp<T>(exp, key);
    uint64_t w[MAXNB];
    uint64_t sg = 0;
#pragma unroll
    for (int p = 0; p < MAXNB; p++) w[p] = 0;
    if (base < n) {
        for (int c = 0; c < 64; c++) {
            const uint64_t e = base + (uint64_t)c * 64 + lane;
            const T xv = e < n ? x[e] : (T)0;
            const F m = gb_mag<T>(xv, shift, compat);
            const uint64_t sm = ballot64(xv < 0);
            sg = (lane == (uint32_t)c) ? sm : sg;
#pragma unroll
            for (int p = 0; p < MAXNB; p++) {
                if (p < nb) {
                    const uint64_t bm = ballot64((m >> (nb - 1 - p)) & 1u);
                    w[p] = (lane == (uint32_t)c) ? bm : w[p];
                }
            }
        }
    }
    // this lane owns blocks b0 (and b1 for u32 streams) of group `lane`
    const uint64_t g = base / 64 + lane;
    const uint64_t b0 = g * (64 / B);
    const bool v0 = b0 < nblk, v1 = (B == 32) && (b0 + 1 < nblk);
    int r0 = nb, r1 = nb;
#pragma unroll
    for (int p = MAXNB - 1; p >= 0; p--) {
        if (p < nb) {
            const uint64_t lo = (B == 32) ? (w[p] & 0xffffffffull) : w[p];
            const uint64_t hi = (B == 32) ? (w[p] >> 32) : 0;
            if (lo) r0 = p;
            if (hi) r1 = p;
        }
    }
    if (!v0) r0 = 255;
    if (!v1) r1 = 255;
    // per-plane word counts and in-wave prefix (ballot + mbcnt)
#pragma unroll
    for (int p = 0; p < MAXNB; p++) {
        if (p < nb) {
            const uint64_t a = ballot64(r0 == p), bq = ballot64(r0 <= p), c2 = ballot64(r1 == p), d = ballot64(r1 <= p);
            if (lane == 0)
                tot[wv][p] = __popcll(a) + __popcll(bq) + __popcll(c2) + __popcll(d);
        }
    }
    __syncthreads();
    if (base >= n) return;
#pragma unroll
    for (int p = 0; p < MAXNB; p++) {
        if (p < nb) {
            const uint64_t a = ballot64(r0 == p), bq = ballot64(r0 <= p), c2 = ballot64(r1 == p), d = ballot64(r1 <= p);
            uint64_t pos = off[(uint64_t)blockIdx.x * nb + p];
            for (uint32_t q = 0; q < wv; q++) pos += tot[q][p];
            pos += lanes_below(a) + lanes_below(bq) + lanes_below(c2) + lanes_below(d);
            S *dst = planes + (uint64_t)p * stride;
            if (r0 == p) dst[pos++] = (S)((B == 32) ? (sg & 0xffffffffull) : sg);
            if (r0 <= p) dst[pos++] = (S)((B == 32) ? (w[p] & 0xffffffffull) : w[p]);
            if (r1 == p) dst[pos++] = (S)(sg >> 32);
            if (r1 <= p) dst[pos++] = (S)(w[p] >> 32);
        }
    }
}

template <class T> size_t gb_workspace_bytes(uint32_t n, int nb) {
    const uint32_t nwg = div_up(div_up(n, kChunk), kWavesPerBlock);
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
    const uint32_t nwg = div_up(div_up(n, kChunk), kWavesPerBlock);
    uint32_t *hist = static_cast<uint32_t *>(ws);
    uint32_t *off = hist + (size_t)nwg * (nb + 1);
    hipLaunchKernelGGL(k_gb_rb<T>, dim3(nwg), dim3(kThreads), 0, s, x, n, exp, key, nb, compat, rb, hist);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gb_scan, dim3(nb), dim3(1024), 0, s, hist, nwg, nb + 1, 0, nb, off, words_per_plane);
    PRODM_CHECK_LAUNCH();
    if (nb <= 32)
        hipLaunchKernelGGL((k_gb_write<T, 32>), dim3(nwg), dim3(kThreads), 0, s, x, n, exp, key, nb, compat, off,
                           planes, stride);
    else
        hipLaunchKernelGGL((k_gb_write<T, 64>), dim3(nwg), dim3(kThreads), 0, s, x, n, exp, key, nb, compat, off,
                           planes, stride);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_gb_encode<float>(const float *, uint32_t, int32_t, const uint32_t *, int, int, uint8_t *,
                                     uint32_t *, uint64_t, uint32_t *, void *, size_t, hipStream_t);
template int launch_gb_encode<double>(const double *, uint32_t, int32_t, const unsigned long long *, int, int,
                                      uint8_t *, uint64_t *, uint64_t, uint32_t *, void *, size_t, hipStream_t);

// ---- GroupedBP progressive decode ----------------------------------------------
// rb histogram per workgroup (decode side; same chunking as the encoder)
template <class T>
__global__ __launch_bounds__(kThreads) void k_gb_rbhist(const uint8_t *__restrict__ rb, uint32_t nblk, int nbins,
                                                        uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    for (int i = threadIdx.x; i < nbins; i += kThreads) h[i] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * blocks_per_wg<T>();
    for (uint32_t i = threadIdx.x; i < blocks_per_wg<T>(); i += kThreads) {
        const uint64_t b = b0 + i;
        if (b < nblk) {
            const int r = rb[b];
            if (r < nbins) atomicAdd(&h[r], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += kThreads) hist[(uint64_t)blockIdx.x * nbins + i] = h[i];
}

template <class S> __device__ inline S load_word(const uint8_t *base, uint64_t byte_off, bool aligned) {
    if (aligned) return *reinterpret_cast<const S *>(base + byte_off);
    S v = 0;
#pragma unroll
    for (int i = 0; i < (int)sizeof(S); i++) v |= (S)base[byte_off + i] << (8 * i);
    return v;
}

template <class T, int MAXK>
__global__ __launch_bounds__(kThreads) void k_gb_decode(PlanePtrs pl, uint32_t n, int32_t exp, const KeyOf<T> *key,
                                                        int start, int k, const uint8_t *__restrict__ rb,
                                                        typename Traits<T>::Stream *__restrict__ signs,
                                                        const uint32_t *__restrict__ off, T *__restrict__ out,
                                                        uint64_t hdr0) {
    using F = typename Traits<T>::Fixed;
    using S = typename Traits<T>::Stream;
    constexpr int B = Traits<T>::kBlock;
    __shared__ uint32_t tot[kWavesPerBlock][64];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / kWave;
    const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + wv;
    const uint64_t base = wave * kChunk;
    const uint32_t nblk = (n - 1) / B + 1;
    const int e0 = level_exp<T>(exp, key);
    const int end = start + k;
    const uint64_t g = base / 64 + lane;
    const uint64_t b0 = g * (64 / B);
    const bool v0 = b0 < nblk, v1 = (B == 32) && (b0 + 1 < nblk);
    const int r0 = v0 ? rb[b0] : 255, r1 = v1 ? rb[b0 + 1] : 255;
    for (int i = 0; i < k; i++) {
        const int p = start + i;
        const uint64_t a = ballot64(r0 == p), bq = ballot64(r0 <= p), c2 = ballot64(r1 == p), d = ballot64(r1 <= p);
        if (lane == 0) tot[wv][i] = __popcll(a) + __popcll(bq) + __popcll(c2) + __popcll(d);
    }
    __syncthreads();
    if (base >= n) return;
    S s0 = v0 ? signs[b0] : 0, s1 = v1 ? signs[b0 + 1] : 0;
    uint64_t wl[MAXK];
#pragma unroll
    for (int i = 0; i < MAXK; i++) wl[i] = 0;
#pragma unroll
    for (int i = 0; i < MAXK; i++) {
        if (i < k) {
            const int p = start + i;
            const uint64_t a = ballot64(r0 == p), bq = ballot64(r0 <= p), c2 = ballot64(r1 == p), d = ballot64(r1 <= p);
            uint64_t pos = off[(uint64_t)blockIdx.x * k + i];
            for (uint32_t q = 0; q < wv; q++) pos += tot[q][i];
            pos += lanes_below(a) + lanes_below(bq) + lanes_below(c2) + lanes_below(d);
            const uint8_t *sb = static_cast<const uint8_t *>(pl.p[i]);
            const uint64_t hb = (i == 0) ? hdr0 : 0;
            const bool al = (hb % sizeof(S)) == 0;
            uint64_t lo = 0, hi = 0;
            if (r0 == p) { s0 = load_word<S>(sb, hb + pos * sizeof(S), al); pos++; }
            if (r0 <= p) { lo = load_word<S>(sb, hb + pos * sizeof(S), al); pos++; }
            if (r1 == p) { s1 = load_word<S>(sb, hb + pos * sizeof(S), al); pos++; }
            if (r1 <= p) { hi = load_word<S>(sb, hb + pos * sizeof(S), al); pos++; }
            wl[i] = (B == 32) ? ((hi << 32) | (lo & 0xffffffffull)) : lo;
        }
    }
    // persist the signs of blocks recorded in this window
    if (v0 && r0 >= start && r0 < end) signs[b0] = s0;
    if (v1 && r1 >= start && r1 < end) signs[b0 + 1] = s1;
    const uint64_t sg = (B == 32) ? (((uint64_t)s1 << 32) | ((uint64_t)s0 & 0xffffffffull)) : (uint64_t)s0;
    // transpose: group c's words live in lane c
    for (int c = 0; c < 64; c++) {
        const uint64_t e = base + (uint64_t)c * 64 + lane;
        if (base + (uint64_t)c * 64 >= n) break;
        F acc = 0;
#pragma unroll
        for (int i = 0; i < MAXK; i++) {
            if (i < k) {
                const uint64_t word = readlane_u64(wl[i], c);
                acc |= (F)((word >> lane) & 1u) << (k - 1 - i);
            }
        }
        const uint64_t sgc = readlane_u64(sg, c);
        const int rr0 = __builtin_amdgcn_readlane(r0, c), rr1 = __builtin_amdgcn_readlane(r1, c);
        const int rmine = (B == 32 && lane >= 32) ? rr1 : rr0;
        T y = dev_ldexp((T)acc, e0 - end);
        if ((sgc >> lane) & 1u) y = -y;
        if (rmine >= end) y = (T)0;
        if (e < n) out[e] = y;
    }
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
    const uint32_t nwg = div_up(nblk, blocks_per_wg<T>());
    const int nbins = start + k;  // rb values >= end never contribute
    if (ws_bytes < (size_t)nwg * (nbins + k) * 4) { set_error("grouped decode: workspace too small"); return kErrArg; }
    uint32_t *hist = static_cast<uint32_t *>(ws);
    uint32_t *off = hist + (size_t)nwg * nbins;
    uint32_t *totals = off + (size_t)nwg * k;
    hipLaunchKernelGGL(k_gb_rbhist<T>, dim3(nwg), dim3(kThreads), 0, s, rb, nblk, nbins, hist);
    PRODM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gb_scan, dim3(k), dim3(1024), 0, s, hist, nwg, nbins, start, k, off, totals);
    PRODM_CHECK_LAUNCH();
    if (k <= 32)
        hipLaunchKernelGGL((k_gb_decode<T, 32>), dim3(nwg), dim3(kThreads), 0, s, streams, n, exp, key, start, k, rb,
                           signs, off, out, hdr0);
    else
        hipLaunchKernelGGL((k_gb_decode<T, 64>), dim3(nwg), dim3(kThreads), 0, s, streams, n, exp, key, start, k, rb,
                           signs, off, out, hdr0);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_gb_decode<float>(const PlanePtrs &, uint32_t, int32_t, const uint32_t *, int, int, const uint8_t *,
                                     uint32_t *, float *, void *, size_t, uint64_t, hipStream_t);
template int launch_gb_decode<double>(const PlanePtrs &, uint32_t, int32_t, const unsigned long long *, int, int,
                                      const uint8_t *, uint64_t *, double *, void *, size_t, uint64_t, hipStream_t);

template <class T> size_t gb_decode_workspace_bytes(uint32_t n, int start, int k) {
    constexpr int B = Traits<T>::kBlock;
    const uint32_t nblk = (n - 1) / B + 1;
    const uint32_t nwg = div_up(nblk, blocks_per_wg<T>());
    return (size_t)nwg * (start + 2 * k) * 4 + 64 * 4 + 256;
}
template size_t gb_decode_workspace_bytes<float>(uint32_t, int, int);
template size_t gb_decode_workspace_bytes<double>(uint32_t, int, int);

}  // namespace prodm

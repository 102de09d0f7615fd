// collect.hip -- SquaredErrorCollector (SquaredErrorCollector.hpp:23-65) on gfx950.
//
// The reference adds, per element, the squared truncation error of every
// bitplane prefix (computed in T) into a double table, sequentially. A
// parallel FP64 sum cannot reproduce that order, so each table entry is
// defined as the EXACTLY ROUNDED sum of the same terms: order independent,
// and identical between this kernel and the CPU restatement.
//
// Exactness: every truncation term is a multiple of Q = max(q^2, min_sub)
// with q = 2^(le - nb - prec) (the finest mantissa step of any element whose
// planes are truncated), and is at most 2^(2 le). Terms are accumulated as
// integers in units of Q in per-thread multi-limb registers, reduced across
// the wave and workgroup with carries, and summed exactly on the host. The
// rare squares x^2 of elements far below the level maximum that are not
// multiples of Q go to a per-(bin, exponent) histogram of mantissa sums.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace prodm {

template <class T> struct SqCfg;
template <> struct SqCfg<float> {
    static constexpr int kLimbs = 2;      // 2^128 > 2^(2*32+46) * 2^12 terms per wave chunk
    static constexpr int kExpBuckets = 256;
    static constexpr int kEmin = -149;
    static constexpr int kMantBits = 23;
};
template <> struct SqCfg<double> {
    static constexpr int kLimbs = 4;      // 2^256 > 2^(2*64+104) * 2^12
    static constexpr int kExpBuckets = 2048;
    static constexpr int kEmin = -1074;
    static constexpr int kMantBits = 52;
};

constexpr int kSqThreads = 256;
constexpr int kSqEPT = 16;  // elements per thread

template <class T> __device__ inline typename Traits<T>::Bits to_bits(T v) {
    typename Traits<T>::Bits b;
    __builtin_memcpy(&b, &v, sizeof(b));
    return b;
}
template <class T> __device__ inline T from_bits(typename Traits<T>::Bits b) {
    T v;
    __builtin_memcpy(&v, &b, sizeof(v));
    return v;
}

// term of bin k for element x, restating the reference loop (:326-356):
// bins s_idx, s_idx-1, ... get the squared residual after clearing mantissa
// bits 0..b (b = b0, b0+1, ..., prec-1); the remaining bins get x^2.
template <class T>
__device__ inline T sq_term(T x, int de, int le, int nb, int k) {
    using Bits = typename Traits<T>::Bits;
    constexpr int P = Traits<T>::kPrec;
    const int exp_diff = le - de + P - nb;
    const int s_idx = exp_diff > 0 ? nb : nb + exp_diff;
    const int b0 = exp_diff > 0 ? exp_diff : 0;
    if (s_idx <= 0 || k > s_idx) return (T)0;
    if (b0 < P && k >= s_idx - (P - 1 - b0)) {
        const int b = b0 + (s_idx - k);
        const Bits mask = ~(((Bits)2 << b) - 1);  // clears bits 0..b
        const T xb = from_bits<T>(to_bits<T>(x) & mask);
        const T d = x - xb;
        return d * d;
    }
    return x * x;
}

template <class T>
__global__ __launch_bounds__(kSqThreads) void k_collect_squared(const T *__restrict__ x, uint64_t n, int nb, int le,
                                                                int log2q, uint64_t *__restrict__ partial,
                                                                unsigned long long *__restrict__ under) {
    constexpr int LM = SqCfg<T>::kLimbs;
    constexpr int MB = SqCfg<T>::kMantBits;
    using Bits = typename Traits<T>::Bits;
    __shared__ uint64_t wacc[kSqThreads / 64][65][LM];
    const uint32_t lane = lane_id(), wv = threadIdx.x / 64;
    const uint64_t base = (uint64_t)blockIdx.x * kSqThreads * kSqEPT;
    T v[kSqEPT];
    int de[kSqEPT];
#pragma unroll
    for (int i = 0; i < kSqEPT; i++) {
        const uint64_t e = base + (uint64_t)i * kSqThreads + threadIdx.x;
        v[i] = e < n ? x[e] : (T)0;
        de[i] = dev_frexp_exp(v[i]);
    }
    for (int k = 0; k <= nb; k++) {
        uint64_t a[LM];
#pragma unroll
        for (int j = 0; j < LM; j++) a[j] = 0;
#pragma unroll
        for (int i = 0; i < kSqEPT; i++) {
            if (v[i] == (T)0) continue;
            const T t = sq_term<T>(v[i], de[i], le, nb, k);
            if (t == (T)0) continue;
            const Bits tb = to_bits<T>(t);
            const int be = (int)((tb >> MB) & ((1u << (8 * sizeof(T) == 32 ? 8 : 11)) - 1));
            const uint64_t M = (uint64_t)(tb & (((Bits)1 << MB) - 1)) | (be ? ((uint64_t)1 << MB) : 0);
            const int E = (be ? be : 1) - 1 + SqCfg<T>::kEmin;  // t = M * 2^E
            const int s = E - log2q;
            if (s < 0) {
                atomicAdd(&under[(size_t)k * SqCfg<T>::kExpBuckets + (be ? be : 1)], (unsigned long long)M);
                continue;
            }
            const int li = s >> 6, off = s & 63;
            const uint64_t plo = M << off;
            const uint64_t phi = off ? (M >> (64 - off)) : 0;
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const uint64_t add = (j == li ? plo : 0) + (j == li + 1 ? phi : 0);
                const uint64_t s1 = a[j] + add;
                const uint64_t c1 = s1 < add;
                const uint64_t s2 = s1 + carry;
                const uint64_t c2 = s2 < carry;
                a[j] = s2;
                carry = c1 | c2;
            }
        }
        // wave reduction with carries
        for (int o = 32; o >= 1; o >>= 1) {
            uint64_t b[LM];
#pragma unroll
            for (int j = 0; j < LM; j++) b[j] = __shfl_xor(a[j], o, 64);
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const uint64_t s1 = a[j] + b[j];
                const uint64_t c1 = s1 < b[j];
                const uint64_t s2 = s1 + carry;
                const uint64_t c2 = s2 < carry;
                a[j] = s2;
                carry = c1 | c2;
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < LM; j++) wacc[wv][k][j] = a[j];
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= nb; k += kSqThreads) {
        uint64_t a[LM];
#pragma unroll
        for (int j = 0; j < LM; j++) a[j] = 0;
        for (int w = 0; w < kSqThreads / 64; w++) {
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const uint64_t bj = wacc[w][k][j];
                const uint64_t s1 = a[j] + bj;
                const uint64_t c1 = s1 < bj;
                const uint64_t s2 = s1 + carry;
                const uint64_t c2 = s2 < carry;
                a[j] = s2;
                carry = c1 | c2;
            }
        }
#pragma unroll
        for (int j = 0; j < LM; j++) partial[((uint64_t)blockIdx.x * (nb + 1) + k) * LM + j] = a[j];
    }
}

// workspace: [header: nwg][partials][underflow histogram]
template <class T> static uint32_t sq_nwg(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kSqThreads * kSqEPT - 1) / (kSqThreads * kSqEPT)); }
template <class T> size_t sq_result_bytes_n(uint64_t n, int nb) {
    return 64 + (size_t)sq_nwg<T>(n) * (nb + 1) * SqCfg<T>::kLimbs * 8 + (size_t)(nb + 1) * SqCfg<T>::kExpBuckets * 8;
}
template <class T> size_t sq_workspace_bytes(uint64_t n, int nb) { return sq_result_bytes_n<T>(n, nb); }
template size_t sq_workspace_bytes<float>(uint64_t, int);
template size_t sq_workspace_bytes<double>(uint64_t, int);

// the host copies the whole result area; its size depends on n, so it is
// encoded in the header written by the launch
static thread_local size_t g_sq_last_bytes = 0;
template <class T> size_t sq_result_bytes(int) { return g_sq_last_bytes; }
template size_t sq_result_bytes<float>(int);
template size_t sq_result_bytes<double>(int);

template <class T> static int log2_quantum(int le, int nb) {
    const long q2 = 2L * ((long)le - nb - Traits<T>::kPrec);
    return (int)std::max<long>(q2, SqCfg<T>::kEmin);
}

template <class T>
int launch_collect_squared(const T *x, uint64_t n, int nb, T maxerr, void *ws, size_t ws_bytes, hipStream_t s) {
    int le = 0;
    std::frexp(maxerr, &le);
    const uint32_t nwg = sq_nwg<T>(n);
    const size_t bytes = sq_result_bytes_n<T>(n, nb);
    if (ws_bytes < bytes) { set_error("squared collector: workspace too small"); return kErrArg; }
    g_sq_last_bytes = bytes;
    char *base = static_cast<char *>(ws);
    uint64_t *hdr = reinterpret_cast<uint64_t *>(base);
    uint64_t *partial = reinterpret_cast<uint64_t *>(base + 64);
    unsigned long long *under =
        reinterpret_cast<unsigned long long *>(base + 64 + (size_t)nwg * (nb + 1) * SqCfg<T>::kLimbs * 8);
    int rc = launch_fill_u64(hdr, 8, 0ull, s);
    if (!rc) rc = launch_fill_u64(hdr, 1, (uint64_t)nwg, s);
    if (!rc) rc = launch_fill_u64(reinterpret_cast<uint64_t *>(under), (size_t)(nb + 1) * SqCfg<T>::kExpBuckets, 0ull, s);
    if (rc) return rc;
    if (n == 0) {
        return launch_fill_u64(partial, (size_t)nwg * (nb + 1) * SqCfg<T>::kLimbs, 0ull, s);
    }
    hipLaunchKernelGGL(k_collect_squared<T>, dim3(nwg), dim3(kSqThreads), 0, s, x, n, nb, le, log2_quantum<T>(le, nb),
                       partial, under);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_collect_squared<float>(const float *, uint64_t, int, float, void *, size_t, hipStream_t);
template int launch_collect_squared<double>(const double *, uint64_t, int, double, void *, size_t, hipStream_t);

// ---- exact host summation and rounding ---------------------------------------------
namespace {
struct Big {
    std::vector<uint64_t> w;  // little endian
    explicit Big(size_t limbs) : w(limbs, 0) {}
    void add_shifted(uint64_t v, int shift) {  // v << shift, shift >= 0
        if (!v) return;
        const size_t li = (size_t)shift / 64;
        const int off = shift % 64;
        uint64_t parts[2] = {v << off, off ? (v >> (64 - off)) : 0};
        uint64_t carry = 0;
        for (size_t j = li; j < w.size(); j++) {
            const uint64_t add = (j - li < 2) ? parts[j - li] : 0;
            if (j - li >= 2 && !carry) break;
            const uint64_t s1 = w[j] + add;
            const uint64_t c1 = s1 < add;
            const uint64_t s2 = s1 + carry;
            const uint64_t c2 = s2 < carry;
            w[j] = s2;
            carry = c1 | c2;
        }
    }
    int bit(long i) const { return i < 0 ? 0 : (int)((w[(size_t)i / 64] >> (i % 64)) & 1u); }
    // value * 2^scale rounded to nearest even
    double to_double(int scale) const {
        long top = -1;
        for (long i = (long)w.size() - 1; i >= 0; i--)
            if (w[i]) { top = i * 64 + 63 - __builtin_clzll(w[i]); break; }
        if (top < 0) return 0.0;
        if (top <= 52) return std::ldexp((double)w[0], scale);
        const long lsb = top - 52;
        uint64_t mant = 0;
        for (long b = top; b >= lsb; b--) mant = (mant << 1) | (uint64_t)bit(b);
        const int rb = bit(lsb - 1);
        int sticky = 0;
        for (long b = lsb - 2; b >= 0 && !sticky; b--) sticky = bit(b);
        if (rb && (sticky || (mant & 1u))) mant++;
        return std::ldexp((double)mant, (int)(scale + lsb));
    }
};
}  // namespace

template <class T> int finish_collect_squared(const void *h, int nb, T maxerr, double *out) {
    constexpr int LM = SqCfg<T>::kLimbs;
    constexpr int NE = SqCfg<T>::kExpBuckets;
    int le = 0;
    std::frexp(maxerr, &le);
    const int log2q = log2_quantum<T>(le, nb);
    const uint8_t *base = static_cast<const uint8_t *>(h);
    uint64_t nwg;
    std::memcpy(&nwg, base, 8);
    const uint64_t *partial = reinterpret_cast<const uint64_t *>(base + 64);
    const uint64_t *under = partial + nwg * (nb + 1) * LM;
    // common scale: 2^emin; window terms sit at 2^log2q = 2^emin * 2^(log2q - emin)
    const int wshift = log2q - SqCfg<T>::kEmin;
    const size_t limbs = (size_t)(wshift + 64 * LM + 64 + 64 + 2 * 1100) / 64 + 2;
    for (int k = 0; k <= nb; k++) {
        Big b(limbs);
        for (uint64_t g = 0; g < nwg; g++)
            for (int j = 0; j < LM; j++) b.add_shifted(partial[(g * (nb + 1) + k) * LM + j], wshift + 64 * j);
        for (int e = 1; e < NE; e++) b.add_shifted(under[(size_t)k * NE + e], e - 1);
        out[k] = b.to_double(SqCfg<T>::kEmin);
    }
    return kOk;
}
template int finish_collect_squared<float>(const void *, int, float, double *);
template int finish_collect_squared<double>(const void *, int, double, double *);

// ---- encoder-fused level errors ------------------------------------------------------
// NegaBinaryBPEncoder::encode(..., level_errors) (NegaBinaryBPEncoder.hpp:65-125,
// collect_level_errors :231-239) and GroupedBPEncoder::encode(..., level_errors)
// (GroupedBPEncoder.hpp:76-145, :324-335), for fp32 data (the reference's
// collectors take float / uint32). Per element, in the shifted units
// s = x 2^(nb - e) (e + 2 for NegaBinary), bin nb - k holds the squared
// value of the digits below the top nb - k planes plus the fraction m, bin
// nb the fraction alone and bin 0 s^2; bins 0 and nb square in float as the
// reference does. Each entry is the exactly rounded sum over elements, then
// scaled by 2^(2 (e - nb)).
//
// Exactness: in shifted units every term of an element with |s| >= 1 is a
// multiple of 2^-46 (the fraction's step is at least 2^-23) and every term
// of an element with 2^-32 <= |s| < 1 a multiple of 2^-110, so terms are
// integers in units of Q = 2^-110 below 2^176 (3 limbs, 2^13 terms per
// workgroup). Terms of smaller elements go to a per-(bin, exponent)
// histogram of 128-bit mantissa sums.
namespace {
constexpr int kFeLimbs = 3;
constexpr int kFeLog2Q = -110;
constexpr int kFeThreads = 256;
constexpr int kFeEPT = 32;
constexpr int kFeExpBuckets = 2048;

__device__ __forceinline__ uint32_t fe_mask(int k) { return k >= 31 ? 0x7fffffffu : ((1u << k) - 1u); }

template <int ENC>
__device__ __forceinline__ double fe_term(float s, float m, uint32_t bits, int nb, int bin) {
    if (ENC == 0) {  // NegaBinary: bits = negabinary digits of (int32) s
        if (bin == nb) { const float mm = m * m; return (double)mm; }
        if (bin == 0) { const float ss = s * s; return (double)ss; }
        const int32_t v = (int32_t)(((bits & fe_mask(nb - bin)) ^ 0xaaaaaaaau) - 0xaaaaaaaau);
        const double d = (double)v + (double)m;
        return d * d;
    }
    // GroupedBP: s = |shifted|, bits = (uint32) s, m = s - bits (float)
    const double md = (double)m;
    if (bin == nb) return md * md;
    if (bin == 0) { const float ss = s * s; return (double)ss; }
    const double d = (double)(bits & fe_mask(nb - bin)) + md;
    return d * d;
}

template <int ENC>
__global__ __launch_bounds__(kFeThreads) void k_fused_errors(const float *__restrict__ x, uint64_t n, int shift,
                                                             int nb, uint64_t *__restrict__ partial,
                                                             unsigned long long *__restrict__ under) {
    constexpr int LM = kFeLimbs;
    __shared__ uint64_t wacc[kFeThreads / 64][33][LM];
    const uint32_t lane = lane_id(), wv = threadIdx.x / 64;
    const uint64_t base = (uint64_t)blockIdx.x * kFeThreads * kFeEPT;
    float s[kFeEPT], m[kFeEPT];
    uint32_t bits[kFeEPT];
#pragma unroll
    for (int i = 0; i < kFeEPT; i++) {
        const uint64_t e = base + (uint64_t)i * kFeThreads + threadIdx.x;
        // unconditional load (past n: element 0, discarded), so that the
        // loads of all elements are in flight together
        const float x0 = x[e < n ? e : 0];
        const float v = e < n ? x0 : 0.0f;
        const float sh = ldexpf(v, shift);
        if (ENC == 0) {
            // (int32) shifted as in the reference (in range: |x| < 2^e)
            const int32_t q = (sh == sh && fabsf(sh) < 2147483648.0f) ? (int32_t)sh : (int32_t)0x80000000;
            bits[i] = ((uint32_t)q + 0xaaaaaaaau) ^ 0xaaaaaaaau;
            s[i] = sh;
            m[i] = sh - (float)q;
        } else {
            const float a = fabsf(sh);
            const uint32_t fp = (a == a && a < 4294967296.0f) ? (uint32_t)a : 0u;
            bits[i] = fp;
            s[i] = a;
            m[i] = a - (float)fp;
        }
    }
    for (int bin = 0; bin <= nb; bin++) {
        uint64_t a[LM];
#pragma unroll
        for (int j = 0; j < LM; j++) a[j] = 0;
#pragma unroll 4
        for (int i = 0; i < kFeEPT; i++) {
            const double t = fe_term<ENC>(s[i], m[i], bits[i], nb, bin);
            if (t == 0.0) continue;
            uint64_t tb;
            __builtin_memcpy(&tb, &t, 8);
            const int be = (int)((tb >> 52) & 0x7ff);
            const uint64_t M = (tb & ((1ull << 52) - 1)) | (be ? (1ull << 52) : 0);
            const int E = (be ? be : 1) - 1075;  // t = M 2^E
            const int sft = E - kFeLog2Q;
            if (sft < 0) {
                // 128-bit bucket: low word with carry into the high word
                unsigned long long *bk = under + ((size_t)bin * kFeExpBuckets + (be ? be : 1)) * 2;
                const unsigned long long old = atomicAdd(bk, (unsigned long long)M);
                if (old + M < old) atomicAdd(bk + 1, 1ull);
                continue;
            }
            const int li = sft >> 6, off = sft & 63;
            const uint64_t plo = M << off;
            const uint64_t phi = off ? (M >> (64 - off)) : 0;
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const uint64_t add = (j == li ? plo : 0) + (j == li + 1 ? phi : 0);
                const uint64_t s1 = a[j] + add;
                const uint64_t c1 = s1 < add;
                const uint64_t s2 = s1 + carry;
                const uint64_t c2 = s2 < carry;
                a[j] = s2;
                carry = c1 | c2;
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            uint64_t b[LM];
#pragma unroll
            for (int j = 0; j < LM; j++) b[j] = __shfl_xor(a[j], o, 64);
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const uint64_t s1 = a[j] + b[j];
                const uint64_t c1 = s1 < b[j];
                const uint64_t s2 = s1 + carry;
                const uint64_t c2 = s2 < carry;
                a[j] = s2;
                carry = c1 | c2;
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < LM; j++) wacc[wv][bin][j] = a[j];
        }
    }
    __syncthreads();
    for (int bin = threadIdx.x; bin <= nb; bin += kFeThreads) {
        uint64_t a[LM];
#pragma unroll
        for (int j = 0; j < LM; j++) a[j] = 0;
        for (int w = 0; w < kFeThreads / 64; w++) {
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const uint64_t bj = wacc[w][bin][j];
                const uint64_t s1 = a[j] + bj;
                const uint64_t c1 = s1 < bj;
                const uint64_t s2 = s1 + carry;
                const uint64_t c2 = s2 < carry;
                a[j] = s2;
                carry = c1 | c2;
            }
        }
#pragma unroll
        for (int j = 0; j < LM; j++) partial[((uint64_t)blockIdx.x * (nb + 1) + bin) * LM + j] = a[j];
    }
}

uint32_t fe_nwg(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kFeThreads * kFeEPT - 1) / (kFeThreads * kFeEPT)); }
size_t fe_partial_bytes(uint64_t n, int nb) { return (size_t)fe_nwg(n) * (nb + 1) * kFeLimbs * 8; }
size_t fe_under_bytes(int nb) { return (size_t)(nb + 1) * kFeExpBuckets * 16; }
}  // namespace

size_t fused_errors_workspace_bytes(uint64_t n, int nb) {
    return fe_partial_bytes(n, nb) + fe_under_bytes(nb);
}

int fused_level_errors(int enc, const float *x, uint64_t n, int32_t exp, int nb, double *out, void *ws,
                       size_t ws_bytes, hipStream_t s) {
    if (nb < 1 || nb > 32 || !out || (enc != 0 && enc != 1)) { set_error("level errors: bad arguments"); return kErrArg; }
    if (n > 0 && !x) { set_error("level errors: null input"); return kErrArg; }
    const size_t pb = fe_partial_bytes(n, nb), ub = fe_under_bytes(nb);
    if (ws_bytes < pb + ub || !ws) { set_error("level errors: workspace too small"); return kErrArg; }
    const int e = enc == 0 ? exp + 2 : exp;  // NegaBinaryBPEncoder.hpp:68
    uint64_t *partial = static_cast<uint64_t *>(ws);
    unsigned long long *under = reinterpret_cast<unsigned long long *>(static_cast<char *>(ws) + pb);
    const uint32_t nwg = fe_nwg(n);
    int rc = launch_fill_u64(reinterpret_cast<uint64_t *>(under), ub / 8, 0ull, s);
    if (!rc && n == 0) rc = launch_fill_u64(partial, pb / 8, 0ull, s);
    if (rc) return rc;
    if (n) {
        if (enc == 0)
            hipLaunchKernelGGL(k_fused_errors<0>, dim3(nwg), dim3(kFeThreads), 0, s, x, n, nb - e, nb, partial, under);
        else
            hipLaunchKernelGGL(k_fused_errors<1>, dim3(nwg), dim3(kFeThreads), 0, s, x, n, nb - e, nb, partial, under);
        PRODM_CHECK_LAUNCH();
    }
    std::vector<uint64_t> host((pb + ub) / 8);
    PRODM_HIP(hipMemcpyAsync(host.data(), ws, pb + ub, hipMemcpyDeviceToHost, s));
    PRODM_HIP(hipStreamSynchronize(s));
    const uint64_t *hp = host.data(), *hu = host.data() + pb / 8;
    const int wshift = kFeLog2Q - (-1074);
    const size_t limbs = (size_t)(2048 + 128 + 64 * kFeLimbs + 64) / 64 + 2;
    for (int k = 0; k <= nb; k++) {
        Big b(limbs);
        for (uint32_t g = 0; g < nwg; g++)
            for (int j = 0; j < kFeLimbs; j++) b.add_shifted(hp[((uint64_t)g * (nb + 1) + k) * kFeLimbs + j], wshift + 64 * j);
        for (int slot = 1; slot < kFeExpBuckets; slot++) {
            b.add_shifted(hu[((size_t)k * kFeExpBuckets + slot) * 2], slot - 1);
            b.add_shifted(hu[((size_t)k * kFeExpBuckets + slot) * 2 + 1], slot - 1 + 64);
        }
        out[k] = std::ldexp(b.to_double(-1074), 2 * (e - nb));
    }
    return kOk;
}

}  // namespace prodm

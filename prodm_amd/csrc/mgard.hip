// mgard.hip -- MGARD orthogonal decompose / recompose / adjoint for gfx950.
//
// One level step on a fine box n with coarse box c = (n >> 1) + 1 per dim:
//   W   = u - I u_c at every node with an odd index (I: multilinear interp.)
//   f   = (R M)_0 (R M)_1 (R M)_2 W                       (FP64 load vector)
//   z   = M_c,0^-1 M_c,1^-1 M_c,2^-1 f                    (FP64 line solves)
//   u_c' = u_c + z                                         (coarse nodal values)
// recompose inverts it; the adjoint is its transpose (see k_adjoint_fine).
// The arithmetic is restated from Ainsworth, Tugluk, Whitney, Klasky
// (SIAM J. Sci. Comput. 2019); ProDM calls it through MGARDx
// (Decomposer/MGARD.hpp:14-35).
//
// Kernels per level:
//   k_load    fine sweep, marching along dim 0 with a (dim1, dim2) tile in LDS:
//             W, coefficient stores, u_c, max |W| and the 3-D load stencil.
//   k_solve_contig   line solves along dim 2 (contiguous): two lines per wave,
//             a segment of each per lane, affine-map scan across lanes.
//   k_solve_strided  line solves along dims 1 / 0: lane per line (coalesced),
//             waves split the line into segments, LDS scan across waves;
//             the last solve applies u_c +/- z.
//   k_interp  recompose: u = I u_c + W.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

namespace prodm {

constexpr int kLoadThreads = 256;

enum LoadMode : int { kDecompose = 0, kRecompose = 1, kAdjointPT = 2 };

__host__ __device__ inline double rweight(int j, int k, uint32_t c) {
    if (c == 1) return k == 0 ? 1.0 : 0.0;
    if (k == 0) return (j == 0 || j == (int)c - 1) ? 1.25 : 2.5;
    if (k == 1 || k == -1) return 1.5;
    return 0.25;
}
// transpose of the interpolation along one dim: h_j = g_2j + (g_2j-1 + g_2j+1)/2
__host__ __device__ inline double pweight(int j, int k, uint32_t c) {
    if (c == 1) return k == 0 ? 1.0 : 0.0;
    if (k == 0) return 1.0;
    if (k == 1 || k == -1) return 0.5;
    return 0.0;
}

// Row pitch (elements) of the FP64 load vector on a coarse grid: rows start
// on 128-byte boundaries, so the line solves' 64-lane accesses are aligned.
__host__ __device__ inline uint32_t f_pitch(uint32_t c2) { return (c2 + 15u) & ~15u; }

template <class T> struct LoadParams {
    uint32_t n[3], c[3];
    const T *in;           // natural-order input box (decompose / adjoint)
    uint64_t is[3];        // its strides
    T *coef_full;          // reordered full-array coefficients
    uint64_t cs[3];
    T *coef_ilv;           // interleaved level buffer
    uint32_t ilv_c[3];     // coarse dims used by the interleave index
    T *uc;                 // decompose: coarse nodal values (dense c)
    double *f;             // load vector (c0 x c1 rows of pitch fp)
    uint32_t fp;           // row pitch of f
    KeyOf<T> *key;         // decompose: level max |W|
    int march;             // coarse planes per workgroup along dim 0
    int hb;                // decompose: hierarchical basis (no load vector)
};

template <class T>
__device__ inline uint64_t coef_index_full(const LoadParams<T> &P, uint32_t i0, uint32_t i1, uint32_t i2) {
    return rpos(i0, P.c[0]) * P.cs[0] + rpos(i1, P.c[1]) * P.cs[1] + rpos(i2, P.c[2]);
}
template <class T>
__device__ inline uint64_t coef_index_ilv(const LoadParams<T> &P, uint32_t i0, uint32_t i1, uint32_t i2) {
    // ilv_index (DirectInterleaver order) with scalar operands: taking the
    // address of the by-value kernel argument would force a scratch copy
    const uint32_t p0 = rpos(i0, P.c[0]), p1 = rpos(i1, P.c[1]), p2 = rpos(i2, P.c[2]);
    const uint32_t n1 = P.n[1], n2 = P.n[2], c0 = P.ilv_c[0], c1 = P.ilv_c[1], c2 = P.ilv_c[2];
    const uint64_t n12 = (uint64_t)n1 * n2;
    const uint64_t plane_in = n12 - (uint64_t)c1 * c2;
    if (p0 < c0) {
        const uint64_t base = (uint64_t)p0 * plane_in;
        if (p1 < c1) return base + (uint64_t)p1 * (n2 - c2) + (p2 - c2);
        return base + (uint64_t)c1 * (n2 - c2) + (uint64_t)(p1 - c1) * n2 + p2;
    }
    return (uint64_t)c0 * plane_in + (uint64_t)(p0 - c0) * n12 + (uint64_t)p1 * n2 + p2;
}

struct DimTile {
    int js, tj, rs, R;  // coarse start, coarse count, fine region start, region length
};
__device__ inline DimTile dim_tile(uint32_t tile, int TJ, uint32_t c, uint32_t n) {
    DimTile t;
    if (n == 1) { t.js = 0; t.tj = 1; t.rs = 0; t.R = 1; return t; }
    t.js = (int)tile * TJ;
    t.tj = min(TJ, (int)c - t.js);
    t.rs = 2 * t.js - 2;
    t.R = 2 * t.tj + 3;
    return t;
}

// One workgroup per coarse tile: kM1 rows (dim 1) x kM2 columns (dim 2),
// marching along dim 0 over kM0 coarse planes. Each fine plane of the tile
// plus a 2-node halo (kN1 x kN2, compile-time) lives in LDS; the next pair of
// planes is prefetched into registers while the current pair is processed.
// Per fine plane: W (in LDS), coefficient / coarse-value / max stores, the
// load stencil along dim 2 (LDS, FP64) and dim 1 (registers, one coarse
// column per thread), and the dim-0 stencil as a running sum over planes.
// Tile shapes: coarse extents are 2^k + 1, so 32-wide tiles would leave a
// one-column tile at the end of every dim (257 = 8 * 32 + 1).
//   SweepStd: 7 x 33 coarse nodes, 256 threads; the 17 x 69 region fits
//     five elements per thread.
//   SweepWide: 7 x 65 coarse nodes, 512 threads (the fp32 decompose sweep of
//     levels with many tiles). The same five elements per thread and waves
//     per SIMD (two workgroups of 54 KB LDS per CU); each fine row is stored
//     in runs of 65 coefficients instead of 33, and the tile re-reads less
//     halo. Finest sweep at 513^3 0.378 -> 0.366 ms. The recompose sweep and
//     the smaller levels (fewer, larger tiles) measured slower with it.
template <int M1, int M2, int THREADS> struct SweepShape {
    static constexpr int kM1 = M1, kM2 = M2, kThreads = THREADS;
    static constexpr int kN1 = 2 * M1 + 3, kN2 = 2 * M2 + 3, kPL = kN1 * kN2;
    static constexpr int kPER = (kPL + THREADS - 1) / THREADS;  // plane elements per thread
    static constexpr int kG2 = kN1 * M2;
    static constexpr int kG2PER = (kG2 + THREADS - 1) / THREADS;
    static_assert(THREADS % 64 == 0 && M1 * M2 <= THREADS, "a thread per coarse column");
    static_assert((THREADS / M2) * kG2PER >= kN1, "the dim-2 stencil rows cover the region");
};
using SweepStd = SweepShape<7, 33, 256>;
using SweepWide = SweepShape<7, 65, 512>;

__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    // workgroups are dispatched round-robin over the 8 XCDs: give XCD x the
    // contiguous range [x q, (x + 1) q) of the first 8 q tiles (a bijection;
    // the remaining tail keeps its index)
    const uint32_t q = nblocks / 8;
    if (bid >= 8 * q) return bid;
    return (bid % 8) * q + bid / 8;
}

// Tiles along dim 0: floor(c0 / march), the last one also taking the
// remainder. Coarse extents are 2^k + 1, so rounding up would leave a last
// tile of a single plane (257 = 16 x 16 + 1) that costs a whole tile's set-up;
// here the last tile marches 17 planes instead. Which workgroup owns a coarse
// plane does not change any result.
__host__ __device__ inline uint32_t march_tiles0(uint32_t n0, uint32_t c0, int march) {
    if (n0 == 1) return 1;
    const uint32_t t = c0 / (uint32_t)march;
    return t ? t : 1;
}

struct MarchTile {
    int js0, tj0, js1, tj1, js2, tj2;  // coarse tile
    int rs1, rs2;                      // fine region start (dims 1, 2)
};
template <class SH>
__device__ inline MarchTile march_tile(uint32_t bt, uint32_t n0, uint32_t n1, uint32_t n2, uint32_t c0, uint32_t c1,
                                       uint32_t c2, int kM0) {
    constexpr int kM1 = SH::kM1, kM2 = SH::kM2;
    const uint32_t n[3] = {n0, n1, n2}, c[3] = {c0, c1, c2};
    const uint32_t t2n = (n[2] == 1) ? 1 : (c[2] + kM2 - 1) / kM2;
    const uint32_t t1n = (n[1] == 1) ? 1 : (c[1] + kM1 - 1) / kM1;
    const uint32_t a2 = bt % t2n, a1 = (bt / t2n) % t1n, a0 = bt / (t2n * t1n);
    MarchTile t;
    t.js2 = (n[2] == 1) ? 0 : (int)a2 * kM2;
    t.tj2 = (n[2] == 1) ? 1 : min(kM2, (int)c[2] - t.js2);
    t.rs2 = (n[2] == 1) ? 0 : 2 * t.js2 - 2;
    t.js1 = (n[1] == 1) ? 0 : (int)a1 * kM1;
    t.tj1 = (n[1] == 1) ? 1 : min(kM1, (int)c[1] - t.js1);
    t.rs1 = (n[1] == 1) ? 0 : 2 * t.js1 - 2;
    t.js0 = (n[0] == 1) ? 0 : (int)a0 * kM0;
    t.tj0 = (n[0] == 1) ? 1 : (a0 + 1 == march_tiles0(n[0], c[0], kM0) ? (int)c[0] - t.js0 : kM0);
    return t;
}

// workgroups (coarse tiles) of one load sweep
template <class T, class SH = SweepStd> __host__ __device__ inline uint32_t load_tiles(const LoadParams<T> &P) {
    const uint32_t t2n = (P.n[2] == 1) ? 1 : (P.c[2] + SH::kM2 - 1) / SH::kM2;
    const uint32_t t1n = (P.n[1] == 1) ? 1 : (P.c[1] + SH::kM1 - 1) / SH::kM1;
    const uint32_t t0n = march_tiles0(P.n[0], P.c[0], P.march);
    return t0n * t1n * t2n;
}

// LDS of a sweep: staged fine planes, W planes and dim-2 stencils. The
// kernels declare it, so that the sweep's forms in one kernel share it.
template <class T, int MODE, class SH = SweepStd> struct SweepLds {
    static constexpr bool kDec = MODE == kDecompose;
    static constexpr int kPL = SH::kPL, kG2 = SH::kG2;
    static constexpr int kPE = kDec ? kPL : 1;  // input planes staged only by decompose
    T E[kPE], O[kPE], Wb[kPL], Wb2[kPE];
    double G2[kG2], G2b[kDec ? kG2 : 1];
};

// the sweep of coarse tile bt (one workgroup). A3: every dim of the level is
// longer than 1 (the tests of inactive dims fold away; same arithmetic).
// LOOPED: called in a loop over tiles (grouped launches); the thread index
// is then opaque to the compiler, which would otherwise hoist the per-slot
// set-up out of the loop and keep it live across tiles (fp32: 159 instead
// of 103 VGPRs, three waves per SIMD instead of four)
template <class T, int MODE, bool ILV, bool A3 = false, class SH = SweepStd, bool LOOPED = false>
__device__ __forceinline__ void load_tile(const LoadParams<T> &P, uint32_t bt, SweepLds<T, MODE, SH> &lds) {
    using K = KeyOf<T>;
    constexpr int kM2 = SH::kM2, kN1 = SH::kN1, kN2 = SH::kN2, kPL = SH::kPL, kPER = SH::kPER;
    constexpr int kG2 = SH::kG2, kG2PER = SH::kG2PER, kThreads = SH::kThreads;
    // buffer accesses through global-address-space pointers: the parameters
    // may come from memory (grouped launches), where the compiler cannot
    // infer the address space and would emit flat_* operations
    using GT = PRODM_GLOBAL T;
    constexpr bool kDec = MODE == kDecompose;
    T *const E = lds.E, *const O = lds.O, *const Wb = lds.Wb, *const Wb2 = lds.Wb2;
    double *const G2 = lds.G2, *const G2b = lds.G2b;
    int tid = threadIdx.x;
    if constexpr (LOOPED) asm volatile("" : "+v"(tid));
    const uint32_t n0 = P.n[0], n1 = P.n[1], n2 = P.n[2], c0 = P.c[0], c1 = P.c[1], c2 = P.c[2];
    const MarchTile M = march_tile<SH>(bt, n0, n1, n2, c0, c1, c2, P.march);
    const bool act0 = A3 || n0 > 1, act1 = A3 || n1 > 1, act2 = A3 || n2 > 1;

    // ---- per-slot constants: slot m of this thread is plane element
    // idx = tid + kThreads m at (r1, r2) of the (kN1 x kN2) region; everything
    // that depends only on (i1, i2) is computed once here, so that a plane
    // step costs a scalar plane offset plus one add per access.
    // offsets are in bytes and 32-bit, so that every access is a scalar
    // plane pointer plus one 32-bit lane offset
    uint32_t in_off[kPER];  // decompose / adjoint: byte offset in an input plane
    uint32_t offA[kPER];    // coefficient byte offset in an even plane (coarse slots: u_c offset)
    uint32_t offB[kPER];    // coefficient byte offset in an odd plane
    uint32_t bits = 0;      // 4 bits per slot: valid, odd i1, odd i2, owned (i1, i2)
    const uint32_t n12 = n1 * n2, plane_in = n12 - c1 * c2;
#pragma unroll
    for (int m = 0; m < kPER; m++) {
        const int idx = tid + kThreads * m;
        const int r1 = idx / kN2, r2 = idx - r1 * kN2;
        const int i1 = M.rs1 + r1, i2 = M.rs2 + r2;
        const bool valid = idx < kPL && i1 >= 0 && i1 < (int)n1 && i2 >= 0 && i2 < (int)n2;
        const int o1 = valid ? (r1 & 1) : 0, o2 = valid ? (r2 & 1) : 0;  // region starts are even
        const bool own1 = !act1 || (i1 >= 2 * M.js1 - 1 && i1 < 2 * (M.js1 + M.tj1) - 1);
        const bool own2 = !act2 || (i2 >= 2 * M.js2 - 1 && i2 < 2 * (M.js2 + M.tj2) - 1);
        bits |= (uint32_t)(valid | (o1 << 1) | (o2 << 2) | ((own1 && own2) << 3)) << (4 * m);
        const uint32_t u1 = valid ? (uint32_t)i1 : 0, u2 = valid ? (uint32_t)i2 : 0;
        in_off[m] = valid ? (uint32_t)(u1 * (uint32_t)P.is[1] + u2) * (uint32_t)sizeof(T) : kOffNone;
        const uint32_t p1 = rpos(u1, c1), p2 = rpos(u2, c2);
        uint32_t a, bb;
        if (ILV) {
            if (!o1 && !o2) a = (u1 >> 1) * c2 + (u2 >> 1);  // coarse in even planes
            else if (p1 < c1) a = p1 * (n2 - c2) + (p2 - c2);
            else a = c1 * (n2 - c2) + (p1 - c1) * n2 + p2;
            bb = p1 * n2 + p2;
        } else {
            a = (!o1 && !o2) ? (u1 >> 1) * c2 + (u2 >> 1) : p1 * (uint32_t)P.cs[1] + p2;
            bb = p1 * (uint32_t)P.cs[1] + p2;
        }
        offA[m] = a * (uint32_t)sizeof(T);
        offB[m] = bb * (uint32_t)sizeof(T);
    }
    // scalar coefficient base of fine plane i0 (even planes come first in a level)
    auto coef_base = [&](int i0) __attribute__((always_inline)) -> uint64_t {
        if (ILV) {
            if (!act0 || !(i0 & 1)) return (uint64_t)(i0 >> 1) * plane_in;
            return (uint64_t)c0 * plane_in + (uint64_t)(i0 >> 1) * n12;
        }
        return (uint64_t)rpos((uint32_t)i0, c0) * P.cs[0];
    };

    // fine plane i0 into registers; slots outside the box read 0
    auto fetch = [&](T (&reg)[kPER], int i0) __attribute__((always_inline)) {
        const bool ok = i0 >= 0 && i0 < (int)n0;
        if constexpr (MODE != kRecompose) {
            // the input plane (invalid slots carry kOffNone)
            const auto r = buf_rsrc((const void *)((const GT *)P.in + (uint64_t)(ok ? i0 : 0) * P.is[0]), ok);
#pragma unroll
            for (int m = 0; m < kPER; m++) reg[m] = buf_load<T>(r, in_off[m]);
        } else {
            // W from the coefficients: odd planes hold every node, even
            // planes all but the coarse ones (W = 0 there). Predicated
            // global loads: measured faster here than buffer loads with
            // masked offsets.
            const bool odd0 = act0 && (i0 & 1);
            const char *pb = (const char *)((const GT *)(ILV ? P.coef_ilv : P.coef_full) + coef_base(i0));
#pragma unroll
            for (int m = 0; m < kPER; m++) {
                const uint32_t b = bits >> (4 * m);
                T v = (T)0;
                if (ok && (b & 1) && (odd0 || (b & 6)))
                    v = *(const GT *)((const PRODM_GLOBAL char *)pb + (odd0 ? offB[m] : offA[m]));
                reg[m] = v;
            }
        }
    };
    auto put = [&](T *dst, const T (&reg)[kPER]) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < kPER; m++) {
            const int idx = tid + kThreads * m;
            if (kThreads * (m + 1) <= kPL || idx < kPL) dst[idx] = reg[m];
        }
    };

    const int cj1 = tid / kM2, cj2 = tid % kM2;  // this thread's coarse column
    const bool has_col = cj1 < M.tj1 && cj2 < M.tj2;
    const int j1 = M.js1 + cj1, j2 = M.js2 + cj2;
    double ra = 0.0, rb = 0.0, rc = 0.0;  // running dim-0 sums (see column)
    T lmax = 0;

    // load-stencil weights along one dim: restriction of the mass matrix,
    // 3/h-scaled, [1/4, 3/2, 5/2, 3/2, 1/4] (centre 5/4 at the two ends);
    // the adjoint uses the interpolation transpose [0, 1/2, 1, 1/2, 0].
    // Taps outside the grid read zeros from the staged planes.
    constexpr double kW2 = MODE == kAdjointPT ? 0.0 : 0.25;
    constexpr double kW1 = MODE == kAdjointPT ? 0.5 : 1.5;
    auto centre = [](int j, uint32_t c) __attribute__((always_inline)) -> double {
        if (MODE == kAdjointPT) return 1.0;
        return (j == 0 || j == (int)c - 1) ? 1.25 : 2.5;
    };
    const double w0_2 = centre(M.js2 + cj2, c2), w0_1 = centre(j1, c1);

    // W of the fine planes [decompose]. The multilinear interpolant of a node
    // is the tree sum of its cell's 8 corners (an even dim repeating the same
    // corner) times 1/8. Per plane the 4-corner part is
    //   S = (P[q] + P[q + d2]) + (P[q + d1] + P[q + d1 + d2]),
    // so on an even plane (both dim-0 corners in the plane itself) the tree
    // is S + S = 2 S exactly and I = S / 4, and on an odd plane I = (S of the
    // plane below + S of the plane above) / 8: the same operations in the
    // same order, with each even plane's S computed once and kept per slot.
    auto own_plane = [&](int i0) __attribute__((always_inline)) -> bool {
        return !act0 || (i0 >= 2 * M.js0 - 1 && i0 < 2 * (M.js0 + M.tj0) - 1);
    };
    // fmax returns the other operand when one is NaN (IEEE maxNum), so NaN
    // coefficients never enter the level max
    auto keep_max = [&](bool st, T w) __attribute__((always_inline)) { lmax = fmax(lmax, st ? fabs(w) : (T)0); };
    // even plane i0 staged in Ev: S per slot into s, W into Wout, coarse
    // values into u_c and coefficients into the level buffer
    auto even_plane = [&](int i0, const T *Ev, T *Wout, T (&s)[kPER]) __attribute__((always_inline)) {
        const bool own0 = own_plane(i0);
        // planes of the halo store nothing (range 0)
        const auto rc = buf_rsrc((const void *)((GT *)(ILV ? P.coef_ilv : P.coef_full) + coef_base(i0)), own0);
        const auto ru = buf_rsrc((const void *)((GT *)P.uc + (uint64_t)(i0 >> 1) * c1 * c2), own0);
#pragma unroll
        for (int m = 0; m < kPER; m++) {
            const int idx = tid + kThreads * m;
            if (kThreads * (m + 1) > kPL && idx >= kPL) continue;
            const uint32_t b = bits >> (4 * m);
            const int d1 = (b & 2) ? 2 * kN2 : 0, d2 = (b & 4) ? 2 : 0;
            const int q = idx - (d1 >> 1) - (d2 >> 1);  // LDS offset of the low interpolation corner
            const T S = (Ev[q] + Ev[q + d2]) + (Ev[q + d1] + Ev[q + d1 + d2]);
            s[m] = S;
            const T cv = Ev[idx];
            const bool coarse = !(b & 6);
            const T w = ((b & 1) && !coarse) ? cv - S * (T)0.25 : (T)0;
            Wout[idx] = w;
            // the coarse value into u_c or the coefficient into the level
            // buffer; slots not stored by this tile take kOffNone
            const bool st = (b & 1) && (b & 8);
            const uint32_t oa = offA[m];
            buf_store<T>(rc, st && !coarse ? oa : kOffNone, w);
            buf_store<T>(ru, st && coarse ? oa : kOffNone, cv);
            keep_max(st && own0 && !coarse, w);
        }
    };
    // odd plane i0 staged in Ov, between even planes with sums sa and sb
    auto odd_plane = [&](int i0, const T *Ov, T *Wout, const T (&sa)[kPER], const T (&sb)[kPER])
        __attribute__((always_inline)) {
        const bool own0 = own_plane(i0);
        const auto rc = buf_rsrc((const void *)((GT *)(ILV ? P.coef_ilv : P.coef_full) + coef_base(i0)), own0);
#pragma unroll
        for (int m = 0; m < kPER; m++) {
            const int idx = tid + kThreads * m;
            if (kThreads * (m + 1) > kPL && idx >= kPL) continue;
            const uint32_t b = bits >> (4 * m);
            const T I = (sa[m] + sb[m]) * (T)0.125;
            const T ov = Ov[idx];  // every slot below kPL holds a staged value
            const T w = (b & 1) ? ov - I : (T)0;
            Wout[idx] = w;
            const bool st = (b & 1) && (b & 8);
            buf_store<T>(rc, st ? offB[m] : kOffNone, w);
            keep_max(st && own0, w);
        }
    };
    // load stencil along dim 2 from Wb into G2[r1][jj2]; this thread's jj2
    // is fixed (tid % kM2), its rows step by kThreads / kM2
    auto stencil2 = [&](const T *Wsrc, double *G) __attribute__((always_inline)) {
        constexpr int kRowStep = kThreads / kM2;
        if (cj2 >= M.tj2 || cj1 >= kRowStep) return;
#pragma unroll
        for (int m = 0; m < kG2PER; m++) {
            const int r1 = cj1 + kRowStep * m;
            if (kThreads * (m + 1) > kG2 && r1 >= kN1) continue;
            const T *row = Wsrc + r1 * kN2;
            double a;
            if (!act2) {
                a = (double)row[0];
            } else {
                const T *t = row + 2 * cj2;  // taps i2 = 2 jj - 2 .. 2 jj + 2
                a = kW2 * (double)t[0];
                a = fma(kW1, (double)t[1], a);
                a = fma(w0_2, (double)t[2], a);
                a = fma(kW1, (double)t[3], a);
                a = fma(kW2, (double)t[4], a);
            }
            G[r1 * kM2 + cj2] = a;
        }
    };
    // dim 1 for this thread's column (rows 2 cj1 .. 2 cj1 + 4 of G2), then
    // the running dim-0 sums. Fine plane i0 contributes to coarse planes j0
    // with |2 j0 - i0| <= 2, in increasing i0. Registers: ra = coarse plane
    // m-1, rb = m, rc = m+1 for the current even plane 2m; after it, m-1 is
    // complete and the ring rotates.
    auto dim1 = [&](const double *G) __attribute__((always_inline)) -> double {
        if (!act1) return G[cj2];
        const double *g = G + 2 * cj1 * kM2 + cj2;
        double h = kW2 * g[0];
        h = fma(kW1, g[kM2], h);
        h = fma(w0_1, g[2 * kM2], h);
        h = fma(kW1, g[3 * kM2], h);
        h = fma(kW2, g[4 * kM2], h);
        return h;
    };
    auto wt0 = [&](int j, int k) __attribute__((always_inline)) -> double {
        if (k == 0) return centre(j, c0);
        if (k == 1 || k == -1) return kW1;
        return kW2;
    };
    auto in_tile0 = [&](int j0) __attribute__((always_inline)) -> bool { return j0 >= M.js0 && j0 < M.js0 + M.tj0; };
    auto store_f = [&](int j0, double v) __attribute__((always_inline)) {
        ((PRODM_GLOBAL double *)P.f)[((uint64_t)j0 * c1 + j1) * P.fp + j2] = v;
    };
    auto column = [&](int i0, const double *G) __attribute__((always_inline)) {
        if (!has_col) return;
        const double h = dim1(G);
        if (!act0) {
            store_f(0, h);
            return;
        }
        const int m = i0 >> 1;
        if (i0 & 1) {
            // odd plane 2m+1 (after the rotation: ra = m, rb = m+1)
            if (in_tile0(m)) ra = fma(wt0(m, 1), h, ra);
            if (in_tile0(m + 1)) rb = fma(wt0(m + 1, -1), h, rb);
        } else {
            // even plane 2m (ra = m-1, rb = m, rc = m+1)
            if (m >= 1 && in_tile0(m - 1)) ra = fma(wt0(m - 1, 2), h, ra);
            if (in_tile0(m)) rb = fma(wt0(m, 0), h, rb);
            if (2 * m + 2 < (int)n0 + 1 && in_tile0(m + 1)) rc = fma(wt0(m + 1, -2), h, rc);
            if (m >= 1 && in_tile0(m - 1)) store_f(m - 1, ra);
            if (i0 == (int)n0 - 1 && in_tile0(m)) store_f(m, rb);
            ra = rb;
            rb = rc;
            rc = 0.0;
        }
    };

    const int i0lo = act0 ? max(0, 2 * M.js0 - 2) : 0;
    const int i0hi = act0 ? min((int)n0 - 1, 2 * (M.js0 + M.tj0 - 1) + 2) : 0;
    T pa[kPER], pb[kPER];
    if constexpr (kDec) {
        // plane i0lo (even), then per step the odd plane i0 and the even
        // plane i0 + 1, two barriers per step. After the first: the columns
        // of the previous step (from G2 / G2b, in plane order), then this
        // step's W planes (E / O into Wb / Wb2). After the second: their
        // dim-2 stencils into G2 / G2b, and the next step's planes from
        // registers into E / O (their loads issued one step earlier).
        T sprev[kPER];
        fetch(pa, i0lo);
        put(E, pa);
        if (i0lo + 1 <= i0hi) { fetch(pa, i0lo + 1); fetch(pb, i0lo + 2); }
        __syncthreads();
        even_plane(i0lo, E, Wb, sprev);
        __syncthreads();
        if (!P.hb) stencil2(Wb, G2);
        if (i0lo + 1 <= i0hi) {
            put(O, pa);
            put(E, pb);
            if (i0lo + 3 <= i0hi) { fetch(pa, i0lo + 3); fetch(pb, i0lo + 4); }
        }
        __syncthreads();
        if (!P.hb) column(i0lo, G2);
#pragma unroll 1
        for (int i0 = i0lo + 1; i0 <= i0hi; i0 += 2) {
            T scur[kPER];
            even_plane(i0 + 1, E, Wb2, scur);
            odd_plane(i0, O, Wb, sprev, scur);
#pragma unroll
            for (int m = 0; m < kPER; m++) sprev[m] = scur[m];
            __syncthreads();
            if (!P.hb) {
                stencil2(Wb, G2);
                stencil2(Wb2, G2b);
            }
            if (i0 + 2 <= i0hi) {
                put(O, pa);
                put(E, pb);
                if (i0 + 4 <= i0hi) { fetch(pa, i0 + 4); fetch(pb, i0 + 5); }
            }
            __syncthreads();
            if (!P.hb) {
                column(i0, G2);
                column(i0 + 1, G2b);
            }
        }
        if (P.key) {
            K mk = abs_key(lmax);
            for (int o = 32; o >= 1; o >>= 1) {
                K q = __shfl_xor(mk, o, 64);
                mk = q > mk ? q : mk;
            }
            if (lane_id() == 0 && mk) atomicMax(P.key, mk);
        }
    } else {
        // recompose: W straight from the coefficients; adjoint: g
        fetch(pa, i0lo);
        for (int i0 = i0lo; i0 <= i0hi; i0++) {
            put(Wb, pa);
            if (i0 + 1 <= i0hi) fetch(pa, i0 + 1);
            __syncthreads();
            stencil2(Wb, G2);
            __syncthreads();
            column(i0, G2);
        }
    }
}

template <class T, int MODE, bool ILV, bool A3, class SH>
__global__ __launch_bounds__(SH::kThreads) void k_load(LoadParams<T> P) {
    __shared__ SweepLds<T, MODE, SH> lds;
    load_tile<T, MODE, ILV, A3, SH>(P, xcd_remap(blockIdx.x, load_tiles<T, SH>(P)), lds);
}

// ---- line solves -----------------------------------------------------------------
// Thomas factors (scaled coarse mass matrix tridiag(1, 4, 1), ends 2) for a
// line of m nodes, stored as inv[0..m) then cp[0..m).
static void thomas_factors(uint32_t m, double *inv, double *cp) {
    for (uint32_t i = 0; i < m; i++) {
        const double b = (i == 0 || i == m - 1) ? 2.0 : 4.0;
        const double w = (i == 0) ? b : b - cp[i - 1];
        inv[i] = 1.0 / w;
        cp[i] = inv[i];
    }
}

enum Epilogue : int { kEpiNone = 0, kEpiAdd = 1, kEpiSub = 2 };

// The factors of interior nodes depend only on the node index (the end
// diagonal 2 enters at node 0 and at the last node) and reach a bitwise
// fixed point after a few nodes: the kernels read the table below nfix and
// use lim for the interior nodes above it and last for node m - 1, the same
// values the table holds. cp[i] = inv[i] (the off-diagonal is 1).
template <class T> struct SolveParams {
    double *f;            // c0 x c1 rows of pitch fp
    uint32_t fp;
    uint32_t c[3];
    const double *inv;    // inv[0..m) of the solve dim
    double lim, last;     // inv[i] for nfix <= i < m - 1, inv[m - 1]
    uint32_t nfix;
    int epi;
    const T *uc_in;       // epilogue input (dense c)
    T *uc_out;            // epilogue output (dense c)
};
template <class T, class P>
__device__ __forceinline__ double thomas_factor(const SolveParams<T> &S, P inv, uint32_t m, uint32_t i) {
    return i < S.nfix ? inv[i] : (i == m - 1 ? S.last : S.lim);
}

// fidx: position in f (padded rows), uidx: position in the dense u_c
template <class T> __device__ inline void epilogue(const SolveParams<T> &S, uint64_t fidx, uint64_t uidx, double z) {
    using GT = PRODM_GLOBAL T;
    if (S.epi == kEpiAdd) ((GT *)S.uc_out)[uidx] = (T)((double)((const GT *)S.uc_in)[uidx] + z);
    else if (S.epi == kEpiSub) ((GT *)S.uc_out)[uidx] = (T)((double)((const GT *)S.uc_in)[uidx] - z);
    else ((PRODM_GLOBAL double *)S.f)[fidx] = z;
}

// contiguous lines (dim 2): a wave's lanes form groups of W (64, or 32 / 16
// for lines of at most 32 / 16 nodes, so that short lines do not leave most
// lanes idle); a group solves kContigLPW lines side by side (two independent
// chains interleave, hiding each other's latency), each lane holding a
// segment of at most SEG nodes of every line. Wave w of the workgroup solves
// the kContigLPW * 64 / W lines from line0 + w * kContigLPW * 64 / W, lane
// group g of it lines + g * kContigLPW + p. The scans stay inside a group: for
// a line that fits a narrower group, the steps a 64-wide scan adds combine
// only with identity elements of idle lanes, so the result is the same.
constexpr int kContigLPW = 2;
template <int W> constexpr uint32_t contig_lines_per_wave() { return kContigLPW * (64 / W); }
// Lines of 64-lane groups are staged through wave-private LDS rows: the
// wave reads and writes each line in 512-byte runs (a lane's segment of
// consecutive nodes would otherwise spread every load instruction over the
// whole line), and each lane moves its segment between LDS and registers.
// LDS operations of one wave complete in order, so the rows need no barrier.
// The kernel owns the rows (contig_rows_doubles each). The grouped kernels
// load directly (WPB 0): rows for their 8-wave workgroups cost more
// occupancy than the staging saves (batched refactor 1.819 -> 1.869 ms).
#ifndef PRODM_CONTIG_STAGE
#define PRODM_CONTIG_STAGE 1
#endif
template <int SEG, int W, int WPB> constexpr bool contig_staged() {
    return PRODM_CONTIG_STAGE && W == 64 && WPB > 0 && WPB * SEG <= 48;  // rows of at most 48 KB per workgroup
}
template <int SEG, int W, int WPB> constexpr int contig_rows_doubles() {
    return contig_staged<SEG, W, WPB>() ? WPB * kContigLPW * 64 * SEG : 0;
}
template <class T, int SEG, int W, int WPB>
__device__ __forceinline__ void solve_contig_lines(const SolveParams<T> &S, uint64_t nlines, uint64_t line0,
                                                   double *lds_rows) {
    constexpr int NL = kContigLPW;
    constexpr bool kStage = contig_staged<SEG, W, WPB>();
    static_assert(W == 16 || W == 32 || W == 64, "lane group width");
    const uint32_t m = S.c[2];
    const uint32_t lane = lane_id(), wv = threadIdx.x / 64;
    const uint32_t li = lane % W, grp = lane / W;
    const uint64_t wfirst = line0 + (uint64_t)wv * contig_lines_per_wave<W>();
    if (wfirst >= nlines) return;
    const uint64_t first = wfirst + (uint64_t)grp * NL;  // this lane group's first line
    const uint32_t seg = (m + W - 1) / W;
    const uint32_t b = li * seg;
    const uint32_t e = min(m, b + seg);
    // Slot i of the lane is node b + i when in[i]. The factors are loaded
    // once; slots past the segment compute on zeros and are discarded by
    // selects (no exec-mask branches), and their loads and stores carry
    // kOffNone. Lines past nlines get zero-range descriptors (W = 64) or
    // kOffNone offsets (narrower groups share one descriptor per wave).
    bool in[SEG];
    double iv[SEG];
    const auto rinv = buf_rsrc((const void *)S.inv, true);
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        const uint32_t k = b + i;
        in[i] = k < e;
        const double t = buf_load<double>(rinv, k < S.nfix && in[i] ? k * 8u : kOffNone);
        iv[i] = k < S.nfix ? t : (k == m - 1 ? S.last : S.lim);
    }
    auto fac = [&](int i) __attribute__((always_inline)) -> double { return iv[i]; };
    __amdgpu_buffer_rsrc_t rl[NL];
    uint32_t lo[NL];  // byte offset of line p inside its descriptor (kOffNone: past nlines)
#pragma unroll
    for (int p = 0; p < NL; p++) {
        const bool ok = first + p < nlines;
        if constexpr (W == 64) {
            rl[p] = buf_rsrc((const void *)((const PRODM_GLOBAL double *)S.f + (ok ? first + p : 0) * S.fp), ok);
            lo[p] = 0;
        } else {
            rl[p] = buf_rsrc((const void *)((const PRODM_GLOBAL double *)S.f + wfirst * S.fp), true);
            lo[p] = ok ? (uint32_t)((first + p - wfirst) * S.fp) * 8u : kOffNone;
        }
    }
    auto off = [&](int p, int i) __attribute__((always_inline)) -> uint32_t {
        return (in[i] && lo[p] != kOffNone) ? lo[p] + (b + i) * 8u : kOffNone;
    };
    double v[NL][SEG];
    if constexpr (kStage) {
        // whole line into this wave's LDS row (m <= 64 SEG), then the segments
#pragma unroll
        for (int p = 0; p < NL; p++) {
            double t[SEG];
#pragma unroll
            for (int i = 0; i < SEG; i++) {
                const uint32_t k = 64u * i + lane;
                t[i] = buf_load<double>(rl[p], k < m ? k * 8u : kOffNone);
            }
#pragma unroll
            for (int i = 0; i < SEG; i++) lds_rows[(wv * NL + p) * 64 * SEG + 64 * i + lane] = t[i];
        }
        wave_lds_sync();
#pragma unroll
        for (int p = 0; p < NL; p++) {
#pragma unroll
            for (int i = 0; i < SEG; i++) v[p][i] = in[i] ? lds_rows[(wv * NL + p) * 64 * SEG + b + i] : 0.0;
        }
    } else {
#pragma unroll
        for (int p = 0; p < NL; p++) {
#pragma unroll
            for (int i = 0; i < SEG; i++) v[p][i] = buf_load<double>(rl[p], off(p, i));
        }
    }
    // forward: y_i = (f_i - y_{i-1}) * inv_i  == affine map y_{i-1} -> y_i
    double A[NL], Bm[NL];
#pragma unroll
    for (int p = 0; p < NL; p++) { A[p] = 0.0; Bm[p] = 1.0; }
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        const double f = fac(i);
#pragma unroll
        for (int p = 0; p < NL; p++) {
            const double nA = (v[p][i] - A[p]) * f, nB = -Bm[p] * f;
            A[p] = in[i] ? nA : A[p];
            Bm[p] = in[i] ? nB : Bm[p];
        }
    }
    for (int o = 1; o < W; o <<= 1) {
#pragma unroll
        for (int p = 0; p < NL; p++) {
            const double pA = __shfl_up(A[p], o, W), pB = __shfl_up(Bm[p], o, W);
            if ((int)li >= o) { A[p] = A[p] + Bm[p] * pA; Bm[p] = Bm[p] * pB; }
        }
    }
    double y[NL];
#pragma unroll
    for (int p = 0; p < NL; p++) {
        y[p] = __shfl_up(A[p], 1, W);
        if (li == 0) y[p] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        const double f = fac(i);
#pragma unroll
        for (int p = 0; p < NL; p++) {
            const double ny = (v[p][i] - y[p]) * f;
            y[p] = in[i] ? ny : y[p];
            v[p][i] = ny;
        }
    }
    // backward: z_i = y_i - cp_i z_{i+1} (z_{m-1} = y_{m-1})
    double Cc[NL], Dm[NL];
#pragma unroll
    for (int p = 0; p < NL; p++) { Cc[p] = 0.0; Dm[p] = 1.0; }
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        const double c = (b + i == m - 1) ? 0.0 : fac(i);
#pragma unroll
        for (int p = 0; p < NL; p++) {
            const double nC = v[p][i] - c * Cc[p], nD = -c * Dm[p];
            Cc[p] = in[i] ? nC : Cc[p];
            Dm[p] = in[i] ? nD : Dm[p];
        }
    }
    for (int o = 1; o < W; o <<= 1) {
#pragma unroll
        for (int p = 0; p < NL; p++) {
            const double nC = __shfl_down(Cc[p], o, W), nD = __shfl_down(Dm[p], o, W);
            if (li + o < W) { Cc[p] = Cc[p] + Dm[p] * nC; Dm[p] = Dm[p] * nD; }
        }
    }
    double z[NL];
#pragma unroll
    for (int p = 0; p < NL; p++) {
        z[p] = __shfl_down(Cc[p], 1, W);
        if (li == W - 1) z[p] = 0.0;
    }
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        const double c = (b + i == m - 1) ? 0.0 : fac(i);
#pragma unroll
        for (int p = 0; p < NL; p++) {
            const double nz = v[p][i] - c * z[p];
            z[p] = in[i] ? nz : z[p];
            v[p][i] = nz;
        }
    }
    if (S.epi == kEpiNone) {
        if constexpr (kStage) {
            wave_lds_sync();  // every lane has read its segment
#pragma unroll
            for (int p = 0; p < NL; p++) {
#pragma unroll
                for (int i = 0; i < SEG; i++)
                    if (in[i]) lds_rows[(wv * NL + p) * 64 * SEG + b + i] = v[p][i];
            }
            wave_lds_sync();
#pragma unroll
            for (int p = 0; p < NL; p++) {
#pragma unroll
                for (int i = 0; i < SEG; i++) {
                    const uint32_t k = 64u * i + lane;
                    const double t = lds_rows[(wv * NL + p) * 64 * SEG + 64 * i + lane];
                    buf_store<double>(rl[p], k < m ? k * 8u : kOffNone, t);
                }
            }
        } else {
#pragma unroll
            for (int p = 0; p < NL; p++)
#pragma unroll
                for (int i = 0; i < SEG; i++) buf_store<double>(rl[p], off(p, i), v[p][i]);
        }
        return;
    }
#pragma unroll
    for (int p = 0; p < NL; p++) {
        if (first + p >= nlines) continue;
        const uint64_t line = first + p;
#pragma unroll
        for (int i = 0; i < SEG; i++)
            if (in[i]) epilogue(S, line * S.fp + b + i, line * m + b + i, v[p][i]);
    }
}
template <class T, int SEG, int W>
__global__ __launch_bounds__(256) void k_solve_contig(SolveParams<T> S, uint64_t nlines) {
    __shared__ double rows[contig_rows_doubles<SEG, W, 4>() > 0 ? contig_rows_doubles<SEG, W, 4>() : 1];
    solve_contig_lines<T, SEG, W, 4>(S, nlines, (uint64_t)blockIdx.x * 4 * contig_lines_per_wave<W>(), rows);
}

// contiguous-solve classes by line length: class c holds lines of up to
// kContigSeg[c] * kContigW[c] nodes (0: longer than the register-resident
// solve holds)
constexpr int kContigSeg[10] = {0, 1, 1, 1, 2, 4, 5, 8, 9, 16};
constexpr int kContigW[10] = {0, 16, 32, 64, 64, 64, 64, 64, 64, 64};
inline int contig_class(uint32_t m) {
    for (int c = 1; c < 10; c++)
        if (m <= (uint32_t)(kContigSeg[c] * kContigW[c])) return c;
    return 0;
}
inline uint32_t contig_class_lines_per_wave(int c) { return kContigLPW * (64 / kContigW[c]); }

// strided lines along dim d (1 or 0) of the coarse grid, lane per line.
// Lines are indexed over padded rows, q = r * fp + j2 (r = j0 for d = 1,
// j1 for d = 0); lanes with j2 >= c2 idle, so each wave's 64 accesses are
// one aligned 512-byte run of f. nw waves split the line into segments of
// at most SEG (nw is a function of m alone, so the scan order is fixed).
// A workgroup may hold several such groups (the grouped launches pack them):
// this wave is wave wv of the group whose waves are wbase .. wbase + nw - 1
// of the workgroup; a wave with gact false (past the packed groups) finds an
// empty segment and only joins the barriers. bid: the group's 64-line group.
template <class T, int SEG>
__device__ __forceinline__ void solve_strided_lines(const SolveParams<T> &S, uint32_t m, uint64_t nlines, int d,
                                                    uint32_t nw, uint32_t bid, uint32_t wv, uint32_t wbase,
                                                    bool gact) {
    __shared__ double sA[8][64], sB[8][64];
    const uint32_t lane = lane_id();
    const uint32_t w0 = wbase + wv;  // this wave's row of the scan buffers
    const uint64_t q = (uint64_t)bid * 64 + lane;
    const uint32_t fp = S.fp, c1 = S.c[1], c2 = S.c[2];
    const uint32_t j2 = (uint32_t)(q % fp);
    const uint64_t r = q / fp;
    const bool act = gact && q < nlines && j2 < c2;
    uint64_t base = 0, stride, ubase = 0, ustride;
    if (d == 1) {
        stride = fp;
        ustride = c2;
        if (act) { base = r * c1 * fp + j2; ubase = r * c1 * c2 + j2; }
    } else {
        stride = (uint64_t)c1 * fp;
        ustride = (uint64_t)c1 * c2;
        if (act) { base = r * fp + j2; ubase = r * c2 + j2; }
    }
    const PRODM_GLOBAL double *fg = (const PRODM_GLOBAL double *)S.f;
    const uint32_t seg = (m + nw - 1) / nw;
    // wave-uniform segment bounds: the factor loads become scalar loads
    const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)(wv * seg));
    const uint32_t e = gact ? (uint32_t)__builtin_amdgcn_readfirstlane((int)min(m, b + seg)) : b;
    double v[SEG];
#pragma unroll
    for (int i = 0; i < SEG; i++) v[i] = (act && b + i < e) ? fg[base + (uint64_t)(b + i) * stride] : 0.0;
    // The segment's factors, lane i holding node b + i's (and, for the
    // backward pass, 0 at node m - 1), read back per node with readlane: no
    // scalar table loads or index tests inside the recurrences.
    constexpr int NF = (SEG + 63) / 64;  // factor registers per lane
    double fl[NF], cl[NF];
#pragma unroll
    for (int r = 0; r < NF; r++) {
        const uint32_t nl = b + 64u * r + lane;
        const bool has = 64u * r + lane < (uint32_t)SEG && nl < e;
        fl[r] = has ? thomas_factor(S, (const PRODM_GLOBAL double *)S.inv, m, nl) : 0.0;
        cl[r] = nl == m - 1 ? 0.0 : fl[r];
    }
    auto fac = [&](int i) __attribute__((always_inline)) -> double { return readlane_f64(fl[i / 64], i % 64); };
    auto cfac = [&](int i) __attribute__((always_inline)) -> double { return readlane_f64(cl[i / 64], i % 64); };
    double A = 0.0, Bm = 1.0;
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (b + i < e) {
            const double iv = fac(i);
            A = (v[i] - A) * iv;
            Bm = -Bm * iv;
        }
    }
    sA[w0][lane] = A;
    sB[w0][lane] = Bm;
    __syncthreads();
    double y = 0.0;
    for (uint32_t w = 0; w < wv; w++) y = sA[wbase + w][lane] + sB[wbase + w][lane] * y;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (b + i < e) {
            y = (v[i] - y) * fac(i);
            v[i] = y;
        }
    }
    double C = 0.0, Dm = 1.0;
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        if (b + i < e) {
            const double c = cfac(i);
            C = v[i] - c * C;
            Dm = -c * Dm;
        }
    }
    sA[w0][lane] = C;
    sB[w0][lane] = Dm;
    __syncthreads();
    double z = 0.0;
    if (gact)
        for (int w = (int)nw - 1; w > (int)wv; w--) z = sA[wbase + w][lane] + sB[wbase + w][lane] * z;
    // each node's result is stored as soon as it is known
    if (S.epi == kEpiNone) {
#pragma unroll
        for (int i = SEG - 1; i >= 0; i--) {
            if (b + i < e) {
                z = v[i] - cfac(i) * z;
                if (act) ((PRODM_GLOBAL double *)S.f)[base + (uint64_t)(b + i) * stride] = z;
            }
        }
        return;
    }
    // With an epilogue, u_c is updated in place: a node's u_c load issued
    // after the previous node's store could not be moved above it (the same
    // buffer), and every node would wait a memory round trip inside the
    // recurrence. The loads run kAhead nodes ahead in a ring of registers:
    // node i's slot is refilled with node i - kAhead's original value, which
    // its own store overwrites only kAhead nodes later.
    using GT = PRODM_GLOBAL T;
    constexpr int kAhead = SEG < 8 ? SEG : 8;
    const GT *uin = (const GT *)S.uc_in;
    GT *uout = (GT *)S.uc_out;
    // unconditional loads (a lane or node outside the line reads element 0,
    // unused): a predicated load would sit in a branch that also takes its
    // value's wait
    auto uload = [&](int i) __attribute__((always_inline)) -> T {
        return uin[(act && b + i < e) ? ubase + (uint64_t)(b + i) * ustride : 0];
    };
    T ub[kAhead];
#pragma unroll
    for (int i = SEG - 1; i >= SEG - kAhead; i--) ub[i % kAhead] = uload(i);
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        const T u = ub[i % kAhead];
        if (i - kAhead >= 0) {
            ub[(i - kAhead) % kAhead] = uload(i - kAhead);
            // keep the load here: the scheduler would otherwise sink it to
            // its use, where its wait (vmcnt also counts the stores) would
            // again serialize the nodes
            __builtin_amdgcn_sched_barrier(0);
        }
        if (b + i < e) {
            z = v[i] - cfac(i) * z;
            if (act)
                uout[ubase + (uint64_t)(b + i) * ustride] = (T)(S.epi == kEpiAdd ? (double)u + z : (double)u - z);
        }
    }
}
template <class T, int SEG>
__global__ __launch_bounds__(512) void k_solve_strided(SolveParams<T> S, uint32_t m, uint64_t nlines, int d) {
    solve_strided_lines<T, SEG>(S, m, nlines, d, blockDim.x / 64, blockIdx.x, threadIdx.x / 64, 0, true);
}

// Lines longer than the register-resident solves hold (dim 2: more than
// 64 x 16 nodes; dims 1 / 0: more than 8 x 72): a lane per line runs the
// Thomas recurrences in node order, the forward sweep writing y in place
// and the backward sweep reading it back (two passes over the line instead
// of one). Lines of dims 1 / 0 are indexed as in the strided solve, so a
// wave's accesses are one 512-byte run; dim-2 lanes sit a row apart.
template <class T>
__global__ __launch_bounds__(256) void k_solve_seq(SolveParams<T> S, uint32_t m, uint64_t nlines, int d) {
    const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nlines) return;
    const uint64_t fp = S.fp, c1 = S.c[1], c2 = S.c[2];
    uint64_t base, stride, ubase, ustride;
    if (d == 2) {
        base = q * fp; stride = 1;
        ubase = q * c2; ustride = 1;
    } else {
        const uint64_t j2 = q % fp, r = q / fp;
        if (j2 >= c2) return;
        if (d == 1) {
            base = r * c1 * fp + j2; stride = fp;
            ubase = r * c1 * c2 + j2; ustride = c2;
        } else {
            base = r * fp + j2; stride = c1 * fp;
            ubase = r * c2 + j2; ustride = c1 * c2;
        }
    }
    PRODM_GLOBAL double *f = (PRODM_GLOBAL double *)S.f;
    const PRODM_CONST double *inv = (const PRODM_CONST double *)S.inv;
    constexpr uint32_t U = 8;  // loads issued ahead of the dependent chain
    double y = 0.0;
    uint32_t i = 0;
    for (; i + U <= m; i += U) {
        double t[U];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) t[k] = f[base + (i + k) * stride];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) {
            y = (t[k] - y) * thomas_factor(S, inv, m, i + k);
            f[base + (i + k) * stride] = y;
        }
    }
    for (; i < m; i++) {
        y = (f[base + i * stride] - y) * thomas_factor(S, inv, m, i);
        f[base + i * stride] = y;
    }
    double z = 0.0;
    uint32_t e = m;  // nodes [0, e) are left
    for (; e >= U; e -= U) {
        double t[U];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) t[k] = f[base + (e - 1 - k) * stride];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) {
            const uint32_t n = e - 1 - k;
            const double c = (n == m - 1) ? 0.0 : thomas_factor(S, inv, m, n);
            z = t[k] - c * z;
            epilogue(S, base + n * stride, ubase + n * ustride, z);
        }
    }
    for (; e > 0; e--) {
        const uint32_t n = e - 1;
        const double c = (n == m - 1) ? 0.0 : thomas_factor(S, inv, m, n);
        z = f[base + n * stride] - c * z;
        epilogue(S, base + n * stride, ubase + n * ustride, z);
    }
}

// ---- recompose interpolation: u = I u_c + W ------------------------------------
// One workgroup per fine tile 8 x 8 x 64; its coarse neighbourhood
// (5 x 5 x 33) is staged in LDS, W is read from the coefficients.
constexpr int kI0 = 8, kI1 = 8, kI2 = 64;
constexpr int kC0 = kI0 / 2 + 1, kC1 = kI1 / 2 + 1, kC2 = kI2 / 2 + 1;

// SUBZ: the level's solves ran ahead (the load vector P.f holds z on the
// coarse grid, rows of pitch P.fp) and the coarse values are u_c - z, formed
// here as each one is staged, as the last solve's epilogue forms them
template <class T, bool ILV, bool SUBZ = false>
__global__ __launch_bounds__(256) void k_interp(LoadParams<T> P, const T *__restrict__ uc, T *__restrict__ out,
                                                uint64_t os0, uint64_t os1) {
    __shared__ T C[kC0 * kC1 * kC2];
    const uint32_t t2n = (P.n[2] + kI2 - 1) / kI2, t1n = (P.n[1] + kI1 - 1) / kI1;
    const uint32_t t0n = (P.n[0] + kI0 - 1) / kI0;
    const uint32_t bt = xcd_remap(blockIdx.x, t0n * t1n * t2n);
    const uint32_t s2 = (bt % t2n) * kI2, s1 = ((bt / t2n) % t1n) * kI1, s0 = (bt / (t2n * t1n)) * kI0;
    const uint32_t *c = P.c;
    // coarse region [s/2, s/2 + C) clipped
    const uint32_t c0s = s0 >> 1, c1s = s1 >> 1, c2s = s2 >> 1;
    const uint32_t e0 = min(c[0] - c0s, (uint32_t)kC0), e1 = min(c[1] - c1s, (uint32_t)kC1),
                   e2 = min(c[2] - c2s, (uint32_t)kC2);
    // every load of the thread issued before its LDS writes (one latency,
    // not one per element)
    constexpr int kStage = (kC0 * kC1 * kC2 + 255) / 256;
    const uint32_t ne = e0 * e1 * e2;
    T sv[kStage];
    uint32_t sl[kStage];
#pragma unroll
    for (int k = 0; k < kStage; k++) {
        const uint32_t idx = threadIdx.x + 256u * k;
        const uint32_t a2 = idx % e2, a1 = (idx / e2) % e1, a0 = idx / (e2 * e1);
        sl[k] = (a0 * kC1 + a1) * kC2 + a2;
        sv[k] = idx < ne ? uc[((uint64_t)(c0s + a0) * c[1] + (c1s + a1)) * c[2] + (c2s + a2)] : (T)0;
        if (SUBZ && idx < ne)
            sv[k] = (T)((double)sv[k] -
                        ((const PRODM_GLOBAL double *)P.f)[((uint64_t)(c0s + a0) * c[1] + (c1s + a1)) * P.fp + (c2s + a2)]);
    }
#pragma unroll
    for (int k = 0; k < kStage; k++)
        if (threadIdx.x + 256u * k < ne) C[sl[k]] = sv[k];
    __syncthreads();
    // a wave handles rows of 64 along dim 2; row and plane offsets are
    // formed in 64 bits, lane offsets inside a row or plane in 32
    const uint32_t lane = threadIdx.x % kI2, wave = threadIdx.x / kI2;
    const uint32_t n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
    const uint32_t i2 = s2 + lane;
    const bool in2 = i2 < n2;
    const uint32_t o2 = (n2 == 1) ? 0 : (i2 & 1);
    const uint32_t p2 = rpos(i2, c[2]);
    const uint32_t n12 = n1 * n2, plane_in = n12 - c[1] * c[2];
    const uint32_t h0 = s0 >> 1;  // (i0 >> 1) of the tile's first plane
    const char *ce = nullptr, *co = nullptr;
    if (ILV) {
        ce = reinterpret_cast<const char *>(P.coef_ilv + (uint64_t)h0 * plane_in);  // even planes (p0 < c0)
        co = reinterpret_cast<const char *>(P.coef_ilv + (uint64_t)c[0] * plane_in + (uint64_t)h0 * n12);  // odd
    }
    char *ot = reinterpret_cast<char *>(out + (uint64_t)s0 * os0 + (uint64_t)s1 * os1 + s2);
    const uint32_t la2 = (i2 >> 1) - c2s, lb2 = la2 + o2;  // LDS columns of the two dim-2 corners
    constexpr int kWaves = 256 / kI2;
    constexpr int kRows = kI0 * kI1 / kWaves;  // rows per wave
    // tiles inside the grid (all but the last along each dim) skip every
    // bounds test; coarse lanes load a harmless in-buffer word and discard it
    const bool full = s0 + kI0 <= n0 && s1 + kI1 <= n1 && s2 + kI2 <= n2;
    auto body = [&](auto full_t) __attribute__((always_inline)) {
        constexpr bool FULL = decltype(full_t)::value;
        T wv[kRows];
#pragma unroll
        for (int q = 0; q < kRows; q++) {
            const uint32_t row = wave + q * kWaves;
            const uint32_t i1 = s1 + row % kI1, i0 = s0 + row / kI1;
            const bool odd0 = n0 > 1 && (i0 & 1), e1 = n1 == 1 || !(i1 & 1);
            const bool coarse = !odd0 && e1 && !o2;
            const bool ok = (FULL || (i0 < n0 && i1 < n1 && in2)) && !coarse;
            T v;
            if (ILV) {
                // DirectInterleaver order inside the level buffer: the plane
                // base in 64 bits, the offset inside the plane in 32
                const uint32_t k0 = (i0 >> 1) - h0, p1 = rpos(i1, c[1]);
                const uint32_t off = !odd0 ? (p1 < c[1] ? p1 * (n2 - c[2]) + (p2 - c[2])
                                                        : c[1] * (n2 - c[2]) + (p1 - c[1]) * n2 + p2)
                                           : p1 * n2 + p2;
                const char *src = !ok ? ce
                                      : odd0 ? co + (uint64_t)k0 * n12 * sizeof(T)
                                             : ce + (uint64_t)k0 * plane_in * sizeof(T);
                v = *reinterpret_cast<const T *>(src + (ok ? off * (uint32_t)sizeof(T) : 0u));
            } else {
                v = P.coef_full[ok ? coef_index_full(P, i0, i1, i2) : 0];
            }
            wv[q] = ok ? v : (T)0;
        }
#pragma unroll
        for (int q = 0; q < kRows; q++) {
            const uint32_t row = wave + q * kWaves;
            const uint32_t r1 = row % kI1, r0 = row / kI1;
            const uint32_t i1 = s1 + r1, i0 = s0 + r0;
            if (!FULL && (i0 >= n0 || i1 >= n1 || !in2)) continue;
            const uint32_t o0 = (n0 == 1) ? 0 : (i0 & 1), o1 = (n1 == 1) ? 0 : (i1 & 1);
            // 8 corners with repeats along even dims (exact doubling): equals
            // the per-case tree sum of the distinct coarse corners
            const uint32_t ra0 = ((i0 >> 1) - c0s) * kC1, rb0 = ra0 + o0 * kC1;
            const uint32_t ra1 = (i1 >> 1) - c1s, rb1 = ra1 + o1;
            const T *Caa = C + (ra0 + ra1) * kC2, *Cab = C + (ra0 + rb1) * kC2;
            const T *Cba = C + (rb0 + ra1) * kC2, *Cbb = C + (rb0 + rb1) * kC2;
            const T I = (((Caa[la2] + Caa[lb2]) + (Cab[la2] + Cab[lb2])) +
                         ((Cba[la2] + Cba[lb2]) + (Cbb[la2] + Cbb[lb2]))) *
                        (T)0.125;
            const T v = (o0 + o1 + o2 == 0) ? Caa[la2] : (T)(wv[q] + I);
            *reinterpret_cast<T *>(ot + ((uint64_t)r0 * os0 + (uint64_t)r1 * os1) * sizeof(T) + lane * (uint32_t)sizeof(T)) = v;
        }
    };
    if (full) body(std::true_type{});
    else body(std::false_type{});
}

// ---- box copies -----------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_copy_box(const T *__restrict__ src, uint64_t ss0, uint64_t ss1, T *__restrict__ dst,
                                                  uint64_t ds0, uint64_t ds1, uint32_t n2) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    dst[(uint64_t)blockIdx.z * ds0 + (uint64_t)blockIdx.y * ds1 + i2] =
        src[(uint64_t)blockIdx.z * ss0 + (uint64_t)blockIdx.y * ss1 + i2];
}
template <class T>
static int copy_box(const T *src, uint64_t ss0, uint64_t ss1, T *dst, uint64_t ds0, uint64_t ds1, const uint32_t *n,
                    hipStream_t s) {
    dim3 grid(div_up(n[2], 256), n[1], n[0]);
    hipLaunchKernelGGL(k_copy_box<T>, grid, dim3(256), 0, s, src, ss0, ss1, dst, ds0, ds1, n[2]);
    PRODM_CHECK_LAUNCH();
    return kOk;
}

// ---- grouped kernels (batched refactor) ---------------------------------------
// Each walks the virtual workgroups of a grouped launch (see GroupLaunch):
// the entry's parameters are read at wave-uniform addresses and the same tile /
// line body as the single-field kernel runs, so every field's results equal
// its single-field decompose bit for bit.
template <class T> struct SolveEntry {
    SolveParams<T> S;
    uint64_t nlines;
    uint32_t m, nw;   // line length, waves per line (strided)
    int32_t d, cls;   // solve dim, contiguous-solve class (dim 2)
};
template <class T> struct Level0Entry {
    const T *src;     // coarsest grid (strides s0, s1)
    uint64_t s0, s1;
    T *dst;           // dense level-0 buffer
    KeyOf<T> *key;
    uint32_t c0, c1, c2, pad;
};

static_assert(SweepStd::kThreads == kLoadThreads, "grouped sweeps run the standard tile shape");
template <class T>
__global__ __launch_bounds__(kLoadThreads) void k_load_grouped(const LoadParams<T> *__restrict__ ps,
                                                               const uint32_t *__restrict__ first, uint32_t nent,
                                                               uint32_t total) {
    __shared__ SweepLds<T, kDecompose> lds;
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t e = group_entry(first, nent, vb);
        const LoadParams<T> &P = ps[e];
        const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)first[e]);
        const uint32_t t = xcd_remap(vb - f0, load_tiles(P));
        if (P.n[0] > 1 && P.n[1] > 1 && P.n[2] > 1) load_tile<T, kDecompose, true, true, SweepStd, true>(P, t, lds);
        else load_tile<T, kDecompose, true, false, SweepStd, true>(P, t, lds);
        __syncthreads();  // LDS planes are reused by the next tile
    }
}

// each entry runs its own contiguous-solve class (lane group width and
// slots per lane), up to the launch's largest class MAXC
template <class T, int C>
__device__ __forceinline__ void contig_class_lines(const SolveEntry<T> &E, uint32_t wpb, uint32_t vbl, double *rows) {
    solve_contig_lines<T, kContigSeg[C], kContigW[C], 0>(
        E.S, E.nlines, (uint64_t)vbl * wpb * contig_lines_per_wave<kContigW[C]>(), rows);
}
template <class T, int C, int MAXC>
__device__ __forceinline__ void contig_dispatch(int c, const SolveEntry<T> &E, uint32_t wpb, uint32_t vbl,
                                                double *rows) {
    if constexpr (C < MAXC) {
        if (c == C) contig_class_lines<T, C>(E, wpb, vbl, rows);
        else contig_dispatch<T, C + 1, MAXC>(c, E, wpb, vbl, rows);
    } else {
        contig_class_lines<T, MAXC>(E, wpb, vbl, rows);  // c == MAXC: no entry exceeds the launch's class
    }
}
template <class T, int MAXC>
__global__ __launch_bounds__(512) void k_contig_grouped(const SolveEntry<T> *__restrict__ ps,
                                                        const uint32_t *__restrict__ first, uint32_t nent,
                                                        uint32_t total) {
    double *const rows = nullptr;  // unstaged
    const uint32_t wpb = blockDim.x / 64;
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t e = group_entry(first, nent, vb);
        const SolveEntry<T> E = uniform_load(ps + e);
        const uint32_t vbl = vb - (uint32_t)__builtin_amdgcn_readfirstlane((int)first[e]);
        contig_dispatch<T, 1, MAXC>(__builtin_amdgcn_readfirstlane(E.cls), E, wpb, vbl, rows);
    }
}

// A workgroup of 64 * WB threads holds floor(WB / nw) of an entry's line
// groups (nw waves each); the scan order of each is the single-field one.
template <class T, int SEG>
__global__ __launch_bounds__(512) void k_strided_grouped(const SolveEntry<T> *__restrict__ ps,
                                                         const uint32_t *__restrict__ first, uint32_t nent,
                                                         uint32_t total) {
    const uint32_t w = threadIdx.x / 64, wb = blockDim.x / 64;
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t e = group_entry(first, nent, vb);
        const SolveEntry<T> E = uniform_load(ps + e);
        const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)first[e]);
        const uint32_t nw = (uint32_t)__builtin_amdgcn_readfirstlane((int)E.nw);
        const uint32_t G = wb / nw, g = w / nw;
        solve_strided_lines<T, SEG>(E.S, E.m, E.nlines, E.d, nw, (vb - f0) * G + g, w - g * nw, g * nw, g < G);
        __syncthreads();  // the cross-wave scan buffers are reused
    }
}

// coarsest grid of every field: dense copy into the level-0 buffer and its
// max |value| (NaN excluded), as copy_box + k_maxabs do for one field
template <class T>
__global__ __launch_bounds__(256) void k_level0_grouped(const Level0Entry<T> *__restrict__ ps,
                                                        const uint32_t *__restrict__ first, uint32_t nent,
                                                        uint32_t total) {
    using K = KeyOf<T>;
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t e = group_entry(first, nent, vb);
        const Level0Entry<T> E = uniform_load(ps + e);
        const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)first[e]);
        const uint32_t idx = (vb - f0) * 256 + threadIdx.x;
        K m = 0;
        if (idx < E.c0 * E.c1 * E.c2) {
            const uint32_t i2 = idx % E.c2, i1 = (idx / E.c2) % E.c1, i0 = idx / (E.c2 * E.c1);
            const T v = ((const PRODM_GLOBAL T *)E.src)[(uint64_t)i0 * E.s0 + (uint64_t)i1 * E.s1 + i2];
            ((PRODM_GLOBAL T *)E.dst)[idx] = v;
            if (v == v) m = abs_key(v);
        }
        for (int o = 32; o >= 1; o >>= 1) {
            const K q = __shfl_xor(m, o, 64);
            m = q > m ? q : m;
        }
        if (lane_id() == 0 && m) atomicMax(E.key, m);
    }
}

// ---- host orchestration -----------------------------------------------------------
namespace {

uint64_t points3(const uint32_t *n) { return (uint64_t)n[0] * n[1] * n[2]; }

// factor tables for every (level, dim), packed: offsets[l][d] into the table
struct Factors {
    std::vector<double> host;
    uint64_t off[33][3];
};
Factors make_factors(const LevelGeom &g) {
    Factors F;
    for (int l = 1; l <= g.L; l++)
        for (int d = 0; d < 3; d++) {
            const uint32_t m = g.ld[l - 1][d];
            F.off[l][d] = F.host.size();
            std::vector<double> inv(m), cp(m);
            thomas_factors(m, inv.data(), cp.data());
            F.host.insert(F.host.end(), inv.begin(), inv.end());
            F.host.insert(F.host.end(), cp.begin(), cp.end());
        }
    return F;
}

struct WsLayout {
    size_t xa, xb, f, tab, total;
};
WsLayout ws_layout(const LevelGeom &g, size_t elem) {
    WsLayout w{};
    const uint64_t p1 = g.L >= 1 ? g.points[g.L - 1] : 1;
    const uint64_t p2 = g.L >= 2 ? g.points[g.L - 2] : 1;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    w.xa = 0;
    w.xb = al(w.xa + p1 * elem);
    w.f = al(w.xb + std::max(p1, p2) * elem);
    const uint64_t f1 = g.L >= 1 ? (uint64_t)g.ld[g.L - 1][0] * g.ld[g.L - 1][1] * f_pitch(g.ld[g.L - 1][2]) : 1;
    w.tab = al(w.f + f1 * sizeof(double));
    size_t tab = 0;
    for (int l = 1; l <= g.L; l++)
        for (int d = 0; d < 3; d++) tab += 2 * g.ld[l - 1][d];
    w.total = al(w.tab + tab * sizeof(double));
    return w;
}

// factor table of level l, dim d on the device and its fixed-point summary
template <class T>
void set_factors(SolveParams<T> &S, const double *tab_dev, const Factors &F, int l, int d, uint32_t m) {
    const double *h = F.host.data() + F.off[l][d];
    S.inv = tab_dev + F.off[l][d];
    S.last = h[m - 1];
    S.nfix = m;
    S.lim = 0.0;
    if (m >= 3) {
        uint32_t j = m - 2;  // h[j .. m - 2] all equal h[m - 2]
        while (j > 0 && h[j - 1] == h[m - 2]) j--;
        S.nfix = j;
        S.lim = h[m - 2];
    }
}


template <class T>
int solve_all(const LevelGeom &g, int l, double *f, const double *tab_dev, const Factors &F, int epi, const T *uc_in,
              T *uc_out, hipStream_t s) {
    const uint32_t *c = g.ld[l - 1];
    const uint32_t *n = g.ld[l];
    // solves run dims 2, 1, 0: the last one executed is the smallest active d
    int last = -1;
    for (int d = 0; d < 3; d++)
        if (n[d] > 1) { last = d; break; }
    for (int d = 2; d >= 0; d--) {
        if (n[d] == 1) continue;
        SolveParams<T> S;
        S.f = f;
        S.fp = f_pitch(c[2]);
        S.c[0] = c[0]; S.c[1] = c[1]; S.c[2] = c[2];
        set_factors(S, tab_dev, F, l, d, c[d]);
        S.epi = (d == last) ? epi : kEpiNone;
        S.uc_in = uc_in;
        S.uc_out = uc_out;
        const uint32_t m = c[d];
        if (d == 2) {
            const uint64_t nlines = (uint64_t)c[0] * c[1];
            const int cls = contig_class(m);
            const uint32_t grid = cls ? div_up(nlines, 4 * contig_class_lines_per_wave(cls)) : 0;
#define PRODM_CONTIG(SEG, W) hipLaunchKernelGGL((k_solve_contig<T, SEG, W>), dim3(grid), dim3(256), 0, s, S, nlines)
            switch (cls) {
            case 1: PRODM_CONTIG(1, 16); break;
            case 2: PRODM_CONTIG(1, 32); break;
            case 3: PRODM_CONTIG(1, 64); break;
            case 4: PRODM_CONTIG(2, 64); break;
            case 5: PRODM_CONTIG(4, 64); break;
            case 6: PRODM_CONTIG(5, 64); break;
            case 7: PRODM_CONTIG(8, 64); break;
            case 8: PRODM_CONTIG(9, 64); break;
            case 9: PRODM_CONTIG(16, 64); break;
            default:
                hipLaunchKernelGGL((k_solve_seq<T>), dim3(div_up(nlines, 256)), dim3(256), 0, s, S, m, nlines, d);
                break;
            }
#undef PRODM_CONTIG
        } else {
            const uint64_t nlines = (uint64_t)(d == 1 ? c[0] : c[1]) * S.fp;
            // up to 8 waves split a line; each lane keeps its segment in registers
            uint32_t nw = std::min<uint32_t>(8, (m + 15) / 16);
            if (nw < 1) nw = 1;
            const uint32_t seg = (m + nw - 1) / nw;
            const uint32_t grid = div_up(nlines, 64);
            if (seg <= 8)
                hipLaunchKernelGGL((k_solve_strided<T, 8>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, d);
            else if (seg <= 16)
                hipLaunchKernelGGL((k_solve_strided<T, 16>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, d);
            else if (seg <= 24)
                hipLaunchKernelGGL((k_solve_strided<T, 24>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, d);
            else if (seg <= 40)
                hipLaunchKernelGGL((k_solve_strided<T, 40>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, d);
            else if (seg <= 72)
                hipLaunchKernelGGL((k_solve_strided<T, 72>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, d);
            else
                hipLaunchKernelGGL((k_solve_seq<T>), dim3(div_up(nlines, 256)), dim3(256), 0, s, S, m, nlines, d);
        }
        PRODM_CHECK_LAUNCH();
    }
    return kOk;
}

template <class T>
LoadParams<T> load_params(const LevelGeom &g, int l, const CoefSink<T> &sink) {
    LoadParams<T> P{};
    for (int d = 0; d < 3; d++) { P.n[d] = g.ld[l][d]; P.c[d] = g.ld[l - 1][d]; P.ilv_c[d] = g.ld[l - 1][d]; }
    P.coef_full = sink.full;
    P.cs[0] = sink.s[0]; P.cs[1] = sink.s[1]; P.cs[2] = 1;
    P.coef_ilv = sink.level[l];
    return P;
}

// The decompose and adjoint sweeps address an input plane, a coefficient
// plane and a coarse-value plane through buffer descriptors of range
// 2^31 - 1 bytes (masked lanes use the offset 2^31), so each plane's byte
// extent must stay below 2^31. The recompose sweep and the interpolation
// form offsets inside a plane in 32 bits; they keep the same limit, so that
// one rule covers a plan's refactor and retrieve.
template <class T> bool sweep_planes_fit(const LoadParams<T> &P) {
    const uint64_t lim = 0x7FFFFFFFull;
    const uint64_t in_plane = ((uint64_t)(P.n[1] - 1) * P.is[1] + P.n[2]) * sizeof(T);
    const uint64_t coef_plane = (uint64_t)P.n[1] * std::max<uint64_t>(P.n[2], P.cs[1]) * sizeof(T);
    return in_plane < lim && coef_plane < lim;
}

// tiles of one sweep of shape SH at march length `march`
template <class T, class SH> uint32_t sweep_tiles(const LoadParams<T> &P, int march) {
    const uint32_t t2n = (P.n[2] == 1) ? 1 : div_up(P.c[2], SH::kM2);
    const uint32_t t1n = (P.n[1] == 1) ? 1 : div_up(P.c[1], SH::kM1);
    return t2n * t1n * march_tiles0(P.n[0], P.c[0], march);
}

template <class T, int MODE, class SH> void launch_sweep(LoadParams<T> P, bool ilv, hipStream_t s) {
    // march length: 16 coarse planes per workgroup, shorter when the level
    // has too few tiles to fill the 256 CUs (short marches = more
    // parallelism). Measured at 513^3, 6 levels: the decompose sweep keeps
    // 16 planes down to 256 workgroups (its sweeps below the finest level
    // took 0.231 ms per step with the 2048 rule, 0.161 ms with this one);
    // the recompose sweep is fastest shortening below 2048.
    int march = 16;
    const uint64_t want = MODE == kDecompose ? 256 : 2048;
    while (march > 2 && sweep_tiles<T, SH>(P, march) < want) march /= 2;
    P.march = march;
    const uint32_t grid = sweep_tiles<T, SH>(P, march);
    const bool a3 = P.n[0] > 1 && P.n[1] > 1 && P.n[2] > 1;
    const dim3 blk(SH::kThreads);
    if (ilv && a3) hipLaunchKernelGGL((k_load<T, MODE, true, true, SH>), dim3(grid), blk, 0, s, P);
    else if (ilv) hipLaunchKernelGGL((k_load<T, MODE, true, false, SH>), dim3(grid), blk, 0, s, P);
    else if (a3) hipLaunchKernelGGL((k_load<T, MODE, false, true, SH>), dim3(grid), blk, 0, s, P);
    else hipLaunchKernelGGL((k_load<T, MODE, false, false, SH>), dim3(grid), blk, 0, s, P);
}

template <class T, int MODE>
int launch_load(LoadParams<T> P, bool ilv, hipStream_t s) {
    if (!sweep_planes_fit(P)) {
        set_error("a fine plane of this level spans 2 GiB or more; the sweeps address planes with 31-bit offsets");
        return kErrUnsupported;
    }
    // the wide tiles for fp32 decompose levels that still give 2048
    // workgroups at full march (513^3: the finest level only); fp64 planes
    // would take 90 KB of LDS per workgroup
    bool wide = false;
    if constexpr (MODE == kDecompose && sizeof(T) == 4) wide = sweep_tiles<T, SweepWide>(P, 16) >= 2048;
    if (wide) {
        if constexpr (MODE == kDecompose && sizeof(T) == 4) launch_sweep<T, MODE, SweepWide>(P, ilv, s);
    } else {
        launch_sweep<T, MODE, SweepStd>(P, ilv, s);
    }
    PRODM_CHECK_LAUNCH();
    return kOk;
}

}  // namespace

size_t mgard_workspace_bytes(const LevelGeom &g, size_t elem) { return ws_layout(g, elem).total; }

int mgard_upload_tables(const LevelGeom &g, size_t elem, void *ws, size_t ws_bytes) {
    const WsLayout w = ws_layout(g, elem);
    if (ws_bytes < w.total) { set_error("workspace too small"); return kErrArg; }
    const Factors F = make_factors(g);
    if (!F.host.empty())
        PRODM_HIP(hipMemcpy(static_cast<char *>(ws) + w.tab, F.host.data(), F.host.size() * 8, hipMemcpyHostToDevice));
    return kOk;
}

template <class T>
int mgard_decompose(const T *in, const uint64_t *in_strides, const LevelGeom &g, const CoefSink<T> &sink,
                    KeyOf<T> *level_keys, void *ws, size_t ws_bytes, hipStream_t s, StageTimer *tm, bool tables_ready,
                    bool hier, const LevelHook *hook) {
    const LevelHook none{nullptr, nullptr};
    const LevelHook &ready = hook ? *hook : none;
    const WsLayout w = ws_layout(g, sizeof(T));
    if (ws_bytes < w.total) { set_error("decompose: workspace too small"); return kErrArg; }
    char *base = static_cast<char *>(ws);
    T *xa = reinterpret_cast<T *>(base + w.xa), *xb = reinterpret_cast<T *>(base + w.xb);
    double *f = reinterpret_cast<double *>(base + w.f);
    double *tab = reinterpret_cast<double *>(base + w.tab);
    const Factors F = make_factors(g);
    // a copy from pageable memory blocks the host until the stream drains:
    // plans upload the tables once (mgard_upload_tables)
    if (!tables_ready && !F.host.empty())
        PRODM_HIP(hipMemcpyAsync(tab, F.host.data(), F.host.size() * 8, hipMemcpyHostToDevice, s));
    const bool ilv = sink.full == nullptr;
    const T *cur = in;
    uint64_t cs[3] = {in_strides[0], in_strides[1], 1};
    T *bufs[2] = {xa, xb};
    int which = 0;
    for (int l = g.L; l >= 1; l--) {
        LoadParams<T> P = load_params(g, l, sink);
        P.in = cur;
        P.is[0] = cs[0]; P.is[1] = cs[1]; P.is[2] = 1;
        T *uc = bufs[which];
        P.uc = uc;
        P.f = f;
        P.fp = f_pitch(P.c[2]);
        P.key = level_keys ? level_keys + l : nullptr;
        P.hb = hier;
        int rc;
        {
            StageScope sc(tm, l == g.L ? kStDecLoadTop : kStDecLoadRest);
            rc = launch_load<T, kDecompose>(P, ilv, s);
        }
        if (!rc) rc = ready(l);
        if (rc) return rc;
        if (!hier) {
            StageScope sc(tm, l == g.L ? kStDecSolveTop : kStDecSolveRest);
            rc = solve_all<T>(g, l, f, tab, F, kEpiAdd, uc, uc, s);
        }
        if (rc) return rc;
        cur = uc;
        cs[0] = (uint64_t)g.ld[l - 1][1] * g.ld[l - 1][2];
        cs[1] = g.ld[l - 1][2];
        which ^= 1;
    }
    // level 0 (coarsest grid) into the sink
    StageScope sc0(tm, kStDecCopy);
    const uint32_t *c0 = g.ld[0];
    if (ilv) {
        int rc;
        // a kernel rather than a copy node, so that the sequence captures
        // into a graph as kernel nodes only
        if (g.L == 0) rc = copy_box<T>(in, in_strides[0], in_strides[1], sink.level[0], (uint64_t)c0[1] * c0[2], c0[2], c0, s);
        else rc = copy_box<T>(cur, cs[0], cs[1], sink.level[0], (uint64_t)c0[1] * c0[2], c0[2], c0, s);
        if (!rc && level_keys) rc = launch_maxabs<T>(sink.level[0], points3(c0), level_keys + 0, s);
        return rc ? rc : ready(0);
    }
    int rc = copy_box<T>(cur, cs[0], cs[1], sink.full, sink.s[0], sink.s[1], c0, s);
    if (rc) return rc;
    if (level_keys) {
        // level-0 max over the dense coarsest grid
        return launch_maxabs<T>(g.L == 0 ? in : cur, points3(c0), level_keys + 0, s);
    }
    return kOk;
}
template int mgard_decompose<float>(const float *, const uint64_t *, const LevelGeom &, const CoefSink<float> &,
                                    uint32_t *, void *, size_t, hipStream_t, StageTimer *, bool, bool,
                                    const LevelHook *);
template int mgard_decompose<double>(const double *, const uint64_t *, const LevelGeom &, const CoefSink<double> &,
                                     unsigned long long *, void *, size_t, hipStream_t, StageTimer *, bool, bool,
                                     const LevelHook *);

template <class T>
int mgard_recompose(const CoefSink<T> &src, const LevelGeom &g, int tables_ready, T *out, const uint64_t *out_strides,
                    void *ws, size_t ws_bytes, hipStream_t s, StageTimer *tm, bool hier, const LevelHook *hook,
                    double *const *f_levels, bool top_solved) {
    const LevelHook none{nullptr, nullptr};
    const LevelHook &need = hook ? *hook : none;
    const WsLayout w = ws_layout(g, sizeof(T));
    if (ws_bytes < w.total) { set_error("recompose: workspace too small"); return kErrArg; }
    char *base = static_cast<char *>(ws);
    T *bufs[2] = {reinterpret_cast<T *>(base + w.xa), reinterpret_cast<T *>(base + w.xb)};
    double *f = reinterpret_cast<double *>(base + w.f);
    double *tab = reinterpret_cast<double *>(base + w.tab);
    const Factors F = make_factors(g);
    if (!tables_ready && !F.host.empty())
        PRODM_HIP(hipMemcpyAsync(tab, F.host.data(), F.host.size() * 8, hipMemcpyHostToDevice, s));
    const bool ilv = src.full == nullptr;
    const uint32_t *c0 = g.ld[0];
    if (int rc = need(0)) return rc;
    if (g.L == 0) {
        if (ilv) return copy_box<T>(src.level[0], (uint64_t)c0[1] * c0[2], c0[2], out, out_strides[0], out_strides[1], c0, s);
        return copy_box<T>(src.full, src.s[0], src.s[1], out, out_strides[0], out_strides[1], c0, s);
    }
    // coarsest grid into a dense buffer
    int which = 0;
    T *cur = bufs[which];
    int rc;
    if (ilv) rc = copy_box<T>(src.level[0], (uint64_t)c0[1] * c0[2], c0[2], cur, (uint64_t)c0[1] * c0[2], c0[2], c0, s);
    else rc = copy_box<T>(src.full, src.s[0], src.s[1], cur, (uint64_t)c0[1] * c0[2], c0[2], c0, s);
    if (rc) return rc;
    for (int l = 1; l <= g.L; l++) {
        if ((rc = need(l))) return rc;
        LoadParams<T> P = load_params(g, l, src);
        double *fl = f_levels ? f_levels[l] : f;
        P.f = fl;
        P.fp = f_pitch(P.c[2]);
        const bool subz = !hier && top_solved && f_levels && ilv && l == g.L;
        if (!hier && !subz) {
            if (!f_levels) {
                StageScope sc(tm, l == g.L ? kStRecLoadTop : kStRecLoadRest);
                rc = launch_load<T, kRecompose>(P, ilv, s);
            }
            if (rc) return rc;
            StageScope sc(tm, l == g.L ? kStRecSolveTop : kStRecSolveRest);
            rc = solve_all<T>(g, l, fl, tab, F, kEpiSub, cur, cur, s);  // u_c = u_c' - z, in place
            if (rc) return rc;
        }
        const uint32_t *n = g.ld[l];
        T *dst;
        uint64_t d0, d1;
        if (l == g.L) { dst = out; d0 = out_strides[0]; d1 = out_strides[1]; }
        else { dst = bufs[which ^ 1]; d0 = (uint64_t)n[1] * n[2]; d1 = n[2]; }
        const uint32_t grid = div_up(n[2], kI2) * div_up(n[1], kI1) * div_up(n[0], kI0);
        {
            StageScope sc(tm, l == g.L ? kStRecInterpTop : kStRecInterpRest);
            if (subz) hipLaunchKernelGGL((k_interp<T, true, true>), dim3(grid), dim3(256), 0, s, P, cur, dst, d0, d1);
            else if (ilv) hipLaunchKernelGGL((k_interp<T, true>), dim3(grid), dim3(256), 0, s, P, cur, dst, d0, d1);
            else hipLaunchKernelGGL((k_interp<T, false>), dim3(grid), dim3(256), 0, s, P, cur, dst, d0, d1);
        }
        PRODM_CHECK_LAUNCH();
        which ^= 1;
        cur = bufs[which];
    }
    return kOk;
}
// Adjoint helpers (adjoint.hip): h = P^T g on the coarse grid of level l
// (the load kernel with the interpolation-transpose weights), copied to
// h_out, then t = S h in place in the workspace's load vector.
int mgard_adjoint_coarse(const double *gfine, const LevelGeom &g, int l, double *h_out, void *ws, size_t ws_bytes,
                         hipStream_t s) {
    const WsLayout w = ws_layout(g, sizeof(double));
    if (ws_bytes < w.total) { set_error("adjoint: workspace too small"); return kErrArg; }
    char *base = static_cast<char *>(ws);
    double *f = reinterpret_cast<double *>(base + w.f);
    double *tab = reinterpret_cast<double *>(base + w.tab);
    const Factors F = make_factors(g);
    if (l == g.L && !F.host.empty())
        PRODM_HIP(hipMemcpyAsync(tab, F.host.data(), F.host.size() * 8, hipMemcpyHostToDevice, s));
    LoadParams<double> P{};
    for (int d = 0; d < 3; d++) { P.n[d] = g.ld[l][d]; P.c[d] = g.ld[l - 1][d]; P.ilv_c[d] = g.ld[l - 1][d]; }
    P.in = gfine;
    P.is[0] = (uint64_t)P.n[1] * P.n[2];
    P.is[1] = P.n[2];
    P.is[2] = 1;
    P.f = f;
    P.fp = f_pitch(P.c[2]);
    int rc = launch_load<double, kAdjointPT>(P, false, s);
    if (rc) return rc;
    const uint32_t *c = g.ld[l - 1];
    rc = copy_box<double>(f, (uint64_t)c[1] * P.fp, P.fp, h_out, (uint64_t)c[1] * c[2], c[2], c, s);
    if (rc) return rc;
    return solve_all<double>(g, l, f, tab, F, kEpiNone, nullptr, nullptr, s);
}
const double *mgard_adjoint_solution(const LevelGeom &g, int l, void *ws, uint32_t *pitch) {
    *pitch = f_pitch(g.ld[l - 1][2]);
    return reinterpret_cast<const double *>(static_cast<char *>(ws) + ws_layout(g, sizeof(double)).f);
}

template int mgard_recompose<float>(const CoefSink<float> &, const LevelGeom &, int, float *, const uint64_t *, void *,
                                    size_t, hipStream_t, StageTimer *, bool, const LevelHook *, double *const *, bool);
template int mgard_recompose<double>(const CoefSink<double> &, const LevelGeom &, int, double *, const uint64_t *,
                                     void *, size_t, hipStream_t, StageTimer *, bool, const LevelHook *,
                                     double *const *, bool);

template <class T> int mgard_solve_ahead(const LevelGeom &g, int l, double *f_l, void *ws, size_t ws_bytes, hipStream_t s) {
    const WsLayout w = ws_layout(g, sizeof(T));
    if (l < 1 || l > g.L || !f_l || ws_bytes < w.total) { set_error("solve ahead: bad level or workspace"); return kErrArg; }
    const double *tab = reinterpret_cast<const double *>(static_cast<const char *>(ws) + w.tab);
    const Factors F = make_factors(g);
    return solve_all<T>(g, l, f_l, tab, F, kEpiNone, (const T *)nullptr, (T *)nullptr, s);
}
template int mgard_solve_ahead<float>(const LevelGeom &, int, double *, void *, size_t, hipStream_t);
template int mgard_solve_ahead<double>(const LevelGeom &, int, double *, void *, size_t, hipStream_t);

size_t mgard_level_f_bytes(const LevelGeom &g, int l) {
    if (l < 1 || l > g.L) return 0;
    const uint32_t *c = g.ld[l - 1];
    return (size_t)c[0] * c[1] * f_pitch(c[2]) * sizeof(double);
}

template <class T> int mgard_recompose_load(const CoefSink<T> &src, const LevelGeom &g, int l, double *f_l, hipStream_t s) {
    if (l < 1 || l > g.L || !f_l) { set_error("recompose load: bad level"); return kErrArg; }
    LoadParams<T> P = load_params(g, l, src);
    P.f = f_l;
    P.fp = f_pitch(P.c[2]);
    return launch_load<T, kRecompose>(P, src.full == nullptr, s);
}
template int mgard_recompose_load<float>(const CoefSink<float> &, const LevelGeom &, int, double *, hipStream_t);
template int mgard_recompose_load<double>(const CoefSink<double> &, const LevelGeom &, int, double *, hipStream_t);

// ---- batched decompose: grouped launches over many fields ---------------------
namespace {

template <class E> void append_entry(GroupLaunch &L, const E &e, uint32_t blocks) {
    const size_t o = L.params.size();
    L.params.resize(o + sizeof(E));
    std::memcpy(L.params.data() + o, &e, sizeof(E));
    if (L.first.empty()) L.first.push_back(0);
    L.first.push_back(L.first.back() + blocks);
    L.nent++;
    L.total = L.first.back();
}

// entries in decreasing work, so that the largest fields are dispatched first
template <class E> void fill_sorted(GroupLaunch &L, std::vector<std::pair<uint32_t, E>> &v) {
    std::stable_sort(v.begin(), v.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    for (auto &x : v)
        if (x.first) append_entry(L, x.second, x.first);
}

int seg_class_strided(uint32_t seg) {
    return seg <= 8 ? 8 : seg <= 16 ? 16 : seg <= 24 ? 24 : seg <= 40 ? 40 : seg <= 72 ? 72 : 0;
}
uint32_t strided_waves(uint32_t m) { return std::max<uint32_t>(1, std::min<uint32_t>(8, (m + 15) / 16)); }

template <class T> int default_march(const LoadParams<T> &P) {
    // shorter marches while the field's level has few tiles (in a grouped
    // launch the fields' workgroups fill the GPU together; the autotuner
    // picks the march per launch)
    int march = 16;
    while (march > 2 && sweep_tiles<T, SweepStd>(P, march) < 2048) march /= 2;
    return march;
}

}  // namespace

template <class T>
int mgard_batched_launches(const std::vector<BatchFieldRef> &fields, const std::vector<GroupShape> &shapes,
                           std::vector<GroupLaunch> &out) {
    out.clear();
    struct State {
        const T *cur;
        uint64_t cs0, cs1;
        int which;
    };
    std::vector<State> st(fields.size());
    int Lmax = 0;
    for (size_t i = 0; i < fields.size(); i++) {
        const BatchFieldRef &f = fields[i];
        if (!f.g || !f.g->ok || f.g->L < 1 || !f.in || !f.ws || !f.keys) {
            set_error("batched decompose: field " + std::to_string(i) + " needs a valid geometry with L >= 1");
            return kErrArg;
        }
        if (f.ws_bytes < ws_layout(*f.g, sizeof(T)).total) {
            set_error("batched decompose: workspace too small for field " + std::to_string(i));
            return kErrArg;
        }
        Lmax = std::max(Lmax, f.g->L);
        st[i] = State{static_cast<const T *>(f.in), (uint64_t)f.g->dims[1] * f.g->dims[2], f.g->dims[2], 0};
    }
    auto make = [&](int kind, uint32_t block) {
        GroupLaunch L;
        L.kind = kind;
        L.elem = sizeof(T);
        L.block = block;
        if (out.size() < shapes.size()) L.shape = shapes[out.size()];
        return L;
    };
    for (int j = 0; j < Lmax; j++) {
        GroupLaunch load = make(kGrpLoad, kLoadThreads);
        std::vector<std::pair<uint32_t, LoadParams<T>>> le;
        std::vector<std::pair<uint32_t, SolveEntry<T>>> se[3];
        for (size_t i = 0; i < fields.size(); i++) {
            const BatchFieldRef &f = fields[i];
            const LevelGeom &g = *f.g;
            const int l = g.L - j;
            if (l < 1) continue;
            const WsLayout w = ws_layout(g, sizeof(T));
            char *base = static_cast<char *>(f.ws);
            T *bufs[2] = {reinterpret_cast<T *>(base + w.xa), reinterpret_cast<T *>(base + w.xb)};
            double *fv = reinterpret_cast<double *>(base + w.f);
            const double *tab = reinterpret_cast<const double *>(base + w.tab);
            CoefSink<T> sink;
            for (int k = 0; k <= g.L; k++) sink.level[k] = static_cast<T *>(f.level[k]);
            State &s = st[i];
            LoadParams<T> P = load_params(g, l, sink);
            P.in = s.cur;
            P.is[0] = s.cs0; P.is[1] = s.cs1; P.is[2] = 1;
            T *uc = bufs[s.which];
            P.uc = uc;
            P.f = fv;
            P.fp = f_pitch(P.c[2]);
            P.key = static_cast<KeyOf<T> *>(f.keys) + l;
            P.hb = f.hier;
            P.march = load.shape.a > 0 ? load.shape.a : default_march(P);
            if (!sweep_planes_fit(P)) {
                set_error("batched decompose: field " + std::to_string(i) +
                          " has a fine plane of 2 GiB or more; the sweeps address planes with 31-bit offsets");
                return kErrUnsupported;
            }
            le.push_back({load_tiles(P), P});
            if (!f.hier) {
                const uint32_t *c = g.ld[l - 1], *n = g.ld[l];
                int last = -1;
                for (int d = 0; d < 3; d++)
                    if (n[d] > 1) { last = d; break; }
                const Factors F = make_factors(g);
                for (int d = 2; d >= 0; d--) {
                    if (n[d] == 1) continue;
                    SolveEntry<T> E{};
                    E.S.f = fv;
                    E.S.fp = f_pitch(c[2]);
                    E.S.c[0] = c[0]; E.S.c[1] = c[1]; E.S.c[2] = c[2];
                    set_factors(E.S, tab, F, l, d, c[d]);
                    E.S.epi = (d == last) ? kEpiAdd : kEpiNone;
                    E.S.uc_in = uc;
                    E.S.uc_out = uc;
                    E.m = c[d];
                    E.d = d;
                    E.nlines = d == 2 ? (uint64_t)c[0] * c[1] : (uint64_t)(d == 1 ? c[0] : c[1]) * E.S.fp;
                    E.nw = d == 2 ? 1 : strided_waves(E.m);
                    E.cls = d == 2 ? contig_class(E.m) : 0;
                    se[d].push_back({0, E});
                }
            }
            s.cur = uc;
            s.cs0 = (uint64_t)g.ld[l - 1][1] * g.ld[l - 1][2];
            s.cs1 = g.ld[l - 1][2];
            s.which ^= 1;
        }
        fill_sorted(load, le);
        out.push_back(std::move(load));
        // line solves along dims 2, 1, 0
        for (int d = 2; d >= 0; d--) {
            if (se[d].empty()) continue;
            if (d == 2) {
                GroupLaunch L = make(kGrpContig, 256);
                const uint32_t wpb = L.shape.a > 0 ? std::min(8, L.shape.a) : 4;
                L.block = 64 * wpb;
                int cmax = 1;
                for (auto &x : se[d]) {
                    const int c = x.second.cls;
                    if (!c) {
                        set_error("batched decompose: a contiguous line of more than 1024 coarse nodes (the grouped "
                                  "solves hold lines in registers)");
                        return kErrUnsupported;
                    }
                    cmax = std::max(cmax, c);
                    x.first = div_up(x.second.nlines, (uint64_t)wpb * contig_class_lines_per_wave(c));
                }
                L.variant = cmax;
                fill_sorted(L, se[d]);
                out.push_back(std::move(L));
            } else {
                GroupLaunch L = make(kGrpStrided, 64);
                uint32_t segmax = 0, nwmax = 1;
                for (auto &x : se[d]) {
                    const uint32_t nw = x.second.nw;
                    segmax = std::max(segmax, (x.second.m + nw - 1) / nw);
                    nwmax = std::max(nwmax, nw);
                }
                // a workgroup of nwmax waves packs floor(nwmax / nw) of an
                // entry's 64-line groups
                for (auto &x : se[d]) x.first = div_up(div_up(x.second.nlines, 64), nwmax / x.second.nw);
                L.variant = seg_class_strided(segmax);
                if (!L.variant) {
                    set_error("batched decompose: a strided line of more than 576 coarse nodes (the grouped solves "
                              "hold lines in registers)");
                    return kErrUnsupported;
                }
                L.block = 64 * nwmax;
                fill_sorted(L, se[d]);
                out.push_back(std::move(L));
            }
        }
    }
    // coarsest grids
    GroupLaunch L0 = make(kGrpLevel0, 256);
    std::vector<std::pair<uint32_t, Level0Entry<T>>> e0;
    for (size_t i = 0; i < fields.size(); i++) {
        const LevelGeom &g = *fields[i].g;
        Level0Entry<T> E{};
        E.src = st[i].cur;
        E.s0 = st[i].cs0;
        E.s1 = st[i].cs1;
        E.dst = static_cast<T *>(fields[i].level[0]);
        E.key = static_cast<KeyOf<T> *>(fields[i].keys);
        E.c0 = g.ld[0][0]; E.c1 = g.ld[0][1]; E.c2 = g.ld[0][2];
        e0.push_back({div_up(points3(g.ld[0]), 256), E});
    }
    fill_sorted(L0, e0);
    out.push_back(std::move(L0));
    return kOk;
}
template int mgard_batched_launches<float>(const std::vector<BatchFieldRef> &, const std::vector<GroupShape> &,
                                           std::vector<GroupLaunch> &);
template int mgard_batched_launches<double>(const std::vector<BatchFieldRef> &, const std::vector<GroupShape> &,
                                            std::vector<GroupLaunch> &);

namespace {
template <class T> int group_run_t(const GroupLaunch &L, hipStream_t s) {
    const uint32_t grid = L.grid();
    if (!grid || !L.nent) return kOk;
    const uint32_t *first = L.d_first;
    const dim3 G(grid), B(L.block);
#define PRODM_GRP(kern, PT) hipLaunchKernelGGL(kern, G, B, 0, s, static_cast<const PT *>(L.d_params), first, L.nent, L.total)
    switch (L.kind) {
    case kGrpLoad: PRODM_GRP((k_load_grouped<T>), LoadParams<T>); break;
    case kGrpContig:
        switch (L.variant) {
        case 1: PRODM_GRP((k_contig_grouped<T, 1>), SolveEntry<T>); break;
        case 2: PRODM_GRP((k_contig_grouped<T, 2>), SolveEntry<T>); break;
        case 3: PRODM_GRP((k_contig_grouped<T, 3>), SolveEntry<T>); break;
        case 4: PRODM_GRP((k_contig_grouped<T, 4>), SolveEntry<T>); break;
        case 5: PRODM_GRP((k_contig_grouped<T, 5>), SolveEntry<T>); break;
        case 6: PRODM_GRP((k_contig_grouped<T, 6>), SolveEntry<T>); break;
        case 7: PRODM_GRP((k_contig_grouped<T, 7>), SolveEntry<T>); break;
        case 8: PRODM_GRP((k_contig_grouped<T, 8>), SolveEntry<T>); break;
        case 9: PRODM_GRP((k_contig_grouped<T, 9>), SolveEntry<T>); break;
        default: set_error("grouped contiguous solve: bad class"); return kErrArg;
        }
        break;
    case kGrpStrided:
        if (L.variant == 8) PRODM_GRP((k_strided_grouped<T, 8>), SolveEntry<T>);
        else if (L.variant == 16) PRODM_GRP((k_strided_grouped<T, 16>), SolveEntry<T>);
        else if (L.variant == 24) PRODM_GRP((k_strided_grouped<T, 24>), SolveEntry<T>);
        else if (L.variant == 40) PRODM_GRP((k_strided_grouped<T, 40>), SolveEntry<T>);
        else if (L.variant == 72) PRODM_GRP((k_strided_grouped<T, 72>), SolveEntry<T>);
        else { set_error("grouped strided solve: bad variant"); return kErrArg; }
        break;
    case kGrpLevel0: PRODM_GRP((k_level0_grouped<T>), Level0Entry<T>); break;
    default: set_error("grouped launch: not an MGARD kind"); return kErrArg;
    }
#undef PRODM_GRP
    PRODM_CHECK_LAUNCH();
    return kOk;
}
}  // namespace

int mgard_group_run(const GroupLaunch &L, hipStream_t s) {
    if (!L.d_params || !L.d_first) { set_error("grouped launch: tables not uploaded"); return kErrState; }
    return L.elem == 8 ? group_run_t<double>(L, s) : group_run_t<float>(L, s);
}

}  // namespace prodm

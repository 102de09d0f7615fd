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
//   k_solve_contig   line solves along dim 2 (contiguous): one wave per line,
//             segment per lane, affine-map scan across lanes.
//   k_solve_strided  line solves along dims 1 / 0: lane per line (coalesced),
//             waves split the line into segments, LDS scan across waves;
//             the last solve applies u_c +/- z.
//   k_interp  recompose: u = I u_c + W.
#include "internal.hpp"

#include <algorithm>
#include <vector>

namespace prodm {

constexpr int kTJ0 = 8, kTJ1 = 8, kTJ2 = 32;          // coarse tile per workgroup
constexpr int kR1 = 2 * kTJ1 + 3, kR2 = 2 * kTJ2 + 3;  // fine region (with halo)
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

template <class T> struct LoadParams {
    uint32_t n[3], c[3];
    const T *in;           // natural-order input box (decompose / adjoint)
    uint64_t is[3];        // its strides
    T *coef_full;          // reordered full-array coefficients
    uint64_t cs[3];
    T *coef_ilv;           // interleaved level buffer
    uint32_t ilv_c[3];     // coarse dims used by the interleave index
    T *uc;                 // decompose: coarse nodal values (dense c)
    double *f;             // load vector (dense c)
    KeyOf<T> *key;         // decompose: level max |W|
};

template <class T>
__device__ inline uint64_t coef_index_full(const LoadParams<T> &P, uint32_t i0, uint32_t i1, uint32_t i2) {
    return rpos(i0, P.c[0]) * P.cs[0] + rpos(i1, P.c[1]) * P.cs[1] + rpos(i2, P.c[2]);
}
template <class T>
__device__ inline uint64_t coef_index_ilv(const LoadParams<T> &P, uint32_t i0, uint32_t i1, uint32_t i2) {
    return ilv_index(rpos(i0, P.c[0]), rpos(i1, P.c[1]), rpos(i2, P.c[2]), P.n, P.ilv_c);
}

// Multilinear interpolant of fine node (r1, r2) of the current plane from
// its coarse neighbours; corners lexicographic over the odd dims, pairwise
// tree sum, then * 1/2, 1/4, 1/8 (same expression as the CPU restatement).
template <class T>
__device__ inline T interp_lds(const T (*A)[kR2], const T (*Bp)[kR2], const T (*Cp)[kR2], int o0, int r1, int r2) {
    // A: plane i0-1, Bp: plane i0+1 (used when o0), Cp: current plane
    const int o1 = r1 & 1, o2 = r2 & 1;
    auto V = [&](int x0, int x1, int x2) -> T {
        const int a1 = o1 ? r1 + 2 * x1 - 1 : r1;
        const int a2 = o2 ? r2 + 2 * x2 - 1 : r2;
        const T(*pl)[kR2] = o0 ? (x0 ? Bp : A) : Cp;
        return pl[a1][a2];
    };
    const int cnt = o0 + o1 + o2;
    if (cnt == 1) {
        if (o2) return (V(0, 0, 0) + V(0, 0, 1)) * (T)0.5;
        if (o1) return (V(0, 0, 0) + V(0, 1, 0)) * (T)0.5;
        return (V(0, 0, 0) + V(1, 0, 0)) * (T)0.5;
    }
    if (cnt == 2) {
        if (!o0) return ((V(0, 0, 0) + V(0, 0, 1)) + (V(0, 1, 0) + V(0, 1, 1))) * (T)0.25;
        if (!o1) return ((V(0, 0, 0) + V(0, 0, 1)) + (V(1, 0, 0) + V(1, 0, 1))) * (T)0.25;
        return ((V(0, 0, 0) + V(0, 1, 0)) + (V(1, 0, 0) + V(1, 1, 0))) * (T)0.25;
    }
    return (((V(0, 0, 0) + V(0, 0, 1)) + (V(0, 1, 0) + V(0, 1, 1))) +
            ((V(1, 0, 0) + V(1, 0, 1)) + (V(1, 1, 0) + V(1, 1, 1)))) *
           (T)0.125;
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

template <class T, int MODE, bool ILV>
__global__ __launch_bounds__(kLoadThreads) void k_load(LoadParams<T> P) {
    using K = KeyOf<T>;
    __shared__ T Ue0[kR1][kR2], Ue1[kR1][kR2], Uo[kR1][kR2], Wb[kR1][kR2];
    __shared__ double g2[kR1][kTJ2];
    const int tid = threadIdx.x;
    const uint32_t t2n = (P.n[2] == 1) ? 1 : (P.c[2] + kTJ2 - 1) / kTJ2;
    const uint32_t t1n = (P.n[1] == 1) ? 1 : (P.c[1] + kTJ1 - 1) / kTJ1;
    const uint32_t bt = blockIdx.x;
    const DimTile D2 = dim_tile(bt % t2n, kTJ2, P.c[2], P.n[2]);
    const DimTile D1 = dim_tile((bt / t2n) % t1n, kTJ1, P.c[1], P.n[1]);
    const DimTile D0 = dim_tile(bt / (t2n * t1n), kTJ0, P.c[0], P.n[0]);
    const int R1 = D1.R, R2 = D2.R, RR = R1 * R2;
    // ownership of fine indices: [2js - 1, 2(js + tj) - 1) clipped
    auto own = [](const DimTile &d, int i, uint32_t n) -> bool {
        if (n == 1) return true;
        return i >= 2 * d.js - 1 && i < 2 * (d.js + d.tj) - 1 && i >= 0 && i < (int)n;
    };
    // thread's coarse column for the accumulation
    const int cj1 = tid / kTJ2, cj2 = tid % kTJ2;
    const bool has_col = cj1 < D1.tj && cj2 < D2.tj;
    double acc[3] = {0.0, 0.0, 0.0};
    T lmax = 0;

    const int i0lo = (P.n[0] == 1) ? 0 : max(0, 2 * D0.js - 2);
    const int i0hi = (P.n[0] == 1) ? 0 : min((int)P.n[0] - 1, 2 * (D0.js + D0.tj - 1) + 2);

    auto load_plane = [&](T (*dst)[kR2], int i0) {
        for (int idx = tid; idx < RR; idx += kLoadThreads) {
            const int r1 = idx / R2, r2 = idx % R2;
            const int i1 = D1.rs + r1, i2 = D2.rs + r2;
            T v = (T)0;
            if (i1 >= 0 && i1 < (int)P.n[1] && i2 >= 0 && i2 < (int)P.n[2]) {
                if (MODE == kRecompose) {
                    const bool coarse = !(i0 & 1) && !(i1 & 1) && !(i2 & 1);
                    if (!coarse) {
                        if (ILV) v = P.coef_ilv[coef_index_ilv(P, i0, i1, i2)];
                        else v = P.coef_full[coef_index_full(P, i0, i1, i2)];
                    }
                } else {
                    v = P.in[(uint64_t)i0 * P.is[0] + (uint64_t)i1 * P.is[1] + (uint64_t)i2];
                }
            }
            dst[r1][r2] = v;
        }
    };

    // W of plane i0 into Wb (decompose), from planes A (i0-1), B (i0+1), C (i0)
    auto make_w = [&](int i0, const T (*A)[kR2], const T (*Bp)[kR2], const T (*Cp)[kR2]) {
        const int o0 = (P.n[0] == 1) ? 0 : (i0 & 1);
        for (int idx = tid; idx < RR; idx += kLoadThreads) {
            const int r1 = idx / R2, r2 = idx % R2;
            const int i1 = D1.rs + r1, i2 = D2.rs + r2;
            const bool inb = i1 >= 0 && i1 < (int)P.n[1] && i2 >= 0 && i2 < (int)P.n[2];
            const bool coarse = !o0 && !(r1 & 1) && !(r2 & 1);
            T w = (T)0;
            if (inb && !coarse) w = Cp[r1][r2] - interp_lds<T>(A, Bp, Cp, o0, r1, r2);
            Wb[r1][r2] = w;
            if (inb && own(D0, i0, P.n[0]) && own(D1, i1, P.n[1]) && own(D2, i2, P.n[2])) {
                if (coarse) {
                    const uint64_t j = ((uint64_t)(i0 >> 1) * P.c[1] + (i1 >> 1)) * P.c[2] + (i2 >> 1);
                    P.uc[j] = Cp[r1][r2];
                } else {
                    if (ILV) P.coef_ilv[coef_index_ilv(P, i0, i1, i2)] = w;
                    else P.coef_full[coef_index_full(P, i0, i1, i2)] = w;
                    if (w == w) lmax = fabs(w) > lmax ? fabs(w) : lmax;
                }
            }
        }
    };

    // in-plane restriction of Wb (dims 2 then 1) and accumulation along dim 0
    auto restrict_plane = [&](int i0) {
        __syncthreads();
        const int tj2 = D2.tj;
        for (int idx = tid; idx < R1 * tj2; idx += kLoadThreads) {
            const int r1 = idx / tj2, jj2 = idx % tj2;
            const int j2 = D2.js + jj2;
            double a = 0.0;
            if (P.n[2] == 1) {
                a += 1.0 * (double)Wb[r1][0];
            } else {
                for (int k = -2; k <= 2; k++) {
                    const int i2 = 2 * j2 + k;
                    if (i2 < 0 || i2 >= (int)P.n[2]) continue;
                    const double wt = (MODE == kAdjointPT) ? pweight(j2, k, P.c[2]) : rweight(j2, k, P.c[2]);
                    if (MODE == kAdjointPT && wt == 0.0) continue;
                    a += wt * (double)Wb[r1][i2 - D2.rs];
                }
            }
            g2[r1][jj2] = a;
        }
        __syncthreads();
        if (has_col) {
            const int j1 = D1.js + cj1;
            double h = 0.0;
            if (P.n[1] == 1) {
                h += 1.0 * g2[0][cj2];
            } else {
                for (int k = -2; k <= 2; k++) {
                    const int i1 = 2 * j1 + k;
                    if (i1 < 0 || i1 >= (int)P.n[1]) continue;
                    const double wt = (MODE == kAdjointPT) ? pweight(j1, k, P.c[1]) : rweight(j1, k, P.c[1]);
                    if (MODE == kAdjointPT && wt == 0.0) continue;
                    h += wt * g2[i1 - D1.rs][cj2];
                }
            }
            // dim 0: plane i0 contributes to coarse j0 with |2 j0 - i0| <= 2
            if (P.n[0] == 1) {
                acc[0] += 1.0 * h;
            } else {
                for (int j0 = (i0 - 1) >> 1; j0 <= (i0 + 2) >> 1; j0++) {
                    if (j0 < D0.js || j0 >= D0.js + D0.tj) continue;
                    const int k = i0 - 2 * j0;
                    if (k < -2 || k > 2) continue;
                    const double wt = (MODE == kAdjointPT) ? pweight(j0, k, P.c[0]) : rweight(j0, k, P.c[0]);
                    if (MODE == kAdjointPT && wt == 0.0) continue;
                    const int sl = j0 % 3;
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        if (q == sl) acc[q] += wt * h;
                }
            }
            // coarse plane j0 is complete once fine plane min(2 j0 + 2, n0 - 1) is in
            const int j2 = D2.js + cj2;
            auto emit = [&](int j0) {
                if (j0 < D0.js || j0 >= D0.js + D0.tj || j0 < 0) return;
                const int sl = j0 % 3;
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < 3; q++)
                    if (q == sl) { v = acc[q]; acc[q] = 0.0; }
                P.f[((uint64_t)j0 * P.c[1] + j1) * P.c[2] + j2] = v;
            };
            if (P.n[0] == 1) {
                emit(0);
            } else if (!(i0 & 1)) {
                if (i0 >= 2) emit((i0 - 2) / 2);
                if (i0 == (int)P.n[0] - 1) emit(i0 / 2);
            }
        }
        __syncthreads();
    };

    if (MODE == kDecompose) {
        if (P.n[0] == 1) {
            load_plane(Ue0, 0);
            __syncthreads();
            make_w(0, Ue0, Ue0, Ue0);
            restrict_plane(0);
        } else {
            load_plane(Ue0, i0lo);
            __syncthreads();
            make_w(i0lo, Ue0, Ue0, Ue0);
            restrict_plane(i0lo);
            T(*E0)[kR2] = Ue0;
            T(*E1)[kR2] = Ue1;
            for (int i0 = i0lo + 1; i0 <= i0hi; i0 += 2) {
                load_plane(Uo, i0);
                load_plane(E1, i0 + 1);
                __syncthreads();
                make_w(i0, E0, E1, Uo);
                restrict_plane(i0);
                make_w(i0 + 1, E1, E1, E1);
                restrict_plane(i0 + 1);
                T(*t)[kR2] = E0;
                E0 = E1;
                E1 = t;
            }
        }
        // level max |W|
        if (P.key) {
            K m = abs_key(lmax);
            for (int o = 32; o >= 1; o >>= 1) {
                K q = __shfl_xor(m, o, 64);
                m = q > m ? q : m;
            }
            if (lane_id() == 0 && m) atomicMax(P.key, m);
        }
    } else {
        // recompose: W straight from the coefficients; adjoint: g
        for (int i0 = i0lo; i0 <= i0hi; i0++) {
            load_plane(Wb, i0);
            restrict_plane(i0);
        }
    }
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

template <class T> struct SolveParams {
    double *f;            // dense c0 x c1 x c2
    uint32_t c[3];
    const double *inv, *cp;
    int epi;
    const T *uc_in;       // epilogue input (dense c)
    T *uc_out;            // epilogue output (dense c)
};

template <class T> __device__ inline void epilogue(const SolveParams<T> &S, uint64_t idx, double z) {
    if (S.epi == kEpiAdd) S.uc_out[idx] = (T)((double)S.uc_in[idx] + z);
    else if (S.epi == kEpiSub) S.uc_out[idx] = (T)((double)S.uc_in[idx] - z);
    else S.f[idx] = z;
}

// contiguous lines (dim 2): one wave per line, lane holds a segment of SEG
template <class T, int SEG>
__global__ __launch_bounds__(256) void k_solve_contig(SolveParams<T> S, uint64_t nlines) {
    const uint32_t m = S.c[2];
    const uint32_t lane = lane_id();
    const uint64_t line = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (line >= nlines) return;
    double *x = S.f + line * m;
    const uint32_t seg = (m + 63) / 64;
    const uint32_t b = lane * seg;
    const uint32_t e = min(m, b + seg);
    double v[SEG];
#pragma unroll
    for (int i = 0; i < SEG; i++) v[i] = (b + i < e) ? x[b + i] : 0.0;
    // forward: y_i = (f_i - y_{i-1}) * inv_i  == affine map y_{i-1} -> y_i
    double A = 0.0, Bm = 1.0;
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (b + i < e) {
            const double iv = S.inv[b + i];
            A = (v[i] - A) * iv;
            Bm = -Bm * iv;
        }
    }
    for (int o = 1; o < 64; o <<= 1) {
        const double pA = __shfl_up(A, o, 64), pB = __shfl_up(Bm, o, 64);
        if ((int)lane >= o) { A = A + Bm * pA; Bm = Bm * pB; }
    }
    double y = __shfl_up(A, 1, 64);
    if (lane == 0) y = 0.0;
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (b + i < e) {
            y = (v[i] - y) * S.inv[b + i];
            v[i] = y;
        }
    }
    // backward: z_i = y_i - cp_i z_{i+1} (z_{m-1} = y_{m-1})
    double C = 0.0, Dm = 1.0;
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        if (b + i < e) {
            const double c = (b + i == m - 1) ? 0.0 : S.cp[b + i];
            C = v[i] - c * C;
            Dm = -c * Dm;
        }
    }
    for (int o = 1; o < 64; o <<= 1) {
        const double nC = __shfl_down(C, o, 64), nD = __shfl_down(Dm, o, 64);
        if (lane + o < 64) { C = C + Dm * nC; Dm = Dm * nD; }
    }
    double z = __shfl_down(C, 1, 64);
    if (lane == 63) z = 0.0;
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        if (b + i < e) {
            const double c = (b + i == m - 1) ? 0.0 : S.cp[b + i];
            z = v[i] - c * z;
            v[i] = z;
        }
    }
#pragma unroll
    for (int i = 0; i < SEG; i++)
        if (b + i < e) epilogue(S, line * m + b + i, v[i]);
}

// strided lines: lines q in [0, nlines), element j at base(q) + j * stride,
// base(q) = (q / inner) * outer + q % inner. Lane per line, W waves per
// workgroup split the line into segments of SEG.
template <class T, int SEG>
__global__ __launch_bounds__(512) void k_solve_strided(SolveParams<T> S, uint32_t m, uint64_t nlines, uint64_t inner,
                                                        uint64_t outer, uint64_t stride) {
    __shared__ double sA[8][64], sB[8][64];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / 64, nw = blockDim.x / 64;
    const uint64_t q = (uint64_t)blockIdx.x * 64 + lane;
    const bool act = q < nlines;
    const uint64_t base = act ? (q / inner) * outer + (q % inner) : 0;
    const uint32_t seg = (m + nw - 1) / nw;
    // wave-uniform segment bounds: the factor loads become scalar loads
    const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)(wv * seg));
    const uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(m, b + seg));
    double v[SEG];
#pragma unroll
    for (int i = 0; i < SEG; i++) v[i] = (act && b + i < e) ? S.f[base + (uint64_t)(b + i) * stride] : 0.0;
    double A = 0.0, Bm = 1.0;
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (b + i < e) {
            const double iv = S.inv[b + i];
            A = (v[i] - A) * iv;
            Bm = -Bm * iv;
        }
    }
    sA[wv][lane] = A;
    sB[wv][lane] = Bm;
    __syncthreads();
    double y = 0.0;
    for (uint32_t w = 0; w < wv; w++) y = sA[w][lane] + sB[w][lane] * y;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (b + i < e) {
            y = (v[i] - y) * S.inv[b + i];
            v[i] = y;
        }
    }
    double C = 0.0, Dm = 1.0;
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        if (b + i < e) {
            const double c = (b + i == m - 1) ? 0.0 : S.cp[b + i];
            C = v[i] - c * C;
            Dm = -c * Dm;
        }
    }
    sA[wv][lane] = C;
    sB[wv][lane] = Dm;
    __syncthreads();
    double z = 0.0;
    for (int w = (int)nw - 1; w > (int)wv; w--) z = sA[w][lane] + sB[w][lane] * z;
#pragma unroll
    for (int i = SEG - 1; i >= 0; i--) {
        if (b + i < e) {
            const double c = (b + i == m - 1) ? 0.0 : S.cp[b + i];
            z = v[i] - c * z;
            v[i] = z;
        }
    }
    if (act) {
#pragma unroll
        for (int i = 0; i < SEG; i++)
            if (b + i < e) epilogue(S, base + (uint64_t)(b + i) * stride, v[i]);
    }
}

// ---- recompose interpolation: u = I u_c + W ------------------------------------
template <class T, bool ILV>
__global__ __launch_bounds__(256) void k_interp(LoadParams<T> P, const T *__restrict__ uc, T *__restrict__ out,
                                                uint64_t os0, uint64_t os1) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    const uint32_t i1 = blockIdx.y;
    const uint32_t i0 = blockIdx.z;
    if (i2 >= P.n[2]) return;
    const uint32_t o0 = (P.n[0] == 1) ? 0 : (i0 & 1), o1 = (P.n[1] == 1) ? 0 : (i1 & 1), o2 = (P.n[2] == 1) ? 0 : (i2 & 1);
    const uint32_t j0 = i0 >> 1, j1 = i1 >> 1, j2 = i2 >> 1;
    const uint32_t *c = P.c;
    auto UC = [&](uint32_t a, uint32_t b, uint32_t d) -> T { return uc[((uint64_t)a * c[1] + b) * c[2] + d]; };
    T v;
    const uint32_t cnt = o0 + o1 + o2;
    if (cnt == 0) {
        v = UC(j0, j1, j2);
    } else {
        T I;
        if (cnt == 1) {
            if (o2) I = (UC(j0, j1, j2) + UC(j0, j1, j2 + 1)) * (T)0.5;
            else if (o1) I = (UC(j0, j1, j2) + UC(j0, j1 + 1, j2)) * (T)0.5;
            else I = (UC(j0, j1, j2) + UC(j0 + 1, j1, j2)) * (T)0.5;
        } else if (cnt == 2) {
            if (!o0) I = ((UC(j0, j1, j2) + UC(j0, j1, j2 + 1)) + (UC(j0, j1 + 1, j2) + UC(j0, j1 + 1, j2 + 1))) * (T)0.25;
            else if (!o1) I = ((UC(j0, j1, j2) + UC(j0, j1, j2 + 1)) + (UC(j0 + 1, j1, j2) + UC(j0 + 1, j1, j2 + 1))) * (T)0.25;
            else I = ((UC(j0, j1, j2) + UC(j0, j1 + 1, j2)) + (UC(j0 + 1, j1, j2) + UC(j0 + 1, j1 + 1, j2))) * (T)0.25;
        } else {
            I = (((UC(j0, j1, j2) + UC(j0, j1, j2 + 1)) + (UC(j0, j1 + 1, j2) + UC(j0, j1 + 1, j2 + 1))) +
                 ((UC(j0 + 1, j1, j2) + UC(j0 + 1, j1, j2 + 1)) + (UC(j0 + 1, j1 + 1, j2) + UC(j0 + 1, j1 + 1, j2 + 1)))) *
                (T)0.125;
        }
        const T w = ILV ? P.coef_ilv[coef_index_ilv(P, i0, i1, i2)] : P.coef_full[coef_index_full(P, i0, i1, i2)];
        v = (T)(w + I);
    }
    out[(uint64_t)i0 * os0 + (uint64_t)i1 * os1 + i2] = v;
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
    w.tab = al(w.f + p1 * sizeof(double));
    size_t tab = 0;
    for (int l = 1; l <= g.L; l++)
        for (int d = 0; d < 3; d++) tab += 2 * g.ld[l - 1][d];
    w.total = al(w.tab + tab * sizeof(double));
    return w;
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
        S.c[0] = c[0]; S.c[1] = c[1]; S.c[2] = c[2];
        S.inv = tab_dev + F.off[l][d];
        S.cp = S.inv + c[d];
        S.epi = (d == last) ? epi : kEpiNone;
        S.uc_in = uc_in;
        S.uc_out = uc_out;
        const uint32_t m = c[d];
        if (d == 2) {
            const uint64_t nlines = (uint64_t)c[0] * c[1];
            const uint32_t grid = div_up(nlines, 4);
            const uint32_t seg = (m + 63) / 64;
            if (seg <= 4) hipLaunchKernelGGL((k_solve_contig<T, 4>), dim3(grid), dim3(256), 0, s, S, nlines);
            else if (seg <= 8) hipLaunchKernelGGL((k_solve_contig<T, 8>), dim3(grid), dim3(256), 0, s, S, nlines);
            else if (seg <= 16) hipLaunchKernelGGL((k_solve_contig<T, 16>), dim3(grid), dim3(256), 0, s, S, nlines);
            else { set_error("line too long for the contiguous solve"); return kErrUnsupported; }
        } else {
            uint64_t nlines, inner, outer, stride;
            if (d == 1) { nlines = (uint64_t)c[0] * c[2]; inner = c[2]; outer = (uint64_t)c[1] * c[2]; stride = c[2]; }
            else { nlines = (uint64_t)c[1] * c[2]; inner = nlines; outer = 0; stride = (uint64_t)c[1] * c[2]; }
            // up to 8 waves split a line; each lane keeps its segment in registers
            uint32_t nw = std::min<uint32_t>(8, (m + 15) / 16);
            if (nw < 1) nw = 1;
            const uint32_t seg = (m + nw - 1) / nw;
            const uint32_t grid = div_up(nlines, 64);
            if (seg <= 8)
                hipLaunchKernelGGL((k_solve_strided<T, 8>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, inner, outer, stride);
            else if (seg <= 16)
                hipLaunchKernelGGL((k_solve_strided<T, 16>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, inner, outer, stride);
            else if (seg <= 24)
                hipLaunchKernelGGL((k_solve_strided<T, 24>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, inner, outer, stride);
            else if (seg <= 40)
                hipLaunchKernelGGL((k_solve_strided<T, 40>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, inner, outer, stride);
            else if (seg <= 72)
                hipLaunchKernelGGL((k_solve_strided<T, 72>), dim3(grid), dim3(64 * nw), 0, s, S, m, nlines, inner, outer, stride);
            else { set_error("line too long for the strided solve"); return kErrUnsupported; }
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

template <class T, int MODE>
int launch_load(const LoadParams<T> &P, bool ilv, hipStream_t s) {
    const uint32_t t2n = (P.n[2] == 1) ? 1 : div_up(P.c[2], kTJ2);
    const uint32_t t1n = (P.n[1] == 1) ? 1 : div_up(P.c[1], kTJ1);
    const uint32_t t0n = (P.n[0] == 1) ? 1 : div_up(P.c[0], kTJ0);
    const uint32_t grid = t2n * t1n * t0n;
    if (ilv) hipLaunchKernelGGL((k_load<T, MODE, true>), dim3(grid), dim3(kLoadThreads), 0, s, P);
    else hipLaunchKernelGGL((k_load<T, MODE, false>), dim3(grid), dim3(kLoadThreads), 0, s, P);
    PRODM_CHECK_LAUNCH();
    return kOk;
}

}  // namespace

size_t mgard_workspace_bytes(const LevelGeom &g, size_t elem) { return ws_layout(g, elem).total; }

template <class T>
int mgard_decompose(const T *in, const uint64_t *in_strides, const LevelGeom &g, const CoefSink<T> &sink,
                    KeyOf<T> *level_keys, void *ws, size_t ws_bytes, hipStream_t s, StageTimer *tm) {
    const WsLayout w = ws_layout(g, sizeof(T));
    if (ws_bytes < w.total) { set_error("decompose: workspace too small"); return kErrArg; }
    char *base = static_cast<char *>(ws);
    T *xa = reinterpret_cast<T *>(base + w.xa), *xb = reinterpret_cast<T *>(base + w.xb);
    double *f = reinterpret_cast<double *>(base + w.f);
    double *tab = reinterpret_cast<double *>(base + w.tab);
    const Factors F = make_factors(g);
    if (!F.host.empty()) PRODM_HIP(hipMemcpyAsync(tab, F.host.data(), F.host.size() * 8, hipMemcpyHostToDevice, s));
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
        P.key = level_keys ? level_keys + l : nullptr;
        int rc;
        {
            StageScope sc(tm, l == g.L ? kStDecLoadTop : kStDecLoadRest);
            rc = launch_load<T, kDecompose>(P, ilv, s);
        }
        if (rc) return rc;
        {
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
        if (g.L == 0) return copy_box<T>(in, in_strides[0], in_strides[1], sink.level[0], (uint64_t)c0[1] * c0[2], c0[2], c0, s);
        PRODM_HIP(hipMemcpyAsync(sink.level[0], cur, points3(c0) * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (level_keys) return launch_maxabs<T>(sink.level[0], points3(c0), level_keys + 0, s);
        return kOk;
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
                                    uint32_t *, void *, size_t, hipStream_t, StageTimer *);
template int mgard_decompose<double>(const double *, const uint64_t *, const LevelGeom &, const CoefSink<double> &,
                                     unsigned long long *, void *, size_t, hipStream_t, StageTimer *);

template <class T>
int mgard_recompose(const CoefSink<T> &src, const LevelGeom &g, int /*unused*/, T *out, const uint64_t *out_strides,
                    void *ws, size_t ws_bytes, hipStream_t s, StageTimer *tm) {
    const WsLayout w = ws_layout(g, sizeof(T));
    if (ws_bytes < w.total) { set_error("recompose: workspace too small"); return kErrArg; }
    char *base = static_cast<char *>(ws);
    T *bufs[2] = {reinterpret_cast<T *>(base + w.xa), reinterpret_cast<T *>(base + w.xb)};
    double *f = reinterpret_cast<double *>(base + w.f);
    double *tab = reinterpret_cast<double *>(base + w.tab);
    const Factors F = make_factors(g);
    if (!F.host.empty()) PRODM_HIP(hipMemcpyAsync(tab, F.host.data(), F.host.size() * 8, hipMemcpyHostToDevice, s));
    const bool ilv = src.full == nullptr;
    const uint32_t *c0 = g.ld[0];
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
        LoadParams<T> P = load_params(g, l, src);
        P.f = f;
        {
            StageScope sc(tm, l == g.L ? kStRecLoadTop : kStRecLoadRest);
            rc = launch_load<T, kRecompose>(P, ilv, s);
        }
        if (rc) return rc;
        {
            StageScope sc(tm, l == g.L ? kStRecSolveTop : kStRecSolveRest);
            rc = solve_all<T>(g, l, f, tab, F, kEpiSub, cur, cur, s);  // u_c = u_c' - z, in place
        }
        if (rc) return rc;
        const uint32_t *n = g.ld[l];
        T *dst;
        uint64_t d0, d1;
        if (l == g.L) { dst = out; d0 = out_strides[0]; d1 = out_strides[1]; }
        else { dst = bufs[which ^ 1]; d0 = (uint64_t)n[1] * n[2]; d1 = n[2]; }
        dim3 grid(div_up(n[2], 256), n[1], n[0]);
        {
            StageScope sc(tm, l == g.L ? kStRecInterpTop : kStRecInterpRest);
            if (ilv) hipLaunchKernelGGL((k_interp<T, true>), grid, dim3(256), 0, s, P, cur, dst, d0, d1);
            else hipLaunchKernelGGL((k_interp<T, false>), grid, dim3(256), 0, s, P, cur, dst, d0, d1);
        }
        PRODM_CHECK_LAUNCH();
        which ^= 1;
        cur = bufs[which];
    }
    return kOk;
}
template int mgard_recompose<float>(const CoefSink<float> &, const LevelGeom &, int, float *, const uint64_t *, void *,
                                    size_t, hipStream_t, StageTimer *);
template int mgard_recompose<double>(const CoefSink<double> &, const LevelGeom &, int, double *, const uint64_t *,
                                     void *, size_t, hipStream_t, StageTimer *);

}  // namespace prodm

// adjoint.hip -- transpose of the MGARD orthogonal recompose (FP64).
//
// One recompose level maps (u_c', W) to u = P (u_c' - S R W) + E W, with P
// the multilinear prolongation (injection at coarse nodes), R = P^T M the
// load operator (M: fine mass matrix, 3/h-scaled [1/2, 2, 1/2] per dim, ends
// 1), S = M_c^-1 the coarse solve and E the embedding of the coefficients.
// Its transpose, given g = dJ/du on the fine grid:
//   h   = P^T g                     (coarse; the next coarser level's g)
//   t   = S h                       (coarse line solves, S symmetric)
//   r   = M P t                     (fine)
//   dJ/dW = g - r                   at the coefficient nodes
// applied from the finest level to the coarsest, where dJ/du_0 = h.
// The result is laid out like a decompose output (coarse nodes first along
// each dim inside every level box). ProDM has no adjoint; this serves
// QoI-guided retrieval (sensitivities of a QoI to each level's coefficients).
#include "internal.hpp"

#include <vector>

namespace prodm {

namespace {

__global__ __launch_bounds__(256) void k_to_double_f32(const float *__restrict__ in, uint64_t s0, uint64_t s1,
                                                        double *__restrict__ out, uint32_t n1, uint32_t n2) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    out[((uint64_t)blockIdx.z * n1 + blockIdx.y) * n2 + i2] = (double)in[(uint64_t)blockIdx.z * s0 + (uint64_t)blockIdx.y * s1 + i2];
}
__global__ __launch_bounds__(256) void k_to_double_f64(const double *__restrict__ in, uint64_t s0, uint64_t s1,
                                                        double *__restrict__ out, uint32_t n1, uint32_t n2) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    out[((uint64_t)blockIdx.z * n1 + blockIdx.y) * n2 + i2] = in[(uint64_t)blockIdx.z * s0 + (uint64_t)blockIdx.y * s1 + i2];
}

// v = P t: multilinear prolongation of the coarse grid t (c, rows of pitch
// tp) to the dense fine grid v (n), same corner formula as the recompose
// interpolation
__global__ __launch_bounds__(256) void k_prolong(const double *__restrict__ t, uint32_t c1, uint32_t tp,
                                                  double *__restrict__ v, uint32_t n0, uint32_t n1, uint32_t n2) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    const uint32_t i1 = blockIdx.y, i0 = blockIdx.z;
    const uint32_t o0 = n0 == 1 ? 0 : (i0 & 1), o1 = n1 == 1 ? 0 : (i1 & 1), o2 = n2 == 1 ? 0 : (i2 & 1);
    const uint32_t a0 = i0 >> 1, b0 = a0 + o0, a1 = i1 >> 1, b1 = a1 + o1, a2 = i2 >> 1, b2 = a2 + o2;
    auto T = [&](uint32_t j0, uint32_t j1, uint32_t j2) { return t[((uint64_t)j0 * c1 + j1) * tp + j2]; };
    const double I = (((T(a0, a1, a2) + T(a0, a1, b2)) + (T(a0, b1, a2) + T(a0, b1, b2))) +
                      ((T(b0, a1, a2) + T(b0, a1, b2)) + (T(b0, b1, a2) + T(b0, b1, b2)))) *
                     0.125;
    v[((uint64_t)i0 * n1 + i1) * n2 + i2] = I;
}

// out = M_d in along dim d of a dense fine grid: [1/2, 2, 1/2], ends 1
__global__ __launch_bounds__(256) void k_mass(const double *__restrict__ in, double *__restrict__ out, uint32_t n0,
                                               uint32_t n1, uint32_t n2, int d) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    const uint32_t i1 = blockIdx.y, i0 = blockIdx.z;
    const uint32_t n[3] = {n0, n1, n2}, i[3] = {i0, i1, i2};
    const uint64_t st[3] = {(uint64_t)n1 * n2, n2, 1};
    const uint64_t k = (uint64_t)i0 * st[0] + (uint64_t)i1 * st[1] + i2;
    const uint32_t id = i[d], nd = n[d];
    double a = (id == 0 || id == nd - 1) ? in[k] : 2.0 * in[k];
    if (id > 0) a += 0.5 * in[k - st[d]];
    if (id + 1 < nd) a += 0.5 * in[k + st[d]];
    out[k] = a;
}

// dJ/dW = g - r at the coefficient nodes of the level box, written at the
// decompose layout position (rpos per dim) of the output array
template <class T>
__global__ __launch_bounds__(256) void k_adj_store(const double *__restrict__ g, const double *__restrict__ r,
                                                    uint32_t n0, uint32_t n1, uint32_t n2, uint32_t c0, uint32_t c1,
                                                    uint32_t c2, T *__restrict__ out, uint64_t os0, uint64_t os1) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    const uint32_t i1 = blockIdx.y, i0 = blockIdx.z;
    const bool coarse = (n0 == 1 || !(i0 & 1)) && (n1 == 1 || !(i1 & 1)) && (n2 == 1 || !(i2 & 1));
    if (coarse) return;
    const uint64_t k = ((uint64_t)i0 * n1 + i1) * n2 + i2;
    out[(uint64_t)rpos(i0, c0) * os0 + (uint64_t)rpos(i1, c1) * os1 + rpos(i2, c2)] = (T)(g[k] - r[k]);
}

// the coarsest grid's adjoint into the level-0 box of the output
template <class T>
__global__ __launch_bounds__(256) void k_adj_store0(const double *__restrict__ h, uint32_t n1, uint32_t n2,
                                                     T *__restrict__ out, uint64_t os0, uint64_t os1) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    if (i2 >= n2) return;
    const uint32_t i1 = blockIdx.y, i0 = blockIdx.z;
    out[(uint64_t)i0 * os0 + (uint64_t)i1 * os1 + i2] = (T)h[((uint64_t)i0 * n1 + i1) * n2 + i2];
}

dim3 grid_of(const uint32_t *n) { return dim3(div_up(n[2], 256), n[1], n[0]); }

struct AdjLayout {
    size_t g, h0, h1, v1, v2, mgard, total;
};
AdjLayout adj_layout(const LevelGeom &g) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const uint64_t pL = g.points[g.L], p1 = g.L >= 1 ? g.points[g.L - 1] : 1;
    AdjLayout a{};
    a.g = 0;
    a.h0 = al(a.g + pL * 8);
    a.h1 = al(a.h0 + p1 * 8);
    a.v1 = al(a.h1 + p1 * 8);
    a.v2 = al(a.v1 + pL * 8);
    a.mgard = al(a.v2 + pL * 8);
    a.total = a.mgard + mgard_workspace_bytes(g, 8);
    return a;
}

}  // namespace

size_t adjoint_workspace_bytes(const LevelGeom &g) { return adj_layout(g).total; }

template <class T>
int mgard_recompose_adjoint(const T *g_in, const uint64_t *in_strides, const LevelGeom &g, T *out,
                            const uint64_t *out_strides, void *ws, size_t ws_bytes, hipStream_t s) {
    const AdjLayout A = adj_layout(g);
    if (ws_bytes < A.total) { set_error("adjoint: workspace too small"); return kErrArg; }
    char *base = static_cast<char *>(ws);
    double *gd = reinterpret_cast<double *>(base + A.g);
    double *hb[2] = {reinterpret_cast<double *>(base + A.h0), reinterpret_cast<double *>(base + A.h1)};
    double *v1 = reinterpret_cast<double *>(base + A.v1), *v2 = reinterpret_cast<double *>(base + A.v2);
    void *mws = base + A.mgard;
    const size_t mws_bytes = A.total - A.mgard;
    if (sizeof(T) == 4)
        hipLaunchKernelGGL(k_to_double_f32, grid_of(g.ld[g.L]), dim3(256), 0, s, reinterpret_cast<const float *>(g_in),
                           in_strides[0], in_strides[1], gd, g.ld[g.L][1], g.ld[g.L][2]);
    else
        hipLaunchKernelGGL(k_to_double_f64, grid_of(g.ld[g.L]), dim3(256), 0, s, reinterpret_cast<const double *>(g_in),
                           in_strides[0], in_strides[1], gd, g.ld[g.L][1], g.ld[g.L][2]);
    PRODM_CHECK_LAUNCH();
    const double *cur = gd;
    int which = 0;
    for (int l = g.L; l >= 1; l--) {
        const uint32_t *n = g.ld[l], *c = g.ld[l - 1];
        double *h = hb[which];
        // h = P^T g and t = S h (the transposed load and the coarse solves)
        if (int rc = mgard_adjoint_coarse(cur, g, l, h, mws, mws_bytes, s)) return rc;
        uint32_t tp = 0;
        const double *t = mgard_adjoint_solution(g, l, mws, &tp);
        // r = M P t, then dJ/dW = g - r at the coefficient nodes
        hipLaunchKernelGGL(k_prolong, grid_of(n), dim3(256), 0, s, t, c[1], tp, v1, n[0], n[1], n[2]);
        hipLaunchKernelGGL(k_mass, grid_of(n), dim3(256), 0, s, v1, v2, n[0], n[1], n[2], 2);
        hipLaunchKernelGGL(k_mass, grid_of(n), dim3(256), 0, s, v2, v1, n[0], n[1], n[2], 1);
        hipLaunchKernelGGL(k_mass, grid_of(n), dim3(256), 0, s, v1, v2, n[0], n[1], n[2], 0);
        hipLaunchKernelGGL((k_adj_store<T>), grid_of(n), dim3(256), 0, s, cur, v2, n[0], n[1], n[2], c[0], c[1], c[2],
                           out, out_strides[0], out_strides[1]);
        PRODM_CHECK_LAUNCH();
        cur = h;
        which ^= 1;
    }
    hipLaunchKernelGGL((k_adj_store0<T>), grid_of(g.ld[0]), dim3(256), 0, s, cur, g.ld[0][1], g.ld[0][2], out,
                       out_strides[0], out_strides[1]);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int mgard_recompose_adjoint<float>(const float *, const uint64_t *, const LevelGeom &, float *,
                                            const uint64_t *, void *, size_t, hipStream_t);
template int mgard_recompose_adjoint<double>(const double *, const uint64_t *, const LevelGeom &, double *,
                                             const uint64_t *, void *, size_t, hipStream_t);

}  // namespace prodm

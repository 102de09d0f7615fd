// interleave.hip -- DirectInterleaver (DirectInterleaver.hpp:12-63) on gfx950.
//
// Every (i, j, k) of box(fine) \ box(coarse), row-major, maps to a closed-form
// position in the level buffer (ilv_index), so gather and scatter are one
// thread per element with coalesced accesses along k. Dims are padded to 3;
// nd = 2 is the natural extension of the reference's 1-D / 3-D branches.
#include "internal.hpp"

namespace prodm {

template <class T, bool REPOS>
__global__ __launch_bounds__(256) void k_interleave(const T *__restrict__ src, T *__restrict__ dst, uint64_t s0,
                                                    uint64_t s1, uint32_t f0, uint32_t f1, uint32_t f2, uint32_t c0,
                                                    uint32_t c1, uint32_t c2) {
    const uint32_t i2 = blockIdx.x * 256 + threadIdx.x;
    const uint32_t i1 = blockIdx.y, i0 = blockIdx.z;
    if (i2 >= f2) return;
    if (i0 < c0 && i1 < c1 && i2 < c2) return;
    const uint32_t f[3] = {f0, f1, f2}, c[3] = {c0, c1, c2};
    const uint64_t idx = ilv_index(i0, i1, i2, f, c);
    const uint64_t a = (uint64_t)i0 * s0 + (uint64_t)i1 * s1 + i2;
    if (REPOS) dst[a] = src[idx];
    else dst[idx] = src[a];
}

template <class T>
int launch_interleave(const T *data, const uint64_t *s, const uint32_t *f, const uint32_t *c, T *buf, bool reposition,
                      hipStream_t st) {
    if (!f[0] || !f[1] || !f[2]) return kOk;
    dim3 grid(div_up(f[2], 256), f[1], f[0]);
    if (reposition)
        hipLaunchKernelGGL((k_interleave<T, true>), grid, dim3(256), 0, st, buf, const_cast<T *>(data), s[0], s[1],
                           f[0], f[1], f[2], c[0], c[1], c[2]);
    else
        hipLaunchKernelGGL((k_interleave<T, false>), grid, dim3(256), 0, st, data, buf, s[0], s[1], f[0], f[1], f[2],
                           c[0], c[1], c[2]);
    PRODM_CHECK_LAUNCH();
    return kOk;
}
template int launch_interleave<float>(const float *, const uint64_t *, const uint32_t *, const uint32_t *, float *,
                                      bool, hipStream_t);
template int launch_interleave<double>(const double *, const uint64_t *, const uint32_t *, const uint32_t *, double *,
                                       bool, hipStream_t);

}  // namespace prodm

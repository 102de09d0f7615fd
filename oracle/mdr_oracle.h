/*
 * mdr_oracle.h -- CPU oracle for the MDR refactor/retrieve hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path (prodm_amd/, include/prodm_mdr.h) never links or calls it.
 *
 * It restates, in plain single-threaded C, the reference algorithms of ProDM's
 * MDR path (file:line citations are into the ProDM sources):
 *   - level geometry          RefactorUtils.hpp:21-59
 *   - max-abs                 RefactorUtils.hpp:68-76
 *   - interleave/reposition   Interleaver/DirectInterleaver.hpp:12-63 (+2-D)
 *   - NegaBinary encode/decode BitplaneEncoder/NegaBinaryBPEncoder.hpp:18-62,132-192,219-260
 *   - GroupedBP encode/decode  BitplaneEncoder/GroupedBPEncoder.hpp:18-73,210-297,337-377
 *   - Max error collector     ErrorCollector/MaxErrorCollector.hpp:14-25
 *   - Squared error collector ErrorCollector/SquaredErrorCollector.hpp:23-65
 *   - MGARD orthogonal decompose/recompose: the arithmetic lives in the
 *     external MGARDx library (Decomposer/MGARD.hpp:5-6,14-35), which is not
 *     available; it is restated from the published mathematics (Ainsworth,
 *     Tugluk, Whitney, Klasky, SIAM J. Sci. Comput. 2019): L2 projection
 *     with a tensor-product mass matrix, coarse nodes first along each dim.
 *   - adjoint (transpose) of recompose, for the QoI extension.
 *
 * Each function is pinned by tests/test_oracle_golden.py (the reference's
 * own fixtures), tests/test_oracle_ref.py (the compiled reference, where
 * it builds) and tests/test_oracle_math.py (closed forms, invariants and
 * brute force). MGARDx's own operation order and rounding cannot be
 * compared (the library is absent): that part is parity unpinned; the
 * mathematics above is pinned.
 */
#ifndef PRODM_MDR_ORACLE_H
#define PRODM_MDR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry (RefactorUtils.hpp:21-59) ------------------------------- */
/* out[l*nd + d] = level dims, l = 0 (coarsest) .. L (finest)            */
void orc_level_dims(const uint32_t *dims, int nd, int L, uint32_t *out);
/* out[l] = elements of level l (box difference)                          */
void orc_level_elements(const uint32_t *level_dims, int nd, int L, uint64_t *out);

/* ---- interleave (DirectInterleaver.hpp:12-63), nd = 1, 2, 3 ------------ */
void orc_interleave_f32(const float *data, const uint32_t *dims, int nd,
                        const uint32_t *fine, const uint32_t *coarse,
                        const uint32_t *strides, float *buffer);
void orc_interleave_f64(const double *data, const uint32_t *dims, int nd,
                        const uint32_t *fine, const uint32_t *coarse,
                        const uint32_t *strides, double *buffer);
void orc_reposition_f32(const float *buffer, const uint32_t *dims, int nd,
                        const uint32_t *fine, const uint32_t *coarse,
                        const uint32_t *strides, float *data);
void orc_reposition_f64(const double *buffer, const uint32_t *dims, int nd,
                        const uint32_t *fine, const uint32_t *coarse,
                        const uint32_t *strides, double *data);

/* ---- max-abs + exponent (RefactorUtils.hpp:68-76, ComposedRefactor.hpp:108-121) */
float orc_maxabs_f32(const float *x, uint64_t n);
double orc_maxabs_f64(const double *x, uint64_t n);
int orc_frexp_exp_f32(float x);
int orc_frexp_exp_f64(double x);

/* ---- NegaBinary bitplanes: planes[p * nwords + b], nwords = ceil(n/B) --- */
int orc_nb_encode_f32(const float *x, uint32_t n, int32_t exp, int nb, uint32_t *planes);
int orc_nb_encode_f64(const double *x, uint32_t n, int32_t exp, int nb, uint64_t *planes);
/* planes: k pointers, planes[i] is plane (start + i)                      */
int orc_nb_decode_f32(const uint32_t *const *planes, uint32_t n, int32_t exp,
                      int start, int k, float *out);
int orc_nb_decode_f64(const uint64_t *const *planes, uint32_t n, int32_t exp,
                      int start, int k, double *out);

/* ---- GroupedBP bitplanes, reference byte format ------------------------ */
/* Upper bound of the concatenated stream bytes for out_cap.               */
uint64_t orc_gb_max_bytes_f32(uint32_t n, int nb);
uint64_t orc_gb_max_bytes_f64(uint32_t n, int nb);
/* Streams are written back to back into out (stream p at sum(sizes[<p])).
 * compat != 0 reproduces the x86 saturation of (int64_t) for |v| >= 2^63. */
int orc_gb_encode_f32(const float *x, uint32_t n, int32_t exp, int nb, int compat,
                      uint8_t *out, uint64_t out_cap, uint32_t *sizes);
int orc_gb_encode_f64(const double *x, uint32_t n, int32_t exp, int nb, int compat,
                      uint8_t *out, uint64_t out_cap, uint32_t *sizes);
/* Progressive decode state, keyed by level explicitly.                    */
typedef struct orc_gb_state orc_gb_state;
orc_gb_state *orc_gb_state_new(int max_levels);
void orc_gb_state_free(orc_gb_state *s);
int orc_gb_decode_f32(orc_gb_state *s, int level, const uint8_t *const *streams,
                      uint32_t n, int32_t exp, int start, int k, float *out);
int orc_gb_decode_f64(orc_gb_state *s, int level, const uint8_t *const *streams,
                      uint32_t n, int32_t exp, int start, int k, double *out);

/* ---- PerBit (PerBitBPEncoder.hpp:72-146, 282-391) ------------------------
 * streams[p]: capacity (2n + 63) / 64 words each; sizes in bytes. Decode
 * state keyed by level; planes [start, start + k) as deltas.              */
typedef struct orc_pb_state orc_pb_state;
orc_pb_state *orc_pb_state_new(int max_levels);
void orc_pb_state_free(orc_pb_state *s);
int orc_pb_encode_f32(const float *x, uint32_t n, int32_t exp, int nb, uint64_t *const *streams, uint32_t *sizes);
int orc_pb_encode_f64(const double *x, uint32_t n, int32_t exp, int nb, uint64_t *const *streams, uint32_t *sizes);
int orc_pb_decode_f32(orc_pb_state *s, int level, const uint64_t *const *streams, uint32_t n, int32_t exp, int start,
                      int k, float *out);
int orc_pb_decode_f64(orc_pb_state *s, int level, const uint64_t *const *streams, uint32_t n, int32_t exp, int start,
                      int k, double *out);

/* ---- error collectors ---------------------------------------------------- */
void orc_collect_max_f32(float bound, int nb, double *out);
void orc_collect_max_f64(double bound, int nb, double *out);
/* exactly rounded sums (order independent); *_seq: reference summation order */
void orc_collect_squared_f32(const float *x, uint64_t n, int nb, float maxerr, double *out);
void orc_collect_squared_f64(const double *x, uint64_t n, int nb, double maxerr, double *out);
void orc_collect_squared_seq_f32(const float *x, uint64_t n, int nb, float maxerr, double *out);
void orc_collect_squared_seq_f64(const double *x, uint64_t n, int nb, double maxerr, double *out);
/* encoder-fused level errors (NegaBinaryBPEncoder.hpp:65-125,231-239;
 * GroupedBPEncoder.hpp:76-145,324-335), fp32 data, nb <= 32: exactly rounded
 * sums of the reference's per-element terms; *_seq: reference order.      */
void orc_nb_level_errors_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out);
void orc_nb_level_errors_seq_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out);
void orc_gb_level_errors_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out);
void orc_gb_level_errors_seq_f32(const float *x, uint64_t n, int32_t exp, int nb, double *out);

/* ---- MGARD orthogonal decompose / recompose (in place, strided box) ----- */
/* strides may be NULL (dense). Returns 0, or -1 for unsupported geometry.  */
int orc_decompose_f32(float *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
int orc_decompose_f64(double *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
int orc_recompose_f32(float *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
int orc_recompose_f64(double *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
/* Hierarchical basis (MGARDHierarchicalDecomposer, Decomposer/MGARD.hpp:42-70):
 * the same steps without the L2 correction.                              */
int orc_decompose_hb_f32(float *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
int orc_decompose_hb_f64(double *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
int orc_recompose_hb_f32(float *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
int orc_recompose_hb_f64(double *data, const uint32_t *dims, int nd, int L, const uint32_t *strides);
/* Adjoint (transpose) of recompose as a linear map on FP64 arrays:
 * given g = dJ/du (full grid), returns dJ/dcoef in the decomposed layout.  */
int orc_recompose_adjoint_f64(double *data, const uint32_t *dims, int nd, int L,
                              const uint32_t *strides);
/* geometry check used by both paths: every active dim odd at every level  */
int orc_geometry_ok(const uint32_t *dims, int nd, int L);

#ifdef __cplusplus
}
#endif
#endif

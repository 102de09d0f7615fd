/*
 * prodm_mdr.h -- C ABI of the MI355X (gfx950) MDR refactor/retrieve library
 * (libprodm_mdr.so). Plain pointers and sizes only; device pointers are HIP
 * device memory, `stream` is a hipStream_t (NULL = default stream). Every
 * entry point is asynchronous on `stream` unless stated otherwise, returns
 * PRODM_OK (0) or a negative status, and never exits the process; the
 * message of the last failure on the calling thread is prodm_last_error().
 *
 * Each entry point names the ProDM interface it replaces (paths relative to
 * the ProDM repository root). The C++ host classes in include/prodm/ (header-only)
 * wrap these calls behind the same component concepts that ComposedRefactor
 * (include/MDR/Refactor/ComposedRefactor.hpp:15-19) and ComposedReconstructor
 * (include/MDR/Reconstructor/ComposedReconstructor.hpp:17-21) are templated on.
 */
#ifndef PRODM_MDR_H
#define PRODM_MDR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRODM_OK 0
#define PRODM_ERR_ARG (-1)
#define PRODM_ERR_GEOMETRY (-2)
#define PRODM_ERR_HIP (-3)
#define PRODM_ERR_NOMEM (-4)
#define PRODM_ERR_STATE (-5)
#define PRODM_ERR_UNSUPPORTED (-6)

#define PRODM_F32 0
#define PRODM_F64 1
#define PRODM_ENC_NEGABINARY 0
#define PRODM_ENC_GROUPED 1
#define PRODM_ENC_PERBIT 2

typedef void *prodm_stream_t;

int prodm_version(void);
const char *prodm_last_error(void);

/* Level geometry, RefactorUtils.hpp:21-59 (compute_level_dims /
 * compute_level_elements). out_dims: (L+1)*ndims, level 0 coarsest.
 * Returns 1 when every dim > 1 stays odd down to level 0 (the supported
 * decomposition geometry), 0 otherwise; host only.                       */
int prodm_level_geometry(const uint32_t *dims, int ndims, int target_level, uint32_t *out_dims,
                         uint64_t *out_elements);

/* ---- Decomposer: MGARDOrthoganalDecomposer<T>::decompose / recompose
 *      (include/MDR/Decomposer/MGARD.hpp:14-35). Out of place on device
 *      arrays with the same dims/strides; strides NULL = dense.          */
size_t prodm_mgard_workspace_bytes(const uint32_t *dims, int ndims, int target_level, int dtype);
int prodm_decompose_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_decompose_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_recompose_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_recompose_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);

/* ---- Decomposer: MGARDHierarchicalDecomposer<T>::decompose / recompose
 *      (include/MDR/Decomposer/MGARD.hpp:42-70): the same level steps
 *      without the L2 correction. Same arguments and layout as above.     */
int prodm_decompose_hierarchical_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                                     const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_decompose_hierarchical_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims,
                                     int target_level, const uint32_t *strides, void *d_ws, size_t ws_bytes,
                                     prodm_stream_t stream);
int prodm_recompose_hierarchical_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                                     const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_recompose_hierarchical_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims,
                                     int target_level, const uint32_t *strides, void *d_ws, size_t ws_bytes,
                                     prodm_stream_t stream);

/* ---- Adjoint (transpose) of recompose, FP64 inside. No counterpart in
 *      ProDM (whose reconstructor only applies recompose,
 *      ComposedReconstructor.hpp:214,250-253); it gives the sensitivity of
 *      a quantity of interest J to every stored coefficient: d_g = dJ/du on
 *      the box (dims, strides), d_out = dJ/dcoefficients in the layout of
 *      prodm_decompose's output (same dims/strides).                      */
size_t prodm_recompose_adjoint_workspace_bytes(const uint32_t *dims, int ndims, int target_level);
int prodm_recompose_adjoint_f32(const float *d_g, float *d_out, const uint32_t *dims, int ndims, int target_level,
                                const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_recompose_adjoint_f64(const double *d_g, double *d_out, const uint32_t *dims, int ndims, int target_level,
                                const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream);

/* ---- Interleaver: DirectInterleaver<T>::interleave / reposition
 *      (include/MDR/Interleaver/DirectInterleaver.hpp:12-63); ndims 1..3. */
int prodm_interleave_f32(const float *d_data, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, float *d_buffer, prodm_stream_t stream);
int prodm_interleave_f64(const double *d_data, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, double *d_buffer, prodm_stream_t stream);
int prodm_reposition_f32(const float *d_buffer, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, float *d_data, prodm_stream_t stream);
int prodm_reposition_f64(const double *d_buffer, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, double *d_data, prodm_stream_t stream);

/* ---- compute_max_abs_value (RefactorUtils.hpp:68-76). Synchronizes the
 *      stream and returns the value on the host.                        */
int prodm_level_maxabs_f32(const float *d_x, uint64_t n, float *h_out, prodm_stream_t stream);
int prodm_level_maxabs_f64(const double *d_x, uint64_t n, double *h_out, prodm_stream_t stream);

/* ---- NegaBinaryBPEncoder<T, S>::encode / progressive_decode
 *      (include/MDR/BitplaneEncoder/NegaBinaryBPEncoder.hpp:18-62,132-192).
 *      Plane p (p = 0 is the most significant) is stored at
 *      d_planes + p * plane_stride_words and holds ceil(n / B) words
 *      (B = 32 for f32/u32, 64 for f64/u64); for u32 words the stride
 *      must be even and >= ceil(n / 32) rounded up to even.
 *      Decode: h_plane_ptrs[i] (host array of k device pointers) is plane
 *      start + i; k == 0 writes zeros.                                    */
int prodm_negabinary_encode_f32(const float *d_x, uint32_t n, int32_t exp, int num_bitplanes, uint32_t *d_planes,
                                uint64_t plane_stride_words, prodm_stream_t stream);
int prodm_negabinary_encode_f64(const double *d_x, uint32_t n, int32_t exp, int num_bitplanes, uint64_t *d_planes,
                                uint64_t plane_stride_words, prodm_stream_t stream);
int prodm_negabinary_decode_f32(const uint32_t *const *h_plane_ptrs, uint32_t n, int32_t exp, int start, int k,
                                float *d_out, prodm_stream_t stream);
int prodm_negabinary_decode_f64(const uint64_t *const *h_plane_ptrs, uint32_t n, int32_t exp, int start, int k,
                                double *d_out, prodm_stream_t stream);

/* ---- GroupedBPEncoder<T, S>::encode / progressive_decode
 *      (include/MDR/BitplaneEncoder/GroupedBPEncoder.hpp:18-73,210-297).
 *      Encode writes the per-block recording planes d_rb[nblocks], the
 *      compacted words of plane p at d_planes + p * plane_stride_words
 *      (stride >= 2 * nblocks) and d_words[p] = words in plane p. The
 *      reference stream p is those words; stream 0 is prefixed with
 *      u32 nblocks and the nblocks rb bytes. compat != 0 reproduces the
 *      x86 result of (int64_t) on |v| >= 2^63 (nb = 64 only).
 *      Decode takes k device pointers to reference-format streams
 *      (stream 0 carries the header when start == 0) and keeps the
 *      per-level state (rb, signs) in `state`, keyed by `level`.         */
size_t prodm_grouped_workspace_bytes(uint32_t n, int num_bitplanes, int dtype);
int prodm_grouped_encode_f32(const float *d_x, uint32_t n, int32_t exp, int num_bitplanes, int compat, uint8_t *d_rb,
                             uint32_t *d_planes, uint64_t plane_stride_words, uint32_t *d_words, void *d_ws,
                             size_t ws_bytes, prodm_stream_t stream);
int prodm_grouped_encode_f64(const double *d_x, uint32_t n, int32_t exp, int num_bitplanes, int compat, uint8_t *d_rb,
                             uint64_t *d_planes, uint64_t plane_stride_words, uint32_t *d_words, void *d_ws,
                             size_t ws_bytes, prodm_stream_t stream);
/* Allocation-free variant of prodm_grouped_encode_f32 (same arguments, same
 * results, no workspace). No counterpart in ProDM: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). The encode's per-workgroup
 * histograms and plane offsets live in an 8 MiB device array of the library
 * that each GPU holds until the process ends (its address is looked up once
 * per device on the first call); n elements with nb planes need
 * prodm_grouped_workspace_bytes(n, nb, PRODM_F32) bytes of it (about 3.8 MB
 * at 118 M elements and 32 planes; the limit is about 264 M elements at 32
 * planes), larger requests return PRODM_ERR_UNSUPPORTED. Calls that overlap on one GPU share the array and
 * race: serialize them.                                                   */
int prodm_grouped_encode_noalloc_f32(const float *d_x, uint32_t n, int32_t exp, int num_bitplanes, int compat,
                                     uint8_t *d_rb, uint32_t *d_planes, uint64_t plane_stride_words, uint32_t *d_words,
                                     prodm_stream_t stream);
typedef struct prodm_grouped_state prodm_grouped_state;
int prodm_grouped_state_create(prodm_grouped_state **out, int max_levels, int dtype);  /* max_levels 1..64 */
int prodm_grouped_state_destroy(prodm_grouped_state *state);
int prodm_grouped_decode_f32(prodm_grouped_state *state, int level, const uint8_t *const *h_stream_ptrs, uint32_t n,
                             int32_t exp, int start, int k, float *d_out, prodm_stream_t stream);
int prodm_grouped_decode_f64(prodm_grouped_state *state, int level, const uint8_t *const *h_stream_ptrs, uint32_t n,
                             int32_t exp, int start, int k, double *d_out, prodm_stream_t stream);

/* ---- PerBitBPEncoder<T, S>::encode / progressive_decode
 *      (include/MDR/BitplaneEncoder/PerBitBPEncoder.hpp:81-143,290-382).
 *      Stream p (p = 0 holds the most significant bit) carries, element
 *      by element, the element's bit and, right after its first 1 bit,
 *      its sign, in 64-bit words filled LSB first. Encode (device) writes
 *      stream p at d_streams + p * stride_words (stride_words >=
 *      prodm_perbit_stream_words(n)) and d_sizes[p] = its bytes (whole
 *      words), as the reference's stream_sizes; compat as for GroupedBP
 *      (x86's (int64_t) on |v| >= 2^63 or NaN). Decode runs on the host
 *      (the format has no index of where an element's bits begin):
 *      h_streams[i] is stream start + i in host memory, h_sizes[i] its
 *      bytes (h_sizes may be NULL: no bound is then checked); the
 *      per-level sign state lives in `state`, keyed by `level`; k == 0
 *      writes zeros. The encoder's level errors (:151-224) are those of
 *      GroupedBP (same per-element terms): prodm_grouped_level_errors_f32. */
uint64_t prodm_perbit_stream_words(uint32_t n);
size_t prodm_perbit_workspace_bytes(uint32_t n, int num_bitplanes, int dtype);
int prodm_perbit_encode_f32(const float *d_x, uint32_t n, int32_t exp, int num_bitplanes, int compat,
                            uint64_t *d_streams, uint64_t stride_words, uint32_t *d_sizes, void *d_ws,
                            size_t ws_bytes, prodm_stream_t stream);
int prodm_perbit_encode_f64(const double *d_x, uint32_t n, int32_t exp, int num_bitplanes, int compat,
                            uint64_t *d_streams, uint64_t stride_words, uint32_t *d_sizes, void *d_ws,
                            size_t ws_bytes, prodm_stream_t stream);
typedef struct prodm_perbit_state prodm_perbit_state;
int prodm_perbit_state_create(prodm_perbit_state **out);
int prodm_perbit_state_destroy(prodm_perbit_state *state);
int prodm_perbit_decode_f32(prodm_perbit_state *state, int level, const uint64_t *const *h_streams,
                            const uint32_t *h_sizes, uint32_t n, int32_t exp, int start, int k, float *h_out);
int prodm_perbit_decode_f64(prodm_perbit_state *state, int level, const uint64_t *const *h_streams,
                            const uint32_t *h_sizes, uint32_t n, int32_t exp, int start, int k, double *h_out);

/* ---- MaxErrorCollector<T>::collect_level_error
 *      (include/MDR/ErrorCollector/MaxErrorCollector.hpp:14-25); host,
 *      closed form, h_out[0..nb].                                        */
void prodm_collect_max_f32(float max_level_error, int num_bitplanes, double *h_out);
void prodm_collect_max_f64(double max_level_error, int num_bitplanes, double *h_out);

/* ---- SquaredErrorCollector<T>::collect_level_error
 *      (include/MDR/ErrorCollector/SquaredErrorCollector.hpp:23-65).
 *      Each table entry is the exactly rounded sum of the reference's
 *      per-element terms. Synchronizes the stream; h_out[0..nb].          */
size_t prodm_collect_squared_workspace_bytes(uint64_t n, int num_bitplanes, int dtype);
int prodm_collect_squared_f32(const float *d_x, uint64_t n, int num_bitplanes, float max_level_error, double *h_out,
                              void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_collect_squared_f64(const double *d_x, uint64_t n, int num_bitplanes, double max_level_error, double *h_out,
                              void *d_ws, size_t ws_bytes, prodm_stream_t stream);

/* ---- encoder-fused level errors: the level_errors of
 *      NegaBinaryBPEncoder::encode(..., level_errors)
 *      (NegaBinaryBPEncoder.hpp:65-125, collect_level_errors :231-239) and
 *      GroupedBPEncoder::encode(..., level_errors) (GroupedBPEncoder.hpp:
 *      76-145, :324-335): entry j is the squared error of keeping the first
 *      j planes of the codec. fp32 data, num_bitplanes <= 32 (the reference
 *      collectors take float / uint32); exp as passed to the encoder. Each
 *      entry is the exactly rounded sum of the reference's per-element terms.
 *      Synchronizes the stream; h_out[0..nb].                               */
size_t prodm_level_errors_workspace_bytes(uint64_t n, int num_bitplanes);
int prodm_negabinary_level_errors_f32(const float *d_x, uint64_t n, int32_t exp, int num_bitplanes, double *h_out,
                                      void *d_ws, size_t ws_bytes, prodm_stream_t stream);
int prodm_grouped_level_errors_f32(const float *d_x, uint64_t n, int32_t exp, int num_bitplanes, double *h_out,
                                   void *d_ws, size_t ws_bytes, prodm_stream_t stream);

/* ---- device-resident refactor / retrieve pipeline: the GPU stages of
 *      ComposedRefactor::refactor (ComposedRefactor.hpp:78-141: decompose,
 *      interleave, max-abs, encode) and of ComposedReconstructor::reconstruct
 *      (ComposedReconstructor.hpp:187-258: decode, reposition, recompose).
 *      Interleave is fused into the decomposition (coefficients are written
 *      straight into per-level buffers) and reposition into the
 *      recomposition. Lossless compression and file I/O stay on the host.
 *      encoder: PRODM_ENC_NEGABINARY, PRODM_ENC_GROUPED or PRODM_ENC_PERBIT
 *      (PerBitBPEncoder: encoded on the device; prodm_mdr_retrieve decodes
 *      its streams on the host, see prodm_perbit_decode_f32, and has no
 *      graph form). A plane of the last two dims must span less than 2 GiB
 *      and a level fewer than 2^32 elements: plan creation refuses larger
 *      geometries (PRODM_ERR_UNSUPPORTED / PRODM_ERR_ARG) before any device
 *      call. Lines of any length are solved.                                 */
typedef struct prodm_mdr_plan prodm_mdr_plan;
int prodm_mdr_plan_create(prodm_mdr_plan **out, const uint32_t *dims, int ndims, int target_level, int num_bitplanes,
                          int dtype, int encoder);
int prodm_mdr_plan_destroy(prodm_mdr_plan *plan);
/* device bytes held by the plan */
uint64_t prodm_mdr_plan_device_bytes(const prodm_mdr_plan *plan);
int prodm_mdr_refactor(prodm_mdr_plan *plan, const void *d_in, prodm_stream_t stream);
/* level results (synchronizes): max |coef|, exponent, element count      */
int prodm_mdr_level_info(prodm_mdr_plan *plan, int level, double *max_abs, int32_t *exp, uint64_t *elements);
/* the plan encoder's level-error table of level l (fp32 plans; see
 * prodm_negabinary_level_errors_f32); synchronizes                        */
int prodm_mdr_level_errors(prodm_mdr_plan *plan, int level, double *h_out, prodm_stream_t stream);
/* device pointer and byte size of plane p of level l (reference stream
 * bytes without the GroupedBP stream-0 header); synchronizes for GroupedBP
 * and PerBit                                                              */
int prodm_mdr_plane(prodm_mdr_plan *plan, int level, int plane, const void **d_ptr, uint64_t *bytes);
/* GroupedBP: device pointer to the rb bytes of a level                   */
int prodm_mdr_level_rb(prodm_mdr_plan *plan, int level, const uint8_t **d_rb, uint64_t *nblocks);
/* device pointer to the interleaved coefficients of a level              */
int prodm_mdr_level_coefficients(prodm_mdr_plan *plan, int level, void **d_ptr);
/* Retrieve planes [0, planes[l]) of every level from the plan's plane
 * storage and recompose to the full grid (d_out dense).                   */
int prodm_mdr_retrieve(prodm_mdr_plan *plan, const uint8_t *planes_per_level, void *d_out, prodm_stream_t stream);
/* Decomposer of a plan: orthogonal (default, MGARDOrthoganalDecomposer,
 * Decomposer/MGARD.hpp:10-40) or hierarchical (MGARD.hpp:42-70).           */
#define PRODM_DECOMP_ORTHOGONAL 0
#define PRODM_DECOMP_HIERARCHICAL 1
int prodm_mdr_plan_set_decomposer(prodm_mdr_plan *plan, int kind);
/* Graph forms of prodm_mdr_refactor / prodm_mdr_retrieve (no counterpart
 * in ProDM, whose path is host code): the first call for a given input
 * pointer (refactor) or (planes per level, output pointer) (retrieve)
 * captures the whole kernel sequence on `stream` into a HIP graph, later
 * calls replay it with one launch. The stream must not be the default
 * stream. Results are identical to the eager calls.                      */
int prodm_mdr_refactor_graph(prodm_mdr_plan *plan, const void *d_in, prodm_stream_t stream);
int prodm_mdr_retrieve_graph(prodm_mdr_plan *plan, const uint8_t *planes_per_level, void *d_out, prodm_stream_t stream);
/* Reference-format byte stream of one plane, copied to host memory
 * (NegaBinary: ceil(n/B) words, NegaBinaryBPEncoder.hpp:241-249; GroupedBP
 * plane 0 carries the u32 nblocks + u8 rb[nblocks] prefix,
 * GroupedBPEncoder.hpp:64-71). h_dst NULL queries *bytes only. Synchronizes
 * the stream. These are the streams BitplaneEncoder::encode returns.     */
int prodm_mdr_plane_stream(prodm_mdr_plan *plan, int level, int plane, uint8_t *h_dst, uint64_t capacity,
                           uint64_t *bytes, prodm_stream_t stream);
/* Inverse: loads one plane's stream (as written by the encoder) into the
 * plan's plane storage, so that prodm_mdr_retrieve decodes stored planes
 * (the decode side of ComposedReconstructor.hpp:236-253). Synchronizes.   */
int prodm_mdr_load_plane_stream(prodm_mdr_plan *plan, int level, int plane, const uint8_t *h_src, uint64_t bytes,
                                prodm_stream_t stream);
/* Sets a level's max |coefficient| (whose exponent drives decoding) for
 * planes loaded from storage: the level error bound of the metadata
 * (divided by 4 for negabinary, ComposedReconstructor.hpp:202-204).       */
int prodm_mdr_set_level_max(prodm_mdr_plan *plan, int level, double max_abs, prodm_stream_t stream);
/* Per-stage timing with HIP events recorded on the launch stream around
 * each kernel group (no counterpart in ProDM, whose Timer wraps whole
 * calls, RefactorUtils.hpp:145-170). Stages: 0 decompose load (finest
 * level), 1 load (other levels), 2 line solves (finest), 3 solves (other),
 * 4 coarsest copy, 5 encode (finest), 6 encode (other), 7 decode (finest),
 * 8 decode (other), 9 recompose load (finest), 10 (other), 11 recompose
 * solves (finest), 12 (other), 13 interpolation (finest), 14 (other).
 * stage_times synchronizes, returns summed ms and launch-group counts since
 * the previous call, and resets.                                           */
/* ---- retrieval planning (host only) --------------------------------------
 * Size interpreters of include/MDR/SizeInterpreter/GreedyBasedSizeInterpreter.hpp:23-250
 * and BasicSizeInterpreter.hpp:11-112 with the estimators of
 * include/MDR/ErrorEstimator/MaxErrorEstimator.hpp:7-68 and
 * SquaredErrorEstimator.hpp:7-68 (C++ in include/prodm/interpreters.hpp).
 * Level l has counts[l] planes: sizes holds counts[l] plane byte sizes per
 * level back to back, errors counts[l]+1 table entries per level. index[l]
 * (planes already held) is updated; out_sizes[l] = bytes to read next.
 * dtype selects the estimator's arithmetic (ProDM instantiates it with the
 * field type). target_level = L of the S-norm / L2 weights (>= nlevels-1). */
#define PRODM_PLAN_GREEDY 0
#define PRODM_PLAN_SIGN_EXCLUDE_GREEDY 1
#define PRODM_PLAN_NEGABINARY_GREEDY 2
#define PRODM_PLAN_INORDER 3
#define PRODM_PLAN_ROUND_ROBIN 4
#define PRODM_EST_MAX_OB 0
#define PRODM_EST_MAX_HB 1
#define PRODM_EST_SNORM 2
#define PRODM_EST_L2_HB 3
int prodm_interpret_retrieve_size(int planner, int estimator, int dtype, int ndims, int target_level, double s,
                                  int nlevels, const int *counts, const uint32_t *sizes, const double *errors,
                                  double tolerance, uint8_t *index, uint32_t *out_sizes);
/* c of MaxErrorEstimatorOB (MaxErrorEstimator.hpp:15-31), in the estimator's
 * arithmetic; 0 for ndims outside 1..3.                                    */
double prodm_max_estimator_ob_constant(int ndims, int dtype);

/* ---- certified amplification and QoI planning (host only; C++ in
 *      include/prodm/qoi.hpp). out[l], l = 0..target_level: the largest
 *      row 1-norm, over every node, of the level-l block of the recompose
 *      operator (exact, from the 1-D L2 projections): a retrieval whose
 *      level-l coefficient errors are at most t_l is within sum_l out[l] t_l
 *      of the field at every node. It replaces the one constant per
 *      dimension count of MaxErrorEstimatorOB (MaxErrorEstimator.hpp:13-40).
 *      decomposer: PRODM_DECOMP_ORTHOGONAL or PRODM_DECOMP_HIERARCHICAL
 *      (all 1).                                                            */
int prodm_recompose_amplification(const uint32_t *dims, int ndims, int target_level, int decomposer, double *out);
/* Greedy planning of several components' planes on a quantity of interest
 * (QoIGreedyBasedSizeInterpreter): the interpreter arguments of
 * prodm_interpret_retrieve_size over ncomp * nlevels (component, level)
 * entries, component-major; errors = per-coefficient error bounds after
 * each plane count (MaxErrorCollector tables); amplification[nlevels] as
 * prodm_recompose_amplification; rounding[ncomp] (NULL: 0) added to each
 * component's bound; sample[npoints * ncomp] reconstructed values (point
 * major) at which the QoI bound is evaluated (ProDM's composition,
 * include/qoi_utils.hpp:18-41). Planes are added until the sampled bound is
 * below the tolerance; index is updated, out_sizes = bytes to read.       */
#define PRODM_QOI_VELOCITY_MAGNITUDE 0
int prodm_qoi_interpret_retrieve_size(int qoi, int ncomp, int nlevels, const double *amplification,
                                      const double *rounding, const int *counts, const uint32_t *sizes,
                                      const double *errors, const double *sample, uint64_t npoints, double tolerance,
                                      uint8_t *index, uint32_t *out_sizes);

/* ---- batched refactor (no counterpart in ProDM, which refactors one field
 *      per ComposedRefactor::refactor call, ComposedRefactor.hpp:78-141):
 *      many fields of uneven shapes refactored together with one grouped
 *      launch per kernel and level step. A batch owns one plan per field:
 *      prodm_mdr_batch_plan returns it, so prodm_mdr_plane,
 *      prodm_mdr_level_info, prodm_mdr_plane_stream and prodm_mdr_retrieve
 *      read and retrieve a batch field. Each field's planes and level maxima
 *      are bit-identical to prodm_mdr_refactor of that field alone.       */
typedef struct prodm_field_desc {
    const void *d_in;       /* dense row-major device field, dims[0] slowest */
    uint32_t dims[3];       /* first ndims used */
    int ndims;
    int target_level;       /* >= 1 */
} prodm_field_desc;
typedef struct prodm_mdr_batch prodm_mdr_batch;
int prodm_mdr_batch_create(prodm_mdr_batch **out, const prodm_field_desc *fields, int nfields, int num_bitplanes,
                           int dtype, int encoder);
int prodm_mdr_batch_destroy(prodm_mdr_batch *batch);
int prodm_mdr_batch_fields(const prodm_mdr_batch *batch);
uint64_t prodm_mdr_batch_device_bytes(const prodm_mdr_batch *batch);
/* the plan holding field f's results (owned by the batch)                 */
prodm_mdr_plan *prodm_mdr_batch_plan(prodm_mdr_batch *batch, int field);
int prodm_mdr_batch_set_input(prodm_mdr_batch *batch, int field, const void *d_in);
int prodm_mdr_batch_set_decomposer(prodm_mdr_batch *batch, int kind);
/* decompose + interleave + level maxima + encode of every field           */
int prodm_mdr_batched_refactor(prodm_mdr_batch *batch, prodm_stream_t stream);
/* the same sequence captured once into a HIP graph and replayed (stream
 * must not be the default stream); inputs or shapes changed since the
 * capture trigger a new capture                                            */
int prodm_mdr_batched_refactor_graph(prodm_mdr_batch *batch, prodm_stream_t stream);
/* Chooses each grouped launch's workgroup count and tile / block shape by
 * timing `reps` runs of every candidate (launch by launch); with cache_path
 * the choice is read from that text file when its signature (fields,
 * dtype, encoder, device) matches, and written there otherwise. Results
 * never depend on the shapes. Synchronizes.                                */
int prodm_mdr_batch_autotune(prodm_mdr_batch *batch, const char *cache_path, int reps, prodm_stream_t stream);
/* one text line per grouped launch: kind, entries, workgroups, shapes     */
int prodm_mdr_batch_describe(prodm_mdr_batch *batch, char *buf, uint64_t capacity, uint64_t *length);

#define PRODM_NUM_STAGES 15
int prodm_mdr_profile(prodm_mdr_plan *plan, int enable);
int prodm_mdr_stage_times(prodm_mdr_plan *plan, double *ms, uint32_t *counts, int nstages);

#ifdef __cplusplus
}
#endif
#endif

// internal.hpp -- launchers shared between the HIP translation units and the
// C ABI layer. All launchers are asynchronous on `stream` and return a Status.
#pragma once

#include <utility>
#include <vector>

#include "common.hpp"

namespace prodm {

struct PlanePtrs {
    const void *p[64];
};

// ---- per-stage timing with HIP events on the launch stream -------------------
enum Stage : int {
    kStDecLoadTop = 0, kStDecLoadRest, kStDecSolveTop, kStDecSolveRest, kStDecCopy, kStEncTop, kStEncRest,
    kStDecodeTop, kStDecodeRest, kStRecLoadTop, kStRecLoadRest, kStRecSolveTop, kStRecSolveRest, kStRecInterpTop,
    kStRecInterpRest, kNumStages
};
struct StageTimer {
    bool on = false;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<int, size_t>> marks;  // (stage, index of the start event; end = +1)
    int open_stage = -1;
    void begin(int stage);
    void end();
    int collect(double *ms, uint32_t *counts, int n);  // synchronizes, then resets
    ~StageTimer();
};
struct StageScope {
    StageTimer *t;
    StageScope(StageTimer *tm, int stage) : t(tm && tm->on ? tm : nullptr) { if (t) t->begin(stage); }
    ~StageScope() { if (t) t->end(); }
};

// ---- fills (graph-safe replacements for hipMemsetAsync) --------------------
int launch_fill_u32(uint32_t *dst, uint64_t count, uint32_t value, hipStream_t s);
int launch_fill_u64(uint64_t *dst, uint64_t count, uint64_t value, hipStream_t s);

// ---- max-abs (RefactorUtils.hpp:68-76) into an order key (atomicMax) --------
template <class T> using KeyOf = typename std::conditional<sizeof(T) == 4, uint32_t, unsigned long long>::type;
template <class T> int launch_maxabs(const T *x, uint64_t n, KeyOf<T> *key, hipStream_t s);

// ---- NegaBinary (NegaBinaryBPEncoder.hpp:18-62, 132-192) ---------------------
// planes: plane p at planes + p * stride (stride even for u32 words, and
// >= nwords rounded up to even). Exponent from `exp` or, if key != nullptr,
// from frexp(value(*key)) on the device.
template <class T>
int launch_nb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb,
                     typename Traits<T>::Stream *planes, uint64_t stride, hipStream_t s);
template <class T>
int launch_nb_decode(const PlanePtrs &planes, uint32_t n, int32_t exp, const KeyOf<T> *key, int start, int k,
                     T *out, hipStream_t s);

// ---- GroupedBP (GroupedBPEncoder.hpp:18-73, 210-297) --------------------------
template <class T> size_t gb_workspace_bytes(uint32_t n, int nb);
// Writes rb[nblocks], compacted plane words (plane p at planes + p * stride,
// capacity 2 * nblocks words each) and words_per_plane[nb] (device).
template <class T>
int launch_gb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb, int compat, uint8_t *rb,
                     typename Traits<T>::Stream *planes, uint64_t stride, uint32_t *words_per_plane, void *ws,
                     size_t ws_bytes, hipStream_t s);
// streams: k device pointers in reference byte format (stream 0 carries the
// 4 + nblocks byte header when start == 0). rb / signs: per-level state.
template <class T>
int launch_gb_decode(const PlanePtrs &streams, uint32_t n, int32_t exp, const KeyOf<T> *key, int start, int k,
                     const uint8_t *rb, typename Traits<T>::Stream *signs, T *out, void *ws, size_t ws_bytes,
                     uint64_t hdr0_bytes, hipStream_t s);
template <class T> size_t gb_decode_workspace_bytes(uint32_t n, int start, int k);

// ---- PerBit (PerBitBPEncoder.hpp:81-143) -------------------------------------
// Stream p (64-bit words) at streams + p * stride, stride >= pb_stream_words(n);
// sizes[p] (device) = its bytes, whole words.
uint64_t pb_stream_words(uint32_t n);
template <class T> size_t pb_workspace_bytes(uint32_t n, int nb);
template <class T>
int launch_pb_encode(const T *x, uint32_t n, int32_t exp, const KeyOf<T> *key, int nb, int compat,
                     uint64_t *streams, uint64_t stride, uint32_t *sizes, void *ws, size_t ws_bytes, hipStream_t s);

// ---- interleave (DirectInterleaver.hpp:12-63) --------------------------------
template <class T>
int launch_interleave(const T *data, const uint64_t *strides3, const uint32_t *fine3, const uint32_t *coarse3, T *buf,
                      bool reposition, hipStream_t s);

// ---- squared error collector (SquaredErrorCollector.hpp:23-65) ---------------
template <class T> size_t sq_workspace_bytes(uint64_t n, int nb);
template <class T>
int launch_collect_squared(const T *x, uint64_t n, int nb, T max_level_error, void *ws, size_t ws_bytes,
                           hipStream_t s);
// finishes on the host after the stream is synchronized: exact rounding
template <class T> int finish_collect_squared(const void *h_ws_copy, int nb, T max_level_error, double *out);
template <class T> size_t sq_result_bytes(int nb);

// ---- encoder-fused level errors (NegaBinaryBPEncoder.hpp:65-125,231-239,
// GroupedBPEncoder.hpp:76-145,324-335), fp32: enc 0 NegaBinary, 1 GroupedBP.
// Exactly rounded table of nb + 1 doubles on the host; synchronizes.
size_t fused_errors_workspace_bytes(uint64_t n, int nb);
int fused_level_errors(int enc, const float *x, uint64_t n, int32_t exp, int nb, double *out, void *ws,
                       size_t ws_bytes, hipStream_t s);

// ---- MGARD ------------------------------------------------------------------
size_t mgard_workspace_bytes(const LevelGeom &g, size_t elem);
// Coefficient addressing of a decompose/recompose: full array (reordered
// positions with strides) or per-level interleaved buffers.
template <class T> struct CoefSink {
    T *full = nullptr;           // full-array mode
    uint64_t s[3] = {0, 0, 1};   // strides of the full array
    T *level[33] = {nullptr};    // interleaved mode: level buffers (0..L)
};
// hier: hierarchical basis (MGARDHierarchicalDecomposer): the same level
// steps without the L2 correction (no load vector, no line solves).
// Thomas factor tables live in the workspace. tables_ready: they were
// uploaded once by mgard_upload_tables (synchronous); otherwise each call
// uploads them on the stream.
int mgard_upload_tables(const LevelGeom &g, size_t elem, void *ws, size_t ws_bytes);
// Level hook: decompose calls it with l once the launches that finish level
// l's coefficients and max key are on the stream; recompose calls it with l
// before the first launch that reads level l (for cross-stream events).
struct LevelHook {
    int (*fn)(void *ctx, int level);
    void *ctx;
    int operator()(int l) const { return fn ? fn(ctx, l) : kOk; }
};
template <class T>
int mgard_decompose(const T *in, const uint64_t *in_strides, const LevelGeom &g, const CoefSink<T> &sink,
                    KeyOf<T> *level_keys, void *ws, size_t ws_bytes, hipStream_t s, StageTimer *tm = nullptr,
                    bool tables_ready = false, bool hier = false, const LevelHook *hook = nullptr);
// f_levels (optional): per level l the load vector of that level, already
// computed by mgard_recompose_load (so those sweeps can run ahead on another
// stream); the recompose then only solves and interpolates. top_solved: the
// finest level's solves also ran ahead (mgard_solve_ahead), so f_levels[L]
// holds z and the interpolation forms u_c - z itself.
template <class T>
int mgard_recompose(const CoefSink<T> &src, const LevelGeom &g, int tables_ready, T *out, const uint64_t *out_strides,
                    void *ws, size_t ws_bytes, hipStream_t s, StageTimer *tm = nullptr, bool hier = false,
                    const LevelHook *hook = nullptr, double *const *f_levels = nullptr, bool top_solved = false);
// the line solves of level l on its load vector f_l (from
// mgard_recompose_load), z left in f_l; they need no coarse values, so they can
// run before the coarser levels are recomposed
template <class T> int mgard_solve_ahead(const LevelGeom &g, int l, double *f_l, void *ws, size_t ws_bytes, hipStream_t s);
// the recompose load sweep of level l (reads only level l's coefficients)
// into f_l (rows of pitch round16(c2)); bytes of such a buffer
template <class T> int mgard_recompose_load(const CoefSink<T> &src, const LevelGeom &g, int l, double *f_l, hipStream_t s);
size_t mgard_level_f_bytes(const LevelGeom &g, int l);
// Transpose of recompose (adjoint.hip): g_in = dJ/du on the level-L box
// (strides in_strides), out = dJ/dcoefficients in the decompose layout.
size_t adjoint_workspace_bytes(const LevelGeom &g);
template <class T>
int mgard_recompose_adjoint(const T *g_in, const uint64_t *in_strides, const LevelGeom &g, T *out,
                            const uint64_t *out_strides, void *ws, size_t ws_bytes, hipStream_t s);
int mgard_adjoint_coarse(const double *gfine, const LevelGeom &g, int l, double *h_out, void *ws, size_t ws_bytes,
                         hipStream_t s);
// t = S h of level l (rows of `pitch` doubles) in the workspace
const double *mgard_adjoint_solution(const LevelGeom &g, int l, void *ws, uint32_t *pitch);

// ---- grouped launches (batched refactor) -----------------------------------
// A grouped launch runs one kernel over the work of many fields: entry e
// (one field's parameters for that kernel) owns the virtual workgroups
// [first[e], first[e + 1]), and grid() workgroups walk all `total` of them
// with stride grid() (a persistent loop when the grid is capped).
enum GroupKind : int { kGrpLoad = 0, kGrpContig = 1, kGrpStrided = 2, kGrpLevel0 = 3, kGrpNbEncode = 4, kGrpZero = 5 };
// Shape knobs of one grouped launch; none changes any result.
struct GroupShape {
    int a = 0;         // load: coarse planes per tile march (0: per-entry default);
                       // contig solve: lines (waves) per workgroup (0: 4)
    uint32_t cap = 0;  // workgroups of the launch (0: one per virtual workgroup)
};
struct GroupLaunch {
    int kind = 0;
    int elem = 4;       // bytes of the field type
    int variant = 0;    // kernel instance (contiguous solve: largest class; strided: register segment length)
    int nb = 0;         // encode: bitplanes
    uint32_t block = 256, nent = 0, total = 0;
    GroupShape shape;
    std::vector<uint8_t> params;   // nent parameter structs of the kind
    std::vector<uint32_t> first;   // nent + 1 prefix offsets
    const void *d_params = nullptr;
    const uint32_t *d_first = nullptr;
    uint32_t grid() const {
        uint64_t g = total;
        if (shape.cap && shape.cap < g) g = ((uint64_t)shape.cap + 7) / 8 * 8;  // whole rounds over the 8 XCDs
        const uint64_t lim = 0xffffffffull / block;
        return (uint32_t)(g < lim ? g : lim / 8 * 8);
    }
};
// One field of a batched decompose: dense input, geometry, interleaved level
// buffers, L+1 max keys and its own MGARD workspace (factor tables uploaded).
struct BatchFieldRef {
    const void *in = nullptr;
    const LevelGeom *g = nullptr;
    void *level[33] = {nullptr};
    void *keys = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    bool hier = false;
};
// The grouped launches of a batched decompose, in stream order: per level
// step j (level L_f - j of every field with L_f - j >= 1) the load sweep and
// the line solves along dims 2, 1, 0, then the coarsest grids. shapes[i]
// shapes launch i (missing entries: defaults). Every field needs L >= 1.
template <class T>
int mgard_batched_launches(const std::vector<BatchFieldRef> &fields, const std::vector<GroupShape> &shapes,
                           std::vector<GroupLaunch> &out);
int mgard_group_run(const GroupLaunch &L, hipStream_t s);
// One level of a batched NegaBinary encode (planes of level buffer x).
struct EncodeRef {
    const void *x = nullptr;
    uint32_t n = 0;
    void *key = nullptr;
    void *planes = nullptr;
    uint64_t stride = 0;
};
template <class T>
int nb_batched_launch(const std::vector<EncodeRef> &levels, int nb, const GroupShape &shape, GroupLaunch &out);
// Zeroes many small buffers (u32 word counts) in one launch: the level max
// keys of a batch before its decompose.
int zero_batched_launch(const std::vector<std::pair<void *, uint32_t>> &buffers, GroupLaunch &out);
int bitplane_group_run(const GroupLaunch &L, hipStream_t s);

}  // namespace prodm

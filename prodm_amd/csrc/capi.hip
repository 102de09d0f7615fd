// capi.hip -- extern "C" entry points of libprodm_mdr.so (include/prodm_mdr.h).
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "plan.hpp"
#include "prodm_mdr.h"

namespace prodm {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const std::string &last_error() { return g_last_error; }

void StageTimer::begin(int stage) {
    if (used + 2 > pool.size()) {
        const size_t grow = std::max<size_t>(64, pool.size());
        for (size_t i = 0; i < grow; i++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
            pool.push_back(e);
        }
    }
    (void)hipEventRecord(pool[used], stream);
    marks.emplace_back(stage, used);
    used += 2;
    open_stage = stage;
}
void StageTimer::end() {
    if (open_stage < 0 || marks.empty()) return;
    (void)hipEventRecord(pool[marks.back().second + 1], stream);
    open_stage = -1;
}
int StageTimer::collect(double *ms, uint32_t *counts, int n) {
    for (int i = 0; i < n; i++) { ms[i] = 0; counts[i] = 0; }
    for (auto &m : marks) {
        float t = 0;
        if (hipEventSynchronize(pool[m.second + 1]) != hipSuccess) return kErrHip;
        if (hipEventElapsedTime(&t, pool[m.second], pool[m.second + 1]) != hipSuccess) return kErrHip;
        if (m.first < n) { ms[m.first] += t; counts[m.first]++; }
    }
    marks.clear();
    used = 0;
    return kOk;
}
StageTimer::~StageTimer() {
    for (auto e : pool) (void)hipEventDestroy(e);
}

LevelGeom make_geom(const uint32_t *dims, int nd, int L) {
    LevelGeom g;
    g.nd = nd;
    g.L = L;
    if (nd < 1 || nd > 3 || L < 0 || L > 32 || !dims) return g;
    for (int d = 0; d < nd; d++) {
        if (dims[d] == 0) return g;
        g.dims[3 - nd + d] = dims[d];
    }
    bool ok = true;
    for (int d = 0; d < 3; d++) {
        uint32_t n = g.dims[d];
        for (int j = 0; j <= L; j++) {
            g.ld[L - j][d] = n;
            if (n > 1 && j < L && (n < 3 || (n & 1u) == 0)) ok = false;
            n = (n == 1) ? 1 : (n >> 1) + 1;
        }
    }
    uint64_t prev = 0;
    for (int l = 0; l <= L; l++) {
        g.points[l] = (uint64_t)g.ld[l][0] * g.ld[l][1] * g.ld[l][2];
        g.elems[l] = g.points[l] - prev;
        prev = g.points[l];
    }
    g.ok = ok;
    return g;
}

// strides of an nd-dim array padded to 3 dims (innermost stride 1)
static void strides3(const LevelGeom &g, const uint32_t *strides, uint64_t *s) {
    s[2] = 1;
    s[1] = g.dims[2];
    s[0] = (uint64_t)g.dims[1] * g.dims[2];
    if (strides)
        for (int d = 0; d < g.nd - 1; d++) s[3 - g.nd + d] = strides[d];
}

// elem > 0 also checks the plane limit of the level sweeps: they address a
// fine plane (the last two dims) with 31-bit byte offsets
static int check_geom(const LevelGeom &g, size_t elem) {
    if (!g.ok) {
        set_error("unsupported geometry: every dim > 1 must be odd (>= 3) at every level above 0, ndims 1..3");
        return kErrGeometry;
    }
    if (elem && (uint64_t)g.ld[g.L][1] * g.ld[g.L][2] * elem >= 0x7FFFFFFFull) {
        set_error("unsupported geometry: a plane of the last two dims spans 2 GiB or more; the level sweeps address "
                  "a plane with 31-bit offsets");
        return kErrUnsupported;
    }
    return kOk;
}

template <class T>
static int decompose_impl(const T *in, T *out, const uint32_t *dims, int nd, int L, const uint32_t *strides, void *ws,
                          size_t ws_bytes, hipStream_t s, bool hier = false) {
    const LevelGeom g = make_geom(dims, nd, L);
    if (int rc = check_geom(g, sizeof(T))) return rc;
    if (!in || !out || in == out) { set_error("decompose: need distinct device input and output"); return kErrArg; }
    uint64_t st[3];
    strides3(g, strides, st);
    CoefSink<T> sink;
    sink.full = out;
    sink.s[0] = st[0]; sink.s[1] = st[1]; sink.s[2] = 1;
    return mgard_decompose<T>(in, st, g, sink, nullptr, ws, ws_bytes, s, nullptr, false, hier);
}

template <class T>
static int recompose_impl(const T *in, T *out, const uint32_t *dims, int nd, int L, const uint32_t *strides, void *ws,
                          size_t ws_bytes, hipStream_t s, bool hier = false) {
    const LevelGeom g = make_geom(dims, nd, L);
    if (int rc = check_geom(g, sizeof(T))) return rc;
    if (!in || !out || in == out) { set_error("recompose: need distinct device input and output"); return kErrArg; }
    uint64_t st[3];
    strides3(g, strides, st);
    CoefSink<T> src;
    src.full = const_cast<T *>(in);
    src.s[0] = st[0]; src.s[1] = st[1]; src.s[2] = 1;
    return mgard_recompose<T>(src, g, 0, out, st, ws, ws_bytes, s, nullptr, hier);
}

template <class T>
static int adjoint_impl(const T *g_in, T *out, const uint32_t *dims, int nd, int L, const uint32_t *strides, void *ws,
                        size_t ws_bytes, hipStream_t s) {
    const LevelGeom g = make_geom(dims, nd, L);
    if (int rc = check_geom(g, sizeof(T))) return rc;
    if (!g_in || !out || (const void *)g_in == (const void *)out) {
        set_error("recompose_adjoint: need distinct device input and output");
        return kErrArg;
    }
    uint64_t st[3];
    strides3(g, strides, st);
    return mgard_recompose_adjoint<T>(g_in, st, g, out, st, ws, ws_bytes, s);
}

template <class T>
static int interleave_impl(const T *data, const uint32_t *dims, int nd, const uint32_t *fine, const uint32_t *coarse,
                           const uint32_t *strides, T *buf, bool repos, hipStream_t s) {
    if (nd < 1 || nd > 3 || !dims || !fine || !coarse) { set_error("interleave: bad arguments"); return kErrArg; }
    LevelGeom g;
    g.nd = nd;
    for (int d = 0; d < nd; d++) g.dims[3 - nd + d] = dims[d];
    uint64_t st[3];
    strides3(g, strides, st);
    uint32_t f3[3] = {1, 1, 1}, c3[3] = {1, 1, 1};
    for (int d = 0; d < nd; d++) { f3[3 - nd + d] = fine[d]; c3[3 - nd + d] = coarse[d]; }
    return launch_interleave<T>(data, st, f3, c3, buf, repos, s);
}

template <class T>
static int maxabs_impl(const T *x, uint64_t n, T *h_out, hipStream_t s) {
    KeyOf<T> *key = nullptr;
    PRODM_HIP(hipMallocAsync((void **)&key, sizeof(KeyOf<T>), s));
    int rc = (sizeof(T) == 4) ? launch_fill_u32(reinterpret_cast<uint32_t *>(key), 1, 0u, s)
                              : launch_fill_u64(reinterpret_cast<uint64_t *>(key), 1, 0ull, s);
    if (!rc) rc = launch_maxabs<T>(x, n, key, s);
    KeyOf<T> hk = 0;
    if (!rc) {
        hipError_t e = hipMemcpyAsync(&hk, key, sizeof(hk), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = kErrHip; }
    }
    (void)hipFreeAsync(key, s);
    if (!rc) *h_out = key_to_value<T>((uint64_t)hk);
    return rc;
}

// ---- grouped decode state -------------------------------------------------------
template <class T> struct GroupedLevel {
    bool seen = false;
    uint32_t n = 0;
    uint8_t *rb = nullptr;
    typename Traits<T>::Stream *signs = nullptr;
};

}  // namespace prodm

using namespace prodm;

struct prodm_grouped_state {
    int dtype;
    int max_levels;
    std::vector<GroupedLevel<float>> lf;
    std::vector<GroupedLevel<double>> ld;
    void *ws = nullptr;
    size_t ws_bytes = 0;
};

namespace prodm {
template <class T> std::vector<GroupedLevel<T>> &levels_of(prodm_grouped_state *s);
template <> std::vector<GroupedLevel<float>> &levels_of<float>(prodm_grouped_state *s) { return s->lf; }
template <> std::vector<GroupedLevel<double>> &levels_of<double>(prodm_grouped_state *s) { return s->ld; }

template <class T>
static int grouped_decode_impl(prodm_grouped_state *st, int level, const uint8_t *const *streams, uint32_t n,
                               int32_t exp, int start, int k, T *out, hipStream_t s) {
    constexpr int B = Traits<T>::kBlock;
    if (!st || level < 0 || level >= st->max_levels) { set_error("grouped decode: bad state or level"); return kErrArg; }
    if ((sizeof(T) == 4) != (st->dtype == PRODM_F32)) { set_error("grouped decode: state dtype mismatch"); return kErrArg; }
    if (n == 0) return kOk;
    if (k < 0 || start < 0 || start + k > B || (k > 0 && !streams)) { set_error("grouped decode: bad plane range"); return kErrArg; }
    PlanePtrs pp{};
    for (int i = 0; i < k; i++) pp.p[i] = streams[i];
    if (k == 0) return launch_gb_decode<T>(pp, n, exp, nullptr, start, 0, nullptr, nullptr, out, nullptr, 0, 0, s);
    GroupedLevel<T> &L = levels_of<T>(st)[level];
    const uint32_t nblk = (n - 1) / B + 1;
    uint64_t hdr = 0;
    if (!L.seen) {
        // first sight of the level: parse the stream-0 header (GroupedBPEncoder.hpp:223-230)
        if (start != 0) { set_error("grouped decode: first retrieval of a level must start at plane 0"); return kErrState; }
        PRODM_HIP(hipMalloc((void **)&L.rb, nblk));
        PRODM_HIP(hipMalloc((void **)&L.signs, (size_t)nblk * sizeof(typename Traits<T>::Stream)));
        PRODM_HIP(hipMemcpyAsync(L.rb, streams[0] + 4, nblk, hipMemcpyDeviceToDevice, s));
        int rc = (sizeof(T) == 4) ? launch_fill_u32(reinterpret_cast<uint32_t *>(L.signs), nblk, 0u, s)
                                  : launch_fill_u64(reinterpret_cast<uint64_t *>(L.signs), nblk, 0ull, s);
        if (rc) return rc;
        L.seen = true;
        L.n = n;
        hdr = 4 + (uint64_t)nblk;
    } else if (L.n != n) {
        set_error("grouped decode: element count differs from the first retrieval of this level");
        return kErrState;
    }
    const size_t need = gb_decode_workspace_bytes<T>(n, start, k);
    if (st->ws_bytes < need) {
        if (st->ws) { PRODM_HIP(hipStreamSynchronize(s)); PRODM_HIP(hipFree(st->ws)); st->ws = nullptr; }
        PRODM_HIP(hipMalloc(&st->ws, need));
        st->ws_bytes = need;
    }
    return launch_gb_decode<T>(pp, n, exp, nullptr, start, k, L.rb, L.signs, out, st->ws, st->ws_bytes, hdr, s);
}

template <class T> static void collect_max_impl(T bound, int nb, double *out) {
    // MaxErrorCollector.hpp:14-25
    int e = 0;
    std::frexp(bound, &e);
    out[0] = bound;
    double err = std::ldexp(1.0, e - 1);
    for (int i = 1; i <= nb; i++) { out[i] = err; err /= 2; }
}

template <class T>
static int collect_squared_impl(const T *x, uint64_t n, int nb, T maxerr, double *h_out, void *ws, size_t ws_bytes,
                                hipStream_t s) {
    if (nb < 1 || nb > Traits<T>::kBlock || !h_out) { set_error("squared collector: bad arguments"); return kErrArg; }
    if (ws_bytes < sq_workspace_bytes<T>(n, nb)) { set_error("squared collector: workspace too small"); return kErrArg; }
    int rc = launch_collect_squared<T>(x, n, nb, maxerr, ws, ws_bytes, s);
    if (rc) return rc;
    std::vector<uint8_t> host(sq_result_bytes<T>(nb));
    PRODM_HIP(hipMemcpyAsync(host.data(), ws, host.size(), hipMemcpyDeviceToHost, s));
    PRODM_HIP(hipStreamSynchronize(s));
    return finish_collect_squared<T>(host.data(), nb, maxerr, h_out);
}

}  // namespace prodm


namespace prodm {

static int plan_alloc(void **p, size_t bytes, uint64_t &total) {
    if (!bytes) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) { set_error(std::string("plan allocation failed: ") + hipGetErrorString(e)); return kErrNoMem; }
    total += bytes;
    return kOk;
}

static void plan_free(prodm_mdr_plan *p) {
    if (p->g_ref) (void)hipGraphExecDestroy(p->g_ref);
    if (p->g_ret) (void)hipGraphExecDestroy(p->g_ret);
    for (hipEvent_t e : p->lev_ev)
        if (e) (void)hipEventDestroy(e);
    if (p->join_ev) (void)hipEventDestroy(p->join_ev);
    if (p->side) (void)hipStreamDestroy(p->side);
    if (p->side2) (void)hipStreamDestroy(p->side2);
    void *ptrs[] = {p->ws, p->levels, p->keys, p->planes, p->rb, p->words, p->signs, p->gbws, p->flev};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
}

static int plan_side(prodm_mdr_plan *p) {
    if (p->side) return kOk;
    PRODM_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    PRODM_HIP(hipStreamCreateWithFlags(&p->side2, hipStreamNonBlocking));
    for (int l = 0; l <= p->g.L; l++) PRODM_HIP(hipEventCreateWithFlags(&p->lev_ev[l], hipEventDisableTiming));
    PRODM_HIP(hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming));
    return kOk;
}

// per-level load vectors of the retrieve pipeline (allocated on first use,
// before any graph capture)
static int plan_flev(prodm_mdr_plan *p) {
    if (p->flev || p->hier || p->g.L < 1) return kOk;
    uint64_t total = 0;
    for (int l = 1; l <= p->g.L; l++) {
        p->flev_off[l] = total;
        total += (mgard_level_f_bytes(p->g, l) + 255) & ~(uint64_t)255;
    }
    return plan_alloc(&p->flev, total, p->device_bytes);
}

template <class F> static int hook_call(void *ctx, int level) { return (*static_cast<F *>(ctx))(level); }

template <class T> static int plan_refactor(prodm_mdr_plan *p, const T *in, hipStream_t s) {
    using S = typename Traits<T>::Stream;
    using K = KeyOf<T>;
    const LevelGeom &g = p->g;
    K *keys = static_cast<K *>(p->keys);
    int rc = (sizeof(T) == 4) ? launch_fill_u32(reinterpret_cast<uint32_t *>(keys), g.L + 1, 0u, s)
                              : launch_fill_u64(reinterpret_cast<uint64_t *>(keys), g.L + 1, 0ull, s);
    if (rc) return rc;
    CoefSink<T> sink;
    for (int l = 0; l <= g.L; l++) sink.level[l] = static_cast<T *>(p->levels) + p->level_off[l];
    const uint64_t st[3] = {(uint64_t)g.dims[1] * g.dims[2], g.dims[2], 1};
    StageTimer *tm = p->timer.on ? &p->timer : nullptr;
    if (tm) tm->stream = s;
    auto encode_level = [&](int l, hipStream_t es) -> int {
        StageScope sc(tm, l == g.L ? kStEncTop : kStEncRest);
        S *planes = static_cast<S *>(p->planes) + p->plane_off[l];
        if (p->encoder == PRODM_ENC_NEGABINARY)
            return launch_nb_encode<T>(sink.level[l], (uint32_t)g.elems[l], 0, keys + l, p->nb, planes,
                                       p->plane_stride[l], es);
        if (p->encoder == PRODM_ENC_PERBIT)
            return launch_pb_encode<T>(sink.level[l], (uint32_t)g.elems[l], 0, keys + l, p->nb, 0,
                                       reinterpret_cast<uint64_t *>(planes), p->plane_stride[l] * sizeof(T) / 8,
                                       p->words + (size_t)l * p->nb, p->gbws, p->gbws_bytes, es);
        return launch_gb_encode<T>(sink.level[l], (uint32_t)g.elems[l], 0, keys + l, p->nb, 0, p->rb + p->rb_off[l],
                                   planes, p->plane_stride[l], p->words + (size_t)l * p->nb, p->gbws, p->gbws_bytes, es);
    };
    // per-stage timing keeps every launch on one stream; otherwise each
    // level's encode runs on the side stream as soon as its coefficients and
    // max are final, beside the decomposition of the coarser levels (also
    // under graph capture: the fork and join become graph edges)
    if (tm) {
        rc = mgard_decompose<T>(in, st, g, sink, keys, p->ws, p->ws_bytes, s, tm, true, p->hier != 0);
        for (int l = g.L; l >= 0 && !rc; l--) rc = encode_level(l, s);
        return rc;
    }
    if ((rc = plan_side(p))) return rc;
    // the finest levels' encodes start once the decomposition reaches level
    // L - 2, whose short, latency-bound launches leave the GPU room for them
    // (beside the finest levels' own solves they would only compete for HBM)
    const int fork = std::max(0, g.L - 2);
    auto ready = [&](int l) -> int {
        if (l > fork) return kOk;
        PRODM_HIP(hipEventRecord(p->lev_ev[l], s));
        PRODM_HIP(hipStreamWaitEvent(p->side, p->lev_ev[l], 0));
        for (int j = (l == fork ? g.L : l); j >= l; j--)
            if (int r = encode_level(j, p->side)) return r;
        return kOk;
    };
    const LevelHook hook{&hook_call<decltype(ready)>, &ready};
    rc = mgard_decompose<T>(in, st, g, sink, keys, p->ws, p->ws_bytes, s, nullptr, true, p->hier != 0, &hook);
    // join the side stream back, also after an error (nothing stays forked)
    const hipError_t e1 = hipEventRecord(p->join_ev, p->side);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(s, p->join_ev, 0) : e1;
    if (rc) return rc;
    if (e2 != hipSuccess) { set_error(std::string("refactor join: ") + hipGetErrorString(e2)); return kErrHip; }
    return kOk;
}

// PerBit planes [0, k[l]) of every level decoded on the host (the streams
// have no index of where an element's bits begin) into the level buffers.
// Synchronizes the stream.
template <class T> static int perbit_decode_levels(prodm_mdr_plan *p, const uint8_t *k, hipStream_t s) {
    const LevelGeom &g = p->g;
    PRODM_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> sizes((size_t)(g.L + 1) * p->nb);
    PRODM_HIP(hipMemcpy(sizes.data(), p->words, sizes.size() * 4, hipMemcpyDeviceToHost));
    prodm_perbit_state *st = nullptr;
    if (int rc = prodm_perbit_state_create(&st)) return rc;
    int rc = kOk;
    std::vector<T> vals;
    std::vector<std::vector<uint64_t>> words;
    for (int l = 0; l <= g.L && !rc; l++) {
        const uint64_t n = g.elems[l];
        double mx;
        int32_t e;
        if ((rc = prodm_mdr_level_info(p, l, &mx, &e, nullptr))) break;
        words.assign(k[l], {});
        std::vector<const uint64_t *> ptrs(k[l] ? k[l] : 1);
        const char *base = static_cast<const char *>(p->planes) + p->elem * p->plane_off[l];
        for (int i = 0; i < k[l]; i++) {
            const uint32_t bytes = sizes[(size_t)l * p->nb + i];
            words[i].resize(bytes / 8 + 1);
            if (bytes)
                if (hipMemcpy(words[i].data(), base + p->elem * p->plane_stride[l] * i, bytes, hipMemcpyDeviceToHost) !=
                    hipSuccess) { set_error("retrieve: plane copy"); rc = kErrHip; break; }
            ptrs[i] = words[i].data();
        }
        if (rc) break;
        vals.resize(n ? n : 1);
        std::vector<uint32_t> sz(k[l] ? k[l] : 1);
        for (int i = 0; i < k[l]; i++) sz[i] = sizes[(size_t)l * p->nb + i];
        rc = sizeof(T) == 4 ? prodm_perbit_decode_f32(st, l, ptrs.data(), sz.data(), (uint32_t)n, e, 0, k[l],
                                                      reinterpret_cast<float *>(vals.data()))
                            : prodm_perbit_decode_f64(st, l, ptrs.data(), sz.data(), (uint32_t)n, e, 0, k[l],
                                                      reinterpret_cast<double *>(vals.data()));
        if (!rc && n &&
            hipMemcpy(static_cast<T *>(p->levels) + p->level_off[l], vals.data(), n * sizeof(T), hipMemcpyHostToDevice) !=
                hipSuccess) { set_error("retrieve: level upload"); rc = kErrHip; }
    }
    prodm_perbit_state_destroy(st);
    return rc;
}

template <class T> static int plan_retrieve(prodm_mdr_plan *p, const uint8_t *k, T *out, hipStream_t s) {
    using S = typename Traits<T>::Stream;
    using K = KeyOf<T>;
    const LevelGeom &g = p->g;
    K *keys = static_cast<K *>(p->keys);
    CoefSink<T> src;
    for (int l = 0; l <= g.L; l++) src.level[l] = static_cast<T *>(p->levels) + p->level_off[l];
    for (int l = 0; l <= g.L; l++)
        if (k[l] > p->nb) { set_error("retrieve: more planes than encoded"); return kErrArg; }
    StageTimer *tm = p->timer.on ? &p->timer : nullptr;
    if (tm) tm->stream = s;
    if (p->encoder == PRODM_ENC_PERBIT) {
        if (p->capturing) { set_error("retrieve: PerBit decodes on the host and has no graph form"); return kErrUnsupported; }
        const uint64_t st[3] = {(uint64_t)g.dims[1] * g.dims[2], g.dims[2], 1};
        if (int rc = perbit_decode_levels<T>(p, k, s)) return rc;
        return mgard_recompose<T>(src, g, 1, out, st, p->ws, p->ws_bytes, s, tm, p->hier != 0);
    }
    auto decode_level = [&](int l, hipStream_t ds) -> int {
        StageScope sc(tm, l == g.L ? kStDecodeTop : kStDecodeRest);
        S *planes = static_cast<S *>(p->planes) + p->plane_off[l];
        PlanePtrs pp{};
        for (int i = 0; i < k[l]; i++) pp.p[i] = planes + (uint64_t)i * p->plane_stride[l];
        const uint32_t n = (uint32_t)g.elems[l];
        if (p->encoder == PRODM_ENC_NEGABINARY) return launch_nb_decode<T>(pp, n, 0, keys + l, 0, k[l], src.level[l], ds);
        S *signs = static_cast<S *>(p->signs) + p->rb_off[l];
        return launch_gb_decode<T>(pp, n, 0, keys + l, 0, k[l], p->rb + p->rb_off[l], signs, src.level[l], p->gbws,
                                   p->gbws_bytes, 0, ds);
    };
    const uint64_t st[3] = {(uint64_t)g.dims[1] * g.dims[2], g.dims[2], 1};
    int rc = kOk;
    if (tm) {
        for (int l = g.L; l >= 0 && !rc; l--) rc = decode_level(l, s);
        if (rc) return rc;
        return mgard_recompose<T>(src, g, 1, out, st, p->ws, p->ws_bytes, s, tm, p->hier != 0);
    }
    // on the side stream, coarsest first: each level's decode and, for the
    // orthogonal decomposer, its recompose load sweep into a per-level load
    // vector (it reads only that level's coefficients); the main stream's
    // chain of solves and interpolations waits for level l just before it.
    // The finest level's decode, load sweep and line solves, the bulk of the
    // bytes, start at once on a second side stream instead of after the
    // coarser levels' short, latency-bound launches (negabinary: its decode
    // needs no workspace, so the two side streams share nothing but the
    // read-only factor tables).
    if ((rc = plan_side(p))) return rc;
    const bool ahead = p->hier == 0 && g.L >= 1;
    double *fl[33] = {nullptr};
    if (ahead) {
        if ((rc = plan_flev(p))) return rc;
        for (int l = 1; l <= g.L; l++) fl[l] = reinterpret_cast<double *>(static_cast<char *>(p->flev) + p->flev_off[l]);
    }
    PRODM_HIP(hipEventRecord(p->join_ev, s));
    PRODM_HIP(hipStreamWaitEvent(p->side, p->join_ev, 0));
    const bool split = p->encoder == PRODM_ENC_NEGABINARY && g.L >= 1;
    if (split) PRODM_HIP(hipStreamWaitEvent(p->side2, p->join_ev, 0));
    for (int i = 0; i <= g.L && !rc; i++) {
        const int l = split ? (i == 0 ? g.L : i - 1) : i;
        hipStream_t ss = (split && l == g.L) ? p->side2 : p->side;
        rc = decode_level(l, ss);
        if (!rc && ahead && l >= 1) rc = mgard_recompose_load<T>(src, g, l, fl[l], ss);
        // the finest level's solves need only its load vector: they run here,
        // beside the recomposition of the coarser levels, and its
        // interpolation forms u_c - z
        if (!rc && ahead && split && l == g.L) rc = mgard_solve_ahead<T>(g, l, fl[l], p->ws, p->ws_bytes, ss);
        if (!rc && hipEventRecord(p->lev_ev[l], ss) != hipSuccess) { set_error("retrieve: event record"); rc = kErrHip; }
    }
    if (rc) {
        // join whatever was forked before reporting the error
        if (hipEventRecord(p->join_ev, p->side) == hipSuccess) (void)hipStreamWaitEvent(s, p->join_ev, 0);
        if (split && hipEventRecord(p->join_ev, p->side2) == hipSuccess) (void)hipStreamWaitEvent(s, p->join_ev, 0);
        return rc;
    }
    auto need = [&](int lv) -> int {
        PRODM_HIP(hipStreamWaitEvent(s, p->lev_ev[lv], 0));
        return kOk;
    };
    const LevelHook hook{&hook_call<decltype(need)>, &need};
    rc = mgard_recompose<T>(src, g, 1, out, st, p->ws, p->ws_bytes, s, nullptr, p->hier != 0, &hook,
                            ahead ? fl : nullptr, ahead && split);
    // every decode is joined (the last level is waited for before its recomposition)
    if (rc) (void)hipStreamWaitEvent(s, p->lev_ev[g.L], 0);
    return rc;
}

}  // namespace prodm

extern "C" {

int prodm_version(void) { return 10000; }
const char *prodm_last_error(void) { return prodm::last_error().c_str(); }

int prodm_level_geometry(const uint32_t *dims, int ndims, int target_level, uint32_t *out_dims, uint64_t *out_elements) {
    const LevelGeom g = make_geom(dims, ndims, target_level);
    if (g.nd == 0) return 0;
    for (int l = 0; l <= target_level; l++) {
        if (out_dims)
            for (int d = 0; d < ndims; d++) out_dims[l * ndims + d] = g.ld[l][3 - ndims + d];
        if (out_elements) out_elements[l] = g.elems[l];
    }
    return g.ok ? 1 : 0;
}

size_t prodm_recompose_adjoint_workspace_bytes(const uint32_t *dims, int ndims, int target_level) {
    const LevelGeom g = make_geom(dims, ndims, target_level);
    return g.ok ? adjoint_workspace_bytes(g) : 0;
}
int prodm_recompose_adjoint_f32(const float *d_g, float *d_out, const uint32_t *dims, int ndims, int target_level,
                                const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return adjoint_impl<float>(d_g, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream);
}
int prodm_recompose_adjoint_f64(const double *d_g, double *d_out, const uint32_t *dims, int ndims, int target_level,
                                const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return adjoint_impl<double>(d_g, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream);
}

size_t prodm_mgard_workspace_bytes(const uint32_t *dims, int ndims, int target_level, int dtype) {
    const LevelGeom g = make_geom(dims, ndims, target_level);
    if (!g.ok) return 0;
    return mgard_workspace_bytes(g, dtype == PRODM_F64 ? 8 : 4);
}

int prodm_decompose_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return decompose_impl<float>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream);
}
int prodm_decompose_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return decompose_impl<double>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream);
}
int prodm_recompose_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return recompose_impl<float>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream);
}
int prodm_recompose_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims, int target_level,
                        const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return recompose_impl<double>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream);
}

int prodm_decompose_hierarchical_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                                     const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return decompose_impl<float>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream, true);
}
int prodm_decompose_hierarchical_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims,
                                     int target_level, const uint32_t *strides, void *d_ws, size_t ws_bytes,
                                     prodm_stream_t stream) {
    return decompose_impl<double>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream, true);
}
int prodm_recompose_hierarchical_f32(const float *d_in, float *d_out, const uint32_t *dims, int ndims, int target_level,
                                     const uint32_t *strides, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return recompose_impl<float>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream, true);
}
int prodm_recompose_hierarchical_f64(const double *d_in, double *d_out, const uint32_t *dims, int ndims,
                                     int target_level, const uint32_t *strides, void *d_ws, size_t ws_bytes,
                                     prodm_stream_t stream) {
    return recompose_impl<double>(d_in, d_out, dims, ndims, target_level, strides, d_ws, ws_bytes, (hipStream_t)stream, true);
}
int prodm_mdr_plan_set_decomposer(prodm_mdr_plan *p, int kind) {
    if (!p || (kind != PRODM_DECOMP_ORTHOGONAL && kind != PRODM_DECOMP_HIERARCHICAL)) {
        set_error("plan_set_decomposer: bad arguments");
        return kErrArg;
    }
    p->hier = kind == PRODM_DECOMP_HIERARCHICAL;
    if (p->g_ref) { (void)hipGraphExecDestroy(p->g_ref); p->g_ref = nullptr; }
    if (p->g_ret) { (void)hipGraphExecDestroy(p->g_ret); p->g_ret = nullptr; }
    return kOk;
}

int prodm_interleave_f32(const float *d_data, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, float *d_buffer, prodm_stream_t stream) {
    return interleave_impl<float>(d_data, dims, ndims, dims_fine, dims_coarse, strides, d_buffer, false, (hipStream_t)stream);
}
int prodm_interleave_f64(const double *d_data, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, double *d_buffer, prodm_stream_t stream) {
    return interleave_impl<double>(d_data, dims, ndims, dims_fine, dims_coarse, strides, d_buffer, false, (hipStream_t)stream);
}
int prodm_reposition_f32(const float *d_buffer, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, float *d_data, prodm_stream_t stream) {
    return interleave_impl<float>(d_data, dims, ndims, dims_fine, dims_coarse, strides, const_cast<float *>(d_buffer), true,
                                  (hipStream_t)stream);
}
int prodm_reposition_f64(const double *d_buffer, const uint32_t *dims, int ndims, const uint32_t *dims_fine,
                         const uint32_t *dims_coarse, const uint32_t *strides, double *d_data, prodm_stream_t stream) {
    return interleave_impl<double>(d_data, dims, ndims, dims_fine, dims_coarse, strides, const_cast<double *>(d_buffer),
                                   true, (hipStream_t)stream);
}

int prodm_level_maxabs_f32(const float *d_x, uint64_t n, float *h_out, prodm_stream_t stream) {
    return maxabs_impl<float>(d_x, n, h_out, (hipStream_t)stream);
}
int prodm_level_maxabs_f64(const double *d_x, uint64_t n, double *h_out, prodm_stream_t stream) {
    return maxabs_impl<double>(d_x, n, h_out, (hipStream_t)stream);
}

int prodm_negabinary_encode_f32(const float *d_x, uint32_t n, int32_t exp, int nb, uint32_t *d_planes,
                                uint64_t stride, prodm_stream_t stream) {
    return launch_nb_encode<float>(d_x, n, exp, nullptr, nb, d_planes, stride, (hipStream_t)stream);
}
int prodm_negabinary_encode_f64(const double *d_x, uint32_t n, int32_t exp, int nb, uint64_t *d_planes,
                                uint64_t stride, prodm_stream_t stream) {
    return launch_nb_encode<double>(d_x, n, exp, nullptr, nb, d_planes, stride, (hipStream_t)stream);
}
int prodm_negabinary_decode_f32(const uint32_t *const *h_plane_ptrs, uint32_t n, int32_t exp, int start, int k,
                                float *d_out, prodm_stream_t stream) {
    if (k < 0 || k > 32 || (k > 0 && !h_plane_ptrs)) { set_error("negabinary decode: bad planes"); return kErrArg; }
    PlanePtrs pp{};
    for (int i = 0; i < k; i++) pp.p[i] = h_plane_ptrs[i];
    return launch_nb_decode<float>(pp, n, exp, nullptr, start, k, d_out, (hipStream_t)stream);
}
int prodm_negabinary_decode_f64(const uint64_t *const *h_plane_ptrs, uint32_t n, int32_t exp, int start, int k,
                                double *d_out, prodm_stream_t stream) {
    if (k < 0 || k > 64 || (k > 0 && !h_plane_ptrs)) { set_error("negabinary decode: bad planes"); return kErrArg; }
    PlanePtrs pp{};
    for (int i = 0; i < k; i++) pp.p[i] = h_plane_ptrs[i];
    return launch_nb_decode<double>(pp, n, exp, nullptr, start, k, d_out, (hipStream_t)stream);
}

size_t prodm_grouped_workspace_bytes(uint32_t n, int nb, int dtype) {
    return dtype == PRODM_F64 ? gb_workspace_bytes<double>(n, nb) : gb_workspace_bytes<float>(n, nb);
}
int prodm_grouped_encode_f32(const float *d_x, uint32_t n, int32_t exp, int nb, int compat, uint8_t *d_rb,
                             uint32_t *d_planes, uint64_t stride, uint32_t *d_words, void *d_ws, size_t ws_bytes,
                             prodm_stream_t stream) {
    return launch_gb_encode<float>(d_x, n, exp, nullptr, nb, compat, d_rb, d_planes, stride, d_words, d_ws, ws_bytes,
                                   (hipStream_t)stream);
}
int prodm_grouped_encode_f64(const double *d_x, uint32_t n, int32_t exp, int nb, int compat, uint8_t *d_rb,
                             uint64_t *d_planes, uint64_t stride, uint32_t *d_words, void *d_ws, size_t ws_bytes,
                             prodm_stream_t stream) {
    return launch_gb_encode<double>(d_x, n, exp, nullptr, nb, compat, d_rb, d_planes, stride, d_words, d_ws, ws_bytes,
                                    (hipStream_t)stream);
}
int prodm_grouped_state_create(prodm_grouped_state **out, int max_levels, int dtype) {
    if (!out || max_levels < 1 || max_levels > 64) { set_error("grouped state: bad arguments"); return kErrArg; }
    prodm_grouped_state *s = new prodm_grouped_state();
    s->dtype = dtype;
    s->max_levels = max_levels;
    s->lf.resize(max_levels);
    s->ld.resize(max_levels);
    *out = s;
    return kOk;
}
int prodm_grouped_state_destroy(prodm_grouped_state *s) {
    if (!s) return kOk;
    (void)hipDeviceSynchronize();
    for (auto &l : s->lf) { if (l.rb) (void)hipFree(l.rb); if (l.signs) (void)hipFree(l.signs); }
    for (auto &l : s->ld) { if (l.rb) (void)hipFree(l.rb); if (l.signs) (void)hipFree(l.signs); }
    if (s->ws) (void)hipFree(s->ws);
    delete s;
    return kOk;
}
int prodm_grouped_decode_f32(prodm_grouped_state *state, int level, const uint8_t *const *h_stream_ptrs, uint32_t n,
                             int32_t exp, int start, int k, float *d_out, prodm_stream_t stream) {
    return grouped_decode_impl<float>(state, level, h_stream_ptrs, n, exp, start, k, d_out, (hipStream_t)stream);
}
int prodm_grouped_decode_f64(prodm_grouped_state *state, int level, const uint8_t *const *h_stream_ptrs, uint32_t n,
                             int32_t exp, int start, int k, double *d_out, prodm_stream_t stream) {
    return grouped_decode_impl<double>(state, level, h_stream_ptrs, n, exp, start, k, d_out, (hipStream_t)stream);
}

uint64_t prodm_perbit_stream_words(uint32_t n) { return pb_stream_words(n); }
size_t prodm_perbit_workspace_bytes(uint32_t n, int nb, int dtype) {
    return dtype == PRODM_F64 ? pb_workspace_bytes<double>(n, nb) : pb_workspace_bytes<float>(n, nb);
}
int prodm_perbit_encode_f32(const float *d_x, uint32_t n, int32_t exp, int nb, int compat, uint64_t *d_streams,
                            uint64_t stride, uint32_t *d_sizes, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return launch_pb_encode<float>(d_x, n, exp, nullptr, nb, compat, d_streams, stride, d_sizes, d_ws, ws_bytes,
                                   (hipStream_t)stream);
}
int prodm_perbit_encode_f64(const double *d_x, uint32_t n, int32_t exp, int nb, int compat, uint64_t *d_streams,
                            uint64_t stride, uint32_t *d_sizes, void *d_ws, size_t ws_bytes, prodm_stream_t stream) {
    return launch_pb_encode<double>(d_x, n, exp, nullptr, nb, compat, d_streams, stride, d_sizes, d_ws, ws_bytes,
                                    (hipStream_t)stream);
}

void prodm_collect_max_f32(float bound, int nb, double *h_out) { collect_max_impl<float>(bound, nb, h_out); }
void prodm_collect_max_f64(double bound, int nb, double *h_out) { collect_max_impl<double>(bound, nb, h_out); }

size_t prodm_collect_squared_workspace_bytes(uint64_t n, int nb, int dtype) {
    return dtype == PRODM_F64 ? sq_workspace_bytes<double>(n, nb) : sq_workspace_bytes<float>(n, nb);
}
int prodm_collect_squared_f32(const float *d_x, uint64_t n, int nb, float maxerr, double *h_out, void *d_ws,
                              size_t ws_bytes, prodm_stream_t stream) {
    return collect_squared_impl<float>(d_x, n, nb, maxerr, h_out, d_ws, ws_bytes, (hipStream_t)stream);
}
int prodm_collect_squared_f64(const double *d_x, uint64_t n, int nb, double maxerr, double *h_out, void *d_ws,
                              size_t ws_bytes, prodm_stream_t stream) {
    return collect_squared_impl<double>(d_x, n, nb, maxerr, h_out, d_ws, ws_bytes, (hipStream_t)stream);
}

size_t prodm_level_errors_workspace_bytes(uint64_t n, int nb) { return fused_errors_workspace_bytes(n, nb); }
int prodm_negabinary_level_errors_f32(const float *d_x, uint64_t n, int32_t exp, int nb, double *h_out, void *d_ws,
                                      size_t ws_bytes, prodm_stream_t stream) {
    return fused_level_errors(0, d_x, n, exp, nb, h_out, d_ws, ws_bytes, (hipStream_t)stream);
}
int prodm_grouped_level_errors_f32(const float *d_x, uint64_t n, int32_t exp, int nb, double *h_out, void *d_ws,
                                   size_t ws_bytes, prodm_stream_t stream) {
    return fused_level_errors(1, d_x, n, exp, nb, h_out, d_ws, ws_bytes, (hipStream_t)stream);
}

// ---- plan ---------------------------------------------------------------------------
int prodm_mdr_plan_create(prodm_mdr_plan **out, const uint32_t *dims, int ndims, int target_level, int nb, int dtype,
                          int encoder) {
    if (!out) { set_error("plan: null output"); return kErrArg; }
    *out = nullptr;
    if (dtype != PRODM_F32 && dtype != PRODM_F64) { set_error("plan: bad dtype"); return kErrArg; }
    if (encoder != PRODM_ENC_NEGABINARY && encoder != PRODM_ENC_GROUPED && encoder != PRODM_ENC_PERBIT) {
        set_error("plan: bad encoder");
        return kErrArg;
    }
    const int B = dtype == PRODM_F64 ? 64 : 32;
    if (nb < 1 || nb > B) { set_error("plan: num_bitplanes out of range"); return kErrArg; }
    if (encoder == PRODM_ENC_NEGABINARY && (nb & 1)) {
        set_error("plan: negabinary needs an even number of bitplanes");  // test_mdr_refactor.cpp:40-43
        return kErrArg;
    }
    prodm_mdr_plan *p = new prodm_mdr_plan();
    p->g = make_geom(dims, ndims, target_level);
    if (int rc = check_geom(p->g, dtype == PRODM_F64 ? 8 : 4)) { delete p; return rc; }
    for (int l = 0; l <= target_level; l++)
        if (p->g.elems[l] > 0xffffffffull) { delete p; set_error("plan: level larger than 2^32 elements"); return kErrArg; }
    p->dtype = dtype;
    p->encoder = encoder;
    p->nb = nb;
    p->elem = dtype == PRODM_F64 ? 8 : 4;
    const LevelGeom &g = p->g;
    const size_t sw = p->elem;  // stream word bytes equal the element bytes
    uint64_t lo = 0, po = 0, ro = 0;
    for (int l = 0; l <= g.L; l++) {
        p->level_off[l] = lo;
        lo += (g.elems[l] + 63) & ~63ull;
        const uint64_t nblk = (g.elems[l] + B - 1) / B;
        uint64_t stride = encoder == PRODM_ENC_NEGABINARY ? nblk : 2 * nblk;
        if (encoder == PRODM_ENC_PERBIT) stride = pb_stream_words((uint32_t)g.elems[l]) * 8 / sw;  // 64-bit words
        stride = (stride + 63) & ~63ull;  // 256-byte aligned planes
        p->plane_stride[l] = stride;
        p->plane_off[l] = po;
        po += stride * nb;
        p->rb_off[l] = ro;
        ro += (nblk + 255) & ~255ull;
    }
    int rc = kOk;
    p->ws_bytes = mgard_workspace_bytes(g, p->elem);
    if (!rc) rc = plan_alloc(&p->ws, p->ws_bytes, p->device_bytes);
    if (!rc) rc = mgard_upload_tables(g, p->elem, p->ws, p->ws_bytes);
    if (!rc) rc = plan_alloc(&p->levels, lo * p->elem, p->device_bytes);
    if (!rc) rc = plan_alloc(&p->keys, 8 * (g.L + 1), p->device_bytes);
    if (!rc) rc = plan_alloc(&p->planes, po * sw, p->device_bytes);
    if (!rc && encoder == PRODM_ENC_PERBIT) {
        rc = plan_alloc((void **)&p->words, 4ull * (g.L + 1) * nb, p->device_bytes);
        size_t need = 0;
        for (int l = 0; l <= g.L; l++) {
            const uint32_t n = (uint32_t)g.elems[l];
            need = std::max(need, dtype == PRODM_F64 ? pb_workspace_bytes<double>(n, nb) : pb_workspace_bytes<float>(n, nb));
        }
        p->gbws_bytes = need;
        if (!rc) rc = plan_alloc(&p->gbws, need, p->device_bytes);
    }
    if (!rc && encoder == PRODM_ENC_GROUPED) {
        rc = plan_alloc((void **)&p->rb, ro, p->device_bytes);
        if (!rc) rc = plan_alloc((void **)&p->words, 4ull * (g.L + 1) * nb, p->device_bytes);
        if (!rc) rc = plan_alloc(&p->signs, ro * sw, p->device_bytes);
        size_t need = 0;
        for (int l = 0; l <= g.L; l++) {
            const uint32_t n = (uint32_t)g.elems[l];
            size_t a = dtype == PRODM_F64 ? gb_workspace_bytes<double>(n, nb) : gb_workspace_bytes<float>(n, nb);
            size_t b = dtype == PRODM_F64 ? gb_decode_workspace_bytes<double>(n, 0, nb) : gb_decode_workspace_bytes<float>(n, 0, nb);
            need = std::max(need, std::max(a, b));
        }
        p->gbws_bytes = need;
        if (!rc) rc = plan_alloc(&p->gbws, need, p->device_bytes);
    }
    if (rc) { plan_free(p); delete p; return rc; }
    *out = p;
    return kOk;
}

int prodm_mdr_plan_destroy(prodm_mdr_plan *p) {
    if (!p) return kOk;
    (void)hipDeviceSynchronize();
    plan_free(p);
    delete p;
    return kOk;
}

uint64_t prodm_mdr_plan_device_bytes(const prodm_mdr_plan *p) { return p ? p->device_bytes : 0; }

int prodm_mdr_refactor(prodm_mdr_plan *p, const void *d_in, prodm_stream_t stream) {
    if (!p || !d_in) { set_error("refactor: null argument"); return kErrArg; }
    if (p->dtype == PRODM_F32) return plan_refactor<float>(p, static_cast<const float *>(d_in), (hipStream_t)stream);
    return plan_refactor<double>(p, static_cast<const double *>(d_in), (hipStream_t)stream);
}

int prodm_mdr_level_info(prodm_mdr_plan *p, int level, double *max_abs, int32_t *exp, uint64_t *elements) {
    if (!p || level < 0 || level > p->g.L) { set_error("level_info: bad level"); return kErrArg; }
    uint64_t key = 0;
    PRODM_HIP(hipDeviceSynchronize());
    PRODM_HIP(hipMemcpy(&key, static_cast<char *>(p->keys) + p->elem * level, p->elem, hipMemcpyDeviceToHost));
    double v;
    int e = 0;
    if (p->dtype == PRODM_F32) { float f = key_to_value<float>(key); v = f; std::frexp(f, &e); }
    else { double d = key_to_value<double>(key); v = d; std::frexp(d, &e); }
    if (max_abs) *max_abs = v;
    if (exp) *exp = e;
    if (elements) *elements = p->g.elems[level];
    return kOk;
}

int prodm_mdr_level_errors(prodm_mdr_plan *p, int level, double *h_out, prodm_stream_t stream) {
    if (!p || level < 0 || level > p->g.L || !h_out) { set_error("level_errors: bad level"); return kErrArg; }
    if (p->dtype != PRODM_F32) { set_error("level_errors: the encoders' level errors are defined for fp32 data"); return kErrUnsupported; }
    if (p->nb > 32) { set_error("level_errors: at most 32 bitplanes"); return kErrUnsupported; }
    double mx;
    int32_t e;
    uint64_t n;
    if (int rc = prodm_mdr_level_info(p, level, &mx, &e, &n)) return rc;
    const size_t need = fused_errors_workspace_bytes(n, p->nb);
    void *ws = nullptr;
    PRODM_HIP(hipMalloc(&ws, need));
    const float *x = static_cast<const float *>(p->levels) + p->level_off[level];
    const int rc = fused_level_errors(p->encoder == PRODM_ENC_NEGABINARY ? 0 : 1, x, n, e, p->nb, h_out, ws, need,
                                      (hipStream_t)stream);
    (void)hipStreamSynchronize((hipStream_t)stream);
    (void)hipFree(ws);
    return rc;
}

int prodm_mdr_plane(prodm_mdr_plan *p, int level, int plane, const void **d_ptr, uint64_t *bytes) {
    if (!p || level < 0 || level > p->g.L || plane < 0 || plane >= p->nb) { set_error("plane: bad index"); return kErrArg; }
    const char *base = static_cast<const char *>(p->planes) + p->elem * (p->plane_off[level] + (uint64_t)plane * p->plane_stride[level]);
    if (d_ptr) *d_ptr = base;
    if (bytes) {
        if (p->encoder == PRODM_ENC_NEGABINARY) {
            const int B = p->dtype == PRODM_F64 ? 64 : 32;
            *bytes = (p->g.elems[level] + B - 1) / B * p->elem;
        } else {
            uint32_t w = 0;
            PRODM_HIP(hipDeviceSynchronize());
            PRODM_HIP(hipMemcpy(&w, p->words + (size_t)level * p->nb + plane, 4, hipMemcpyDeviceToHost));
            *bytes = p->encoder == PRODM_ENC_PERBIT ? (uint64_t)w : (uint64_t)w * p->elem;  // PerBit keeps bytes
        }
    }
    return kOk;
}

int prodm_mdr_level_rb(prodm_mdr_plan *p, int level, const uint8_t **d_rb, uint64_t *nblocks) {
    if (!p || level < 0 || level > p->g.L || p->encoder != PRODM_ENC_GROUPED) { set_error("level_rb: bad level or encoder"); return kErrArg; }
    const int B = p->dtype == PRODM_F64 ? 64 : 32;
    if (d_rb) *d_rb = p->rb + p->rb_off[level];
    if (nblocks) *nblocks = (p->g.elems[level] + B - 1) / B;
    return kOk;
}

int prodm_mdr_level_coefficients(prodm_mdr_plan *p, int level, void **d_ptr) {
    if (!p || level < 0 || level > p->g.L) { set_error("level_coefficients: bad level"); return kErrArg; }
    *d_ptr = static_cast<char *>(p->levels) + p->elem * p->level_off[level];
    return kOk;
}

int prodm_mdr_profile(prodm_mdr_plan *p, int enable) {
    if (!p) { set_error("profile: null plan"); return kErrArg; }
    double ms[kNumStages];
    uint32_t cnt[kNumStages];
    int rc = p->timer.collect(ms, cnt, kNumStages);  // drop pending marks
    p->timer.on = enable != 0;
    return rc;
}

int prodm_mdr_stage_times(prodm_mdr_plan *p, double *ms, uint32_t *counts, int nstages) {
    if (!p || !ms || !counts || nstages < 1) { set_error("stage_times: bad arguments"); return kErrArg; }
    std::vector<double> m(kNumStages);
    std::vector<uint32_t> c(kNumStages);
    int rc = p->timer.collect(m.data(), c.data(), kNumStages);
    for (int i = 0; i < nstages; i++) { ms[i] = i < kNumStages ? m[i] : 0; counts[i] = i < kNumStages ? c[i] : 0; }
    return rc;
}

int prodm_mdr_plane_stream(prodm_mdr_plan *p, int level, int plane, uint8_t *h_dst, uint64_t capacity,
                           uint64_t *bytes, prodm_stream_t stream) {
    if (!p || level < 0 || level > p->g.L || plane < 0 || plane >= p->nb || !bytes) {
        set_error("plane_stream: bad arguments");
        return kErrArg;
    }
    hipStream_t s = (hipStream_t)stream;
    PRODM_HIP(hipStreamSynchronize(s));
    const void *src = nullptr;
    uint64_t body = 0;
    if (int rc = prodm_mdr_plane(p, level, plane, &src, &body)) return rc;
    const int B = p->dtype == PRODM_F64 ? 64 : 32;
    const uint64_t nblk = (p->g.elems[level] + B - 1) / B;
    const uint64_t head = (p->encoder == PRODM_ENC_GROUPED && plane == 0) ? 4 + nblk : 0;
    *bytes = head + body;
    if (!h_dst) return kOk;
    if (capacity < *bytes) { set_error("plane_stream: destination too small"); return kErrArg; }
    if (head) {
        const uint32_t nb32 = (uint32_t)nblk;
        std::memcpy(h_dst, &nb32, 4);
        PRODM_HIP(hipMemcpyAsync(h_dst + 4, p->rb + p->rb_off[level], nblk, hipMemcpyDeviceToHost, s));
    }
    if (body) PRODM_HIP(hipMemcpyAsync(h_dst + head, src, body, hipMemcpyDeviceToHost, s));
    PRODM_HIP(hipStreamSynchronize(s));
    return kOk;
}

int prodm_mdr_load_plane_stream(prodm_mdr_plan *p, int level, int plane, const uint8_t *h_src, uint64_t bytes,
                                prodm_stream_t stream) {
    if (!p || level < 0 || level > p->g.L || plane < 0 || plane >= p->nb || (!h_src && bytes)) {
        set_error("load_plane_stream: bad arguments");
        return kErrArg;
    }
    hipStream_t s = (hipStream_t)stream;
    const int B = p->dtype == PRODM_F64 ? 64 : 32;
    const uint64_t nblk = (p->g.elems[level] + B - 1) / B;
    char *slot = static_cast<char *>(p->planes) + p->elem * (p->plane_off[level] + (uint64_t)plane * p->plane_stride[level]);
    const uint64_t cap = p->elem * p->plane_stride[level];
    uint64_t head = 0;
    if (p->encoder == PRODM_ENC_NEGABINARY) {
        if (bytes != nblk * p->elem) { set_error("load_plane_stream: negabinary plane size mismatch"); return kErrArg; }
    } else if (p->encoder == PRODM_ENC_PERBIT) {
        if (bytes % 8) { set_error("load_plane_stream: PerBit streams are whole 64-bit words"); return kErrArg; }
        if (bytes > cap) { set_error("load_plane_stream: plane larger than its storage"); return kErrArg; }
        const uint32_t b32 = (uint32_t)bytes;
        PRODM_HIP(hipMemcpyAsync(p->words + (size_t)level * p->nb + plane, &b32, 4, hipMemcpyHostToDevice, s));
    } else {
        if (plane == 0) {
            uint32_t hdr = 0;
            if (bytes < 4) { set_error("load_plane_stream: truncated header"); return kErrArg; }
            std::memcpy(&hdr, h_src, 4);
            if (hdr != nblk || bytes < 4 + nblk) { set_error("load_plane_stream: header does not match the level"); return kErrArg; }
            head = 4 + nblk;
            PRODM_HIP(hipMemcpyAsync(p->rb + p->rb_off[level], h_src + 4, nblk, hipMemcpyHostToDevice, s));
        }
        if ((bytes - head) % p->elem) { set_error("load_plane_stream: not a whole number of words"); return kErrArg; }
    }
    if (bytes - head > cap) { set_error("load_plane_stream: plane larger than its storage"); return kErrArg; }
    if (bytes > head) PRODM_HIP(hipMemcpyAsync(slot, h_src + head, bytes - head, hipMemcpyHostToDevice, s));
    PRODM_HIP(hipStreamSynchronize(s));
    return kOk;
}

int prodm_mdr_set_level_max(prodm_mdr_plan *p, int level, double max_abs, prodm_stream_t stream) {
    if (!p || level < 0 || level > p->g.L || !(max_abs >= 0)) { set_error("set_level_max: bad arguments"); return kErrArg; }
    hipStream_t s = (hipStream_t)stream;
    if (p->dtype == PRODM_F32) {
        const float f = (float)max_abs;
        uint32_t k;
        std::memcpy(&k, &f, 4);
        return launch_fill_u32(reinterpret_cast<uint32_t *>(p->keys) + level, 1, k, s);
    }
    uint64_t k;
    std::memcpy(&k, &max_abs, 8);
    return launch_fill_u64(reinterpret_cast<uint64_t *>(p->keys) + level, 1, k, s);
}

// Capture `body` on stream s into a graph executable (kernel nodes only).

int prodm_mdr_refactor_graph(prodm_mdr_plan *p, const void *d_in, prodm_stream_t stream) {
    if (!p || !d_in) { set_error("refactor_graph: null argument"); return kErrArg; }
    hipStream_t s = (hipStream_t)stream;
    if (!p->g_ref || p->g_ref_in != d_in) {
        if (int r = prodm::plan_side(p)) return r;  // streams and events exist before the capture
        const bool on = p->timer.on;
        p->timer.on = false;  // no timing events in the graph
        p->capturing = true;
        const int rc = prodm::graph_capture(s, &p->g_ref, [&] { return prodm_mdr_refactor(p, d_in, stream); });
        p->capturing = false;
        p->timer.on = on;
        if (rc) return rc;
        p->g_ref_in = d_in;
    }
    PRODM_HIP(hipGraphLaunch(p->g_ref, s));
    return kOk;
}

int prodm_mdr_retrieve_graph(prodm_mdr_plan *p, const uint8_t *k, void *d_out, prodm_stream_t stream) {
    if (!p || !k || !d_out) { set_error("retrieve_graph: null argument"); return kErrArg; }
    if (p->encoder == PRODM_ENC_PERBIT) {
        set_error("retrieve_graph: PerBit decodes on the host and has no graph form");
        return kErrUnsupported;
    }
    hipStream_t s = (hipStream_t)stream;
    if (!p->g_ret || p->g_ret_out != d_out || std::memcmp(p->g_ret_k, k, p->g.L + 1) != 0) {
        if (int r = prodm::plan_side(p)) return r;
        if (int r = prodm::plan_flev(p)) return r;
        const bool on = p->timer.on;
        p->timer.on = false;
        p->capturing = true;
        const int rc = prodm::graph_capture(s, &p->g_ret, [&] { return prodm_mdr_retrieve(p, k, d_out, stream); });
        p->capturing = false;
        p->timer.on = on;
        if (rc) return rc;
        p->g_ret_out = d_out;
        std::memcpy(p->g_ret_k, k, p->g.L + 1);
    }
    PRODM_HIP(hipGraphLaunch(p->g_ret, s));
    return kOk;
}

int prodm_mdr_retrieve(prodm_mdr_plan *p, const uint8_t *k, void *d_out, prodm_stream_t stream) {
    if (!p || !k || !d_out) { set_error("retrieve: null argument"); return kErrArg; }
    if (p->dtype == PRODM_F32) return plan_retrieve<float>(p, k, static_cast<float *>(d_out), (hipStream_t)stream);
    return plan_retrieve<double>(p, k, static_cast<double *>(d_out), (hipStream_t)stream);
}

}  // extern "C"

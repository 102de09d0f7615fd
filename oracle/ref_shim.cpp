// ref_shim.cpp -- C entry points over the ProDM reference headers, compiled
// in place from the reference tree by oracle/Makefile (target `ref`) into
// oracle/_ref/libref.so. TEST INFRASTRUCTURE ONLY: it is the compiled
// reference used to pin the oracle and to write golden fixtures; nothing in
// the product links it. Only the headers that build without MGARDx are used
// (encoders, collectors, interleaver, estimators, interpreters, compressors).
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <queue>
#include <sstream>
#include <vector>

#include "MDR/RefactorUtils.hpp"
#include "MDR/Interleaver/DirectInterleaver.hpp"
#include "MDR/BitplaneEncoder/GroupedBPEncoder.hpp"
#include "MDR/BitplaneEncoder/NegaBinaryBPEncoder.hpp"
#include "MDR/BitplaneEncoder/PerBitBPEncoder.hpp"
#include "MDR/ErrorCollector/MaxErrorCollector.hpp"
#include "MDR/ErrorCollector/SquaredErrorCollector.hpp"
#include "MDR/ErrorEstimator/MaxErrorEstimator.hpp"
#include "MDR/ErrorEstimator/SquaredErrorEstimator.hpp"
#include "MDR/SizeInterpreter/GreedyBasedSizeInterpreter.hpp"
#include "MDR/SizeInterpreter/BasicSizeInterpreter.hpp"
#include "MDR/LosslessCompressor/LevelCompressor.hpp"

namespace {
// the reference interpreters print progress to std::cout; keep it quiet
struct Mute {
    std::streambuf *old;
    std::ostringstream sink;
    Mute() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~Mute() { std::cout.rdbuf(old); }
};

template <class T>
std::vector<uint32_t> vec(const uint32_t *p, int n) { return std::vector<uint32_t>(p, p + n); }

template <class Enc, class TD, class TS>
int nb_encode(const TD *x, int32_t n, int32_t exp, int nb, TS *planes) {
    Enc enc;
    std::vector<uint32_t> sizes;
    auto streams = enc.encode(x, n, exp, (uint8_t)nb, sizes);
    uint32_t nw = (n - 1) / (8 * sizeof(TS)) + 1;
    for (int p = 0; p < nb; p++) {
        if (sizes[p] != nw * sizeof(TS)) return -1;
        memcpy(planes + (size_t)p * nw, streams[p], sizes[p]);
        free(streams[p]);
    }
    return 0;
}

template <class Enc, class TD, class TS>
int nb_decode(const TS *const *planes, int32_t n, int32_t exp, int start, int k, TD *out) {
    Enc enc;
    std::vector<const uint8_t *> streams;
    for (int i = 0; i < k; i++) streams.push_back(reinterpret_cast<const uint8_t *>(planes[i]));
    TD *r = enc.progressive_decode(streams, n, exp, (uint8_t)start, (uint8_t)k, 0);
    memcpy(out, r, sizeof(TD) * n);
    free(r);
    return 0;
}

template <class Enc, class TD>
int gb_encode(const TD *x, int32_t n, int32_t exp, int nb, uint8_t *out, uint64_t cap, uint32_t *sizes) {
    Enc enc;
    std::vector<uint32_t> sz;
    auto streams = enc.encode(x, n, exp, (uint8_t)nb, sz);
    uint64_t off = 0;
    int rc = 0;
    for (int p = 0; p < nb; p++) {
        sizes[p] = sz[p];
        if (off + sz[p] > cap) rc = -1;
        else memcpy(out + off, streams[p], sz[p]);
        off += sz[p];
        free(streams[p]);
    }
    return rc;
}

template <class Enc, class TD>
int gb_decode(void *h, int level, const uint8_t *const *streams, int32_t n, int32_t exp, int start, int k,
              TD *out) {
    Enc *enc = static_cast<Enc *>(h);
    std::vector<const uint8_t *> s(streams, streams + k);
    TD *r = enc->progressive_decode(s, n, exp, (uint8_t)start, (uint8_t)k, level);
    memcpy(out, r, sizeof(TD) * n);
    free(r);
    return 0;
}

template <class TD, class TS>
int pb_encode(const TD *x, int32_t n, int32_t exp, int nb, uint8_t *out, uint64_t cap, uint32_t *sizes) {
    MDR::PerBitBPEncoder<TD, TS> enc;
    std::vector<uint32_t> sz;
    auto streams = enc.encode(x, n, exp, (uint8_t)nb, sz);
    uint64_t off = 0;
    int rc = 0;
    for (int p = 0; p < nb; p++) {
        sizes[p] = sz[p];
        if (off + sz[p] > cap) rc = -1;
        else memcpy(out + off, streams[p], sz[p]);
        off += sz[p];
        free(streams[p]);
    }
    return rc;
}
std::vector<std::vector<uint32_t>> unflat_u32(const uint32_t *p, int L, const int *counts) {
    std::vector<std::vector<uint32_t>> v;
    for (int i = 0; i < L; i++) { v.emplace_back(p, p + counts[i]); p += counts[i]; }
    return v;
}
std::vector<std::vector<double>> unflat_f64(const double *p, int L, const int *counts) {
    std::vector<std::vector<double>> v;
    for (int i = 0; i < L; i++) { v.emplace_back(p, p + counts[i] + 1); p += counts[i] + 1; }
    return v;
}
}  // namespace

extern "C" {

void ref_level_dims(const uint32_t *dims, int nd, int L, uint32_t *out) {
    auto ld = MDR::compute_level_dims(std::vector<uint32_t>(dims, dims + nd), L);
    for (int l = 0; l <= L; l++)
        for (int d = 0; d < nd; d++) out[l * nd + d] = ld[l][d];
}

void ref_level_elements(const uint32_t *dims, int nd, int L, uint32_t *out) {
    auto ld = MDR::compute_level_dims(std::vector<uint32_t>(dims, dims + nd), L);
    auto le = MDR::compute_level_elements(ld, L);
    for (int l = 0; l <= L; l++) out[l] = le[l];
}

float ref_maxabs_f32(const float *x, uint32_t n) { return MDR::compute_max_abs_value(x, n); }
double ref_maxabs_f64(const double *x, uint32_t n) { return MDR::compute_max_abs_value(x, n); }

// nd must be 1 or 3 (the reference exits for other ranks)
void ref_interleave_f32(const float *data, const uint32_t *dims, int nd, const uint32_t *fine,
                        const uint32_t *coarse, const uint32_t *strides, float *buf) {
    MDR::DirectInterleaver<float> il;
    il.interleave(data, vec<float>(dims, nd), vec<float>(fine, nd), vec<float>(coarse, nd), buf,
                  strides ? vec<float>(strides, nd) : std::vector<uint32_t>());
}
void ref_interleave_f64(const double *data, const uint32_t *dims, int nd, const uint32_t *fine,
                        const uint32_t *coarse, const uint32_t *strides, double *buf) {
    MDR::DirectInterleaver<double> il;
    il.interleave(data, vec<double>(dims, nd), vec<double>(fine, nd), vec<double>(coarse, nd), buf,
                  strides ? vec<double>(strides, nd) : std::vector<uint32_t>());
}
void ref_reposition_f32(const float *buf, const uint32_t *dims, int nd, const uint32_t *fine,
                        const uint32_t *coarse, const uint32_t *strides, float *data) {
    MDR::DirectInterleaver<float> il;
    il.reposition(buf, vec<float>(dims, nd), vec<float>(fine, nd), vec<float>(coarse, nd), data,
                  strides ? vec<float>(strides, nd) : std::vector<uint32_t>());
}
void ref_reposition_f64(const double *buf, const uint32_t *dims, int nd, const uint32_t *fine,
                        const uint32_t *coarse, const uint32_t *strides, double *data) {
    MDR::DirectInterleaver<double> il;
    il.reposition(buf, vec<double>(dims, nd), vec<double>(fine, nd), vec<double>(coarse, nd), data,
                  strides ? vec<double>(strides, nd) : std::vector<uint32_t>());
}

// n must be >= the block size (the reference's `n - block_size` wraps otherwise)
int ref_nb_encode_f32(const float *x, int32_t n, int32_t exp, int nb, uint32_t *planes) {
    return nb_encode<MDR::NegaBinaryBPEncoder<float, uint32_t>>(x, n, exp, nb, planes);
}
int ref_nb_encode_f64(const double *x, int32_t n, int32_t exp, int nb, uint64_t *planes) {
    return nb_encode<MDR::NegaBinaryBPEncoder<double, uint64_t>>(x, n, exp, nb, planes);
}
int ref_nb_decode_f32(const uint32_t *const *planes, int32_t n, int32_t exp, int start, int k, float *out) {
    return nb_decode<MDR::NegaBinaryBPEncoder<float, uint32_t>>(planes, n, exp, start, k, out);
}
int ref_nb_decode_f64(const uint64_t *const *planes, int32_t n, int32_t exp, int start, int k, double *out) {
    return nb_decode<MDR::NegaBinaryBPEncoder<double, uint64_t>>(planes, n, exp, start, k, out);
}

int ref_gb_encode_f32(const float *x, int32_t n, int32_t exp, int nb, uint8_t *out, uint64_t cap,
                      uint32_t *sizes) {
    return gb_encode<MDR::GroupedBPEncoder<float, uint32_t>>(x, n, exp, nb, out, cap, sizes);
}
int ref_gb_encode_f64(const double *x, int32_t n, int32_t exp, int nb, uint8_t *out, uint64_t cap,
                      uint32_t *sizes) {
    return gb_encode<MDR::GroupedBPEncoder<double, uint64_t>>(x, n, exp, nb, out, cap, sizes);
}
void *ref_gb_new_f32() { return new MDR::GroupedBPEncoder<float, uint32_t>(); }
void *ref_gb_new_f64() { return new MDR::GroupedBPEncoder<double, uint64_t>(); }
void ref_gb_free_f32(void *h) { delete static_cast<MDR::GroupedBPEncoder<float, uint32_t> *>(h); }
void ref_gb_free_f64(void *h) { delete static_cast<MDR::GroupedBPEncoder<double, uint64_t> *>(h); }
int ref_gb_decode_f32(void *h, int level, const uint8_t *const *streams, int32_t n, int32_t exp, int start,
                      int k, float *out) {
    return gb_decode<MDR::GroupedBPEncoder<float, uint32_t>>(h, level, streams, n, exp, start, k, out);
}
int ref_gb_decode_f64(void *h, int level, const uint8_t *const *streams, int32_t n, int32_t exp, int start,
                      int k, double *out) {
    return gb_decode<MDR::GroupedBPEncoder<double, uint64_t>>(h, level, streams, n, exp, start, k, out);
}

// PerBitBPEncoder (PerBitBPEncoder.hpp:72-146, 282-391): streams packed back
// to back into out (cap bytes), sizes[p] bytes each; decoder handles keep the
// per-level sign state
int ref_pb_encode_f32(const float *x, int32_t n, int32_t exp, int nb, uint8_t *out, uint64_t cap, uint32_t *sizes) {
    return pb_encode<float, uint32_t>(x, n, exp, nb, out, cap, sizes);
}
int ref_pb_encode_f64(const double *x, int32_t n, int32_t exp, int nb, uint8_t *out, uint64_t cap, uint32_t *sizes) {
    return pb_encode<double, uint64_t>(x, n, exp, nb, out, cap, sizes);
}
void *ref_pb_new_f32() { return new MDR::PerBitBPEncoder<float, uint32_t>(); }
void *ref_pb_new_f64() { return new MDR::PerBitBPEncoder<double, uint64_t>(); }
void ref_pb_free_f32(void *h) { delete static_cast<MDR::PerBitBPEncoder<float, uint32_t> *>(h); }
void ref_pb_free_f64(void *h) { delete static_cast<MDR::PerBitBPEncoder<double, uint64_t> *>(h); }
int ref_pb_decode_f32(void *h, int level, const uint8_t *const *streams, int32_t n, int32_t exp, int start, int k,
                      float *out) {
    return gb_decode<MDR::PerBitBPEncoder<float, uint32_t>, float>(h, level, streams, n, exp, start, k, out);
}
int ref_pb_decode_f64(void *h, int level, const uint8_t *const *streams, int32_t n, int32_t exp, int start, int k,
                      double *out) {
    return gb_decode<MDR::PerBitBPEncoder<double, uint64_t>, double>(h, level, streams, n, exp, start, k, out);
}

// the encoders' level-error overloads (NegaBinaryBPEncoder.hpp:65-125,
// GroupedBPEncoder.hpp:76-145); n >= the block size
void ref_nb_level_errors_f32(const float *x, int32_t n, int32_t exp, int nb, double *out) {
    MDR::NegaBinaryBPEncoder<float, uint32_t> enc;
    std::vector<uint32_t> sizes;
    std::vector<double> le;
    auto streams = enc.encode(x, n, exp, (uint8_t)nb, sizes, le);
    for (auto *p : streams) free(p);
    std::copy(le.begin(), le.end(), out);
}
void ref_gb_level_errors_f32(const float *x, int32_t n, int32_t exp, int nb, double *out) {
    MDR::GroupedBPEncoder<float, uint32_t> enc;
    std::vector<uint32_t> sizes;
    std::vector<double> le;
    auto streams = enc.encode(x, n, exp, (uint8_t)nb, sizes, le);
    for (auto *p : streams) free(p);
    std::copy(le.begin(), le.end(), out);
}

void ref_pb_level_errors_f32(const float *x, int32_t n, int32_t exp, int nb, double *out) {
    MDR::PerBitBPEncoder<float, uint32_t> enc;
    std::vector<uint32_t> sizes;
    std::vector<double> le;
    auto streams = enc.encode(x, n, exp, (uint8_t)nb, sizes, le);
    for (auto *p : streams) free(p);
    std::copy(le.begin(), le.end(), out);
}

void ref_collect_max_f32(float bound, int nb, double *out) {
    MDR::MaxErrorCollector<float> c;
    auto v = c.collect_level_error(nullptr, 0, nb, bound);
    std::copy(v.begin(), v.end(), out);
}
void ref_collect_max_f64(double bound, int nb, double *out) {
    MDR::MaxErrorCollector<double> c;
    auto v = c.collect_level_error(nullptr, 0, nb, bound);
    std::copy(v.begin(), v.end(), out);
}
void ref_collect_squared_f32(const float *x, uint64_t n, int nb, float maxerr, double *out) {
    MDR::SquaredErrorCollector<float> c;
    auto v = c.collect_level_error(x, n, nb, maxerr);
    std::copy(v.begin(), v.end(), out);
}

double ref_max_estimator_ob_c(int nd) {
    // c is private; recover it from estimate_error(1, level)
    MDR::MaxErrorEstimatorOB<float> e(nd);
    return e.estimate_error(1.0f, 0);
}

// Interpreters. kind: 0 greedy, 1 sign-exclude greedy, 2 negabinary greedy,
// 3 in-order, 4 round-robin. est: 0 MaxOB(nd), 1 MaxHB, 2 SNorm(nd, L, s)
// counts[l] = planes of level l; sizes flattened; errors flattened with
// counts[l] + 1 entries per level; index in/out.
int ref_interpret(int kind, int est, int nd, double s_param, int L, const int *counts,
                  const uint32_t *sizes, const double *errors, double tol, uint8_t *index,
                  uint32_t *retrieve) {
    Mute mute;
    auto ls = unflat_u32(sizes, L, counts);
    auto le = unflat_f64(errors, L, counts);
    std::vector<uint8_t> idx(index, index + L);
    std::vector<uint32_t> r;
    auto run = [&](auto estimator) {
        using E = decltype(estimator);
        switch (kind) {
            case 0: r = MDR::GreedyBasedSizeInterpreter<E>(estimator).interpret_retrieve_size(ls, le, tol, idx); break;
            case 1: r = MDR::SignExcludeGreedyBasedSizeInterpreter<E>(estimator).interpret_retrieve_size(ls, le, tol, idx); break;
            case 2: r = MDR::NegaBinaryGreedyBasedSizeInterpreter<E>(estimator).interpret_retrieve_size(ls, le, tol, idx); break;
            case 3: r = MDR::InorderSizeInterpreter<E>(estimator).interpret_retrieve_size(ls, le, tol, idx); break;
            case 4: r = MDR::RoundRobinSizeInterpreter<E>(estimator).interpret_retrieve_size(ls, le, tol, idx); break;
        }
    };
    if (est == 0) run(MDR::MaxErrorEstimatorOB<float>(nd));
    else if (est == 1) run(MDR::MaxErrorEstimatorHB<float>());
    else run(MDR::SNormErrorEstimator<float>(nd, L - 1, (float)s_param));
    for (int l = 0; l < L; l++) { index[l] = idx[l]; retrieve[l] = r[l]; }
    return 0;
}

// Level compressors. kind 0: Default, 1: Adaptive(latter). Streams are
// concatenated in `in` with `sizes`; compressed streams are written back to
// back into out; returns stopping index (or -1 if out is too small).
int ref_compress_level(int kind, int latter, int nb, const uint8_t *in, const uint32_t *sizes, uint8_t *out,
                       uint64_t cap, uint32_t *out_sizes) {
    std::vector<uint8_t *> streams;
    std::vector<uint32_t> sz(sizes, sizes + nb);
    const uint8_t *p = in;
    for (int i = 0; i < nb; i++) {
        uint8_t *b = (uint8_t *)malloc(sizes[i] ? sizes[i] : 1);
        memcpy(b, p, sizes[i]);
        p += sizes[i];
        streams.push_back(b);
    }
    uint8_t stop;
    if (kind == 0) stop = MDR::DefaultLevelCompressor().compress_level(streams, sz);
    else stop = MDR::AdaptiveLevelCompressor(latter).compress_level(streams, sz);
    uint64_t off = 0;
    int rc = stop;
    for (int i = 0; i < nb; i++) {
        out_sizes[i] = sz[i];
        if (off + sz[i] > cap) rc = -1;
        else memcpy(out + off, streams[i], sz[i]);
        off += sz[i];
        free(streams[i]);
    }
    return rc;
}

}  // extern "C"

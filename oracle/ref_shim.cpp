// ref_shim.cpp -- C entry points over the ProDM reference headers, compiled
// in place from the reference tree by oracle/Makefile (target `ref`) into
// oracle/_ref/libref.so. TEST INFRASTRUCTURE ONLY: it is the compiled
// reference used to pin the oracle and to write golden fixtures; nothing in
// the product links it. Only the headers that build without MGARDx are used
// (encoders, collectors, interleaver, estimators, interpreters, compressors,
// serialize/deserialize, the concatenated-level file writer and retriever).
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
#include "MDR/Retriever/FileRetriever.hpp"
#include "MDR/Writer/FileWriter.hpp"

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

// ---- ProDM's file layout, read and written by ProDM's own code ------------
namespace {
template <class T>
int read_mdr(const char *metadata_path, const char *const *level_paths, int nfiles, int latter, const uint8_t *k1,
             uint32_t *dims, int *ndims, int *nlevels, double *bounds, int *nplanes, uint32_t *level_sizes_out,
             uint8_t *stopping, uint32_t *level_num_out, int *negabinary, uint8_t *out, uint64_t cap,
             uint64_t *raw_sizes, uint64_t *retrieved) {
    std::vector<std::string> files(level_paths, level_paths + nfiles);
    MDR::ConcatLevelFileRetriever retriever(metadata_path, files);
    // the field order of ComposedReconstructor::load_metadata (ComposedReconstructor.hpp:133-145)
    uint8_t *metadata = retriever.load_metadata();
    uint8_t const *pos = metadata;
    std::vector<uint32_t> dimensions;
    std::vector<T> level_error_bounds;
    std::vector<std::vector<uint32_t>> level_sizes;
    std::vector<uint8_t> stopping_indices;
    std::vector<uint32_t> level_num;
    uint8_t num_dims = *(pos++);
    MDR::deserialize(pos, num_dims, dimensions);
    uint8_t num_levels = *(pos++);
    MDR::deserialize(pos, num_levels, level_error_bounds);
    MDR::deserialize(pos, num_levels, level_sizes);
    MDR::deserialize(pos, num_levels, stopping_indices);
    MDR::deserialize(pos, num_levels, level_num);
    const int nega = *(pos++);
    free(metadata);
    if (num_levels != nfiles || num_dims > 3) return -1;
    *ndims = num_dims;
    *nlevels = num_levels;
    *negabinary = nega;
    for (int d = 0; d < num_dims; d++) dims[d] = dimensions[d];
    const int nb = level_sizes.empty() ? 0 : (int)level_sizes[0].size();
    *nplanes = nb;
    for (int l = 0; l < num_levels; l++) {
        if ((int)level_sizes[l].size() != nb) return -2;
        bounds[l] = (double)level_error_bounds[l];
        stopping[l] = stopping_indices[l];
        level_num_out[l] = level_num[l];
        for (int p = 0; p < nb; p++) level_sizes_out[l * nb + p] = level_sizes[l][p];
    }
    // two progressive steps through the retriever's per-level offsets, as a
    // reconstructor reads: planes [0, k1[l]), then the rest
    std::vector<std::vector<std::vector<uint8_t>>> planes(num_levels, std::vector<std::vector<uint8_t>>(nb));
    std::vector<uint8_t> prev(num_levels, 0), held(num_levels);
    for (int step = 0; step < 2; step++) {
        std::vector<uint32_t> rs(num_levels, 0);
        for (int l = 0; l < num_levels; l++) {
            held[l] = step == 0 ? std::min<int>(k1[l], nb) : nb;
            for (int p = prev[l]; p < held[l]; p++) rs[l] += level_sizes[l][p];
        }
        auto comps = retriever.retrieve_level_components(level_sizes, rs, prev, held);
        for (int l = 0; l < num_levels; l++) {
            const int k = held[l] - prev[l];
            if (k <= 0) continue;
            std::vector<uint64_t> raw(k);
            for (int i = 0; i < k; i++) {  // decompressed length: the frame's size_t prefix
                const int p = prev[l] + i;
                size_t len = level_sizes[l][p];
                if (p <= stopping_indices[l] || p >= latter) memcpy(&len, comps[l][i], sizeof(size_t));
                raw[i] = len;
            }
            MDR::AdaptiveLevelCompressor c(latter);
            c.decompress_level(comps[l], level_sizes[l], prev[l], (uint8_t)k, stopping_indices[l]);
            for (int i = 0; i < k; i++) planes[l][prev[l] + i].assign(comps[l][i], comps[l][i] + raw[i]);
            c.decompress_release();
        }
        prev = held;
    }
    *retrieved = retriever.get_retrieved_size();
    retriever.release();
    uint64_t off = 0;
    for (int l = 0; l < num_levels; l++)
        for (int p = 0; p < nb; p++) {
            const auto &v = planes[l][p];
            raw_sizes[l * nb + p] = v.size();
            if (off + v.size() > cap) return -3;
            if (!v.empty()) memcpy(out + off, v.data(), v.size());
            off += v.size();
        }
    return 0;
}

// streams: level-major, plane order, back to back, raw sizes in sizes
template <class T>
int write_mdr(const char *metadata_path, const char *const *level_paths, int ndims, const uint32_t *dims,
              int nlevels, int nb, const double *bounds, int negabinary, int latter, const uint8_t *streams,
              const uint32_t *sizes) {
    std::vector<uint32_t> dimensions(dims, dims + ndims);
    std::vector<T> level_error_bounds;
    std::vector<std::vector<uint8_t *>> level_components(nlevels);
    std::vector<std::vector<uint32_t>> level_sizes(nlevels);
    std::vector<uint8_t> stopping_indices;
    const uint8_t *p = streams;
    for (int l = 0; l < nlevels; l++) {
        level_error_bounds.push_back((T)bounds[l]);
        for (int q = 0; q < nb; q++) {
            const uint32_t n = sizes[l * nb + q];
            uint8_t *b = (uint8_t *)malloc(n ? n : 1);
            memcpy(b, p, n);
            p += n;
            level_components[l].push_back(b);
            level_sizes[l].push_back(n);
        }
        stopping_indices.push_back(MDR::AdaptiveLevelCompressor(latter).compress_level(level_components[l], level_sizes[l]));
    }
    MDR::ConcatLevelFileWriter writer(metadata_path, std::vector<std::string>(level_paths, level_paths + nlevels));
    std::vector<uint32_t> level_num = writer.write_level_components(level_components, level_sizes);
    // the field order of ComposedRefactor::write_metadata (ComposedRefactor.hpp:48-67)
    using MDR::get_size;
    const uint32_t metadata_size = sizeof(uint8_t) + get_size(dimensions) + sizeof(uint8_t) +
                                   get_size(level_error_bounds) + get_size(level_sizes) +
                                   get_size(stopping_indices) + get_size(level_num) + 1;
    uint8_t *metadata = (uint8_t *)malloc(metadata_size);
    uint8_t *pos = metadata;
    *(pos++) = (uint8_t)dimensions.size();
    MDR::serialize(dimensions, pos);
    *(pos++) = (uint8_t)level_error_bounds.size();
    MDR::serialize(level_error_bounds, pos);
    MDR::serialize(level_sizes, pos);
    MDR::serialize(stopping_indices, pos);
    MDR::serialize(level_num, pos);
    *(pos++) = (uint8_t)negabinary;
    writer.write_metadata(metadata, metadata_size);
    free(metadata);
    for (auto &lv : level_components)
        for (uint8_t *b : lv) free(b);
    return 0;
}
}  // namespace

extern "C" {
int ref_read_mdr(const char *metadata_path, const char *const *level_paths, int nfiles, int dtype, int latter,
                 const uint8_t *k1, uint32_t *dims, int *ndims, int *nlevels, double *bounds, int *nplanes,
                 uint32_t *level_sizes, uint8_t *stopping, uint32_t *level_num, int *negabinary, uint8_t *out,
                 uint64_t cap, uint64_t *raw_sizes, uint64_t *retrieved) {
    return dtype ? read_mdr<double>(metadata_path, level_paths, nfiles, latter, k1, dims, ndims, nlevels, bounds,
                                    nplanes, level_sizes, stopping, level_num, negabinary, out, cap, raw_sizes, retrieved)
                 : read_mdr<float>(metadata_path, level_paths, nfiles, latter, k1, dims, ndims, nlevels, bounds,
                                   nplanes, level_sizes, stopping, level_num, negabinary, out, cap, raw_sizes, retrieved);
}
int ref_write_mdr(const char *metadata_path, const char *const *level_paths, int ndims, const uint32_t *dims,
                  int nlevels, int nb, int dtype, const double *bounds, int negabinary, int latter,
                  const uint8_t *streams, const uint32_t *sizes) {
    return dtype ? write_mdr<double>(metadata_path, level_paths, ndims, dims, nlevels, nb, bounds, negabinary, latter,
                                     streams, sizes)
                 : write_mdr<float>(metadata_path, level_paths, ndims, dims, nlevels, nb, bounds, negabinary, latter,
                                    streams, sizes);
}
}  // extern "C"

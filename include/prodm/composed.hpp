// composed.hpp -- composers with the template shapes of ProDM's
// ComposedRefactor (Refactor/ComposedRefactor.hpp:15-141) and
// ComposedReconstructor (Reconstructor/ComposedReconstructor.hpp:17-297).
//
// They drive any components with ProDM's interfaces (the GPU ones of
// components.hpp, or ProDM's CPU ones) in ProDM's order and write / read
// ProDM's file layout (io.hpp), so files interoperate both ways.
//
// Differences from ProDM, each at a point where ProDM's behaviour is
// undefined or unwired:
//  * A squared-error collector's tables are collected at refactor and
//    appended to metadata.bin after ProDM's last field (ProDM's reader stops
//    before them), so an L2 estimator has its tables at reconstruct; ProDM
//    leaves them empty (ComposedRefactor.hpp:116-117, ComposedReconstructor.hpp:42-44).
//  * The progressive update of the previous field works for 1-, 2- and 3-D
//    fields (ProDM: 1-D and 3-D, ComposedReconstructor.hpp:217-234).
#ifndef PRODM_COMPOSED_HPP
#define PRODM_COMPOSED_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../prodm_mdr.h"
#include "estimators.hpp"
#include "io.hpp"

namespace prodm {

// Collectors that produce squared-error tables (their tables are stored).
template <class C> struct is_squared_collector : std::false_type {};
template <class T> class GpuSquaredErrorCollector;
template <class T> struct is_squared_collector<GpuSquaredErrorCollector<T>> : std::true_type {};
template <class T> class GpuEncoderErrorCollector;
template <class T> struct is_squared_collector<GpuEncoderErrorCollector<T>> : std::true_type {};

inline std::vector<std::vector<uint32_t>> level_dims_of(const std::vector<uint32_t> &dims, int L) {
    std::vector<uint32_t> flat((L + 1) * dims.size());
    std::vector<uint64_t> el(L + 1);
    prodm_level_geometry(dims.data(), (int)dims.size(), L, flat.data(), el.data());
    std::vector<std::vector<uint32_t>> out(L + 1);
    for (int l = 0; l <= L; l++) out[l].assign(flat.begin() + l * dims.size(), flat.begin() + (l + 1) * dims.size());
    return out;
}
inline std::vector<uint64_t> level_elements_of(const std::vector<uint32_t> &dims, int L) {
    std::vector<uint32_t> flat((L + 1) * dims.size());
    std::vector<uint64_t> el(L + 1);
    prodm_level_geometry(dims.data(), (int)dims.size(), L, flat.data(), el.data());
    return el;
}

// Metadata bytes from a retriever; the size is known when the retriever
// reports it (ProDM's retriever returns the bytes only).
template <class R> auto load_bytes_impl(R &r, size_t *size, int) -> decltype(r.load_metadata(size)) {
    return r.load_metadata(size);
}
template <class R> uint8_t *load_bytes_impl(R &r, size_t *size, long) {
    *size = 0;  // unknown: the optional trailing block is not read
    return r.load_metadata();
}
template <class R> uint8_t *load_bytes(R &r, size_t *size) { return load_bytes_impl(r, size, 0); }

template <class T, class Decomposer, class Interleaver, class Encoder, class Compressor, class ErrorCollector,
          class Writer>
class ComposedRefactor {
  public:
    ComposedRefactor(Decomposer decomposer, Interleaver interleaver, Encoder encoder, Compressor compressor,
                     ErrorCollector collector, Writer writer)
        : decomposer(decomposer), interleaver(interleaver), encoder(encoder), compressor(compressor),
          collector(collector), writer(writer) {}

    void refactor(T const *data_, const std::vector<uint32_t> &dims, uint8_t target_level, uint8_t num_bitplanes) {
        dimensions = dims;
        data.assign(data_, data_ + product_of(dims));
        if (refactor(target_level, num_bitplanes)) level_num = writer.write_level_components(level_components, level_sizes);
        write_metadata();
        for (auto &lv : level_components)
            for (uint8_t *p : lv) std::free(p);
        level_components.clear();
    }

    void write_metadata() const {
        MdrMetadata<T> m;
        m.dims = dimensions;
        m.level_error_bounds = level_error_bounds;
        m.level_sizes = level_sizes;
        m.stopping_indices = stopping_indices;
        m.level_num = level_num;
        m.negabinary = negabinary;
        m.level_squared_errors = level_squared_errors;
        const auto bytes = m.serialize_bytes();
        writer.write_metadata(bytes.data(), (uint32_t)bytes.size());
    }
    void print() const {
        std::printf("composed refactor:\n  ");
        decomposer.print();
        std::printf("  ");
        interleaver.print();
        std::printf("  ");
        encoder.print();
    }

    bool negabinary = false;

  private:
    static uint64_t product_of(const std::vector<uint32_t> &d) {
        uint64_t n = 1;
        for (uint32_t v : d) n *= v;
        return n;
    }
    bool refactor(uint8_t target_level, uint8_t nb) {
        // ProDM's limit (ComposedRefactor.hpp:79-83), in its uint8 arithmetic
        const uint8_t max_level = (uint8_t)(int)(std::log2((double)*std::min_element(dimensions.begin(), dimensions.end())) - 1);
        if (target_level > max_level) {
            std::fprintf(stderr, "Target level is higher than %d\n", (int)max_level);
            return false;
        }
        decomposer.decompose(data.data(), dimensions, target_level);
        level_error_bounds.clear();
        level_squared_errors.clear();
        level_components.clear();
        level_sizes.clear();
        stopping_indices.clear();
        const auto ld = level_dims_of(dimensions, target_level);
        const auto el = level_elements_of(dimensions, target_level);
        const std::vector<uint32_t> zero(dimensions.size(), 0);
        for (int l = 0; l <= target_level; l++) {
            std::vector<T> buf(el[l] ? el[l] : 1);
            interleaver.interleave(data.data(), dimensions, ld[l], l ? ld[l - 1] : zero, buf.data());
            T m = 0;
            for (uint64_t i = 0; i < el[l]; i++) {
                const T v = std::fabs(buf[i]);
                if (v > m) m = v;
            }
            level_error_bounds.push_back(negabinary ? m * 4 : m);
            if constexpr (is_squared_collector<ErrorCollector>::value)
                level_squared_errors.push_back(collector.collect_level_error(buf.data(), el[l], nb, m));
            int e = 0;
            std::frexp(m, &e);
            std::vector<uint32_t> sizes;
            auto streams = encoder.encode(buf.data(), (int32_t)el[l], e, nb, sizes);
            stopping_indices.push_back(compressor.compress_level(streams, sizes));
            level_components.push_back(streams);
            level_sizes.push_back(sizes);
        }
        return true;
    }

    Decomposer decomposer;
    Interleaver interleaver;
    Encoder encoder;
    Compressor compressor;
    ErrorCollector collector;
    Writer writer;
    std::vector<T> data;
    std::vector<uint32_t> dimensions;
    std::vector<T> level_error_bounds;
    std::vector<uint8_t> stopping_indices;
    std::vector<std::vector<uint8_t *>> level_components;
    std::vector<std::vector<uint32_t>> level_sizes;
    std::vector<uint32_t> level_num;
    std::vector<std::vector<double>> level_squared_errors;
};

template <class T, class Decomposer, class Interleaver, class Encoder, class Compressor, class SizeInterpreter,
          class ErrorEstimator, class Retriever>
class ComposedReconstructor {
  public:
    ComposedReconstructor(Decomposer decomposer, Interleaver interleaver, Encoder encoder, Compressor compressor,
                          SizeInterpreter interpreter, Retriever retriever)
        : decomposer(decomposer), interleaver(interleaver), encoder(encoder), compressor(compressor),
          interpreter(interpreter), retriever(retriever) {}

    T *reconstruct(double tolerance) { return reconstruct(tolerance, -1); }
    // Plans with the interpreter, reads the new planes, decodes them and
    // updates the field (ComposedReconstructor.hpp:27-100). A field whose
    // finest levels received no planes is returned at the resolution of the
    // finest level that did (get_current_dimensions), as in ProDM.
    T *reconstruct(double tolerance, int max_level) {
        const int L = (int)level_error_bounds.size() - 1;
        std::vector<std::vector<double>> errs;
        if constexpr (std::is_base_of<MaxErrorEstimator<T>, ErrorEstimator>::value) {
            for (int l = 0; l <= L; l++) {
                std::vector<double> t(level_sizes[l].size() + 1);
                if constexpr (std::is_same<T, float>::value)
                    prodm_collect_max_f32(level_error_bounds[l], (int)level_sizes[l].size(), t.data());
                else
                    prodm_collect_max_f64(level_error_bounds[l], (int)level_sizes[l].size(), t.data());
                errs.push_back(t);
            }
        } else if constexpr (std::is_base_of<SquaredErrorEstimator<T>, ErrorEstimator>::value) {
            if (level_squared_errors.size() != (size_t)L + 1) {
                std::fprintf(stderr, "prodm: metadata holds no squared-error tables (refactor with a squared collector)\n");
                return nullptr;
            }
            errs = level_squared_errors;
        } else {
            static_assert(std::is_base_of<MaxErrorEstimator<T>, ErrorEstimator>::value ||
                              std::is_base_of<SquaredErrorEstimator<T>, ErrorEstimator>::value,
                          "ErrorEstimator must derive from MaxErrorEstimator<T> or SquaredErrorEstimator<T>");
        }
        const std::vector<uint8_t> prev(level_num_bitplanes);
        if (max_level == -1 || max_level >= (int)level_num_bitplanes.size()) {
            auto rs = interpreter.interpret_retrieve_size(level_sizes, errs, tolerance, level_num_bitplanes);
            level_components = retriever.retrieve_level_components(level_sizes, rs, prev, level_num_bitplanes);
        } else {
            std::vector<std::vector<uint32_t>> s(level_sizes.begin(), level_sizes.begin() + max_level + 1);
            std::vector<std::vector<double>> e(errs.begin(), errs.begin() + max_level + 1);
            std::vector<uint8_t> k(level_num_bitplanes.begin(), level_num_bitplanes.begin() + max_level + 1);
            auto rs = interpreter.interpret_retrieve_size(s, e, tolerance, k);
            level_components = retriever.retrieve_level_components(s, rs, prev, k);
            for (int l = 0; l <= max_level; l++) level_num_bitplanes[l] = k[l];
        }
        int skipped = 0;
        for (int i = 0; i <= L; i++)
            if (level_num_bitplanes[L - i] != 0) {
                skipped = i;
                break;
            }
        const int level = L - skipped;
        update(level, prev);
        retriever.release();
        current_level = level;
        return data.data();
    }
    T *progressive_reconstruct(double tolerance, int max_level = -1) {
        reconstruct(tolerance, max_level);
        return data.data();
    }
    // Recomposes a reduced-resolution result to the full grid with the
    // missing coefficients zero (ComposedReconstructor.hpp:124-131).
    T *recompose_to_full() {
        const int L = (int)level_num.size() - 1;
        clear_outside(current_dimensions);
        decomposer.recompose(data.data(), dimensions, L - current_level, strides);
        return data.data();
    }
    void load_metadata() {
        size_t size = 0;
        uint8_t *raw = load_bytes(retriever, &size);
        MdrMetadata<T> m;
        m.parse(raw, size);
        std::free(raw);
        dimensions = m.dims;
        level_error_bounds = m.level_error_bounds;
        level_sizes = m.level_sizes;
        stopping_indices = m.stopping_indices;
        level_num = m.level_num;
        negabinary = m.negabinary;
        level_squared_errors = m.level_squared_errors;
        const size_t nl = level_error_bounds.size();
        level_num_bitplanes.assign(nl, 0);
        strides.assign(dimensions.size(), 0);
        uint32_t st = 1;
        for (int d = (int)dimensions.size() - 1; d >= 0; d--) {
            strides[d] = st;
            st *= dimensions[d];
        }
        data.assign(st, 0);
        current_level = -1;
    }
    const std::vector<uint32_t> &get_dimensions() const { return dimensions; }
    const std::vector<uint32_t> &get_current_dimensions() const { return current_dimensions; }
    int get_reconstruct_level() const { return current_level; }
    size_t get_retrieved_size() const { return retriever.get_retrieved_size(); }
    void print() const {
        std::printf("composed reconstructor:\n  ");
        decomposer.print();
        std::printf("  ");
        interleaver.print();
        std::printf("  ");
        encoder.print();
        std::printf("  ");
        interpreter.print();
        std::printf("  ");
        retriever.print();
    }

  private:
    template <class F> void for_box(const std::vector<uint32_t> &box, F f) const {
        const int nd = (int)box.size();
        std::vector<uint32_t> i(nd, 0);
        if (product_of(box) == 0) return;
        for (;;) {
            uint64_t off = 0;
            for (int d = 0; d < nd; d++) off += (uint64_t)i[d] * strides[d];
            f(off, i);
            int d = nd - 1;
            while (d >= 0 && ++i[d] == box[d]) i[d--] = 0;
            if (d < 0) return;
        }
    }
    static uint64_t product_of(const std::vector<uint32_t> &d) {
        uint64_t n = 1;
        for (uint32_t v : d) n *= v;
        return n;
    }
    void clear_outside(const std::vector<uint32_t> &inner) {
        for_box(dimensions, [&](uint64_t off, const std::vector<uint32_t> &i) {
            bool in = true;
            for (size_t d = 0; d < i.size(); d++) in = in && i[d] < inner[d];
            if (!in) data[off] = 0;
        });
    }
    T *decode_level(int l, const std::vector<uint8_t> &prev, uint64_t n) {
        const int k = level_num_bitplanes[l] - prev[l];
        compressor.decompress_level(level_components[l], level_sizes[l], prev[l], (uint8_t)k, stopping_indices[l]);
        int e = 0;
        std::frexp(negabinary ? level_error_bounds[l] / 4 : level_error_bounds[l], &e);
        T *r = encoder.progressive_decode(level_components[l], (int32_t)n, e, prev[l], (uint8_t)k, l);
        compressor.decompress_release();
        return r;
    }
    // ComposedReconstructor.hpp:187-258: decode the new planes of the held
    // levels as a delta, recompose it to the held resolution and add the
    // previous field; place the levels new to this call; recompose upward.
    void update(int target, const std::vector<uint8_t> &prev) {
        const int L = (int)level_num.size() - 1;
        const auto ld = level_dims_of(dimensions, L);
        const auto &box = ld[target];
        const auto el = level_elements_of(dimensions, L);
        std::vector<T> previous(data);
        std::fill(data.begin(), data.end(), (T)0);
        const std::vector<uint32_t> zero(dimensions.size(), 0);
        for (int l = 0; l <= current_level; l++) {
            if (level_num_bitplanes[l] - prev[l] > 0) {
                T *d = decode_level(l, prev, el[l]);
                interleaver.reposition(d, box, ld[l], l ? ld[l - 1] : zero, data.data(), strides);
                std::free(d);
            }
        }
        if (current_level >= 0) {
            if (current_level) decomposer.recompose(data.data(), current_dimensions, current_level, strides);
            for_box(current_dimensions, [&](uint64_t off, const std::vector<uint32_t> &) { data[off] += previous[off]; });
        }
        for (int l = current_level + 1; l <= target; l++) {
            T *d = decode_level(l, prev, el[l]);
            interleaver.reposition(d, box, ld[l], l ? ld[l - 1] : zero, data.data(), strides);
            std::free(d);
        }
        const int up = current_level >= 0 ? target - current_level : target;
        if (up > 0) decomposer.recompose(data.data(), box, up, strides);
        current_dimensions = box;
    }

    Decomposer decomposer;
    Interleaver interleaver;
    Encoder encoder;
    Compressor compressor;
    SizeInterpreter interpreter;
    Retriever retriever;
    std::vector<T> data;
    std::vector<uint32_t> dimensions, current_dimensions;
    std::vector<T> level_error_bounds;
    std::vector<uint8_t> level_num_bitplanes, stopping_indices;
    std::vector<std::vector<const uint8_t *>> level_components;
    std::vector<std::vector<uint32_t>> level_sizes;
    std::vector<uint32_t> level_num;
    std::vector<std::vector<double>> level_squared_errors;
    int current_level = -1;
    std::vector<uint32_t> strides;
    bool negabinary = true;
};

}  // namespace prodm
#endif

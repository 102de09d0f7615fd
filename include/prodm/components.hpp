// components.hpp -- GPU components with ProDM's component interfaces.
//
// Each class has the method names, signatures and ownership rules of the
// ProDM component it replaces, so it can be slotted into a composer
// (ProDM's own ComposedRefactor / ComposedReconstructor, or prodm's in
// composed.hpp) in place of the CPU class:
//
//   GpuMGARDOrthogonalDecomposer<T>  MDR::MGARDOrthoganalDecomposer<T>  (Decomposer/MGARD.hpp:10-40)
//   GpuMGARDHierarchicalDecomposer<T> MDR::MGARDHierarchicalDecomposer<T> (Decomposer/MGARD.hpp:42-70)
//   GpuDirectInterleaver<T>          MDR::DirectInterleaver<T>          (Interleaver/DirectInterleaver.hpp:9-67)
//   GpuNegaBinaryBPEncoder<T, S>     MDR::NegaBinaryBPEncoder<T, S>     (BitplaneEncoder/NegaBinaryBPEncoder.hpp:12-262)
//   GpuGroupedBPEncoder<T, S>        MDR::GroupedBPEncoder<T, S>        (BitplaneEncoder/GroupedBPEncoder.hpp:12-380)
//   GpuPerBitBPEncoder<T, S>         MDR::PerBitBPEncoder<T, S>         (BitplaneEncoder/PerBitBPEncoder.hpp:70-403)
//   MaxErrorCollector<T>             MDR::MaxErrorCollector<T>          (ErrorCollector/MaxErrorCollector.hpp:9-30)
//   GpuSquaredErrorCollector<T>      MDR::SquaredErrorCollector<T>      (ErrorCollector/SquaredErrorCollector.hpp:14-70)
//   GpuEncoderErrorCollector<T>      the encoders' level_errors overloads (NegaBinaryBPEncoder.hpp:65-125,
//                                    GroupedBPEncoder.hpp:76-145) behind the collector interface
//
// The interfaces take host pointers (ProDM's components work in host
// memory), so every call copies its operands to the device and back. The
// device-resident path without those copies is gpu_mdr.hpp.
//
// Buffers returned by encode / decode are malloc'd and released by the
// caller with free(), as in ProDM (BitplaneEncoderInterface.hpp:16-20).
#ifndef PRODM_COMPONENTS_HPP
#define PRODM_COMPONENTS_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <type_traits>
#include <vector>

#include "device.hpp"

namespace prodm {

// ---------------------------------------------------------------- decomposer
namespace detail {
// HIER: hierarchical basis (MGARDHierarchicalDecomposer), else orthogonal
template <class T, bool HIER>
class GpuMGARDDecomposer {
  public:
    explicit GpuMGARDDecomposer(Context ctx = make_context()) : ctx_(std::move(ctx)) {}
    void decompose(T *data, const std::vector<uint32_t> &dims, uint32_t target_level,
                   std::vector<uint32_t> strides = std::vector<uint32_t>()) const {
        run(true, data, dims, target_level, strides);
    }
    void recompose(T *data, const std::vector<uint32_t> &dims, uint32_t target_level,
                   std::vector<uint32_t> strides = std::vector<uint32_t>()) const {
        run(false, data, dims, target_level, strides);
    }
    void print() const {
        std::printf(HIER ? "MGARD hierarchical decomposer (MI355X)\n" : "MGARD orthogonal decomposer (MI355X)\n");
    }

  private:
    void run(bool forward, T *data, const std::vector<uint32_t> &dims, uint32_t L,
             const std::vector<uint32_t> &strides) const {
        if (L == 0) return;
        DeviceContext &c = *ctx_;
        const uint64_t n = span_of(dims, strides);
        T *a = c.a.as<T>(n), *b = c.b.as<T>(n);
        to_device(a, data, n, c.stream);
        if (!strides.empty())  // elements outside the box pass through
            hip_check(hipMemcpyAsync(b, a, n * sizeof(T), hipMemcpyDeviceToDevice, c.stream), "hipMemcpyAsync");
        const int nd = (int)dims.size();
        const size_t wsb = prodm_mgard_workspace_bytes(dims.data(), nd, (int)L, dtype_of<T>());
        void *ws = c.ws.get(wsb);
        const uint32_t *st = strides.empty() ? nullptr : strides.data();
        int rc;
        if constexpr (std::is_same<T, float>::value) {
            if (HIER)
                rc = forward ? prodm_decompose_hierarchical_f32(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream)
                             : prodm_recompose_hierarchical_f32(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream);
            else
                rc = forward ? prodm_decompose_f32(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream)
                             : prodm_recompose_f32(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream);
        } else {
            if (HIER)
                rc = forward ? prodm_decompose_hierarchical_f64(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream)
                             : prodm_recompose_hierarchical_f64(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream);
            else
                rc = forward ? prodm_decompose_f64(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream)
                             : prodm_recompose_f64(a, b, dims.data(), nd, (int)L, st, ws, wsb, c.stream);
        }
        check(rc, forward ? "decompose" : "recompose");
        to_host(data, b, n, c.stream);
    }
    Context ctx_;
};
}  // namespace detail

template <class T>
class GpuMGARDOrthogonalDecomposer : public detail::GpuMGARDDecomposer<T, false> {
  public:
    using detail::GpuMGARDDecomposer<T, false>::GpuMGARDDecomposer;
};
template <class T>
class GpuMGARDHierarchicalDecomposer : public detail::GpuMGARDDecomposer<T, true> {
  public:
    using detail::GpuMGARDDecomposer<T, true>::GpuMGARDDecomposer;
};

// ---------------------------------------------------------------- interleaver
template <class T>
class GpuDirectInterleaver {
  public:
    explicit GpuDirectInterleaver(Context ctx = make_context()) : ctx_(std::move(ctx)) {}
    void interleave(T const *data, const std::vector<uint32_t> &dims, const std::vector<uint32_t> &dims_fine,
                    const std::vector<uint32_t> &dims_coarse, T *buffer,
                    std::vector<uint32_t> strides = std::vector<uint32_t>()) const {
        DeviceContext &c = *ctx_;
        const uint64_t n = span_of(dims, strides), m = product(dims_fine) - product(dims_coarse);
        T *a = c.a.as<T>(n), *b = c.b.as<T>(m);
        to_device(a, data, n, c.stream);
        const uint32_t *st = strides.empty() ? nullptr : strides.data();
        const int nd = (int)dims.size();
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_interleave_f32(a, dims.data(), nd, dims_fine.data(), dims_coarse.data(), st, b, c.stream);
        else
            rc = prodm_interleave_f64(a, dims.data(), nd, dims_fine.data(), dims_coarse.data(), st, b, c.stream);
        check(rc, "interleave");
        to_host(buffer, b, m, c.stream);
    }
    void reposition(T const *buffer, const std::vector<uint32_t> &dims, const std::vector<uint32_t> &dims_fine,
                    const std::vector<uint32_t> &dims_coarse, T *data,
                    std::vector<uint32_t> strides = std::vector<uint32_t>()) const {
        DeviceContext &c = *ctx_;
        const uint64_t n = span_of(dims, strides), m = product(dims_fine) - product(dims_coarse);
        T *a = c.a.as<T>(n), *b = c.b.as<T>(m);
        to_device(a, data, n, c.stream);
        to_device(b, buffer, m, c.stream);
        const uint32_t *st = strides.empty() ? nullptr : strides.data();
        const int nd = (int)dims.size();
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_reposition_f32(b, dims.data(), nd, dims_fine.data(), dims_coarse.data(), st, a, c.stream);
        else
            rc = prodm_reposition_f64(b, dims.data(), nd, dims_fine.data(), dims_coarse.data(), st, a, c.stream);
        check(rc, "reposition");
        to_host(data, a, n, c.stream);
    }
    void print() const { std::printf("direct interleaver (MI355X)\n"); }

  private:
    Context ctx_;
};

namespace detail {
template <class T, class S> void check_stream_type() {
    static_assert((std::is_same<T, float>::value && std::is_same<S, uint32_t>::value) ||
                      (std::is_same<T, double>::value && std::is_same<S, uint64_t>::value),
                  "supported pairs: <float, uint32_t> and <double, uint64_t>");
}
}  // namespace detail

// ---------------------------------------------------------------- negabinary
template <class T, class S>
class GpuNegaBinaryBPEncoder {
  public:
    explicit GpuNegaBinaryBPEncoder(Context ctx = make_context()) : ctx_(std::move(ctx)) {
        detail::check_stream_type<T, S>();
    }
    std::vector<uint8_t *> encode(T const *data, int32_t n, int32_t exp, uint8_t num_bitplanes,
                                  std::vector<uint32_t> &stream_sizes) const {
        DeviceContext &c = *ctx_;
        const uint64_t nw = words(n), stride = (nw + 1) & ~1ull;
        T *x = c.a.as<T>(n);
        S *planes = c.b.as<S>(stride * num_bitplanes);
        to_device(x, data, (uint64_t)n, c.stream);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_negabinary_encode_f32(x, (uint32_t)n, exp, num_bitplanes, planes, stride, c.stream);
        else
            rc = prodm_negabinary_encode_f64(x, (uint32_t)n, exp, num_bitplanes, planes, stride, c.stream);
        check(rc, "negabinary encode");
        std::vector<S> host(stride * num_bitplanes);
        to_host(host.data(), planes, host.size(), c.stream);
        std::vector<uint8_t *> streams(num_bitplanes);
        stream_sizes.assign(num_bitplanes, (uint32_t)(nw * sizeof(S)));
        for (int p = 0; p < num_bitplanes; p++) {
            streams[p] = static_cast<uint8_t *>(std::malloc(nw * sizeof(S)));
            std::memcpy(streams[p], host.data() + p * stride, nw * sizeof(S));
        }
        return streams;
    }
    // encode with the per-plane level errors (NegaBinaryBPEncoder.hpp:65-125):
    // level_errors[j] = squared error of keeping the first j planes. fp32
    // only, as the reference's collect_level_errors (:231-239).
    std::vector<uint8_t *> encode(T const *data, int32_t n, int32_t exp, uint8_t num_bitplanes,
                                  std::vector<uint32_t> &stream_sizes, std::vector<double> &level_errors) const {
        static_assert(std::is_same<T, float>::value, "the encoder's level errors are defined for float data");
        std::vector<uint8_t *> streams = encode(data, n, exp, num_bitplanes, stream_sizes);
        DeviceContext &c = *ctx_;
        const size_t wsb = prodm_level_errors_workspace_bytes((uint64_t)n, num_bitplanes);
        level_errors.assign(num_bitplanes + 1, 0.0);
        check(prodm_negabinary_level_errors_f32(c.a.as<float>(n), (uint64_t)n, exp, num_bitplanes,
                                                level_errors.data(), c.ws.get(wsb), wsb, c.stream),
              "negabinary level errors");
        return streams;
    }
    T *decode(const std::vector<uint8_t const *> &streams, int32_t n, int exp, uint8_t num_bitplanes) {
        return progressive_decode(streams, n, exp, 0, num_bitplanes, 0);
    }
    T *progressive_decode(const std::vector<uint8_t const *> &streams, int32_t n, int exp, uint8_t start,
                          uint8_t k, int /*level*/) {
        DeviceContext &c = *ctx_;
        const uint64_t nw = words(n), stride = (nw + 1) & ~1ull;
        S *planes = c.b.as<S>(stride * (k ? k : 1));
        std::vector<const S *> ptrs(k);
        for (int i = 0; i < k; i++) {
            to_device(planes + i * stride, reinterpret_cast<const S *>(streams[i]), nw, c.stream);
            ptrs[i] = planes + i * stride;
        }
        T *out = c.a.as<T>(n);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_negabinary_decode_f32(ptrs.data(), (uint32_t)n, exp, start, k, out, c.stream);
        else
            rc = prodm_negabinary_decode_f64(ptrs.data(), (uint32_t)n, exp, start, k, out, c.stream);
        check(rc, "negabinary decode");
        T *h = static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
        to_host(h, out, (uint64_t)n, c.stream);
        return h;
    }
    void print() const { std::printf("negabinary bitplane encoder (MI355X)\n"); }

  private:
    static uint64_t words(int32_t n) { return n > 0 ? ((uint64_t)n - 1) / (8 * sizeof(S)) + 1 : 0; }
    Context ctx_;
};

// ---------------------------------------------------------------- grouped
template <class T, class S>
class GpuGroupedBPEncoder {
  public:
    explicit GpuGroupedBPEncoder(Context ctx = make_context(), bool x86_compat = false)
        : ctx_(std::move(ctx)), st_(std::make_shared<State>()), compat_(x86_compat) {
        detail::check_stream_type<T, S>();
    }
    std::vector<uint8_t *> encode(T const *data, int32_t n, int32_t exp, uint8_t num_bitplanes,
                                  std::vector<uint32_t> &stream_sizes) const {
        DeviceContext &c = *ctx_;
        const uint64_t nblk = blocks(n), stride = 2 * nblk + 64;
        T *x = c.a.as<T>(n);
        S *planes = c.b.as<S>(stride * num_bitplanes);
        uint8_t *rb = c.aux.as<uint8_t>(nblk + 4 * num_bitplanes + 256);
        uint32_t *wcount = reinterpret_cast<uint32_t *>(rb + ((nblk + 255) & ~255ull));
        const size_t wsb = prodm_grouped_workspace_bytes((uint32_t)n, num_bitplanes, dtype_of<T>());
        void *ws = c.ws.get(wsb);
        to_device(x, data, (uint64_t)n, c.stream);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_grouped_encode_f32(x, (uint32_t)n, exp, num_bitplanes, compat_, rb, planes, stride, wcount, ws,
                                          wsb, c.stream);
        else
            rc = prodm_grouped_encode_f64(x, (uint32_t)n, exp, num_bitplanes, compat_, rb, planes, stride, wcount, ws,
                                          wsb, c.stream);
        check(rc, "grouped encode");
        std::vector<uint32_t> w(num_bitplanes);
        std::vector<uint8_t> hrb(nblk);
        to_host(w.data(), wcount, w.size(), c.stream);
        to_host(hrb.data(), rb, nblk, c.stream);
        std::vector<uint8_t *> streams(num_bitplanes);
        stream_sizes.assign(num_bitplanes, 0);
        for (int p = 0; p < num_bitplanes; p++) {
            const uint64_t head = p == 0 ? 4 + nblk : 0, body = (uint64_t)w[p] * sizeof(S);
            uint8_t *buf = static_cast<uint8_t *>(std::malloc(head + body + 1));
            if (head) {
                const uint32_t nb32 = (uint32_t)nblk;
                std::memcpy(buf, &nb32, 4);
                std::memcpy(buf + 4, hrb.data(), nblk);
            }
            to_host(buf + head, reinterpret_cast<const uint8_t *>(planes + p * stride), body, c.stream);
            streams[p] = buf;
            stream_sizes[p] = (uint32_t)(head + body);
        }
        return streams;
    }
    // encode with the per-plane level errors (GroupedBPEncoder.hpp:76-145):
    // level_errors[j] = squared error of keeping the first j planes. fp32
    // only, as the reference's collect_level_errors (:324-335).
    std::vector<uint8_t *> encode(T const *data, int32_t n, int32_t exp, uint8_t num_bitplanes,
                                  std::vector<uint32_t> &stream_sizes, std::vector<double> &level_errors) const {
        static_assert(std::is_same<T, float>::value, "the encoder's level errors are defined for float data");
        std::vector<uint8_t *> streams = encode(data, n, exp, num_bitplanes, stream_sizes);
        DeviceContext &c = *ctx_;
        const size_t wsb = prodm_level_errors_workspace_bytes((uint64_t)n, num_bitplanes);
        level_errors.assign(num_bitplanes + 1, 0.0);
        check(prodm_grouped_level_errors_f32(c.a.as<float>(n), (uint64_t)n, exp, num_bitplanes, level_errors.data(),
                                             c.ws.get(wsb), wsb, c.stream),
              "grouped level errors");
        return streams;
    }
    T *decode(const std::vector<uint8_t const *> &streams, int32_t n, int exp, uint8_t num_bitplanes) {
        GpuGroupedBPEncoder fresh(ctx_, compat_);
        return fresh.progressive_decode(streams, n, exp, 0, num_bitplanes, 0);
    }
    // Per-level state is keyed by `level` (ProDM keys it by call order,
    // GroupedBPEncoder.hpp:223,233; the explicit key also serves levels
    // first decoded with zero planes).
    T *progressive_decode(const std::vector<uint8_t const *> &streams, int32_t n, int exp, uint8_t start,
                          uint8_t k, int level) {
        DeviceContext &c = *ctx_;
        State &s = *st_;
        if (!s.dev) check(prodm_grouped_state_create(&s.dev, 64, dtype_of<T>()), "grouped state");  // levels 0..63
        const uint64_t nblk = blocks(n);
        auto &hist = s.rb_hist[level];
        if (k && start == 0) {  // stream 0 carries u32 nblocks + rb[nblocks]
            uint32_t hdr;
            std::memcpy(&hdr, streams[0], 4);
            if (hdr != nblk) fatal("grouped decode: stream header does not match the level size");
            hist.assign(256, 0);
            for (uint64_t b = 0; b < nblk; b++) hist[streams[0][4 + b]]++;
        }
        if (k && hist.empty()) fatal("grouped decode: plane 0 of this level was never decoded");
        // stream sizes follow from the recording-plane histogram
        std::vector<uint64_t> size(k);
        uint64_t total = 0, below = 0;
        for (int p = 0; p < start + k; p++) {
            below += p < 256 ? hist[p] : 0;  // blocks with rb <= p
            if (p >= start) {
                const uint64_t sz = (below + hist[p]) * sizeof(S) + (p == 0 ? 4 + nblk : 0);
                size[p - start] = sz;
                total += (sz + 255) & ~255ull;
            }
        }
        uint8_t *dev = c.b.as<uint8_t>(total + 256);
        std::vector<const uint8_t *> ptrs(k);
        uint64_t off = 0;
        for (int i = 0; i < k; i++) {
            to_device(dev + off, streams[i], size[i], c.stream);
            ptrs[i] = dev + off;
            off += (size[i] + 255) & ~255ull;
        }
        T *out = c.a.as<T>(n);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_grouped_decode_f32(s.dev, level, ptrs.data(), (uint32_t)n, exp, start, k, out, c.stream);
        else
            rc = prodm_grouped_decode_f64(s.dev, level, ptrs.data(), (uint32_t)n, exp, start, k, out, c.stream);
        check(rc, "grouped decode");
        T *h = static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
        to_host(h, out, (uint64_t)n, c.stream);
        return h;
    }
    void print() const { std::printf("grouped bitplane encoder (MI355X)\n"); }

  private:
    struct State {
        prodm_grouped_state *dev = nullptr;
        std::map<int, std::vector<uint64_t>> rb_hist;
        ~State() {
            if (dev) prodm_grouped_state_destroy(dev);
        }
    };
    static uint64_t blocks(int32_t n) { return n > 0 ? ((uint64_t)n - 1) / (8 * sizeof(S)) + 1 : 0; }
    Context ctx_;
    std::shared_ptr<State> st_;
    int compat_;
};

// ---------------------------------------------------------------- per-bit
// Encode on the GPU; progressive_decode on the host (the streams carry no
// index of where an element's bits begin, so each is parsed from its
// start), with the per-level sign state keyed by `level`.
template <class T, class S>
class GpuPerBitBPEncoder {
  public:
    explicit GpuPerBitBPEncoder(Context ctx = make_context(), bool x86_compat = false)
        : ctx_(std::move(ctx)), st_(std::make_shared<State>()), compat_(x86_compat) {
        detail::check_stream_type<T, S>();
    }
    std::vector<uint8_t *> encode(T const *data, int32_t n, int32_t exp, uint8_t num_bitplanes,
                                  std::vector<uint32_t> &stream_sizes) const {
        DeviceContext &c = *ctx_;
        const uint64_t stride = prodm_perbit_stream_words((uint32_t)n);
        T *x = c.a.as<T>(n);
        uint64_t *planes = c.b.as<uint64_t>(stride * num_bitplanes);
        uint32_t *sizes = c.aux.as<uint32_t>(num_bitplanes);
        const size_t wsb = prodm_perbit_workspace_bytes((uint32_t)n, num_bitplanes, dtype_of<T>());
        void *ws = c.ws.get(wsb);
        to_device(x, data, (uint64_t)n, c.stream);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_perbit_encode_f32(x, (uint32_t)n, exp, num_bitplanes, compat_, planes, stride, sizes, ws, wsb,
                                         c.stream);
        else
            rc = prodm_perbit_encode_f64(x, (uint32_t)n, exp, num_bitplanes, compat_, planes, stride, sizes, ws, wsb,
                                         c.stream);
        check(rc, "perbit encode");
        stream_sizes.assign(num_bitplanes, 0);
        to_host(stream_sizes.data(), sizes, num_bitplanes, c.stream);
        std::vector<uint8_t *> streams(num_bitplanes);
        for (int p = 0; p < num_bitplanes; p++) {
            streams[p] = static_cast<uint8_t *>(std::malloc(stream_sizes[p] + 8));
            to_host(streams[p], reinterpret_cast<const uint8_t *>(planes + p * stride), stream_sizes[p], c.stream);
        }
        return streams;
    }
    // encode with the per-plane level errors (PerBitBPEncoder.hpp:151-226):
    // the per-element terms are GroupedBP's (:387-397), fp32 only.
    std::vector<uint8_t *> encode(T const *data, int32_t n, int32_t exp, uint8_t num_bitplanes,
                                  std::vector<uint32_t> &stream_sizes, std::vector<double> &level_errors) const {
        static_assert(std::is_same<T, float>::value, "the encoder's level errors are defined for float data");
        std::vector<uint8_t *> streams = encode(data, n, exp, num_bitplanes, stream_sizes);
        DeviceContext &c = *ctx_;
        const size_t wsb = prodm_level_errors_workspace_bytes((uint64_t)n, num_bitplanes);
        level_errors.assign(num_bitplanes + 1, 0.0);
        check(prodm_grouped_level_errors_f32(c.a.as<float>(n), (uint64_t)n, exp, num_bitplanes, level_errors.data(),
                                             c.ws.get(wsb), wsb, c.stream),
              "perbit level errors");
        return streams;
    }
    T *decode(const std::vector<uint8_t const *> &streams, int32_t n, int exp, uint8_t num_bitplanes) {
        GpuPerBitBPEncoder fresh(ctx_, compat_);
        return fresh.progressive_decode(streams, n, exp, 0, num_bitplanes, 0);
    }
    T *progressive_decode(const std::vector<uint8_t const *> &streams, int32_t n, int exp, uint8_t start, uint8_t k,
                          int level) {
        State &s = *st_;
        if (!s.h) check(prodm_perbit_state_create(&s.h), "perbit state");
        std::vector<const uint64_t *> ptrs(k);
        for (int i = 0; i < k; i++) ptrs[i] = reinterpret_cast<const uint64_t *>(streams[i]);
        T *h = static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_perbit_decode_f32(s.h, level, ptrs.data(), nullptr, (uint32_t)n, exp, start, k, h);
        else
            rc = prodm_perbit_decode_f64(s.h, level, ptrs.data(), nullptr, (uint32_t)n, exp, start, k, h);
        check(rc, "perbit decode");
        return h;
    }
    void print() const { std::printf("per-bit bitplane encoder (MI355X encode, host decode)\n"); }

  private:
    struct State {
        prodm_perbit_state *h = nullptr;
        ~State() {
            if (h) prodm_perbit_state_destroy(h);
        }
    };
    Context ctx_;
    std::shared_ptr<State> st_;
    int compat_;
};

// ---------------------------------------------------------------- collectors
template <class T>
class MaxErrorCollector {
  public:
    std::vector<double> collect_level_error(T const *, size_t, int num_bitplanes, T max_level_error) const {
        std::vector<double> t(num_bitplanes + 1);
        if constexpr (std::is_same<T, float>::value)
            prodm_collect_max_f32(max_level_error, num_bitplanes, t.data());
        else
            prodm_collect_max_f64(max_level_error, num_bitplanes, t.data());
        return t;
    }
    void print() const { std::printf("max error collector\n"); }
};

// Each entry is the exactly rounded sum of ProDM's per-element terms; ProDM
// sums them sequentially in FP64, which differs in the last ulps.
template <class T>
class GpuSquaredErrorCollector {
  public:
    explicit GpuSquaredErrorCollector(Context ctx = make_context()) : ctx_(std::move(ctx)) {}
    std::vector<double> collect_level_error(T const *data, size_t n, int num_bitplanes, T max_level_error) const {
        DeviceContext &c = *ctx_;
        T *x = c.a.as<T>(n);
        to_device(x, data, (uint64_t)n, c.stream);
        return collect_device(x, n, num_bitplanes, max_level_error);
    }
    std::vector<double> collect_device(const T *d_x, size_t n, int num_bitplanes, T max_level_error) const {
        DeviceContext &c = *ctx_;
        std::vector<double> t(num_bitplanes + 1);
        const size_t wsb = prodm_collect_squared_workspace_bytes(n, num_bitplanes, dtype_of<T>());
        void *ws = c.ws.get(wsb);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_collect_squared_f32(d_x, n, num_bitplanes, max_level_error, t.data(), ws, wsb, c.stream);
        else
            rc = prodm_collect_squared_f64(d_x, n, num_bitplanes, max_level_error, t.data(), ws, wsb, c.stream);
        check(rc, "squared error collector");
        return t;
    }
    void print() const { std::printf("squared error collector (MI355X)\n"); }

  private:
    Context ctx_;
};

// Squared-error tables consistent with the bitplane codec that stores the
// level: entry j = Σ (x − decode_j(x))² with the encoder's own truncation
// after j planes, the level_errors of NegaBinaryBPEncoder::encode
// (NegaBinaryBPEncoder.hpp:65-125) or GroupedBPEncoder::encode
// (GroupedBPEncoder.hpp:76-145; PerBit streams truncate the same way).
// SquaredErrorCollector (SquaredErrorCollector.hpp:23-65) models
// sign-magnitude truncation at the exponent of the level maximum, which for
// negabinary streams (aligned two exponents higher, digits of both signs)
// under-states the error of a plane count by about two planes, so an L2
// retrieval planned on it misses its target; with these tables the planned
// estimate is the coefficients' actual squared error. fp32 data, at most 32
// planes, as the encoders' collectors. Same interface as the collectors.
template <class T>
class GpuEncoderErrorCollector {
    static_assert(std::is_same<T, float>::value, "the encoders' level errors are defined for float data");

  public:
    explicit GpuEncoderErrorCollector(int encoder = PRODM_ENC_NEGABINARY, Context ctx = make_context())
        : encoder_(encoder), ctx_(std::move(ctx)) {}
    std::vector<double> collect_level_error(T const *data, size_t n, int num_bitplanes, T max_level_error) const {
        DeviceContext &c = *ctx_;
        T *x = c.a.as<T>(n);
        to_device(x, data, (uint64_t)n, c.stream);
        return collect_device(x, n, num_bitplanes, max_level_error);
    }
    std::vector<double> collect_device(const T *d_x, size_t n, int num_bitplanes, T max_level_error) const {
        DeviceContext &c = *ctx_;
        int e = 0;
        std::frexp(max_level_error, &e);  // the exponent the encoder is given
        std::vector<double> t(num_bitplanes + 1);
        const size_t wsb = prodm_level_errors_workspace_bytes((uint64_t)n, num_bitplanes);
        void *ws = c.ws.get(wsb);
        const int rc = encoder_ == PRODM_ENC_NEGABINARY
                           ? prodm_negabinary_level_errors_f32(d_x, n, e, num_bitplanes, t.data(), ws, wsb, c.stream)
                           : prodm_grouped_level_errors_f32(d_x, n, e, num_bitplanes, t.data(), ws, wsb, c.stream);
        check(rc, "encoder level errors");
        return t;
    }
    void print() const {
        std::printf("%s level-error collector (MI355X)\n", encoder_ == PRODM_ENC_NEGABINARY ? "negabinary" : "grouped");
    }

  private:
    int encoder_;
    Context ctx_;
};

}  // namespace prodm
#endif

// gpu_mdr.hpp -- device-resident refactor and reconstruct (host C++ over the
// prodm_mdr_plan C ABI).
//
// GpuRefactor runs the GPU stages of ComposedRefactor::refactor
// (Refactor/ComposedRefactor.hpp:78-141: decompose, interleave, max-abs,
// encode) without leaving the device, then copies the plane streams out once
// and hands them to the host lossless stage and the writer. GpuReconstructor
// reads ProDM's files, plans with a size interpreter, decompresses on the
// host, and decodes and recomposes on the device
// (Reconstructor/ComposedReconstructor.hpp:27-258). Both use the same files
// as ComposedRefactor / ComposedReconstructor and ProDM.
//
// The reconstructor always returns the full grid: levels without planes
// contribute zero coefficients (ProDM returns a coarser grid in that case and
// offers recompose_to_full, ComposedReconstructor.hpp:76-90,124-131), and each
// call decodes the held planes of every level in one pass rather than adding
// per-call deltas.
#ifndef PRODM_GPU_MDR_HPP
#define PRODM_GPU_MDR_HPP

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "composed.hpp"
#include "device.hpp"
#include "io.hpp"

namespace prodm {

class Plan {
  public:
    Plan(const std::vector<uint32_t> &dims, int L, int nb, int dtype, int encoder) {
        check(prodm_mdr_plan_create(&p_, dims.data(), (int)dims.size(), L, nb, dtype, encoder), "plan");
    }
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan() {
        if (p_) prodm_mdr_plan_destroy(p_);
    }
    prodm_mdr_plan *get() const { return p_; }

  private:
    prodm_mdr_plan *p_ = nullptr;
};

template <class T, class Compressor, class Writer>
class GpuRefactor {
  public:
    // encoder: PRODM_ENC_NEGABINARY (ProDM's `negabinary = true`) or PRODM_ENC_GROUPED;
    // decomposer: PRODM_DECOMP_ORTHOGONAL or PRODM_DECOMP_HIERARCHICAL.
    GpuRefactor(int encoder, Compressor compressor, Writer writer, bool collect_squared = false,
                int decomposer = PRODM_DECOMP_ORTHOGONAL)
        : encoder_(encoder), compressor_(compressor), writer_(writer), squared_(collect_squared),
          decomposer_(decomposer) {}

    // Host input (copied to the device) ...
    void refactor(T const *data, const std::vector<uint32_t> &dims, uint8_t target_level, uint8_t num_bitplanes) {
        prepare(dims, target_level, num_bitplanes);
        T *d = in_.as<T>(product(dims));
        to_device(d, data, product(dims), stream_);
        refactor_device(d, dims, target_level, num_bitplanes);
    }
    // ... or a device array (dense, dims[0] slowest).
    void refactor_device(const T *d_data, const std::vector<uint32_t> &dims, uint8_t target_level,
                         uint8_t num_bitplanes) {
        prepare(dims, target_level, num_bitplanes);
        check(prodm_mdr_refactor(plan_->get(), d_data, stream_), "refactor");
        const int L = target_level;
        MdrMetadata<T> m;
        m.dims = dims;
        m.negabinary = encoder_ == PRODM_ENC_NEGABINARY;
        std::vector<std::vector<uint8_t *>> comps(L + 1);
        m.level_sizes.assign(L + 1, {});
        for (int l = 0; l <= L; l++) {
            double mx;
            int32_t e;
            uint64_t n;
            check(prodm_mdr_level_info(plan_->get(), l, &mx, &e, &n), "level info");
            const T tmx = (T)mx;
            m.level_error_bounds.push_back(m.negabinary ? tmx * 4 : tmx);
            if (squared_) m.level_squared_errors.push_back(collect_squared(l, n, num_bitplanes, tmx));
            for (int p = 0; p < num_bitplanes; p++) {
                uint64_t bytes = 0;
                check(prodm_mdr_plane_stream(plan_->get(), l, p, nullptr, 0, &bytes, stream_), "plane size");
                uint8_t *buf = static_cast<uint8_t *>(std::malloc(bytes ? bytes : 1));
                check(prodm_mdr_plane_stream(plan_->get(), l, p, buf, bytes, &bytes, stream_), "plane stream");
                comps[l].push_back(buf);
                m.level_sizes[l].push_back((uint32_t)bytes);
            }
            m.stopping_indices.push_back(compressor_.compress_level(comps[l], m.level_sizes[l]));
        }
        m.level_num = writer_.write_level_components(comps, m.level_sizes);
        const auto bytes = m.serialize_bytes();
        writer_.write_metadata(bytes.data(), (uint32_t)bytes.size());
        for (auto &lv : comps)
            for (uint8_t *b : lv) std::free(b);
    }
    prodm_mdr_plan *plan() const { return plan_ ? plan_->get() : nullptr; }

  private:
    void prepare(const std::vector<uint32_t> &dims, int L, int nb) {
        if (plan_ && dims == dims_ && L == L_ && nb == nb_) return;
        plan_.reset();
        plan_ = std::make_shared<Plan>(dims, L, nb, dtype_of<T>(), encoder_);
        check(prodm_mdr_plan_set_decomposer(plan_->get(), decomposer_), "plan decomposer");
        dims_ = dims;
        L_ = L;
        nb_ = nb;
    }
    std::vector<double> collect_squared(int l, uint64_t n, int nb, T mx) {
        void *coef = nullptr;
        check(prodm_mdr_level_coefficients(plan_->get(), l, &coef), "level coefficients");
        const size_t wsb = prodm_collect_squared_workspace_bytes(n, nb, dtype_of<T>());
        std::vector<double> t(nb + 1);
        int rc;
        if constexpr (std::is_same<T, float>::value)
            rc = prodm_collect_squared_f32(static_cast<const float *>(coef), n, nb, mx, t.data(), ws_.get(wsb), wsb, stream_);
        else
            rc = prodm_collect_squared_f64(static_cast<const double *>(coef), n, nb, mx, t.data(), ws_.get(wsb), wsb,
                                           stream_);
        check(rc, "squared error collector");
        return t;
    }
    int encoder_;
    Compressor compressor_;
    Writer writer_;
    bool squared_;
    int decomposer_;
    hipStream_t stream_ = nullptr;
    std::shared_ptr<Plan> plan_;
    std::vector<uint32_t> dims_;
    int L_ = -1, nb_ = -1;
    DeviceArena in_, ws_;
};

template <class T, class Compressor, class SizeInterpreter, class ErrorEstimator, class Retriever>
class GpuReconstructor {
  public:
    // decomposer: the one the files were refactored with (ProDM's metadata
    // does not record it)
    GpuReconstructor(Compressor compressor, SizeInterpreter interpreter, Retriever retriever,
                     int decomposer = PRODM_DECOMP_ORTHOGONAL)
        : compressor_(compressor), interpreter_(interpreter), retriever_(retriever), decomposer_(decomposer) {}

    void load_metadata() {
        size_t size = 0;
        uint8_t *raw = load_bytes(retriever_, &size);
        meta_.parse(raw, size);
        std::free(raw);
        const int L = (int)meta_.level_error_bounds.size() - 1;
        const int nb = (int)meta_.level_sizes[0].size();
        for (const auto &s : meta_.level_sizes)
            if ((int)s.size() != nb) fatal("reconstruct: levels with different plane counts");
        plan_ = std::make_shared<Plan>(meta_.dims, L, nb, dtype_of<T>(),
                                       meta_.negabinary ? PRODM_ENC_NEGABINARY : PRODM_ENC_GROUPED);
        check(prodm_mdr_plan_set_decomposer(plan_->get(), decomposer_), "plan decomposer");
        for (int l = 0; l <= L; l++) {
            const double b = meta_.negabinary ? (double)(meta_.level_error_bounds[l] / 4) : (double)meta_.level_error_bounds[l];
            check(prodm_mdr_set_level_max(plan_->get(), l, b, stream_), "set level max");
        }
        held_.assign(L + 1, 0);
        host_.assign(product(meta_.dims), 0);
        d_out_ = out_.as<T>(product(meta_.dims));
    }

    T *reconstruct(double tolerance) {
        const int L = (int)meta_.level_error_bounds.size() - 1;
        std::vector<std::vector<double>> errs;
        if constexpr (std::is_base_of<MaxErrorEstimator<T>, ErrorEstimator>::value) {
            for (int l = 0; l <= L; l++) {
                std::vector<double> t(meta_.level_sizes[l].size() + 1);
                if constexpr (std::is_same<T, float>::value)
                    prodm_collect_max_f32(meta_.level_error_bounds[l], (int)meta_.level_sizes[l].size(), t.data());
                else
                    prodm_collect_max_f64(meta_.level_error_bounds[l], (int)meta_.level_sizes[l].size(), t.data());
                errs.push_back(t);
            }
        } else {
            static_assert(std::is_base_of<SquaredErrorEstimator<T>, ErrorEstimator>::value,
                          "ErrorEstimator must derive from MaxErrorEstimator<T> or SquaredErrorEstimator<T>");
            if (meta_.level_squared_errors.size() != (size_t)L + 1) {
                std::fprintf(stderr, "prodm: metadata holds no squared-error tables\n");
                return nullptr;
            }
            errs = meta_.level_squared_errors;
        }
        const std::vector<uint8_t> prev(held_);
        auto rs = interpreter_.interpret_retrieve_size(meta_.level_sizes, errs, tolerance, held_);
        auto comps = retriever_.retrieve_level_components(meta_.level_sizes, rs, prev, held_);
        for (int l = 0; l <= L; l++) {
            const int k = held_[l] - prev[l];
            if (k <= 0) continue;
            compressor_.decompress_level(comps[l], meta_.level_sizes[l], prev[l], (uint8_t)k, meta_.stopping_indices[l]);
            for (int i = 0; i < k; i++) {
                const int p = prev[l] + i;
                const uint64_t n = raw_size(l, p, comps[l][i]);
                check(prodm_mdr_load_plane_stream(plan_->get(), l, p, comps[l][i], n, stream_), "load plane");
            }
            compressor_.decompress_release();
        }
        retriever_.release();
        check(prodm_mdr_retrieve(plan_->get(), held_.data(), d_out_, stream_), "retrieve");
        to_host(host_.data(), d_out_, host_.size(), stream_);
        return host_.data();
    }
    T *progressive_reconstruct(double tolerance) { return reconstruct(tolerance); }
    const T *device_data() const { return d_out_; }
    const std::vector<uint32_t> &get_dimensions() const { return meta_.dims; }
    const std::vector<uint8_t> &get_level_num_bitplanes() const { return held_; }
    size_t get_retrieved_size() const { return retriever_.get_retrieved_size(); }

  private:
    // Byte size of a decompressed plane: fixed for negabinary; GroupedBP
    // plane sizes follow from the recording planes of stream 0.
    uint64_t raw_size(int l, int p, const uint8_t *stream) {
        const int B = sizeof(T) * 8;
        uint64_t n = 0;
        {
            std::vector<uint32_t> flat((meta_.level_error_bounds.size()) * meta_.dims.size());
            std::vector<uint64_t> el(meta_.level_error_bounds.size());
            prodm_level_geometry(meta_.dims.data(), (int)meta_.dims.size(), (int)el.size() - 1, flat.data(), el.data());
            n = el[l];
        }
        const uint64_t nblk = (n + B - 1) / B;
        if (meta_.negabinary) return nblk * sizeof(T);
        auto &h = hist_[l];
        if (p == 0) {
            h.assign(256, 0);
            for (uint64_t b = 0; b < nblk; b++) h[stream[4 + b]]++;
        }
        uint64_t below = 0;
        for (int q = 0; q <= p; q++) below += h[q];
        return (below + h[p]) * sizeof(T) + (p == 0 ? 4 + nblk : 0);
    }
    Compressor compressor_;
    SizeInterpreter interpreter_;
    Retriever retriever_;
    int decomposer_;
    MdrMetadata<T> meta_;
    std::shared_ptr<Plan> plan_;
    std::vector<uint8_t> held_;
    std::vector<T> host_;
    std::map<int, std::vector<uint64_t>> hist_;
    hipStream_t stream_ = nullptr;
    DeviceArena out_;
    T *d_out_ = nullptr;
};

}  // namespace prodm
#endif

// gpu_mdr.hpp -- device-resident refactor and reconstruct (host C++ over the
// prodm_mdr_plan C ABI).
//
// GpuRefactor runs the GPU stages of ComposedRefactor::refactor
// (Refactor/ComposedRefactor.hpp:78-141: decompose, interleave, max-abs,
// encode) without leaving the device, then copies the plane streams out once
// (overlapped with the host lossless stage, level by level) and hands them
// to the compressor and the writer. GpuReconstructor
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

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "composed.hpp"
#include "compressors.hpp"  // parallel_for
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

// Squared-error tables a refactor stores for L2 retrieval: none, those of
// SquaredErrorCollector (ErrorCollector/SquaredErrorCollector.hpp:23-65,
// GpuSquaredErrorCollector), or those of the encoder's own truncation
// (GpuEncoderErrorCollector: fp32, the tables an L2 retrieval of negabinary
// streams must be planned on to meet its target).
enum SquaredTables { kNoSquaredTables = 0, kCollectorTables = 1, kEncoderTables = 2 };

// Grow-only pinned host buffer (hipHostMalloc): copies to and from it run at
// the link's rate instead of through the driver's bounce buffers.
struct PinnedHostBuffer {
    uint8_t *p = nullptr;
    uint64_t cap = 0;
    uint8_t *get(uint64_t n) {
        if (n > cap) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            hip_check(hipHostMalloc((void **)&p, n ? n : 1, hipHostMallocDefault), "hipHostMalloc");
            cap = n;
        }
        return p;
    }
    PinnedHostBuffer() = default;
    PinnedHostBuffer(const PinnedHostBuffer &) = delete;
    PinnedHostBuffer &operator=(const PinnedHostBuffer &) = delete;
    ~PinnedHostBuffer() {
        if (p) (void)hipHostFree(p);
    }
};

template <class T, class Compressor, class Writer>
class GpuRefactor {
  public:
    // encoder: PRODM_ENC_NEGABINARY (ProDM's `negabinary = true`), PRODM_ENC_GROUPED
    // or PRODM_ENC_PERBIT; squared_tables: a SquaredTables value (true =
    // kCollectorTables); decomposer: PRODM_DECOMP_ORTHOGONAL or
    // PRODM_DECOMP_HIERARCHICAL.
    GpuRefactor(int encoder, Compressor compressor, Writer writer, int squared_tables = kNoSquaredTables,
                int decomposer = PRODM_DECOMP_ORTHOGONAL)
        : encoder_(encoder), compressor_(compressor), writer_(writer), squared_(squared_tables),
          decomposer_(decomposer) {
        if (squared_ == kEncoderTables && !std::is_same<T, float>::value)
            fatal("GpuRefactor: the encoder's squared-error tables are defined for float data");
    }

    // Host input (copied to the device) ...
    void refactor(T const *data, const std::vector<uint32_t> &dims, uint8_t target_level, uint8_t num_bitplanes) {
        const auto t0 = std::chrono::steady_clock::now();
        prepare(dims, target_level, num_bitplanes);
        T *d = in_.as<T>(product(dims));
        to_device(d, data, product(dims), stream_);
        hip_check(hipStreamSynchronize(stream_), "upload");
        const double up = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        refactor_device(d, dims, target_level, num_bitplanes);
        times_.device += up;
    }
    // ... or a device array (dense, dims[0] slowest).
    //
    // The plane streams of all levels are copied to pinned host memory in one
    // pass of asynchronous copies (a strided copy per level), coarsest level
    // first, each level followed by an event; level l is compressed on host
    // threads as soon as its event completes, while the copies of the finer
    // levels continue.
    void refactor_device(const T *d_data, const std::vector<uint32_t> &dims, uint8_t target_level,
                         uint8_t num_bitplanes) {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        prepare(dims, target_level, num_bitplanes);
        check(prodm_mdr_refactor(plan_->get(), d_data, stream_), "refactor");
        const int L = target_level, nb = num_bitplanes;
        const bool grouped = encoder_ == PRODM_ENC_GROUPED;
        MdrMetadata<T> m;
        m.dims = dims;
        m.negabinary = encoder_ == PRODM_ENC_NEGABINARY;
        std::vector<std::vector<uint8_t *>> comps(L + 1);
        m.level_sizes.assign(L + 1, {});
        // plane layout (level info synchronizes: the device stages are done)
        std::vector<std::vector<const uint8_t *>> src(L + 1);
        std::vector<uint64_t> head(L + 1, 0), off(L + 2, 0);
        std::vector<const uint8_t *> rb(L + 1, nullptr);
        for (int l = 0; l <= L; l++) {
            double mx;
            int32_t e;
            uint64_t n;
            check(prodm_mdr_level_info(plan_->get(), l, &mx, &e, &n), "level info");
            const T tmx = (T)mx;
            m.level_error_bounds.push_back(m.negabinary ? tmx * 4 : tmx);
            if (squared_ == kCollectorTables) m.level_squared_errors.push_back(collect_squared(l, n, nb, tmx));
            if (squared_ == kEncoderTables) {  // the plan encoder's table (PerBit: GroupedBP's, same terms)
                std::vector<double> t(nb + 1);
                check(prodm_mdr_level_errors(plan_->get(), l, t.data(), stream_), "level errors");
                m.level_squared_errors.push_back(t);
            }
            if (grouped) {  // stream 0 carries u32 nblocks + u8 rb[nblocks] (GroupedBPEncoder.hpp:64-71)
                uint64_t nblk = 0;
                check(prodm_mdr_level_rb(plan_->get(), l, &rb[l], &nblk), "level rb");
                head[l] = 4 + nblk;
            }
            uint64_t sum = head[l];
            for (int p = 0; p < nb; p++) {
                const void *d = nullptr;
                uint64_t bytes = 0;
                check(prodm_mdr_plane(plan_->get(), l, p, &d, &bytes), "plane");
                src[l].push_back(static_cast<const uint8_t *>(d));
                m.level_sizes[l].push_back((uint32_t)(bytes + (p == 0 ? head[l] : 0)));
                sum += bytes;
            }
            off[l + 1] = off[l] + ((sum + 255) & ~255ull);
        }
        const auto t1 = clock::now();
        uint8_t *h = pinned_->get(off[L + 1]);
        events_->ensure(L + 1);
        for (int l = 0; l <= L; l++) {
            uint8_t *dst = h + off[l];
            if (encoder_ != PRODM_ENC_NEGABINARY) {  // planes of their own sizes (GroupedBP, PerBit)
                if (grouped) {
                    const uint32_t nblk = (uint32_t)(head[l] - 4);
                    std::memcpy(dst, &nblk, 4);
                    hip_check(hipMemcpyAsync(dst + 4, rb[l], nblk, hipMemcpyDeviceToHost, stream_), "rb copy");
                    dst += head[l];
                }
                for (int p = 0; p < nb; p++) {
                    const uint64_t bytes = m.level_sizes[l][p] - (p == 0 ? head[l] : 0);
                    if (bytes) hip_check(hipMemcpyAsync(dst, src[l][p], bytes, hipMemcpyDeviceToHost, stream_), "plane copy");
                    dst += bytes;
                }
            } else if (m.level_sizes[l][0]) {  // equal planes at a fixed stride: one strided copy
                const uint64_t w = m.level_sizes[l][0];
                const uint64_t pitch = nb > 1 ? (uint64_t)(src[l][1] - src[l][0]) : w;
                hip_check(hipMemcpy2DAsync(dst, w, src[l][0], pitch, w, nb, hipMemcpyDeviceToHost, stream_),
                          "plane copy");
            }
            hip_check(hipEventRecord(events_->at(l), stream_), "event record");
        }
        for (int l = 0; l <= L; l++) {
            hip_check(hipEventSynchronize(events_->at(l)), "event wait");
            comps[l].assign(nb, nullptr);
            std::vector<uint64_t> at(nb, off[l]);
            for (int p = 1; p < nb; p++) at[p] = at[p - 1] + m.level_sizes[l][p - 1];
            parallel_for(nb, default_threads(), [&](int p) {
                const uint32_t bytes = m.level_sizes[l][p];
                comps[l][p] = static_cast<uint8_t *>(std::malloc(bytes ? bytes : 1));
                std::memcpy(comps[l][p], h + at[p], bytes);
            });
            m.stopping_indices.push_back(compressor_.compress_level(comps[l], m.level_sizes[l]));
        }
        const auto t2 = clock::now();
        m.level_num = writer_.write_level_components(comps, m.level_sizes);
        const auto bytes = m.serialize_bytes();
        writer_.write_metadata(bytes.data(), (uint32_t)bytes.size());
        for (auto &lv : comps)
            for (uint8_t *b : lv) std::free(b);
        const auto t3 = clock::now();
        times_ = {std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
                  std::chrono::duration<double>(t3 - t2).count()};
    }
    prodm_mdr_plan *plan() const { return plan_ ? plan_->get() : nullptr; }
    // seconds of the last refactor_device: device stages (to the level
    // results), copy-out + lossless stage, file writes
    struct Times {
        double device = 0, host = 0, write = 0;
    };
    Times times() const { return times_; }

  private:
    using Pinned = PinnedHostBuffer;
    // per-level events
    struct Events {
        std::vector<hipEvent_t> ev;
        void ensure(size_t n) {
            while (ev.size() < n) {
                hipEvent_t e;
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
                ev.push_back(e);
            }
        }
        hipEvent_t at(size_t i) const { return ev[i]; }
        ~Events() {
            for (auto e : ev) (void)hipEventDestroy(e);
        }
    };
    std::shared_ptr<Pinned> pinned_ = std::make_shared<Pinned>();
    std::shared_ptr<Events> events_ = std::make_shared<Events>();
    Times times_;

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
    int squared_;
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
    // decomposer: the one the files were refactored with; encoder: the
    // bitplane encoder, or -1 for the metadata's flag (negabinary, else
    // grouped). ProDM's metadata records neither the decomposer nor which
    // non-negabinary encoder wrote the streams: PerBit files need
    // PRODM_ENC_PERBIT here.
    GpuReconstructor(Compressor compressor, SizeInterpreter interpreter, Retriever retriever,
                     int decomposer = PRODM_DECOMP_ORTHOGONAL, int encoder = -1)
        : compressor_(compressor), interpreter_(interpreter), retriever_(retriever), decomposer_(decomposer),
          encoder_(encoder) {}

    void load_metadata() {
        size_t size = 0;
        uint8_t *raw = load_bytes(retriever_, &size);
        meta_.parse(raw, size);
        std::free(raw);
        const int L = (int)meta_.level_error_bounds.size() - 1;
        const int nb = (int)meta_.level_sizes[0].size();
        for (const auto &s : meta_.level_sizes)
            if ((int)s.size() != nb) fatal("reconstruct: levels with different plane counts");
        const int enc = plan_encoder();
        if ((enc == PRODM_ENC_NEGABINARY) != (bool)meta_.negabinary) fatal("reconstruct: encoder does not match the metadata");
        plan_ = std::make_shared<Plan>(meta_.dims, L, nb, dtype_of<T>(), enc);
        check(prodm_mdr_plan_set_decomposer(plan_->get(), decomposer_), "plan decomposer");
        for (int l = 0; l <= L; l++) {
            const double b = meta_.negabinary ? (double)(meta_.level_error_bounds[l] / 4) : (double)meta_.level_error_bounds[l];
            check(prodm_mdr_set_level_max(plan_->get(), l, b, stream_), "set level max");
        }
        held_.assign(L + 1, 0);
        host_n_ = product(meta_.dims);
        host_ = reinterpret_cast<T *>(host_buf_->get(host_n_ * sizeof(T)));
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
                const uint64_t n = plan_encoder() == PRODM_ENC_PERBIT ? compressor_.decompressed_sizes()[i]
                                                                      : raw_size(l, p, comps[l][i]);
                check(prodm_mdr_load_plane_stream(plan_->get(), l, p, comps[l][i], n, stream_), "load plane");
            }
            compressor_.decompress_release();
        }
        retriever_.release();
        check(prodm_mdr_retrieve(plan_->get(), held_.data(), d_out_, stream_), "retrieve");
        // the result into pinned memory (540 MB at 513^3)
        to_host(host_, d_out_, host_n_, stream_);
        return host_;
    }
    T *progressive_reconstruct(double tolerance) { return reconstruct(tolerance); }
    const T *device_data() const { return d_out_; }
    const std::vector<uint32_t> &get_dimensions() const { return meta_.dims; }
    const std::vector<uint8_t> &get_level_num_bitplanes() const { return held_; }
    size_t get_retrieved_size() const { return retriever_.get_retrieved_size(); }

  private:
    int plan_encoder() const {
        return encoder_ >= 0 ? encoder_ : meta_.negabinary ? PRODM_ENC_NEGABINARY : PRODM_ENC_GROUPED;
    }
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
    int encoder_;
    MdrMetadata<T> meta_;
    std::shared_ptr<Plan> plan_;
    std::vector<uint8_t> held_;
    std::shared_ptr<PinnedHostBuffer> host_buf_ = std::make_shared<PinnedHostBuffer>();
    T *host_ = nullptr;
    uint64_t host_n_ = 0;
    std::map<int, std::vector<uint64_t>> hist_;
    hipStream_t stream_ = nullptr;
    DeviceArena out_;
    T *d_out_ = nullptr;
};

}  // namespace prodm
#endif

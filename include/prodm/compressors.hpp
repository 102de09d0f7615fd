// compressors.hpp -- host lossless stage: per-plane ZSTD with ProDM's
// stream format and level-compressor interface (host C++, libzstd).
//
//   ZSTD stream            [size_t original_len][zstd frame, level 3]
//                          (LosslessCompressor/ZSTD.hpp:8-28)
//   DefaultLevelCompressor every plane compressed
//                          (LosslessCompressor/DefaultLevelCompressor.hpp:9-46)
//   AdaptiveLevelCompressor planes 0..stop compressed, stop = first plane
//                          i >= 1 whose ratio is below 1.05 (that plane stays
//                          compressed), planes from max(latter, stop + 1) on
//                          compressed, the rest stored raw
//                          (LosslessCompressor/AdaptiveLevelCompressor.hpp:9-65)
//
// Planes are compressed in parallel on host threads. Each plane's bytes are
// what the sequential ProDM compressor writes (ZSTD is deterministic per
// input and level), and the adaptive compressor makes ProDM's sequential
// decision. The destination capacity is
// ZSTD_compressBound (ProDM sizes it by 1.2x the input, ZSTD.hpp:12-16).
#ifndef PRODM_COMPRESSORS_HPP
#define PRODM_COMPRESSORS_HPP

#include <zstd.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace prodm {

namespace zstd {
constexpr int kLevel = 3;

// Returns the stream size; *out is malloc'd.
inline uint32_t compress(const uint8_t *data, uint32_t n, uint8_t **out) {
    const size_t cap = ZSTD_compressBound(n);
    *out = static_cast<uint8_t *>(std::malloc(sizeof(size_t) + cap));
    const size_t len = n;
    std::memcpy(*out, &len, sizeof(size_t));
    const size_t r = ZSTD_compress(*out + sizeof(size_t), cap, data, n, kLevel);
    if (ZSTD_isError(r)) {
        std::fprintf(stderr, "prodm: ZSTD_compress: %s\n", ZSTD_getErrorName(r));
        std::exit(-1);
    }
    return (uint32_t)(r + sizeof(size_t));
}

inline uint32_t decompress(const uint8_t *in, uint32_t n, uint8_t **out) {
    size_t len;
    std::memcpy(&len, in, sizeof(size_t));
    *out = static_cast<uint8_t *>(std::malloc(len ? len : 1));
    const size_t r = ZSTD_decompress(*out, len, in + sizeof(size_t), n - sizeof(size_t));
    if (ZSTD_isError(r) || r != len) {
        std::fprintf(stderr, "prodm: ZSTD_decompress: %s\n", ZSTD_isError(r) ? ZSTD_getErrorName(r) : "short frame");
        std::exit(-1);
    }
    return (uint32_t)len;
}
}  // namespace zstd

// Runs f(i) for i in [0, n) on up to `threads` host threads.
template <class F>
void parallel_for(int n, int threads, F f) {
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&] {
            for (int i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto &th : pool) th.join();
}

inline int default_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::min(16u, h ? h : 1u);
}

class DefaultLevelCompressor {
  public:
    explicit DefaultLevelCompressor(int threads = default_threads()) : threads_(threads) {}
    DefaultLevelCompressor(const DefaultLevelCompressor &o) : threads_(o.threads_) {}
    ~DefaultLevelCompressor() { decompress_release(); }
    uint8_t compress_level(std::vector<uint8_t *> &streams, std::vector<uint32_t> &sizes) const {
        parallel_for((int)streams.size(), threads_, [&](int i) {
            uint8_t *c = nullptr;
            const uint32_t n = zstd::compress(streams[i], sizes[i], &c);
            std::free(streams[i]);
            streams[i] = c;
            sizes[i] = n;
        });
        return 0;
    }
    void decompress_level(std::vector<const uint8_t *> &streams, const std::vector<uint32_t> &sizes,
                          uint8_t start, uint8_t k, uint8_t /*stopping_index*/) {
        std::vector<uint8_t *> out(k, nullptr);
        raw_.assign(k, 0);
        parallel_for(k, threads_, [&](int i) { raw_[i] = zstd::decompress(streams[i], sizes[start + i], &out[i]); });
        for (int i = 0; i < k; i++) {
            buffer_.push_back(out[i]);
            streams[i] = out[i];
        }
    }
    void decompress_release() {
        for (uint8_t *b : buffer_) std::free(b);
        buffer_.clear();
    }
    void print() const { std::printf("default level compressor (ZSTD, %d threads)\n", threads_); }
    // byte sizes of the streams the last decompress_level returned
    const std::vector<uint32_t> &decompressed_sizes() const { return raw_; }

  private:
    int threads_;
    std::vector<uint8_t *> buffer_;
    std::vector<uint32_t> raw_;
};

class AdaptiveLevelCompressor {
  public:
    explicit AdaptiveLevelCompressor(int latter = 26, int threads = default_threads())
        : latter_(latter), threads_(threads) {}
    AdaptiveLevelCompressor(const AdaptiveLevelCompressor &o) : latter_(o.latter_), threads_(o.threads_) {}
    ~AdaptiveLevelCompressor() { decompress_release(); }
    // ProDM compresses planes in order until the first plane i >= 1 whose
    // ratio is below 1.05 (the stopping index), then the planes from
    // max(latter, stop + 1) on. So every plane at or above `latter` ends up
    // compressed whatever the stop, and a plane below it only when it is at
    // or before the stop. Host threads take planes in index order; a plane
    // below `latter` that lies past a stop already found is skipped, so the
    // only compression discarded is that of planes already in flight when
    // the stop is found (fewer than the thread count). The stop is the first
    // low-ratio plane: every plane before it is compressed.
    uint8_t compress_level(std::vector<uint8_t *> &streams, std::vector<uint32_t> &sizes) const {
        const int nb = (int)streams.size();
        std::vector<uint8_t *> c(nb, nullptr);
        std::vector<uint32_t> cs(nb, 0);
        std::vector<char> low(nb, 0);
        std::atomic<int> stop_seen(nb);
        parallel_for(nb, threads_, [&](int i) {
            if (i < latter_ && i > stop_seen.load()) return;
            cs[i] = zstd::compress(streams[i], sizes[i], &c[i]);
            const float ratio = (float)(sizes[i] * 1.0 / cs[i]);  // a float, compared in double
            if (i >= 1 && ratio < 1.05) {
                low[i] = 1;
                int cur = stop_seen.load();
                while (i < cur && !stop_seen.compare_exchange_weak(cur, i)) {
                }
            }
        });
        int stop = nb;
        for (int i = 1; i < nb; i++)
            if (low[i]) {
                stop = i;
                break;
            }
        const int latter_start = stop < latter_ ? latter_ : stop + 1;
        for (int i = 0; i < nb; i++) {
            if (i <= stop || i >= latter_start) {
                std::free(streams[i]);
                streams[i] = c[i];
                sizes[i] = cs[i];
            } else {
                std::free(c[i]);  // in flight past the stop (or never compressed: null)
            }
        }
        return (uint8_t)stop;
    }
    void decompress_level(std::vector<const uint8_t *> &streams, const std::vector<uint32_t> &sizes,
                          uint8_t start, uint8_t k, uint8_t stopping_index) {
        std::vector<uint8_t *> out(k, nullptr);
        raw_.assign(k, 0);
        parallel_for(k, threads_, [&](int i) {
            const int p = start + i;
            raw_[i] = (p <= stopping_index || p >= latter_) ? zstd::decompress(streams[i], sizes[p], &out[i]) : sizes[p];
        });
        for (int i = 0; i < k; i++)
            if (out[i]) {
                buffer_.push_back(out[i]);
                streams[i] = out[i];
            }
    }
    void decompress_release() {
        for (uint8_t *b : buffer_) std::free(b);
        buffer_.clear();
    }
    void print() const { std::printf("adaptive level compressor (ZSTD, latter %d, %d threads)\n", latter_, threads_); }
    // byte sizes of the streams the last decompress_level returned
    const std::vector<uint32_t> &decompressed_sizes() const { return raw_; }

  private:
    int latter_, threads_;
    std::vector<uint8_t *> buffer_;
    std::vector<uint32_t> raw_;
};

}  // namespace prodm
#endif

// host_capi.cpp -- C entry points of libprodm_host.so: the host lossless
// stage of include/prodm/compressors.hpp (per-plane ZSTD), for callers that
// bind C rather than C++ (counterpart of ProDM's
// LosslessCompressor/{Default,Adaptive}LevelCompressor.hpp).
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "prodm/compressors.hpp"

namespace {
template <class C>
int compress(C &c, int nb, const uint8_t *in, const uint32_t *sizes, uint8_t *out, uint64_t cap, uint32_t *out_sizes) {
    std::vector<uint8_t *> streams(nb);
    std::vector<uint32_t> sz(sizes, sizes + nb);
    for (int i = 0; i < nb; i++) {
        streams[i] = static_cast<uint8_t *>(std::malloc(sizes[i] ? sizes[i] : 1));
        std::memcpy(streams[i], in, sizes[i]);
        in += sizes[i];
    }
    const int stop = c.compress_level(streams, sz);
    uint64_t off = 0;
    int rc = stop;
    for (int i = 0; i < nb; i++) {
        out_sizes[i] = sz[i];
        if (off + sz[i] > cap) rc = -1;
        else std::memcpy(out + off, streams[i], sz[i]);
        off += sz[i];
        std::free(streams[i]);
    }
    return rc;
}

// kind 0 decompresses every plane; kind 1 the planes <= stop or >= latter.
template <class C>
int decompress(C &c, int kind, int latter, int nb, int stop, const uint8_t *in, const uint32_t *sizes, int start,
               int k, uint8_t *out, uint64_t cap, uint32_t *out_sizes) {
    std::vector<const uint8_t *> streams;
    std::vector<uint64_t> raw(k);
    const uint8_t *p = in;
    for (int i = 0; i < start; i++) p += sizes[i];
    for (int i = 0; i < k; i++) {
        const int q = start + i;
        const bool packed = kind == 0 || q <= stop || q >= latter;
        size_t len = sizes[q];
        if (packed) std::memcpy(&len, p, sizeof(size_t));  // [size_t original_len][frame]
        raw[i] = len;
        streams.push_back(p);
        p += sizes[q];
    }
    c.decompress_level(streams, std::vector<uint32_t>(sizes, sizes + nb), (uint8_t)start, (uint8_t)k, (uint8_t)stop);
    uint64_t off = 0;
    int rc = 0;
    for (int i = 0; i < k; i++) {
        if (off + raw[i] > cap) {
            rc = -1;
            break;
        }
        std::memcpy(out + off, streams[i], raw[i]);
        out_sizes[i] = (uint32_t)raw[i];
        off += raw[i];
    }
    c.decompress_release();
    return rc;
}
}  // namespace

extern "C" {

/* kind 0: DefaultLevelCompressor, 1: AdaptiveLevelCompressor(latter).
 * Streams back to back in `in` (sizes[nb]); compressed streams are written
 * back to back into out (capacity cap), sizes into out_sizes. Returns the
 * stopping index, or -1 when out is too small.                             */
int prodm_host_compress_level(int kind, int latter, int threads, int nb, const uint8_t *in, const uint32_t *sizes,
                              uint8_t *out, uint64_t cap, uint32_t *out_sizes) {
    if (kind == 0) {
        prodm::DefaultLevelCompressor c(threads);
        return compress(c, nb, in, sizes, out, cap, out_sizes);
    }
    prodm::AdaptiveLevelCompressor c(latter, threads);
    return compress(c, nb, in, sizes, out, cap, out_sizes);
}

/* Decompresses planes [start, start+k) of a level (in: all stored streams
 * back to back, sizes[nb] their stored sizes) back to back into out;
 * raw_sizes[i] receives the size of plane start+i. Returns 0, or -1 when
 * out is too small.                                                        */
int prodm_host_decompress_level(int kind, int latter, int threads, int nb, int stopping_index, const uint8_t *in,
                                const uint32_t *sizes, int start, int k, uint8_t *out, uint64_t cap,
                                uint32_t *raw_sizes) {
    if (kind == 0) {
        prodm::DefaultLevelCompressor c(threads);
        return decompress(c, 0, latter, nb, stopping_index, in, sizes, start, k, out, cap, raw_sizes);
    }
    prodm::AdaptiveLevelCompressor c(latter, threads);
    return decompress(c, 1, latter, nb, stopping_index, in, sizes, start, k, out, cap, raw_sizes);
}

}  // extern "C"

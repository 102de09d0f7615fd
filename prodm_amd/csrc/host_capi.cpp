// host_capi.cpp -- C entry points of libprodm_host.so (include/prodm_host.h):
// the host lossless stage of include/prodm/compressors.hpp (per-plane ZSTD;
// ProDM's LosslessCompressor/{Default,Adaptive}LevelCompressor.hpp) and
// ProDM's file layout through include/prodm/io.hpp (metadata.bin +
// level_<l>.bin; ComposedRefactor.hpp:48-67, ComposedReconstructor.hpp:133-154,
// Writer/FileWriter.hpp, Retriever/FileRetriever.hpp), for callers that bind
// C rather than C++.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "prodm/compressors.hpp"
#include "prodm/io.hpp"
#include "prodm_host.h"

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

namespace {
template <class T>
int write_mdr(const char *metadata_path, const char *const *level_paths, int ndims, const uint32_t *dims, int nlevels,
              int nb, const double *bounds, int negabinary, int latter, int threads, const uint8_t *streams,
              const uint32_t *sizes) {
    prodm::MdrMetadata<T> m;
    m.dims.assign(dims, dims + ndims);
    m.negabinary = negabinary != 0;
    std::vector<std::vector<uint8_t *>> comps(nlevels);
    m.level_sizes.assign(nlevels, {});
    const uint8_t *p = streams;
    prodm::AdaptiveLevelCompressor c(latter, threads);
    for (int l = 0; l < nlevels; l++) {
        m.level_error_bounds.push_back((T)bounds[l]);
        for (int q = 0; q < nb; q++) {
            const uint32_t n = sizes[l * nb + q];
            uint8_t *b = static_cast<uint8_t *>(std::malloc(n ? n : 1));
            std::memcpy(b, p, n);
            p += n;
            comps[l].push_back(b);
            m.level_sizes[l].push_back(n);
        }
        m.stopping_indices.push_back(c.compress_level(comps[l], m.level_sizes[l]));
    }
    prodm::ConcatLevelFileWriter w(metadata_path, std::vector<std::string>(level_paths, level_paths + nlevels));
    m.level_num = w.write_level_components(comps, m.level_sizes);
    const auto bytes = m.serialize_bytes();
    w.write_metadata(bytes.data(), (uint32_t)bytes.size());
    for (auto &lv : comps)
        for (uint8_t *b : lv) std::free(b);
    return 0;
}

template <class T>
int read_mdr(const char *metadata_path, const char *const *level_paths, int nfiles, int latter, int threads,
             const uint8_t *k1, uint32_t *dims, int *ndims, int *nlevels, double *bounds, int *nplanes,
             uint32_t *level_sizes, uint8_t *stopping, uint32_t *level_num, int *negabinary, uint8_t *out, uint64_t cap,
             uint64_t *raw_sizes, uint64_t *retrieved) {
    prodm::ConcatLevelFileRetriever r(metadata_path, std::vector<std::string>(level_paths, level_paths + nfiles));
    size_t size = 0;
    uint8_t *raw = r.load_metadata(&size);
    prodm::MdrMetadata<T> m;
    m.parse(raw, size);
    std::free(raw);
    const int L1 = (int)m.level_error_bounds.size();
    if (L1 != nfiles || m.dims.size() > 3) return -1;
    const int nb = L1 ? (int)m.level_sizes[0].size() : 0;
    *ndims = (int)m.dims.size();
    *nlevels = L1;
    *nplanes = nb;
    *negabinary = m.negabinary;
    for (size_t d = 0; d < m.dims.size(); d++) dims[d] = m.dims[d];
    for (int l = 0; l < L1; l++) {
        if ((int)m.level_sizes[l].size() != nb) return -2;
        bounds[l] = (double)m.level_error_bounds[l];
        stopping[l] = m.stopping_indices[l];
        level_num[l] = m.level_num[l];
        for (int p = 0; p < nb; p++) level_sizes[l * nb + p] = m.level_sizes[l][p];
    }
    std::vector<std::vector<std::vector<uint8_t>>> planes(L1, std::vector<std::vector<uint8_t>>(nb));
    std::vector<uint8_t> prev(L1, 0), held(L1);
    prodm::AdaptiveLevelCompressor c(latter, threads);
    for (int step = 0; step < 2; step++) {
        std::vector<uint32_t> rs(L1, 0);
        for (int l = 0; l < L1; l++) {
            held[l] = step == 0 ? (uint8_t)std::min<int>(k1[l], nb) : (uint8_t)nb;
            for (int p = prev[l]; p < held[l]; p++) rs[l] += m.level_sizes[l][p];
        }
        auto comps = r.retrieve_level_components(m.level_sizes, rs, prev, held);
        for (int l = 0; l < L1; l++) {
            const int k = held[l] - prev[l];
            if (k <= 0) continue;
            c.decompress_level(comps[l], m.level_sizes[l], prev[l], (uint8_t)k, m.stopping_indices[l]);
            const auto &n = c.decompressed_sizes();
            for (int i = 0; i < k; i++) planes[l][prev[l] + i].assign(comps[l][i], comps[l][i] + n[i]);
            c.decompress_release();
        }
        prev = held;
    }
    *retrieved = r.get_retrieved_size();
    uint64_t off = 0;
    for (int l = 0; l < L1; l++)
        for (int p = 0; p < nb; p++) {
            const auto &v = planes[l][p];
            raw_sizes[l * nb + p] = v.size();
            if (off + v.size() > cap) return -3;
            if (!v.empty()) std::memcpy(out + off, v.data(), v.size());
            off += v.size();
        }
    return 0;
}
}  // namespace

extern "C" {

int prodm_host_write_mdr(const char *metadata_path, const char *const *level_paths, int ndims, const uint32_t *dims,
                         int nlevels, int nb, int dtype, const double *bounds, int negabinary, int latter, int threads,
                         const uint8_t *streams, const uint32_t *sizes) {
    if (!metadata_path || !level_paths || ndims < 1 || ndims > 3 || nlevels < 1 || nb < 1 || nb > 64 || !bounds ||
        !streams || !sizes)
        return -1;
    return dtype ? write_mdr<double>(metadata_path, level_paths, ndims, dims, nlevels, nb, bounds, negabinary, latter,
                                     threads, streams, sizes)
                 : write_mdr<float>(metadata_path, level_paths, ndims, dims, nlevels, nb, bounds, negabinary, latter,
                                    threads, streams, sizes);
}

int prodm_host_read_mdr(const char *metadata_path, const char *const *level_paths, int nfiles, int dtype, int latter,
                        int threads, const uint8_t *k1, uint32_t *dims, int *ndims, int *nlevels, double *bounds,
                        int *nplanes, uint32_t *level_sizes, uint8_t *stopping, uint32_t *level_num, int *negabinary,
                        uint8_t *out, uint64_t cap, uint64_t *raw_sizes, uint64_t *retrieved) {
    if (!metadata_path || !level_paths || nfiles < 1 || nfiles > 64 || !k1 || !out) return -1;
    return dtype ? read_mdr<double>(metadata_path, level_paths, nfiles, latter, threads, k1, dims, ndims, nlevels,
                                    bounds, nplanes, level_sizes, stopping, level_num, negabinary, out, cap, raw_sizes,
                                    retrieved)
                 : read_mdr<float>(metadata_path, level_paths, nfiles, latter, threads, k1, dims, ndims, nlevels, bounds,
                                   nplanes, level_sizes, stopping, level_num, negabinary, out, cap, raw_sizes, retrieved);
}

}  // extern "C"

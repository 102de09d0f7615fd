/*
 * prodm_host.h -- C ABI of libprodm_host.so: the host stages around the GPU
 * library (include/prodm_mdr.h), for callers that bind C rather than the
 * header-only C++ of include/prodm/. Host memory only; no GPU.
 *
 *  - level compressors: ProDM's LosslessCompressor/DefaultLevelCompressor.hpp
 *    and AdaptiveLevelCompressor.hpp (per-plane ZSTD, stream format
 *    [size_t original length][zstd frame, level 3], ZSTD.hpp:8-28), planes
 *    compressed on `threads` host threads with ProDM's bytes and decisions;
 *  - ProDM's file layout: metadata.bin (ComposedRefactor::write_metadata,
 *    Refactor/ComposedRefactor.hpp:48-67; read in the order of
 *    ComposedReconstructor::load_metadata, Reconstructor/ComposedReconstructor.hpp:133-154)
 *    and level_<l>.bin, the level's stored plane streams back to back
 *    (Writer/FileWriter.hpp:13-39, Retriever/FileRetriever.hpp:15-81).
 */
#ifndef PRODM_HOST_H
#define PRODM_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kind 0: DefaultLevelCompressor, 1: AdaptiveLevelCompressor(latter).
 * Streams back to back in `in` (sizes[nb]); compressed streams are written
 * back to back into out (capacity cap), sizes into out_sizes. Returns the
 * stopping index (LevelCompressorInterface.hpp:14-20), or -1 when out is
 * too small.                                                               */
int prodm_host_compress_level(int kind, int latter, int threads, int nb, const uint8_t *in, const uint32_t *sizes,
                              uint8_t *out, uint64_t cap, uint32_t *out_sizes);
/* Decompresses planes [start, start+k) of a level (in: all stored streams
 * back to back, sizes[nb] their stored sizes) back to back into out;
 * raw_sizes[i] receives the size of plane start+i. Returns 0, or -1 when
 * out is too small.                                                        */
int prodm_host_decompress_level(int kind, int latter, int threads, int nb, int stopping_index, const uint8_t *in,
                                const uint32_t *sizes, int start, int k, uint8_t *out, uint64_t cap,
                                uint32_t *raw_sizes);
/* Writes ProDM's files for nlevels levels of nb plane streams each (raw
 * streams level-major, back to back, sizes[nlevels * nb]): each level
 * through AdaptiveLevelCompressor(latter), the level files, then
 * metadata.bin with the level bounds (dtype 0: stored as float, 1: double;
 * ProDM stores T) and the negabinary flag. Returns 0 or -1.                */
int prodm_host_write_mdr(const char *metadata_path, const char *const *level_paths, int ndims, const uint32_t *dims,
                         int nlevels, int nb, int dtype, const double *bounds, int negabinary, int latter, int threads,
                         const uint8_t *streams, const uint32_t *sizes);
/* Reads ProDM's files: the metadata fields (dims[3], bounds[64],
 * level_sizes[64 * 64] level-major, stopping[64], level_num[64]), then every
 * plane through the concatenated-level retriever in two progressive steps
 * (planes [0, k1[l]) of each level, then the rest), decompressed with
 * AdaptiveLevelCompressor(latter): raw streams level-major back to back into
 * out (capacity cap), their sizes in raw_sizes; *retrieved = the bytes the
 * retriever reports. Returns 0, or a negative value on a malformed file or
 * a short buffer.                                                          */
int prodm_host_read_mdr(const char *metadata_path, const char *const *level_paths, int nfiles, int dtype, int latter,
                        int threads, const uint8_t *k1, uint32_t *dims, int *ndims, int *nlevels, double *bounds,
                        int *nplanes, uint32_t *level_sizes, uint8_t *stopping, uint32_t *level_num, int *negabinary,
                        uint8_t *out, uint64_t cap, uint64_t *raw_sizes, uint64_t *retrieved);

#ifdef __cplusplus
}
#endif
#endif

// io.hpp -- ProDM's on-disk layout: metadata serialization helpers and the
// concatenated per-level file writer / retriever (host C++).
//
//   level_i.bin   the level's plane streams back to back, plane order
//                 (Writer/FileWriter.hpp:13-39)
//   metadata.bin  u8 ndims | u32 dims[ndims] | u8 nlevels | T bounds[nlevels]
//                 | per level: u32 count, u32 stream_sizes[count]
//                 | u8 stopping_indices[nlevels] | u32 level_num[nlevels]
//                 | u8 negabinary
//                 (Refactor/ComposedRefactor.hpp:48-67, RefactorUtils.hpp:79-125)
//
// The retriever keeps per-level read offsets and hands out pointers to the
// newly read planes (Retriever/FileRetriever.hpp:15-81); the retrieved byte
// count is 64-bit here (ProDM's is 32-bit).
#ifndef PRODM_IO_HPP
#define PRODM_IO_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace prodm {

template <class T> uint32_t serialized_size(const std::vector<T> &v) { return (uint32_t)(v.size() * sizeof(T)); }
template <class T> uint32_t serialized_size(const std::vector<std::vector<T>> &v) {
    uint32_t n = 0;
    for (const auto &x : v) n += sizeof(uint32_t) + (uint32_t)(x.size() * sizeof(T));
    return n;
}
template <class T> void serialize(const std::vector<T> &v, uint8_t *&pos) {
    if (!v.empty()) std::memcpy(pos, v.data(), v.size() * sizeof(T));
    pos += v.size() * sizeof(T);
}
template <class T> void serialize(const std::vector<std::vector<T>> &v, uint8_t *&pos) {
    for (const auto &x : v) {
        const uint32_t n = (uint32_t)x.size();
        std::memcpy(pos, &n, sizeof n);
        pos += sizeof n;
        serialize(x, pos);
    }
}
template <class T> void deserialize(const uint8_t *&pos, uint32_t count, std::vector<T> &v) {
    v.resize(count);
    if (count) std::memcpy(v.data(), pos, count * sizeof(T));
    pos += count * sizeof(T);
}
template <class T> void deserialize(const uint8_t *&pos, uint32_t levels, std::vector<std::vector<T>> &v) {
    v.assign(levels, {});
    for (uint32_t l = 0; l < levels; l++) {
        uint32_t n;
        std::memcpy(&n, pos, sizeof n);
        pos += sizeof n;
        deserialize(pos, n, v[l]);
    }
}

constexpr uint32_t kSquaredErrorTag = 0x4c325351u;  // trailing metadata block of L2 tables

// metadata.bin contents (ComposedRefactor.hpp:48-67 / ComposedReconstructor.hpp:133-154).
// The optional squared-error tables follow ProDM's last field, where
// ProDM's reader does not look.
template <class T>
struct MdrMetadata {
    std::vector<uint32_t> dims;
    std::vector<T> level_error_bounds;
    std::vector<std::vector<uint32_t>> level_sizes;
    std::vector<uint8_t> stopping_indices;
    std::vector<uint32_t> level_num;
    bool negabinary = false;
    std::vector<std::vector<double>> level_squared_errors;  // empty: not collected

    std::vector<uint8_t> serialize_bytes() const {
        uint32_t size = 1 + serialized_size(dims) + 1 + serialized_size(level_error_bounds) +
                        serialized_size(level_sizes) + serialized_size(stopping_indices) + serialized_size(level_num) + 1;
        const bool sq = !level_squared_errors.empty();
        if (sq) size += sizeof(uint32_t) + serialized_size(level_squared_errors);
        std::vector<uint8_t> m(size);
        uint8_t *pos = m.data();
        *pos++ = (uint8_t)dims.size();
        serialize(dims, pos);
        *pos++ = (uint8_t)level_error_bounds.size();
        serialize(level_error_bounds, pos);
        serialize(level_sizes, pos);
        serialize(stopping_indices, pos);
        serialize(level_num, pos);
        *pos++ = (uint8_t)negabinary;
        if (sq) {
            std::memcpy(pos, &kSquaredErrorTag, 4);
            pos += 4;
            serialize(level_squared_errors, pos);
        }
        return m;
    }
    // `size` bounds the optional trailing block; ProDM's own fields are read
    // as ProDM reads them.
    void parse(const uint8_t *m, size_t size) {
        const uint8_t *pos = m;
        const uint8_t nd = *pos++;
        deserialize(pos, nd, dims);
        const uint8_t nl = *pos++;
        deserialize(pos, nl, level_error_bounds);
        deserialize(pos, nl, level_sizes);
        deserialize(pos, nl, stopping_indices);
        deserialize(pos, nl, level_num);
        negabinary = *pos++ != 0;
        level_squared_errors.clear();
        uint32_t tag = 0;
        if ((size_t)(pos - m) + 4 <= size) std::memcpy(&tag, pos, 4);
        if (tag == kSquaredErrorTag) {
            pos += 4;
            deserialize(pos, nl, level_squared_errors);
        }
    }
};

inline void write_file(const std::string &path, const uint8_t *data, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "prodm: cannot write %s\n", path.c_str());
        std::exit(-1);
    }
    if (n && std::fwrite(data, 1, n, f) != n) {
        std::fprintf(stderr, "prodm: short write to %s\n", path.c_str());
        std::exit(-1);
    }
    std::fclose(f);
}

class ConcatLevelFileWriter {
  public:
    ConcatLevelFileWriter(const std::string &metadata_file, const std::vector<std::string> &level_files)
        : metadata_file_(metadata_file), level_files_(level_files) {}
    std::vector<uint32_t> write_level_components(const std::vector<std::vector<uint8_t *>> &components,
                                                 const std::vector<std::vector<uint32_t>> &sizes) const {
        std::vector<uint32_t> level_num;
        for (size_t l = 0; l < components.size(); l++) {
            FILE *f = std::fopen(level_files_[l].c_str(), "wb");
            if (!f) {
                std::fprintf(stderr, "prodm: cannot write %s\n", level_files_[l].c_str());
                std::exit(-1);
            }
            for (size_t p = 0; p < components[l].size(); p++)
                if (sizes[l][p] && std::fwrite(components[l][p], 1, sizes[l][p], f) != sizes[l][p]) {
                    std::fprintf(stderr, "prodm: short write to %s\n", level_files_[l].c_str());
                    std::exit(-1);
                }
            std::fclose(f);
            level_num.push_back(1);
        }
        return level_num;
    }
    void write_metadata(const uint8_t *metadata, uint32_t size) const { write_file(metadata_file_, metadata, size); }
    void print() const { std::printf("concatenated level file writer\n"); }

  private:
    std::string metadata_file_;
    std::vector<std::string> level_files_;
};

class ConcatLevelFileRetriever {
  public:
    ConcatLevelFileRetriever(const std::string &metadata_file, const std::vector<std::string> &level_files)
        : metadata_file_(metadata_file), level_files_(level_files), offsets_(level_files.size(), 0) {}
    ConcatLevelFileRetriever(const ConcatLevelFileRetriever &o)
        : metadata_file_(o.metadata_file_), level_files_(o.level_files_), offsets_(o.offsets_),
          retrieved_(o.retrieved_) {}
    ~ConcatLevelFileRetriever() { release(); }

    std::vector<std::vector<const uint8_t *>> retrieve_level_components(
        const std::vector<std::vector<uint32_t>> &sizes, const std::vector<uint32_t> &retrieve_sizes,
        const std::vector<uint8_t> &prev_planes, const std::vector<uint8_t> &planes) {
        release();
        uint64_t total = 0;
        for (size_t l = 0; l < retrieve_sizes.size(); l++) {
            uint8_t *buf = static_cast<uint8_t *>(std::malloc(retrieve_sizes[l] ? retrieve_sizes[l] : 1));
            if (retrieve_sizes[l]) {
                FILE *f = std::fopen(level_files_[l].c_str(), "rb");
                if (!f || std::fseek(f, (long)offsets_[l], SEEK_SET) ||
                    std::fread(buf, 1, retrieve_sizes[l], f) != retrieve_sizes[l]) {
                    std::fprintf(stderr, "prodm: cannot read %u bytes of %s at %llu\n", retrieve_sizes[l],
                                 level_files_[l].c_str(), (unsigned long long)offsets_[l]);
                    std::exit(-1);
                }
                std::fclose(f);
            }
            held_.push_back(buf);
            offsets_[l] += retrieve_sizes[l];
            total += offsets_[l];
        }
        retrieved_ = total;
        std::vector<std::vector<const uint8_t *>> out;
        for (size_t l = 0; l < planes.size(); l++) {
            const uint8_t *pos = l < held_.size() ? held_[l] : nullptr;
            std::vector<const uint8_t *> lv;
            for (int p = prev_planes[l]; p < planes[l]; p++) {
                lv.push_back(pos);
                pos += sizes[l][p];
            }
            out.push_back(lv);
        }
        return out;
    }
    uint8_t *load_metadata() const { return load_metadata(nullptr); }
    uint8_t *load_metadata(size_t *size) const {
        FILE *f = std::fopen(metadata_file_.c_str(), "rb");
        if (!f) {
            std::fprintf(stderr, "prodm: cannot read %s\n", metadata_file_.c_str());
            std::exit(-1);
        }
        std::fseek(f, 0, SEEK_END);
        const long n = std::ftell(f);
        std::rewind(f);
        uint8_t *m = static_cast<uint8_t *>(std::malloc(n > 0 ? n : 1));
        if (n > 0 && std::fread(m, 1, n, f) != (size_t)n) {
            std::fprintf(stderr, "prodm: short read of %s\n", metadata_file_.c_str());
            std::exit(-1);
        }
        std::fclose(f);
        if (size) *size = (size_t)(n > 0 ? n : 0);
        return m;
    }
    void release() {
        for (uint8_t *b : held_) std::free(b);
        held_.clear();
    }
    size_t get_retrieved_size() const { return retrieved_; }
    std::vector<uint64_t> get_offsets() const { return offsets_; }
    void print() const { std::printf("concatenated level file retriever\n"); }

  private:
    std::string metadata_file_;
    std::vector<std::string> level_files_;
    std::vector<uint64_t> offsets_;
    std::vector<uint8_t *> held_;
    uint64_t retrieved_ = 0;
};

}  // namespace prodm
#endif

// prodm_refactor -- refactor a raw binary field into ProDM's files on the GPU
// (counterpart of ProDM's test/test_mdr_refactor.cpp:34-86).
//
//   prodm_refactor <data file> <target level> <num bitplanes> <num dims> <dims...>
//                  [--f64] [--grouped | --perbit] [--composed] [--squared | --squared-encoder]
//                  [--hierarchical] [--out DIR]
//
// Writes DIR/metadata.bin and DIR/level_<l>.bin (default DIR refactored_data).
// Default: float, negabinary encoder, AdaptiveLevelCompressor(64), the
// device-resident GpuRefactor; --composed drives the GPU components through
// prodm::ComposedRefactor instead (same files). --hierarchical uses the
// hierarchical basis (ProDM's drivers use MGARDHierarchicalDecomposer);
// --perbit the PerBit encoder (with --hierarchical, the GE applications'
// pairing). --squared stores SquaredErrorCollector's squared-error tables for
// L2 retrieval; --squared-encoder (float) stores the encoder's own truncation
// tables instead, which an L2 retrieval of negabinary streams needs to meet
// its target (SquaredErrorCollector models sign-magnitude truncation).
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "prodm/prodm.hpp"

using namespace prodm;

// The raw field, read into pinned host memory: its upload then runs at the
// link's rate instead of being staged through the driver's bounce buffers
// (540 MB at 513^3).
template <class T> class RawField {
  public:
    RawField(const std::string &path, size_t n) {
        hip_check(hipHostMalloc((void **)&p_, (n ? n : 1) * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
        FILE *f = std::fopen(path.c_str(), "rb");
        const bool ok = f && std::fread(p_, sizeof(T), n, f) == n;
        if (f) std::fclose(f);
        if (!ok) fatal("cannot read " + std::to_string(n) + " values from " + path);
    }
    ~RawField() {
        if (p_) (void)hipHostFree(p_);
    }
    RawField(const RawField &) = delete;
    RawField &operator=(const RawField &) = delete;
    const T *data() const { return p_; }

  private:
    T *p_ = nullptr;
};

template <class T, class S>
static int run(const std::string &file, const std::vector<uint32_t> &dims, int L, int nb, bool grouped, bool perbit,
               bool composed, int squared, bool hier, const std::string &dir) {
    const RawField<T> data(file, product(dims));
    std::vector<std::string> files;
    for (int l = 0; l <= L; l++) files.push_back(dir + "/level_" + std::to_string(l) + ".bin");
    ConcatLevelFileWriter writer(dir + "/metadata.bin", files);
    const auto t0 = std::chrono::steady_clock::now();
    if (composed) {
        Context ctx = make_context();
        auto go2 = [&](auto decomposer, auto encoder, auto collector) {
            ComposedRefactor<T, decltype(decomposer), GpuDirectInterleaver<T>, decltype(encoder),
                             AdaptiveLevelCompressor, decltype(collector), ConcatLevelFileWriter>
                r{decomposer, GpuDirectInterleaver<T>(ctx), encoder, AdaptiveLevelCompressor(64), collector, writer};
            r.negabinary = !grouped && !perbit;
            r.refactor(data.data(), dims, L, nb);
        };
        auto go = [&](auto encoder, auto collector) {
            if (hier) go2(GpuMGARDHierarchicalDecomposer<T>(ctx), encoder, collector);
            else go2(GpuMGARDOrthogonalDecomposer<T>(ctx), encoder, collector);
        };
        auto with = [&](auto encoder, int enc) {
            if (squared == kCollectorTables) {
                go(encoder, GpuSquaredErrorCollector<T>(ctx));
            } else if (squared == kEncoderTables) {
                if constexpr (std::is_same<T, float>::value) go(encoder, GpuEncoderErrorCollector<T>(enc, ctx));
                else fatal("--squared-encoder needs float data");
            } else {
                go(encoder, MaxErrorCollector<T>());
            }
        };
        if (perbit) with(GpuPerBitBPEncoder<T, S>(ctx), PRODM_ENC_PERBIT);
        else if (grouped) with(GpuGroupedBPEncoder<T, S>(ctx), PRODM_ENC_GROUPED);
        else with(GpuNegaBinaryBPEncoder<T, S>(ctx), PRODM_ENC_NEGABINARY);
    } else {
        GpuRefactor<T, AdaptiveLevelCompressor, ConcatLevelFileWriter> r(
            perbit ? PRODM_ENC_PERBIT : grouped ? PRODM_ENC_GROUPED : PRODM_ENC_NEGABINARY, AdaptiveLevelCompressor(64), writer, squared,
            hier ? PRODM_DECOMP_HIERARCHICAL : PRODM_DECOMP_ORTHOGONAL);
        r.refactor(data.data(), dims, L, nb);
        const auto t = r.times();
        std::printf("Stages: device %.6fs (upload + GPU stages), copy-out + lossless %.6fs, files %.6fs\n", t.device,
                    t.host, t.write);
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Refactor time: %.6fs (%s, %s)\n", s, composed ? "composed GPU components" : "device-resident",
                perbit ? "per-bit" : grouped ? "grouped" : "negabinary");
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s <file> <target level> <bitplanes> <ndims> <dims...> [--f64] [--grouped] "
                             "[--perbit] [--composed] [--squared | --squared-encoder] [--hierarchical] [--out DIR]\n",
                     argv[0]);
        return 2;
    }
    int a = 1;
    const std::string file = argv[a++];
    const int L = std::atoi(argv[a++]);
    int nb = std::atoi(argv[a++]);
    const int nd = std::atoi(argv[a++]);
    std::vector<uint32_t> dims;
    for (int i = 0; i < nd && a < argc; i++) dims.push_back((uint32_t)std::atoi(argv[a++]));
    bool f64 = false, grouped = false, perbit = false, composed = false, hier = false;
    int squared = kNoSquaredTables;
    std::string dir = "refactored_data";
    for (; a < argc; a++) {
        const std::string o = argv[a];
        if (o == "--f64") f64 = true;
        else if (o == "--grouped") grouped = true;
        else if (o == "--perbit") perbit = true;
        else if (o == "--composed") composed = true;
        else if (o == "--squared") squared = kCollectorTables;
        else if (o == "--squared-encoder") squared = kEncoderTables;
        else if (o == "--hierarchical") hier = true;
        else if (o == "--out" && a + 1 < argc) dir = argv[++a];
        else fatal("unknown option " + o);
    }
    if (!grouped && !perbit && (nb & 1)) {
        nb += 1;  // negabinary needs an even plane count (test_mdr_refactor.cpp:40-43)
        std::printf("Change to %d bitplanes for negabinary encoding\n", nb);
    }
    const int maxnb = f64 ? 64 : 32;
    if (nb > maxnb) nb = maxnb;
    mkdir(dir.c_str(), 0755);
    return f64 ? run<double, uint64_t>(file, dims, L, nb, grouped, perbit, composed, squared, hier, dir)
               : run<float, uint32_t>(file, dims, L, nb, grouped, perbit, composed, squared, hier, dir);
}

// test_cpp_pipeline.cpp -- the host C++ layer end to end on the GPU.
//
//  1. prodm::ComposedRefactor driving the GPU components and
//     prodm::GpuRefactor (device-resident) write byte-identical files, for
//     the negabinary and the grouped encoder.
//  2. prodm::ComposedReconstructor and prodm::GpuReconstructor read them
//     progressively; each result meets the requested L-infinity tolerance
//     and the two agree to rounding.
//  3. The negabinary encoder's squared-error tables (composed and
//     device-resident refactors write the same files) drive S-norm
//     reconstructions that meet each relative L2 target 1e-2 .. 1e-6.
//  4. GPU components against each other: encode -> progressive decode
//     equals one-shot decode (negabinary, grouped, per-bit); interleave ->
//     reposition is the identity.
//  5. The per-bit encoder with the hierarchical decomposer: the composers
//     and GpuRefactor write byte-identical files; the composed and the
//     device-resident reconstructions meet each L-infinity tolerance.
//
// Usage: test_cpp_pipeline <scratch dir>. Exit status 0 on success.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "prodm/prodm.hpp"

using namespace prodm;

static int failures = 0;
#define EXPECT(cond, ...)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
            failures++;                                    \
        }                                                  \
    } while (0)

static std::vector<float> field(const std::vector<uint32_t> &d) {
    std::vector<float> u((size_t)d[0] * d[1] * d[2]);
    size_t i = 0;
    for (uint32_t a = 0; a < d[0]; a++)
        for (uint32_t b = 0; b < d[1]; b++)
            for (uint32_t c = 0; c < d[2]; c++) {
                const double x = a / (d[0] - 1.0), y = b / (d[1] - 1.0), z = c / (d[2] - 1.0);
                double v = std::sin(2 * M_PI * 3 * x) * std::cos(2 * M_PI * 5 * y) * std::sin(2 * M_PI * 2 * z);
                v += 1.5 * std::exp(-((x - .3) * (x - .3) + (y - .6) * (y - .6) + (z - .4) * (z - .4)) / (2 * .05 * .05));
                v += 1e-3 * std::sin(12345.678 * (double)i);
                u[i++] = (float)v;
            }
    return u;
}

static std::vector<std::string> level_files(const std::string &dir, int L) {
    std::vector<std::string> f;
    for (int l = 0; l <= L; l++) f.push_back(dir + "/level_" + std::to_string(l) + ".bin");
    return f;
}

static std::vector<uint8_t> slurp(const std::string &p) {
    std::vector<uint8_t> b;
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return b;
    std::fseek(f, 0, SEEK_END);
    b.resize(std::ftell(f));
    std::rewind(f);
    if (!b.empty() && std::fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
    std::fclose(f);
    return b;
}

static double linf(const std::vector<float> &a, const float *b) {
    double m = 0;
    for (size_t i = 0; i < a.size(); i++) m = std::max(m, std::fabs((double)a[i] - b[i]));
    return m;
}

static double range_of(const std::vector<float> &a) {
    float lo = a[0], hi = a[0];
    for (float v : a) lo = std::min(lo, v), hi = std::max(hi, v);
    return (double)hi - lo;
}

template <class Encoder>
static void roundtrip_files(const std::string &root, const std::vector<float> &u, const std::vector<uint32_t> &dims,
                            int L, int nb, bool nega, Encoder enc) {
    const std::string a = root + (nega ? "/composed_nb" : "/composed_gb"), b = root + (nega ? "/gpu_nb" : "/gpu_gb");
    mkdir(a.c_str(), 0755);
    mkdir(b.c_str(), 0755);
    Context ctx = make_context();
    {
        ComposedRefactor<float, GpuMGARDOrthogonalDecomposer<float>, GpuDirectInterleaver<float>, Encoder,
                         AdaptiveLevelCompressor, MaxErrorCollector<float>, ConcatLevelFileWriter>
            r{GpuMGARDOrthogonalDecomposer<float>(ctx), GpuDirectInterleaver<float>(ctx), enc,
              AdaptiveLevelCompressor(64), MaxErrorCollector<float>(),
              ConcatLevelFileWriter(a + "/metadata.bin", level_files(a, L))};
        r.negabinary = nega;
        r.refactor(u.data(), dims, L, nb);
    }
    {
        GpuRefactor<float, AdaptiveLevelCompressor, ConcatLevelFileWriter> r(
            nega ? PRODM_ENC_NEGABINARY : PRODM_ENC_GROUPED, AdaptiveLevelCompressor(64),
            ConcatLevelFileWriter(b + "/metadata.bin", level_files(b, L)));
        r.refactor(u.data(), dims, L, nb);
    }
    EXPECT(slurp(a + "/metadata.bin") == slurp(b + "/metadata.bin") && !slurp(a + "/metadata.bin").empty(),
           "metadata differs (%s)", nega ? "negabinary" : "grouped");
    for (int l = 0; l <= L; l++)
        EXPECT(slurp(level_files(a, L)[l]) == slurp(level_files(b, L)[l]), "level %d file differs (%s)", l,
               nega ? "negabinary" : "grouped");

    using Est = MaxErrorEstimatorOB<float>;
    using Interp = GreedyBasedSizeInterpreter<Est>;
    ComposedReconstructor<float, GpuMGARDOrthogonalDecomposer<float>, GpuDirectInterleaver<float>, Encoder,
                          AdaptiveLevelCompressor, Interp, Est, ConcatLevelFileRetriever>
        rc{GpuMGARDOrthogonalDecomposer<float>(ctx), GpuDirectInterleaver<float>(ctx), enc, AdaptiveLevelCompressor(64),
           Interp(Est(3)), ConcatLevelFileRetriever(a + "/metadata.bin", level_files(a, L))};
    rc.load_metadata();
    GpuReconstructor<float, AdaptiveLevelCompressor, Interp, Est, ConcatLevelFileRetriever> rg(
        AdaptiveLevelCompressor(64), Interp(Est(3)), ConcatLevelFileRetriever(b + "/metadata.bin", level_files(b, L)));
    rg.load_metadata();
    const double R = range_of(u);
    for (double rel : {1e-1, 1e-2, 1e-3, 1e-4, 1e-5}) {
        const double tol = rel * R;
        float *x = rc.progressive_reconstruct(tol);
        std::vector<float> full(x, x + u.size());
        const bool reduced = rc.get_current_dimensions() != dims;
        if (reduced) {
            float *y = rc.recompose_to_full();
            full.assign(y, y + u.size());
        }
        const float *g = rg.progressive_reconstruct(tol);
        const double ec = linf(u, full.data()), eg = linf(u, g), d = linf(full, g);
        std::printf("%s tol %.1e: composed err %.3e (%s), gpu err %.3e, retrieved %zu / %zu bytes, |composed-gpu| %.2e\n",
                    nega ? "negabinary" : "grouped", rel, ec / R, reduced ? "coarse grid" : "full grid", eg / R,
                    rc.get_retrieved_size(), rg.get_retrieved_size(), d);
        if (!reduced) EXPECT(ec <= tol, "composed error %.3e above %.3e", ec, tol);
        EXPECT(eg <= tol, "gpu error %.3e above %.3e", eg, tol);
        EXPECT(rc.get_retrieved_size() == rg.get_retrieved_size(), "retrieved bytes differ");
        if (!reduced) EXPECT(d <= 1e-5 * R, "composed and gpu results differ by %.3e", d);
    }
}

static void squared_path(const std::string &root, const std::vector<float> &u, const std::vector<uint32_t> &dims, int L,
                         int nb) {
    // Tables of the negabinary encoder's own truncation (GpuEncoderErrorCollector):
    // the composed refactor with that collector and GpuRefactor(kEncoderTables)
    // write the same files, and an S-norm retrieval planned on them meets each
    // relative L2 target. ProDM's SquaredErrorCollector tables (the composed
    // refactor with GpuSquaredErrorCollector) are stored as well; planned on
    // them, negabinary retrievals can miss the target, so those results are
    // only reported.
    const std::string a = root + "/composed_sq", b = root + "/gpu_sq", c = root + "/gpu_sqc";
    mkdir(a.c_str(), 0755);
    mkdir(b.c_str(), 0755);
    mkdir(c.c_str(), 0755);
    Context ctx = make_context();
    {
        ComposedRefactor<float, GpuMGARDOrthogonalDecomposer<float>, GpuDirectInterleaver<float>,
                         GpuNegaBinaryBPEncoder<float, uint32_t>, DefaultLevelCompressor, GpuEncoderErrorCollector<float>,
                         ConcatLevelFileWriter>
            r{GpuMGARDOrthogonalDecomposer<float>(ctx), GpuDirectInterleaver<float>(ctx),
              GpuNegaBinaryBPEncoder<float, uint32_t>(ctx), DefaultLevelCompressor(),
              GpuEncoderErrorCollector<float>(PRODM_ENC_NEGABINARY, ctx),
              ConcatLevelFileWriter(a + "/metadata.bin", level_files(a, L))};
        r.negabinary = true;
        r.refactor(u.data(), dims, L, nb);
    }
    {
        GpuRefactor<float, DefaultLevelCompressor, ConcatLevelFileWriter> r(
            PRODM_ENC_NEGABINARY, DefaultLevelCompressor(), ConcatLevelFileWriter(b + "/metadata.bin", level_files(b, L)),
            kEncoderTables);
        r.refactor(u.data(), dims, L, nb);
    }
    {
        GpuRefactor<float, DefaultLevelCompressor, ConcatLevelFileWriter> r(
            PRODM_ENC_NEGABINARY, DefaultLevelCompressor(), ConcatLevelFileWriter(c + "/metadata.bin", level_files(c, L)),
            kCollectorTables);
        r.refactor(u.data(), dims, L, nb);
    }
    EXPECT(slurp(a + "/metadata.bin") == slurp(b + "/metadata.bin") && !slurp(a + "/metadata.bin").empty(),
           "squared-table metadata differs (composed vs device-resident)");
    for (int l = 0; l <= L; l++)
        EXPECT(slurp(level_files(a, L)[l]) == slurp(level_files(b, L)[l]), "squared-table level %d file differs", l);
    using Est = SNormErrorEstimator<float>;
    using Interp = NegaBinaryGreedyBasedSizeInterpreter<Est>;
    double norm2 = 0;
    for (float v : u) norm2 += (double)v * v;
    for (const std::string &dir : {b, c}) {
        GpuReconstructor<float, DefaultLevelCompressor, Interp, Est, ConcatLevelFileRetriever> rg(
            DefaultLevelCompressor(), Interp(Est(3, L, 0)),
            ConcatLevelFileRetriever(dir + "/metadata.bin", level_files(dir, L)));
        rg.load_metadata();
        const bool encoder_tables = dir == b;
        for (double rel : {1e-2, 1e-3, 1e-4, 1e-5, 1e-6}) {
            // the S-norm estimate sums squared coefficient errors weighted by
            // level volume, in units of the fine grid: ask for (rel ||u||)^2
            const double tol = rel * rel * norm2;
            const float *g = rg.progressive_reconstruct(tol);
            EXPECT(g != nullptr, "no squared tables");
            if (!g) return;
            double e2 = 0;
            for (size_t i = 0; i < u.size(); i++) e2 += ((double)u[i] - g[i]) * ((double)u[i] - g[i]);
            const double got = std::sqrt(e2 / norm2);
            std::printf("L2 target %.0e (%s tables): actual relative L2 %.3e, retrieved %zu bytes\n", rel,
                        encoder_tables ? "encoder" : "SquaredErrorCollector", got, rg.get_retrieved_size());
            if (encoder_tables) EXPECT(got <= rel, "relative L2 %.3e above the target %.0e", got, rel);
        }
    }
}

// ProDM's own driver pairing (test/test_mdr_refactor.cpp:64-70,
// test_mdr_reconstructor.cpp:70-83): hierarchical decomposer, negabinary,
// MaxErrorEstimatorHB with the sign-excluding greedy planner.
static void hierarchical_path(const std::string &root, const std::vector<float> &u, const std::vector<uint32_t> &dims,
                              int L, int nb) {
    const std::string a = root + "/composed_hb", b = root + "/gpu_hb";
    mkdir(a.c_str(), 0755);
    mkdir(b.c_str(), 0755);
    Context ctx = make_context();
    {
        ComposedRefactor<float, GpuMGARDHierarchicalDecomposer<float>, GpuDirectInterleaver<float>,
                         GpuNegaBinaryBPEncoder<float, uint32_t>, AdaptiveLevelCompressor, MaxErrorCollector<float>,
                         ConcatLevelFileWriter>
            r{GpuMGARDHierarchicalDecomposer<float>(ctx), GpuDirectInterleaver<float>(ctx),
              GpuNegaBinaryBPEncoder<float, uint32_t>(ctx), AdaptiveLevelCompressor(64), MaxErrorCollector<float>(),
              ConcatLevelFileWriter(a + "/metadata.bin", level_files(a, L))};
        r.negabinary = true;
        r.refactor(u.data(), dims, L, nb);
    }
    {
        GpuRefactor<float, AdaptiveLevelCompressor, ConcatLevelFileWriter> r(
            PRODM_ENC_NEGABINARY, AdaptiveLevelCompressor(64), ConcatLevelFileWriter(b + "/metadata.bin", level_files(b, L)),
            kNoSquaredTables, PRODM_DECOMP_HIERARCHICAL);
        r.refactor(u.data(), dims, L, nb);
    }
    EXPECT(slurp(a + "/metadata.bin") == slurp(b + "/metadata.bin"), "hierarchical metadata differs");
    for (int l = 0; l <= L; l++)
        EXPECT(slurp(level_files(a, L)[l]) == slurp(level_files(b, L)[l]), "hierarchical level %d file differs", l);
    using Est = MaxErrorEstimatorHB<float>;
    using Interp = SignExcludeGreedyBasedSizeInterpreter<Est>;
    GpuReconstructor<float, AdaptiveLevelCompressor, Interp, Est, ConcatLevelFileRetriever> rg(
        AdaptiveLevelCompressor(64), Interp(Est()), ConcatLevelFileRetriever(b + "/metadata.bin", level_files(b, L)),
        PRODM_DECOMP_HIERARCHICAL);
    rg.load_metadata();
    const double R = range_of(u);
    for (double rel : {1e-1, 1e-3, 1e-5}) {
        const float *g = rg.progressive_reconstruct(rel * R);
        const double e = linf(u, g);
        std::printf("hierarchical tol %.1e: gpu err %.3e, retrieved %zu bytes\n", rel, e / R, rg.get_retrieved_size());
        EXPECT(e <= rel * R, "hierarchical error %.3e above %.3e", e, rel * R);
    }
}

// The GE applications' pairing (Synthesizer4GE.hpp:123): hierarchical
// decomposer with the per-bit encoder, composed from the GPU components;
// the progressive reconstruction meets each L-infinity tolerance.
static void perbit_path(const std::string &root, const std::vector<float> &u, const std::vector<uint32_t> &dims,
                        int L, int nb) {
    const std::string a = root + "/composed_pb", b = root + "/gpu_pb";
    mkdir(a.c_str(), 0755);
    mkdir(b.c_str(), 0755);
    Context ctx = make_context();
    {  // the device-resident path with the PerBit encoder writes the same files
        GpuRefactor<float, AdaptiveLevelCompressor, ConcatLevelFileWriter> r(
            PRODM_ENC_PERBIT, AdaptiveLevelCompressor(64), ConcatLevelFileWriter(b + "/metadata.bin", level_files(b, L)),
            kNoSquaredTables, PRODM_DECOMP_HIERARCHICAL);
        r.refactor(u.data(), dims, L, nb);
    }
    {
        ComposedRefactor<float, GpuMGARDHierarchicalDecomposer<float>, GpuDirectInterleaver<float>,
                         GpuPerBitBPEncoder<float, uint32_t>, AdaptiveLevelCompressor, MaxErrorCollector<float>,
                         ConcatLevelFileWriter>
            r{GpuMGARDHierarchicalDecomposer<float>(ctx), GpuDirectInterleaver<float>(ctx),
              GpuPerBitBPEncoder<float, uint32_t>(ctx), AdaptiveLevelCompressor(64), MaxErrorCollector<float>(),
              ConcatLevelFileWriter(a + "/metadata.bin", level_files(a, L))};
        r.refactor(u.data(), dims, L, nb);
    }
    using Est = MaxErrorEstimatorHB<float>;
    using Interp = SignExcludeGreedyBasedSizeInterpreter<Est>;
    ComposedReconstructor<float, GpuMGARDHierarchicalDecomposer<float>, GpuDirectInterleaver<float>,
                          GpuPerBitBPEncoder<float, uint32_t>, AdaptiveLevelCompressor, Interp, Est,
                          ConcatLevelFileRetriever>
        rc{GpuMGARDHierarchicalDecomposer<float>(ctx), GpuDirectInterleaver<float>(ctx),
           GpuPerBitBPEncoder<float, uint32_t>(ctx), AdaptiveLevelCompressor(64), Interp(Est()),
           ConcatLevelFileRetriever(a + "/metadata.bin", level_files(a, L))};
    rc.load_metadata();
    EXPECT(slurp(a + "/metadata.bin") == slurp(b + "/metadata.bin"), "per-bit metadata differs");
    for (int l = 0; l <= L; l++)
        EXPECT(slurp(level_files(a, L)[l]) == slurp(level_files(b, L)[l]), "per-bit level %d file differs", l);
    GpuReconstructor<float, AdaptiveLevelCompressor, Interp, Est, ConcatLevelFileRetriever> rg(
        AdaptiveLevelCompressor(64), Interp(Est()), ConcatLevelFileRetriever(b + "/metadata.bin", level_files(b, L)),
        PRODM_DECOMP_HIERARCHICAL, PRODM_ENC_PERBIT);
    rg.load_metadata();
    const double R = range_of(u);
    for (double rel : {1e-1, 1e-3, 1e-5}) {
        const float *g = rg.progressive_reconstruct(rel * R);
        const double e = linf(u, g);
        std::printf("per-bit tol %.1e: gpu err %.3e, retrieved %zu bytes\n", rel, e / R, rg.get_retrieved_size());
        EXPECT(e <= rel * R, "per-bit gpu error %.3e above %.3e", e, rel * R);
    }
    // recompose_to_full overwrites the held field (ComposedReconstructor.hpp:123-131),
    // so a coarse-grid result is only reported
    int checked = 0;
    for (double rel : {1e-1, 1e-2, 1e-3, 1e-4, 1e-5}) {
        float *x = rc.progressive_reconstruct(rel * R);
        const bool reduced = rc.get_current_dimensions() != dims;
        const double e = reduced ? -1 : linf(u, x);
        std::printf("per-bit tol %.1e: composed err %.3e (%s), retrieved %zu bytes\n", rel, e / R,
                    reduced ? "coarse grid" : "full grid", rc.get_retrieved_size());
        if (!reduced) EXPECT(e <= rel * R, "per-bit error %.3e above %.3e", e, rel * R);
        checked += !reduced;
    }
    EXPECT(checked >= 3, "per-bit: only %d tolerances reached the full grid", checked);
}

template <class Encoder>
static void component_checks(const std::vector<float> &u, int nb, Encoder enc) {
    const int n = (int)u.size() / 7 + 5;
    float mx = 0;
    for (int i = 0; i < n; i++) mx = std::max(mx, std::fabs(u[i]));
    int e;
    std::frexp(mx, &e);
    std::vector<uint32_t> sizes;
    auto streams = enc.encode(u.data(), n, e, nb, sizes);
    std::vector<const uint8_t *> cs(streams.begin(), streams.end());
    float *one = enc.decode(cs, n, e, nb);
    std::vector<double> acc(n, 0.0);
    const int cuts[] = {0, 3, nb / 2, nb};
    for (int c = 0; c < 3; c++) {
        std::vector<const uint8_t *> part(cs.begin() + cuts[c], cs.begin() + cuts[c + 1]);
        float *d = enc.progressive_decode(part, n, e, cuts[c], cuts[c + 1] - cuts[c], 5);
        for (int i = 0; i < n; i++) acc[i] += d[i];
        std::free(d);
    }
    double worst = 0;
    for (int i = 0; i < n; i++) worst = std::max(worst, std::fabs(acc[i] - one[i]));
    EXPECT(worst <= std::ldexp(1.0, e - nb + 3), "progressive deltas differ from one-shot decode by %.3e", worst);
    std::free(one);
    // the level-error overload: the same streams, and entry j = squared error
    // of decoding the first j planes (float-squared bins 0 and nb: 2^-22)
    std::vector<uint32_t> sizes2;
    std::vector<double> le;
    auto streams2 = enc.encode(u.data(), n, e, nb, sizes2, le);
    EXPECT(sizes2 == sizes && (int)le.size() == nb + 1, "level-error encode changed the streams");
    for (int p = 0; p < nb && sizes2 == sizes; p++)
        EXPECT(std::memcmp(streams[p], streams2[p], sizes[p]) == 0, "level-error encode changed plane %d", p);
    for (int j : {0, 1, nb / 2, nb}) {
        double want = 0;
        if (j == 0) {
            for (int i = 0; i < n; i++) want += (double)u[i] * u[i];
        } else {
            std::vector<const uint8_t *> part(cs.begin(), cs.begin() + j);
            // a fresh level key per prefix: GroupedBP decode state is progressive per level
            float *d = enc.progressive_decode(part, n, e, 0, j, 10 + j);
            for (int i = 0; i < n; i++) want += ((double)u[i] - d[i]) * ((double)u[i] - d[i]);
            std::free(d);
        }
        const double tol = (j == 0 || j == nb) ? std::ldexp(1.0, -22) : 1e-9;
        EXPECT(std::fabs(le[j] - want) <= tol * want + 1e-300, "level error %d: %.6e vs %.6e", j, le[j], want);
    }
    for (auto *s : streams2) std::free(s);
    for (auto *s : streams) std::free(s);
}

int main(int argc, char **argv) {
    const std::string root = argc > 1 ? argv[1] : "cpp_pipeline_out";
    mkdir(root.c_str(), 0755);
    const std::vector<uint32_t> dims = {65, 65, 65};
    const int L = 4, nb = 32;
    const auto u = field(dims);
    Context ctx = make_context();
    roundtrip_files(root, u, dims, L, nb, true, GpuNegaBinaryBPEncoder<float, uint32_t>(ctx));
    roundtrip_files(root, u, dims, L, nb, false, GpuGroupedBPEncoder<float, uint32_t>(ctx));
    squared_path(root, u, dims, L, nb);
    hierarchical_path(root, u, dims, L, nb);
    component_checks(u, nb, GpuNegaBinaryBPEncoder<float, uint32_t>(ctx));
    component_checks(u, nb, GpuGroupedBPEncoder<float, uint32_t>(ctx));
    component_checks(u, nb, GpuPerBitBPEncoder<float, uint32_t>(ctx));
    perbit_path(root, u, dims, L, nb);
    {  // interleave -> reposition is the identity on every level
        GpuDirectInterleaver<float> ilv(ctx);
        const auto ld = level_dims_of(dims, L);
        std::vector<float> back(u.size(), 0.f);
        const std::vector<uint32_t> zero(3, 0);
        const auto el = level_elements_of(dims, L);
        for (int l = 0; l <= L; l++) {
            std::vector<float> buf(el[l]);
            ilv.interleave(u.data(), dims, ld[l], l ? ld[l - 1] : zero, buf.data());
            ilv.reposition(buf.data(), dims, ld[l], l ? ld[l - 1] : zero, back.data());
        }
        EXPECT(back == u, "interleave/reposition is not the identity");
    }
    std::printf(failures ? "FAILED (%d)\n" : "PASSED\n", failures);
    return failures ? 1 : 0;
}

// prodm_reconstruct -- progressive retrieval from ProDM's files on the GPU
// (counterpart of ProDM's test/test_mdr_reconstructor.cpp:44-88).
//
//   prodm_reconstruct <original data file> <num tolerances> <tolerances...>
//                     [--f64] [--l2] [--hierarchical] [--perbit] [--in DIR]
//
// Reads DIR/metadata.bin (default DIR refactored_data). For each tolerance:
// plans with the greedy interpreter (L-infinity: MaxErrorEstimatorOB;
// --l2: S-norm estimator with the negabinary greedy interpreter, needs files
// written with --squared; --hierarchical: files written with the hierarchical
// basis, planned with MaxErrorEstimatorHB and the sign-excluding greedy
// planner as ProDM's driver does; --perbit: files written with the PerBit
// encoder, which ProDM's metadata does not record), reads and decodes the new planes,
// recomposes, and prints the retrieved bytes and the error against the
// original.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "prodm/prodm.hpp"

using namespace prodm;

template <class T>
static std::vector<T> read_raw(const std::string &path, size_t &n) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) fatal("cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    n = std::ftell(f) / sizeof(T);
    std::rewind(f);
    std::vector<T> v(n);
    if (std::fread(v.data(), sizeof(T), n, f) != n) fatal("short read of " + path);
    std::fclose(f);
    return v;
}

template <class T, class R>
static void evaluate(const std::vector<T> &u, const std::vector<double> &tols, R &r) {
    double lo = u[0], hi = u[0], n2 = 0;
    for (T v : u) lo = std::min(lo, (double)v), hi = std::max(hi, (double)v), n2 += (double)v * v;
    for (double t : tols) {
        const auto t0 = std::chrono::steady_clock::now();
        const T *x = r.progressive_reconstruct(t);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!x) fatal("reconstruction failed");
        double linf = 0, e2 = 0;
        for (size_t i = 0; i < u.size(); i++) {
            const double d = (double)u[i] - x[i];
            linf = std::max(linf, std::fabs(d));
            e2 += d * d;
        }
        std::printf("tolerance %.3e: reconstruct %.6fs, retrieved %zu bytes, max error %.6e (rel %.3e), "
                    "L2 rel %.6e, PSNR %.2f\n",
                    t, s, r.get_retrieved_size(), linf, linf / (hi - lo), std::sqrt(e2 / n2),
                    20 * std::log10((hi - lo) / std::sqrt(e2 / u.size())));
    }
}

template <class T>
static int run(const std::string &file, const std::vector<double> &tols, bool l2, bool hier, bool perbit,
               const std::string &dir) {
    const int enc = perbit ? PRODM_ENC_PERBIT : -1;
    size_t n = 0;
    const auto u = read_raw<T>(file, n);
    size_t msize = 0;
    std::vector<std::string> files;
    {
        ConcatLevelFileRetriever probe(dir + "/metadata.bin", {});
        uint8_t *m = probe.load_metadata(&msize);
        const int nd = m[0];
        const int nl = m[nd * sizeof(uint32_t) + 1];
        std::free(m);
        for (int l = 0; l < nl; l++) files.push_back(dir + "/level_" + std::to_string(l) + ".bin");
    }
    ConcatLevelFileRetriever retriever(dir + "/metadata.bin", files);
    if (hier) {
        using E = MaxErrorEstimatorHB<T>;
        GpuReconstructor<T, AdaptiveLevelCompressor, SignExcludeGreedyBasedSizeInterpreter<E>, E,
                         ConcatLevelFileRetriever>
            r(AdaptiveLevelCompressor(64), SignExcludeGreedyBasedSizeInterpreter<E>(E()), retriever,
              PRODM_DECOMP_HIERARCHICAL, enc);
        r.load_metadata();
        evaluate(u, tols, r);
    } else if (l2) {
        using E = SNormErrorEstimator<T>;
        const int L = (int)files.size() - 1;
        GpuReconstructor<T, AdaptiveLevelCompressor, NegaBinaryGreedyBasedSizeInterpreter<E>, E, ConcatLevelFileRetriever>
            r(AdaptiveLevelCompressor(64), NegaBinaryGreedyBasedSizeInterpreter<E>(E(3, L, 0)), retriever,
              PRODM_DECOMP_ORTHOGONAL, enc);
        r.load_metadata();
        evaluate(u, tols, r);
    } else {
        using E = MaxErrorEstimatorOB<T>;
        GpuReconstructor<T, AdaptiveLevelCompressor, GreedyBasedSizeInterpreter<E>, E, ConcatLevelFileRetriever> r(
            AdaptiveLevelCompressor(64), GreedyBasedSizeInterpreter<E>(E(3)), retriever, PRODM_DECOMP_ORTHOGONAL, enc);
        r.load_metadata();
        evaluate(u, tols, r);
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <original file> <num tolerances> <tolerances...> [--f64] [--l2] [--in DIR]\n",
                     argv[0]);
        return 2;
    }
    int a = 1;
    const std::string file = argv[a++];
    const int nt = std::atoi(argv[a++]);
    std::vector<double> tols;
    for (int i = 0; i < nt && a < argc; i++) tols.push_back(std::atof(argv[a++]));
    bool f64 = false, l2 = false, hier = false, perbit = false;
    std::string dir = "refactored_data";
    for (; a < argc; a++) {
        const std::string o = argv[a];
        if (o == "--f64") f64 = true;
        else if (o == "--l2") l2 = true;
        else if (o == "--hierarchical") hier = true;
        else if (o == "--perbit") perbit = true;
        else if (o == "--in" && a + 1 < argc) dir = argv[++a];
        else fatal("unknown option " + o);
    }
    return f64 ? run<double>(file, tols, l2, hier, perbit, dir) : run<float>(file, tols, l2, hier, perbit, dir);
}

// interpreters.hpp -- retrieval planners ("size interpreters", host C++).
//
// Counterparts of ProDM's include/MDR/SizeInterpreter/GreedyBasedSizeInterpreter.hpp:23-250
// and BasicSizeInterpreter.hpp:11-112. Given, per level, the byte size of each
// bitplane and the error table (entry k = error after k planes), and the
// number of planes each level already holds (`index`, updated in place), a
// planner returns how many more bytes to read per level so that the
// estimated error falls below `tolerance`.
//
// The selection order, tie breaking (a binary heap keyed on gain per byte)
// and the T-precision estimator arithmetic follow ProDM's planners, so the
// same tables give the same plan. Two guards differ where ProDM reads past
// the table: a level whose planes are all held already is never offered.
#ifndef PRODM_INTERPRETERS_HPP
#define PRODM_INTERPRETERS_HPP

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <queue>
#include <vector>

namespace prodm {

using LevelSizes = std::vector<std::vector<uint32_t>>;
using LevelErrors = std::vector<std::vector<double>>;

namespace detail {
struct Gain {
    double per_byte;
    int level;
    int planes;  // consecutive planes taken together (negabinary planner)
};
struct GainLess {
    bool operator()(const Gain &a, const Gain &b) const { return a.per_byte < b.per_byte; }
};
using GainHeap = std::priority_queue<Gain, std::vector<Gain>, GainLess>;

template <class E>
double total_error(const E &est, const LevelErrors &err, const std::vector<uint8_t> &index) {
    double acc = 0;
    for (size_t l = 0; l < err.size(); l++) acc += est.estimate_error(err[l][index[l]], (int)l);
    return acc;
}
}  // namespace detail

// Greedy on error gain per byte (GreedyBasedSizeInterpreter.hpp:23-70).
template <class E>
class GreedyBasedSizeInterpreter {
  public:
    explicit GreedyBasedSizeInterpreter(const E &e) : est_(e) {}
    std::vector<uint32_t> interpret_retrieve_size(const LevelSizes &sizes, const LevelErrors &err, double tolerance,
                                                  std::vector<uint8_t> &index) const {
        const int L = (int)sizes.size();
        std::vector<uint32_t> out(L, 0);
        double acc = detail::total_error(est_, err, index);
        detail::GainHeap heap;
        for (int l = 0; l < L; l++)
            if (index[l] < sizes[l].size()) heap.push(offer(acc, err, sizes, index, l));
        while (!heap.empty()) {
            const int l = heap.top().level;
            heap.pop();
            const int j = index[l];
            out[l] += sizes[l][j];
            acc -= est_.estimate_error(err[l][j], l);
            acc += est_.estimate_error(err[l][j + 1], l);
            index[l]++;
            const bool met = acc < tolerance;
            if (index[l] < sizes[l].size()) heap.push(offer(acc, err, sizes, index, l));
            if (met) break;
        }
        return out;
    }
    void print() const { std::printf("greedy size interpreter\n"); }

  private:
    detail::Gain offer(double acc, const LevelErrors &err, const LevelSizes &sizes, const std::vector<uint8_t> &index,
                       int l) const {
        const int j = index[l];
        const double g = est_.estimate_error_gain(acc, err[l][j], err[l][j + 1], l);
        return {g / sizes[l][j], l, 1};
    }
    E est_;
};

// Greedy that first takes plane 0 (the sign-carrying plane) of every level
// still at index 0, stopping at the first prefix of levels whose best case
// meets the tolerance (GreedyBasedSizeInterpreter.hpp:72-148).
template <class E>
class SignExcludeGreedyBasedSizeInterpreter {
  public:
    explicit SignExcludeGreedyBasedSizeInterpreter(const E &e) : est_(e) {}
    std::vector<uint32_t> interpret_retrieve_size(const LevelSizes &sizes, const LevelErrors &err, double tolerance,
                                                  std::vector<uint8_t> &index) const {
        const int L = (int)sizes.size();
        std::vector<uint32_t> out(L, 0);
        double acc = detail::total_error(est_, err, index);
        if (acc < tolerance) return out;
        detail::GainHeap heap;
        double best = acc;  // error with every plane of the levels seen so far
        for (int l = 0; l < L; l++) {
            best -= est_.estimate_error(err[l][index[l]], l);
            best += est_.estimate_error(err[l].back(), l);
            if (index[l] == 0 && !sizes[l].empty()) {
                out[l] += sizes[l][0];
                acc -= est_.estimate_error(err[l][0], l);
                acc += est_.estimate_error(err[l][1], l);
                index[l]++;
            }
            if (index[l] < sizes[l].size()) {
                const int j = index[l];
                heap.push({est_.estimate_error_gain(acc, err[l][j], err[l][j + 1], l) / sizes[l][j], l, 1});
            }
            if (best < tolerance) break;
        }
        bool met = acc < tolerance;
        while (!met && !heap.empty()) {
            const int l = heap.top().level;
            heap.pop();
            const int j = index[l];
            out[l] += sizes[l][j];
            acc -= est_.estimate_error(err[l][j], l);
            acc += est_.estimate_error(err[l][j + 1], l);
            met = acc < tolerance;
            index[l]++;
            if (index[l] < sizes[l].size()) {
                const int k = index[l];
                heap.push({est_.estimate_error_gain(acc, err[l][k], err[l][k + 1], l) / sizes[l][k], l, 1});
            }
        }
        return out;
    }
    void print() const { std::printf("sign-excluding greedy size interpreter\n"); }

  private:
    E est_;
};

// Greedy for negabinary planes: an offer may span several consecutive planes
// when that raises the gain per byte, since a single negabinary plane can
// increase the error (GreedyBasedSizeInterpreter.hpp:161-248).
template <class E>
class NegaBinaryGreedyBasedSizeInterpreter {
  public:
    explicit NegaBinaryGreedyBasedSizeInterpreter(const E &e) : est_(e) {}
    std::vector<uint32_t> interpret_retrieve_size(const LevelSizes &sizes, const LevelErrors &err, double tolerance,
                                                  std::vector<uint8_t> &index) const {
        const int L = (int)sizes.size();
        std::vector<uint32_t> out(L, 0);
        double acc = detail::total_error(est_, err, index);
        detail::GainHeap heap;
        for (int l = 0; l < L; l++)
            if (index[l] < sizes[l].size()) heap.push(offer(acc, index[l], l, err[l], sizes[l]));
        bool met = acc < tolerance;
        while (!met && !heap.empty()) {
            const detail::Gain top = heap.top();
            heap.pop();
            const int l = top.level, j = index[l], k = top.planes;
            for (int q = 0; q < k; q++) out[l] += sizes[l][j + q];
            acc -= est_.estimate_error(err[l][j], l);
            acc += est_.estimate_error(err[l][j + k], l);
            met = acc < tolerance;
            index[l] += k;
            if (index[l] < sizes[l].size()) heap.push(offer(acc, index[l], l, err[l], sizes[l]));
        }
        return out;
    }
    void print() const { std::printf("negabinary greedy size interpreter\n"); }

  private:
    // Extend the run j, j+1, ... while the gain per byte does not drop (or is
    // not yet positive); runs stop one plane short of the last plane, as in
    // ProDM's planner.
    detail::Gain offer(double acc, int j, int l, const std::vector<double> &e, const std::vector<uint32_t> &s) const {
        double gain = est_.estimate_error_gain(acc, e[j], e[j + 1], l);
        uint64_t bytes = s[j];
        double eff = gain / (uint32_t)bytes;
        int run = 1;
        const int left = (int)s.size() - j;
        for (int q = 2; q < left; q++) {
            const double g = est_.estimate_error_gain(acc, e[j], e[j + q], l);
            const uint32_t b = (uint32_t)bytes + s[j + q - 1];
            const double f = g / b;
            if (eff > 0 && eff > f) break;
            gain = g;
            eff = f;
            bytes = b;
            run = q;
        }
        return {eff, l, run};
    }
    E est_;
};

// Levels in order, planes in order (BasicSizeInterpreter.hpp:11-44).
template <class E>
class InorderSizeInterpreter {
  public:
    explicit InorderSizeInterpreter(const E &e) : est_(e) {}
    std::vector<uint32_t> interpret_retrieve_size(const LevelSizes &sizes, const LevelErrors &err, double tolerance,
                                                  std::vector<uint8_t> &index) const {
        const int L = (int)sizes.size();
        std::vector<uint32_t> out(L, 0);
        double acc = detail::total_error(est_, err, index);
        for (int l = 0; l < L; l++) {
            for (int j = index[l]; j < (int)sizes[l].size(); j++) {
                out[l] += sizes[l][j];
                index[l]++;
                acc -= est_.estimate_error(err[l][j], l);
                acc += est_.estimate_error(err[l][j + 1], l);
                if (acc < tolerance) return out;
            }
        }
        return out;
    }
    void print() const { std::printf("in-order size interpreter\n"); }

  private:
    E est_;
};

// One plane per level per round, resuming at the first level that is behind
// its predecessor (BasicSizeInterpreter.hpp:46-110). ProDM's round loop
// reuses its counter for the level's plane index; the number of rounds it
// runs is reproduced here as `round`.
template <class E>
class RoundRobinSizeInterpreter {
  public:
    explicit RoundRobinSizeInterpreter(const E &e) : est_(e) {}
    std::vector<uint32_t> interpret_retrieve_size(const LevelSizes &sizes, const LevelErrors &err, double tolerance,
                                                  std::vector<uint8_t> &index) const {
        const int L = (int)sizes.size();
        std::vector<uint32_t> out(L, 0);
        double acc = detail::total_error(est_, err, index);
        int widest = 0;
        for (int l = 0; l < L; l++) widest = std::max(widest, (int)sizes[l].size());
        int first = 0;
        for (int l = 1; l < L; l++)
            if (index[l] < index[l - 1]) {
                first = l;
                break;
            }
        auto take = [&](int l) {
            const int j = index[l];
            out[l] += sizes[l][j];
            index[l]++;
            acc -= est_.estimate_error(err[l][j], l);
            acc += est_.estimate_error(err[l][j + 1], l);
            return acc < tolerance;
        };
        for (int l = first; l < L; l++)
            if (index[l] < sizes[l].size() && take(l)) return out;
        for (int round = 0; round < widest; round++) {
            for (int l = 0; l < L; l++) {
                round = index[l];
                if (index[l] >= sizes[l].size()) continue;
                if (take(l)) return out;
            }
        }
        return out;
    }
    void print() const { std::printf("round-robin size interpreter\n"); }

  private:
    E est_;
};

}  // namespace prodm
#endif

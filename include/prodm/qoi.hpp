// qoi.hpp -- certified L-infinity amplification of the recompose operator and
// retrieval planning on a quantity of interest (host C++).
//
// Recompose maps the stored coefficients c to the field u = R c. Its level-l
// block is R_l = I^(L-l) S_l: S_l reconstructs the level-l grid from the
// level-l coefficients (coarse values zero), I interpolates one level up.
// For the orthogonal basis S_l = (Id - Q) on the coefficient nodes, with
// Q = I M_c^-1 I^T M_f the L2 projection onto the coarser grid, a Kronecker
// product of 1-D projections P_d (the correction is applied one dimension at
// a time). I has non-negative rows that sum to 1, and a level-l grid node
// reproduces its own value, so
//     A_l = max_x sum_(j in level l) |R(x, j)| = max_y s_l(y),
//     s_l(y) = sum_(j coefficient node) |delta(y, j) - prod_d P_d(y_d, j_d)|,
// and s_l(y) has a closed form in three 1-D row sums of each P_d (all
// columns, even columns, diagonal). A_l is exact (the largest row 1-norm of
// the level-l block over every node), not a sampled estimate: a retrieval
// whose level-l coefficient errors are at most t_l satisfies
// |u - u~| <= sum_l A_l t_l at every node, in exact arithmetic.
// Hierarchical basis: S_l = identity on the coefficient nodes, A_l = 1.
// The coarsest level passes through interpolation only: A_0 = 1.
//
// ProDM's L-infinity estimator for the orthogonal basis uses one constant
// per dimension count (MaxErrorEstimatorOB, ErrorEstimator/MaxErrorEstimator.hpp:13-40)
// and multiplies it by 4; AmplificationMaxErrorEstimator uses A_l instead.
// The QoI planner follows the error composition of ProDM's GE application
// (include/qoi_utils.hpp:18-41, app/GE/sequential/halving_Vtot.cpp:31-84).
#ifndef PRODM_QOI_HPP
#define PRODM_QOI_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <vector>

#include "estimators.hpp"
#include "interpreters.hpp"

namespace prodm {

namespace detail {
// Row sums of the 1-D L2 projection P = I (I^T M I)^-1 I^T M onto the coarse
// grid of n_f = 2m + 1 fine nodes (unit spacing; it cancels): per fine row i,
// sum_j |P(i,j)|, sum_(j even) |P(i,j)| and P(i,i). Row i of P is
// (M I M_c^-1 I^T e_i)^T: one tridiagonal solve per row, O(n_f n_c) in all.
struct ProjectionRows {
    std::vector<double> all, even, diag;
};
inline ProjectionRows projection_rows(uint32_t nf) {
    ProjectionRows r;
    if (nf == 1) {  // a dimension that is not refined: P = identity
        r.all = r.even = r.diag = {1.0};
        return r;
    }
    if (nf % 2 == 0) throw std::invalid_argument("projection_rows: even extent");
    const uint32_t nc = (nf + 1) / 2;
    // coarse mass matrix I^T M I (spacing 2): diagonal 2/3 at the ends, 4/3
    // inside, off-diagonal 1/3; Thomas factors once
    std::vector<double> cp(nc), dinv(nc);
    {
        double prev = 0;
        for (uint32_t k = 0; k < nc; k++) {
            const double a = (k == 0 || k == nc - 1) ? 2.0 / 3 : 4.0 / 3, b = 1.0 / 3;
            const double d = a - (k ? b * prev : 0.0);
            dinv[k] = 1.0 / d;
            prev = b / d;
            cp[k] = prev;
        }
    }
    std::vector<double> v(nc), y(nf), row(nf);
    r.all.resize(nf);
    r.even.resize(nf);
    r.diag.resize(nf);
    for (uint32_t i = 0; i < nf; i++) {
        std::fill(v.begin(), v.end(), 0.0);
        if (i % 2 == 0) v[i / 2] = 1.0;
        else v[i / 2] = v[i / 2 + 1] = 0.5;
        for (uint32_t k = 0; k < nc; k++) v[k] = (v[k] - (k ? (1.0 / 3) * v[k - 1] : 0.0)) * dinv[k];
        for (uint32_t k = nc - 1; k-- > 0;) v[k] -= cp[k] * v[k + 1];
        for (uint32_t j = 0; j < nf; j++) y[j] = (j % 2 == 0) ? v[j / 2] : 0.5 * (v[j / 2] + v[j / 2 + 1]);
        double a = 0, e = 0;
        for (uint32_t j = 0; j < nf; j++) {  // fine mass (unit spacing): 1/3, 2/3 inside, 1/6 off
            const double dj = (j == 0 || j == nf - 1) ? 1.0 / 3 : 2.0 / 3;
            double s = dj * y[j];
            if (j > 0) s += y[j - 1] / 6;
            if (j + 1 < nf) s += y[j + 1] / 6;
            row[j] = s;
            a += std::fabs(s);
            if (j % 2 == 0) e += std::fabs(s);
        }
        r.all[i] = a;
        r.even[i] = e;
        r.diag[i] = row[i];
    }
    return r;
}
}  // namespace detail

// A_l for l = 0..L of a field with `dims` (dims[0] slowest, 1 to 3 dims)
// decomposed over L levels; hierarchical: the basis without correction.
inline std::vector<double> recompose_amplification(const std::vector<uint32_t> &dims, int L, bool hierarchical) {
    if (dims.empty() || dims.size() > 3 || L < 0) throw std::invalid_argument("recompose_amplification: 1 to 3 dims");
    std::vector<double> A(L + 1, 1.0);
    std::vector<uint32_t> n(3, 1);
    for (size_t d = 0; d < dims.size(); d++) n[3 - dims.size() + d] = dims[d];
    std::vector<std::vector<uint32_t>> lv(L + 1);
    lv[L] = n;
    for (int l = L; l > 0; l--) {
        lv[l - 1] = lv[l];
        for (auto &v : lv[l - 1]) v = v > 1 ? (v >> 1) + 1 : 1;
    }
    if (hierarchical) return A;
    for (int l = 1; l <= L; l++) {
        const auto &g = lv[l];
        detail::ProjectionRows p[3];
        for (int d = 0; d < 3; d++) p[d] = detail::projection_rows(g[d]);
        double best = 0;
        for (uint32_t a = 0; a < g[0]; a++)
            for (uint32_t b = 0; b < g[1]; b++) {
                const double A01 = p[0].all[a] * p[1].all[b], B01 = p[0].even[a] * p[1].even[b];
                const double P01 = p[0].diag[a] * p[1].diag[b];
                const bool ev01 = (g[0] == 1 || a % 2 == 0) && (g[1] == 1 || b % 2 == 0);
                for (uint32_t c = 0; c < g[2]; c++) {
                    const double pp = P01 * p[2].diag[c];
                    const double self = std::fabs(1.0 - pp) - std::fabs(pp);
                    double s = A01 * p[2].all[c] + self - B01 * p[2].even[c];
                    if (ev01 && (g[2] == 1 || c % 2 == 0)) s -= self;  // y is not a coefficient column
                    best = std::max(best, s);
                }
            }
        A[l] = best * (1.0 + 1e-12);  // FP64 rounding of the sums
    }
    return A;
}

// L-infinity estimator with the exact per-level amplification: the error of
// the field is at most sum_l A_l t_l when the level tables t_l bound each
// coefficient's error (MaxErrorCollector's tables do for the negabinary
// streams with the x4 bound and for GroupedBP/PerBit with the plain bound;
// the decoder's float conversion adds at most 2^-21 of the level maximum).
// The arithmetic is in T, as ProDM's estimators.
template <class T>
class AmplificationMaxErrorEstimator : public MaxErrorEstimator<T> {
  public:
    AmplificationMaxErrorEstimator() = default;
    explicit AmplificationMaxErrorEstimator(const std::vector<double> &amplification) {
        for (double a : amplification) {  // rounded up, so the estimate stays a bound
            T t = (T)a;
            if ((double)t < a) t = std::nextafter(t, std::numeric_limits<T>::infinity());
            a_.push_back(t);
        }
    }
    AmplificationMaxErrorEstimator(const std::vector<uint32_t> &dims, int target_level, bool hierarchical = false)
        : AmplificationMaxErrorEstimator(recompose_amplification(dims, target_level, hierarchical)) {}
    T estimate_error(T error, int level) const { return a_[level] * error; }
    T estimate_error(T data, T reconstructed, int level) const { return a_[level] * (data - reconstructed); }
    T estimate_error_gain(T, T current, T next, int level) const { return a_[level] * (current - next); }
    const std::vector<T> &amplification() const { return a_; }
    void print() const { std::printf("L-infinity error estimator, exact recompose amplification per level\n"); }

  private:
    std::vector<T> a_;
};

// |V| = sqrt(sum_i V_i^2) from per-component bounds E_i at the reconstructed
// values v: e(|V|^2) <= sum_i 2 |v_i| E_i + E_i^2 (compute_bound_x_square),
// then the square root (compute_bound_square_root_x), qoi_utils.hpp:18-41.
struct VelocityMagnitude {
    static double bound(const double *v, const double *E, int ncomp) {
        double v2 = 0, e2 = 0;
        for (int i = 0; i < ncomp; i++) {
            v2 += v[i] * v[i];
            e2 += 2 * std::fabs(v[i]) * E[i] + E[i] * E[i];
        }
        if (v2 == 0) return std::sqrt(e2);
        if (v2 > e2) return e2 / (std::sqrt(v2 - e2) + std::sqrt(v2));
        return e2 / std::sqrt(v2);
    }
};

// Retrieval planning on a QoI of several refactored components. The
// "levels" of interpret_retrieve_size are the (component, level) pairs,
// component-major (component i, level l at i * (L + 1) + l), with the same
// meaning as in ProDM's planners (SizeInterpreter/SizeInterpreterInterface.hpp:13):
// sizes[k] the plane bytes, errors[k] the coefficient error bound after each
// plane count (MaxErrorCollector tables), index[k] the planes held (updated).
// The component error bound is E_i = sum_l A_l t_(i,l)[k] + r_i (r_i: an
// allowance for the decoder's and the recompose's rounding); the
// QoI bound is evaluated at a sample of reconstructed values (set_sample:
// npoints x ncomp, point-major), and planes are added greedily, the largest
// drop of the sampled maximum per byte first, until it is below the
// tolerance. The sample should hold the points where the bound is largest;
// a caller re-evaluates the bound over the whole reconstructed field after
// the retrieval (it holds at every node when that check passes).
template <class QoI = VelocityMagnitude>
class QoIGreedyBasedSizeInterpreter {
  public:
    QoIGreedyBasedSizeInterpreter(int ncomp, const std::vector<double> &amplification,
                                  const std::vector<double> &rounding = {})
        : ncomp_(ncomp), a_(amplification), rounding_(rounding) {
        if (ncomp < 1) throw std::invalid_argument("QoIGreedyBasedSizeInterpreter: no components");
        rounding_.resize(ncomp, 0.0);
    }
    void set_sample(const std::vector<double> &values) {
        if (values.size() % ncomp_) throw std::invalid_argument("QoIGreedyBasedSizeInterpreter: sample size");
        sample_ = values;
    }
    std::vector<uint32_t> interpret_retrieve_size(const LevelSizes &sizes, const LevelErrors &err, double tolerance,
                                                  std::vector<uint8_t> &index) const {
        const int nl = (int)a_.size();
        if ((int)sizes.size() != ncomp_ * nl || err.size() != sizes.size() || index.size() != sizes.size())
            throw std::invalid_argument("QoIGreedyBasedSizeInterpreter: expected components x levels tables");
        std::vector<uint32_t> out(sizes.size(), 0);
        std::vector<double> E(ncomp_);
        auto comp_bound = [&](int i) {
            double e = rounding_[i];
            for (int l = 0; l < nl; l++) e += a_[l] * err[i * nl + l][index[i * nl + l]];
            return e;
        };
        for (int i = 0; i < ncomp_; i++) E[i] = comp_bound(i);
        const size_t np = sample_.size() / ncomp_;
        auto worst = [&](const std::vector<double> &e) {
            double m = 0;
            for (size_t p = 0; p < np; p++) m = std::max(m, QoI::bound(&sample_[p * ncomp_], e.data(), ncomp_));
            return m;
        };
        double cur = worst(E);
        while (cur >= tolerance) {
            int bi = -1, bl = -1;
            double best = -1, bval = cur;
            for (int i = 0; i < ncomp_; i++)
                for (int l = 0; l < nl; l++) {
                    const int k = i * nl + l;
                    if (index[k] >= sizes[k].size()) continue;
                    std::vector<double> e2(E);
                    e2[i] += a_[l] * (err[k][index[k] + 1] - err[k][index[k]]);
                    const double val = worst(e2);
                    const double gain = (cur - val) / std::max<uint32_t>(sizes[k][index[k]], 1u);
                    if (gain > best) best = gain, bi = i, bl = l, bval = val;
                }
            if (bi < 0) break;  // every plane is held
            const int k = bi * nl + bl;
            out[k] += sizes[k][index[k]];
            index[k]++;
            E[bi] = comp_bound(bi);
            cur = bval;
        }
        return out;
    }
    void print() const { std::printf("greedy QoI size interpreter (%d components)\n", ncomp_); }

  private:
    int ncomp_;
    std::vector<double> a_;
    std::vector<double> rounding_;
    std::vector<double> sample_;
};

}  // namespace prodm
#endif

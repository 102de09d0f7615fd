// estimators.hpp -- error estimators of the retrieval planner (host C++).
//
// Counterparts of ProDM's include/MDR/ErrorEstimator/MaxErrorEstimator.hpp:7-68
// and SquaredErrorEstimator.hpp:7-68. An estimator maps a per-level error
// table entry to its contribution to the error of the reconstructed field.
// The arithmetic is carried out in T exactly as in ProDM (the table entry is
// narrowed to T before scaling), so a planner built on these classes picks
// the same bitplanes as ProDM's for the same tables.
#ifndef PRODM_ESTIMATORS_HPP
#define PRODM_ESTIMATORS_HPP

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace prodm {

// Tag bases: ComposedReconstructor selects the L-infinity or the L2 path by
// std::is_base_of on these (ComposedReconstructor.hpp:33,42 in ProDM).
template <class T>
struct MaxErrorEstimator {};
template <class T>
struct SquaredErrorEstimator {};

// L-infinity estimator of the orthogonal (L2-projection) basis: the error of
// the field is at most c times the sum of the per-level coefficient errors,
// with c = 1 + (operator norm of the interpolation/correction) per dimension
// count (MaxErrorEstimator.hpp:15-31), times 4 for the two extra negabinary
// bitplanes.
template <class T>
class MaxErrorEstimatorOB : public MaxErrorEstimator<T> {
  public:
    explicit MaxErrorEstimatorOB(int num_dims = 1) {
        double base;
        if (num_dims == 1) base = 1.0 + std::sqrt(3.0) / 2;
        else if (num_dims == 2) base = 1.0 + 9.0 / 4;
        else if (num_dims == 3) base = 1.0 + 21.0 * std::sqrt(3.0) / 8;
        else throw std::invalid_argument("MaxErrorEstimatorOB: 1 to 3 dimensions");
        c_ = (T)base;
        c_ *= 4;
    }
    T estimate_error(T error, int) const { return c_ * error; }
    T estimate_error(T data, T reconstructed, int) const { return c_ * (data - reconstructed); }
    T estimate_error_gain(T, T current, T next, int) const { return c_ * (current - next); }
    T constant() const { return c_; }
    void print() const { std::printf("L-infinity error estimator, orthogonal basis (c = %g)\n", (double)c_); }

  private:
    T c_ = 0;
};

// L-infinity estimator of the hierarchical basis: c = 1 (MaxErrorEstimator.hpp:50-68).
template <class T>
class MaxErrorEstimatorHB : public MaxErrorEstimator<T> {
  public:
    T estimate_error(T error, int) const { return error; }
    T estimate_error(T data, T reconstructed, int) const { return data - reconstructed; }
    T estimate_error_gain(T, T current, T next, int) const { return current - next; }
    void print() const { std::printf("L-infinity error estimator, hierarchical basis\n"); }
};

// Squared-L2 estimator of the hierarchical basis: level l weighs
// vol(P_l) = 2^(d (L - l)), doubled for (e1 + e2)^2 <= 2 (e1^2 + e2^2)
// (SquaredErrorEstimator.hpp:10-37).
template <class T>
class L2ErrorEstimator_HB : public SquaredErrorEstimator<T> {
  public:
    L2ErrorEstimator_HB(int num_dims = 1, int target_level = 0) : w_(target_level + 1) {
        for (int l = 0; l <= target_level; l++) w_[l] = (T)std::pow(2.0, (double)(num_dims * (target_level - l)));
    }
    T estimate_error(T error, int level) const { return w_[level] * error * 2; }
    T estimate_error(T data, T reconstructed, int level) const { return w_[level] * (data - reconstructed) * 2; }
    T estimate_error_gain(T, T current, T next, int level) const { return w_[level] * (current - next); }
    void print() const { std::printf("squared-L2 error estimator, hierarchical basis\n"); }

  private:
    std::vector<T> w_;
};

// S-norm estimator of the orthogonal basis: level l weighs 2^(2 s l) vol(P_l)
// (SquaredErrorEstimator.hpp:39-65); s = 0 is the squared L2 norm.
template <class T>
class SNormErrorEstimator : public SquaredErrorEstimator<T> {
  public:
    SNormErrorEstimator(int num_dims = 1, int target_level = 0, T s = 0) : s_(s), w_(target_level + 1) {
        for (int l = 0; l <= target_level; l++) {
            const T e = 2 * s * l + (T)(num_dims * (target_level - l));
            w_[l] = (T)std::pow(2.0, (double)e);
        }
    }
    T estimate_error(T error, int level) const { return w_[level] * error; }
    T estimate_error(T data, T reconstructed, int level) const { return w_[level] * (data - reconstructed); }
    T estimate_error_gain(T, T current, T next, int level) const { return w_[level] * (current - next); }
    void print() const { std::printf("S-norm error estimator (s = %g)\n", (double)s_); }

  private:
    T s_ = 0;
    std::vector<T> w_;
};

}  // namespace prodm
#endif

// planner.cpp -- C ABI over the host retrieval planners of
// include/prodm/interpreters.hpp and the estimators of
// include/prodm/estimators.hpp (host code, no device work).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "prodm/estimators.hpp"
#include "prodm/interpreters.hpp"
#include "prodm/qoi.hpp"
#include "prodm_mdr.h"

namespace prodm {
void set_error(const std::string &msg);
}

namespace {

void fail(const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    prodm::set_error(buf);
}

template <class E>
std::vector<uint32_t> run_planner(int planner, const E &est, const prodm::LevelSizes &s, const prodm::LevelErrors &e,
                                  double tol, std::vector<uint8_t> &idx, bool &ok) {
    ok = true;
    switch (planner) {
        case 0: return prodm::GreedyBasedSizeInterpreter<E>(est).interpret_retrieve_size(s, e, tol, idx);
        case 1: return prodm::SignExcludeGreedyBasedSizeInterpreter<E>(est).interpret_retrieve_size(s, e, tol, idx);
        case 2: return prodm::NegaBinaryGreedyBasedSizeInterpreter<E>(est).interpret_retrieve_size(s, e, tol, idx);
        case 3: return prodm::InorderSizeInterpreter<E>(est).interpret_retrieve_size(s, e, tol, idx);
        case 4: return prodm::RoundRobinSizeInterpreter<E>(est).interpret_retrieve_size(s, e, tol, idx);
    }
    ok = false;
    return {};
}

template <class T>
int plan(int planner, int estimator, int ndims, int target_level, double s_param, const prodm::LevelSizes &s,
         const prodm::LevelErrors &e, double tol, std::vector<uint8_t> &idx, std::vector<uint32_t> &out) {
    bool ok = false;
    switch (estimator) {
        case PRODM_EST_MAX_OB: out = run_planner(planner, prodm::MaxErrorEstimatorOB<T>(ndims), s, e, tol, idx, ok); break;
        case PRODM_EST_MAX_HB: out = run_planner(planner, prodm::MaxErrorEstimatorHB<T>(), s, e, tol, idx, ok); break;
        case PRODM_EST_SNORM:
            out = run_planner(planner, prodm::SNormErrorEstimator<T>(ndims, target_level, (T)s_param), s, e, tol, idx, ok);
            break;
        case PRODM_EST_L2_HB:
            out = run_planner(planner, prodm::L2ErrorEstimator_HB<T>(ndims, target_level), s, e, tol, idx, ok);
            break;
        default: fail("unknown estimator %d", estimator); return PRODM_ERR_ARG;
    }
    if (!ok) {
        fail("unknown planner %d", planner);
        return PRODM_ERR_ARG;
    }
    return PRODM_OK;
}

}  // namespace

extern "C" int prodm_interpret_retrieve_size(int planner, int estimator, int dtype, int ndims, int target_level,
                                             double s_param, int nlevels, const int *counts, const uint32_t *sizes,
                                             const double *errors, double tolerance, uint8_t *index,
                                             uint32_t *out_sizes) {
    if (nlevels <= 0 || !counts || !sizes || !errors || !index || !out_sizes) {
        fail("prodm_interpret_retrieve_size: bad arguments");
        return PRODM_ERR_ARG;
    }
    if ((estimator == PRODM_EST_MAX_OB && (ndims < 1 || ndims > 3)) ||
        ((estimator == PRODM_EST_SNORM || estimator == PRODM_EST_L2_HB) && (target_level < nlevels - 1))) {
        fail("prodm_interpret_retrieve_size: estimator needs 1..3 dims and target_level >= nlevels-1");
        return PRODM_ERR_ARG;
    }
    prodm::LevelSizes s(nlevels);
    prodm::LevelErrors e(nlevels);
    std::vector<uint8_t> idx(index, index + nlevels);
    for (int l = 0; l < nlevels; l++) {
        if (counts[l] < 0 || counts[l] > 255 || idx[l] > counts[l]) {
            fail("prodm_interpret_retrieve_size: level %d: bad plane count or index", l);
            return PRODM_ERR_ARG;
        }
        s[l].assign(sizes, sizes + counts[l]);
        e[l].assign(errors, errors + counts[l] + 1);
        sizes += counts[l];
        errors += counts[l] + 1;
    }
    std::vector<uint32_t> out;
    int rc;
    try {
        rc = dtype == PRODM_F64 ? plan<double>(planner, estimator, ndims, target_level, s_param, s, e, tolerance, idx, out)
                                : plan<float>(planner, estimator, ndims, target_level, s_param, s, e, tolerance, idx, out);
    } catch (const std::exception &ex) {
        fail("prodm_interpret_retrieve_size: %s", ex.what());
        return PRODM_ERR_ARG;
    }
    if (rc != PRODM_OK) return rc;
    std::memcpy(index, idx.data(), nlevels);
    std::memcpy(out_sizes, out.data(), sizeof(uint32_t) * nlevels);
    return PRODM_OK;
}

extern "C" double prodm_max_estimator_ob_constant(int ndims, int dtype) {
    if (ndims < 1 || ndims > 3) return 0.0;
    return dtype == PRODM_F64 ? prodm::MaxErrorEstimatorOB<double>(ndims).constant()
                              : (double)prodm::MaxErrorEstimatorOB<float>(ndims).constant();
}

extern "C" int prodm_recompose_amplification(const uint32_t *dims, int ndims, int target_level, int decomposer,
                                             double *out) {
    if (!dims || !out || ndims < 1 || ndims > 3 || target_level < 0 ||
        (decomposer != PRODM_DECOMP_ORTHOGONAL && decomposer != PRODM_DECOMP_HIERARCHICAL)) {
        fail("prodm_recompose_amplification: bad arguments");
        return PRODM_ERR_ARG;
    }
    std::vector<uint32_t> flat((target_level + 1) * ndims);
    std::vector<uint64_t> el(target_level + 1);
    if (!prodm_level_geometry(dims, ndims, target_level, flat.data(), el.data())) {
        fail("prodm_recompose_amplification: dims do not stay odd down to level 0");
        return PRODM_ERR_GEOMETRY;
    }
    try {
        const auto a = prodm::recompose_amplification(std::vector<uint32_t>(dims, dims + ndims), target_level,
                                                      decomposer == PRODM_DECOMP_HIERARCHICAL);
        std::memcpy(out, a.data(), sizeof(double) * a.size());
    } catch (const std::exception &ex) {
        fail("prodm_recompose_amplification: %s", ex.what());
        return PRODM_ERR_ARG;
    }
    return PRODM_OK;
}

extern "C" int prodm_qoi_interpret_retrieve_size(int qoi, int ncomp, int nlevels, const double *amplification,
                                                 const double *rounding, const int *counts, const uint32_t *sizes,
                                                 const double *errors, const double *sample, uint64_t npoints,
                                                 double tolerance, uint8_t *index, uint32_t *out_sizes) {
    if (qoi != PRODM_QOI_VELOCITY_MAGNITUDE || ncomp < 1 || nlevels < 1 || !amplification || !counts || !sizes ||
        !errors || (npoints && !sample) || !index || !out_sizes) {
        fail("prodm_qoi_interpret_retrieve_size: bad arguments");
        return PRODM_ERR_ARG;
    }
    const int nk = ncomp * nlevels;
    prodm::LevelSizes s(nk);
    prodm::LevelErrors e(nk);
    std::vector<uint8_t> idx(index, index + nk);
    for (int k = 0; k < nk; k++) {
        if (counts[k] < 0 || counts[k] > 255 || idx[k] > counts[k]) {
            fail("prodm_qoi_interpret_retrieve_size: entry %d: bad plane count or index", k);
            return PRODM_ERR_ARG;
        }
        s[k].assign(sizes, sizes + counts[k]);
        e[k].assign(errors, errors + counts[k] + 1);
        sizes += counts[k];
        errors += counts[k] + 1;
    }
    std::vector<uint32_t> out;
    try {
        prodm::QoIGreedyBasedSizeInterpreter<prodm::VelocityMagnitude> q(
            ncomp, std::vector<double>(amplification, amplification + nlevels),
            rounding ? std::vector<double>(rounding, rounding + ncomp) : std::vector<double>());
        q.set_sample(std::vector<double>(sample, sample + npoints * ncomp));
        out = q.interpret_retrieve_size(s, e, tolerance, idx);
    } catch (const std::exception &ex) {
        fail("prodm_qoi_interpret_retrieve_size: %s", ex.what());
        return PRODM_ERR_ARG;
    }
    std::memcpy(index, idx.data(), nk);
    std::memcpy(out_sizes, out.data(), sizeof(uint32_t) * nk);
    return PRODM_OK;
}

// plan.hpp -- the device-resident refactor / retrieve plan behind the
// prodm_mdr_plan handle of the C ABI (one field: geometry, level buffers,
// plane storage, max keys, MGARD workspace and the GroupedBP state).
#pragma once

#include "internal.hpp"
#include "prodm_mdr.h"

struct prodm_mdr_plan {
    prodm::LevelGeom g;
    int dtype = PRODM_F32, encoder = PRODM_ENC_NEGABINARY, nb = 32;
    size_t elem = 4;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    void *levels = nullptr;          // interleaved level buffers
    uint64_t level_off[33] = {0};    // element offsets
    void *keys = nullptr;            // L+1 max keys
    void *planes = nullptr;          // plane storage
    uint64_t plane_off[33] = {0};    // word offsets per level
    uint64_t plane_stride[33] = {0}; // words per plane
    uint8_t *rb = nullptr;           // GroupedBP recording planes
    uint64_t rb_off[33] = {0};
    uint32_t *words = nullptr;       // GroupedBP words / PerBit bytes per plane, (L+1) * nb
    void *signs = nullptr;           // GroupedBP decode signs
    void *gbws = nullptr;
    size_t gbws_bytes = 0;
    uint64_t device_bytes = 0;
    prodm::StageTimer timer;
    int hier = 0;  // PRODM_DECOMP_HIERARCHICAL
    // captured graphs: refactor of one input pointer, retrieve of one
    // (planes per level, output pointer) pair
    hipGraphExec_t g_ref = nullptr, g_ret = nullptr;
    const void *g_ref_in = nullptr;
    void *g_ret_out = nullptr;
    uint8_t g_ret_k[33] = {0};
    // second stream: the encode of a level overlaps the decomposition of the
    // coarser ones, and the decode of a level the recomposition of the finer
    // ones (events per level, created on first use)
    hipStream_t side = nullptr;
    hipStream_t side2 = nullptr;  // retrieve: the finest level's decode and load sweep
    hipEvent_t lev_ev[33] = {};
    hipEvent_t join_ev = nullptr;
    bool capturing = false;  // a graph capture is in progress (the PerBit retrieve has no graph form)
    // per-level load vectors of the retrieve pipeline (allocated on first use)
    void *flev = nullptr;
    uint64_t flev_off[33] = {0};
};

namespace prodm {
// Captures the launches that body() issues on stream s (not the default
// stream) into an executable graph, replacing *exec.
template <class F> inline int graph_capture(hipStream_t s, hipGraphExec_t *exec, F body) {
    if (!s) { set_error("graph capture needs a non-default stream"); return kErrArg; }
    if (*exec) { (void)hipGraphExecDestroy(*exec); *exec = nullptr; }
    PRODM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { set_error(std::string("graph capture: ") + hipGetErrorString(e)); return kErrHip; }
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) { *exec = nullptr; set_error(std::string("graph instantiate: ") + hipGetErrorString(ei)); return kErrHip; }
    return kOk;
}
}  // namespace prodm

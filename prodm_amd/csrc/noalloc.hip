// noalloc.hip -- prodm_grouped_encode_noalloc_f32: the GroupedBP encode of
// prodm_grouped_encode_f32 without a workspace argument and without any
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The encode's kernels (k_gb_rb, k_gb_scan, k_gb_write) keep per-workgroup
// recording-plane histograms and the scanned plane offsets in a scratch
// buffer. Here that buffer is `scratch`, a file-scope device array of
// kScratchBytes that every GPU the library runs on holds until the process
// ends. Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and passed to the kernels
// as the workspace; the host-side name `scratch` is never used as an address.
// Each call overwrites the part it uses before reading it (every histogram
// and offset is written by the call itself), so nothing needs clearing.
// Calls that overlap on one GPU (two streams at once) share the array and
// race; serialize them.
#include <mutex>
#include <string>

#include "internal.hpp"
#include "prodm_mdr.h"

namespace prodm {
namespace {

constexpr size_t kScratchBytes = size_t(8) << 20;  // 8 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on first use.
int scratch_address(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("grouped encode (no allocation): no current HIP device");
        return kErrHip;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("grouped encode (no allocation): hipGetSymbolAddress: ") + hipGetErrorString(e));
            return kErrHip;
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return kOk;
}

}  // namespace
}  // namespace prodm

using namespace prodm;

extern "C" int prodm_grouped_encode_noalloc_f32(const float *d_x, uint32_t n, int32_t exp, int num_bitplanes,
                                                int compat, uint8_t *d_rb, uint32_t *d_planes,
                                                uint64_t plane_stride_words, uint32_t *d_words,
                                                prodm_stream_t stream) {
    if (n == 0 || num_bitplanes < 1 || num_bitplanes > 32) {
        set_error("grouped encode (no allocation): bad n or num_bitplanes");
        return kErrArg;
    }
    const size_t need = gb_workspace_bytes<float>(n, num_bitplanes);
    if (need > kScratchBytes) {
        set_error("grouped encode (no allocation): " + std::to_string(n) + " elements with " +
                  std::to_string(num_bitplanes) + " planes need " + std::to_string(need) +
                  " bytes of scratch, above its " + std::to_string(kScratchBytes) + " bytes");
        return kErrUnsupported;
    }
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != kOk) return rc;
    return launch_gb_encode<float>(d_x, n, exp, nullptr, num_bitplanes, compat, d_rb, d_planes, plane_stride_words,
                                   d_words, ws, kScratchBytes, (hipStream_t)stream);
}

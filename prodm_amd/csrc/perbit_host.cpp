// perbit_host.cpp -- progressive decode of PerBit streams
// (PerBitBPEncoder.hpp:290-382), host code.
//
// An element takes one bit of a stream, or two where its first 1 bit is
// (the bit, then its sign), and the streams carry no index of where each
// element's bits begin: the position of element e in stream i depends on
// the bits of every element before it. Each stream is therefore parsed
// from its start, element by element, on the host; the encoder's GPU pass
// (bitplane.hip) is where the throughput of this codec lies.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "prodm_mdr.h"

namespace prodm {
void set_error(const std::string &msg);
}

// Per-level sign state, keyed by `level` (the reference keys it by call
// order, PerBitBPEncoder.hpp:304-309; the explicit key also serves levels
// first decoded with zero planes).
struct prodm_perbit_state {
    struct Level {
        uint32_t n = 0;
        std::vector<uint8_t> sign, flag;  // sign recorded, sign known
    };
    std::map<int, Level> levels;
};

namespace {

int fail(const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    prodm::set_error(buf);
    return -1;
}

// LSB-first bit reader over 64-bit words (PerBitBPEncoder.hpp:41-67),
// bounded by the stream's size when one is given
struct BitReader {
    const uint64_t *w = nullptr;
    uint64_t nwords = ~0ull, next = 0, buf = 0;
    int left = 0;
    bool overrun = false;
    int get() {
        if (!left) {
            if (next >= nwords) {
                overrun = true;
                return 0;
            }
            std::memcpy(&buf, reinterpret_cast<const char *>(w) + 8 * next, 8);  // any alignment
            next++;
            left = 64;
        }
        const int b = (int)(buf & 1u);
        buf >>= 1;
        left--;
        return b;
    }
};

template <class T>
int perbit_decode(prodm_perbit_state *s, int level, const uint64_t *const *streams, const uint32_t *sizes, uint32_t n,
                  int32_t exp, int start, int k, T *out) {
    using F = typename std::conditional<std::is_same<T, double>::value, uint64_t, uint32_t>::type;
    if (!s || (n && !out)) return fail("perbit decode: null state or output");
    if (k == 0) {  // (:296-299)
        for (uint32_t e = 0; e < n; e++) out[e] = 0;
        return 0;
    }
    if (k < 0 || start < 0 || start + k > (int)(8 * sizeof(F))) return fail("perbit decode: bad plane range");
    if (!streams) return fail("perbit decode: null stream array");
    for (int i = 0; i < k; i++)
        if (!streams[i] && n) return fail("perbit decode: null stream %d", i);
    prodm_perbit_state::Level &L = s->levels[level];
    if (L.sign.empty() && L.n == 0) {
        L.n = n;
        L.sign.assign(n, 0);
        L.flag.assign(n, 0);
    } else if (L.n != n) {
        return fail("perbit decode: level %d has %u elements, not %u", level, L.n, n);
    }
    std::vector<BitReader> rd(k);
    for (int i = 0; i < k; i++) {
        rd[i].w = streams[i];
        if (sizes) rd[i].nwords = sizes[i] / 8;
    }
    const int end = start + k;
    for (uint32_t e = 0; e < n; e++) {
        F fp = 0;
        int sign;
        if (L.flag[e]) {
            sign = L.sign[e];
            for (int i = 0; i < k; i++) fp |= (F)rd[i].get() << (k - 1 - i);
        } else {
            sign = 0;
            for (int i = 0; i < k; i++) {
                const int bit = rd[i].get();
                fp |= (F)bit << (k - 1 - i);
                if (bit && !L.flag[e]) {
                    sign = rd[i].get();
                    L.flag[e] = 1;
                }
            }
            L.sign[e] = (uint8_t)sign;
        }
        const T y = std::ldexp((T)fp, exp - end);
        out[e] = sign ? -y : y;
    }
    for (int i = 0; i < k; i++)
        if (rd[i].overrun) return fail("perbit decode: stream %d is shorter than its content", i);
    return 0;
}

}  // namespace

extern "C" {

int prodm_perbit_state_create(prodm_perbit_state **out) {
    if (!out) return fail("perbit state: null output");
    *out = new prodm_perbit_state();
    return 0;
}
int prodm_perbit_state_destroy(prodm_perbit_state *s) {
    delete s;
    return 0;
}
int prodm_perbit_decode_f32(prodm_perbit_state *s, int level, const uint64_t *const *h_streams,
                            const uint32_t *h_sizes, uint32_t n, int32_t exp, int start, int k, float *h_out) {
    return perbit_decode<float>(s, level, h_streams, h_sizes, n, exp, start, k, h_out);
}
int prodm_perbit_decode_f64(prodm_perbit_state *s, int level, const uint64_t *const *h_streams,
                            const uint32_t *h_sizes, uint32_t n, int32_t exp, int start, int k, double *h_out) {
    return perbit_decode<double>(s, level, h_streams, h_sizes, n, exp, start, k, h_out);
}

}  // extern "C"

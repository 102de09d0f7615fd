// batched.hip -- refactor of many fields of uneven shapes with grouped launches.
//
// ProDM refactors one field per ComposedRefactor::refactor call
// (ComposedRefactor.hpp:78-141); a batch of fields there is a loop of such
// calls. Here a batch owns one plan per field (the same buffers, so every
// plan accessor and prodm_mdr_retrieve work on a batch field) and runs the
// GPU stages of all fields together: one grouped launch per kernel and level
// step (MGARD load sweep, the line solves of dims 2 / 1 / 0, the coarsest
// grids), one grouped NegaBinary encode over every level of every field, and
// one zero-fill of all level keys. A launch walks a table of (field, tile)
// entries, largest fields first; its workgroup count and tile / block shapes
// come from an autotuner (timed per launch, cached in a text file). The
// whole sequence can be captured into one HIP graph. No shape knob changes
// the arithmetic of any tile, so each field's planes and level maxima are
// bit-identical to its single-field plan refactor.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"
#include "plan.hpp"
#include "prodm_mdr.h"

using namespace prodm;

struct prodm_mdr_batch {
    int dtype = PRODM_F32, encoder = PRODM_ENC_NEGABINARY, nb = 32, hier = 0;
    std::vector<prodm_mdr_plan *> plans;
    std::vector<const void *> inputs;
    std::vector<GroupShape> shapes;      // per launch (tuned or default)
    std::vector<GroupLaunch> launches;   // built from plans, inputs and shapes
    void *tables = nullptr;              // device copy of every launch's entries and prefix
    size_t tables_bytes = 0;
    bool built = false;
    hipGraphExec_t graph = nullptr;
    uint64_t device_bytes = 0;
};

namespace prodm {
namespace {

void invalidate(prodm_mdr_batch *b) {
    b->built = false;
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
}

template <class T> int build_launches(prodm_mdr_batch *b, const std::vector<GroupShape> &shapes,
                                      std::vector<GroupLaunch> &out) {
    out.clear();
    // level keys of every field to zero first (the load sweeps take atomic maxima)
    GroupLaunch zero;
    std::vector<std::pair<void *, uint32_t>> keys;
    for (prodm_mdr_plan *p : b->plans) keys.push_back({p->keys, (uint32_t)((p->g.L + 1) * sizeof(T) / 4)});
    if (int rc = zero_batched_launch(keys, zero)) return rc;
    if (!shapes.empty()) zero.shape = shapes[0];
    out.push_back(std::move(zero));
    std::vector<BatchFieldRef> refs(b->plans.size());
    for (size_t i = 0; i < b->plans.size(); i++) {
        prodm_mdr_plan *p = b->plans[i];
        BatchFieldRef &r = refs[i];
        r.in = b->inputs[i];
        r.g = &p->g;
        for (int l = 0; l <= p->g.L; l++) r.level[l] = static_cast<T *>(p->levels) + p->level_off[l];
        r.keys = p->keys;
        r.ws = p->ws;
        r.ws_bytes = p->ws_bytes;
        r.hier = b->hier != 0;
    }
    std::vector<GroupShape> mshapes(shapes.size() > 1 ? shapes.begin() + 1 : shapes.end(), shapes.end());
    std::vector<GroupLaunch> ml;
    if (int rc = mgard_batched_launches<T>(refs, mshapes, ml)) return rc;
    for (auto &L : ml) out.push_back(std::move(L));
    if (b->encoder == PRODM_ENC_NEGABINARY) {
        using S = typename Traits<T>::Stream;
        std::vector<EncodeRef> lv;
        for (prodm_mdr_plan *p : b->plans)
            for (int l = 0; l <= p->g.L; l++) {
                EncodeRef e;
                e.x = static_cast<T *>(p->levels) + p->level_off[l];
                e.n = (uint32_t)p->g.elems[l];
                e.key = static_cast<KeyOf<T> *>(p->keys) + l;
                e.planes = static_cast<S *>(p->planes) + p->plane_off[l];
                e.stride = p->plane_stride[l];
                lv.push_back(e);
            }
        GroupLaunch enc;
        const GroupShape es = out.size() < shapes.size() ? shapes[out.size()] : GroupShape{};
        if (int rc = nb_batched_launch<T>(lv, b->nb, es, enc)) return rc;
        out.push_back(std::move(enc));
    }
    return kOk;
}

// all entries and prefixes in one device buffer (synchronous upload)
int upload(prodm_mdr_batch *b, std::vector<GroupLaunch> &ls) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t total = 0;
    for (auto &L : ls) {
        total = al(total + L.params.size());
        total = al(total + L.first.size() * 4);
    }
    std::vector<uint8_t> host(total);
    std::vector<std::pair<size_t, size_t>> offs;
    size_t o = 0;
    for (auto &L : ls) {
        const size_t po = o;
        if (!L.params.empty()) std::memcpy(host.data() + po, L.params.data(), L.params.size());
        o = al(o + L.params.size());
        const size_t fo = o;
        if (!L.first.empty()) std::memcpy(host.data() + fo, L.first.data(), L.first.size() * 4);
        o = al(o + L.first.size() * 4);
        offs.push_back({po, fo});
    }
    if (total > b->tables_bytes) {
        if (b->tables) {
            PRODM_HIP(hipDeviceSynchronize());
            PRODM_HIP(hipFree(b->tables));
            b->device_bytes -= b->tables_bytes;
            b->tables = nullptr;
            b->tables_bytes = 0;
        }
        PRODM_HIP(hipMalloc(&b->tables, std::max<size_t>(total, 256)));
        b->tables_bytes = std::max<size_t>(total, 256);
        b->device_bytes += b->tables_bytes;
    } else {
        PRODM_HIP(hipDeviceSynchronize());  // launches in flight may still read the old tables
    }
    if (total) PRODM_HIP(hipMemcpy(b->tables, host.data(), total, hipMemcpyHostToDevice));
    char *base = static_cast<char *>(b->tables);
    for (size_t i = 0; i < ls.size(); i++) {
        ls[i].d_params = base + offs[i].first;
        ls[i].d_first = reinterpret_cast<const uint32_t *>(base + offs[i].second);
    }
    return kOk;
}

int build(prodm_mdr_batch *b, const std::vector<GroupShape> &shapes) {
    std::vector<GroupLaunch> ls;
    int rc = b->dtype == PRODM_F64 ? build_launches<double>(b, shapes, ls) : build_launches<float>(b, shapes, ls);
    if (rc) return rc;
    if ((rc = upload(b, ls))) return rc;
    invalidate(b);
    b->launches = std::move(ls);
    b->shapes.resize(b->launches.size());
    for (size_t i = 0; i < b->launches.size(); i++) b->shapes[i] = b->launches[i].shape;
    b->built = true;
    return kOk;
}

int run_launch(const GroupLaunch &L, hipStream_t s) {
    if (L.kind == kGrpNbEncode || L.kind == kGrpZero) return bitplane_group_run(L, s);
    return mgard_group_run(L, s);
}

template <class T> int grouped_encode_fields(prodm_mdr_batch *b, hipStream_t s) {
    // GroupedBP and PerBit: the plan's per-level encode (count or rb
    // histogram, scan, write)
    using S = typename Traits<T>::Stream;
    for (prodm_mdr_plan *p : b->plans)
        for (int l = p->g.L; l >= 0; l--) {
            const T *x = static_cast<const T *>(p->levels) + p->level_off[l];
            S *planes = static_cast<S *>(p->planes) + p->plane_off[l];
            const KeyOf<T> *key = static_cast<const KeyOf<T> *>(p->keys) + l;
            const uint32_t n = (uint32_t)p->g.elems[l];
            const int rc =
                b->encoder == PRODM_ENC_PERBIT
                    ? launch_pb_encode<T>(x, n, 0, key, p->nb, 0, reinterpret_cast<uint64_t *>(planes),
                                          p->plane_stride[l] * sizeof(T) / 8, p->words + (size_t)l * p->nb, p->gbws,
                                          p->gbws_bytes, s)
                    : launch_gb_encode<T>(x, n, 0, key, p->nb, 0, p->rb + p->rb_off[l], planes, p->plane_stride[l],
                                          p->words + (size_t)l * p->nb, p->gbws, p->gbws_bytes, s);
            if (rc) return rc;
        }
    return kOk;
}

int run_all(prodm_mdr_batch *b, hipStream_t s) {
    for (const GroupLaunch &L : b->launches)
        if (int rc = run_launch(L, s)) return rc;
    if (b->encoder == PRODM_ENC_GROUPED || b->encoder == PRODM_ENC_PERBIT)
        return b->dtype == PRODM_F64 ? grouped_encode_fields<double>(b, s) : grouped_encode_fields<float>(b, s);
    return kOk;
}

// ---- autotuner -------------------------------------------------------------------
std::vector<GroupShape> candidates(int kind) {
    std::vector<GroupShape> c;
    auto add = [&](int a, uint32_t cap) { GroupShape g; g.a = a; g.cap = cap; c.push_back(g); };
    switch (kind) {
    case kGrpLoad:
        for (int a : {0, 2, 4, 8, 16})
            for (uint32_t cap : {0u, 1024u, 2048u}) add(a, cap);
        break;
    case kGrpContig:
        for (int a : {1, 2, 4, 8})
            for (uint32_t cap : {0u, 2048u}) add(a, cap);
        break;
    case kGrpStrided:
    case kGrpNbEncode:
        for (uint32_t cap : {0u, 1024u, 2048u, 4096u}) add(0, cap);
        break;
    default: add(0, 0); break;
    }
    return c;
}

uint64_t fnv(uint64_t h, const void *p, size_t n) {
    const uint8_t *c = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    return h;
}
uint64_t signature(const prodm_mdr_batch *b) {
    uint64_t h = 1469598103934665603ull;
    const int head[5] = {b->dtype, b->encoder, b->nb, b->hier, (int)b->plans.size()};
    h = fnv(h, head, sizeof(head));
    for (const prodm_mdr_plan *p : b->plans) {
        const int g[5] = {(int)p->g.dims[0], (int)p->g.dims[1], (int)p->g.dims[2], p->g.L, p->g.nd};
        h = fnv(h, g, sizeof(g));
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        h = fnv(h, prop.gcnArchName, strnlen(prop.gcnArchName, sizeof(prop.gcnArchName)));
        h = fnv(h, &prop.multiProcessorCount, sizeof(int));
    }
    return h;
}

bool load_cache(const char *path, uint64_t sig, size_t nlaunch, std::vector<GroupShape> &shapes) {
    std::ifstream f(path);
    if (!f) return false;
    std::string tag;
    int ver = 0;
    if (!(f >> tag >> ver) || tag != "prodm-batch-shapes" || ver != 1) return false;
    std::string key;
    unsigned long long s = 0;
    size_t n = 0;
    if (!(f >> key >> std::hex >> s >> std::dec) || key != "signature" || s != sig) return false;
    if (!(f >> key >> n) || key != "launches" || n != nlaunch) return false;
    std::vector<GroupShape> out(n);
    for (size_t i = 0; i < n; i++) {
        size_t idx = 0;
        int a = 0;
        unsigned cap = 0;
        if (!(f >> key >> idx >> a >> cap) || key != "launch" || idx != i) return false;
        out[i].a = a;
        out[i].cap = cap;
    }
    shapes = out;
    return true;
}

int save_cache(const char *path, uint64_t sig, const std::vector<GroupShape> &shapes) {
    std::ofstream f(path, std::ios::trunc);
    if (!f) { set_error(std::string("autotune: cannot write ") + path); return kErrArg; }
    f << "prodm-batch-shapes 1\n";
    f << "signature " << std::hex << (unsigned long long)sig << std::dec << "\n";
    f << "launches " << shapes.size() << "\n";
    for (size_t i = 0; i < shapes.size(); i++) f << "launch " << i << " " << shapes[i].a << " " << shapes[i].cap << "\n";
    return f ? kOk : kErrArg;
}

// average ms of `reps` back-to-back runs of launch i
int time_launch(prodm_mdr_batch *b, size_t i, hipStream_t s, int reps, hipEvent_t e0, hipEvent_t e1, float *ms) {
    if (int rc = run_launch(b->launches[i], s)) return rc;  // warm-up
    PRODM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++)
        if (int rc = run_launch(b->launches[i], s)) return rc;
    PRODM_HIP(hipEventRecord(e1, s));
    PRODM_HIP(hipEventSynchronize(e1));
    float t = 0;
    PRODM_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return kOk;
}

int autotune(prodm_mdr_batch *b, hipStream_t s, int reps) {
    if (!b->built)
        if (int rc = build(b, b->shapes)) return rc;
    hipEvent_t e0, e1;
    PRODM_HIP(hipEventCreate(&e0));
    PRODM_HIP(hipEventCreate(&e1));
    int rc = run_all(b, s);  // every buffer holds the data of a real refactor
    std::vector<GroupShape> best = b->shapes;
    for (size_t i = 0; !rc && i < b->launches.size(); i++) {
        const std::vector<GroupShape> cand = candidates(b->launches[i].kind);
        if (cand.size() < 2) continue;
        float tbest = 0;
        GroupShape gbest = best[i];
        bool have = false;
        for (const GroupShape &c : cand) {
            std::vector<GroupShape> trial = best;
            trial[i] = c;
            if ((rc = build(b, trial))) break;
            float t = 0;
            if ((rc = time_launch(b, i, s, reps, e0, e1, &t))) break;
            if (!have || t < tbest) { tbest = t; gbest = c; have = true; }
        }
        best[i] = gbest;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return build(b, best);
}

}  // namespace
}  // namespace prodm

extern "C" {

int prodm_mdr_batch_create(prodm_mdr_batch **out, const prodm_field_desc *fields, int nfields, int nb, int dtype,
                           int encoder) {
    if (!out) { set_error("batch: null output"); return kErrArg; }
    *out = nullptr;
    if (!fields || nfields < 1) { set_error("batch: no fields"); return kErrArg; }
    prodm_mdr_batch *b = new prodm_mdr_batch();
    b->dtype = dtype;
    b->encoder = encoder;
    b->nb = nb;
    int rc = kOk;
    for (int i = 0; i < nfields && !rc; i++) {
        const prodm_field_desc &f = fields[i];
        if (!f.d_in) { set_error("batch: field " + std::to_string(i) + " has no input"); rc = kErrArg; break; }
        if (f.target_level < 1) { set_error("batch: field " + std::to_string(i) + " needs target_level >= 1"); rc = kErrArg; break; }
        prodm_mdr_plan *p = nullptr;
        rc = prodm_mdr_plan_create(&p, f.dims, f.ndims, f.target_level, nb, dtype, encoder);
        if (rc) {
            set_error("batch: field " + std::to_string(i) + ": " + last_error());
            break;
        }
        b->plans.push_back(p);
        b->inputs.push_back(f.d_in);
        b->device_bytes += prodm_mdr_plan_device_bytes(p);
    }
    if (rc) { prodm_mdr_batch_destroy(b); return rc; }
    *out = b;
    return kOk;
}

int prodm_mdr_batch_destroy(prodm_mdr_batch *b) {
    if (!b) return kOk;
    (void)hipDeviceSynchronize();
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->tables) (void)hipFree(b->tables);
    for (prodm_mdr_plan *p : b->plans) prodm_mdr_plan_destroy(p);
    delete b;
    return kOk;
}

int prodm_mdr_batch_fields(const prodm_mdr_batch *b) { return b ? (int)b->plans.size() : 0; }

uint64_t prodm_mdr_batch_device_bytes(const prodm_mdr_batch *b) { return b ? b->device_bytes : 0; }

prodm_mdr_plan *prodm_mdr_batch_plan(prodm_mdr_batch *b, int field) {
    if (!b || field < 0 || field >= (int)b->plans.size()) { set_error("batch: bad field index"); return nullptr; }
    return b->plans[field];
}

int prodm_mdr_batch_set_input(prodm_mdr_batch *b, int field, const void *d_in) {
    if (!b || field < 0 || field >= (int)b->plans.size() || !d_in) { set_error("batch: bad field or input"); return kErrArg; }
    if (b->inputs[field] != d_in) {
        b->inputs[field] = d_in;
        invalidate(b);
    }
    return kOk;
}

int prodm_mdr_batch_set_decomposer(prodm_mdr_batch *b, int kind) {
    if (!b || (kind != PRODM_DECOMP_ORTHOGONAL && kind != PRODM_DECOMP_HIERARCHICAL)) {
        set_error("batch: bad decomposer");
        return kErrArg;
    }
    for (prodm_mdr_plan *p : b->plans)
        if (int rc = prodm_mdr_plan_set_decomposer(p, kind)) return rc;
    if (b->hier != (kind == PRODM_DECOMP_HIERARCHICAL)) {
        b->hier = kind == PRODM_DECOMP_HIERARCHICAL;
        invalidate(b);
    }
    return kOk;
}

int prodm_mdr_batched_refactor(prodm_mdr_batch *b, prodm_stream_t stream) {
    if (!b) { set_error("batched refactor: null batch"); return kErrArg; }
    if (!b->built)
        if (int rc = build(b, b->shapes)) return rc;
    return run_all(b, (hipStream_t)stream);
}

int prodm_mdr_batched_refactor_graph(prodm_mdr_batch *b, prodm_stream_t stream) {
    if (!b) { set_error("batched refactor: null batch"); return kErrArg; }
    hipStream_t s = (hipStream_t)stream;
    if (!b->built)
        if (int rc = build(b, b->shapes)) return rc;
    if (!b->graph) {
        hipGraphExec_t g = nullptr;
        const int rc = graph_capture(s, &g, [&] { return run_all(b, s); });
        if (rc) return rc;
        b->graph = g;
    }
    PRODM_HIP(hipGraphLaunch(b->graph, s));
    return kOk;
}

int prodm_mdr_batch_autotune(prodm_mdr_batch *b, const char *cache_path, int reps, prodm_stream_t stream) {
    if (!b) { set_error("autotune: null batch"); return kErrArg; }
    if (reps < 1) reps = 3;
    hipStream_t s = (hipStream_t)stream;
    if (!b->built)
        if (int rc = build(b, b->shapes)) return rc;
    const uint64_t sig = signature(b);
    std::vector<GroupShape> cached;
    if (cache_path && load_cache(cache_path, sig, b->launches.size(), cached)) return build(b, cached);
    if (int rc = autotune(b, s, reps)) return rc;
    if (cache_path) return save_cache(cache_path, sig, b->shapes);
    return kOk;
}

int prodm_mdr_batch_describe(prodm_mdr_batch *b, char *buf, uint64_t capacity, uint64_t *length) {
    if (!b) { set_error("describe: null batch"); return kErrArg; }
    if (!b->built)
        if (int rc = build(b, b->shapes)) return rc;
    static const char *names[] = {"load", "solve_contig", "solve_strided", "coarsest", "nb_encode", "zero_keys"};
    std::ostringstream o;
    for (size_t i = 0; i < b->launches.size(); i++) {
        const GroupLaunch &L = b->launches[i];
        o << i << " " << names[L.kind] << " entries=" << L.nent << " virtual=" << L.total << " grid=" << L.grid()
          << " block=" << L.block << " a=" << L.shape.a << " cap=" << L.shape.cap << " variant=" << L.variant << "\n";
    }
    const std::string t = o.str();
    if (length) *length = t.size();
    if (buf && capacity) {
        const size_t n = std::min<size_t>(capacity - 1, t.size());
        std::memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return kOk;
}

}  // extern "C"

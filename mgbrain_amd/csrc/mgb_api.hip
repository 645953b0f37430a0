// mgb_api.hip -- the C ABI (include/mgbrain_hip.h) over the HIP kernels.
//
// Host-side orchestration of one single-GPU simulator: device network build,
// format choice, constant upload, the windowed step loop (optionally replayed
// from a hipGraph) and read-back.  Reference counterparts are cited per entry
// point in the header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "mgb_build.h"
#include "mgb_internal.h"
#include "mgb_rng.h"

namespace mgb {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

}  // namespace mgb

using namespace mgb;

struct mgb_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    int32_t n = 0, npad = 0, D = 1, min_delay = 0, window = 1;
    int64_t S = 0;
    int mode = kRingClass;
    int wf = MGB_WF_CLASS;
    int wide_kind = kWideF32;
    int stdp = 0;
    int use_graph = 0;
    int prop_grid = 0, prop_block = 256;
    int32_t split = 1;                 // waves per spike row in k_prop / k_stdp
    int64_t rec_cap = 0;
    int64_t launches = 0;
    float scale_e = 1.f, scale_i = 1.f;
    float sc[6] = {0, 0, 0, 0, 0, 0};
    int64_t n_plastic = 0;
    // device
    Ctl* ctl = nullptr;
    uint8_t *type = nullptr, *cset = nullptr;
    uint16_t *tbl = nullptr, *stream_idx = nullptr;
    float *rate = nullptr, *V = nullptr, *Ie = nullptr, *Ii = nullptr;
    int32_t *refr = nullptr, *last = nullptr, *cnt = nullptr;   // fused STDP: last is [3][npad]
    CSet* csets = nullptr;
    int32_t* stream_base = nullptr;
    uint64_t* stream_seed = nullptr;
    int64_t* row_ptr = nullptr;
    uint32_t* syn = nullptr;
    int32_t* col = nullptr;
    uint8_t* dly = nullptr;
    float* w32 = nullptr;
    uint16_t* w16 = nullptr;
    uint8_t* w8 = nullptr;
    float* row_scale = nullptr;
    int64_t* in_ptr = nullptr;
    uint32_t* in_pos = nullptr;
    int32_t* in_src = nullptr;
    uint32_t* ring = nullptr;
    int32_t *ring_e = nullptr, *ring_i = nullptr;
    uint32_t* tbl_off = nullptr;
    float* seq = nullptr;
    float* tbl_w = nullptr;
    float *ltp = nullptr, *ltd = nullptr;
    int32_t lut_n = 1;
    int32_t lut_lds = 0;               // fused STDP: table entries the fused kernel keeps in LDS
    uint32_t* win_buf = nullptr;
    uint32_t win_stride = 0;
    uint64_t* rec = nullptr;
    // fused tiled path
    int tiled = 0;
    int32_t B = 0, G = 0, chunk = 8, tblock = 512;
    int32_t trw = 1;                   // fused path: ring words per (slot, target)
    int32_t nseg = 1;
    std::vector<int32_t> h_seg_lo;     // [nseg+1] segment boundaries (independent networks)
    std::vector<int32_t> h_seg_tile0;  // [nseg+1] first tile of each segment
    std::vector<uint64_t> h_seg_obase; // [nseg+1] first offset entry of each segment
    int32_t* tile_lo = nullptr;        // [G+1]
    uint32_t* tile_of = nullptr;       // [n] tile of each neuron
    uint16_t* tile_seg = nullptr;      // [G]
    int32_t *seg_lo = nullptr, *seg_tile0 = nullptr;   // [nseg+1]
    uint32_t* seg_kbase = nullptr;     // [nseg+1]
    uint64_t* seg_obase = nullptr;     // [nseg+1]
    uint64_t n_offs = 0, n_keys = 0;
    uint32_t* tcnt = nullptr;          // [3][nseg]
    int tkind = kTClass;               // synapse word kind of the fused kernel
    uint32_t* tperm = nullptr;         // fused STDP: CSR slot of each tile-order position (~0u: padding)
    uint64_t n_tslots = 0;             // length of the (padded) tile-order arrays
    float2* seg_scale = nullptr;       // fixed-point formats, several segments: (scale, 1 / scale)
    uint16_t* seg_of = nullptr;        // segment of each neuron (with seg_scale)
    int32_t *hist = nullptr, *tf = nullptr, *cntf = nullptr;   // fused STDP: deferred dendrite updates
    int32_t* flist = nullptr;          // fused STDP: [2][npad][2] neurons to flush
    uint32_t* fcnt = nullptr;          // [2]
    int32_t stdp_isi = 1;              // fused STDP: shortest inter-spike interval of plastic synapses' ends
    uint32_t* syn_t = nullptr;
    uint16_t* syn16 = nullptr;         // kTClass16 words in tile order (padded sub-rows)
    int64_t* prow = nullptr;           // first slot of each padded row of syn16 / syn_t / w_t
    float* tile_cc = nullptr;          // CLASS kinds: class cache of each tile
    uint32_t* tile_ccls = nullptr;
    uint8_t* tile_lean = nullptr;      // every neuron of the tile an LIFB neuron of its two classes
    float* w_t = nullptr;              // kTF32 / kTStdp / kTStdpB weights in tile order
    uint32_t* row_aux = nullptr;       // kTStdpB: row scale bits | has-plastic << 31 per source
    uint8_t* tile_pl = nullptr;        // kTStdpB: 1 per tile with a plastic in-synapse
    uint16_t* offs16 = nullptr;
    uint32_t* offs32 = nullptr;
    uint32_t* ring_t = nullptr;
    size_t ring_slab = 0;              // words per tile in ring_t (RW * D * B rounded up to 16 bytes)
    // persistent launches of the fused kernel (not MGB_PERSIST_OFF, every workgroup resident)
    int persist_req = MGB_PERSIST_AUTO;
    int persist = 0;                   // 1: persistent launches, 2: with list regions
    uint32_t* bar = nullptr;           // [G * kBarStride] grid-wait flags
    uint32_t* bar_err = nullptr;       // set by a wait that timed out
    uint64_t spin_limit = 0;           // wall-clock ticks
    bool failed = false;               // a persistent launch timed out: results void
    uint32_t* h_err = nullptr;         // pinned host copy of *bar_err, written at the end of each run
    uint32_t* h_err_dev = nullptr;     // its address on the device (the run's last kernel writes it)
    bool err_pending = false;          // h_err is due with the stream's completion
    int xcd_map = -1;                  // fused launches: XCD-contiguous tile order (-1: not computed yet)
    int region = 0;                    // persistent launches: per-tile list regions (TiledArgs::region)
    uint64_t tile_hash = 0;            // FNV-1a of the tile boundaries (the ring slices' layout)
    // host mirrors for read-back
    std::vector<float> class_w;        // [n_tbl][2]
    std::vector<uint16_t> h_tbl;       // [n]
    int32_t n_tbl = 0;
    // graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int32_t graph_windows = 0;
    // per-kernel profiling (opts.profile)
    int profile = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Span {
        int kind;
        hipEvent_t a, b;
        int32_t count;      // launches of the span's kind it covers (graph replays: windows)
    };
    std::vector<Span> spans;
    // bookkeeping
    std::vector<void*> allocs;
    int64_t bytes = 0;
};

namespace {

mgb_status dalloc(mgb_sim* s, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return e == hipErrorOutOfMemory ? MGB_ERR_OOM : MGB_ERR_HIP;
    }
    s->allocs.push_back(*p);
    s->bytes += (int64_t)bytes;
    return MGB_OK;
}

template <class T>
mgb_status dalloc_n(mgb_sim* s, T** p, size_t count) {
    return dalloc(s, reinterpret_cast<void**>(p), count * sizeof(T));
}

void dfree(mgb_sim* s, void* p) {
    if (!p) return;
    auto it = std::find(s->allocs.begin(), s->allocs.end(), p);
    if (it != s->allocs.end()) s->allocs.erase(it);
    (void)hipFree(p);
}

// Temporary device buffers freed on scope exit.
struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    mgb_status get(T** p, size_t count) {
        void* q = nullptr;
        size_t bytes = std::max<size_t>(count * sizeof(T), 4);
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) {
            set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? MGB_ERR_OOM : MGB_ERR_HIP;
        }
        ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return MGB_OK;
    }
    void release(void* p) {
        auto it = std::find(ptrs.begin(), ptrs.end(), p);
        if (it != ptrs.end()) {
            (void)hipFree(p);
            ptrs.erase(it);
        }
    }
};

#define MGB_TRY(x)                       \
    do {                                 \
        mgb_status st_ = (x);            \
        if (st_ != MGB_OK) return st_;   \
    } while (0)

int bits_for(uint64_t v) {   // bits to represent 0..v
    int b = 1;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

void free_all(mgb_sim* s) {
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    s->ev_pool.clear();
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (void* p : s->allocs) (void)hipFree(p);
    s->allocs.clear();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->h_err) (void)hipHostFree(s->h_err);
}

// Per-neuron host inputs common to both construction paths.
struct HostNeurons {
    int32_t n = 0;
    std::vector<uint8_t> type, cset;
    std::vector<uint16_t> stream;
    std::vector<float> rate, v0;
    std::vector<float> lifconst;       // [n_csets][30]
    std::vector<int32_t> stream_base;
    std::vector<uint64_t> stream_seed;
    std::vector<float> fp8_scale;      // [n] or empty
    float stdpconst[6] = {0.1f, -0.01f, 17.f, 34.f, 40.f, 0.f};
    float step = 1e-4f;
    int stdp = 0;
};

// Device list of synapses in reference list order.
struct DevList {
    uint64_t S = 0;
    int32_t* src = nullptr;
    int32_t* tar = nullptr;
    float* weight = nullptr;
    float* delay = nullptr;
    uint8_t* plastic = nullptr;   // may be null (no STDP)
};

mgb_status init_device(mgb_sim* s, const mgb_opts* o) {
    int ndev = 0;
    MGB_HIP_TRY(hipGetDeviceCount(&ndev));
    if (ndev <= 0) {
        set_error("no HIP device visible");
        return MGB_ERR_HIP;
    }
    s->device = o ? o->device : 0;
    if (s->device < 0 || s->device >= ndev) {
        set_error("device %d out of range (%d devices)", s->device, ndev);
        return MGB_ERR_INVALID;
    }
    MGB_HIP_TRY(hipSetDevice(s->device));
    MGB_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    MGB_HIP_TRY(hipEventCreate(&s->ev0));
    MGB_HIP_TRY(hipEventCreate(&s->ev1));
    return MGB_OK;
}

mgb_status validate_neurons(const HostNeurons& h) {
    if (h.n <= 0 || h.n > kMaxNeurons) {
        set_error("neuron count %d outside [1, %d]", h.n, kMaxNeurons);
        return MGB_ERR_INVALID;
    }
    if (!(h.step > 0.f)) {
        set_error("time step must be positive");
        return MGB_ERR_INVALID;
    }
    const size_t ncs = h.lifconst.size() / 30;
    if (ncs == 0 || ncs > 255) {
        set_error("constant sets must number 1..255");
        return MGB_ERR_INVALID;
    }
    for (int32_t i = 0; i < h.n; ++i) {
        if (h.type[i] > MGB_LIFB) {
            set_error("neuron %d has unknown type %d", i, (int)h.type[i]);
            return MGB_ERR_INVALID;
        }
        if (h.cset[i] >= ncs) {
            set_error("neuron %d uses constant set %d of %zu", i, (int)h.cset[i], ncs);
            return MGB_ERR_INVALID;
        }
        if (h.stream[i] >= h.stream_seed.size()) {
            set_error("neuron %d uses Poisson stream %d of %zu", i, (int)h.stream[i],
                      h.stream_seed.size());
            return MGB_ERR_INVALID;
        }
    }
    for (size_t k = 0; k < ncs; ++k) {
        const float* c = &h.lifconst[30 * k];
        if (!(c[0] > 0.f)) {
            set_error("constant set %zu has non-positive DT (LIFECONSTS[0])", k);
            return MGB_ERR_INVALID;
        }
    }
    return MGB_OK;
}

// STDP tables: ltp[k] = A_LTP*expf(-k/TAU_LTP), ltd[k] = A_LTD*expf(-k/TAU_LTD)
// (mgcore.cu:351,355 evaluated on the host); the last entry is the zero plateau.
mgb_status build_stdp_lut(const float* sc, std::vector<float>& ltp, std::vector<float>& ltd) {
    const int32_t cap = 1 << 22;
    ltp.assign(1, 0.f);
    ltd.assign(1, 0.f);
    for (int32_t k = 1; k < cap; ++k) {
        const float x = (float)(-k);
        const float e1 = expf(x / sc[2]);
        const float e2 = expf(x / sc[3]);
        ltp.push_back(sc[0] * e1);
        ltd.push_back(sc[1] * e2);
        if (e1 == 0.f && e2 == 0.f) return MGB_OK;
    }
    set_error("STDP time constants too long for the exponential table (%d steps)", cap);
    return MGB_ERR_UNSUPPORTED;
}

uint64_t host_u64(const void* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

// Tile shape of the fused path.  Tiles never straddle a segment boundary and
// never mix neuron types (one update path per tile).  Within a segment they are
// balanced by weight — a neuron weighs its in-degree plus half the mean
// in-degree of the neurons that have one (its own update) — so tiles of neurons
// without input (Poisson) are larger and the others share the synapses evenly.
// A segment of Nq neurons gets at most max(ceil(Nq / B), its type runs) tiles
// (B the requested mean tile size, default max(256, n / 256)) of at most the
// LDS limit; the smallest per-tile weight cap meeting that is found by bisection.
mgb_status choose_tiles(mgb_sim* s, const mgb_opts* o, int32_t n, int32_t D, const std::vector<uint8_t>& type,
                        const std::vector<uint32_t>& indeg) {
    int lds_max = 65536;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device);
    lds_max = std::min(lds_max, 160 * 1024);
    // the kernel's static LDS (block scan sums, event counter) shares the limit
    // (per target: the difference over 4 targets, which no alignment rounds; STDP keeps at
    // least one entry of each table in LDS)
    const int lut1 = stdp_kind(s->tkind) ? 1 : 0;
    const int64_t per_b =
        (int64_t)(tiled_lds_bytes(D, 4, s->tkind, s->trw) - tiled_lds_bytes(D, 0, s->tkind, s->trw)) / 4;
    int32_t bmax =
        (int32_t)((lds_max - (int64_t)tiled_lds_bytes(D, 0, s->tkind, s->trw, lut1) - kTiledStaticLds) / per_b);
    while (bmax > 1 && (int64_t)tiled_lds_bytes(D, bmax, s->tkind, s->trw, lut1) + kTiledStaticLds > lds_max)
        --bmax;
    // wide words hold an 11-bit tile-local target; CLASS tiles of at most 1024
    // targets take 16-bit words (else 32-bit words with the global id)
    bmax = std::max<int32_t>(1, std::min<int32_t>(bmax, s->tkind == kTClass ? 1024 : 2048));
    int32_t Breq = (o && o->tile_size > 0) ? o->tile_size : std::max<int32_t>(256, (n + 255) / 256);
    Breq = std::max<int32_t>(1, Breq);
    const std::vector<int32_t>& lo = s->h_seg_lo;
    const int32_t nseg = (int32_t)lo.size() - 1;
    std::vector<double> wt((size_t)n);
    double sum_in = 0.0;
    int64_t cnt_in = 0;
    for (int32_t t = 0; t < n; ++t) {
        const uint32_t d = indeg[t] + indeg[(size_t)n + t];
        if (d) {
            sum_in += d;
            ++cnt_in;
        }
    }
    const double base = cnt_in ? 0.5 * sum_in / (double)cnt_in : 1.0;
    for (int32_t t = 0; t < n; ++t) wt[t] = (double)(indeg[t] + indeg[(size_t)n + t]) + base;
    // tiles of [a, b) under weight cap T; their starts appended to `cuts`
    auto greedy = [&](int32_t a, int32_t b, double T, std::vector<int32_t>* cuts) {
        int32_t count = 0, start = a;
        double acc = 0.0;
        for (int32_t i = a; i < b; ++i) {
            if (i > start && (type[i] != type[i - 1] || i - start >= bmax || acc + wt[i] > T)) {
                if (cuts) cuts->push_back(start);
                ++count;
                start = i;
                acc = 0.0;
            }
            acc += wt[i];
        }
        if (b > a) {
            if (cuts) cuts->push_back(start);
            ++count;
        }
        return count;
    };
    std::vector<int32_t> tlo;
    std::vector<int32_t> tile0(nseg + 1, 0);
    for (int32_t q = 0; q < nseg; ++q) {
        const int32_t a = lo[q], b = lo[q + 1];
        tile0[q] = (int32_t)tlo.size();
        if (b <= a) continue;
        const int32_t target = (int32_t)(((int64_t)(b - a) + Breq - 1) / Breq);
        double t_lo = 0.0, t_hi = 0.0;
        for (int32_t i = a; i < b; ++i) {
            t_lo = std::max(t_lo, wt[i]);
            t_hi += wt[i];
        }
        if (greedy(a, b, t_lo, nullptr) > target) {
            for (int it = 0; it < 40 && t_hi - t_lo > 1e-9 * t_hi; ++it) {
                const double mid = 0.5 * (t_lo + t_hi);
                if (greedy(a, b, mid, nullptr) > target) t_lo = mid;
                else t_hi = mid;
            }
        } else {
            t_hi = t_lo;
        }
        greedy(a, b, t_hi, &tlo);
    }
    tile0[nseg] = (int32_t)tlo.size();
    std::vector<uint16_t> tseg(tlo.size());
    for (int32_t q = 0; q < nseg; ++q)
        for (int32_t g = tile0[q]; g < tile0[q + 1]; ++g) tseg[g] = (uint16_t)q;
    tlo.push_back(n);
    uint64_t hsh = 1469598103934665603ull;
    for (int32_t v : tlo) hsh = (hsh ^ (uint64_t)(uint32_t)v) * 1099511628211ull;
    s->tile_hash = hsh;
    int32_t B = 1;
    for (size_t g = 0; g + 1 < tlo.size(); ++g) B = std::max(B, tlo[g + 1] - tlo[g]);
    s->B = B;
    std::vector<uint32_t> tof((size_t)n);
    for (size_t g = 0; g + 1 < tlo.size(); ++g)
        for (int32_t t = tlo[g]; t < tlo[g + 1]; ++t) tof[t] = (uint32_t)g;
    s->chunk = (o && (o->chunk == 1 || o->chunk == 2 || o->chunk == 4 || o->chunk == 16)) ? o->chunk : 8;
    s->tblock = (o && o->prop_block >= 64 && o->prop_block <= 1024) ? (o->prop_block / 64) * 64 : 512;
    std::vector<uint32_t> kbase(nseg + 1, 0);
    std::vector<uint64_t> obase(nseg + 1, 0);
    uint64_t kb = 0, ob = 0;
    for (int32_t q = 0; q < nseg; ++q) {
        const int32_t nq = lo[q + 1] - lo[q];
        const int32_t gq = tile0[q + 1] - tile0[q];
        kbase[q] = (uint32_t)kb;
        obase[q] = ob;
        kb += (uint64_t)nq * (uint64_t)gq;
        ob += (uint64_t)nq * (uint64_t)(gq + 1);
        if (kb >= 0xFFFFFFFFull) {
            set_error("tiled propagation: %d neurons x %d tiles exceed 32-bit keys", n, tile0[nseg]);
            return MGB_ERR_UNSUPPORTED;
        }
    }
    kbase[nseg] = (uint32_t)kb;
    obase[nseg] = ob;
    s->h_seg_tile0 = tile0;
    s->h_seg_obase = obase;
    s->G = (int32_t)tseg.size();
    s->nseg = nseg;
    s->n_offs = ob;
    s->n_keys = kb;
    hipStream_t st = s->stream;
    MGB_TRY(dalloc_n(s, &s->tile_lo, tlo.size()));
    MGB_TRY(dalloc_n(s, &s->tile_of, (size_t)n));
    MGB_HIP_TRY(hipMemcpyAsync(s->tile_of, tof.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
    MGB_TRY(dalloc_n(s, &s->tile_seg, tseg.size()));
    MGB_TRY(dalloc_n(s, &s->seg_lo, lo.size()));
    MGB_TRY(dalloc_n(s, &s->seg_tile0, tile0.size()));
    MGB_TRY(dalloc_n(s, &s->seg_kbase, kbase.size()));
    MGB_TRY(dalloc_n(s, &s->seg_obase, obase.size()));
    MGB_HIP_TRY(hipMemcpyAsync(s->tile_lo, tlo.data(), 4 * tlo.size(), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->tile_seg, tseg.data(), 2 * tseg.size(), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->seg_lo, lo.data(), 4 * lo.size(), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->seg_tile0, tile0.data(), 4 * tile0.size(), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->seg_kbase, kbase.data(), 4 * kbase.size(), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->seg_obase, obase.data(), 8 * obase.size(), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipStreamSynchronize(st));   // host vectors go out of scope
    const size_t lds = tiled_lds_bytes(D, B, s->tkind, s->trw);
    if (lds > 65536) MGB_HIP_TRY(set_tiled_lds_limit(s->tkind, lds));
    return MGB_OK;
}

TileTabs tile_tabs(const mgb_sim* s, const uint16_t* seg_dev) {
    TileTabs tt{};
    tt.nseg = s->nseg; tt.B = s->B;
    tt.seg_of = s->nseg > 1 ? seg_dev : nullptr;
    tt.seg_lo = s->seg_lo; tt.seg_tile0 = s->seg_tile0;
    tt.seg_kbase = s->seg_kbase; tt.seg_obase = s->seg_obase;
    tt.tile_of = s->tile_of; tt.tile_lo = s->tile_lo;
    return tt;
}

// Per-(source, tile) offsets of the tile-sorted copy (sorted_keys: its sort keys).
mgb_status build_tile_offsets(mgb_sim* s, int32_t n, const uint32_t* sorted_keys, const TileTabs& tt) {
    hipStream_t st = s->stream;
    uint32_t max_row = 0;
    {
        std::vector<int64_t> rp((size_t)n + 1);
        MGB_HIP_TRY(hipMemcpyAsync(rp.data(), s->row_ptr, 8 * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
        MGB_HIP_TRY(hipStreamSynchronize(st));
        for (int32_t m = 0; m < n; ++m) max_row = std::max<uint32_t>(max_row, (uint32_t)(rp[m + 1] - rp[m]));
    }
    if (max_row >= (1u << 26)) {
        set_error("tiled propagation: a CSR row of %u synapses exceeds 2^26", max_row);
        return MGB_ERR_UNSUPPORTED;
    }
    const size_t no = (size_t)s->n_offs;
    if (max_row <= 0xFFFFu) MGB_TRY(dalloc_n(s, &s->offs16, no));
    else MGB_TRY(dalloc_n(s, &s->offs32, no));
    MGB_HIP_TRY(tile_offsets(sorted_keys, s->row_ptr, tt, s->n_offs, s->offs16, s->offs32, st));
    return MGB_OK;
}

// Moves one tile-order array into the padded layout (`total` elements, zero
// padding) and frees the unpadded one (`s_old` elements).
template <class T>
mgb_status repad(mgb_sim* s, T** arr, uint64_t total, uint64_t s_old, const uint32_t* deo, const uint32_t* den) {
    hipStream_t st = s->stream;
    T* to = nullptr;
    MGB_TRY(dalloc_n(s, &to, (size_t)total));
    MGB_HIP_TRY(hipMemsetAsync(to, 0, sizeof(T) * (size_t)total, st));
    MGB_HIP_TRY(pad_copy<T>(*arr, s->row_ptr, s->prow, deo, den, tile_tabs(s, nullptr), s->n_offs, to, st));
    MGB_HIP_TRY(hipStreamSynchronize(st));
    dfree(s, *arr);
    s->bytes -= (int64_t)(sizeof(T) * s_old);
    *arr = to;
    return MGB_OK;
}

// The tile-order copy (syn16, or syn_t and w_t) with every (source, tile)
// sub-row starting at a multiple of kSubAlign words, so that a work item of the
// fused kernel is aligned vector loads.  Offset entry t of row m becomes the
// sub-row's start inside the padded row, which begins at prow[m]; its low bits
// hold the padding after sub-row t - 1, so that sub-row t is (e[t + 1] & ~7) -
// (e[t] & ~7) - (e[t + 1] & 7) words long.  Padding words are zero;
// 2 * kSubAlign words of slack follow the last row.
mgb_status pad_subrows(mgb_sim* s, int32_t n, uint32_t* pmap = nullptr) {
    hipStream_t st = s->stream;
    const size_t no = (size_t)s->n_offs;
    std::vector<uint32_t> eo(no), en(no);
    if (s->offs16) {
        std::vector<uint16_t> t(no);
        MGB_HIP_TRY(hipMemcpyAsync(t.data(), s->offs16, 2 * no, hipMemcpyDeviceToHost, st));
        MGB_HIP_TRY(hipStreamSynchronize(st));
        for (size_t e = 0; e < no; ++e) eo[e] = t[e];
    } else {
        MGB_HIP_TRY(hipMemcpyAsync(eo.data(), s->offs32, 4 * no, hipMemcpyDeviceToHost, st));
        MGB_HIP_TRY(hipStreamSynchronize(st));
    }
    const std::vector<int32_t>& lo = s->h_seg_lo;
    const std::vector<int32_t>& t0 = s->h_seg_tile0;
    const std::vector<uint64_t>& ob = s->h_seg_obase;
    std::vector<int64_t> prow((size_t)n + 1, 0);
    constexpr uint32_t A = (uint32_t)kSubAlign;
    uint64_t p = 0, s_old = 0;
    uint32_t emax = 0;
    for (size_t q = 0; q + 1 < lo.size(); ++q) {
        const uint32_t gq = (uint32_t)(t0[q + 1] - t0[q]);
        for (int32_t m = lo[q]; m < lo[q + 1]; ++m) {
            prow[m] = (int64_t)p;
            const size_t e0 = ob[q] + (size_t)(m - lo[q]) * (gq + 1);
            uint32_t r = 0, pad = 0;
            for (uint32_t t = 0; t < gq; ++t) {
                const uint32_t len = eo[e0 + t + 1] - eo[e0 + t];
                en[e0 + t] = r | pad;
                pad = (A - len % A) % A;
                r += len + pad;
                s_old += len;
            }
            en[e0 + gq] = r | pad;
            emax = std::max(emax, r | pad);
            p += r;
        }
    }
    prow[n] = (int64_t)p;
    const uint64_t total = p + 2 * A;
    if (total > 0xFFFFFFFFull) {
        set_error("tiled propagation: %llu padded synapse slots exceed 32-bit positions",
                  (unsigned long long)total);
        return MGB_ERR_UNSUPPORTED;
    }
    uint32_t *deo = nullptr, *den = nullptr;
    MGB_TRY(dalloc_n(s, &s->prow, (size_t)n + 1));
    MGB_TRY(dalloc_n(s, &deo, no));
    MGB_TRY(dalloc_n(s, &den, no));
    MGB_HIP_TRY(hipMemcpyAsync(s->prow, prow.data(), 8 * ((size_t)n + 1), hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(deo, eo.data(), 4 * no, hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(den, en.data(), 4 * no, hipMemcpyHostToDevice, st));
    s->n_tslots = total;
    if (pmap) MGB_HIP_TRY(pad_map(s->row_ptr, s->prow, deo, den, tile_tabs(s, nullptr), no, pmap, st));
    if (s->syn16) MGB_TRY(repad(s, &s->syn16, total, s_old, deo, den));
    if (s->syn_t) MGB_TRY(repad(s, &s->syn_t, total, s_old, deo, den));
    if (s->w_t) {
        uint32_t* w = reinterpret_cast<uint32_t*>(s->w_t);    // moved as 32-bit words
        MGB_TRY(repad(s, &w, total, s_old, deo, den));
        s->w_t = reinterpret_cast<float*>(w);
    }
    dfree(s, deo);
    s->bytes -= (int64_t)(4 * no);
    if (emax <= 0xFFFFu) {
        std::vector<uint16_t> t(no);
        for (size_t e = 0; e < no; ++e) t[e] = (uint16_t)en[e];
        if (!s->offs16) {
            dfree(s, s->offs32);
            s->bytes -= (int64_t)(4 * no);
            s->offs32 = nullptr;
            MGB_TRY(dalloc_n(s, &s->offs16, no));
        }
        MGB_HIP_TRY(hipMemcpyAsync(s->offs16, t.data(), 2 * no, hipMemcpyHostToDevice, st));
        MGB_HIP_TRY(hipStreamSynchronize(st));
        dfree(s, den);
        s->bytes -= (int64_t)(4 * no);
    } else {
        s->bytes -= (int64_t)((s->offs16 ? 2 : 4) * no);
        dfree(s, s->offs16);
        dfree(s, s->offs32);
        s->offs16 = nullptr;
        s->offs32 = den;
    }
    return MGB_OK;
}

// Candidate segments: maximal runs of equal Poisson-stream index in id order,
// when every stream occupies one run.  They become segments of the fused path
// if no synapse joins two of them (checked by k_stats).
std::vector<int32_t> stream_runs(const HostNeurons& h, std::vector<uint16_t>& run_of) {
    std::vector<int32_t> lo{0};
    run_of.assign((size_t)h.n, 0);
    std::vector<uint8_t> seen(h.stream_seed.size() + 1, 0);
    if (h.n > 0) seen[h.stream[0]] = 1;
    for (int32_t i = 1; i < h.n; ++i) {
        if (h.stream[i] != h.stream[i - 1]) {
            if (seen[h.stream[i]] || lo.size() >= 0xFFFFu) return {0, h.n};
            seen[h.stream[i]] = 1;
            lo.push_back(i);
        }
        run_of[i] = (uint16_t)(lo.size() - 1);
    }
    lo.push_back(h.n);
    return lo;
}

// Builds device CSR + tables from a device synapse list; frees nothing of `L`.
mgb_status build_network(mgb_sim* s, const HostNeurons& h, const DevList& L, int wf_req,
                         const mgb_opts* o) {
    const int32_t n = h.n;
    const uint64_t S = L.S;
    hipStream_t st = s->stream;
    s->S = (int64_t)S;
    if (S >= 0xFFFFFFFFull) {
        set_error("%llu synapses exceed the 2^32-1 limit of one simulator",
                  (unsigned long long)S);
        return MGB_ERR_UNSUPPORTED;
    }
    Scratch X;
    // ---- statistics + delay quantisation --------------------------------------
    uint8_t* dq = nullptr;
    uint32_t *flags = nullptr, *wcls = nullptr, *indeg = nullptr, *in_pl = nullptr, *absmax = nullptr;
    int32_t* mm = nullptr;
    unsigned long long* npl = nullptr;
    MGB_TRY(X.get(&dq, S));
    MGB_TRY(X.get(&flags, 1));
    MGB_TRY(X.get(&mm, 2));
    MGB_TRY(X.get(&wcls, 2 * (size_t)n));
    MGB_TRY(X.get(&indeg, 2 * (size_t)n));
    MGB_TRY(X.get(&in_pl, (size_t)n));
    MGB_TRY(X.get(&npl, 1));
    std::vector<uint16_t> run_of;
    std::vector<int32_t> runs = stream_runs(h, run_of);
    const size_t nruns = runs.size() - 1;                  // max |w| per run of streams
    MGB_TRY(X.get(&absmax, std::max<size_t>(nruns, 1)));
    uint16_t* seg_dev = nullptr;
    if (runs.size() > 2) {
        MGB_TRY(X.get(&seg_dev, (size_t)n));
        MGB_HIP_TRY(hipMemcpyAsync(seg_dev, run_of.data(), 2 * (size_t)n, hipMemcpyHostToDevice, st));
    }
    const int32_t mm_init[2] = {0x7FFFFFFF, 0};
    MGB_HIP_TRY(hipMemsetAsync(flags, 0, 4, st));
    MGB_HIP_TRY(hipMemcpyAsync(mm, mm_init, 8, hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemsetAsync(wcls, 0xFF, 8 * (size_t)n, st));
    MGB_HIP_TRY(hipMemsetAsync(indeg, 0, 8 * (size_t)n, st));
    MGB_HIP_TRY(hipMemsetAsync(in_pl, 0, 4 * (size_t)n, st));
    MGB_HIP_TRY(hipMemsetAsync(absmax, 0, 4 * std::max<size_t>(nruns, 1), st));
    MGB_HIP_TRY(hipMemsetAsync(npl, 0, 8, st));
    if (S > 0) {
        StatArgs a{};
        a.S = S; a.n = n; a.step = h.step;
        a.src = L.src; a.tar = L.tar; a.weight = L.weight; a.delay = L.delay;
        a.plastic = h.stdp ? L.plastic : nullptr;
        a.seg = seg_dev;                                   // null for one run: absmax[0]
        a.dq = dq; a.flags = flags; a.min_d = mm; a.max_d = mm + 1;
        a.wcls = wcls; a.indeg = indeg; a.in_pl = in_pl; a.absmax = absmax; a.n_plastic = npl;
        MGB_HIP_TRY(run_stats(a, st));
    }
    uint32_t hflags = 0;
    std::vector<uint32_t> habs(std::max<size_t>(nruns, 1), 0u);
    int32_t hmm[2];
    unsigned long long hnpl = 0;
    MGB_HIP_TRY(hipMemcpyAsync(&hflags, flags, 4, hipMemcpyDeviceToHost, st));
    MGB_HIP_TRY(hipMemcpyAsync(hmm, mm, 8, hipMemcpyDeviceToHost, st));
    MGB_HIP_TRY(hipMemcpyAsync(habs.data(), absmax, 4 * habs.size(), hipMemcpyDeviceToHost, st));
    MGB_HIP_TRY(hipMemcpyAsync(&hnpl, npl, 8, hipMemcpyDeviceToHost, st));
    std::vector<uint32_t> h_wcls(2 * (size_t)n), h_indeg(2 * (size_t)n);
    MGB_HIP_TRY(hipMemcpyAsync(h_wcls.data(), wcls, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
    MGB_HIP_TRY(hipMemcpyAsync(h_indeg.data(), indeg, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
    MGB_HIP_TRY(hipStreamSynchronize(st));
    if (hflags & kFlagBadId) {
        set_error("a synapse references a neuron id outside [0, %d)", n);
        return MGB_ERR_INVALID;
    }
    if (hflags & kFlagDelayLow) {
        set_error("a synapse delay rounds to less than one step (the reference divides by "
                  "min_delay = 0, mgsim.cpp:802)");
        return MGB_ERR_INVALID;
    }
    if (hflags & kFlagDelayHigh) {
        set_error("a synapse delay exceeds %d steps", kWideMaxDelay);
        return MGB_ERR_UNSUPPORTED;
    }
    s->h_seg_lo = (hflags & kFlagCrossSeg) ? std::vector<int32_t>{0, n} : runs;
    const int32_t max_d = S ? hmm[1] : 0;
    s->min_delay = S ? hmm[0] : 0;
    s->D = max_d + 1;                                   // mgsim.cpp:71
    // spike rows of ~256 synapses per wave in the direct kernels
    s->split = (int32_t)std::max<int64_t>(1, std::min<int64_t>(32, (int64_t)(S / (uint64_t)n) / 256));
    s->n_plastic = (int64_t)hnpl;
    uint32_t max_deg_cls = 0, max_deg_tot = 0;
    for (int32_t t = 0; t < n; ++t) {
        max_deg_cls = std::max({max_deg_cls, h_indeg[t], h_indeg[(size_t)n + t]});
        max_deg_tot = std::max(max_deg_tot, h_indeg[t] + h_indeg[(size_t)n + t]);
    }
    // ---- format choice ----------------------------------------------------------
    const bool class_ok = !(hflags & kFlagMixed) && !h.stdp && max_d <= kClassMaxDelay &&
                          max_deg_cls <= 0xFFFFu;
    int wf = wf_req;
    if (wf == MGB_WF_AUTO) wf = class_ok ? MGB_WF_CLASS : MGB_WF_F32;
    if (wf == MGB_WF_CLASS && !class_ok) {
        set_error("CLASS weight format needs one weight per (target, sign), no STDP, delays <= %d "
                  "steps and in-degree <= 65535", kClassMaxDelay);
        return MGB_ERR_UNSUPPORTED;
    }
    if (wf < MGB_WF_AUTO || wf > MGB_WF_FP8) {
        set_error("unknown weight format %d", wf);
        return MGB_ERR_INVALID;
    }
    s->wf = wf;
    s->mode = (wf == MGB_WF_CLASS) ? kRingClass : kRingFixed;
    s->wide_kind = wf == MGB_WF_BF16 ? kWideBF16 : (wf == MGB_WF_FP8 ? kWideFP8 : kWideF32);

    // ---- stable sort by source ---------------------------------------------------
    uint32_t *kin = nullptr, *kout = nullptr, *vin = nullptr, *vout = nullptr;
    MGB_TRY(X.get(&kin, S));
    MGB_TRY(X.get(&kout, S));
    MGB_TRY(X.get(&vin, S));
    MGB_TRY(X.get(&vout, S));
    const bool cls = s->mode == kRingClass;
    MGB_HIP_TRY(make_sort_input(L.src, L.tar, L.weight, dq, cls, kin, vin, S, st));
    const int kbits = bits_for((uint64_t)n);
    size_t tb = 0;
    void* temp = nullptr;
    if (S > 0) {
        MGB_HIP_TRY(sort_pairs(nullptr, &tb, kin, kout, vin, vout, S, kbits, st));
        MGB_TRY(X.get(reinterpret_cast<uint8_t**>(&temp), tb));
        MGB_HIP_TRY(sort_pairs(temp, &tb, kin, kout, vin, vout, S, kbits, st));
    }
    MGB_TRY(dalloc_n(s, &s->row_ptr, (size_t)n + 1));
    MGB_HIP_TRY(row_bounds(kout, S, n, s->row_ptr, st));

    if (cls) {
        MGB_TRY(dalloc_n(s, &s->syn, S));
        if (S) MGB_HIP_TRY(hipMemcpyAsync(s->syn, vout, S * 4, hipMemcpyDeviceToDevice, st));
        // class tables: unique (exc weight, inh weight) per target
        std::map<std::pair<uint32_t, uint32_t>, int32_t> ids;
        std::vector<uint32_t> tmax_e, tmax_i;
        s->h_tbl.assign((size_t)n, 0);
        std::vector<std::pair<uint32_t, uint32_t>> keys;
        for (int32_t t = 0; t < n; ++t) {
            const auto key = std::make_pair(h_wcls[t], h_wcls[(size_t)n + t]);
            auto it = ids.find(key);
            int32_t id;
            if (it == ids.end()) {
                id = (int32_t)keys.size();
                ids.emplace(key, id);
                keys.push_back(key);
                tmax_e.push_back(0);
                tmax_i.push_back(0);
            } else {
                id = it->second;
            }
            s->h_tbl[t] = (uint16_t)id;
            tmax_e[id] = std::max(tmax_e[id], h_indeg[t]);
            tmax_i[id] = std::max(tmax_i[id], h_indeg[(size_t)n + t]);
        }
        if (keys.size() > 0xFFFFu) {
            set_error("too many weight classes (%zu)", keys.size());
            return MGB_ERR_UNSUPPORTED;
        }
        s->n_tbl = (int32_t)keys.size();
        std::vector<uint32_t> offs(2 * keys.size());
        std::vector<float> seq;
        s->class_w.assign(2 * keys.size(), 0.f);
        for (size_t c = 0; c < keys.size(); ++c) {
            for (int side = 0; side < 2; ++side) {
                const uint32_t bits = side ? keys[c].second : keys[c].first;
                // at least kClassCache entries: the fused kernel caches that many
                const uint32_t len = std::max<uint32_t>((side ? tmax_i[c] : tmax_e[c]) + 1, kClassCache);
                float w = 0.f;
                if (bits != kWSentinel) std::memcpy(&w, &bits, 4);
                s->class_w[2 * c + side] = w;
                offs[2 * c + side] = (uint32_t)seq.size();
                // seq[k] = k-fold sequential FP32 sum of w: the value k float
                // atomicAdds of w into a zeroed buffer leave (mgcore.cu:112-115)
                float acc = 0.f;
                seq.push_back(acc);
                for (uint32_t k = 1; k < len; ++k) {
                    acc = acc + w;
                    seq.push_back(acc);
                }
            }
        }
        MGB_TRY(dalloc_n(s, &s->tbl, (size_t)n));
        MGB_TRY(dalloc_n(s, &s->tbl_off, offs.size()));
        MGB_TRY(dalloc_n(s, &s->seq, seq.size()));
        MGB_TRY(dalloc_n(s, &s->tbl_w, s->class_w.size()));
        MGB_HIP_TRY(hipMemcpyAsync(s->tbl_w, s->class_w.data(), 4 * s->class_w.size(), hipMemcpyHostToDevice, st));
        MGB_HIP_TRY(hipMemcpyAsync(s->tbl, s->h_tbl.data(), 2 * (size_t)n, hipMemcpyHostToDevice, st));
        MGB_HIP_TRY(hipMemcpyAsync(s->tbl_off, offs.data(), 4 * offs.size(), hipMemcpyHostToDevice, st));
        MGB_HIP_TRY(hipMemcpyAsync(s->seq, seq.data(), 4 * seq.size(), hipMemcpyHostToDevice, st));
        const int prop = o ? o->propagation : MGB_PROP_AUTO;
        s->tiled = (prop == MGB_PROP_AUTO || prop == MGB_PROP_TILED) ? 1 : 0;
        s->tkind = kTClass;
        if (s->tiled) {
            MGB_TRY(choose_tiles(s, o, n, s->D, h.type, h_indeg));
            // tile-sorted copy: stable sort by (source, target tile) keeps list order
            // inside every (source, tile) sub-row
            const TileTabs tt = tile_tabs(s, seg_dev);
            MGB_HIP_TRY(make_tile_keys(L.src, L.tar, L.weight, dq, tt, kin, vin, S, st));
            const int tbits = bits_for(s->n_keys);
            if (S > 0) {
                size_t tb2 = 0;
                MGB_HIP_TRY(sort_pairs(nullptr, &tb2, kin, kout, vin, vout, S, tbits, st));
                if (tb2 > tb) {
                    X.release(temp);
                    MGB_TRY(X.get(reinterpret_cast<uint8_t**>(&temp), tb2));
                    tb = tb2;
                }
                MGB_HIP_TRY(sort_pairs(temp, &tb2, kin, kout, vin, vout, S, tbits, st));
            }
            if (s->B <= 1024) {
                // 2 B per synapse: the target is stored relative to its tile
                s->tkind = kTClass16;
                MGB_TRY(dalloc_n(s, &s->syn16, S));
                MGB_HIP_TRY(class16_words(vout, S, tt, s->syn16, st));
            } else {
                MGB_TRY(dalloc_n(s, &s->syn_t, S));
                if (S) MGB_HIP_TRY(hipMemcpyAsync(s->syn_t, vout, S * 4, hipMemcpyDeviceToDevice, st));
            }
            MGB_TRY(build_tile_offsets(s, n, kout, tt));
            MGB_TRY(pad_subrows(s, n));
            MGB_TRY(dalloc_n(s, &s->tile_cc, (size_t)s->G * 4 * kClassCache));
            MGB_TRY(dalloc_n(s, &s->tile_ccls, (size_t)s->G * 2));
            MGB_HIP_TRY(launch_tile_ccache(s->tile_lo, s->G, s->tbl, s->tbl_off, s->seq, s->tile_cc, s->tile_ccls,
                                           st));
        }
        MGB_HIP_TRY(hipStreamSynchronize(st));
    } else {
        // fixed-point scale per segment (a batched network gets the scale it has
        // alone): |ring value| <= in-degree * max |w| < 2^31
        const std::vector<int32_t>& slo = s->h_seg_lo;
        const size_t nseg = slo.size() - 1;
        std::vector<float2> seg_sc(nseg);
        for (size_t q = 0; q < nseg; ++q) {
            uint32_t deg = 0, abits = 0;
            for (int32_t t = slo[q]; t < slo[q + 1]; ++t) deg = std::max(deg, h_indeg[t] + h_indeg[(size_t)n + t]);
            if (nseg == nruns) abits = habs[q];
            else for (uint32_t b : habs) abits = std::max(abits, b);      // one segment
            float wabs = 0.f;
            std::memcpy(&wabs, &abits, 4);
            if (h.stdp) wabs = std::max({wabs, std::fabs(h.stdpconst[4]), std::fabs(h.stdpconst[5])});
            if (s->wide_kind != kWideF32) wabs *= 1.0625f;   // BF16 / FP8 rounding headroom
            const double bound = (double)std::max<uint32_t>(deg, 1) * (double)wabs;
            int e = 30;
            if (bound > 0) e = (int)std::floor(std::log2(2147483520.0 / bound));
            e = std::max(-100, std::min(e, 100));
            seg_sc[q] = make_float2(std::ldexp(1.0f, e), std::ldexp(1.0f, -e));
        }
        s->scale_e = s->scale_i = seg_sc[0].x;
        if (nseg > 1) {
            MGB_TRY(dalloc_n(s, &s->seg_scale, nseg));
            MGB_TRY(dalloc_n(s, &s->seg_of, (size_t)n));
            MGB_HIP_TRY(hipMemcpyAsync(s->seg_scale, seg_sc.data(), sizeof(float2) * nseg, hipMemcpyHostToDevice, st));
            MGB_HIP_TRY(hipMemcpyAsync(s->seg_of, run_of.data(), 2 * (size_t)n, hipMemcpyHostToDevice, st));
        }
        MGB_TRY(dalloc_n(s, &s->col, S));
        MGB_TRY(dalloc_n(s, &s->dly, S));
        MGB_TRY(dalloc_n(s, &s->w32, S));
        if (s->wide_kind == kWideBF16) MGB_TRY(dalloc_n(s, &s->w16, S));
        if (s->wide_kind == kWideFP8) MGB_TRY(dalloc_n(s, &s->w8, S));
        if (s->wide_kind != kWideF32) {
            MGB_TRY(dalloc_n(s, &s->row_scale, (size_t)n));
            if (s->wide_kind == kWideFP8 && !h.fp8_scale.empty()) {
                MGB_HIP_TRY(hipMemcpyAsync(s->row_scale, h.fp8_scale.data(), 4 * (size_t)n,
                                           hipMemcpyHostToDevice, st));
            } else {
                // per-source-row scale: the row's largest |w| maps onto 448 (FP8) or 1 (BF16)
                uint32_t* rmax = nullptr;
                MGB_TRY(X.get(&rmax, (size_t)n));
                MGB_HIP_TRY(row_scales(L.src, L.weight, S, n, s->wide_kind == kWideFP8 ? 448.f : 1.f, rmax,
                                       s->row_scale, st));
            }
        }
        uint32_t* pos_of = nullptr;
        if (h.stdp) MGB_TRY(X.get(&pos_of, S));
        GatherArgs g{};
        g.S = S; g.idx = vout; g.src = L.src; g.tar = L.tar; g.weight = L.weight; g.dq = dq;
        g.plastic = h.stdp ? L.plastic : nullptr;
        g.col = s->col; g.dly = s->dly; g.w32 = s->w32; g.w16 = s->w16; g.w8 = s->w8;
        g.row_scale = s->row_scale; g.pos_of = pos_of;
        MGB_HIP_TRY(gather_wide(g, st));
        // fused path for delays that fit the 5-bit field; with STDP, FP32 weights or
        // BF16 static ones
        const int prop = o ? o->propagation : MGB_PROP_AUTO;
        const bool tiled_ok = max_d <= kClassMaxDelay && (!h.stdp || s->wide_kind != kWideFP8);
        if (prop == MGB_PROP_TILED && !tiled_ok) {
            set_error("tiled propagation of the wide formats needs delays <= %d steps (and FP32 or BF16 weights "
                      "with STDP)", kClassMaxDelay);
            return MGB_ERR_UNSUPPORTED;
        }
        s->tiled = (tiled_ok && prop != MGB_PROP_DIRECT) ? 1 : 0;
        if (s->tiled) {
            s->tkind = h.stdp ? (s->wide_kind == kWideBF16 ? kTStdpB : kTStdp)
                              : (s->wide_kind == kWideBF16 ? kTBF16 : (s->wide_kind == kWideFP8 ? kTFP8 : kTF32));
            // one fixed-point ring when every integrating neuron is LIFB (it integrates
            // the sum of both sides)
            bool lifb_only = true;
            for (int32_t i = 0; i < n && lifb_only; ++i)
                lifb_only = h.type[i] != MGB_LIF && h.type[i] != MGB_LIFE;
            s->trw = lifb_only ? 1 : 2;
            MGB_TRY(choose_tiles(s, o, n, s->D, h.type, h_indeg));
            const TileTabs tt = tile_tabs(s, seg_dev);
            // kout holds each CSR slot's source (the keys of the sort by source)
            MGB_HIP_TRY(make_tile_keys_csr(kout, s->col, tt, kin, vin, S, st));
            const int tbits = bits_for(s->n_keys);
            if (S > 0) {
                size_t tb2 = 0;
                MGB_HIP_TRY(sort_pairs(nullptr, &tb2, kin, kout, vin, vout, S, tbits, st));
                if (tb2 > tb) {
                    X.release(temp);
                    MGB_TRY(X.get(reinterpret_cast<uint8_t**>(&temp), tb2));
                    tb = tb2;
                }
                MGB_HIP_TRY(sort_pairs(temp, &tb2, kin, kout, vin, vout, S, tbits, st));
            }
            MGB_TRY(dalloc_n(s, &s->syn_t, S));
            if (s->tkind == kTF32 || stdp_kind(s->tkind)) MGB_TRY(dalloc_n(s, &s->w_t, S));
            MGB_HIP_TRY(tile_gather_wide(vout, S, s->col, s->dly, s->w16, s->w8, s->w32, tt, s->tkind, s->syn_t,
                                         s->w_t, st));
            MGB_TRY(build_tile_offsets(s, n, kout, tt));
            uint32_t* pmap = nullptr;          // STDP: padded position of each tile-order position
            if (stdp_kind(s->tkind)) MGB_TRY(X.get(&pmap, std::max<size_t>(S, 1)));
            MGB_TRY(pad_subrows(s, n, pmap));
            if (stdp_kind(s->tkind)) {
                // plastic weights live in (padded) tile order: keep the CSR slot of each
                // position (read-back; ~0u at padding) and map the CSR slots of the
                // in-synapse lists below to padded tile positions
                MGB_TRY(dalloc_n(s, &s->tperm, (size_t)s->n_tslots));
                MGB_HIP_TRY(hipMemsetAsync(s->tperm, 0xFF, 4 * (size_t)s->n_tslots, st));
                MGB_HIP_TRY(scatter_u32(pmap, vout, S, s->tperm, st));
                MGB_HIP_TRY(invert_slots(vout, S, vin, st));
                MGB_HIP_TRY(remap_u32(vin, pmap, S, st));
                MGB_HIP_TRY(remap_u32(pos_of, vin, S, st));
            }
        }
        if (h.stdp) {
            // incoming CSR of plastic synapses in list order (dend_offs/dend_refs)
            const int ibits = bits_for((uint64_t)n);
            MGB_HIP_TRY(make_in_keys(L.tar, L.plastic, n, kin, vin, S, st));
            if (S > 0) {
                size_t tb2 = 0;
                MGB_HIP_TRY(sort_pairs(nullptr, &tb2, kin, kout, vin, vout, S, ibits, st));
                if (tb2 > tb) {
                    X.release(temp);
                    MGB_TRY(X.get(reinterpret_cast<uint8_t**>(&temp), tb2));
                    tb = tb2;
                }
                MGB_HIP_TRY(sort_pairs(temp, &tb2, kin, kout, vin, vout, S, ibits, st));
            }
            MGB_TRY(dalloc_n(s, &s->in_ptr, (size_t)n + 1));
            MGB_HIP_TRY(row_bounds(kout, S, n, s->in_ptr, st));
            MGB_TRY(dalloc_n(s, &s->in_pos, (size_t)s->n_plastic));
            MGB_TRY(dalloc_n(s, &s->in_src, (size_t)s->n_plastic));
            MGB_HIP_TRY(in_fill(vout, pos_of, L.src, (uint64_t)s->n_plastic, s->in_pos, s->in_src, st));
            if (s->tiled && s->tkind == kTStdpB) {
                MGB_TRY(dalloc_n(s, &s->row_aux, (size_t)n));
                MGB_TRY(dalloc_n(s, &s->tile_pl, (size_t)s->G));
                MGB_HIP_TRY(row_aux(s->row_scale, n, L.src, L.plastic, S, s->row_aux, st));
                MGB_HIP_TRY(tile_plastic(s->tile_lo, s->G, s->in_ptr, s->tile_pl, st));
            }
            if (s->tiled && stdp_kind(s->tkind)) {
                // window cap: the shortest inter-spike interval of any end of a plastic
                // synapse (1 step for Poisson neurons, refractory steps + 1 for the
                // integrating types; neurons of other types never fire)
                std::vector<int32_t> isi((size_t)n + 1, kMaxWindow);   // isi[n]: the result
                for (int32_t i = 0; i < n; ++i) {
                    const int ty = h.type[i];
                    if (ty == MGB_POISSON) {
                        isi[i] = 1;
                    } else if (ty == MGB_LIF || ty == MGB_LIFE || ty == MGB_LIFB) {
                        const float* c = &h.lifconst[30 * (size_t)h.cset[i]];
                        const int32_t refr = (int32_t)roundf(c[8] / c[0]) - 1;   // as setup_state
                        isi[i] = std::max(refr + 1, 1);
                    }
                }
                int32_t* disi = nullptr;
                MGB_TRY(X.get(&disi, (size_t)n + 1));
                MGB_HIP_TRY(hipMemcpyAsync(disi, isi.data(), 4 * ((size_t)n + 1), hipMemcpyHostToDevice, st));
                MGB_HIP_TRY(plastic_isi(s->in_ptr, s->in_src, disi, n, (uint64_t)s->n_plastic, disi + n, st));
                MGB_HIP_TRY(hipMemcpyAsync(&s->stdp_isi, disi + n, 4, hipMemcpyDeviceToHost, st));
                MGB_HIP_TRY(hipStreamSynchronize(st));   // isi goes out of scope
            }
        }
        MGB_HIP_TRY(hipStreamSynchronize(st));
    }
    return MGB_OK;
}

void update_persist(mgb_sim* s);

mgb_status setup_state(mgb_sim* s, const HostNeurons& h, const mgb_opts* o) {
    const int32_t n = h.n;
    hipStream_t st = s->stream;
    s->n = n;
    s->npad = (n + kNeuBlock - 1) / kNeuBlock * kNeuBlock;
    s->stdp = h.stdp;
    std::memcpy(s->sc, h.stdpconst, sizeof(s->sc));
    const size_t np = (size_t)s->npad;
    MGB_TRY(dalloc_n(s, &s->ctl, 1));
    MGB_TRY(dalloc_n(s, &s->type, np));
    MGB_TRY(dalloc_n(s, &s->cset, np));
    MGB_TRY(dalloc_n(s, &s->stream_idx, np));
    MGB_TRY(dalloc_n(s, &s->rate, np));
    MGB_TRY(dalloc_n(s, &s->V, np));
    MGB_TRY(dalloc_n(s, &s->Ie, np));
    MGB_TRY(dalloc_n(s, &s->Ii, np));
    MGB_TRY(dalloc_n(s, &s->refr, np));
    const size_t nlast = (s->tiled && stdp_kind(s->tkind)) ? 3 : 1;   // fused STDP: three rotating buffers
    MGB_TRY(dalloc_n(s, &s->last, nlast * np));
    MGB_TRY(dalloc_n(s, &s->cnt, np));
    MGB_HIP_TRY(hipMemsetAsync(s->ctl, 0, sizeof(Ctl), st));
    MGB_HIP_TRY(hipMemcpyAsync(s->type, h.type.data(), (size_t)n, hipMemcpyHostToDevice, st));
    if (s->tiled && s->tile_cc) {
        // CLASS kinds: which tiles hold only LIFB neurons of their two cached classes
        MGB_TRY(dalloc_n(s, &s->tile_lean, (size_t)s->G));
        MGB_HIP_TRY(launch_tile_lean(s->tile_lo, s->G, s->type, s->tbl, s->tile_lean, st));
    }
    MGB_HIP_TRY(hipMemcpyAsync(s->cset, h.cset.data(), (size_t)n, hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->stream_idx, h.stream.data(), 2 * (size_t)n, hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->rate, h.rate.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->V, h.v0.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemsetAsync(s->Ie, 0, 4 * np, st));
    MGB_HIP_TRY(hipMemsetAsync(s->Ii, 0, 4 * np, st));
    MGB_HIP_TRY(hipMemsetAsync(s->refr, 0, 4 * np, st));
    MGB_HIP_TRY(hipMemsetAsync(s->last, 0, 4 * nlast * np, st));   // mgnet.cpp:16 zero-initialised
    if (nlast == 3) {
        MGB_TRY(dalloc_n(s, &s->hist, np * kStdpHist));
        MGB_TRY(dalloc_n(s, &s->tf, np));
        MGB_TRY(dalloc_n(s, &s->cntf, np));
        MGB_HIP_TRY(hipMemsetAsync(s->hist, 0x80, 4 * np * kStdpHist, st));   // kHistNone
        MGB_HIP_TRY(hipMemsetAsync(s->tf, 0xFF, 4 * np, st));     // -1: no update deferred yet
        MGB_HIP_TRY(hipMemsetAsync(s->cntf, 0, 4 * np, st));
        MGB_TRY(dalloc_n(s, &s->flist, 2 * 2 * np));
        MGB_TRY(dalloc_n(s, &s->fcnt, 2));
        MGB_HIP_TRY(hipMemsetAsync(s->fcnt, 0, 8, st));
    }
    MGB_HIP_TRY(hipMemsetAsync(s->cnt, 0, 4 * np, st));
    // constants: LIFECONSTS + refractory steps (mgcore.cu:153)
    const size_t ncs = h.lifconst.size() / 30;
    std::vector<CSet> cs(ncs);
    for (size_t k = 0; k < ncs; ++k) {
        std::memset(&cs[k], 0, sizeof(CSet));
        std::memcpy(cs[k].c, &h.lifconst[30 * k], 30 * sizeof(float));
        const int32_t refr = (int32_t)roundf(cs[k].c[8] / cs[k].c[0]) - 1;
        std::memcpy(&cs[k].c[30], &refr, 4);
    }
    MGB_TRY(dalloc_n(s, &s->csets, ncs));
    MGB_HIP_TRY(hipMemcpyAsync(s->csets, cs.data(), sizeof(CSet) * ncs, hipMemcpyHostToDevice, st));
    MGB_TRY(dalloc_n(s, &s->stream_base, h.stream_base.size()));
    MGB_TRY(dalloc_n(s, &s->stream_seed, h.stream_seed.size()));
    MGB_HIP_TRY(hipMemcpyAsync(s->stream_base, h.stream_base.data(), 4 * h.stream_base.size(),
                               hipMemcpyHostToDevice, st));
    MGB_HIP_TRY(hipMemcpyAsync(s->stream_seed, h.stream_seed.data(), 8 * h.stream_seed.size(),
                               hipMemcpyHostToDevice, st));
    // ring
    if (s->tiled) {
        s->ring_slab = ((size_t)s->D * (size_t)s->B * (size_t)s->trw + 3u) & ~(size_t)3u;
        const size_t rt = (size_t)s->G * s->ring_slab;
        MGB_TRY(dalloc_n(s, &s->ring_t, rt));
        MGB_HIP_TRY(hipMemsetAsync(s->ring_t, 0, 4 * rt, st));
    } else if (s->mode == kRingClass) {
        MGB_TRY(dalloc_n(s, &s->ring, (size_t)s->D * np));
        MGB_HIP_TRY(hipMemsetAsync(s->ring, 0, 4 * (size_t)s->D * np, st));
    } else {
        MGB_TRY(dalloc_n(s, &s->ring_e, (size_t)s->D * np));
        MGB_TRY(dalloc_n(s, &s->ring_i, (size_t)s->D * np));
        MGB_HIP_TRY(hipMemsetAsync(s->ring_e, 0, 4 * (size_t)s->D * np, st));
        MGB_HIP_TRY(hipMemsetAsync(s->ring_i, 0, 4 * (size_t)s->D * np, st));
    }
    // window
    int32_t win = s->S ? s->min_delay : kMaxWindow;
    if (o && o->window > 0) win = std::min(win, o->window);
    win = std::min(win, kMaxWindow);
    // STDP: one step per window on the direct path; the fused path takes windows no
    // longer than the shortest inter-spike interval of the plastic synapses' ends
    if (h.stdp) win = s->tiled ? std::min(win, s->stdp_isi) : 1;
    s->window = std::max(win, 1);
    s->win_stride = (uint32_t)(np * (size_t)s->window);
    MGB_TRY(dalloc_n(s, &s->win_buf, (s->tiled ? 3 : 2) * (size_t)s->win_stride));
    if (s->tiled) {
        MGB_TRY(dalloc_n(s, &s->tcnt, 3 * (size_t)s->nseg));
        MGB_HIP_TRY(hipMemsetAsync(s->tcnt, 0, 12 * (size_t)s->nseg, st));
    }
    // record
    s->rec_cap = o ? std::max<int64_t>(o->record_capacity, 0) : 0;
    if (s->rec_cap > 0) MGB_TRY(dalloc_n(s, &s->rec, (size_t)s->rec_cap));
    // STDP tables
    if (h.stdp) {
        std::vector<float> ltp, ltd;
        MGB_TRY(build_stdp_lut(h.stdpconst, ltp, ltd));
        s->lut_n = (int32_t)ltp.size();
        // one allocation, LTD after LTP (the fused kernel reads both through one buffer
        // resource)
        MGB_TRY(dalloc_n(s, &s->ltp, 2 * ltp.size()));
        s->ltd = s->ltp + ltp.size();
        MGB_HIP_TRY(hipMemcpyAsync(s->ltp, ltp.data(), 4 * ltp.size(), hipMemcpyHostToDevice, st));
        MGB_HIP_TRY(hipMemcpyAsync(s->ltd, ltd.data(), 4 * ltd.size(), hipMemcpyHostToDevice, st));
        if (s->tiled && stdp_kind(s->tkind)) {
            // the LDS left beside the tile's data holds the tables' leading entries
            int lds_max = 65536;
            (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device);
            lds_max = std::min(lds_max, 160 * 1024);
            const int64_t room =
                (int64_t)lds_max - kTiledStaticLds - (int64_t)tiled_lds_bytes(s->D, s->B, s->tkind, s->trw);
            // (at least one entry: the tiles leave room for it)
            s->lut_lds = (int32_t)std::max<int64_t>(1, std::min<int64_t>(s->lut_n, room / 8));
            const size_t lds = tiled_lds_bytes(s->D, s->B, s->tkind, s->trw, s->lut_lds);
            if (lds > 65536) MGB_HIP_TRY(set_tiled_lds_limit(s->tkind, lds));
        }
    }
    s->use_graph = o ? o->use_graph : 0;
    s->profile = o ? o->profile : 0;
    s->prop_block = (o && o->prop_block >= 64 && o->prop_block <= 1024) ? (o->prop_block / 64) * 64 : 256;
    s->prop_grid = (o && o->prop_grid > 0) ? o->prop_grid : 0;
    if (s->tiled && !stdp_kind(s->tkind)) {
        // persistent launches: wait flags (reset before each launch) and the wait's bound
        MGB_TRY(dalloc_n(s, &s->bar, (size_t)s->G * kBarStride));
        MGB_TRY(dalloc_n(s, &s->bar_err, 1));
        MGB_HIP_TRY(hipMemsetAsync(s->bar, 0, 4 * (size_t)s->G * kBarStride, st));
        MGB_HIP_TRY(hipMemsetAsync(s->bar_err, 0, 4, st));
        s->persist_req = o ? o->persistent : MGB_PERSIST_AUTO;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) != hipSuccess || khz <= 0)
            khz = 100000;
        const uint64_t us = (o && o->spin_limit_us > 0) ? (uint64_t)o->spin_limit_us : 2000000ull;
        s->spin_limit = us * (uint64_t)khz / 1000ull;
        MGB_HIP_TRY(hipHostMalloc((void**)&s->h_err, 4, hipHostMallocMapped));
        *s->h_err = 0u;
        MGB_HIP_TRY(hipHostGetDevicePointer((void**)&s->h_err_dev, s->h_err, 0));
    }
    if (s->tiled) update_persist(s);
    MGB_HIP_TRY(hipStreamSynchronize(st));
    return MGB_OK;
}

NeuArgs neu_args(const mgb_sim* s) {
    NeuArgs a{};
    a.n = s->n; a.npad = s->npad; a.D = s->D;
    a.type = s->type; a.cset = s->cset; a.tbl = s->tbl; a.stream = s->stream_idx;
    a.rate = s->rate; a.V = s->V; a.Ie = s->Ie; a.Ii = s->Ii;
    a.refr = s->refr; a.last = s->last; a.cnt = s->cnt;
    a.csets = s->csets; a.stream_base = s->stream_base; a.stream_seed = s->stream_seed;
    a.ring = s->ring; a.tbl_off = s->tbl_off; a.seq = s->seq; a.tbl_w = s->tbl_w;
    a.ring_e = s->ring_e; a.ring_i = s->ring_i;
    a.inv_e = 1.0f / s->scale_e; a.inv_i = 1.0f / s->scale_i;
    a.seg_scale = s->seg_scale; a.seg_of = s->seg_of;
    a.ctl = s->ctl; a.win_buf = s->win_buf; a.win_stride = s->win_stride;
    return a;
}

PropArgs prop_args(const mgb_sim* s) {
    PropArgs a{};
    a.n = s->n; a.npad = s->npad; a.D = s->D; a.split = s->split;
    a.row_ptr = s->row_ptr; a.syn = s->syn; a.col = s->col; a.dly = s->dly;
    a.w32 = s->w32; a.w16 = s->w16; a.w8 = s->w8; a.row_scale = s->row_scale;
    a.ring = s->ring; a.ring_e = s->ring_e; a.ring_i = s->ring_i;
    a.scale_e = s->scale_e; a.scale_i = s->scale_i;
    a.seg_scale = s->seg_scale; a.seg_of = s->seg_of;
    a.ctl = s->ctl; a.win_buf = s->win_buf; a.win_stride = s->win_stride;
    a.rec = s->rec; a.rec_cap = (uint64_t)s->rec_cap;
    return a;
}

StdpArgs stdp_args(const mgb_sim* s) {
    StdpArgs a{};
    a.n = s->n; a.split = s->split;
    a.row_ptr = s->row_ptr; a.col = s->col; a.dly = s->dly; a.w32 = s->w32;
    a.in_ptr = s->in_ptr; a.in_pos = s->in_pos; a.in_src = s->in_src; a.last = s->last;
    a.ltp = s->ltp; a.ltd = s->ltd; a.lut_n = s->lut_n;
    a.w_max = s->sc[4]; a.w_min = s->sc[5];
    a.ctl = s->ctl; a.win_buf = s->win_buf; a.win_stride = s->win_stride;
    return a;
}

// Every workgroup of a fused launch of `a` (persistent form or not) fits on the
// device at once.
bool resident(const mgb_sim* s, const TiledArgs& a, bool persistent) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess) return false;
    return (int64_t)a.G <= (int64_t)cus * tiled_blocks_per_cu(a, s->tblock, persistent);
}

TiledArgs tiled_args(const mgb_sim* s);

// The launch form of the current launch shape, computed once per shape: the
// XCD-contiguous tile order when every workgroup is resident, and persistent launches
// when asked for (not MGB_PERSIST_OFF), the kernel has the form (not STDP) and every
// workgroup of its persistent form is resident.
void update_persist(mgb_sim* s) {
    s->persist = 0;
    s->xcd_map = -1;
    s->region = 0;
    if (!s->tiled) return;
    TiledArgs a = tiled_args(s);
    s->xcd_map = a.xcd_map;
    if (stdp_kind(s->tkind) || s->persist_req == MGB_PERSIST_OFF || !s->bar) return;
    // per-tile list regions (the kernel's variant with them) when asked for, or by
    // default in networks large enough for windows of many spikes (in small ones the
    // variant without them, which preloads the neuron state, is the faster), when one
    // thread per tile can scan their counts and the counts fit in LDS beside the tile's
    // data with every workgroup still resident
    const bool want = s->persist_req == MGB_PERSIST_REGIONS ||
                      (s->persist_req == MGB_PERSIST_AUTO && s->n >= kRegionMinNeurons);
    if (want && s->tkind == kTClass16 && s->G <= kRegionMaxG && s->G <= s->tblock) {
        a.region = 1;
        if (set_tiled_pw_lds_limit(a) == hipSuccess && resident(s, a, true)) {
            s->region = a.region;
            s->persist = 2;
            return;
        }
        (void)hipGetLastError();
        a.region = 0;
    }
    s->persist = resident(s, a, true) ? 1 : 0;
}

TiledArgs tiled_args(const mgb_sim* s) {
    TiledArgs a{};
    a.n = s->n; a.D = s->D; a.B = s->B; a.G = s->G; a.chunk = s->chunk;
    a.kind = s->tkind;
    a.scale_e = s->scale_e; a.scale_i = s->scale_i;
    a.inv_e = 1.0f / s->scale_e; a.inv_i = 1.0f / s->scale_i;
    a.seg_scale = s->seg_scale;
    a.w_t = s->w_t; a.row_scale = s->row_scale;
    a.row_aux = s->row_aux; a.tile_pl = s->tile_pl;
    a.nseg = s->nseg; a.window = s->window;
    a.tile_lo = s->tile_lo; a.tile_seg = s->tile_seg; a.seg_lo = s->seg_lo;
    a.seg_tile0 = s->seg_tile0; a.seg_obase = s->seg_obase; a.tcnt = s->tcnt;
    a.type = s->type; a.cset = s->cset; a.tbl = s->tbl; a.stream = s->stream_idx;
    a.rate = s->rate; a.V = s->V; a.Ie = s->Ie; a.Ii = s->Ii;
    a.refr = s->refr; a.last = s->last; a.cnt = s->cnt;
    a.csets = s->csets; a.stream_base = s->stream_base; a.stream_seed = s->stream_seed;
    a.tbl_off = s->tbl_off; a.seq = s->seq; a.tbl_w = s->tbl_w;
    a.row_ptr = s->row_ptr; a.syn_t = s->syn_t; a.syn16 = s->syn16; a.offs16 = s->offs16; a.offs32 = s->offs32;
    a.prow = s->prow;
    a.tile_cc = s->tile_cc; a.tile_ccls = s->tile_ccls; a.tile_lean = s->tile_lean;
    a.ring_t = s->ring_t; a.ring_slab = s->ring_slab; a.rw = s->trw;
    a.ctl = s->ctl; a.win_buf = s->win_buf; a.win_stride = s->win_stride;
    a.rec = s->rec; a.rec_cap = (uint64_t)s->rec_cap;
    if (stdp_kind(s->tkind)) {
        a.last3 = s->last; a.last_stride = (size_t)s->npad;
        a.hist = s->hist; a.tf = s->tf; a.cntf = s->cntf;
        a.flist = s->flist; a.fcnt = s->fcnt; a.flist_cap = (uint32_t)s->npad;
        a.din_ptr = s->in_ptr; a.din_pos = s->in_pos; a.din_src = s->in_src;
        a.ltp = s->ltp; a.ltd = s->ltd; a.lut_n = s->lut_n; a.lut_lds = s->lut_lds;
        a.w_max = s->sc[4]; a.w_min = s->sc[5];
    }
    a.bar = s->bar; a.bar_err = s->bar_err; a.spin_limit = s->spin_limit;
    a.region = s->region;
    a.xcd_map = s->xcd_map >= 0 ? s->xcd_map : (resident(s, a, false) ? 1 : 0);
    return a;
}

int prop_grid(const mgb_sim* s, int32_t nst) {
    if (s->prop_grid > 0) return s->prop_grid;
    // waves >= expected spikes x split; capped so an idle window costs little
    const int64_t waves_per_block = s->prop_block / 64;
    int64_t spikes_max = (int64_t)s->n * (int64_t)nst * (int64_t)s->split;
    int64_t g = (spikes_max + waves_per_block - 1) / waves_per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

enum { kSpanNeuron = 0, kSpanProp = 1, kSpanStdp = 2, kSpanGraph = 3 };

mgb_status next_event(mgb_sim* s, hipEvent_t* e) {
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t ne = nullptr;
        MGB_HIP_TRY(hipEventCreate(&ne));
        s->ev_pool.push_back(ne);
    }
    *e = s->ev_pool[s->ev_used++];
    return MGB_OK;
}

// Brackets one kernel launch with events on the handle's stream when profiling.
struct SpanGuard {
    mgb_sim* s;
    int kind;
    bool on;
    int32_t count = 1;
    hipEvent_t a = nullptr;
    mgb_status begin() {
        if (!on) return MGB_OK;
        MGB_TRY(next_event(s, &a));
        MGB_HIP_TRY(hipEventRecord(a, s->stream));
        return MGB_OK;
    }
    mgb_status end() {
        if (!on) return MGB_OK;
        hipEvent_t b = nullptr;
        MGB_TRY(next_event(s, &b));
        MGB_HIP_TRY(hipEventRecord(b, s->stream));
        s->spans.push_back({kind, a, b, count});
        return MGB_OK;
    }
};

// Issues the kernels of the windows covering `nsteps` steps after base_step.
mgb_status issue_windows(mgb_sim* s, int32_t nsteps, bool finalize, bool capturing = false) {
    const NeuArgs na = neu_args(s);
    const PropArgs pa = prop_args(s);
    const StdpArgs sa = stdp_args(s);
    const bool prof = s->profile && !capturing;
    int32_t off = 0, wk = 0;
    if (s->tiled) {
        const TiledArgs ta = tiled_args(s);
        const int32_t nfull = nsteps / s->window;
        if (s->persist && nfull >= 2) {
            // the run's whole windows in one launch
            SpanGuard g1{s, kSpanProp, prof, nfull};
            MGB_TRY(g1.begin());
            MGB_HIP_TRY(launch_tiled_persistent(ta, 0, 0, s->window, nfull, s->tblock, s->stream));
            MGB_TRY(g1.end());
            s->launches += 2;
            off = nfull * s->window;
            wk = nfull;
        }
        while (off < nsteps) {
            const int32_t nst = std::min(s->window, nsteps - off);
            // fused STDP: the flush kernel after every second window and after the last
            TiledArgs tw = ta;
            tw.flush_next = stdp_kind(s->tkind) && ((wk & 1) != 0 || off + nst >= nsteps) ? 1 : 0;
            SpanGuard g1{s, kSpanProp, prof};
            MGB_TRY(g1.begin());
            MGB_HIP_TRY(launch_tiled(tw, off, wk, nst, s->tblock, s->stream));
            MGB_TRY(g1.end());
            s->launches += 1;
            if (tw.flush_next) {
                SpanGuard g2{s, kSpanStdp, prof};
                MGB_TRY(g2.begin());
                MGB_HIP_TRY(launch_stdp_flush(ta, off, wk, s->stream));
                MGB_TRY(g2.end());
                s->launches += 1;
            }
            off += nst;
            ++wk;
        }
        if (finalize) {
            MGB_HIP_TRY(launch_finalize_tiled(s->ctl, nsteps, wk, s->persist ? s->bar_err : nullptr,
                                              s->persist ? s->h_err_dev : nullptr, s->stream));
            s->launches += 1;
        }
        return MGB_OK;
    }
    while (off < nsteps) {
        const int32_t nst = std::min(s->window, nsteps - off);
        SpanGuard g0{s, kSpanNeuron, prof};
        MGB_TRY(g0.begin());
        MGB_HIP_TRY(launch_neurons(s->mode, na, off, wk, nst, s->stream));
        MGB_TRY(g0.end());
        SpanGuard g1{s, kSpanProp, prof};
        MGB_TRY(g1.begin());
        MGB_HIP_TRY(launch_prop(s->mode, s->wide_kind, pa, off, wk, prop_grid(s, nst),
                                s->prop_block, s->stream));
        MGB_TRY(g1.end());
        s->launches += 2;
        if (s->stdp) {
            SpanGuard g2{s, kSpanStdp, prof};
            MGB_TRY(g2.begin());
            MGB_HIP_TRY(launch_stdp(sa, off, wk, prop_grid(s, nst), s->stream));
            MGB_TRY(g2.end());
            s->launches += 1;
        }
        off += nst;
        ++wk;
    }
    if (finalize) {
        MGB_HIP_TRY(launch_finalize(s->ctl, nsteps, wk, s->stream));
        s->launches += 1;
    }
    return MGB_OK;
}

mgb_status ensure_graph(mgb_sim* s, int32_t steps) {
    if (s->gexec) return MGB_OK;
    MGB_HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int64_t before = s->launches;
    mgb_status st = issue_windows(s, steps, true, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    s->launches = before;
    if (st != MGB_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return MGB_ERR_HIP;
    }
    s->graph = g;
    MGB_HIP_TRY(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
    s->graph_windows = steps;
    return MGB_OK;
}

mgb_status copy_from_host(mgb_sim* s, Scratch& X, void** dst, const void* src, size_t bytes) {
    uint8_t* p = nullptr;
    MGB_TRY(X.get(&p, bytes));
    if (bytes) MGB_HIP_TRY(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s->stream));
    *dst = p;
    return MGB_OK;
}

mgb_status validate_opts(const mgb_opts* o) {
    if (!o) return MGB_OK;
    if (o->weight_format < MGB_WF_AUTO || o->weight_format > MGB_WF_FP8 || o->window < 0 ||
        o->record_capacity < 0 || o->propagation < MGB_PROP_AUTO || o->propagation > MGB_PROP_TILED ||
        o->tile_size < 0 || o->device < 0 || o->persistent < MGB_PERSIST_AUTO || o->persistent > MGB_PERSIST_LISTS ||
        o->spin_limit_us < 0) {
        set_error("invalid mgb_opts");
        return MGB_ERR_INVALID;
    }
    return MGB_OK;
}

mgb_status create_common(mgb_sim* s, const HostNeurons& h, const DevList& L, const mgb_opts* o) {
    MGB_TRY(validate_neurons(h));
    MGB_TRY(build_network(s, h, L, o ? o->weight_format : MGB_WF_AUTO, o));
    MGB_TRY(setup_state(s, h, o));
    return MGB_OK;
}

}  // namespace

// Device memory of mgb_read_weights_noalloc, one copy per GPU, held until the
// process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns on each device (scratch_on).
constexpr uint64_t kNoAllocSlots = 1ull << 24;   // floats: 64 MiB
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocSlots];

extern "C" {

int32_t mgb_abi_version(void) { return MGB_ABI_VERSION; }

const char* mgb_last_error(void) { return g_err.c_str(); }

mgb_status mgb_create(const mgb_net_desc* d, const mgb_opts* o, mgb_sim** out) {
    if (!d || !out) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    *out = nullptr;
    MGB_TRY(validate_opts(o));
    if (d->n_neurons <= 0 || !d->type || !d->lifconst || d->n_csets < 1 ||
        (d->n_synapses > 0 && (!d->syn_src || !d->syn_tar || !d->syn_weight || !d->syn_delay)) ||
        d->n_synapses < 0) {
        set_error("incomplete network description");
        return MGB_ERR_INVALID;
    }
    HostNeurons h;
    const int32_t n = d->n_neurons;
    h.n = n;
    h.step = d->step;
    h.stdp = d->stdp ? 1 : 0;
    h.lifconst.assign(d->lifconst, d->lifconst + 30 * (size_t)d->n_csets);
    if (d->stdpconst) std::memcpy(h.stdpconst, d->stdpconst, sizeof(h.stdpconst));
    h.type.resize(n);
    h.cset.assign(n, 0);
    h.stream.assign(n, 0);
    h.rate.assign(n, 0.f);
    h.v0.resize(n);
    for (int32_t i = 0; i < n; ++i) {
        if (d->type[i] < 0 || d->type[i] > MGB_LIFB) {
            set_error("neuron %d has unknown type %d", i, d->type[i]);
            return MGB_ERR_INVALID;
        }
        h.type[i] = (uint8_t)d->type[i];
        if (d->cset) h.cset[i] = d->cset[i];
        if (d->stream) h.stream[i] = d->stream[i];
        if (d->rate) h.rate[i] = d->rate[i];
    }
    const int32_t ns = std::max(d->n_streams, 1);
    h.stream_base.assign(ns, 0);
    h.stream_seed.assign(ns, d->seed);
    for (int32_t k = 0; k < ns; ++k) {
        if (d->stream_base && k < d->n_streams) h.stream_base[k] = d->stream_base[k];
        if (d->stream_seed && k < d->n_streams) h.stream_seed[k] = d->stream_seed[k];
    }
    for (int32_t i = 0; i < n; ++i) {
        if (h.cset[i] >= (uint32_t)d->n_csets) {
            set_error("neuron %d uses constant set %d of %d", i, (int)h.cset[i], d->n_csets);
            return MGB_ERR_INVALID;
        }
        h.v0[i] = d->v_init ? d->v_init[i] : h.lifconst[30 * (size_t)h.cset[i] + 1];
    }
    if (d->fp8_scale) h.fp8_scale.assign(d->fp8_scale, d->fp8_scale + n);
    // host-side checks first, so invalid networks fail before any device work
    MGB_TRY(validate_neurons(h));
    for (int64_t k = 0; k < d->n_synapses; ++k) {
        const int32_t a = d->syn_src[k], b = d->syn_tar[k];
        if (a < 0 || a >= n || b < 0 || b >= n) {
            set_error("synapse %lld references neuron (%d -> %d) outside [0, %d)", (long long)k, a, b, n);
            return MGB_ERR_INVALID;
        }
        const int q = (int)roundf(d->syn_delay[k] / d->step);
        if (q < 1) {
            set_error("synapse %lld: delay %g s rounds to %d steps; at least one step is required (the "
                      "reference divides by min_delay = 0, mgsim.cpp:802)",
                      (long long)k, (double)d->syn_delay[k], q);
            return MGB_ERR_INVALID;
        }
        if (q > kWideMaxDelay) {
            set_error("synapse %lld: delay of %d steps exceeds %d", (long long)k, q, kWideMaxDelay);
            return MGB_ERR_UNSUPPORTED;
        }
    }

    mgb_sim* s = new (std::nothrow) mgb_sim();
    if (!s) {
        set_error("host allocation failed");
        return MGB_ERR_OOM;
    }
    mgb_status st = init_device(s, o);
    if (st == MGB_OK) {
        Scratch X;
        DevList L;
        L.S = (uint64_t)d->n_synapses;
        const size_t S = (size_t)L.S;
        void *a = nullptr, *b = nullptr, *c = nullptr, *e = nullptr, *p = nullptr;
        st = copy_from_host(s, X, &a, d->syn_src, 4 * S);
        if (st == MGB_OK) st = copy_from_host(s, X, &b, d->syn_tar, 4 * S);
        if (st == MGB_OK) st = copy_from_host(s, X, &c, d->syn_weight, 4 * S);
        if (st == MGB_OK) st = copy_from_host(s, X, &e, d->syn_delay, 4 * S);
        std::vector<uint8_t> pl;
        if (st == MGB_OK && h.stdp) {
            if (d->syn_plastic) pl.assign(d->syn_plastic, d->syn_plastic + S);
            else pl.assign(S, 1);
            st = copy_from_host(s, X, &p, pl.data(), S);
        }
        if (st == MGB_OK) {
            L.src = (int32_t*)a; L.tar = (int32_t*)b; L.weight = (float*)c; L.delay = (float*)e;
            L.plastic = (uint8_t*)p;
            st = create_common(s, h, L, o);
        }
    }
    if (st != MGB_OK) {
        free_all(s);
        delete s;
        return st;
    }
    *out = s;
    return MGB_OK;
}

namespace {

// One population / projection of a (possibly batched) generator description,
// with indices rebased onto the merged network.
struct FlatPop {
    mgb_gen_pop p;
    uint64_t net_seed;
};
struct FlatProj {
    mgb_gen_proj j;
    float dw, dt;
};

// Networks b = 0..nn-1 are laid out one after another: neuron ids, constant
// sets, Poisson streams and populations of network b follow those of b-1.
// Every draw (Poisson, connectors, initial V) is keyed by the network's own
// seeds and by ids local to its streams and populations, so each network
// evolves exactly as it would alone.
mgb_status create_generated(const mgb_gen_desc* gs, int32_t nn, const mgb_opts* o, mgb_sim** out) {
    if (!gs || nn <= 0 || !out) {
        set_error("incomplete generator description");
        return MGB_ERR_INVALID;
    }
    *out = nullptr;
    MGB_TRY(validate_opts(o));
    HostNeurons h;
    h.step = gs[0].step;
    h.stdp = gs[0].stdp ? 1 : 0;
    if (gs[0].stdpconst) std::memcpy(h.stdpconst, gs[0].stdpconst, sizeof(h.stdpconst));
    std::vector<FlatPop> pops;
    std::vector<FlatProj> projs;
    for (int32_t b = 0; b < nn; ++b) {
        const mgb_gen_desc* g = &gs[b];
        if (g->n_pops <= 0 || !g->pops || (g->n_projs > 0 && !g->projs) || !g->lifconst || g->n_csets < 1) {
            set_error("network %d: incomplete generator description", b);
            return MGB_ERR_INVALID;
        }
        float sc[6];
        std::memcpy(sc, h.stdpconst, sizeof(sc));
        if (g->stdpconst) std::memcpy(sc, g->stdpconst, sizeof(sc));
        if (g->step != h.step || (g->stdp ? 1 : 0) != h.stdp ||
            (h.stdp && std::memcmp(sc, h.stdpconst, sizeof(sc)) != 0)) {
            set_error("network %d: the networks of a batch must share the time step and the STDP "
                      "switch and constants", b);
            return MGB_ERR_INVALID;
        }
        const int32_t cs0 = (int32_t)(h.lifconst.size() / 30);
        if (cs0 + g->n_csets > 255) {
            set_error("more than 255 constant sets in the batch");
            return MGB_ERR_UNSUPPORTED;
        }
        h.lifconst.insert(h.lifconst.end(), g->lifconst, g->lifconst + 30 * (size_t)g->n_csets);
        const int32_t st0 = (int32_t)h.stream_seed.size();
        const int32_t nst = std::max(g->n_streams, 1);
        if (st0 + nst > 65535) {
            set_error("more than 65535 Poisson streams in the batch");
            return MGB_ERR_UNSUPPORTED;
        }
        for (int32_t k = 0; k < nst; ++k)
            h.stream_seed.push_back((g->stream_seed && k < g->n_streams) ? g->stream_seed[k] : g->seed);
        const int32_t p0 = (int32_t)pops.size();
        for (int32_t p = 0; p < g->n_pops; ++p) {
            const mgb_gen_pop& P = g->pops[p];
            if (P.num < 0 || P.stream < 0 || P.stream >= nst || P.cset < 0 || P.cset >= g->n_csets ||
                P.type < 0 || P.type > MGB_LIFB) {
                set_error("network %d: population %d is invalid", b, p);
                return MGB_ERR_INVALID;
            }
            FlatPop f{P, g->net_seed};
            f.p.cset += cs0;
            f.p.stream += st0;
            pops.push_back(f);
        }
        for (int32_t j = 0; j < g->n_projs; ++j) {
            const mgb_gen_proj& J = g->projs[j];
            if (J.src < 0 || J.src >= g->n_pops || J.tar < 0 || J.tar >= g->n_pops) {
                set_error("network %d: projection %d references an unknown population", b, j);
                return MGB_ERR_INVALID;
            }
            FlatProj f{J, g->dw, g->dt};
            f.j.src += p0;
            f.j.tar += p0;
            projs.push_back(f);
        }
    }
    const int32_t npops = (int32_t)pops.size(), nprojs = (int32_t)projs.size();
    const int32_t nst = (int32_t)h.stream_seed.size();
    h.stream_base.assign(nst, -1);
    std::vector<int32_t> first(npops + 1, 0);
    int64_t total = 0;
    for (int32_t p = 0; p < npops; ++p) {
        first[p] = (int32_t)total;
        total += pops[p].p.num;
        if (total > kMaxNeurons) {
            set_error("too many neurons");
            return MGB_ERR_UNSUPPORTED;
        }
    }
    first[npops] = (int32_t)total;
    const int32_t n = (int32_t)total;
    h.n = n;
    h.type.resize(n);
    h.cset.resize(n);
    h.stream.resize(n);
    h.rate.resize(n);
    h.v0.resize(n);
    for (int32_t p = 0; p < npops; ++p) {
        const mgb_gen_pop& P = pops[p].p;
        if (h.stream_base[P.stream] < 0) h.stream_base[P.stream] = first[p];
    }
    for (int32_t k = 0; k < nst; ++k)
        if (h.stream_base[k] < 0) h.stream_base[k] = 0;
    for (int32_t p = 0; p < npops; ++p) {
        const mgb_gen_pop& P = pops[p].p;
        for (int32_t j = 0; j < P.num; ++j) {
            const int32_t i = first[p] + j;
            h.type[i] = (uint8_t)P.type;
            h.cset[i] = (uint8_t)P.cset;
            h.stream[i] = (uint16_t)P.stream;
            h.rate[i] = P.rate;
            h.v0[i] = vinit_draw(pops[p].net_seed, (uint32_t)(i - h.stream_base[P.stream]),
                                 P.v_min, P.v_max);
        }
    }
    // list sizes
    std::vector<uint64_t> base(nprojs + 1, 0);
    mgb_sim* s = new (std::nothrow) mgb_sim();
    if (!s) {
        set_error("host allocation failed");
        return MGB_ERR_OOM;
    }
    mgb_status st = init_device(s, o);
    Scratch X;
    std::vector<std::vector<int64_t>> bern_off(nprojs);
    std::vector<uint32_t*> bern_cnt(nprojs, nullptr);
    for (int32_t j = 0; j < nprojs && st == MGB_OK; ++j) {
        const mgb_gen_proj& J = projs[j].j;
        const int32_t sn = pops[J.src].p.num, tn = pops[J.tar].p.num;
        uint64_t cntj = 0;
        if (J.kind == MGB_CONN_FIXED_INDEGREE) {
            if (J.k < 0 || (J.k > 0 && sn == 0)) {
                set_error("projection %d: invalid in-degree", j);
                st = MGB_ERR_INVALID;
                break;
            }
            cntj = (uint64_t)tn * (uint64_t)J.k;
        } else if (J.kind == MGB_CONN_BERNOULLI) {
            GenProjArgs a{};
            a.sn = sn; a.tn = tn; a.index = (uint32_t)J.index; a.seed = J.seed;
            a.thr = (uint64_t)((double)J.p * 4294967296.0);
            st = X.get(&bern_cnt[j], (size_t)std::max(sn, 1));
            if (st != MGB_OK) break;
            hipError_t e = gen_bern(a, bern_cnt[j], nullptr, false, s->stream);
            if (e != hipSuccess) {
                set_error("Bernoulli count kernel failed: %s", hipGetErrorString(e));
                st = MGB_ERR_HIP;
                break;
            }
            std::vector<uint32_t> rc((size_t)sn);
            if (sn) {
                e = hipMemcpyAsync(rc.data(), bern_cnt[j], 4 * (size_t)sn, hipMemcpyDeviceToHost, s->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
                if (e != hipSuccess) {
                    set_error("copy failed: %s", hipGetErrorString(e));
                    st = MGB_ERR_HIP;
                    break;
                }
            }
            bern_off[j].assign((size_t)sn + 1, 0);
            for (int32_t m = 0; m < sn; ++m) bern_off[j][m + 1] = bern_off[j][m] + rc[m];
            cntj = (uint64_t)bern_off[j][sn];
        } else {
            set_error("projection %d: connector kind %d is not generated on the device", j, J.kind);
            st = MGB_ERR_UNSUPPORTED;
            break;
        }
        base[j + 1] = base[j] + cntj;
    }
    if (st == MGB_OK && base[nprojs] >= 0xFFFFFFFFull) {
        set_error("%llu synapses exceed the 2^32-1 limit of one simulator",
                  (unsigned long long)base[nprojs]);
        st = MGB_ERR_UNSUPPORTED;
    }
    if (st == MGB_OK) {
        DevList L;
        L.S = base[nprojs];
        st = X.get(&L.src, L.S);
        if (st == MGB_OK) st = X.get(&L.tar, L.S);
        if (st == MGB_OK) st = X.get(&L.weight, L.S);
        if (st == MGB_OK) st = X.get(&L.delay, L.S);
        if (st == MGB_OK) st = X.get(&L.plastic, L.S);
        for (int32_t j = 0; j < nprojs && st == MGB_OK; ++j) {
            const mgb_gen_proj& J = projs[j].j;
            GenProjArgs a{};
            a.sn = pops[J.src].p.num; a.tn = pops[J.tar].p.num;
            a.s0 = first[J.src]; a.t0 = first[J.tar]; a.k = J.k;
            a.index = (uint32_t)J.index; a.seed = J.seed;
            a.thr = (uint64_t)((double)J.p * 4294967296.0);
            a.wmin = J.wmin; a.wmax = J.wmax; a.dmin = J.dmin; a.dmax = J.dmax;
            a.dw = projs[j].dw; a.dt = projs[j].dt; a.plastic_flag = J.plastic ? 1 : 0;
            a.base = base[j];
            a.src = L.src; a.tar = L.tar; a.weight = L.weight; a.delay = L.delay; a.plastic = L.plastic;
            hipError_t e;
            if (J.kind == MGB_CONN_FIXED_INDEGREE) {
                e = gen_fixk(a, s->stream);
            } else {
                int64_t* off = nullptr;
                st = X.get(&off, bern_off[j].size());
                if (st != MGB_OK) break;
                e = hipMemcpyAsync(off, bern_off[j].data(), 8 * bern_off[j].size(), hipMemcpyHostToDevice,
                                   s->stream);
                if (e == hipSuccess) e = gen_bern(a, bern_cnt[j], off, true, s->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            }
            if (e != hipSuccess) {
                set_error("generator kernel failed: %s", hipGetErrorString(e));
                st = MGB_ERR_HIP;
            }
        }
        if (st == MGB_OK) st = create_common(s, h, L, o);
    }
    if (st != MGB_OK) {
        free_all(s);
        delete s;
        return st;
    }
    *out = s;
    return MGB_OK;
}

}  // namespace

mgb_status mgb_create_generated(const mgb_gen_desc* g, const mgb_opts* o, mgb_sim** out) {
    return create_generated(g, 1, o, out);
}

mgb_status mgb_create_batch_generated(const mgb_gen_desc* nets, int32_t n_nets, const mgb_opts* o,
                                      mgb_sim** out) {
    return create_generated(nets, n_nets, o, out);
}

mgb_status mgb_run(mgb_sim* s, int32_t nsteps) {
    if (!s || nsteps < 0) {
        set_error("invalid arguments to mgb_run");
        return MGB_ERR_INVALID;
    }
    if (s->failed) {
        set_error("an earlier persistent launch of this simulator timed out; its state is void");
        return MGB_ERR_HIP;
    }
    MGB_HIP_TRY(hipSetDevice(s->device));
    if (s->profile) {
        // the previous run's events may still be pending; reuse the pool after it ends
        MGB_HIP_TRY(hipStreamSynchronize(s->stream));
        s->ev_used = 0;
        s->spans.clear();
    }
    MGB_HIP_TRY(hipEventRecord(s->ev0, s->stream));
    int32_t done = 0;
    if (s->use_graph) {
        // graph block: a whole number of windows, about 500 steps
        const int32_t blk = std::max(1, 500 / s->window) * s->window;
        if (nsteps >= blk) {
            MGB_TRY(ensure_graph(s, blk));
            while (nsteps - done >= blk) {
                // a fused-path replay is its windows' fused launches (+ one finalize)
                SpanGuard gg{s, kSpanGraph, s->profile != 0 && s->tiled != 0, blk / s->window};
                MGB_TRY(gg.begin());
                MGB_HIP_TRY(hipGraphLaunch(s->gexec, s->stream));
                MGB_TRY(gg.end());
                s->launches += 1;
                done += blk;
            }
        }
    }
    if (nsteps - done > 0) MGB_TRY(issue_windows(s, nsteps - done, true));
    // the waits' error word comes back with the run (the finalize kernel writes it to
    // pinned host memory; mgb_sync reads it)
    if (s->persist && s->h_err) s->err_pending = true;
    MGB_HIP_TRY(hipEventRecord(s->ev1, s->stream));
    s->timed = true;
    return MGB_OK;
}

mgb_status mgb_sync(mgb_sim* s) {
    if (!s) {
        set_error("null handle");
        return MGB_ERR_INVALID;
    }
    MGB_HIP_TRY(hipSetDevice(s->device));
    MGB_HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->err_pending) {
        s->failed = s->failed || *s->h_err != 0u;
        s->err_pending = false;
    }
    if (s->failed) {
        set_error("a persistent launch's wait between windows exceeded its bound (%llu clock ticks): a workgroup "
                  "was not resident or a window ran longer than the bound; the simulator's state is void",
                  (unsigned long long)s->spin_limit);
        return MGB_ERR_HIP;
    }
    return MGB_OK;
}

mgb_status mgb_set_launch(mgb_sim* s, int32_t block, int32_t chunk) {
    if (!s || block < 0 || (block > 0 && (block < 64 || block > 1024 || block % 64 != 0)) ||
        !(chunk == 0 || chunk == 1 || chunk == 2 || chunk == 4 || chunk == 8 || chunk == 16)) {
        set_error("invalid launch shape (block %d, chunk %d)", block, chunk);
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    if (block > 0) {
        if (s->tiled) s->tblock = block;
        else s->prop_block = block;
    }
    if (chunk > 0) s->chunk = chunk;
    update_persist(s);
    if (s->gexec) {
        (void)hipGraphExecDestroy(s->gexec);
        s->gexec = nullptr;
    }
    if (s->graph) {
        (void)hipGraphDestroy(s->graph);
        s->graph = nullptr;
    }
    return MGB_OK;
}

mgb_status mgb_get_info(mgb_sim* s, mgb_info* out) {
    if (!s || !out) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    Ctl c;
    MGB_HIP_TRY(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->n_neurons = s->n;
    out->ring_len = s->D;
    out->min_delay = s->min_delay;
    out->window = s->window;
    out->n_synapses = s->S;
    out->step = c.base_step;
    out->syn_events = (int64_t)c.se;
    out->spikes_recorded = (int64_t)c.rec_cnt;
    out->weight_format = s->wf;
    out->n_classes = s->n_tbl;
    out->fix_scale_exc = s->scale_e;
    out->fix_scale_inh = s->scale_i;
    out->device_bytes = s->bytes;
    out->launches = s->launches;
    out->spikes = (int64_t)c.rec_cnt;
    out->propagation = s->tiled ? MGB_PROP_TILED : MGB_PROP_DIRECT;
    out->tile_size = s->B;
    out->tiles = s->G;
    out->block = s->tiled ? s->tblock : s->prop_block;
    out->segments = std::max<int32_t>(1, (int32_t)s->h_seg_lo.size() - 1);
    // (kTStdpB: 4, plus the FP32 weight of a plastic synapse)
    out->word_bytes = !s->tiled ? 0 : (s->tkind == kTClass16 ? 2 : (s->tkind == kTF32 || s->tkind == kTStdp ? 8 : 4));
    out->persistent = s->persist;
    out->last_run_ms = 0.0;
    if (s->timed) {
        float ms = 0.f;
        MGB_HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        out->last_run_ms = ms;
    }
    for (const auto& sp : s->spans) {
        float ms = 0.f;
        MGB_HIP_TRY(hipEventElapsedTime(&ms, sp.a, sp.b));
        if (sp.kind == kSpanNeuron) {
            out->neuron_ms += ms;
            out->neuron_launches += 1;
        } else if (sp.kind == kSpanProp || sp.kind == kSpanGraph) {
            out->prop_ms += ms;
            out->prop_launches += sp.count;
        } else {
            out->stdp_ms += ms;
            out->stdp_launches += 1;
        }
    }
    return MGB_OK;
}

mgb_status mgb_read_spikes(mgb_sim* s, int64_t cap, int32_t* steps, int32_t* ids, int64_t* n_out) {
    if (!s || cap < 0 || (cap > 0 && (!steps || !ids)) || !n_out) {
        set_error("invalid arguments to mgb_read_spikes");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    Ctl c;
    MGB_HIP_TRY(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const uint64_t have = std::min<uint64_t>(c.rec_cnt, (uint64_t)s->rec_cap);
    std::vector<uint64_t> r(have);
    if (have) MGB_HIP_TRY(hipMemcpy(r.data(), s->rec, 8 * have, hipMemcpyDeviceToHost));
    std::sort(r.begin(), r.end());
    const uint64_t w = std::min<uint64_t>(have, (uint64_t)cap);
    for (uint64_t k = 0; k < w; ++k) {
        steps[k] = (int32_t)(r[k] >> 32);
        ids[k] = (int32_t)(r[k] & 0xFFFFFFFFu);
    }
    *n_out = (int64_t)w;
    if (c.rec_cnt > (uint64_t)s->rec_cap || c.rec_overflow || w < have) {
        set_error("spike record truncated: %llu spikes, capacity %lld, caller capacity %lld",
                  (unsigned long long)c.rec_cnt, (long long)s->rec_cap, (long long)cap);
        return MGB_ERR_CAPACITY;
    }
    return MGB_OK;
}

mgb_status mgb_clear_spikes(mgb_sim* s) {
    if (!s) {
        set_error("null handle");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    Ctl c;
    MGB_HIP_TRY(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    c.rec_cnt = 0;
    c.rec_overflow = 0;
    MGB_HIP_TRY(hipMemcpy(s->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice));
    return MGB_OK;
}

mgb_status mgb_read_state(mgb_sim* s, int32_t field, void* out) {
    if (!s || !out) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    const void* src = nullptr;
    switch (field) {
        case MGB_F_V: src = s->V; break;
        case MGB_F_IEXC: src = s->Ie; break;
        case MGB_F_IINH: src = s->Ii; break;
        case MGB_F_REFR: src = s->refr; break;
        case MGB_F_LAST: src = s->last; break;
        case MGB_F_CNT: src = s->cnt; break;
        default:
            set_error("unknown state field %d", field);
            return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    if (field == MGB_F_LAST && s->tiled && stdp_kind(s->tkind)) {
        // the next window reads buffer base_window % 3
        Ctl c;
        MGB_HIP_TRY(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
        src = s->last + (size_t)(c.base_window % 3) * (size_t)s->npad;
    }
    MGB_HIP_TRY(hipMemcpy(out, src, 4 * (size_t)s->n, hipMemcpyDeviceToHost));
    return MGB_OK;
}

// ---- checkpoint -------------------------------------------------------------

namespace {

constexpr uint64_t kCkptMagic = 0x3154504B4342474Dull;   // "MGBCKPT1"

// Layout fields a checkpoint must match (the simulator it is loaded into must
// come from the same network and options).
struct CkptHeader {
    uint64_t magic;
    int32_t abi, n, npad, D, window, tiled, tkind, B, G, nseg, mode, wide_kind, stdp;
    int64_t S, n_rec;
    uint64_t payload;
    uint64_t tiling;        // hash of the tile boundaries (the ring slices' layout)
};

struct CkptSeg {
    void* p;
    size_t bytes;
};

// Every device array a run changes, in a fixed order.
std::vector<CkptSeg> ckpt_segments(const mgb_sim* s, uint64_t n_rec) {
    const size_t np = (size_t)s->npad;
    const size_t nlast = (s->tiled && stdp_kind(s->tkind)) ? 3 : 1;
    std::vector<CkptSeg> v = {{s->V, 4 * np}, {s->Ie, 4 * np}, {s->Ii, 4 * np},
                              {s->refr, 4 * np}, {s->last, 4 * nlast * np}, {s->cnt, 4 * np},
                              {s->ctl, sizeof(Ctl)},
                              {s->win_buf, 4 * (size_t)(s->tiled ? 3 : 2) * (size_t)s->win_stride}};
    if (s->tiled) {
        v.push_back({s->ring_t, 4 * (size_t)s->G * s->ring_slab});
        v.push_back({s->tcnt, 12 * (size_t)s->nseg});
    } else if (s->mode == kRingClass) {
        v.push_back({s->ring, 4 * (size_t)s->D * np});
    } else {
        v.push_back({s->ring_e, 4 * (size_t)s->D * np});
        v.push_back({s->ring_i, 4 * (size_t)s->D * np});
    }
    if (s->stdp) {
        if (stdp_kind(s->tkind) && s->tiled) v.push_back({s->w_t, 4 * (size_t)s->n_tslots});
        else v.push_back({s->w32, 4 * (size_t)s->S});
    }
    if (s->hist) {
        v.push_back({s->hist, 4 * np * kStdpHist});
        v.push_back({s->tf, 4 * np});
        v.push_back({s->cntf, 4 * np});
        v.push_back({s->fcnt, 8});
    }
    if (n_rec) v.push_back({s->rec, 8 * (size_t)n_rec});
    return v;
}

CkptHeader ckpt_header(const mgb_sim* s, uint64_t n_rec) {
    CkptHeader h{};
    h.magic = kCkptMagic;
    h.abi = MGB_ABI_VERSION;
    h.n = s->n; h.npad = s->npad; h.D = s->D; h.window = s->window; h.tiled = s->tiled;
    h.tkind = s->tkind; h.B = s->B; h.G = s->G; h.nseg = s->nseg; h.mode = s->mode;
    h.wide_kind = s->wide_kind; h.stdp = s->stdp; h.S = s->S; h.n_rec = (int64_t)n_rec;
    h.tiling = s->tile_hash;
    for (const CkptSeg& g : ckpt_segments(s, n_rec)) h.payload += g.bytes;
    return h;
}

mgb_status ckpt_records(mgb_sim* s, uint64_t* n_rec) {
    Ctl c;
    MGB_HIP_TRY(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    *n_rec = std::min<uint64_t>(c.rec_cnt, (uint64_t)s->rec_cap);
    return MGB_OK;
}

}  // namespace

mgb_status mgb_checkpoint_size(mgb_sim* s, int64_t* bytes) {
    if (!s || !bytes) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    uint64_t n_rec = 0;
    MGB_TRY(ckpt_records(s, &n_rec));
    *bytes = (int64_t)(sizeof(CkptHeader) + ckpt_header(s, n_rec).payload);
    return MGB_OK;
}

mgb_status mgb_checkpoint_save(mgb_sim* s, void* buf, int64_t bytes) {
    if (!s || !buf) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    uint64_t n_rec = 0;
    MGB_TRY(ckpt_records(s, &n_rec));
    const CkptHeader h = ckpt_header(s, n_rec);
    if (bytes < (int64_t)(sizeof(h) + h.payload)) {
        set_error("checkpoint buffer of %lld bytes, %llu needed", (long long)bytes,
                  (unsigned long long)(sizeof(h) + h.payload));
        return MGB_ERR_CAPACITY;
    }
    uint8_t* out = static_cast<uint8_t*>(buf);
    std::memcpy(out, &h, sizeof(h));
    size_t off = sizeof(h);
    for (const CkptSeg& g : ckpt_segments(s, n_rec)) {
        if (g.bytes) MGB_HIP_TRY(hipMemcpy(out + off, g.p, g.bytes, hipMemcpyDeviceToHost));
        off += g.bytes;
    }
    return MGB_OK;
}

mgb_status mgb_checkpoint_load(mgb_sim* s, const void* buf, int64_t bytes) {
    if (!s || !buf || bytes < (int64_t)sizeof(CkptHeader)) {
        set_error("invalid checkpoint arguments");
        return MGB_ERR_INVALID;
    }
    CkptHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (h.magic != kCkptMagic || h.abi != MGB_ABI_VERSION) {
        set_error("not a checkpoint of this library version");
        return MGB_ERR_INVALID;
    }
    if ((uint64_t)h.n_rec > (uint64_t)s->rec_cap) {
        set_error("checkpoint holds %lld recorded spikes, record capacity is %lld", (long long)h.n_rec,
                  (long long)s->rec_cap);
        return MGB_ERR_CAPACITY;
    }
    const CkptHeader want = ckpt_header(s, (uint64_t)h.n_rec);
    if (h.n != want.n || h.npad != want.npad || h.D != want.D || h.window != want.window ||
        h.tiled != want.tiled || h.tkind != want.tkind || h.B != want.B || h.G != want.G ||
        h.nseg != want.nseg || h.mode != want.mode || h.wide_kind != want.wide_kind ||
        h.stdp != want.stdp || h.S != want.S || h.payload != want.payload || h.tiling != want.tiling) {
        set_error("checkpoint layout differs from this simulator (build it from the same network and "
                  "options)");
        return MGB_ERR_INVALID;
    }
    if (bytes < (int64_t)(sizeof(h) + h.payload)) {
        set_error("checkpoint truncated");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    const uint8_t* in = static_cast<const uint8_t*>(buf);
    size_t off = sizeof(h);
    for (const CkptSeg& g : ckpt_segments(s, (uint64_t)h.n_rec)) {
        if (g.bytes) MGB_HIP_TRY(hipMemcpy(g.p, in + off, g.bytes, hipMemcpyHostToDevice));
        off += g.bytes;
    }
    // the persistent waits' flags hold window indices of the run before (epochs)
    if (s->bar) MGB_HIP_TRY(hipMemset(s->bar, 0, 4 * (size_t)s->G * kBarStride));
    return MGB_OK;
}

mgb_status mgb_fire_counts(mgb_sim* s, int32_t* out) { return mgb_read_state(s, MGB_F_CNT, out); }

mgb_status mgb_read_tiles(mgb_sim* s, int32_t* tile_lo) {
    if (!s || !tile_lo) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    if (!s->tiled) {
        set_error("mgb_read_tiles: the simulator runs the direct path (no tiles)");
        return MGB_ERR_UNSUPPORTED;
    }
    MGB_HIP_TRY(hipSetDevice(s->device));
    MGB_HIP_TRY(hipStreamSynchronize(s->stream));
    MGB_HIP_TRY(hipMemcpy(tile_lo, s->tile_lo, 4 * ((size_t)s->G + 1), hipMemcpyDeviceToHost));
    return MGB_OK;
}

namespace {

// Fused STDP: the kernel applies a window's STDP in the next window's launch.
// Writes to w32 (CSR order) the weights the reference holds after the last
// step run: a copy of w_t (in tmp, n_tslots floats) with the pending window's
// updates applied.
hipError_t current_weights_into(mgb_sim* s, float* tmp) {
    const size_t S = (size_t)s->n_tslots;
    TiledArgs ta = tiled_args(s);
    ta.w_t = tmp;
    ta.apply_only = 1;
    hipError_t e = hipMemcpyAsync(tmp, s->w_t, S * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = launch_tiled(ta, 0, 0, 1, s->tblock, s->stream);
    if (e == hipSuccess) e = launch_stdp_flush(ta, 0, 0, s->stream);
    // the list this used is empty before the next window appends to it
    if (e == hipSuccess) e = hipMemsetAsync(s->fcnt, 0, 8, s->stream);
    if (e == hipSuccess) e = scatter_weights(s->tperm, tmp, (uint64_t)S, s->w32, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e;
}

mgb_status current_weights(mgb_sim* s) {
    const size_t S = (size_t)s->n_tslots;
    float* tmp = nullptr;
    hipError_t e = hipMalloc(&tmp, std::max<size_t>(S, 1) * 4);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) for the weight read-back failed: %s", S * 4, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? MGB_ERR_OOM : MGB_ERR_HIP;
    }
    e = current_weights_into(s, tmp);
    const hipError_t f = hipFree(tmp);
    if (e == hipSuccess) e = f;
    if (e != hipSuccess) {
        set_error("weight read-back: %s", hipGetErrorString(e));
        return MGB_ERR_HIP;
    }
    return MGB_OK;
}

std::mutex g_scratch_mu;
float* g_scratch_at[kNoAllocDevices] = {};

// `device` must be current.
mgb_status scratch_on(int device, float** out) {
    if (device < 0 || device >= kNoAllocDevices) {
        set_error("device %d outside the %d devices the no-allocation read-back supports", device,
                  kNoAllocDevices);
        return MGB_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_at[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
            return MGB_ERR_HIP;
        }
        g_scratch_at[device] = static_cast<float*>(p);
    }
    *out = g_scratch_at[device];
    return MGB_OK;
}

}  // namespace

mgb_status mgb_read_csr(mgb_sim* s, int64_t* row_ptr, int32_t* col, int32_t* delay, float* weight) {
    if (!s) {
        set_error("null handle");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    const size_t S = (size_t)s->S;
    if (row_ptr) MGB_HIP_TRY(hipMemcpy(row_ptr, s->row_ptr, 8 * ((size_t)s->n + 1), hipMemcpyDeviceToHost));
    if (!col && !delay && !weight) return MGB_OK;
    if (s->mode == kRingClass) {
        std::vector<uint32_t> syn(S);
        if (S) MGB_HIP_TRY(hipMemcpy(syn.data(), s->syn, 4 * S, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < S; ++k) {
            const uint32_t x = syn[k];
            const int32_t t = (int32_t)(x & kIdMask);
            if (col) col[k] = t;
            if (delay) delay[k] = (int32_t)((x >> 26) & 31u) + 1;
            if (weight) weight[k] = s->class_w[2 * (size_t)s->h_tbl[t] + (x >> 31)];
        }
        return MGB_OK;
    }
    if (col && S) MGB_HIP_TRY(hipMemcpy(col, s->col, 4 * S, hipMemcpyDeviceToHost));
    std::vector<uint8_t> dl;
    if (delay || (weight && s->wide_kind != kWideF32)) {
        dl.resize(S);
        if (S) MGB_HIP_TRY(hipMemcpy(dl.data(), s->dly, S, hipMemcpyDeviceToHost));
    }
    if (delay)
        for (size_t k = 0; k < S; ++k) delay[k] = (int32_t)(dl[k] & 127u);
    if (weight && S) {
        if (s->tiled && stdp_kind(s->tkind)) MGB_TRY(current_weights(s));
        MGB_HIP_TRY(hipMemcpy(weight, s->w32, 4 * S, hipMemcpyDeviceToHost));
        if (s->wide_kind == kWideBF16) {
            std::vector<uint16_t> w16(S);
            std::vector<float> rs((size_t)s->n);
            std::vector<int64_t> rp((size_t)s->n + 1);
            MGB_HIP_TRY(hipMemcpy(w16.data(), s->w16, 2 * S, hipMemcpyDeviceToHost));
            MGB_HIP_TRY(hipMemcpy(rs.data(), s->row_scale, 4 * (size_t)s->n, hipMemcpyDeviceToHost));
            MGB_HIP_TRY(hipMemcpy(rp.data(), s->row_ptr, 8 * ((size_t)s->n + 1), hipMemcpyDeviceToHost));
            for (int32_t i = 0; i < s->n; ++i)
                for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
                    if (dl[k] & 128u) continue;
                    const uint32_t b = (uint32_t)w16[k] << 16;
                    float v;
                    std::memcpy(&v, &b, 4);
                    weight[k] = v * rs[i];                  // as the kernels decode it
                }
        } else if (s->wide_kind == kWideFP8) {
            std::vector<uint8_t> w8(S);
            std::vector<float> rs((size_t)s->n);
            std::vector<int64_t> rp((size_t)s->n + 1);
            MGB_HIP_TRY(hipMemcpy(w8.data(), s->w8, S, hipMemcpyDeviceToHost));
            MGB_HIP_TRY(hipMemcpy(rs.data(), s->row_scale, 4 * (size_t)s->n, hipMemcpyDeviceToHost));
            MGB_HIP_TRY(hipMemcpy(rp.data(), s->row_ptr, 8 * ((size_t)s->n + 1), hipMemcpyDeviceToHost));
            for (int32_t i = 0; i < s->n; ++i)
                for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
                    if (dl[k] & 128u) continue;
                    const uint8_t b = w8[k];
                    const uint32_t sg = (uint32_t)(b >> 7) << 31, e = (b >> 3) & 15u, m = b & 7u;
                    float v;
                    if (e == 0) {
                        v = (float)m * 0x1p-9f;
                        uint32_t u;
                        std::memcpy(&u, &v, 4);
                        u |= sg;
                        std::memcpy(&v, &u, 4);
                    } else {
                        const uint32_t u = sg | ((e + 120u) << 23) | (m << 20);
                        std::memcpy(&v, &u, 4);
                    }
                    weight[k] = v * rs[i];
                }
        }
    }
    return MGB_OK;
}

mgb_status mgb_read_weights(mgb_sim* s, float* out) {
    if (!out) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    return mgb_read_csr(s, nullptr, nullptr, nullptr, out);
}

mgb_status mgb_read_weights_noalloc(mgb_sim* s, float* out) {
    if (!s || !out) {
        set_error("null argument");
        return MGB_ERR_INVALID;
    }
    MGB_TRY(mgb_sync(s));
    const uint64_t S = (uint64_t)s->S;
    if (S == 0) return MGB_OK;
    const bool pending = s->tiled && stdp_kind(s->tkind);   // fused STDP: updates still to apply
    const bool decode = s->mode == kRingClass || s->wide_kind != kWideF32;
    uint64_t need = 0;
    if (pending) need = s->n_tslots;
    if (decode) need = std::max(need, S);
    if (need > kNoAllocSlots) {
        set_error("mgb_read_weights_noalloc: the network needs %llu scratch floats, the device scratch holds %llu "
                  "(use mgb_read_weights)", (unsigned long long)need, (unsigned long long)kNoAllocSlots);
        return MGB_ERR_UNSUPPORTED;
    }
    float* scr = nullptr;
    if (need) MGB_TRY(scratch_on(s->device, &scr));
    if (pending) {
        const hipError_t e = current_weights_into(s, scr);
        if (e != hipSuccess) {
            set_error("weight read-back: %s", hipGetErrorString(e));
            return MGB_ERR_HIP;
        }
    }
    const float* from = s->w32;
    if (s->mode == kRingClass) {
        MGB_HIP_TRY(class_weights(s->syn, s->tbl, s->tbl_w, S, scr, s->stream));
        from = scr;
    } else if (decode) {
        MGB_HIP_TRY(wide_weights(s->row_ptr, s->n, s->dly, s->w32, s->wide_kind == kWideBF16 ? s->w16 : nullptr,
                                 s->wide_kind == kWideFP8 ? s->w8 : nullptr, s->row_scale, scr, s->stream));
        from = scr;
    }
    MGB_HIP_TRY(hipMemcpyAsync(out, from, 4 * S, hipMemcpyDeviceToHost, s->stream));
    MGB_HIP_TRY(hipStreamSynchronize(s->stream));
    return MGB_OK;
}

mgb_status mgb_destroy(mgb_sim* s) {
    if (!s) return MGB_OK;
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return MGB_OK;
}

}  // extern "C"

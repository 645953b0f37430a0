// mgb_internal.h -- device data layout and kernel launchers of the HIP step.
//
// HBM layout of one simulator (all device arrays, SoA, neuron ids 0..n-1 in the
// reference's id order; npad = n rounded up to 256):
//   neurons   type u8, cset u8, tbl u16, stream u16, rate f32, V f32, Ie f32,
//             Ii f32, refr i32, last i32, cnt i32                    [npad]
//   CSR       row_ptr i64 [n+1]; CLASS format: syn u32 [S] = target (26 bits)
//             | delay-1 (5 bits) | inhibitory bit; wide formats: col i32 [S],
//             dly u8 [S] (bit 7 = plastic), w f32|bf16|fp8 [S]
//   STDP      in_ptr i64 [n+1], in_pos u32 [Sp], in_src i32 [Sp] (plastic only)
//   ring      CLASS: u32 [D][npad] (excitatory count | inhibitory count << 16);
//             wide: i32 [D][npad] twice (fixed-point excitatory / inhibitory)
//   window    spike lists u32 [2][npad*window] (id | step-in-window << 26)
//   record    u64 [cap] (step << 32 | id)
// The ring is slot-major so the neuron kernel reads one coalesced row per step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mgbrain_hip.h"

namespace mgb {

void set_error(const char* fmt, ...);

#define MGB_HIP_TRY(call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::mgb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                    \
            return e_ == hipErrorOutOfMemory ? MGB_ERR_OOM : MGB_ERR_HIP;                 \
        }                                                                                  \
    } while (0)

constexpr int kNeuBlock = 256;
constexpr int kMaxWindow = 63;        // step-in-window field is 6 bits
constexpr uint32_t kIdMask = 0x3FFFFFFu;
constexpr int kMaxNeurons = 1 << 26;
constexpr int kClassMaxDelay = 32;    // 5-bit delay-1 field
constexpr int kWideMaxDelay = 127;    // 7-bit delay field

// Device control block (one per simulator).
struct Ctl {
    int64_t base_step;      // first step of the current launch block
    int64_t base_window;    // windows issued before the block (spike-list parity)
    uint32_t win_cnt[2];    // spikes appended to each window list
    int32_t rec_overflow;
    int32_t pad;
    uint64_t rec_cnt;       // spikes recorded (may exceed capacity)
    uint64_t se;            // synaptic events
    // fused tiled path: three spike lists (write k%3, read (k-1)%3, clear (k+1)%3);
    // their per-segment counters live in TiledArgs::tcnt
    int64_t tbase[3];       // first step of the window that wrote each list
};


// Constant set: LIFECONSTS[30] (mgcore.cu:32-63) + refractory steps in c[30].
struct CSet {
    float c[32];
};

struct NeuArgs {
    int32_t n, npad, D;
    const uint8_t* type;
    const uint8_t* cset;
    const uint16_t* tbl;
    const uint16_t* stream;
    const float* rate;
    float* V;
    float* Ie;
    float* Ii;
    int32_t* refr;
    int32_t* last;
    int32_t* cnt;
    const CSet* csets;
    const int32_t* stream_base;
    const uint64_t* stream_seed;
    // CLASS ring + class tables
    uint32_t* ring;
    const uint32_t* tbl_off;   // [n_tbl][2]: offsets of exc / inh seq-sum tables
    const float* seq;          // seq[off + k] = k-fold sequential FP32 sum of the class weight
    const float* tbl_w;        // [n_tbl][2]: the exc / inh class weights
    // fixed-point ring
    int32_t* ring_e;
    int32_t* ring_i;
    float inv_e, inv_i;
    // batched networks with fixed-point formats: each segment's (scale, 1 / scale)
    // (a network's own, as if it ran alone), segment of each neuron; null for one
    const float2* seg_scale;
    const uint16_t* seg_of;
    Ctl* ctl;
    uint32_t* win_buf;
    uint32_t win_stride;
};

struct PropArgs {
    int32_t n, npad, D;
    int32_t split;             // waves sharing one spike's row
    const int64_t* row_ptr;
    const uint32_t* syn;       // CLASS format
    const int32_t* col;        // wide formats
    const uint8_t* dly;
    const float* w32;
    const uint16_t* w16;
    const uint8_t* w8;
    const float* row_scale;    // FP8: per-source scale
    uint32_t* ring;
    int32_t* ring_e;
    int32_t* ring_i;
    float scale_e, scale_i;
    const float2* seg_scale;   // per segment (scale, 1 / scale); null: scale_e / scale_i
    const uint16_t* seg_of;
    Ctl* ctl;
    const uint32_t* win_buf;
    uint32_t win_stride;
    uint64_t* rec;
    uint64_t rec_cap;
};

struct StdpArgs {
    int32_t n;
    int32_t split;             // waves sharing one spike's out- / in-row
    const int64_t* row_ptr;
    const int32_t* col;
    const uint8_t* dly;
    float* w32;
    const int64_t* in_ptr;
    const uint32_t* in_pos;
    const int32_t* in_src;
    const int32_t* last;
    const float* ltp;          // ltp[k] = A_LTP * expf(-k / TAU_LTP), k >= 1
    const float* ltd;          // ltd[k] = A_LTD * expf(-k / TAU_LTD)
    int32_t lut_n;
    float w_max, w_min;
    Ctl* ctl;
    const uint32_t* win_buf;
    uint32_t win_stride;
};

// Fused tiled window kernel: workgroup g owns targets [tile_lo[g], tile_lo[g+1])
// (at most B) and keeps their ring slice ring_t[g][D][B] in LDS while it (1)
// adds the previous window's spikes arriving in the tile (tile-sorted CSR copy
// syn_t, per-row tile offsets) and (2) advances the tile's neurons.
//
// Segments: contiguous id ranges with no synapse between them (independent
// networks of a batch; one segment for a single network).  Every tile lies in
// one segment, each segment has its own spike lists, and a source row of
// segment q holds offsets for q's tiles only: offs[seg_obase[q] +
// (src - seg_lo[q]) * (Gq + 1) + local tile], Gq = tiles of q.
// Synapse word kinds of the fused kernel.  kTClass16 is the CLASS word of a
// tiling with B <= 1024: tile-local target (10 bits) | delay-1 << 10 | sign << 15.
// kTStdp: the kTF32 word with a plastic flag at bit 16, FP32 weights in tile
// order (w_t, updated in place), pair-based STDP inside the kernel.
// kTStdpB: kTStdp with BF16 static weights: bits 16-31 of the word hold the BF16
// code of a static synapse's w / row scale, or 0xFFFF for a plastic synapse,
// whose FP32 weight is in w_t (static positions of w_t are never read).
enum TiledKind { kTClass = 0, kTBF16 = 1, kTFP8 = 2, kTF32 = 3, kTClass16 = 4, kTStdp = 5, kTStdpB = 6 };
// The fused kinds with STDP.
inline bool stdp_kind(int k) { return k == kTStdp || k == kTStdpB; }

struct TiledArgs {
    int32_t n, D, B, G;
    int32_t chunk;          // synapses per work item in the propagation phase
    int32_t kind;           // TiledKind
    float scale_e, scale_i, inv_e, inv_i;   // fixed-point scales of the wide kinds
    const float2* seg_scale;                // per segment (scale, 1 / scale); null: the above
    float* w_t;             // kTF32 / kTStdp / kTStdpB: weights in tile order
    const float* row_scale; // kTFP8 / kTBF16: per-source scale
    // kTStdpB: per source, its row scale's bits | 1 << 31 when its row holds a
    // plastic synapse; per tile, 1 when a plastic synapse targets it (the kernel
    // loads w_t only for spikes of such rows into such tiles)
    const uint32_t* row_aux;
    const uint8_t* tile_pl;
    int32_t nseg;
    int32_t window;         // spike-list entries reserved per neuron (steps per window)
    const int32_t* tile_lo;     // [G+1]
    const uint16_t* tile_seg;   // [G]
    const int32_t* seg_lo;      // [nseg+1] first neuron of each segment
    const int32_t* seg_tile0;   // [nseg+1] first tile of each segment
    const uint64_t* seg_obase;  // [nseg+1] first offset entry of each segment
    uint32_t* tcnt;             // [3][nseg] entries in each segment's spike lists
    const uint8_t* type;
    const uint8_t* cset;
    const uint16_t* tbl;
    const uint16_t* stream;
    const float* rate;
    float* V;
    float* Ie;
    float* Ii;
    int32_t* refr;
    int32_t* last;
    int32_t* cnt;
    const CSet* csets;
    const int32_t* stream_base;
    const uint64_t* stream_seed;
    const uint32_t* tbl_off;
    const float* seq;
    const float* tbl_w;
    const int64_t* row_ptr;
    const uint32_t* syn_t;
    const uint16_t* syn16;      // kTClass16 words
    const uint16_t* offs16;
    const uint32_t* offs32;
    const int64_t* prow;    // first slot of each padded row of syn16 / syn_t / w_t
    const float* tile_cc;   // CLASS kinds: [G][2][2][kClassCache] class cache of each tile (k_tile_ccache)
    const uint32_t* tile_ccls;   // [G][2] its classes
    const uint8_t* tile_lean;    // [G] 1: every neuron an LIFB neuron of one of those classes
    uint32_t* ring_t;
    size_t ring_slab;       // words per tile in ring_t: RW * D * B rounded up to 16 bytes
    Ctl* ctl;
    uint32_t* win_buf;
    uint32_t win_stride;
    uint64_t* rec;
    uint64_t rec_cap;
    // kTStdp / kTStdpB.  Window k reads its neurons' last spike steps from last3 buffer
    // k % 3 and writes them to (k + 1) % 3, so buffers (k - 1) % 3 and k % 3
    // hold window k-1's start and end values for the whole launch.  Dendrite
    // updates are deferred: neuron j keeps its last kStdpHist spike steps
    // (hist, newest first: a spike moves the row down one entry), its in-synapses are up to date
    // through step tf[j], and cntf[j] is its fire count at tf[j].
    int32_t* last3;             // [3][last_stride]
    size_t last_stride;
    int32_t* hist;              // [n][kStdpHist] newest first, kHistNone past the spikes
    int32_t* tf;                // [n]
    int32_t* cntf;              // [n]
    // neurons to flush after window k: list k % 2 of (neuron, old tf) pairs,
    // fcnt[k % 2] entries; window k clears fcnt[(k + 1) % 2]
    int32_t* flist;             // [2][flist_cap][2]
    uint32_t* fcnt;             // [2]
    uint32_t flist_cap;
    const int64_t* din_ptr;     // [n+1] plastic in-synapses of each target
    const uint32_t* din_pos;    // their tile-order positions
    const int32_t* din_src;     // their sources
    const float* ltp;           // ltp[k] = A_LTP * expf(-k / TAU_LTP)
    const float* ltd;           // ltd[k] = A_LTD * expf(-k / TAU_LTD); ltd == ltp + lut_n
    int32_t lut_n;
    int32_t lut_lds;            // leading entries of each table copied to LDS (>= 1)
    float w_max, w_min;
    // Fused STDP read-back: the launch applies the pending window's updates to w_t
    // (a copy) and lists every neuron with deferred dendrite updates for the
    // flush kernel, touching no ring, neuron state, spike list or tf / cntf
    int32_t apply_only;
    // Fused STDP: 1 when k_stdp_flush follows this window's launch (the host runs it
    // after every second window of a run and after its last), else the window lists
    // no neuron for flushing
    int32_t flush_next;
    // ring words per (slot, target): 1 for CLASS counts and for the wide kinds when
    // every integrating neuron is LIFB (one fixed-point sum), else 2 (exc, inh)
    int32_t rw;
    // 1: workgroup b runs tile (b % kXcds) * (G / kXcds) + ... (a contiguous run of
    // tiles per XCD), set when every workgroup of the launch is resident at once
    int32_t xcd_map;
    // persistent launches of 16-bit CLASS words with G <= kRegionMaxG and G <= the block
    // size (and the counts fitting in LDS), launched as the kernel's list-region variant:
    // windows followed by another window of the same launch write their spikes into the
    // tile's own part of the list (tile_lo * window on) and publish the count with the
    // wait's flag, instead of reserving list space with a returning global atomic
    int32_t region;
    // persistent launches: one flag per tile (kBarStride words apart; a 64-bit word:
    // window epoch | spike count << 32) for the wait between windows, its error word,
    // and the wait's bound in wall-clock ticks
    uint32_t* bar;
    uint32_t* bar_err;
    uint64_t spin_limit;
};

enum RingMode { kRingClass = 0, kRingFixed = 1 };
enum WideKind { kWideF32 = 0, kWideBF16 = 1, kWideFP8 = 2 };

// sim kernels (mgb_kernels.hip)
hipError_t launch_neurons(int mode, const NeuArgs& a, int32_t off, int32_t wk, int32_t nsteps,
                          hipStream_t st);
hipError_t launch_prop(int mode, int wide_kind, const PropArgs& a, int32_t off, int32_t wk,
                       int grid, int block, hipStream_t st);
hipError_t launch_stdp(const StdpArgs& a, int32_t off, int32_t wk, int grid, hipStream_t st);
hipError_t launch_finalize(Ctl* ctl, int32_t steps, int32_t windows, hipStream_t st);
constexpr int kTiledBatch = 1024;   // spikes staged in LDS per propagation batch
constexpr int kTiledItems = 16384;  // work items staged in LDS per propagation round
constexpr int kTiledStaticLds = 1024; // bytes reserved for the fused kernel's static LDS
constexpr int kClassCache = 32;     // k-fold class sums cached in LDS (tables hold >= this many)
constexpr int kStdpHist = 16;       // spike steps kept per neuron for deferred dendrite updates
// A history entry past the neuron's spikes (memset byte 0x80): below every step and
// every lower bound of a history scan (>= -1).
constexpr int32_t kHistNone = (int32_t)0x80808080;
// Every (source, tile) sub-row of the tile-order copy starts at a multiple of
// kSubAlign words (16 B of CLASS16, 32 B of 32-bit words), so that a
// work item of 2-16 words is aligned vector loads
constexpr int kSubAlign = 8;
constexpr int kXcds = 8;            // MI355X: accelerator complex dies, one L2 each
constexpr int kBarStride = 32;      // words between the flags of a persistent launch's grid wait
constexpr int kRegionMaxG = 512;    // persistent launches: most tiles whose lists go to per-tile regions
constexpr int kRegionMinNeurons = 32768; // MGB_PERSIST_AUTO: list regions from this many neurons on
// Word 3 of a raw buffer resource on gfx9 (CDNA): data format 32, no swizzle; loads past
// the resource's byte count return 0.
constexpr int kBufferDword3 = 0x00020000;
// Dynamic LDS of the fused kernel (rw ring words per slot and target); kTStdp also
// caches lut_lds entries of each STDP table.
size_t tiled_lds_bytes(int D, int B, int kind, int rw, int lut_lds = 0);
// Class cache of every tile of the CLASS kinds (cc [G][2][2][kClassCache], cls [G][2]).
hipError_t launch_tile_ccache(const int32_t* tile_lo, int32_t G, const uint16_t* tbl, const uint32_t* tbl_off,
                              const float* seq, float* cc, uint32_t* cls, hipStream_t st);
// lean[g] = 1 when every neuron of tile g is an LIFB neuron of one of the tile's two
// cached classes (needs the neuron types, uploaded after the tiles are built).
hipError_t launch_tile_lean(const int32_t* tile_lo, int32_t G, const uint8_t* type, const uint16_t* tbl,
                            uint8_t* lean, hipStream_t st);
// Workgroups of `block` threads of the fused kernel for `a` that one CU holds at once.
int tiled_blocks_per_cu(const TiledArgs& a, int block, bool persistent = false);
hipError_t launch_tiled(const TiledArgs& a, int32_t off, int32_t wk, int32_t nsteps, int block,
                        hipStream_t st);
// Dynamic LDS of the persistent form for `a` (with a.region, the list regions' counts),
// and the persistent function's limit raised to it.
size_t tiled_pw_lds_bytes(const TiledArgs& a);
hipError_t set_tiled_pw_lds_limit(const TiledArgs& a);
// Persistent form (not kTStdp): nwin windows of nsteps steps in one launch; every
// workgroup must be resident (tiled_blocks_per_cu(a, block, true) x CUs >= G).  The
// wait after window k publishes and waits for epoch k + 1 (the window's index over
// the simulator's life), so flags need no reset between launches, only when the
// window index goes back (a restored checkpoint).
hipError_t launch_tiled_persistent(const TiledArgs& a, int32_t off, int32_t wk, int32_t nsteps, int32_t nwin,
                                   int block, hipStream_t st);
// err, h_err: the persistent waits' error word and its pinned host copy (or null)
hipError_t launch_finalize_tiled(Ctl* ctl, int32_t steps, int32_t windows, const uint32_t* err, uint32_t* h_err,
                                 hipStream_t st);
hipError_t set_tiled_lds_limit(int kind, size_t bytes);
// kTStdp: brings the plastic in-synapses of the neurons the fused kernel listed
// in window base_window + wk up to date (flush of deferred dendrite updates).
hipError_t launch_stdp_flush(const TiledArgs& a, int32_t off, int32_t wk, hipStream_t st);

}  // namespace mgb

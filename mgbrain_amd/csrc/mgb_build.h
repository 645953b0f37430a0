// mgb_build.h -- device-side network construction entry points (mgb_build.hip).
#pragma once
#include "mgb_internal.h"

namespace mgb {

constexpr uint32_t kWSentinel = 0xFFFFFFFFu;   // "no synapse of this class yet"
constexpr uint32_t kFlagBadId = 1u;
constexpr uint32_t kFlagDelayLow = 2u;         // quantised delay < 1 step
constexpr uint32_t kFlagDelayHigh = 4u;        // quantised delay > kWideMaxDelay
constexpr uint32_t kFlagMixed = 8u;            // a target sees two weights in one class
constexpr uint32_t kFlagCrossSeg = 16u;        // a synapse joins two candidate segments

struct StatArgs {
    uint64_t S;
    int32_t n;
    float step;
    const int32_t* src;
    const int32_t* tar;
    const float* weight;
    const float* delay;
    const uint8_t* plastic;
    const uint16_t* seg;    // [n] candidate segment of each neuron (NULL: not checked)
    uint8_t* dq;            // [S] quantised delay
    uint32_t* flags;
    int32_t* min_d;
    int32_t* max_d;
    uint32_t* wcls;         // [2][n] class weight bits per target
    uint32_t* indeg;        // [2][n] in-degree per target and class
    uint32_t* in_pl;        // [n] plastic in-degree
    uint32_t* absmax;       // max |w| (float bits)
    unsigned long long* n_plastic;
};

// Segment / tile tables of the tiled layout (device pointers; see TiledArgs).
struct TileTabs {
    int32_t nseg, B;
    const uint16_t* seg_of;     // [n] segment of each neuron (NULL: one segment)
    const int32_t* seg_lo;      // [nseg+1]
    const int32_t* seg_tile0;   // [nseg+1]
    const uint32_t* seg_kbase;  // [nseg+1] first sort key of each segment
    const uint64_t* seg_obase;  // [nseg+1] first offset entry of each segment
    const uint32_t* tile_of;    // [n] tile of each neuron
    const int32_t* tile_lo;     // [G+1] first neuron of each tile
};

struct GatherArgs {
    uint64_t S;
    const uint32_t* idx;    // sorted list indices
    const int32_t* src;
    const int32_t* tar;
    const float* weight;
    const uint8_t* dq;
    const uint8_t* plastic;
    int32_t* col;
    uint8_t* dly;
    float* w32;
    uint16_t* w16;
    uint8_t* w8;
    const float* row_scale;
    uint32_t* pos_of;       // [S] CSR slot of each list index (NULL if unused)
};

struct GenProjArgs {
    int32_t sn, tn, s0, t0, k;
    uint32_t index;         // projection index in its network (Philox counter word)
    uint64_t seed;
    uint64_t thr;           // Bernoulli threshold p * 2^32
    float wmin, wmax, dmin, dmax, dw, dt;
    int32_t plastic_flag;
    uint64_t base;          // list offset of the projection
    int32_t* src;
    int32_t* tar;
    float* weight;
    float* delay;
    uint8_t* plastic;
};

int blocks_for(uint64_t work, int block);
hipError_t run_stats(const StatArgs& a, hipStream_t st);
hipError_t make_sort_input(const int32_t* src, const int32_t* tar, const float* w, const uint8_t* dq,
                           bool class_payload, uint32_t* keys, uint32_t* vals, uint64_t S,
                           hipStream_t st);
hipError_t make_tile_keys(const int32_t* src, const int32_t* tar, const float* w, const uint8_t* dq,
                          const TileTabs& tt, uint32_t* keys, uint32_t* vals, uint64_t S, hipStream_t st);
hipError_t class16_words(const uint32_t* w32, uint64_t S, const TileTabs& tt, uint16_t* out, hipStream_t st);
hipError_t make_tile_keys_csr(const uint32_t* src, const int32_t* col, const TileTabs& tt, uint32_t* keys,
                              uint32_t* vals, uint64_t S, hipStream_t st);
hipError_t tile_gather_wide(const uint32_t* slot, uint64_t S, const int32_t* col, const uint8_t* dly,
                            const uint16_t* w16, const uint8_t* w8, const float* w32, const TileTabs& tt, int kind,
                            uint32_t* syn_t, float* w_t, hipStream_t st);
hipError_t invert_slots(const uint32_t* slot, uint64_t S, uint32_t* inv, hipStream_t st);
hipError_t remap_u32(uint32_t* a, const uint32_t* map, uint64_t np, hipStream_t st);
hipError_t scatter_weights(const uint32_t* slot, const float* w_t, uint64_t S, float* w32, hipStream_t st);
// kTStdpB: out[i] = bits of row_scale[i] | 1 << 31 when source i has a plastic
// synapse (src / plastic: the synapse list); pl[g] = 1 when a plastic synapse
// (in_ptr: their in-CSR) targets tile g.
hipError_t row_aux(const float* row_scale, int32_t n, const int32_t* src, const uint8_t* plastic, uint64_t S,
                   uint32_t* out, hipStream_t st);
hipError_t tile_plastic(const int32_t* tile_lo, int32_t G, const int64_t* in_ptr, uint8_t* pl, hipStream_t st);
// Weights of every CSR slot as mgb_read_csr decodes them: CLASS words through the
// targets' class tables; BF16 (w16) or FP8 (w8) codes times the row scales, plastic
// slots from w32.
hipError_t class_weights(const uint32_t* syn, const uint16_t* tbl, const float* tbl_w, uint64_t S, float* out,
                         hipStream_t st);
hipError_t wide_weights(const int64_t* row_ptr, int32_t n, const uint8_t* dly, const float* w32,
                        const uint16_t* w16, const uint8_t* w8, const float* row_scale, float* out,
                        hipStream_t st);
hipError_t plastic_isi(const int64_t* in_ptr, const int32_t* in_src, const int32_t* isi, int32_t n, uint64_t np,
                       int32_t* out, hipStream_t st);
hipError_t tile_offsets(const uint32_t* sorted_keys, const int64_t* row_ptr, const TileTabs& tt,
                        uint64_t total, uint16_t* offs16, uint32_t* offs32, hipStream_t st);
// pmap[k]: the padded position of unpadded tile-order position k.
hipError_t pad_map(const int64_t* row_ptr, const int64_t* prow, const uint32_t* eo, const uint32_t* en,
                   const TileTabs& tt, uint64_t total, uint32_t* pmap, hipStream_t st);
// out[idx[k]] = val[k], k < S.
hipError_t scatter_u32(const uint32_t* idx, const uint32_t* val, uint64_t S, uint32_t* out, hipStream_t st);
// Tile-order array `from` (unpadded offsets eo) into `to` (padded rows prow, offsets en).
template <class T>
hipError_t pad_copy(const T* from, const int64_t* row_ptr, const int64_t* prow, const uint32_t* eo, const uint32_t* en,
                    const TileTabs& tt, uint64_t total, T* to, hipStream_t st);
hipError_t row_scales(const int32_t* src, const float* w, uint64_t S, int32_t n, float top,
                      uint32_t* rmax_tmp, float* scale, hipStream_t st);
hipError_t make_in_keys(const int32_t* tar, const uint8_t* plastic, int32_t n, uint32_t* keys,
                        uint32_t* vals, uint64_t S, hipStream_t st);
hipError_t sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* kin, uint32_t* kout,
                      const uint32_t* vin, uint32_t* vout, uint64_t S, int bits, hipStream_t st);
hipError_t row_bounds(const uint32_t* sorted_keys, uint64_t S, int32_t n, int64_t* out,
                      hipStream_t st);
hipError_t gather_wide(const GatherArgs& a, hipStream_t st);
hipError_t in_fill(const uint32_t* sorted_idx, const uint32_t* pos_of, const int32_t* src,
                   uint64_t np, uint32_t* in_pos, int32_t* in_src, hipStream_t st);
hipError_t gen_fixk(const GenProjArgs& g, hipStream_t st);
hipError_t gen_bern(const GenProjArgs& g, uint32_t* row_cnt, const int64_t* row_off, bool fill,
                    hipStream_t st);

}  // namespace mgb

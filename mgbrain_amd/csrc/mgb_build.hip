// mgb_build.hip -- device-side network construction (gfx950).
//
// Replaces the host staging of MGSimulator::build_multinet (mgsim.cpp:8-144) and
// copy_subnet_gpu (mgnet.cpp:92-146) for npart = 1:
//   k_stats       validates ids, quantises delays round(delay/STEP) (mgsim.cpp:49),
//                 and gathers per-target class statistics;
//   radix sort    stable LSD sort of (source, payload) -- the reference's CSR keeps
//                 each source's synapses in synapse-list order (mgsim.cpp:56,
//                 114-131), which a stable sort by source reproduces exactly;
//   k_bounds      row_ptr by lower_bound on the sorted keys (axon_offs);
//   k_class_keys  CLASS-format payload (target | delay-1 | sign) sorted with the key;
//   k_gather_wide per-slot arrays of the wide formats;
//   k_gen_*       the new Bernoulli / fixed in-degree connectors generated on the
//                 device, bit-identical to the host generator (same Philox draws).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "mgb_build.h"
#include "mgb_rng.h"

namespace mgb {
namespace {

__device__ __forceinline__ bool wave_all_equal(uint32_t v) {
    const uint32_t first = __builtin_amdgcn_readfirstlane(v);
    return __all(v == first);
}

__global__ void k_stats(StatArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t total = a.S;
    // every lane of a wave runs the same number of iterations (wave-uniform loop)
    const uint64_t iters = (total + stride - 1) / stride;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t it = 0; it < iters; ++it, i += stride) {
        const bool in = i < total;
        uint32_t key = 0xFFFFFFFFu;    // (target, class) of this lane; ~0 = none
        uint32_t aseg = 0u, abits = 0u; // segment and |w| bits for the max |w| per segment
        if (in) {
            const int32_t s = a.src[i], t = a.tar[i];
            if ((uint32_t)s >= (uint32_t)a.n || (uint32_t)t >= (uint32_t)a.n) {
                atomicOr(a.flags, kFlagBadId);
            } else {
                if (a.seg && a.seg[s] != a.seg[t]) atomicOr(a.flags, kFlagCrossSeg);
                const float q = a.delay[i] / a.step;                 // mgsim.cpp:49
                const int d = (int)roundf(q);
                if (d < 1) atomicOr(a.flags, kFlagDelayLow);
                if (d > kWideMaxDelay) atomicOr(a.flags, kFlagDelayHigh);
                const int dc = d < 0 ? 0 : (d > 255 ? 255 : d);
                a.dq[i] = (uint8_t)dc;
                atomicMin(a.min_d, dc);
                atomicMax(a.max_d, dc);
                const float w = a.weight[i];
                const uint32_t cls = (w > 0.f) ? 0u : 1u;            // mgcore.cu:292-295
                const uint32_t wb = __float_as_uint(w);
                uint32_t* wc = a.wcls + (size_t)cls * (size_t)a.n + (size_t)t;
                const uint32_t cur = *wc;
                if (cur != wb) {
                    const uint32_t old = atomicCAS(wc, kWSentinel, wb);
                    if (old != kWSentinel && old != wb) atomicOr(a.flags, kFlagMixed);
                }
                aseg = a.seg ? a.seg[t] : 0u;
                abits = __float_as_uint(fabsf(w));
                if (a.plastic && a.plastic[i]) {
                    atomicAdd(a.n_plastic, 1ull);
                    atomicAdd(&a.in_pl[t], 1u);
                }
                key = ((uint32_t)t << 1) | cls;
            }
        }
        // max |w| per segment: one atomic per wave where its lanes share the segment
        if (wave_all_equal(aseg)) {
            uint32_t m = abits;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
            if ((threadIdx.x & 63u) == 0u && m) atomicMax(&a.absmax[aseg], m);
        } else if (abits) {
            atomicMax(&a.absmax[aseg], abits);
        }
        // in-degree per (target, class); target-major lists hit one address per wave
        if (wave_all_equal(key)) {
            if (key != 0xFFFFFFFFu && (threadIdx.x & 63u) == 0u) {
                const uint32_t c = (uint32_t)__popcll(__ballot(true));
                atomicAdd(&a.indeg[(size_t)(key & 1u) * (size_t)a.n + (key >> 1)], c);
            }
        } else if (key != 0xFFFFFFFFu) {
            atomicAdd(&a.indeg[(size_t)(key & 1u) * (size_t)a.n + (key >> 1)], 1u);
        }
    }
}

__global__ void k_iota_keys(const int32_t* src, uint32_t* keys, uint32_t* vals, uint64_t S) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S;
         i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = (uint32_t)src[i];
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_class_keys(const int32_t* src, const int32_t* tar, const float* w,
                             const uint8_t* dq, uint32_t* keys, uint32_t* vals, uint64_t S) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S;
         i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = (uint32_t)src[i];
        const uint32_t inh = (w[i] > 0.f) ? 0u : 1u;
        vals[i] = (uint32_t)tar[i] | ((uint32_t)(dq[i] - 1u) << 26) | (inh << 31);
    }
}

// Tile-sorted copy of the CLASS CSR.  Source s of segment q (which has Gq
// tiles of B targets from seg_lo[q]) gets keys seg_kbase[q] + (s - seg_lo[q]) *
// Gq + local tile of the target: sorted, they order each row by tile and keep
// the rows in source order.
__global__ void k_tile_keys(const int32_t* src, const int32_t* tar, const float* w, const uint8_t* dq,
                            TileTabs tt, uint32_t* keys, uint32_t* vals, uint64_t S) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t s = src[i], t = tar[i];
        const int32_t q = tt.seg_of ? (int32_t)tt.seg_of[s] : 0;
        const int32_t lo = tt.seg_lo[q];
        const uint32_t gq = (uint32_t)(tt.seg_tile0[q + 1] - tt.seg_tile0[q]);
        keys[i] = tt.seg_kbase[q] + (uint32_t)(s - lo) * gq + (tt.tile_of[t] - (uint32_t)tt.seg_tile0[q]);
        const uint32_t inh = (w[i] > 0.f) ? 0u : 1u;
        vals[i] = (uint32_t)t | ((uint32_t)(dq[i] - 1u) << 26) | (inh << 31);
    }
}

// 16-bit CLASS words (tiles of at most 1024 targets): tile-local target |
// delay-1 << 10 | inhibitory << 15, from the 32-bit words in tile order.
__global__ void k_class16(const uint32_t* w32, uint64_t S, TileTabs tt, uint16_t* out) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = w32[k];
        const int32_t t = (int32_t)(w & 0x3FFFFFFu);
        const uint32_t loc = (uint32_t)(t - tt.tile_lo[tt.tile_of[t]]);
        out[k] = (uint16_t)(loc | (((w >> 26) & 31u) << 10) | ((w >> 31) << 15));
    }
}

// Tile keys of the CSR slots of a wide-format network: slot k (source src[k],
// target col[k]); the value is the slot itself.
__global__ void k_tile_keys_csr(const uint32_t* src, const int32_t* col, TileTabs tt, uint32_t* keys,
                                uint32_t* vals, uint64_t S) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t s = (int32_t)src[k], t = col[k];
        const int32_t q = tt.seg_of ? (int32_t)tt.seg_of[s] : 0;
        const int32_t lo = tt.seg_lo[q];
        const uint32_t gq = (uint32_t)(tt.seg_tile0[q + 1] - tt.seg_tile0[q]);
        keys[k] = tt.seg_kbase[q] + (uint32_t)(s - lo) * gq + (tt.tile_of[t] - (uint32_t)tt.seg_tile0[q]);
        vals[k] = (uint32_t)k;
    }
}

// Fused-path words of the wide formats in tile order (slot[k]: the CSR slot at
// tile-order position k): local target | delay-1 << 11 | BF16 / FP8 bits << 16
// (kTStdp: the plastic flag << 16; kTStdpB: 0xFFFF for a plastic synapse, else its
// BF16 bits); FP32 weights go to w_t.
__global__ void k_tile_gather_wide(const uint32_t* slot, uint64_t S, const int32_t* col, const uint8_t* dly,
                                   const uint16_t* w16, const uint8_t* w8, const float* w32, TileTabs tt,
                                   int kind, uint32_t* syn_t, float* w_t) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = slot[k];
        const int32_t t = col[p];
        const uint32_t loc = (uint32_t)(t - tt.tile_lo[tt.tile_of[t]]);
        const uint32_t d = (uint32_t)(dly[p] & 127u);
        uint32_t payload = 0u;
        if (kind == kTBF16) payload = w16[p];
        else if (kind == kTFP8) payload = w8[p];
        else if (kind == kTStdp) payload = (uint32_t)(dly[p] >> 7);
        else if (kind == kTStdpB) payload = (dly[p] & 128u) ? 0xFFFFu : (uint32_t)w16[p];
        syn_t[k] = loc | ((d - 1u) << 11) | (payload << 16);
        if (kind == kTF32 || kind == kTStdp || kind == kTStdpB) w_t[k] = w32[p];
    }
}

// kTStdpB: out[i] = row scale bits of source i, then bit 31 set where the source's
// row holds a plastic synapse (list order: src, plastic).
__global__ void k_row_aux(const float* row_scale, int32_t n, uint32_t* out) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = __float_as_uint(row_scale[i]) & 0x7FFFFFFFu;
}

__global__ void k_row_aux_pl(const int32_t* src, const uint8_t* plastic, uint64_t S, uint32_t* out) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x)
        if (plastic[k]) atomicOr(&out[src[k]], 0x80000000u);
}

// kTStdpB: pl[g] = 1 when a plastic synapse targets a neuron of tile g.
__global__ void k_tile_pl(const int32_t* tile_lo, int32_t G, const int64_t* in_ptr, uint8_t* pl) {
    const int32_t g = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g < G) pl[g] = in_ptr[tile_lo[g + 1]] > in_ptr[tile_lo[g]] ? 1 : 0;
}

// inv[slot[k]] = k: the tile-order position of every CSR slot.
__global__ void k_invert(const uint32_t* slot, uint64_t S, uint32_t* inv) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x)
        inv[slot[k]] = (uint32_t)k;
}

// a[p] = map[a[p]].
__global__ void k_remap(uint32_t* a, const uint32_t* map, uint64_t np) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
         p += (uint64_t)gridDim.x * blockDim.x)
        a[p] = map[a[p]];
}

// w32[slot[k]] = w_t[k]: the CSR-order copy of tile-order weights (slot
// 0xFFFFFFFF: a padding position).
__global__ void k_scatter_w(const uint32_t* slot, const float* w_t, uint64_t S, float* w32) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = slot[k];
        if (p != 0xFFFFFFFFu) w32[p] = w_t[k];
    }
}

// CLASS format: out[k] = the class weight of synapse word k (target's table,
// sign bit), as the host read-back decodes it.
__global__ void k_class_weights(const uint32_t* syn, const uint16_t* tbl, const float* tbl_w, uint64_t S,
                                float* out) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = syn[k];
        out[k] = tbl_w[2u * (uint32_t)tbl[x & kIdMask] + (x >> 31)];
    }
}

// OCP FP8 E4M3 (bias 7, no infinities) to FP32, as the kernels decode it.
__device__ __forceinline__ float fp8_value(uint8_t b) {
    const uint32_t s = (uint32_t)(b >> 7) << 31;
    const uint32_t e = (b >> 3) & 15u, m = b & 7u;
    if (e == 0) return __uint_as_float(__float_as_uint((float)m * 0x1p-9f) | s);
    return __uint_as_float(s | ((e + 120u) << 23) | (m << 20));
}

// BF16 / FP8 formats: the weight of every CSR slot (plastic slots: their FP32
// weight; others: the stored code times the source row's scale).  One wave per
// source row.
__global__ void k_wide_weights(const int64_t* row_ptr, int32_t n, const uint8_t* dly, const float* w32,
                               const uint16_t* w16, const uint8_t* w8, const float* row_scale, float* out) {
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    const int32_t waves = (int32_t)(gridDim.x * (blockDim.x >> 6));
    for (int32_t i = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); i < n; i += waves) {
        const float rs = row_scale[i];
        for (int64_t k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += 64) {
            float v;
            if (dly[k] & 128u) v = w32[k];
            else if (w16) v = __uint_as_float((uint32_t)w16[k] << 16) * rs;
            else v = fp8_value(w8[k]) * rs;
            out[k] = v;
        }
    }
}

// out[idx[k]] = val[k].
__global__ void k_scatter_u32(const uint32_t* idx, const uint32_t* val, uint64_t S, uint32_t* out) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S;
         k += (uint64_t)gridDim.x * blockDim.x)
        out[idx[k]] = val[k];
}

// Smallest isi[] over the ends of plastic synapses (in-CSR in_ptr / in_src):
// targets with a plastic in-synapse, then the sources of all of them.
__global__ void k_plastic_isi(const int64_t* in_ptr, const int32_t* in_src, const int32_t* isi, int32_t n,
                              uint64_t np, int32_t* out) {
    int32_t m = 0x7FFFFFFF;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t j = t0; j < (uint64_t)n; j += stride)
        if (in_ptr[j + 1] > in_ptr[j]) m = min(m, isi[j]);
    for (uint64_t p = t0; p < np; p += stride) m = min(m, isi[in_src[p]]);
    for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63u) == 0u && m != 0x7FFFFFFF) atomicMin(out, m);
}

// Offset entry e = seg_obase[q] + (m - seg_lo[q]) * (Gq + 1) + t, t in [0, Gq]:
// first slot of (source m, local tile t) relative to row m (t = Gq: the row length).
template <class T>
__global__ void k_tile_offsets(const uint32_t* keys, const int64_t* row_ptr, TileTabs tt, uint64_t total,
                               T* offs) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = tt.nseg;          // seg_obase[lo] <= e < seg_obase[hi]
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (tt.seg_obase[mid] <= e) lo = mid;
            else hi = mid;
        }
        const int32_t q = lo;
        const uint64_t r = e - tt.seg_obase[q];
        const uint64_t gq = (uint64_t)(tt.seg_tile0[q + 1] - tt.seg_tile0[q]);
        const uint64_t rel = r / (gq + 1);
        const uint32_t m = (uint32_t)tt.seg_lo[q] + (uint32_t)rel;
        const uint64_t key = (uint64_t)tt.seg_kbase[q] + rel * gq + (r - rel * (gq + 1));
        // row m's keys lie in [row_ptr[m], row_ptr[m+1])
        uint64_t a = (uint64_t)row_ptr[m], b = (uint64_t)row_ptr[m + 1];
        const uint64_t a0 = a;
        while (a < b) {
            const uint64_t mid = (a + b) >> 1;
            if ((uint64_t)keys[mid] < key) a = mid + 1;
            else b = mid;
        }
        offs[e] = (T)(a - a0);
    }
}

// Copy of a tile-order array into the padded layout: offset entry e (source m,
// local tile t < Gq) moves its sub-row from row_ptr[m] + eo[e] (the unpadded
// offsets) to prow[m] + (en[e] & ~(kSubAlign - 1)).
template <class T>
__global__ void k_pad_copy(const T* from_arr, const int64_t* row_ptr, const int64_t* prow, const uint32_t* eo,
                           const uint32_t* en, TileTabs tt, uint64_t total, T* to_arr) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = tt.nseg;          // seg_obase[lo] <= e < seg_obase[hi]
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (tt.seg_obase[mid] <= e) lo = mid;
            else hi = mid;
        }
        const int32_t q = lo;
        const uint64_t r = e - tt.seg_obase[q];
        const uint64_t gq = (uint64_t)(tt.seg_tile0[q + 1] - tt.seg_tile0[q]);
        const uint64_t rel = r / (gq + 1);
        if (r - rel * (gq + 1) == gq) continue;   // the row's end entry
        const int32_t m = tt.seg_lo[q] + (int32_t)rel;
        const uint32_t len = eo[e + 1] - eo[e];
        const T* from = from_arr + row_ptr[m] + eo[e];
        T* to = to_arr + prow[m] + (en[e] & ~(uint32_t)(kSubAlign - 1));
        for (uint32_t j = 0; j < len; ++j) to[j] = from[j];
    }
}

__global__ void k_bounds(const uint32_t* keys, uint64_t S, int32_t n, int64_t* out) {
    const int32_t m = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (m > n) return;
    uint64_t lo = 0, hi = S;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < (uint32_t)m) lo = mid + 1;
        else hi = mid;
    }
    out[m] = (int64_t)lo;
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

// FP32 -> OCP FP8 E4M3 (round to nearest even, saturate to +-448).
__device__ __forceinline__ uint8_t f32_to_fp8e4m3(float f) {
    const uint32_t sgn = (__float_as_uint(f) >> 24) & 0x80u;
    float a = fabsf(f);
    if (!(a == a)) return (uint8_t)(sgn | 0x7Fu);
    if (a >= 448.f) return (uint8_t)(sgn | 0x7Eu);
    if (a < 0x1p-6f) {                                  // subnormal grid 2^-9
        const float q = rintf(a * 0x1p9f);
        const uint32_t m = (uint32_t)q;                 // 0..8 (8 -> min normal)
        return (uint8_t)(sgn | (m >= 8u ? 0x08u : m));
    }
    const uint32_t u = __float_as_uint(a);
    const uint32_t r = u + 0x7FFFFu + ((u >> 20) & 1u);   // RNE to 3 mantissa bits
    const uint32_t e = (r >> 23) - 120u, m = (r >> 20) & 7u;
    if (e > 15u || (e == 15u && m == 7u)) return (uint8_t)(sgn | 0x7Eu);
    return (uint8_t)(sgn | (e << 3) | m);
}

__global__ void k_gather_wide(GatherArgs a) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.S;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = a.idx[k];
        a.col[k] = a.tar[i];
        const bool pl = a.plastic && a.plastic[i];
        a.dly[k] = (uint8_t)(a.dq[i] | (pl ? 0x80u : 0u));
        const float w = a.weight[i];
        a.w32[k] = w;
        if (a.w16) a.w16[k] = f32_to_bf16_rne(w / a.row_scale[a.src[i]]);
        if (a.w8) a.w8[k] = f32_to_fp8e4m3(w / a.row_scale[a.src[i]]);
        if (a.pos_of) a.pos_of[i] = (uint32_t)k;
    }
}

__global__ void k_row_absmax(const int32_t* src, const float* w, uint64_t S, uint32_t* rmax) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicMax(&rmax[src[i]], __float_as_uint(fabsf(w[i])));
}

// Scale of each source row: its largest |w| maps onto `top` (E4M3's largest normal
// 448 for FP8, 1 for BF16), so a row of equal weights is stored exactly.
__global__ void k_row_scale(const uint32_t* rmax, int32_t n, float top, float* scale) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float m = __uint_as_float(rmax[i]);
    scale[i] = m > 0.f ? m / top : 1.f;
}

__global__ void k_in_keys(const int32_t* tar, const uint8_t* plastic, int32_t n, uint32_t* keys,
                          uint32_t* vals, uint64_t S) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S;
         i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = plastic[i] ? (uint32_t)tar[i] : (uint32_t)n;
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_in_fill(const uint32_t* sorted_idx, const uint32_t* pos_of, const int32_t* src,
                          uint64_t np, uint32_t* in_pos, int32_t* in_src) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = sorted_idx[p];
        in_pos[p] = pos_of[i];
        in_src[p] = src[i];
    }
}

// Fixed in-degree: list index base + n*K + k, one Philox call per 4 synapses.
__global__ void k_gen_fixk(GenProjArgs g) {
    const uint64_t total = (uint64_t)g.tn * (uint64_t)g.k;
    const uint64_t nq = (total + 3) >> 2;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
         q += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k0 = (uint32_t)g.seed, k1 = (uint32_t)(g.seed >> 32);
        const U32x4 o = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), g.index, kTagFixK, k0, k1);
        U32x4 ow = {{0u, 0u, 0u, 0u}}, od = {{0u, 0u, 0u, 0u}};
        if (g.wmin != g.wmax) ow = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), g.index, kTagFixKW, k0, k1);
        if (g.dmin != g.dmax) od = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), g.index, kTagFixKD, k0, k1);
        for (uint32_t j = 0; j < 4; ++j) {
            const uint64_t f = 4 * q + j;
            if (f >= total) break;
            const uint64_t n = f / (uint64_t)g.k;
            const uint32_t m = (uint32_t)(((uint64_t)pick4(o, j) * (uint64_t)g.sn) >> 32);
            const uint64_t at = g.base + f;
            g.src[at] = g.s0 + (int32_t)m;
            g.tar[at] = g.t0 + (int32_t)n;
            g.weight[at] = grid_weight(pick4(ow, j), g.wmin, g.wmax, g.dw);
            g.delay[at] = grid_delay(pick4(od, j), g.dmin, g.dmax, g.dt);
            g.plastic[at] = (uint8_t)g.plastic_flag;
        }
    }
}

// Bernoulli: one wave per source row, lanes over targets (row-major list order).
template <bool kFill>
__global__ void k_gen_bern(GenProjArgs g, uint32_t* row_cnt, const int64_t* row_off) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t k0 = (uint32_t)g.seed, k1 = (uint32_t)(g.seed >> 32);
    for (uint32_t m = wave; m < (uint32_t)g.sn; m += nwaves) {
        uint64_t at = kFill ? (uint64_t)row_off[m] : 0ull;
        uint32_t count = 0;
        for (uint32_t nb = 0; nb < (uint32_t)g.tn; nb += 64) {
            const uint32_t n = nb + lane;
            const uint64_t f = (uint64_t)m * (uint64_t)g.tn + n;
            bool hit = false;
            if (n < (uint32_t)g.tn) hit = (uint64_t)philox_pick(g.seed, f, g.index, kTagBern) < g.thr;
            const uint64_t mask = __ballot(hit);
            if (kFill && hit) {
                const uint64_t p = g.base + at + (uint64_t)__builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                g.src[p] = g.s0 + (int32_t)m;
                g.tar[p] = g.t0 + (int32_t)n;
                g.weight[p] = grid_weight(philox_pick(g.seed, f, g.index, kTagBernW), g.wmin, g.wmax, g.dw);
                g.delay[p] = grid_delay(philox_pick(g.seed, f, g.index, kTagBernD), g.dmin, g.dmax, g.dt);
                g.plastic[p] = (uint8_t)g.plastic_flag;
            }
            at += (uint64_t)__popcll(mask);
            count += (uint32_t)__popcll(mask);
        }
        if (!kFill && lane == 0) row_cnt[m] = count;
    }
    (void)k0; (void)k1;
}

}  // namespace

int blocks_for(uint64_t work, int block) {
    uint64_t b = (work + (uint64_t)block - 1) / (uint64_t)block;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

hipError_t run_stats(const StatArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_stats, dim3(blocks_for(a.S, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t make_sort_input(const int32_t* src, const int32_t* tar, const float* w, const uint8_t* dq,
                           bool class_payload, uint32_t* keys, uint32_t* vals, uint64_t S,
                           hipStream_t st) {
    if (S == 0) return hipSuccess;
    if (class_payload)
        hipLaunchKernelGGL(k_class_keys, dim3(blocks_for(S, 256)), dim3(256), 0, st, src, tar, w, dq,
                           keys, vals, S);
    else
        hipLaunchKernelGGL(k_iota_keys, dim3(blocks_for(S, 256)), dim3(256), 0, st, src, keys, vals, S);
    return hipGetLastError();
}

hipError_t make_tile_keys(const int32_t* src, const int32_t* tar, const float* w, const uint8_t* dq,
                          const TileTabs& tt, uint32_t* keys, uint32_t* vals, uint64_t S, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tile_keys, dim3(blocks_for(S, 256)), dim3(256), 0, st, src, tar, w, dq, tt, keys,
                       vals, S);
    return hipGetLastError();
}

hipError_t class16_words(const uint32_t* w32, uint64_t S, const TileTabs& tt, uint16_t* out, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_class16, dim3(blocks_for(S, 256)), dim3(256), 0, st, w32, S, tt, out);
    return hipGetLastError();
}

hipError_t make_tile_keys_csr(const uint32_t* src, const int32_t* col, const TileTabs& tt, uint32_t* keys,
                              uint32_t* vals, uint64_t S, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tile_keys_csr, dim3(blocks_for(S, 256)), dim3(256), 0, st, src, col, tt, keys, vals, S);
    return hipGetLastError();
}

hipError_t tile_gather_wide(const uint32_t* slot, uint64_t S, const int32_t* col, const uint8_t* dly,
                            const uint16_t* w16, const uint8_t* w8, const float* w32, const TileTabs& tt, int kind,
                            uint32_t* syn_t, float* w_t, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tile_gather_wide, dim3(blocks_for(S, 256)), dim3(256), 0, st, slot, S, col, dly, w16, w8,
                       w32, tt, kind, syn_t, w_t);
    return hipGetLastError();
}

hipError_t invert_slots(const uint32_t* slot, uint64_t S, uint32_t* inv, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_invert, dim3(blocks_for(S, 256)), dim3(256), 0, st, slot, S, inv);
    return hipGetLastError();
}

hipError_t remap_u32(uint32_t* a, const uint32_t* map, uint64_t np, hipStream_t st) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_remap, dim3(blocks_for(np, 256)), dim3(256), 0, st, a, map, np);
    return hipGetLastError();
}

hipError_t row_aux(const float* row_scale, int32_t n, const int32_t* src, const uint8_t* plastic, uint64_t S,
                   uint32_t* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_row_aux, dim3((n + 255) / 256), dim3(256), 0, st, row_scale, n, out);
    if (S) hipLaunchKernelGGL(k_row_aux_pl, dim3(blocks_for(S, 256)), dim3(256), 0, st, src, plastic, S, out);
    return hipGetLastError();
}

hipError_t tile_plastic(const int32_t* tile_lo, int32_t G, const int64_t* in_ptr, uint8_t* pl, hipStream_t st) {
    if (G <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tile_pl, dim3((G + 255) / 256), dim3(256), 0, st, tile_lo, G, in_ptr, pl);
    return hipGetLastError();
}

hipError_t scatter_weights(const uint32_t* slot, const float* w_t, uint64_t S, float* w32, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_w, dim3(blocks_for(S, 256)), dim3(256), 0, st, slot, w_t, S, w32);
    return hipGetLastError();
}

hipError_t class_weights(const uint32_t* syn, const uint16_t* tbl, const float* tbl_w, uint64_t S, float* out,
                         hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_class_weights, dim3(blocks_for(S, 256)), dim3(256), 0, st, syn, tbl, tbl_w, S, out);
    return hipGetLastError();
}

hipError_t wide_weights(const int64_t* row_ptr, int32_t n, const uint8_t* dly, const float* w32,
                        const uint16_t* w16, const uint8_t* w8, const float* row_scale, float* out,
                        hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_wide_weights, dim3(blocks_for((uint64_t)n * 64u, 256)), dim3(256), 0, st, row_ptr, n,
                       dly, w32, w16, w8, row_scale, out);
    return hipGetLastError();
}

hipError_t plastic_isi(const int64_t* in_ptr, const int32_t* in_src, const int32_t* isi, int32_t n, uint64_t np,
                       int32_t* out, hipStream_t st) {
    const uint64_t work = std::max<uint64_t>((uint64_t)n, np);
    hipLaunchKernelGGL(k_plastic_isi, dim3(blocks_for(std::max<uint64_t>(work, 1), 256)), dim3(256), 0, st, in_ptr,
                       in_src, isi, n, np, out);
    return hipGetLastError();
}

hipError_t tile_offsets(const uint32_t* sorted_keys, const int64_t* row_ptr, const TileTabs& tt,
                        uint64_t total, uint16_t* offs16, uint32_t* offs32, hipStream_t st) {
    if (total == 0) return hipSuccess;
    if (offs16)
        hipLaunchKernelGGL(k_tile_offsets<uint16_t>, dim3(blocks_for(total, 256)), dim3(256), 0, st,
                           sorted_keys, row_ptr, tt, total, offs16);
    else
        hipLaunchKernelGGL(k_tile_offsets<uint32_t>, dim3(blocks_for(total, 256)), dim3(256), 0, st,
                           sorted_keys, row_ptr, tt, total, offs32);
    return hipGetLastError();
}

// pmap[row_ptr[m] + eo[e] + j] = prow[m] + (en[e] & ~(kSubAlign - 1)) + j: the
// padded position of every unpadded tile-order position.
__global__ void k_pad_map(const int64_t* row_ptr, const int64_t* prow, const uint32_t* eo, const uint32_t* en,
                          TileTabs tt, uint64_t total, uint32_t* pmap) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = tt.nseg;
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (tt.seg_obase[mid] <= e) lo = mid;
            else hi = mid;
        }
        const int32_t q = lo;
        const uint64_t r = e - tt.seg_obase[q];
        const uint64_t gq = (uint64_t)(tt.seg_tile0[q + 1] - tt.seg_tile0[q]);
        const uint64_t rel = r / (gq + 1);
        if (r - rel * (gq + 1) == gq) continue;
        const int32_t m = tt.seg_lo[q] + (int32_t)rel;
        const uint32_t len = eo[e + 1] - eo[e];
        const uint64_t from = (uint64_t)row_ptr[m] + eo[e];
        const uint64_t to = (uint64_t)prow[m] + (en[e] & ~(uint32_t)(kSubAlign - 1));
        for (uint32_t j = 0; j < len; ++j) pmap[from + j] = (uint32_t)(to + j);
    }
}

hipError_t pad_map(const int64_t* row_ptr, const int64_t* prow, const uint32_t* eo, const uint32_t* en,
                   const TileTabs& tt, uint64_t total, uint32_t* pmap, hipStream_t st) {
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad_map, dim3(blocks_for(total, 256)), dim3(256), 0, st, row_ptr, prow, eo, en, tt, total,
                       pmap);
    return hipGetLastError();
}

hipError_t scatter_u32(const uint32_t* idx, const uint32_t* val, uint64_t S, uint32_t* out, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_u32, dim3(blocks_for(S, 256)), dim3(256), 0, st, idx, val, S, out);
    return hipGetLastError();
}

template <class T>
hipError_t pad_copy(const T* from, const int64_t* row_ptr, const int64_t* prow, const uint32_t* eo, const uint32_t* en,
                    const TileTabs& tt, uint64_t total, T* to, hipStream_t st) {
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad_copy<T>, dim3(blocks_for(total, 256)), dim3(256), 0, st, from, row_ptr, prow, eo, en,
                       tt, total, to);
    return hipGetLastError();
}
template hipError_t pad_copy<uint16_t>(const uint16_t*, const int64_t*, const int64_t*, const uint32_t*,
                                       const uint32_t*, const TileTabs&, uint64_t, uint16_t*, hipStream_t);
template hipError_t pad_copy<uint32_t>(const uint32_t*, const int64_t*, const int64_t*, const uint32_t*,
                                       const uint32_t*, const TileTabs&, uint64_t, uint32_t*, hipStream_t);

hipError_t row_scales(const int32_t* src, const float* w, uint64_t S, int32_t n, float top,
                      uint32_t* rmax_tmp, float* scale, hipStream_t st) {
    hipError_t e = hipMemsetAsync(rmax_tmp, 0, 4 * (size_t)n, st);
    if (e != hipSuccess) return e;
    if (S) hipLaunchKernelGGL(k_row_absmax, dim3(blocks_for(S, 256)), dim3(256), 0, st, src, w, S, rmax_tmp);
    hipLaunchKernelGGL(k_row_scale, dim3((n + 255) / 256), dim3(256), 0, st, rmax_tmp, n, top, scale);
    return hipGetLastError();
}

hipError_t make_in_keys(const int32_t* tar, const uint8_t* plastic, int32_t n, uint32_t* keys,
                        uint32_t* vals, uint64_t S, hipStream_t st) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_in_keys, dim3(blocks_for(S, 256)), dim3(256), 0, st, tar, plastic, n, keys,
                       vals, S);
    return hipGetLastError();
}

hipError_t sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* kin, uint32_t* kout,
                      const uint32_t* vin, uint32_t* vout, uint64_t S, int bits, hipStream_t st) {
    size_t sz = temp ? *temp_bytes : 0;
    hipError_t e = rocprim::radix_sort_pairs(temp, sz, kin, kout, vin, vout, (size_t)S, 0u,
                                             (unsigned)bits, st);
    if (!temp) *temp_bytes = sz;
    return e;
}

hipError_t row_bounds(const uint32_t* sorted_keys, uint64_t S, int32_t n, int64_t* out,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_bounds, dim3((n + 1 + 255) / 256), dim3(256), 0, st, sorted_keys, S, n, out);
    return hipGetLastError();
}

hipError_t gather_wide(const GatherArgs& a, hipStream_t st) {
    if (a.S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_wide, dim3(blocks_for(a.S, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t in_fill(const uint32_t* sorted_idx, const uint32_t* pos_of, const int32_t* src,
                   uint64_t np, uint32_t* in_pos, int32_t* in_src, hipStream_t st) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_in_fill, dim3(blocks_for(np, 256)), dim3(256), 0, st, sorted_idx, pos_of,
                       src, np, in_pos, in_src);
    return hipGetLastError();
}

hipError_t gen_fixk(const GenProjArgs& g, hipStream_t st) {
    const uint64_t nq = ((uint64_t)g.tn * (uint64_t)g.k + 3) >> 2;
    if (nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_fixk, dim3(blocks_for(nq, 256)), dim3(256), 0, st, g);
    return hipGetLastError();
}

hipError_t gen_bern(const GenProjArgs& g, uint32_t* row_cnt, const int64_t* row_off, bool fill,
                    hipStream_t st) {
    if (g.sn == 0) return hipSuccess;
    const int grid = blocks_for((uint64_t)g.sn * 64ull, 256);
    if (fill)
        hipLaunchKernelGGL(k_gen_bern<true>, dim3(grid), dim3(256), 0, st, g, row_cnt, row_off);
    else
        hipLaunchKernelGGL(k_gen_bern<false>, dim3(grid), dim3(256), 0, st, g, row_cnt, row_off);
    return hipGetLastError();
}

}  // namespace mgb

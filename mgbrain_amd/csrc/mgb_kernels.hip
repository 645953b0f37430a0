// mgb_kernels.hip -- per-step kernels of the MI355X simulation step (gfx950).
//
// One "window" is w <= min_delay consecutive steps.  A spike emitted inside the
// window reaches its target at least min_delay steps later, so all neuron
// updates of a window depend only on spikes of earlier windows:
//   k_neurons   advances every neuron through the w steps with its state in
//               registers (one HBM read/write of the state per window), reads
//               and clears one ring row per step, and appends spikes to the
//               window's list with one wave-ballot + one atomic per wave-step;
//   k_prop      walks the CSR rows of the window's spikes and adds each
//               synapse into the ring row of its arrival step with an integer
//               atomic (order-independent, so results are bit-reproducible);
//   k_stdp      (STDP runs use w = 1) pair-based weight update of the spikers'
//               out- and in-synapses.
// Reference: mgsim_neus_core (mgcore.cu:116-234), mgsim_syn_fast_core
// (:271-305), mgsim_stdp_core/stdp_cal_weight (:346-412), ordering (:535-554).
//
// Floating-point expressions follow the reference's source with the
// contraction nvcc applies by default written as explicit fmaf(); the library
// is built with -ffp-contract=off so nothing else is fused.
#include "mgb_internal.h"
#include "mgb_rng.h"

namespace mgb {
namespace {

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}

// OCP FP8 E4M3 (bias 7, no infinities) to FP32.
__device__ __forceinline__ float fp8e4m3_to_f32(uint8_t b) {
    const uint32_t s = (uint32_t)(b >> 7) << 31;
    const uint32_t e = (b >> 3) & 15u, m = b & 7u;
    if (e == 0) return __uint_as_float(__float_as_uint((float)m * 0x1p-9f) | s);
    return __uint_as_float(s | ((e + 120u) << 23) | (m << 20));
}

enum : int { T_RAW = 0, T_LIF0 = 1, T_POISSON = 2, T_LIF = 3, T_LIFE = 4, T_LIFB = 5 };

// Registers of one neuron across the steps of a window.
struct NState {
    float V, Ie, Ii, rate;
    int32_t refr, last, cnt, refr_steps;
    float c3, c4, c12, c13, c14, c15, c16, c17;
    uint32_t off_e, off_i;       // CLASS seq-sum table offsets
    uint32_t cls;                // CLASS weight class
    float we, wi;                // CLASS weights
    uint32_t key0, key1, lid;    // Poisson stream
    U32x4 rnd;
};

__device__ __forceinline__ bool integrates(int type) {
    return type == T_LIF || type == T_LIFB || type == T_LIFE;
}

// Loads the neuron's registers (state, constants, stream); `last` holds the
// last spike steps.
template <class A>
__device__ __forceinline__ void neuron_load(const A& a, int i, int type, bool class_tables, NState& st,
                                            const int32_t* last) {
    st.V = st.Ie = st.Ii = st.rate = 0.f;
    st.refr = st.refr_steps = 0;
    st.c3 = st.c4 = st.c12 = st.c13 = st.c14 = st.c15 = st.c16 = st.c17 = 0.f;
    st.off_e = st.off_i = st.key0 = st.key1 = st.lid = st.cls = 0u;
    st.we = st.wi = 0.f;
    st.rnd = U32x4{{0u, 0u, 0u, 0u}};
    st.last = last[i];
    st.cnt = a.cnt[i];
    // the per-neuron words that index tables, and V / refr, are loaded before the
    // branch on the type (one round trip with the type's), the table entries after it
    const uint32_t cs = a.cset[i];
    const uint32_t tb = class_tables ? (uint32_t)a.tbl[i] : 0u;
    const uint32_t sidx = a.stream ? a.stream[i] : 0u;
    const float v0 = a.V[i];
    const int32_t r0 = a.refr[i];
    if (type == T_POISSON) {
        st.rate = a.rate[i];
        const uint64_t seed = a.stream_seed[sidx];
        st.key0 = (uint32_t)seed;
        st.key1 = (uint32_t)(seed >> 32);
        st.lid = (uint32_t)(i - a.stream_base[sidx]);
    } else if (integrates(type)) {
        const CSet& C = a.csets[cs];
        st.c3 = C.c[3];
        st.c4 = C.c[4];
        st.refr_steps = __float_as_int(C.c[30]);
        if (type == T_LIF) {
            st.c12 = C.c[12]; st.c13 = C.c[13]; st.c14 = C.c[14];
            st.c15 = C.c[15]; st.c16 = C.c[16]; st.c17 = C.c[17];
        }
        st.V = v0;
        st.refr = r0;
        if (type != T_LIFB) {
            st.Ie = a.Ie[i];
            st.Ii = a.Ii[i];
        }
        if (class_tables) {
            const uint32_t t = tb;
            st.cls = t;
            st.off_e = a.tbl_off[2 * t];
            st.off_i = a.tbl_off[2 * t + 1];
            st.we = a.tbl_w[2 * t];
            st.wi = a.tbl_w[2 * t + 1];
        }
    }
}

template <class A>
__device__ __forceinline__ void neuron_store(const A& a, int i, int type, const NState& st, int32_t* last) {
    last[i] = st.last;
    a.cnt[i] = st.cnt;
    if (integrates(type)) {
        a.V[i] = st.V;
        a.refr[i] = st.refr;
        if (type != T_LIFB) {
            a.Ie[i] = st.Ie;
            a.Ii[i] = st.Ii;
        }
    }
}

// Poisson draw of step s (w = step index inside the window), mgcore.cu:121-131.
__device__ __forceinline__ bool poisson_step(NState& st, int32_t s, int32_t w) {
    if (w == 0 || (s & 3) == 0)
        st.rnd = philox4x32_10(st.lid, (uint32_t)s >> 2, 0u, kTagPoisson, st.key0, st.key1);
    const float u = u01_open0(pick4(st.rnd, (uint32_t)s & 3u));
    const bool fired = st.rate > u;
    if (fired) {
        ++st.cnt;
        st.last = s;
    }
    return fired;
}

// One step of an integrating neuron given this step's ring input (be, bi).
// Inputs of a step from the fixed-point ring sums (scale 1 / inv on both sides):
// LIFB integrates be + bi only, so it gets their integer sum converted once.
__device__ __forceinline__ void fixed_input(int type, int32_t qe, int32_t qi, float inv, float& be, float& bi) {
    if (type == T_LIFB) {
        be = (float)(qe + qi) * inv;
        bi = 0.0f;
    } else {
        be = (float)qe * inv;
        bi = (float)qi * inv;
    }
}

// LIFB step (mgcore.cu:164-196) given this step's input sum be + bi.
__device__ __forceinline__ bool integrate_lifb(NState& st, int32_t s, float input) {
    if (st.refr > 0) {
        --st.refr;
        return false;
    }
    if (st.V >= st.c4) {
        ++st.cnt; st.V = st.c3; st.last = s; st.refr = st.refr_steps;
        return true;
    }
    st.V = st.V + fmaf(-0.005f, st.V, input);
    return false;
}

__device__ __forceinline__ bool integrate_step(int type, NState& st, int32_t s, float be, float bi) {
    bool fired = false;
    if (type == T_LIFB) {
        fired = integrate_lifb(st, s, be + bi);
    } else if (type == T_LIFE) {                                 // :197-233
        st.Ie = fmaf(st.Ie, 0.98f, be);
        st.Ii = fmaf(st.Ii, 0.99f, -bi);
        if (st.refr > 0) {
            --st.refr;
        } else if (st.V >= st.c4) {
            fired = true; ++st.cnt; st.V = st.c3; st.last = s; st.refr = st.refr_steps;
        } else {
            const float V = st.V;
            const float t0 = -0.06f - V;
            const float t1 = fmaf(st.Ie, 0.0f - V, t0);
            const float t2 = fmaf(st.Ii, -0.08f - V, t1);
            const float t3 = t2 + 0.02f;
            st.V = fmaf(t3, 0.005f, V);
        }
    } else {                                                     // LIF :132-163
        if (st.refr > 0) {
            --st.refr;
        } else {
            st.V = fmaf(st.Ii, st.c13, fmaf(st.Ie, st.c12, fmaf(st.c17, st.V, st.c14)));
            st.Ie *= st.c15;
            st.Ii *= st.c16;
            if (st.V >= st.c4) {
                fired = true; ++st.cnt; st.V = st.c3; st.last = s; st.refr = st.refr_steps;
            } else {
                st.Ie += be;
                st.Ii += bi;
            }
        }
    }
    return fired;
}

// k-fold sequential FP32 sum of w: what k float atomicAdds of w into a zeroed
// buffer leave (mgcore.cu:112-115).  Short sums are added in registers, long
// ones read from the host-built table.
__device__ __forceinline__ float class_sum(uint32_t k, float w, const float* seq, uint32_t off) {
    if (k > 24u) return seq[off + k];
    float acc = 0.f;
    for (uint32_t j = 0; j < k; ++j) acc = acc + w;
    return acc;
}

// CLASS ring word -> (excitatory, inhibitory) input.
__device__ __forceinline__ void class_input(uint32_t x, const float* seq, const NState& st, float& be,
                                            float& bi) {
    be = class_sum(x & 0xFFFFu, st.we, seq, st.off_e);
    bi = class_sum(x >> 16, st.wi, seq, st.off_i);
}

template <int MODE>
__global__ __launch_bounds__(kNeuBlock) void k_neurons(NeuArgs a, int32_t off, int32_t wk,
                                                      int32_t nsteps) {
    Ctl* ctl = a.ctl;
    const int64_t s0 = ctl->base_step + off;
    const int par = (int)((ctl->base_window + wk) & 1);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // fold the previous window's spike count into the record counter
        ctl->rec_cnt += ctl->win_cnt[par ^ 1];
        ctl->win_cnt[par ^ 1] = 0u;
    }
    const int i = (int)(blockIdx.x * kNeuBlock + threadIdx.x);
    const bool active = i < a.n;
    const int type = active ? (int)a.type[i] : T_RAW;
    uint32_t* wbuf = a.win_buf + (size_t)par * a.win_stride;
    NState st;
    if (active) neuron_load(a, i, type, MODE == kRingClass, st, a.last);
    const bool integ = integrates(type);
    int32_t slot = (int32_t)(s0 % a.D);
    for (int32_t w = 0; w < nsteps; ++w) {
        const int32_t s = (int32_t)(s0 + w);
        bool fired = false;
        if (type == T_POISSON) {
            fired = poisson_step(st, s, w);
        } else if (integ) {
            const size_t loc = (size_t)slot * (size_t)a.npad + (size_t)i;
            float be, bi;
            if (MODE == kRingClass) {
                const uint32_t x = a.ring[loc];
                if (x != 0u) a.ring[loc] = 0u;                       // clearInput :99-103
                class_input(x, a.seq, st, be, bi);
            } else {
                const int32_t qe = a.ring_e[loc], qi = a.ring_i[loc];
                if (qe != 0) a.ring_e[loc] = 0;
                if (qi != 0) a.ring_i[loc] = 0;
                const float inv = a.seg_scale ? a.seg_scale[a.seg_of[i]].y : a.inv_e;   // inv_e == inv_i
                fixed_input(type, qe, qi, inv, be, bi);
            }
            fired = integrate_step(type, st, s, be, bi);
        }
        // wave-ballot compaction of this step's spikes
        const uint64_t m = __ballot(fired);
        if (m != 0ull) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            uint32_t base = 0u;
            if (lane_id() == leader) base = atomicAdd(&ctl->win_cnt[par], (uint32_t)__popcll(m));
            base = __shfl(base, (int)leader);
            if (fired) wbuf[base + rank_below(m)] = (uint32_t)i | ((uint32_t)w << 26);
        }
        slot = (slot + 1 == a.D) ? 0 : slot + 1;
    }
    if (active) neuron_store(a, i, type, st, a.last);
}

// Propagation: one wave per spike, lanes stride the spiker's CSR row.
template <int MODE, int WK>
__global__ void k_prop(PropArgs a, int32_t off, int32_t wk) {
    __shared__ unsigned long long se_blk;
    Ctl* ctl = a.ctl;
    const int64_t s0 = ctl->base_step + off;
    const int par = (int)((ctl->base_window + wk) & 1);
    const uint32_t nspk = ctl->win_cnt[par];
    const uint64_t rec_base = ctl->rec_cnt;
    const uint32_t* wbuf = a.win_buf + (size_t)par * a.win_stride;
    if (threadIdx.x == 0) se_blk = 0ull;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const int32_t D = a.D;
    // each spike's row is shared by `split` waves: part f takes lanes' chunks
    // beg + f*64 + lane, stepping split*64
    const uint32_t F = (uint32_t)a.split;
    const int64_t stride = 64 * (int64_t)F;
    unsigned long long se_local = 0ull;
    for (uint32_t vw = wave; vw < nspk * F; vw += nwaves) {
        const uint32_t e = vw / F, f = vw - e * F;
        const uint32_t ent = wbuf[e];
        const int32_t i = (int32_t)(ent & kIdMask);
        const int32_t s = (int32_t)(s0 + (ent >> 26));
        const int64_t beg = a.row_ptr[i] + 64 * (int64_t)f, end = a.row_ptr[i + 1];
        if (lane == 0 && f == 0) {
            se_local += (unsigned long long)(end - beg);   // f == 0: the whole row
            if (a.rec) {
                const uint64_t r = rec_base + e;
                if (r < a.rec_cap) a.rec[r] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)i;
                else ctl->rec_overflow = 1;
            }
        }
        const int32_t sD = s % D;
        if (MODE == kRingClass) {
            int64_t k = beg + lane;
            for (; k + 3 * stride < end; k += 4 * stride) {
                uint32_t x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(&a.syn[k + stride * u]);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int32_t slot = sD + (int32_t)((x[u] >> 26) & 31u) + 1;
                    slot -= (slot >= D) ? D : 0;
                    atomicAdd(&a.ring[(size_t)slot * (size_t)a.npad + (x[u] & kIdMask)],
                              (x[u] >> 31) ? 0x10000u : 1u);
                }
            }
            for (; k < end; k += stride) {
                const uint32_t x = __builtin_nontemporal_load(&a.syn[k]);
                int32_t slot = sD + (int32_t)((x >> 26) & 31u) + 1;
                slot -= (slot >= D) ? D : 0;
                atomicAdd(&a.ring[(size_t)slot * (size_t)a.npad + (x & kIdMask)],
                          (x >> 31) ? 0x10000u : 1u);
            }
        } else {
            const float rs = (WK == kWideF32) ? 1.0f : a.row_scale[i];
            const float sc = a.seg_scale ? a.seg_scale[a.seg_of[i]].x : a.scale_e;   // scale_e == scale_i
            // four slots per lane in flight: clamped indices keep the loads in the
            // row; slots past its end add nothing
            for (int64_t k0 = beg + lane; k0 < end; k0 += 4 * stride) {
                int32_t tar[4];
                uint32_t dl[4];
                float w32[4], wn[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t k = min(k0 + u * stride, end - 1);
                    tar[u] = a.col[k];
                    dl[u] = a.dly[k];
                    w32[u] = a.w32[k];
                    if (WK == kWideBF16) wn[u] = bf16_to_f32(a.w16[k]) * rs;
                    else if (WK == kWideFP8) wn[u] = fp8e4m3_to_f32(a.w8[k]) * rs;
                    else wn[u] = w32[u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (k0 + u * stride >= end) break;
                    int32_t slot = sD + (int32_t)(dl[u] & 127u);
                    slot -= (slot >= D) ? D : 0;
                    const float w = (dl[u] & 128u) ? w32[u] : wn[u];   // plastic weights stay FP32
                    const size_t loc = (size_t)slot * (size_t)a.npad + (size_t)tar[u];
                    if (w > 0.f) atomicAdd(&a.ring_e[loc], (int32_t)rintf(w * sc));
                    else atomicAdd(&a.ring_i[loc], (int32_t)rintf(w * sc));
                }
            }
        }
    }
    if (lane == 0 && se_local) atomicAdd(&se_blk, se_local);
    __syncthreads();
    if (threadIdx.x == 0 && se_blk) atomicAdd((unsigned long long*)&ctl->se, se_blk);
}

// Weight change of a pair dt = t_pre - t_post apart (mgcore.cu:346-360): the
// A_LTP branch for dt < 0, A_LTD for dt > 0, none at 0; tables indexed by |dt|.
__device__ __forceinline__ float stdp_dw(const float* ltp, const float* ltd, int32_t lut_n, int32_t dt) {
    if (dt < 0) return ltp[min(-dt, lut_n - 1)];
    if (dt > 0) return ltd[min(dt, lut_n - 1)];
    return 0.0f;
}

__device__ __forceinline__ float stdp_lut(const StdpArgs& a, int32_t dt) {
    return stdp_dw(a.ltp, a.ltd, a.lut_n, dt);
}

__device__ __forceinline__ float clipw(float w, float wmax, float wmin) {   // mgcore.cu:362
    return (w > wmax) ? wmax : ((w < wmin) ? wmin : w);
}

// STDP (window of one step): mgcore.cu:365-412 over plastic synapses.
__global__ void k_stdp(StdpArgs a, int32_t off, int32_t wk) {
    Ctl* ctl = a.ctl;
    const int par = (int)((ctl->base_window + wk) & 1);
    const uint32_t nspk = ctl->win_cnt[par];
    const uint32_t* wbuf = a.win_buf + (size_t)par * a.win_stride;
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t F = (uint32_t)a.split;           // waves per spike, as in k_prop
    const int64_t stride = 64 * (int64_t)F;
    for (uint32_t vw = wave; vw < nspk * F; vw += nwaves) {
        const uint32_t e = vw / F, f = vw - e * F;
        const int32_t i = (int32_t)(wbuf[e] & kIdMask);
        const int32_t li = a.last[i];
        const int64_t beg = a.row_ptr[i] + 64 * (int64_t)f, end = a.row_ptr[i + 1];
        for (int64_t k0 = beg + lane; k0 < end; k0 += 4 * stride) {  // axon loop :372-391
            int32_t lt[4];
            float wv[4];
            bool pl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {                           // loads of four slots in flight
                const int64_t k = min(k0 + u * stride, end - 1);
                pl[u] = (k0 + u * stride < end) && (a.dly[k] & 128u);
                lt[u] = a.last[a.col[k]];
                wv[u] = a.w32[k];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (pl[u]) a.w32[k0 + u * stride] = clipw(wv[u] + stdp_lut(a, li - lt[u]), a.w_max, a.w_min);
        }
        const int64_t pb = a.in_ptr[i] + 64 * (int64_t)f, pe = a.in_ptr[i + 1];
        for (int64_t p0 = pb + lane; p0 < pe; p0 += 4 * stride) {    // dendrite loop :392-411
            uint32_t kk[4];
            int32_t ls[4];
            float wv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t p = min(p0 + u * stride, pe - 1);
                kk[u] = a.in_pos[p];
                ls[u] = a.last[a.in_src[p]];
                wv[u] = a.w32[kk[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (p0 + u * stride < pe) a.w32[kk[u]] = clipw(wv[u] + stdp_lut(a, ls[u] - li), a.w_max, a.w_min);
        }
    }
}

__global__ void k_finalize(Ctl* ctl, int32_t steps, int32_t windows) {
    const int par = (int)((ctl->base_window + windows - 1) & 1);
    if (windows > 0) {
        ctl->rec_cnt += ctl->win_cnt[par];
        ctl->win_cnt[par] = 0u;
    }
    ctl->base_step += steps;
    ctl->base_window += windows;
}


// Inclusive block-wide prefix sum (blockDim.x a multiple of 64, <= 1024).
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* wsum) {
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o);
        if (lane >= (uint32_t)o) v += t;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    if (wid == 0) {
        uint32_t x = lane < nw ? wsum[lane] : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(x, o);
            if (lane >= (uint32_t)o) x += t;
        }
        if (lane < nw) wsum[lane] = x;
    }
    __syncthreads();
    if (wid > 0) v += wsum[wid - 1];
    return v;
}

// One tile's ring slice (rw halves of D x B words) from HBM into LDS.  Rows
// [r0, r0 + nz) mod D were consumed by the previous window and not stored
// back: they are cleared instead of loaded.  8 loads per thread in flight
// (clamped indices keep every load in bounds; the extra values are not stored).
__device__ __forceinline__ void ring_slice_in(uint32_t* ring, const uint32_t* gring, int32_t D, int32_t B,
                                              int32_t r0, int32_t nz, int RW) {
    const int32_t tid = (int32_t)threadIdx.x, bd = (int32_t)blockDim.x;
    auto copy_in = [&](uint32_t* rr, const uint32_t* gr, int32_t k0, int32_t k1) {
        for (int32_t kb = k0 + tid; kb < k1; kb += 8 * bd) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = gr[min(kb + u * bd, k1 - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (kb + u * bd < k1) rr[kb + u * bd] = v[u];
        }
    };
#pragma unroll
    for (int hh = 0; hh < RW; ++hh) {
        uint32_t* rr = ring + (size_t)hh * D * B;
        const uint32_t* gr = gring + (size_t)hh * D * B;
        const int32_t e = r0 + nz;
        const int32_t z0 = (e <= D) ? r0 : 0, z1 = (e <= D) ? e : e - D;   // zero rows [z0, z1)
        if (e <= D) {
            copy_in(rr, gr, 0, z0 * B);
            for (int32_t k = z0 * B + tid; k < z1 * B; k += bd) rr[k] = 0u;
            copy_in(rr, gr, z1 * B, D * B);
        } else {
            for (int32_t k = tid; k < z1 * B; k += bd) rr[k] = 0u;
            copy_in(rr, gr, z1 * B, r0 * B);
            for (int32_t k = r0 * B + tid; k < D * B; k += bd) rr[k] = 0u;
        }
    }
}

// LDS -> the tile's ring slice, except the min(nsteps, D) rows from rcur that
// this window consumed (cleared, or never read for neurons that do not integrate).
__device__ __forceinline__ void ring_slice_out(const uint32_t* ring, uint32_t* gring, int32_t D, int32_t B,
                                               int32_t rcur, int32_t nsteps, int RW) {
    const int32_t tid = (int32_t)threadIdx.x, bd = (int32_t)blockDim.x;
    const int32_t e = rcur + min(nsteps, D);
#pragma unroll
    for (int hh = 0; hh < RW; ++hh) {
        const uint32_t* rr = ring + (size_t)hh * D * B;
        uint32_t* gr = gring + (size_t)hh * D * B;
        if (e <= D) {            // rows [rcur, e) consumed
            for (int32_t k = tid; k < rcur * B; k += bd) gr[k] = rr[k];
            for (int32_t k = e * B + tid; k < D * B; k += bd) gr[k] = rr[k];
        } else {                 // rows [rcur, D) and [0, e - D) consumed
            for (int32_t k = (e - D) * B + tid; k < rcur * B; k += bd) gr[k] = rr[k];
        }
    }
}

// Words that workgroups of one persistent launch pass to each other between
// windows (spike-list entries and counts) are written and read at agent scope,
// which on gfx950 bypasses the XCD-private L2 (sc1).
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t ld_agent64(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide wait between two windows of a persistent launch.  Each thread first
// waits for its own stores to complete (s_waitcnt vmcnt(0): the agent-scope
// stores above are then visible at agent scope), then the workgroup publishes
// `epoch` and `count` (the window's spikes of the tile, when its list went to
// the tile's region) in its flag, one 64-bit word per tile kBarStride words apart
// (a cache line each), and waits until every workgroup's flag holds at least
// `epoch`, one flag per thread; with `cnt`, the counts of tiles [ct0, ct1) land in
// cnt[h] (0 for the others).  The wait is bounded: past `limit` clock ticks, or
// once another workgroup has reported a timeout, it sets *err and returns false,
// so every workgroup leaves the kernel (the host turns *err into an error).
__device__ __forceinline__ bool grid_wait(uint32_t* bar, uint32_t G, uint32_t g, uint32_t epoch, uint32_t count,
                                          uint32_t* cnt, uint32_t ct0, uint32_t ct1, uint32_t* err, uint64_t limit) {
    __builtin_amdgcn_s_waitcnt(0x0F70);    // vmcnt(0)
    __syncthreads();
    uint64_t* flag = reinterpret_cast<uint64_t*>(bar);
    constexpr size_t kStride64 = kBarStride / 2;
    if (threadIdx.x == 0) st_agent64(&flag[(size_t)g * kStride64], ((uint64_t)count << 32) | epoch);
    int bad = 0;
    for (uint32_t h = threadIdx.x; h < G && !bad; h += blockDim.x) {
        const uint64_t c0 = (uint64_t)wall_clock64();
        uint64_t v = ld_agent64(&flag[(size_t)h * kStride64]);
        while ((int32_t)((uint32_t)v - epoch) < 0) {
            if (ld_agent(err) != 0u) {
                bad = 1;
                break;
            }
            if ((uint64_t)wall_clock64() - c0 > limit) {
                st_agent(err, 1u);
                bad = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            v = ld_agent64(&flag[(size_t)h * kStride64]);
        }
        if (cnt) cnt[h] = (h >= ct0 && h < ct1) ? (uint32_t)(v >> 32) : 0u;
    }
    return __syncthreads_or(bad) == 0;
}

// One window of the fused tiled step (CLASS format).  Workgroup g (targets
// [tile_lo[g], tile_lo[g+1]) of segment tile_seg[g]):
//   1. loads its ring slice ring_t[g] (D x B words) into LDS;
//   2. adds the previous window's spikes of its segment: for each spiker, the sub-row of its
//      tile-sorted CSR row that targets this tile, split into work items of
//      `chunk` synapses balanced over the block, one LDS atomic per synapse;
//   3. advances the tile's neurons through this window's steps reading and
//      clearing their LDS ring rows, appending spikes (+ record, synaptic
//      events) with one wave ballot per step;
//   4. writes the ring slice back.
// Spike lists rotate over three buffers: window k writes k%3, reads (k-1)%3
// and clears the counter of (k+1)%3, which no workgroup touches in window k.
//
// WK selects the synapse word.  kTClass: target | delay-1 << 26 | inhibitory
// << 31, arrivals counted per class (16 + 16 bits of one ring word).  Wide
// formats: local target (11 bits) | delay-1 << 11 | BF16 (kTBF16) or FP8
// (kTFP8, per-source scale) weight bits << 16, or the weight in w_t (kTF32);
// arrivals add rintf(w * scale) into int32 exc / inh rings, as k_prop does.
//
// kTStdp: the kTF32 word with a plastic flag at bit 16 and pair-based STDP
// (mgcore.cu:346-412) on w_t.  kTStdpB: the same with BF16 static weights
// (times the source row's scale) in the word and 0xFFFF there for a plastic
// synapse; w_t is read only for spikes of rows with a plastic synapse into
// tiles with a plastic in-synapse (row_aux, tile_pl).  A synapse k -> j lives in the tile of its
// target, so every read and update of it happens in one workgroup.  The host
// caps the window at the shortest inter-spike interval of the plastic
// synapses' ends: k and j fire at most once in a window.  Axon updates (k's
// spike at t2: dt = t2 - j's last spike at t2) are made where the weight is
// read for propagation.  Dendrite updates (j's spike at t1: dt = k's last
// spike at t1 - t1) are deferred: neuron j keeps its last kStdpHist spike
// steps (newest first), and a synapse is brought up to date, replaying j's spikes in step
// order, (a) where it is read at k's next spike t2 — j's spikes after
// max(k's previous spike, tf[j]) and before t2, before the weight is
// propagated, which is the reference's order (each step: propagation, then
// STDP) — and (b) when j is flushed: after j has fired kStdpHist - 1 times since
// tf[j], all of j's plastic in-synapses are brought up to date through the
// previous window and tf[j] moves there, so the history always holds the spikes
// needed.  k's last spike 0 counts as "never" (a spike at step 0 then still
// replays; dt = 0 only clips, which is idempotent).  Window k's neurons read
// their last spike steps from last3 buffer k % 3 and write them to (k+1) % 3.
// With a.apply_only the launch brings every plastic synapse up to date through
// the pending window in a.w_t (a copy), touching nothing else: the weights
// the reference holds after the last step run.
//
// PW (persistent): the launch runs nwin windows of nsteps steps each, every
// workgroup resident for all of them (the host checks occupancy).  The ring
// slice stays in LDS and, for tiles of at most one neuron per thread, the
// neuron state in registers, from the first window to the last; the window
// lists pass between workgroups through agent-scope accesses and a grid_wait
// between windows.  Not with STDP.
template <int CH, int WK, bool PW, bool RG = false>
__global__ __launch_bounds__(1024) void k_tiled(TiledArgs a, int32_t off, int32_t wk, int32_t nsteps,
                                                int32_t nwin) {
    static_assert(!(PW && (WK == kTStdp || WK == kTStdpB)), "the persistent form has no STDP");
    static_assert(!RG || (PW && WK == kTClass16), "list regions: persistent form of 16-bit CLASS words only");
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ uint32_t wsum[16];
    __shared__ unsigned long long se_blk, lsp_r;
    __shared__ uint32_t lsp_n, lsp_g;
    // CLASS: the first kClassCache k-fold sums of the classes of the tile's
    // last and first neurons (one class per tile in population-ordered networks)
    __shared__ float ccache[2][2][kClassCache];
    __shared__ uint32_t ccls[2];
    const int32_t D = a.D, B = a.B, bd = (int32_t)blockDim.x;
    const int32_t tid0 = (int32_t)threadIdx.x;
    // the workgroup's tile.  Workgroups go round-robin to the kXcds XCDs
    // (blockIdx.x % kXcds), each with its own L2; neighbouring tiles share the
    // cache lines at their sub-row boundaries and their row offsets, so each XCD
    // takes one contiguous run of tiles (a bijection onto [0, G))
    // (only when the whole grid is resident: a later workgroup waits for a CU of
    // its own XCD, and tiles of unequal cost are then better spread over XCDs)
    const int32_t g = [&] {
        const int32_t b = (int32_t)blockIdx.x, G = (int32_t)gridDim.x;
        if (!a.xcd_map) return b;
        const int32_t x = b % kXcds, qx = G / kXcds, rx = G % kXcds;
        return x * qx + min(x, rx) + b / kXcds;
    }();
    constexpr bool kCls = WK == kTClass || WK == kTClass16;  // CLASS arrival counts
    constexpr bool kMix = WK == kTStdpB;                    // BF16 static weights beside plastic ones
    constexpr bool kPl = WK == kTStdp || kMix;              // STDP on w_t
    constexpr bool kAux = WK == kTFP8 || WK == kTBF16 || kPl;
    const int RW = kCls ? 1 : a.rw;                         // ring words per (slot, target)
    uint32_t* ring = lds;                                   // [RW][D][B]
    // (STDP: rounded up to 16 bytes, for the history rows' vector reads below)
    uint32_t* sp_beg = ring + (kPl ? (((size_t)RW * D * B + 3) & ~(size_t)3) : (size_t)RW * D * B);
    uint32_t* sp_len = sp_beg + kTiledBatch;                // len | slot-of-spike << 26
    uint32_t* sp_start = sp_len + kTiledBatch;              // exclusive prefix of work items
    constexpr uint32_t kItems = kTiledItems;   // work items staged per round
    uint16_t* itm = reinterpret_cast<uint16_t*>(sp_start + kTiledBatch);   // [kItems] spike of item
    // [kTiledBatch] BF16/FP8: row scale bits; STDP: the source's last spike before the window
    uint32_t* sp_aux = reinterpret_cast<uint32_t*>(itm + kItems);
    // [kTiledBatch] kTStdpB: row scale bits | weights-to-load << 31
    uint32_t* sp_aux2 = sp_aux + kTiledBatch;
    // STDP, per tile neuron: spike history [B][kStdpHist] (newest first, kHistNone past
    // the neuron's spikes; 64-byte rows), fire count, tf
    int32_t* hL = reinterpret_cast<int32_t*>(sp_aux + (kMix ? 2 : 1) * kTiledBatch);
    int32_t* cL = hL + (size_t)B * kStdpHist;
    int32_t* tfL = cL + B;
    // the first lut_lds entries of the LTP and LTD tables
    float* lutL = reinterpret_cast<float*>(tfL + B);
    // STDP weight change for dt = t_pre - t_post (stdp_dw) from the LDS copy where it holds |dt|
    // (lut_lds >= 1; entry 0 of both tables is 0, so dt = 0 gives 0), else from HBM.  The HBM
    // read is a buffer load over both tables (ltd follows ltp), so that the compiler keeps an
    // LDS load and a buffer load instead of one flat load through a selected pointer
    const __amdgpu_buffer_rsrc_t lut_rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(kPl ? a.ltp : nullptr), (short)0, kPl ? 8 * a.lut_n : 0, kBufferDword3);
    auto dw = [&](int32_t dt) -> float {
        const int32_t k = dt < 0 ? -dt : dt;
        float v = lutL[(dt < 0 ? 0 : a.lut_lds) + min(k, a.lut_lds - 1)];
        if (k >= a.lut_lds)
            v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                lut_rs, 4 * ((dt < 0 ? 0 : a.lut_n) + min(k, a.lut_n - 1)), 0, 0));
        return v;
    };
    const bool full = !kPl || a.apply_only == 0;
    Ctl* ctl = a.ctl;
    // fixed-point (scale, 1 / scale) of the tile's segment (scale_e == scale_i)
    const float2 fsc = a.seg_scale ? a.seg_scale[a.tile_seg[g]] : make_float2(a.scale_e, a.inv_e);
    const int32_t t0 = a.tile_lo[g];
    const int32_t tn = a.tile_lo[g + 1] - t0;
    // kTStdpB: a tile without plastic in-synapses never loads w_t
    const uint32_t plmask = (kMix && a.tile_pl[g]) ? 0xFFFFFFFFu : 0x7FFFFFFFu;
    // the tile's segment (one segment: all of it known without loads)
    const bool one = a.nseg == 1;
    const int32_t q = one ? 0 : (int32_t)a.tile_seg[g];
    const int32_t lt = one ? g : g - a.seg_tile0[q];        // tile index inside the segment
    const int32_t slo = one ? 0 : a.seg_lo[q];
    const int32_t nq = one ? a.n : a.seg_lo[q + 1] - slo;    // >= 1: tiles cover neurons
    const size_t G1 = one ? (size_t)a.G + 1 : (size_t)(a.seg_tile0[q + 1] - a.seg_tile0[q] + 1);
    const size_t obase = one ? (size_t)g : (size_t)a.seg_obase[q] + (size_t)lt;
    const size_t lbase = (size_t)slo * (size_t)a.window;    // segment's part of each spike list
    uint32_t* tcnt = a.tcnt + q;                            // tcnt[b * nseg]
    const int32_t nseg = a.nseg;
    const int64_t sbase = ctl->base_step + off;
    const int64_t wbase = ctl->base_window + wk;
    const int32_t nw = PW ? nwin : 1;
    // persistent launches with per-tile list regions (a.region): between windows the
    // tiles' spike counts, then their exclusive prefix, and the tiles' list offsets
    // (tile h's part of a list starts at tile_lo[h] * window), [G] words each of LDS
    // after the staging area (the launch brings them)
    uint32_t* rg_pre = reinterpret_cast<uint32_t*>(itm + kItems) + (kAux ? kTiledBatch : 0);
    uint32_t* rg_off = rg_pre + (PW ? a.G : 0);
    // (a variant of the persistent form of the 16-bit CLASS words only: in the other
    // kinds' persistent forms the code alone, never run, cost their windows 5 %, and
    // it keeps the neuron-state preload out of the variant, which would spill)
    constexpr bool kRg = RG;
    const bool rgn_on = kRg && a.region != 0;
    if (rgn_on)
        for (int32_t h = tid0; h < a.G; h += bd) rg_off[h] = (uint32_t)a.tile_lo[h] * (uint32_t)a.window;
    // every window's list but a launch's last goes to the regions; whether the previous
    // window's did (thread-uniform)
    bool rg_prev = false;
    const uint32_t gt0 = one ? 0u : (uint32_t)a.seg_tile0[q];               // the segment's tiles
    const uint32_t gt1 = one ? (uint32_t)a.G : (uint32_t)a.seg_tile0[q + 1];
    // (thread 0) synaptic events and spikes of the windows that wrote regions, added
    // to the run's counters by the launch's last window
    unsigned long long se_acc = 0ull, spk_acc = 0ull;
    // a tile of at most one neuron per thread: the neuron's state and out-degree are
    // loaded at the start of each window (after the ring copy, which they would
    // delay), their latency hidden by the propagation phase
    // (not where registers would spill: with STDP at 16 synapses a work item or with
    // BF16 static weights, in the persistent form of the 32-bit CLASS words and in the
    // list-region variant)
    const bool pre = !(kRg || (PW && WK == kTClass)) && tn <= bd && full && !(kPl && (CH >= 16 || kMix));
    uint32_t* gring = a.ring_t + (size_t)g * a.ring_slab;
    for (int32_t kw = 0; kw < nw; ++kw) {
        // (persistent form: the thread index is opaque to the compiler in each window,
        // so that the per-thread addresses derived from it are formed inside the window
        // instead of being hoisted out of the loop and held in registers for all of it)
        int32_t tid = tid0;
        if constexpr (PW) asm volatile("" : "+v"(tid));
        const int64_t s0 = sbase + (int64_t)kw * nsteps;
        const int64_t kwin = wbase + kw;
        const int bcur = (int)(kwin % 3), bprev = (int)((kwin + 2) % 3), bnext = (int)((kwin + 1) % 3);
        // The previous window's list: its start step, spike count and this thread's
        // first entry.  In a launch's first window the window's index comes from ctl,
        // and the words of all three lists load with it (one round trip instead of
        // two, the index then picks one; the other lists' words, which other
        // workgroups may be writing, are not used).  Later windows of a persistent
        // launch know the index and read the list other workgroups just wrote.
        const uint32_t li = min((uint32_t)tid, (uint32_t)nq * (uint32_t)a.window - 1u);
        const uint32_t* L = a.win_buf + (size_t)bprev * a.win_stride + lbase;
        // the previous window's list is in the tiles' regions
        const bool rgc = rgn_on && kw > 0 && rg_prev;
        // entry e (< the list's count) of the previous window's list: in the regions,
        // tile h (the last one whose prefix is at most e) holds it at e - prefix
        auto rentry = [&](uint32_t e) -> uint32_t {
            uint32_t lo = gt0, hi = gt1 - 1u;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (rg_pre[mid] <= e) lo = mid;
                else hi = mid - 1u;
            }
            return ld_agent(&a.win_buf[(size_t)bprev * a.win_stride + rg_off[lo] + (e - rg_pre[lo])]);
        };
        auto lent = [&](uint32_t e) -> uint32_t {
            if (rgc) return rentry(e);
            return PW ? ld_agent(&L[e]) : L[e];
        };
        int64_t sprev;
        uint32_t E, ent0;
        if (kw == 0) {
            const size_t ws = a.win_stride;
            const int64_t tb0 = ctl->tbase[0], tb1 = ctl->tbase[1], tb2 = ctl->tbase[2];
            const uint32_t n0 = tcnt[0], n1 = tcnt[nseg], n2 = tcnt[2 * nseg];
            const uint32_t l0 = a.win_buf[lbase + li], l1 = a.win_buf[ws + lbase + li];
            const uint32_t l2 = a.win_buf[2 * ws + lbase + li];
            sprev = bprev == 0 ? tb0 : (bprev == 1 ? tb1 : tb2);
            E = bprev == 0 ? n0 : (bprev == 1 ? n1 : n2);
            ent0 = bprev == 0 ? l0 : (bprev == 1 ? l1 : l2);
        } else if (rgc) {
            // the counts came with the wait (rg_pre); the list is laid out by their prefix
            sprev = s0 - nsteps;
            const uint32_t c = (uint32_t)tid < (uint32_t)a.G ? rg_pre[tid] : 0u;
            const uint32_t incl = block_incl_scan(c, wsum);
            E = wsum[(blockDim.x >> 6) - 1];
            if ((uint32_t)tid < (uint32_t)a.G) rg_pre[tid] = incl - c;
            __syncthreads();
            ent0 = (uint32_t)tid < E ? rentry((uint32_t)tid) : 0u;
        } else {
            sprev = s0 - nsteps;
            E = ld_agent(&tcnt[bprev * nseg]);
            ent0 = ld_agent(&L[li]);
        }
        // this window's list goes to the tiles' regions (another window of the launch follows)
        const bool rgp = rgn_on && kw + 1 < nw;
        rg_prev = rgp;
        if (full && lt == 0 && tid == 0) {
            if (PW) st_agent(&tcnt[bnext * nseg], 0u);
            else tcnt[bnext * nseg] = 0u;
        }
        if (kPl && full && g == 0 && tid == 0) a.fcnt[(kwin + 1) & 1] = 0u;
        if (full && g == 0 && tid == 0) ctl->tbase[bcur] = s0;
        if (tid == 0) {
            se_blk = 0ull;
            lsp_n = 0u;
        }
        const int32_t* lastA = kPl ? a.last3 + (size_t)bprev * a.last_stride : nullptr;   // window k-1 start
        const int32_t* lastB = kPl ? a.last3 + (size_t)bcur * a.last_stride : a.last;     // window k-1 end
        int32_t* lastC = kPl ? a.last3 + (size_t)bnext * a.last_stride : a.last;          // window k end
        if (kPl) {
            for (int32_t j = tid; j < tn; j += bd) {
                cL[j] = a.cnt[t0 + j];
                tfL[j] = a.tf[t0 + j];
            }
            const int32_t* hg = a.hist + (size_t)t0 * kStdpHist;
            for (int32_t k = tid; k < tn * kStdpHist; k += bd) hL[k] = hg[k];
            for (int32_t k = tid; k < a.lut_lds; k += bd) {
                lutL[k] = a.ltp[k];
                lutL[a.lut_lds + k] = a.ltd[k];
            }
        }
        if (kCls && kw == 0 && tid < 4 * kClassCache) {
            // (precomputed per tile by k_tile_ccache: one load instead of a chain of three)
            ccache[tid / (2 * kClassCache)][(tid / kClassCache) & 1][tid % kClassCache] =
                a.tile_cc[(size_t)g * 4 * kClassCache + tid];
            if (tid < 2) ccls[tid] = a.tile_ccls[2 * g + tid];
        }
        // 1. ring slice -> LDS.  The slots the previous window consumed ([sprev, s0)
        // mod D) were cleared by it and not stored back (step 4): they are zero here.
        // 0 before the first window; at most D (a network without synapses has D = 1
        // and windows of up to kMaxWindow steps)
        const int32_t nprev = (int32_t)min(s0 - sprev, (int64_t)D);
        const int32_t rprev = (int32_t)(sprev % D);
        // metadata of list entry e: first slot of its sub-row in this tile, length |
        // ring slot of its spike step << 26, BF16 / FP8 row scale bits or (STDP)
        // the source's last spike before the window; kTStdpB: row scale bits |
        // (the row holds a plastic synapse and the tile a plastic in-synapse) << 31
        auto meta = [&](uint32_t ent, uint32_t& beg, uint32_t& lm, uint32_t& aux, uint32_t& aux2) {
            const uint32_t src = ent & kIdMask;
            const size_t ob = obase + (size_t)(src - (uint32_t)slo) * G1;
            uint32_t o0, o1;
            if (a.offs16) { o0 = a.offs16[ob]; o1 = a.offs16[ob + 1]; }
            else { o0 = a.offs32[ob]; o1 = a.offs32[ob + 1]; }
            const int32_t sD = (int32_t)((sprev + (int64_t)(ent >> 26)) % D);
            // padded sub-rows: starts are multiples of kSubAlign, the low bits of the next
            // entry hold the padding
            constexpr uint32_t am = (uint32_t)kSubAlign - 1u;
            beg = (uint32_t)a.prow[src] + (o0 & ~am);
            lm = ((o1 & ~am) - (o0 & ~am) - (o1 & am)) | ((uint32_t)sD << 26);
            if (WK == kTFP8 || WK == kTBF16) aux = __float_as_uint(a.row_scale[src]);
            else if (kPl) aux = (uint32_t)lastA[src];
            else aux = 0u;
            aux2 = kMix ? (a.row_aux[src] & plmask) : 0u;
        };
        // the first batch's list entries are loaded before their count is known (the
        // index clamped into the segment's part of the list, which holds nq * window
        // entries), and their metadata loads before the ring copy, so that these
        // latencies overlap; entries past the count are never used
        uint32_t pbeg = 0u, plm = 0u, paux = 0u, paux2 = 0u;
        if ((uint32_t)tid < min((uint32_t)bd, E)) meta(ent0, pbeg, plm, paux, paux2);
        // windows of more than one batch (not with STDP, whose registers are spent): the
        // entries of batch b + 2 load during batch b and their metadata during batch b + 1,
        // so neither latency is waited for between batches (ent1: the next batch's entries)
        constexpr bool kAhead = !kPl;
        uint32_t ent1 = 0u;
        if (kAhead && E > (uint32_t)bd && (uint32_t)tid < E - (uint32_t)bd) ent1 = lent((uint32_t)bd + (uint32_t)tid);
        if (full && kw == 0) ring_slice_in(ring, gring, D, tn, rprev, nprev, RW);
        NState pst;
        int ptype = T_RAW;
        int32_t pdeg = 0;
        // CLASS kinds: every neuron of the tile (one per thread) an LIFB neuron of one of
        // the tile's two cached classes (precomputed per tile)
        const bool tile_lean = kCls && pre && a.tile_lean && a.tile_lean[g] != 0;
        if (pre && tid < tn) {
            ptype = (int)a.type[t0 + tid];
            neuron_load(a, t0 + tid, ptype, kCls, pst, lastB);
            if (ptype != T_RAW && ptype != T_LIF0) pdeg = (int32_t)(a.row_ptr[t0 + tid + 1] - a.row_ptr[t0 + tid]);
        }
        __syncthreads();

        // 2. previous window's spikes into the tile
        for (uint32_t b0 = 0; b0 < E; b0 += (uint32_t)bd) {
            const uint32_t nb = min((uint32_t)bd, E - b0);
            uint32_t items = 0u;
            if ((uint32_t)tid < nb) {
                uint32_t beg = pbeg, lm = plm, aux = paux, aux2 = paux2;
                if (!kAhead && b0 > 0) meta(lent(b0 + (uint32_t)tid), beg, lm, aux, aux2);
                sp_beg[tid] = beg;
                sp_len[tid] = lm;
                if (kAux) sp_aux[tid] = aux;
                if (kMix) sp_aux2[tid] = aux2;
                items = ((lm & kIdMask) + CH - 1) / CH;
            }
            if constexpr (kAhead) {
                // the next batch's metadata from its entries (loaded a batch ago), and the
                // entries of the batch after it
                const uint32_t b1 = b0 + (uint32_t)bd, b2 = b1 + (uint32_t)bd;
                if (b1 < E && (uint32_t)tid < E - b1) meta(ent1, pbeg, plm, paux, paux2);
                if (b2 < E && (uint32_t)tid < E - b2) ent1 = lent(b2 + (uint32_t)tid);
            }
            const uint32_t incl = block_incl_scan(items, wsum);
            const uint32_t total = wsum[(blockDim.x >> 6) - 1];
            if ((uint32_t)tid < nb) sp_start[tid] = incl - items;
            for (uint32_t r0 = 0; r0 < total; r0 += kItems) {
                const uint32_t rn = min(kItems, total - r0);
                // item table: each spike writes its own work items of this round
                if ((uint32_t)tid < nb) {
                    const uint32_t q0 = max(incl - items, r0), q1 = min(incl, r0 + rn);
                    for (uint32_t q = q0; q < q1; ++q) itm[q - r0] = (uint16_t)tid;
                }
                __syncthreads();
                // U items per thread in flight: U * CH synapse words loaded before the
                // first LDS atomic
                // (STDP: at most 8 synapses, whose STDP state then fits in the registers)
                constexpr int U = CH >= 16 ? 1 : (kPl ? (CH >= 8 ? 1 : (CH == 1 ? 4 : 8 / CH)) : 16 / CH);
                for (uint32_t it = (uint32_t)tid; it < rn; it += (uint32_t)(U * bd)) {
                    // branch-free body, so that all U * CH loads are in flight before the
                    // first wait: lanes past the round's items re-read the last item and
                    // elements past an item's end re-read its last element; both add 0
                    uint32_t cnt[U], kb[U];
                    int32_t sD[U], t2[U], la[U];
                    float rs[U];
                    bool lw[U];              // kTStdpB: the item's weights are loaded
    #pragma unroll
                    for (int h = 0; h < U; ++h) {
                        const uint32_t q0 = it + (uint32_t)(h * bd);
                        const uint32_t q = min(q0, rn - 1u);
                        const uint32_t j = itm[q];
                        const uint32_t c = r0 + q - sp_start[j];
                        const uint32_t meta = sp_len[j];
                        const uint32_t len = meta & kIdMask;
                        sD[h] = (int32_t)(meta >> 26);
                        kb[h] = sp_beg[j] + c * CH;
                        cnt[h] = (q0 < rn) ? min((uint32_t)CH, len - c * CH) : 0u;
                        rs[h] = (WK == kTFP8 || WK == kTBF16) ? __uint_as_float(sp_aux[j]) : 1.0f;
                        lw[h] = true;
                        if (kMix) {
                            const uint32_t a2 = sp_aux2[j];
                            rs[h] = __uint_as_float(a2 & 0x7FFFFFFFu);
                            lw[h] = (a2 >> 31) != 0u;
                        }
                        if (kPl) {
                            // the spike's step: the window is shorter than the ring, so its
                            // slot determines it
                            const int32_t dd = sD[h] - rprev;
                            t2[h] = (int32_t)sprev + dd + (dd < 0 ? D : 0);
                            la[h] = (int32_t)sp_aux[j];
                        }
                    }
                    uint32_t x[U][CH];
                    float wv[U][CH];
    #pragma unroll
                    for (int h = 0; h < U; ++h) {
                        if constexpr (WK == kTClass16 && CH >= 2) {
                            // the item's CH words as aligned vector loads (sub-rows start at
                            // multiples of kSubAlign words; the words past the sub-row's end
                            // are padding, the next sub-row or the slack after the last one)
                            const uint16_t* p16 = a.syn16 + kb[h];
                            if constexpr (CH == 2) {
                                const uint32_t v = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p16));
                                x[h][0] = v & 0xFFFFu;
                                x[h][1] = v >> 16;
                            } else if constexpr (CH == 4) {
                                const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p16));
                                x[h][0] = v.x & 0xFFFFu;
                                x[h][1] = v.x >> 16;
                                x[h][2] = v.y & 0xFFFFu;
                                x[h][3] = v.y >> 16;
                            } else {
    #pragma unroll
                                for (int v8 = 0; v8 < CH / 8; ++v8) {
                                    const u32x4 v =
                                        __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p16) + v8);
                                    x[h][8 * v8 + 0] = v.x & 0xFFFFu;
                                    x[h][8 * v8 + 1] = v.x >> 16;
                                    x[h][8 * v8 + 2] = v.y & 0xFFFFu;
                                    x[h][8 * v8 + 3] = v.y >> 16;
                                    x[h][8 * v8 + 4] = v.z & 0xFFFFu;
                                    x[h][8 * v8 + 5] = v.z >> 16;
                                    x[h][8 * v8 + 6] = v.w & 0xFFFFu;
                                    x[h][8 * v8 + 7] = v.w >> 16;
                                }
                            }
                            continue;
                        }
                        if constexpr (WK != kTClass16 && CH >= 2) {
                            // 32-bit words and FP32 weights: aligned vector loads likewise
                            // (STDP rewrites its weights below: cached loads; kTStdpB loads
                            // them only for rows and tiles with plastic synapses)
                            const uint32_t* p32 = a.syn_t + kb[h];
                            const u32x2* f2 = reinterpret_cast<const u32x2*>(a.w_t + kb[h]);
                            const u32x4* f4 = reinterpret_cast<const u32x4*>(a.w_t + kb[h]);
                            if constexpr (CH == 2) {
                                const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p32));
                                x[h][0] = v.x;
                                x[h][1] = v.y;
                                if (WK == kTF32 || kPl) {
                                    u32x2 f = {0u, 0u};
                                    if (lw[h]) f = kPl ? *f2 : __builtin_nontemporal_load(f2);
                                    wv[h][0] = __uint_as_float(f.x);
                                    wv[h][1] = __uint_as_float(f.y);
                                }
                            } else {
    #pragma unroll
                                for (int v4 = 0; v4 < CH / 4; ++v4) {
                                    const u32x4 v =
                                        __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p32) + v4);
                                    x[h][4 * v4 + 0] = v.x;
                                    x[h][4 * v4 + 1] = v.y;
                                    x[h][4 * v4 + 2] = v.z;
                                    x[h][4 * v4 + 3] = v.w;
                                    if (WK == kTF32 || kPl) {
                                        u32x4 f = {0u, 0u, 0u, 0u};
                                        if (lw[h]) f = kPl ? f4[v4] : __builtin_nontemporal_load(f4 + v4);
                                        wv[h][4 * v4 + 0] = __uint_as_float(f.x);
                                        wv[h][4 * v4 + 1] = __uint_as_float(f.y);
                                        wv[h][4 * v4 + 2] = __uint_as_float(f.z);
                                        wv[h][4 * v4 + 3] = __uint_as_float(f.w);
                                    }
                                }
                            }
                            continue;
                        }
                        const uint32_t last = max(cnt[h], 1u) - 1u;
    #pragma unroll
                        for (int u = 0; u < CH; ++u) {
                            const uint32_t k = kb[h] + min((uint32_t)u, last);
                            if (WK == kTClass16) x[h][u] = __builtin_nontemporal_load(&a.syn16[k]);
                            else x[h][u] = __builtin_nontemporal_load(&a.syn_t[k]);
                            if (WK == kTF32) wv[h][u] = __builtin_nontemporal_load(&a.w_t[k]);
                            else if (kPl) wv[h][u] = lw[h] ? a.w_t[k] : 0.0f;   // rewritten below: cached load
                        }
                    }
                    // STDP, per plastic element: j's history newest first gives its last spike
                    // at or before t2 (for the axon update) and the range [mlo, mhi] of
                    // deferred dendrite updates in (lo, t2); the axon table loads of all
                    // elements are then issued together
                    int32_t rng[U][CH];      // mlo | mhi << 8 (mhi < mlo: none); -1: not plastic
                    int32_t adt[U][CH];      // dt of the axon update
                    float ax[U][CH];
                    if (kPl) {
    #pragma unroll
                        for (int h = 0; h < U; ++h)
    #pragma unroll
                            for (int u = 0; u < CH; ++u) {
                                const uint32_t xw = x[h][u];
                                const int32_t loc = (int32_t)(xw & 0x7FFu);
                                int32_t lj = 0, mlo = kStdpHist, mhi = 0;
                                rng[h][u] = -1;
                                // (words past the item's end are padding or another tile's)
                                const bool plw = kMix ? (xw >> 16) == 0xFFFFu : ((xw >> 16) & 1u) != 0u;
                                if ((uint32_t)u < cnt[h] && plw) {
                                    // j fires at most once a window: only its newest spike can be
                                    // at or after t2 (then the one before is j's last spike before
                                    // t2), and the entries after lo are a prefix of the history,
                                    // read four at a time
                                    const int4* hv = reinterpret_cast<const int4*>(hL + loc * kStdpHist);
                                    const int32_t lo = max(la[h] == 0 ? -1 : la[h], tfL[loc]);
                                    const int4 q = hv[0];
                                    lj = q.x <= t2[h] ? q.x : q.y;
                                    if (lj == kHistNone) lj = 0;       // none before t2: never fired
                                    int32_t ngt = (q.x > lo) + (q.y > lo) + (q.z > lo) + (q.w > lo);
                                    for (int b4 = 1; b4 < kStdpHist / 4 && ngt == 4 * b4; ++b4) {
                                        const int4 r = hv[b4];
                                        ngt += (r.x > lo) + (r.y > lo) + (r.z > lo) + (r.w > lo);
                                    }
                                    mlo = q.x >= t2[h] ? 1 : 0;        // (a spike at t2: dt = 0)
                                    if (ngt > mlo) mhi = ngt - 1;
                                    else mlo = kStdpHist;
                                    rng[h][u] = mlo | (mhi << 8);
                                }
                                adt[h][u] = t2[h] - lj;
                            }
    #pragma unroll
                        for (int h = 0; h < U; ++h)
    #pragma unroll
                            for (int u = 0; u < CH; ++u)
                                ax[h][u] = rng[h][u] >= 0 ? dw(adt[h][u]) : 0.0f;
                    }
    #pragma unroll
                    for (int h = 0; h < U; ++h)
    #pragma unroll
                        for (int u = 0; u < CH; ++u) {
                            // words past the item's end (padding, another tile's words) and
                            // lanes past the round's items: a distinct in-slice entry per
                            // lane, adding 0
                            const bool valid = (uint32_t)u < cnt[h];
                            const uint32_t xw = x[h][u];
                            if (WK == kTClass) {
                                int32_t slot = sD[h] + (int32_t)((xw >> 26) & 31u) + 1;
                                slot -= (slot >= D) ? D : 0;
                                const int32_t loc = (int32_t)(xw & kIdMask) - t0;
                                const int32_t at = valid ? slot * tn + loc : (tid % tn);
                                atomicAdd(&ring[at], valid ? ((xw >> 31) ? 0x10000u : 1u) : 0u);
                            } else if (WK == kTClass16) {
                                int32_t slot = sD[h] + (int32_t)((xw >> 10) & 31u) + 1;
                                slot -= (slot >= D) ? D : 0;
                                const int32_t loc = (int32_t)(xw & 0x3FFu);
                                const int32_t at = valid ? slot * tn + loc : (tid % tn);
                                atomicAdd(&ring[at], valid ? ((xw >> 15) ? 0x10000u : 1u) : 0u);
                            } else {
                                int32_t slot = sD[h] + (int32_t)((xw >> 11) & 31u) + 1;
                                slot -= (slot >= D) ? D : 0;
                                const int32_t loc = (int32_t)(xw & 0x7FFu);
                                float w;
                                if (WK == kTBF16) w = bf16_to_f32((uint16_t)(xw >> 16)) * rs[h];
                                else if (WK == kTFP8) w = fp8e4m3_to_f32((uint8_t)(xw >> 16)) * rs[h];
                                else if (kMix && (xw >> 16) != 0xFFFFu) w = bf16_to_f32((uint16_t)(xw >> 16)) * rs[h];
                                else w = wv[h][u];
                                const bool pl = kPl && rng[h][u] >= 0;
                                if (pl) {
                                    // deferred dendrite updates before t2, oldest first
                                    const int32_t* hj = hL + loc * kStdpHist;
                                    const int32_t mlo = rng[h][u] & 0xFF, mhi = rng[h][u] >> 8;
                                    for (int32_t m = mhi; m >= mlo; --m) {
                                        const int32_t e = hj[m];
                                        w = clipw(w + dw(la[h] - e), a.w_max, a.w_min);
                                    }
                                }
                                const bool exc = w > 0.f;
                                const int32_t v = (int32_t)rintf(w * fsc.x);
                                // (one ring: both signs into one fixed-point sum)
                                const int32_t at = valid ? ((exc || RW == 1 ? 0 : D) + slot) * tn + loc : (tid % tn);
                                if (full) atomicAdd(reinterpret_cast<int32_t*>(ring) + at, valid ? v : 0);
                                if (pl && valid)                           // axon update of t2
                                    a.w_t[kb[h] + (uint32_t)u] = clipw(w + ax[h][u], a.w_max, a.w_min);
                            }
                        }
                }
                __syncthreads();
            }
            __syncthreads();   // wsum / sp_* are rewritten by the next batch
        }
        if (kPl) {
            // 2b. when a flush follows this window: neurons that fired kStdpHist - 2 times
            // since tf (for apply_only: all that fired since tf) go to this window's flush
            // list (k_stdp_flush, launched next).  A flush follows at least every second
            // window and a neuron fires at most once a window, so one not listed here
            // (at most kStdpHist - 3) is listed at the next flush with at most
            // kStdpHist - 1, and its history then holds every spike since tf plus the one
            // of that window
            uint32_t* fc = a.fcnt + (kwin & 1);
            int32_t* fli = a.flist + (size_t)(kwin & 1) * a.flist_cap * 2;
            for (int32_t j = tid; j < tn; j += bd) {
                const int32_t n_since = cL[j] - a.cntf[t0 + j];
                if (full ? (a.flush_next && n_since >= kStdpHist - 2) : n_since > 0) {
                    const uint32_t e = atomicAdd(fc, 1u);
                    fli[2 * e] = t0 + j;
                    fli[2 * e + 1] = tfL[j];
                    if (full) {
                        a.tf[t0 + j] = (int32_t)(s0 - 1);
                        a.cntf[t0 + j] = cL[j];
                    }
                }
            }
            if (!full) return;
        }

        // 3. neurons of the tile through this window's steps.  Spikes go to an LDS
        // list (the staging area of step 2, free now; one LDS atomic per wave and
        // step) that joins the segment's window list with one global atomic at the
        // end; lanes that find it full append to HBM directly.
        uint32_t* lsp = sp_beg;
        uint32_t* wbuf = a.win_buf + (size_t)bcur * a.win_stride + lbase;
        uint32_t* rgw = a.win_buf + (size_t)bcur * a.win_stride + (size_t)t0 * (size_t)a.window;   // (rgp)
        const int32_t slot0 = (int32_t)(s0 % D);
        unsigned long long se_local = 0ull;
        for (int32_t j0 = 0; j0 < tn; j0 += bd) {
            // a wave without neurons here skips the steps (its instructions would still
            // take the SIMD's cycles)
            if (j0 + (tid & ~63) >= tn) continue;
            const int32_t j = j0 + tid;
            const bool active = j < tn;
            const int32_t i = t0 + j;
            const int type = active ? (pre ? ptype : (int)a.type[i]) : T_RAW;
            NState st;
            int32_t outdeg = 0;
            if (pre) {
                st = pst;
                outdeg = pdeg;
            } else if (active) {
                neuron_load(a, i, type, kCls, st, lastB);
                if (type != T_RAW && type != T_LIF0) outdeg = (int32_t)(a.row_ptr[i + 1] - a.row_ptr[i]);
            }
            const bool integ = integrates(type);
            // wave-uniform fast path: every active lane an LIFB neuron of a cached class
            int32_t csl = -1;
            if (kCls && active && integ) csl = (st.cls == ccls[0]) ? 0 : ((st.cls == ccls[1]) ? 1 : -1);
            // spikes of step w: one wave ballot; the wave's leader reserves their places in
            // the LDS list (one LDS atomic per wave and step)
            auto append = [&](bool fired, int32_t w, int32_t s) {
                const uint64_t m = __ballot(fired);
                if (m == 0ull) return;
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                const uint32_t pc = (uint32_t)__popcll(m);
                uint32_t lb = 0u;
                if (lane_id() == leader) lb = atomicAdd(&lsp_n, pc);
                lb = __shfl(lb, (int)leader);
                const uint32_t cap = (uint32_t)kTiledBatch;
                const uint32_t over0 = max(lb, cap);                 // first index past the LDS list
                if (rgp) {
                    // the tile's region holds every entry at its place in the tile's list
                    // (written now, complete by the wait); the LDS list keeps the first cap
                    // for the records, and records past them take their places here
                    const uint32_t r = lb + rank_below(m);
                    if (fired) st_agent(&rgw[r], (uint32_t)i | ((uint32_t)w << 26));
                    if (a.rec && lb + pc > cap) {
                        unsigned long long rb = 0ull;
                        if (lane_id() == leader)
                            rb = atomicAdd((unsigned long long*)&ctl->rec_cnt, (unsigned long long)(lb + pc - over0));
                        rb = __shfl(rb, (int)leader);
                        if (fired && r >= cap) {
                            const uint64_t at = rb + (r - over0);
                            if (at < a.rec_cap) a.rec[at] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)i;
                            else ctl->rec_overflow = 1;
                        }
                    }
                } else if (lb + pc > cap) {
                    uint32_t gb = 0u;
                    unsigned long long rb = 0ull;
                    if (lane_id() == leader) {
                        gb = atomicAdd(&tcnt[bcur * nseg], lb + pc - over0);
                        rb = atomicAdd((unsigned long long*)&ctl->rec_cnt, (unsigned long long)(lb + pc - over0));
                    }
                    gb = __shfl(gb, (int)leader);
                    rb = __shfl(rb, (int)leader);
                    const uint32_t o = lb + rank_below(m) - over0;
                    if (fired && lb + rank_below(m) >= cap) {
                        if (PW) st_agent(&wbuf[gb + o], (uint32_t)i | ((uint32_t)w << 26));
                        else wbuf[gb + o] = (uint32_t)i | ((uint32_t)w << 26);
                        if (a.rec) {
                            const uint64_t at = rb + o;
                            if (at < a.rec_cap) a.rec[at] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)i;
                            else ctl->rec_overflow = 1;
                        }
                    }
                }
                if (fired) {
                    const uint32_t r = lb + rank_below(m);
                    if (r < cap) lsp[r] = (uint32_t)i | ((uint32_t)w << 26);
                    se_local += (unsigned long long)outdeg;
                }
            };
            if (tile_lean) {
                // every neuron an LIFB neuron of a cached class (tile_lean): the steps
                // (mgcore.cu:164-196) as selects, the next step's ring word loaded ahead,
                // spikes kept as a bit per step; the consumed rows are cleared and the
                // spikes appended after the steps
                const float* cc = &ccache[csl < 0 ? 0 : csl][0][0];
                int32_t slot = slot0;
                uint32_t x = active ? ring[slot * tn + j] : 0u;
                uint64_t fmask = 0ull;
                for (int32_t w = 0; w < nsteps; ++w) {
                    const uint32_t ke = x & 0xFFFFu, ki = x >> 16;
                    float be = cc[min(ke, (uint32_t)kClassCache - 1u)];
                    float bi = cc[kClassCache + min(ki, (uint32_t)kClassCache - 1u)];
                    if (ke >= (uint32_t)kClassCache) be = a.seq[st.off_e + ke];
                    if (ki >= (uint32_t)kClassCache) bi = a.seq[st.off_i + ki];
                    slot = (slot + 1 == D) ? 0 : slot + 1;
                    x = (active && w + 1 < nsteps) ? ring[slot * tn + j] : 0u;
                    const bool ref = st.refr > 0;
                    const bool fire = active && !ref && st.V >= st.c4;
                    const float vn = st.V + fmaf(-0.005f, st.V, be + bi);
                    st.V = fire ? st.c3 : (ref ? st.V : vn);
                    st.refr = fire ? st.refr_steps : (ref ? st.refr - 1 : st.refr);
                    fmask |= (uint64_t)fire << w;
                }
                if (fmask) {
                    st.cnt += __popcll(fmask);
                    st.last = (int32_t)(s0 + 63 - __clzll((long long)fmask));
                }
                for (uint64_t fm = fmask; __any(fm != 0ull);) {   // one step's spikes per lane and pass
                    const bool f = fm != 0ull;
                    const int32_t w = f ? __ffsll((unsigned long long)fm) - 1 : 0;
                    fm &= fm - 1ull;
                    append(f, w, (int32_t)(s0 + w));
                }
            } else {
                const bool lean = kCls && __all(!active || (type == T_LIFB && csl >= 0));
                int32_t slot = slot0;
                for (int32_t w = 0; w < nsteps; ++w) {
                    const int32_t s = (int32_t)(s0 + w);
                    bool fired = false;
                    // class sums of a cached class: one LDS read per side (the table for
                    // counts past the cache)
                    auto cached_input = [&](uint32_t x, float& be, float& bi) {
                        const uint32_t ke = x & 0xFFFFu, ki = x >> 16;
                        const float* cc = &ccache[csl][0][0];
                        be = cc[min(ke, (uint32_t)kClassCache - 1u)];
                        bi = cc[kClassCache + min(ki, (uint32_t)kClassCache - 1u)];
                        if (ke >= (uint32_t)kClassCache) be = a.seq[st.off_e + ke];
                        if (ki >= (uint32_t)kClassCache) bi = a.seq[st.off_i + ki];
                    };
                    if (lean) {
                        if (active) {
                            uint32_t* r = &ring[slot * tn + j];
                            const uint32_t x = *r;
                            *r = 0u;                                          // clearInput
                            float be, bi;
                            cached_input(x, be, bi);
                            fired = integrate_lifb(st, s, be + bi);
                        }
                    } else if (type == T_POISSON) {
                        fired = poisson_step(st, s, w);
                    } else if (integ) {
                        float be, bi;
                        uint32_t* r = &ring[slot * tn + j];
                        if (kCls) {
                            const uint32_t x = *r;
                            *r = 0u;                                          // clearInput
                            if (csl >= 0) cached_input(x, be, bi);
                            else class_input(x, a.seq, st, be, bi);
                        } else {
                            uint32_t* ri = r + (size_t)D * tn;
                            const int32_t qe = (int32_t)*r, qi = RW == 2 ? (int32_t)*ri : 0;
                            *r = 0u;
                            if (RW == 2) *ri = 0u;
                            fixed_input(type, qe, qi, fsc.y, be, bi);
                        }
                        fired = integrate_step(type, st, s, be, bi);
                    }
                    if (PW && active && !integ) {
                        // the slice stays in LDS across windows: the consumed row of a neuron
                        // that does not read its input is cleared here (a per-window launch
                        // reloads consumed rows as zeros)
                        ring[slot * tn + j] = 0u;
                        if (!kCls && RW == 2) ring[((size_t)D + slot) * tn + j] = 0u;
                    }
                    if (kPl && fired) {
                        // the history moves down one entry (the oldest drops out), in LDS and HBM
                        int4* hq = reinterpret_cast<int4*>(hL + (size_t)j * kStdpHist);
                        const int4 q0 = hq[0], q1 = hq[1], q2 = hq[2], q3 = hq[3];
                        const int4 n0 = make_int4(s, q0.x, q0.y, q0.z), n1 = make_int4(q0.w, q1.x, q1.y, q1.z);
                        const int4 n2 = make_int4(q1.w, q2.x, q2.y, q2.z), n3 = make_int4(q2.w, q3.x, q3.y, q3.z);
                        hq[0] = n0; hq[1] = n1; hq[2] = n2; hq[3] = n3;
                        int4* hg = reinterpret_cast<int4*>(a.hist + (size_t)i * kStdpHist);
                        hg[0] = n0; hg[1] = n1; hg[2] = n2; hg[3] = n3;
                    }
                    append(fired, w, s);
                    slot = (slot + 1 == D) ? 0 : slot + 1;
                }
            }
            if (active) neuron_store(a, i, type, st, lastC);
        }
        if (se_local) atomicAdd(&se_blk, se_local);
        __syncthreads();
        if (PW && tile_lean && kw + 1 < nw) {
            // the lean steps leave the rows they consumed for here, after every wave's
            // last read; the slice stays in LDS for the next window
            const int32_t nr = min(nsteps, D);
            int32_t sl = slot0;
            for (int32_t r = 0; r < nr; ++r) {
                for (int32_t jj = tid; jj < tn; jj += bd) ring[sl * tn + jj] = 0u;
                sl = (sl + 1 == D) ? 0 : sl + 1;
            }
        }
        const uint32_t nl = min(lsp_n, (uint32_t)kTiledBatch);
        if (tid == 0) {
            if (rgp) {
                // region window: no list space to reserve; the counters wait for the
                // launch's last window (records still take their places now)
                se_acc += se_blk;
                if (!a.rec) spk_acc += lsp_n;
                else if (nl) lsp_r = atomicAdd((unsigned long long*)&ctl->rec_cnt, (unsigned long long)nl);
            } else {
                if (se_blk + se_acc) atomicAdd((unsigned long long*)&ctl->se, se_blk + se_acc);
                if (spk_acc) atomicAdd((unsigned long long*)&ctl->rec_cnt, spk_acc);
                se_acc = 0ull;
                spk_acc = 0ull;
                if (nl) {
                    lsp_g = atomicAdd(&tcnt[bcur * nseg], nl);
                    lsp_r = atomicAdd((unsigned long long*)&ctl->rec_cnt, (unsigned long long)nl);
                }
            }
        }
        // 4. LDS -> ring slice, except the slots this window consumed (cleared above,
        // or never read for neurons that do not integrate); meanwhile the list's
        // offset comes back from the atomic above
        if (!PW || kw + 1 == nw) ring_slice_out(ring, gring, D, tn, (int32_t)(s0 % D), nsteps, RW);
        __syncthreads();
        for (uint32_t k = (uint32_t)tid; k < ((rgp && !a.rec) ? 0u : nl); k += (uint32_t)bd) {
            const uint32_t ent = lsp[k];
            if (rgp) {
                // (the entry is in the region already)
            } else if (PW) {
                st_agent(&wbuf[lsp_g + k], ent);
            } else {
                wbuf[lsp_g + k] = ent;
            }
            if (a.rec) {
                const uint64_t at = lsp_r + k;
                if (at < a.rec_cap)
                    a.rec[at] = ((uint64_t)(uint32_t)(s0 + (int64_t)(ent >> 26)) << 32) | (ent & kIdMask);
                else ctl->rec_overflow = 1;
            }
        }
        if constexpr (PW) {
            if (kw + 1 < nw &&
                !grid_wait(a.bar, (uint32_t)a.G, (uint32_t)g, (uint32_t)(kwin + 1), rgp ? lsp_n : 0u,
                           rgn_on ? rg_pre : nullptr, gt0, gt1, a.bar_err, a.spin_limit))
                return;
        }
    }   // windows
}

// Flush of deferred dendrite updates (fused STDP): for each neuron j the fused
// kernel listed in window k, every plastic in-synapse k -> j replays j's spikes
// in (max(k's last spike, old tf[j]), end of window k-1], oldest first, with
// dt = k's last spike - t1 (k's last spike 0 counts as "never", as in the fused
// kernel).  The list is staged in LDS 256 entries at a time; their in-synapses
// are cut into units of one per thread, dealt round-robin over the grid.
__global__ __launch_bounds__(256) void k_stdp_flush(TiledArgs a, int32_t off, int32_t wk) {
    __shared__ int32_t ej[256], etf[256];
    __shared__ int64_t epb[256], epe[256];
    __shared__ uint32_t eui[256], wsum[16];                // inclusive prefix of units
    const Ctl* ctl = a.ctl;
    const uint32_t tid0 = threadIdx.x;
    // both lists' counts and first entries load with the window index (one round
    // trip instead of three); the index's parity then picks one
    const uint32_t nf0 = a.fcnt[0], nf1 = a.fcnt[1];
    const uint32_t i0 = min(tid0, a.flist_cap - 1u);
    const int32_t* fl1 = a.flist + (size_t)a.flist_cap * 2;
    const int32_t j0 = a.flist[2 * i0], tf0 = a.flist[2 * i0 + 1];
    const int32_t j1 = fl1[2 * i0], tf1 = fl1[2 * i0 + 1];
    const int64_t kwin = ctl->base_window + wk;
    const bool par = (kwin & 1) != 0;
    const uint32_t nf = par ? nf1 : nf0;
    const int32_t* fli = par ? fl1 : a.flist;
    const int32_t hz = (int32_t)(ctl->base_step + off - 1);
    const int32_t* lastB = a.last3 + (size_t)(kwin % 3) * a.last_stride;   // end of window k-1
    const uint32_t bd = blockDim.x, G = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
    uint64_t ubase = 0;                                     // units of the chunks before
    for (uint32_t f0 = 0; f0 < nf; f0 += bd) {
        const uint32_t nc = min(bd, nf - f0);
        uint32_t nu = 0;
        if (tid < nc) {
            const bool first = f0 == 0;
            const int32_t j = first ? (par ? j1 : j0) : fli[2 * (f0 + tid)];
            ej[tid] = j;
            etf[tid] = first ? (par ? tf1 : tf0) : fli[2 * (f0 + tid) + 1];
            const int64_t pb = a.din_ptr[j], pe = a.din_ptr[j + 1];
            epb[tid] = pb;
            epe[tid] = pe;
            nu = (uint32_t)((pe - pb + bd - 1) / bd);
        }
        const uint32_t incl = block_incl_scan(nu, wsum);
        eui[tid] = incl;
        const uint32_t total = wsum[(bd >> 6) - 1];
        __syncthreads();
        for (uint32_t u = (uint32_t)((b + G - (uint32_t)(ubase % G)) % G); u < total; u += G) {
            uint32_t lo_f = 0, hi_f = nc - 1;               // entry f: eui[f-1] <= u < eui[f]
            while (lo_f < hi_f) {
                const uint32_t mid = (lo_f + hi_f) >> 1;
                if (eui[mid] <= u) lo_f = mid + 1;
                else hi_f = mid;
            }
            const uint32_t f = lo_f;
            const uint32_t uf = u - (f ? eui[f - 1] : 0u);
            const int64_t p = epb[f] + (int64_t)uf * bd + tid;
            if (p >= epe[f]) continue;
            const int32_t j = ej[f];
            const uint32_t kk = a.din_pos[p];
            float w = a.w_t[kk];                            // (with lk: one round trip)
            const int32_t lk = lastB[a.din_src[p]];
            const int32_t lo = max(lk == 0 ? -1 : lk, etf[f]);
            // j's history newest first (kHistNone past its spikes; a spike of window k, after
            // hz, leads it): the newest nr entries are after lo; their table values are
            // loaded together, then applied oldest first
            const int4* hj = reinterpret_cast<const int4*>(a.hist + (size_t)j * kStdpHist);
            int32_t hv[kStdpHist];
#pragma unroll
            for (int b4 = 0; b4 < kStdpHist / 4; ++b4) {
                const int4 q = hj[b4];
                hv[4 * b4] = q.x;
                hv[4 * b4 + 1] = q.y;
                hv[4 * b4 + 2] = q.z;
                hv[4 * b4 + 3] = q.w;
            }
            int32_t nr = 0;
#pragma unroll
            for (int m = 0; m < kStdpHist; ++m) nr += (hv[m] > lo && m == nr) ? 1 : 0;
            if (nr == 0) continue;
            float dv[kStdpHist];
#pragma unroll
            for (int m = 0; m < kStdpHist; ++m)
                dv[m] = (m < nr && hv[m] <= hz) ? stdp_dw(a.ltp, a.ltd, a.lut_n, lk - hv[m]) : 0.0f;
#pragma unroll
            for (int m = kStdpHist - 1; m >= 0; --m)
                if (m < nr && hv[m] <= hz) w = clipw(w + dv[m], a.w_max, a.w_min);
            a.w_t[kk] = w;
        }
        ubase += total;
        __syncthreads();                                    // the staging arrays are reused
    }
}

// End of a run: the step and window counters move on, and the persistent waits'
// error word goes to the host (pinned memory the host reads after the run).
__global__ void k_finalize_tiled(Ctl* ctl, int32_t steps, int32_t windows, const uint32_t* err, uint32_t* h_err) {
    ctl->base_step += steps;
    ctl->base_window += windows;
    if (err && h_err) __hip_atomic_store(h_err, *err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

hipError_t launch_neurons(int mode, const NeuArgs& a, int32_t off, int32_t wk, int32_t nsteps,
                          hipStream_t st) {
    const int grid = (a.npad + kNeuBlock - 1) / kNeuBlock;
    if (mode == kRingClass)
        hipLaunchKernelGGL(k_neurons<kRingClass>, dim3(grid), dim3(kNeuBlock), 0, st, a, off, wk,
                           nsteps);
    else
        hipLaunchKernelGGL(k_neurons<kRingFixed>, dim3(grid), dim3(kNeuBlock), 0, st, a, off, wk,
                           nsteps);
    return hipGetLastError();
}

hipError_t launch_prop(int mode, int wide_kind, const PropArgs& a, int32_t off, int32_t wk,
                       int grid, int block, hipStream_t st) {
    if (mode == kRingClass)
        hipLaunchKernelGGL((k_prop<kRingClass, kWideF32>), dim3(grid), dim3(block), 0, st, a, off,
                           wk);
    else if (wide_kind == kWideF32)
        hipLaunchKernelGGL((k_prop<kRingFixed, kWideF32>), dim3(grid), dim3(block), 0, st, a, off,
                           wk);
    else if (wide_kind == kWideBF16)
        hipLaunchKernelGGL((k_prop<kRingFixed, kWideBF16>), dim3(grid), dim3(block), 0, st, a, off,
                           wk);
    else
        hipLaunchKernelGGL((k_prop<kRingFixed, kWideFP8>), dim3(grid), dim3(block), 0, st, a, off,
                           wk);
    return hipGetLastError();
}

hipError_t launch_stdp(const StdpArgs& a, int32_t off, int32_t wk, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_stdp, dim3(grid), dim3(256), 0, st, a, off, wk);
    return hipGetLastError();
}

hipError_t launch_finalize(Ctl* ctl, int32_t steps, int32_t windows, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1), 0, st, ctl, steps, windows);
    return hipGetLastError();
}

size_t tiled_lds_bytes(int D, int B, int kind, int rw, int lut_lds) {
    const bool pl = stdp_kind(kind);
    size_t rwords = (size_t)rw * (size_t)D * (size_t)B;
    if (pl) rwords = (rwords + 3) & ~(size_t)3;             // 16-byte aligned history rows
    return sizeof(uint32_t) * (rwords + 3 * (size_t)kTiledBatch) +
           sizeof(uint16_t) * (size_t)kTiledItems +
           ((kind == kTFP8 || kind == kTBF16 || pl) ? sizeof(uint32_t) * kTiledBatch : 0) +
           (kind == kTStdpB ? sizeof(uint32_t) * kTiledBatch : 0) +
           (pl ? (kStdpHist + 2) * sizeof(int32_t) * (size_t)B + 2 * sizeof(float) * (size_t)lut_lds : 0);
}

namespace {
template <int CH, int WK, bool PW, bool RG>
const void* tiled_fn() {
    return reinterpret_cast<const void*>(&k_tiled<CH, WK, PW, RG>);
}

template <int WK, bool PW, bool RG = false>
const void* tiled_fn_chunk(int chunk) {
    switch (chunk) {
        case 1: return tiled_fn<1, WK, PW, RG>();
        case 2: return tiled_fn<2, WK, PW, RG>();
        case 4: return tiled_fn<4, WK, PW, RG>();
        case 16: return tiled_fn<16, WK, PW, RG>();
        default: return tiled_fn<8, WK, PW, RG>();
    }
}

// The fused kernel for (kind, chunk); the persistent form exists for every kind without
// STDP, and for the 16-bit CLASS words also with per-tile list regions (rg).
const void* tiled_fn_of(int kind, int chunk, bool pw, bool rg = false) {
    if (rg) return pw && kind == kTClass16 ? tiled_fn_chunk<kTClass16, true, true>(chunk) : nullptr;
    switch (kind) {
        case kTBF16: return pw ? tiled_fn_chunk<kTBF16, true>(chunk) : tiled_fn_chunk<kTBF16, false>(chunk);
        case kTFP8: return pw ? tiled_fn_chunk<kTFP8, true>(chunk) : tiled_fn_chunk<kTFP8, false>(chunk);
        case kTF32: return pw ? tiled_fn_chunk<kTF32, true>(chunk) : tiled_fn_chunk<kTF32, false>(chunk);
        case kTClass16:
            return pw ? tiled_fn_chunk<kTClass16, true>(chunk) : tiled_fn_chunk<kTClass16, false>(chunk);
        case kTStdp: return pw ? nullptr : tiled_fn_chunk<kTStdp, false>(chunk);
        case kTStdpB: return pw ? nullptr : tiled_fn_chunk<kTStdpB, false>(chunk);
        default: return pw ? tiled_fn_chunk<kTClass, true>(chunk) : tiled_fn_chunk<kTClass, false>(chunk);
    }
}

hipError_t launch_fn(const void* f, const TiledArgs& a, int32_t off, int32_t wk, int32_t nsteps, int32_t nwin,
                     int block, size_t lds, hipStream_t st) {
    if (!f) return hipErrorInvalidDeviceFunction;
    TiledArgs aa = a;
    void* args[] = {&aa, &off, &wk, &nsteps, &nwin};
    return hipLaunchKernel(f, dim3(a.G), dim3(block), args, lds, st);
}
}  // namespace

// The fused kernel's class cache of every tile (CLASS kinds): cc[g][c][side][k] =
// the k-fold sum of side (excitatory, inhibitory) of class cls[g][c], where c = 0
// is the class of the tile's last neuron and c = 1 that of its first.
__global__ void k_tile_ccache(const int32_t* tile_lo, const uint16_t* tbl, const uint32_t* tbl_off,
                              const float* seq, float* cc, uint32_t* cls) {
    const int32_t g = (int32_t)blockIdx.x, tid = (int32_t)threadIdx.x;
    const int32_t c = tid / (2 * kClassCache), side = (tid / kClassCache) & 1, k = tid % kClassCache;
    const int32_t t0 = tile_lo[g], t1 = tile_lo[g + 1];
    const uint32_t cl = tbl[c == 0 ? t1 - 1 : t0];
    cc[(size_t)g * 4 * kClassCache + tid] = seq[tbl_off[2 * cl + side] + k];
    if (side == 0 && k == 0) cls[2 * g + c] = cl;
}

// The class cache holds the classes of the tile's last and first neurons.
__global__ void k_tile_lean(const int32_t* tile_lo, const uint8_t* type, const uint16_t* tbl, uint8_t* lean) {
    const int32_t g = (int32_t)blockIdx.x, tid = (int32_t)threadIdx.x;
    const int32_t t0 = tile_lo[g], t1 = tile_lo[g + 1];
    const uint32_t cl0 = tbl[t1 - 1], cl1 = tbl[t0];
    int ok = 1;
    for (int32_t i = t0 + tid; i < t1; i += (int32_t)blockDim.x)
        ok &= type[i] == T_LIFB && (tbl[i] == cl0 || tbl[i] == cl1);
    ok = __syncthreads_and(ok);
    if (tid == 0) lean[g] = ok ? 1 : 0;
}

hipError_t launch_tile_ccache(const int32_t* tile_lo, int32_t G, const uint16_t* tbl, const uint32_t* tbl_off,
                              const float* seq, float* cc, uint32_t* cls, hipStream_t st) {
    if (G <= 0) return hipSuccess;
    if (!tile_lo || !tbl || !tbl_off || !seq || !cc || !cls) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tile_ccache, dim3(G), dim3(4 * kClassCache), 0, st, tile_lo, tbl, tbl_off, seq, cc, cls);
    return hipGetLastError();
}

hipError_t launch_tile_lean(const int32_t* tile_lo, int32_t G, const uint8_t* type, const uint16_t* tbl,
                            uint8_t* lean, hipStream_t st) {
    if (G <= 0) return hipSuccess;
    if (!tile_lo || !type || !tbl || !lean) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tile_lean, dim3(G), dim3(256), 0, st, tile_lo, type, tbl, lean);
    return hipGetLastError();
}

hipError_t launch_tiled(const TiledArgs& a, int32_t off, int32_t wk, int32_t nsteps, int block,
                        hipStream_t st) {
    const size_t lds = tiled_lds_bytes(a.D, a.B, a.kind, a.rw, stdp_kind(a.kind) ? a.lut_lds : 0);
    return launch_fn(tiled_fn_of(a.kind, a.chunk, false), a, off, wk, nsteps, 1, block, lds, st);
}

// Dynamic LDS of the persistent form: the window's, plus the list regions' counts.
size_t tiled_pw_lds_bytes(const TiledArgs& a) {
    return tiled_lds_bytes(a.D, a.B, a.kind, a.rw, 0) + (a.region ? 2 * sizeof(uint32_t) * (size_t)a.G : 0);
}

hipError_t set_tiled_pw_lds_limit(const TiledArgs& a) {
    const void* f = tiled_fn_of(a.kind, a.chunk, true, a.region != 0);
    if (!f) return hipErrorInvalidDeviceFunction;
    const size_t lds = tiled_pw_lds_bytes(a);
    return lds > 65536 ? hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) : hipSuccess;
}

hipError_t launch_tiled_persistent(const TiledArgs& a, int32_t off, int32_t wk, int32_t nsteps, int32_t nwin,
                                   int block, hipStream_t st) {
    if (stdp_kind(a.kind) || !a.bar || !a.bar_err) return hipErrorInvalidValue;
    if (a.region && (a.G > kRegionMaxG || a.G > block)) return hipErrorInvalidValue;
    const void* f = tiled_fn_of(a.kind, a.chunk, true, a.region != 0);
    if (!f) return hipErrorInvalidDeviceFunction;
    return launch_fn(f, a, off, wk, nsteps, nwin, block, tiled_pw_lds_bytes(a), st);
}

int tiled_blocks_per_cu(const TiledArgs& a, int block, bool persistent) {
    const size_t lds = persistent ? tiled_pw_lds_bytes(a)
                                  : tiled_lds_bytes(a.D, a.B, a.kind, a.rw, stdp_kind(a.kind) ? a.lut_lds : 0);
    const void* f = tiled_fn_of(a.kind, a.chunk, persistent, persistent && a.region != 0);
    if (!f) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

hipError_t set_tiled_lds_limit(int kind, size_t bytes) {
    for (int pw = 0; pw < 3; ++pw)     // per window, persistent, persistent with list regions
        for (int chunk : {1, 2, 4, 8, 16}) {
            const void* f = tiled_fn_of(kind, chunk, pw != 0, pw == 2);
            if (!f) continue;
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e == hipSuccess) continue;
            // the persistent form's own static LDS may not leave room for these tiles: it
            // is then never resident, and never launched
            if (pw) {
                (void)hipGetLastError();
                continue;
            }
            return e;
        }
    return hipSuccess;
}

hipError_t launch_stdp_flush(const TiledArgs& a, int32_t off, int32_t wk, hipStream_t st) {
    hipLaunchKernelGGL(k_stdp_flush, dim3(1024), dim3(256), 0, st, a, off, wk);
    return hipGetLastError();
}

hipError_t launch_finalize_tiled(Ctl* ctl, int32_t steps, int32_t windows, const uint32_t* err, uint32_t* h_err,
                                 hipStream_t st) {
    hipLaunchKernelGGL(k_finalize_tiled, dim3(1), dim3(1), 0, st, ctl, steps, windows, err, h_err);
    return hipGetLastError();
}

}  // namespace mgb

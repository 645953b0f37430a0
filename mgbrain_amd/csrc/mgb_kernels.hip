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

    float V = 0.f, Ie = 0.f, Ii = 0.f, rate = 0.f;
    int32_t refr = 0, last = 0, cnt = 0;
    float c3 = 0.f, c4 = 0.f, c12 = 0.f, c13 = 0.f, c14 = 0.f, c15 = 0.f, c16 = 0.f, c17 = 0.f;
    int32_t refr_steps = 0;
    uint32_t off_e = 0, off_i = 0;
    uint32_t key0 = 0, key1 = 0, lid = 0;
    if (active) {
        last = a.last[i];
        cnt = a.cnt[i];
        if (type == T_POISSON) {
            rate = a.rate[i];
            const uint32_t st = a.stream ? a.stream[i] : 0u;
            const uint64_t seed = a.stream_seed[st];
            key0 = (uint32_t)seed;
            key1 = (uint32_t)(seed >> 32);
            lid = (uint32_t)(i - a.stream_base[st]);
        } else if (type == T_LIF || type == T_LIFB || type == T_LIFE) {
            const CSet& C = a.csets[a.cset[i]];
            c3 = C.c[3];
            c4 = C.c[4];
            refr_steps = __float_as_int(C.c[30]);
            if (type == T_LIF) {
                c12 = C.c[12]; c13 = C.c[13]; c14 = C.c[14];
                c15 = C.c[15]; c16 = C.c[16]; c17 = C.c[17];
            }
            V = a.V[i];
            refr = a.refr[i];
            if (type != T_LIFB) {
                Ie = a.Ie[i];
                Ii = a.Ii[i];
            }
            if (MODE == kRingClass) {
                const uint32_t t = a.tbl[i];
                off_e = a.tbl_off[2 * t];
                off_i = a.tbl_off[2 * t + 1];
            }
        }
    }
    const bool integrates = type == T_LIF || type == T_LIFB || type == T_LIFE;
    int32_t slot = (int32_t)(s0 % a.D);
    U32x4 rnd = {{0u, 0u, 0u, 0u}};

    for (int32_t w = 0; w < nsteps; ++w) {
        const int32_t s = (int32_t)(s0 + w);
        bool fired = false;
        if (type == T_POISSON) {                                     // mgcore.cu:121-131
            if (w == 0 || (s & 3) == 0)
                rnd = philox4x32_10(lid, (uint32_t)s >> 2, 0u, kTagPoisson, key0, key1);
            const float u = u01_open0(pick4(rnd, (uint32_t)s & 3u));
            fired = rate > u;
            if (fired) { ++cnt; last = s; }
        } else if (integrates) {
            const size_t loc = (size_t)slot * (size_t)a.npad + (size_t)i;
            float be, bi;
            if (MODE == kRingClass) {
                const uint32_t x = a.ring[loc];
                if (x != 0u) {
                    a.ring[loc] = 0u;                                // clearInput :99-103
                    const uint32_t ce = x & 0xFFFFu, ci = x >> 16;
                    be = ce ? a.seq[off_e + ce] : 0.f;
                    bi = ci ? a.seq[off_i + ci] : 0.f;
                } else {
                    be = 0.f;
                    bi = 0.f;
                }
            } else {
                const int32_t qe = a.ring_e[loc], qi = a.ring_i[loc];
                if (qe != 0) a.ring_e[loc] = 0;
                if (qi != 0) a.ring_i[loc] = 0;
                be = (float)qe * a.inv_e;
                bi = (float)qi * a.inv_i;
            }
            if (type == T_LIFB) {                                    // :164-196
                if (refr > 0) {
                    --refr;
                } else if (V >= c4) {
                    fired = true; ++cnt; V = c3; last = s; refr = refr_steps;
                } else {
                    V = V + fmaf(-0.005f, V, be + bi);
                }
            } else if (type == T_LIFE) {                             // :197-233
                Ie = fmaf(Ie, 0.98f, be);
                Ii = fmaf(Ii, 0.99f, -bi);
                if (refr > 0) {
                    --refr;
                } else if (V >= c4) {
                    fired = true; ++cnt; V = c3; last = s; refr = refr_steps;
                } else {
                    const float t0 = -0.06f - V;
                    const float t1 = fmaf(Ie, 0.0f - V, t0);
                    const float t2 = fmaf(Ii, -0.08f - V, t1);
                    const float t3 = t2 + 0.02f;
                    V = fmaf(t3, 0.005f, V);
                }
            } else {                                                 // LIF :132-163
                if (refr > 0) {
                    --refr;
                } else {
                    V = fmaf(Ii, c13, fmaf(Ie, c12, fmaf(c17, V, c14)));
                    Ie *= c15;
                    Ii *= c16;
                    if (V >= c4) {
                        fired = true; ++cnt; V = c3; last = s; refr = refr_steps;
                    } else {
                        Ie += be;
                        Ii += bi;
                    }
                }
            }
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
    if (active) {
        a.last[i] = last;
        a.cnt[i] = cnt;
        if (integrates) {
            a.V[i] = V;
            a.refr[i] = refr;
            if (type != T_LIFB) {
                a.Ie[i] = Ie;
                a.Ii[i] = Ii;
            }
        }
    }
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
    unsigned long long se_local = 0ull;
    for (uint32_t e = wave; e < nspk; e += nwaves) {
        const uint32_t ent = wbuf[e];
        const int32_t i = (int32_t)(ent & kIdMask);
        const int32_t s = (int32_t)(s0 + (ent >> 26));
        const int64_t beg = a.row_ptr[i], end = a.row_ptr[i + 1];
        if (lane == 0) {
            se_local += (unsigned long long)(end - beg);
            if (a.rec) {
                const uint64_t r = rec_base + e;
                if (r < a.rec_cap) a.rec[r] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)i;
                else ctl->rec_overflow = 1;
            }
        }
        const int32_t sD = s % D;
        if (MODE == kRingClass) {
            int64_t k = beg + lane;
            for (; k + 192 < end; k += 256) {
                uint32_t x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(&a.syn[k + 64 * u]);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int32_t slot = sD + (int32_t)((x[u] >> 26) & 31u) + 1;
                    slot -= (slot >= D) ? D : 0;
                    atomicAdd(&a.ring[(size_t)slot * (size_t)a.npad + (x[u] & kIdMask)],
                              (x[u] >> 31) ? 0x10000u : 1u);
                }
            }
            for (; k < end; k += 64) {
                const uint32_t x = __builtin_nontemporal_load(&a.syn[k]);
                int32_t slot = sD + (int32_t)((x >> 26) & 31u) + 1;
                slot -= (slot >= D) ? D : 0;
                atomicAdd(&a.ring[(size_t)slot * (size_t)a.npad + (x & kIdMask)],
                          (x >> 31) ? 0x10000u : 1u);
            }
        } else {
            const float rs = (WK == kWideFP8) ? a.row_scale[i] : 1.0f;
            for (int64_t k = beg + lane; k < end; k += 64) {
                const int32_t tar = a.col[k];
                const uint32_t dl = a.dly[k];
                int32_t slot = sD + (int32_t)(dl & 127u);
                slot -= (slot >= D) ? D : 0;
                float w;
                if (WK == kWideF32 || (dl & 128u)) w = a.w32[k];     // plastic weights stay FP32
                else if (WK == kWideBF16) w = bf16_to_f32(a.w16[k]);
                else w = fp8e4m3_to_f32(a.w8[k]) * rs;
                const size_t loc = (size_t)slot * (size_t)a.npad + (size_t)tar;
                if (w > 0.f) atomicAdd(&a.ring_e[loc], (int32_t)rintf(w * a.scale_e));
                else atomicAdd(&a.ring_i[loc], (int32_t)rintf(w * a.scale_i));
            }
        }
    }
    if (lane == 0 && se_local) atomicAdd(&se_blk, se_local);
    __syncthreads();
    if (threadIdx.x == 0 && se_blk) atomicAdd((unsigned long long*)&ctl->se, se_blk);
}

__device__ __forceinline__ float stdp_lut(const StdpArgs& a, int32_t dt) {
    if (dt < 0) return a.ltp[min(-dt, a.lut_n - 1)];
    if (dt > 0) return a.ltd[min(dt, a.lut_n - 1)];
    return 0.0f;
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
    for (uint32_t e = wave; e < nspk; e += nwaves) {
        const int32_t i = (int32_t)(wbuf[e] & kIdMask);
        const int32_t li = a.last[i];
        const int64_t beg = a.row_ptr[i], end = a.row_ptr[i + 1];
        for (int64_t k = beg + lane; k < end; k += 64) {             // axon loop :372-391
            if (!(a.dly[k] & 128u)) continue;
            const int32_t dt = li - a.last[a.col[k]];
            a.w32[k] = clipw(a.w32[k] + stdp_lut(a, dt), a.w_max, a.w_min);
        }
        const int64_t pb = a.in_ptr[i], pe = a.in_ptr[i + 1];
        for (int64_t p = pb + lane; p < pe; p += 64) {               // dendrite loop :392-411
            const uint32_t k = a.in_pos[p];
            const int32_t dt = a.last[a.in_src[p]] - li;
            a.w32[k] = clipw(a.w32[k] + stdp_lut(a, dt), a.w_max, a.w_min);
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

}  // namespace mgb

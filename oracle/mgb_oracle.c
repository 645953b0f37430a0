/*
 * mgb_oracle.c -- CPU restatement of MGBrain's single-GPU simulation step.
 *
 * TEST INFRASTRUCTURE ONLY (see mgb_oracle.h).  Build with -ffp-contract=off:
 * every fused multiply-add below is written as an explicit fmaf(), mirroring
 * the contraction nvcc applies by default (--fmad=true) to the reference's
 * float expressions; that choice cannot be checked against nvcc here and is
 * stated in the file that describes the design.
 *
 * Reference paths are relative to the MGBrain source tree.
 */
#include "mgb_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ========================================================================== */
/* Philox4x32-10 (Salmon et al., SC'11, section 5; constants from the paper)  */
/* ========================================================================== */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ========================================================================== */
/* MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998)                     */
/* ========================================================================== */
void orc_mt_seed(orc_mt19937* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->mti = 624;
}

uint32_t orc_mt_next(orc_mt19937* g) {
    if (g->mti >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908B0DFu;
            g->mt[i] = v;
        }
        g->mti = 0;
    }
    uint32_t y = g->mt[g->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9D2C5680u;
    y ^= (y << 15) & 0xEFC60000u;
    y ^= y >> 18;
    return y;
}

/* ========================================================================== */
/* glibc random(3), TYPE_3: r[i] = r[i-31] + r[i-3] (mod 2^32), output r>>1,   */
/* state seeded by the Park-Miller LCG 16807 (Schrage form), 310 discarded.   */
/* ========================================================================== */
void orc_glibc_srand(orc_glibc_rand* g, uint32_t seed) {
    int32_t r[34];
    int32_t word = (int32_t)(seed == 0 ? 1u : seed);
    r[0] = word;
    for (int i = 1; i < 31; ++i) {
        int32_t hi = word / 127773, lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = word;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    /* window holds r[i-31 .. i-1] for the next i = 34 */
    for (int i = 0; i < 31; ++i) g->r[i] = (uint32_t)r[i + 3];
    g->idx = 0;
    for (int i = 0; i < 310; ++i) (void)orc_glibc_rand_next(g);
}

int32_t orc_glibc_rand_next(orc_glibc_rand* g) {
    uint32_t v = g->r[g->idx] + g->r[(g->idx + 28) % 31];
    g->r[g->idx] = v;
    g->idx = (g->idx + 1) % 31;
    return (int32_t)(v >> 1);
}

/* ========================================================================== */
/* Network generation: model/netgen.cpp:19-229 (+ the two new connectors)     */
/* ========================================================================== */
#define TAG_POISSON 0x504F4953u /* "POIS" */
#define TAG_BERN    0x4245524Eu /* "BERN" */
#define TAG_BERN_W  0x42455257u
#define TAG_BERN_D  0x42455244u
#define TAG_FIXK    0x4649584Bu /* "FIXK" */
#define TAG_FIXK_W  0x46495857u
#define TAG_FIXK_D  0x46495844u

typedef struct {
    orc_synlist* s;
    int64_t cap;
} synbuf;

static int synbuf_push(synbuf* b, int32_t src, int32_t tar, float w, float d, uint8_t plastic) {
    orc_synlist* s = b->s;
    if (s->n_syn == b->cap) {
        int64_t nc = b->cap ? b->cap * 2 : 1024;
        int32_t* a = (int32_t*)realloc(s->src, (size_t)nc * sizeof(int32_t));
        if (!a) return -1;
        s->src = a;
        a = (int32_t*)realloc(s->tar, (size_t)nc * sizeof(int32_t));
        if (!a) return -1;
        s->tar = a;
        float* f = (float*)realloc(s->weight, (size_t)nc * sizeof(float));
        if (!f) return -1;
        s->weight = f;
        f = (float*)realloc(s->delay, (size_t)nc * sizeof(float));
        if (!f) return -1;
        s->delay = f;
        uint8_t* u = (uint8_t*)realloc(s->plastic, (size_t)nc);
        if (!u) return -1;
        s->plastic = u;
        b->cap = nc;
    }
    s->src[s->n_syn] = src;
    s->tar[s->n_syn] = tar;
    s->weight[s->n_syn] = w;
    s->delay[s->n_syn] = d;
    s->plastic[s->n_syn] = plastic;
    s->n_syn++;
    return 0;
}

/* netgen.cpp:19-31 */
static float ref_rand_weight(orc_glibc_rand* g, float floor_, float ceil_, float dw) {
    if (ceil_ == floor_) return floor_;
    int range = (int)((ceil_ - floor_) / dw) + 1;
    int r = orc_glibc_rand_next(g) % range;
    if (floor_ < 0) return floor_ + (float)r * dw;
    return floor_ - (float)r * dw;
}

/* netgen.cpp:32-38 */
static float ref_rand_delay(orc_glibc_rand* g, float floor_, float ceil_, float dt) {
    if (ceil_ == floor_) return floor_;
    int range = (int)((ceil_ - floor_) / dt) + 1;
    int r = orc_glibc_rand_next(g) % range;
    return floor_ + (float)r * dt;
}

static uint32_t philox_pick(uint64_t seed, uint64_t f, uint32_t proj, uint32_t tag) {
    uint32_t ctr[4] = {(uint32_t)(f >> 2), (uint32_t)(f >> 34), proj, tag};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[f & 3u];
}

/* Weight / delay draws of the new connectors: the same value grid as
 * rand_weight / rand_delay, with the index taken from a Philox stream. */
static float philox_weight(uint64_t seed, uint64_t f, uint32_t proj, uint32_t tag,
                           float floor_, float ceil_, float dw) {
    if (ceil_ == floor_) return floor_;
    int range = (int)((ceil_ - floor_) / dw) + 1;
    int r = (int)(philox_pick(seed, f, proj, tag) % (uint32_t)range);
    if (floor_ < 0) return floor_ + (float)r * dw;
    return floor_ - (float)r * dw;
}

static float philox_delay(uint64_t seed, uint64_t f, uint32_t proj, uint32_t tag,
                          float floor_, float ceil_, float dt) {
    if (ceil_ == floor_) return floor_;
    int range = (int)((ceil_ - floor_) / dt) + 1;
    int r = (int)(philox_pick(seed, f, proj, tag) % (uint32_t)range);
    return floor_ + (float)r * dt;
}

/* Fixed in-degree draw k of local target n in projection j: local source,
 * weight and delay (seconds).  Target-major: draw f = n * K + k. */
static void fixk_draw(uint64_t seed, uint32_t j, int32_t sn, int32_t K, uint64_t n, uint64_t k, float wmin,
                      float wmax, float dmin, float dmax, float dw, float dt, int32_t* m, float* w, float* d) {
    const uint64_t f = n * (uint64_t)K + k;
    const uint32_t x = philox_pick(seed, f, j, TAG_FIXK);
    *m = (int32_t)(((uint64_t)x * (uint64_t)sn) >> 32);
    *w = philox_weight(seed, f, j, TAG_FIXK_W, wmin, wmax, dw);
    *d = philox_delay(seed, f, j, TAG_FIXK_D, dmin, dmax, dt);
}

void orc_fixk_targets(uint64_t philox_seed, int32_t j, int32_t sn, int32_t K, float wmin, float wmax, float dmin,
                      float dmax, float dw, float dt, float step, const int32_t* tloc, int32_t ntl, int32_t* src,
                      float* w, int32_t* dsteps) {
    for (int32_t q = 0; q < ntl; ++q)
        for (int32_t k = 0; k < K; ++k) {
            float d;
            const size_t o = (size_t)q * (size_t)K + (size_t)k;
            fixk_draw(philox_seed, (uint32_t)j, sn, K, (uint64_t)tloc[q], (uint64_t)k, wmin, wmax, dmin, dmax, dw,
                      dt, &src[o], &w[o], &d);
            dsteps[o] = (int)roundf(d / step);                 /* mgsim.cpp:49, as orc_build_csr */
        }
}

int orc_gen_net(const orc_pop* pops, int32_t n_pops, const orc_proj* pros, int32_t n_pros,
                uint32_t mt_seed, uint32_t rand_seed, uint64_t philox_seed,
                float dw, float dt, orc_synlist* out) {
    memset(out, 0, sizeof(*out));
    synbuf b = {out, 0};
    int32_t* first = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_pops + 1));
    if (!first) return -1;
    first[0] = 0;                                  /* netgen.cpp:56-68 */
    for (int p = 0; p < n_pops; ++p) first[p + 1] = first[p] + pops[p].num;

    orc_mt19937 mt;                                /* netgen.cpp:47 */
    orc_mt_seed(&mt, mt_seed);
    orc_glibc_rand gr;
    orc_glibc_srand(&gr, rand_seed);
    int rc = 0;

    for (int j = 0; j < n_pros && rc == 0; ++j) {
        const orc_proj* pr = &pros[j];
        int32_t sn = pops[pr->src].num, tn = pops[pr->tar].num;
        int32_t s0 = first[pr->src], t0 = first[pr->tar];
        uint8_t pl = (uint8_t)(pr->plastic != 0);
        if (pr->kind == ORC_CONN_REF) {
            if (pr->ctype == 1.0f) {               /* full: netgen.cpp:84-111 */
                for (int32_t m = 0; m < sn && rc == 0; ++m)
                    for (int32_t n = 0; n < tn && rc == 0; ++n) {
                        float w = ref_rand_weight(&gr, pr->wmin, pr->wmax, dw);
                        float d = ref_rand_delay(&gr, pr->dmin, pr->dmax, dt);
                        rc = synbuf_push(&b, s0 + m, t0 + n, w, d, pl);
                    }
            } else if (pr->ctype == 0.0f) {        /* one-to-one: netgen.cpp:112-126 */
                for (int32_t k = 0; k < sn && rc == 0; ++k) {
                    float w = ref_rand_weight(&gr, pr->wmin, pr->wmax, dw);
                    float d = ref_rand_delay(&gr, pr->dmin, pr->dmax, dt);
                    rc = synbuf_push(&b, s0 + k, t0 + k, w, d, pl);
                }
            } else if (pr->ctype < 1.0f && pr->ctype > 0.0f) {
                /* quota generator: netgen.cpp:164-227 (SHUFFLE_GEN off) */
                uint64_t max_num = (uint64_t)sn * (uint64_t)tn;
                uint64_t gen_num = (uint64_t)roundf((float)max_num * pr->ctype);
                double lim = (double)gen_num * 1.1;
                uint64_t count = 0;
                /* genlist is zero-initialised (std::vector<size_t>(gen_num), :172) and the
                 * emission loop walks all gen_num entries (:210), so when the 1.1p acceptance
                 * runs out of pairs before gen_num, the gen_num - count unfilled entries emit
                 * synapses from the projection's first source to its first target, each
                 * drawing its weight and delay in turn. */
                uint64_t* genlist = (uint64_t*)calloc((size_t)(gen_num ? gen_num : 1), sizeof(uint64_t));
                if (!genlist) { rc = -1; break; }
                for (uint64_t m = 0; m < (uint64_t)sn; ++m)
                    for (uint64_t n = 0; n < (uint64_t)tn; ++n)
                        if (count < gen_num && (double)((uint64_t)orc_mt_next(&mt) % max_num) < lim)
                            genlist[count++] = m * (uint64_t)tn + n;
                for (uint64_t t = 0; t < gen_num && rc == 0; ++t) {
                    int32_t m = (int32_t)(genlist[t] / (uint64_t)tn);
                    int32_t n = (int32_t)(genlist[t] % (uint64_t)tn);
                    float w = ref_rand_weight(&gr, pr->wmin, pr->wmax, dw);
                    float d = ref_rand_delay(&gr, pr->dmin, pr->dmax, dt);
                    rc = synbuf_push(&b, s0 + m, t0 + n, w, d, pl);
                }
                free(genlist);
            }
        } else if (pr->kind == ORC_CONN_REF_SHUFFLE && pr->ctype < 1.0f && pr->ctype > 0.0f) {
            /* Config::SHUFFLE_GEN quota: netgen.cpp:127-163, std::random_shuffle of
             * libstdc++ (swap element i with element rand() % (i + 1)) */
            uint64_t len = (uint64_t)sn * (uint64_t)tn;
            uint64_t gen_num = (uint64_t)roundf((float)len * pr->ctype);
            uint8_t* gl = (uint8_t*)malloc((size_t)(len ? len : 1));
            if (!gl) { rc = -1; break; }
            for (uint64_t k = 0; k < len; ++k) gl[k] = k < gen_num;
            for (uint64_t i = 1; i < len; ++i) {
                uint64_t jj = (uint64_t)((long)orc_glibc_rand_next(&gr) % (long)(i + 1));
                if (jj != i) { uint8_t t = gl[i]; gl[i] = gl[jj]; gl[jj] = t; }
            }
            for (uint64_t m = 0; m < (uint64_t)sn && rc == 0; ++m)
                for (uint64_t n = 0; n < (uint64_t)tn && rc == 0; ++n)
                    if (gl[m * (uint64_t)tn + n]) {
                        float w = ref_rand_weight(&gr, pr->wmin, pr->wmax, dw);
                        float d = ref_rand_delay(&gr, pr->dmin, pr->dmax, dt);
                        rc = synbuf_push(&b, s0 + (int32_t)m, t0 + (int32_t)n, w, d, pl);
                    }
            free(gl);
        } else if (pr->kind == ORC_CONN_BERNOULLI) {
            /* new: independent Bernoulli(p) per ordered pair, source-major */
            uint64_t thr = (uint64_t)((double)pr->p * 4294967296.0);
            for (uint64_t m = 0; m < (uint64_t)sn && rc == 0; ++m)
                for (uint64_t n = 0; n < (uint64_t)tn && rc == 0; ++n) {
                    uint64_t f = m * (uint64_t)tn + n;
                    if ((uint64_t)philox_pick(philox_seed, f, (uint32_t)j, TAG_BERN) >= thr) continue;
                    float w = philox_weight(philox_seed, f, (uint32_t)j, TAG_BERN_W, pr->wmin, pr->wmax, dw);
                    float d = philox_delay(philox_seed, f, (uint32_t)j, TAG_BERN_D, pr->dmin, pr->dmax, dt);
                    rc = synbuf_push(&b, s0 + (int32_t)m, t0 + (int32_t)n, w, d, pl);
                }
        } else if (pr->kind == ORC_CONN_FIXED_INDEGREE) {
            /* new: K sources per target drawn with replacement, target-major */
            for (uint64_t n = 0; n < (uint64_t)tn && rc == 0; ++n)
                for (uint64_t k = 0; k < (uint64_t)pr->k && rc == 0; ++k) {
                    int32_t m;
                    float w, d;
                    fixk_draw(philox_seed, (uint32_t)j, sn, pr->k, n, k, pr->wmin, pr->wmax, pr->dmin, pr->dmax, dw,
                              dt, &m, &w, &d);
                    rc = synbuf_push(&b, s0 + m, t0 + (int32_t)n, w, d, pl);
                }
        }
    }
    free(first);
    if (rc) orc_free_synlist(out);
    return rc;
}

void orc_free_synlist(orc_synlist* s) {
    free(s->src); free(s->tar); free(s->weight); free(s->delay); free(s->plastic);
    memset(s, 0, sizeof(*s));
}

/* ========================================================================== */
/* CSR build: simulator/mgsim/adj/mgsim.cpp:8-144 for npart = 1               */
/* ========================================================================== */
int orc_build_csr(int32_t n, int64_t s, const int32_t* src, const int32_t* tar,
                  const float* weight, const float* delay_s, const uint8_t* plastic,
                  float step, orc_csr* c) {
    memset(c, 0, sizeof(*c));
    c->n = n; c->s = s;
    c->row_ptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    c->in_ptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    size_t ss = (size_t)(s ? s : 1);
    c->col = (int32_t*)malloc(ss * sizeof(int32_t));
    c->dly = (int32_t*)malloc(ss * sizeof(int32_t));
    c->w = (float*)malloc(ss * sizeof(float));
    c->plastic = (uint8_t*)malloc(ss);
    c->perm = (int64_t*)malloc(ss * sizeof(int64_t));
    c->in_pos = (int64_t*)malloc(ss * sizeof(int64_t));
    int64_t* cur = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t* pos_of = (int64_t*)malloc(ss * sizeof(int64_t));
    if (!c->row_ptr || !c->in_ptr || !c->col || !c->dly || !c->w || !c->plastic || !c->perm ||
        !c->in_pos || !cur || !pos_of) {
        free(cur); free(pos_of); orc_free_csr(c); return -1;
    }
    int max_delay = 0, min_delay = INT_MAX;
    for (int64_t i = 0; i < s; ++i) {
        if (src[i] < 0 || src[i] >= n || tar[i] < 0 || tar[i] >= n) {
            free(cur); free(pos_of); orc_free_csr(c); return -2;
        }
        c->row_ptr[src[i] + 1]++;
        c->in_ptr[tar[i] + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) {
        c->row_ptr[i + 1] += c->row_ptr[i];
        c->in_ptr[i + 1] += c->in_ptr[i];
    }
    memcpy(cur, c->row_ptr, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < s; ++i) {
        int d = (int)roundf(delay_s[i] / step);           /* mgsim.cpp:49 */
        if (d > max_delay) max_delay = d;                   /* mgsim.cpp:50 */
        if (d < min_delay) min_delay = d;                   /* mgsim.cpp:51 */
        int64_t p = cur[src[i]]++;                          /* mgsim.cpp:56, 114-131 */
        c->col[p] = tar[i];
        c->dly[p] = d;
        c->w[p] = weight[i];
        c->plastic[p] = plastic ? plastic[i] : 1;
        c->perm[p] = i;
        pos_of[i] = p;
    }
    memcpy(cur, c->in_ptr, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < s; ++i) c->in_pos[cur[tar[i]]++] = pos_of[i];   /* mgsim.cpp:57 */
    c->max_delay = max_delay + 1;                           /* mgsim.cpp:71 */
    c->min_delay = min_delay;
    free(cur); free(pos_of);
    return 0;
}

int64_t orc_in_csr(int32_t n, int64_t s, const int32_t* col, const uint8_t* plastic, int64_t* in_ptr,
                   int64_t* in_pos) {
    memset(in_ptr, 0, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < s; ++k)
        if (plastic[k]) in_ptr[col[k] + 1]++;
    for (int32_t i = 0; i < n; ++i) in_ptr[i + 1] += in_ptr[i];
    int64_t* cur = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
    if (!cur) return -1;
    memcpy(cur, in_ptr, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < s; ++k)
        if (plastic[k]) in_pos[cur[col[k]]++] = k;
    free(cur);
    return in_ptr[n];
}

void orc_free_csr(orc_csr* c) {
    free(c->row_ptr); free(c->col); free(c->dly); free(c->w); free(c->plastic);
    free(c->perm); free(c->in_ptr); free(c->in_pos);
    memset(c, 0, sizeof(*c));
}

/* ========================================================================== */
/* Simulation                                                                  */
/* ========================================================================== */
enum { T_RAW = 0, T_LIF0 = 1, T_POISSON = 2, T_LIF = 3, T_LIFE = 4, T_LIFB = 5 };

struct orc_sim {
    int32_t n, D, acc, stdp, n_cset;
    int64_t s, step, se;
    uint64_t seed;
    const orc_csr* csr;
    int32_t *type, *cset, *refr_steps;
    float *rate, *lc, sc[6];
    float inv_exc, inv_inh, scale_exc, scale_inh;
    /* FP32 / fixed-point state */
    float *V, *Ie, *Ii, *bexc, *binh, *w;
    int32_t *qexc, *qinh;
    /* FP64 state */
    double *Vd, *Ied, *Iid, *bexcd, *binhd, *wd;
    int32_t *refr, *last, *cnt, *src_of;
    uint8_t* fired;
    /* orc_run_mt: every CSR row stably sorted by target (col, delay, weight) */
    int32_t *mt_col, *mt_dly;
    float* mt_w;
};

float orc_poisson_u(uint64_t seed, uint32_t id, uint32_t step) {
    uint32_t ctr[4] = {id, step >> 2, 0u, TAG_POISSON};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    uint32_t x = out[step & 3u];
    return (float)((x >> 8) + 1u) * 0x1p-24f;              /* (0, 1] like curand_uniform */
}

float orc_vinit(uint64_t seed, uint32_t local_id, float lo, float hi) {
    if (hi == lo) return lo;
    uint32_t ctr[4] = {local_id, 0u, 0u, 0x56494E49u /* "VINI" */};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    float u = (float)(out[0] >> 8) * 0x1p-24f;
    float span = hi - lo;
    float t = u * span;
    return lo + t;
}

/* mgcore.cu:346-364 with glibc expf in place of CUDA's expf */
float orc_stdp_dw(const float* sc, int32_t dt) {
    if (dt < 0) return sc[0] * expf((float)dt / sc[2]);
    if (dt > 0) return sc[1] * expf((float)(-dt) / sc[3]);
    return 0.0f;
}

static inline float clipw(float w, const float* sc) {     /* mgcore.cu:362 */
    return (w > sc[4]) ? sc[4] : ((w < sc[5]) ? sc[5] : w);
}

static inline double clipwd(double w, const float* sc) {
    return (w > (double)sc[4]) ? (double)sc[4] : ((w < (double)sc[5]) ? (double)sc[5] : w);
}

static double stdp_dw_d(const float* sc, int32_t dt) {
    if (dt < 0) return (double)sc[0] * exp((double)dt / (double)sc[2]);
    if (dt > 0) return (double)sc[1] * exp(-(double)dt / (double)sc[3]);
    return 0.0;
}

orc_sim* orc_create(const orc_desc* d) {
    orc_sim* S = (orc_sim*)calloc(1, sizeof(orc_sim));
    if (!S) return NULL;
    const orc_csr* c = d->csr;
    int32_t n = d->n;
    S->n = n; S->D = c->max_delay; S->acc = d->acc; S->stdp = d->stdp;
    S->s = c->s; S->seed = d->seed; S->csr = c;
    S->n_cset = d->n_cset > 0 ? d->n_cset : 1;
    size_t nn = (size_t)(n ? n : 1), ring = nn * (size_t)S->D, ss = (size_t)(c->s ? c->s : 1);
    S->type = (int32_t*)malloc(nn * sizeof(int32_t));
    S->cset = (int32_t*)calloc(nn, sizeof(int32_t));
    S->rate = (float*)malloc(nn * sizeof(float));
    S->lc = (float*)malloc(sizeof(float) * 30 * (size_t)S->n_cset);
    S->refr_steps = (int32_t*)malloc(sizeof(int32_t) * (size_t)S->n_cset);
    S->refr = (int32_t*)calloc(nn, sizeof(int32_t));
    S->last = (int32_t*)calloc(nn, sizeof(int32_t));       /* mgnet.cpp:16: zero */
    S->cnt = (int32_t*)calloc(nn, sizeof(int32_t));
    S->fired = (uint8_t*)calloc(nn, 1);
    S->src_of = (int32_t*)malloc(ss * sizeof(int32_t));
    memcpy(S->type, d->type, nn * sizeof(int32_t));
    memcpy(S->rate, d->rate, nn * sizeof(float));
    if (d->cset) memcpy(S->cset, d->cset, nn * sizeof(int32_t));
    memcpy(S->lc, d->lifconst, sizeof(float) * 30 * (size_t)S->n_cset);
    memcpy(S->sc, d->stdpconst, sizeof(float) * 6);
    for (int k = 0; k < S->n_cset; ++k)                    /* mgcore.cu:153 */
        S->refr_steps[k] = (int32_t)roundf(S->lc[30 * k + 8] / S->lc[30 * k + 0]) - 1;
    for (int32_t i = 0; i < n; ++i)
        for (int64_t k = c->row_ptr[i]; k < c->row_ptr[i + 1]; ++k) S->src_of[k] = i;
    if (d->acc == ORC_ACC_F64) {
        S->Vd = (double*)malloc(nn * sizeof(double));
        S->Ied = (double*)calloc(nn, sizeof(double));
        S->Iid = (double*)calloc(nn, sizeof(double));
        S->bexcd = (double*)calloc(ring, sizeof(double));
        S->binhd = (double*)calloc(ring, sizeof(double));
        S->wd = (double*)malloc(ss * sizeof(double));
        for (int32_t i = 0; i < n; ++i) S->Vd[i] = (double)d->v0[i];
        for (int64_t k = 0; k < c->s; ++k) S->wd[k] = (double)c->w[k];
    } else {
        S->V = (float*)malloc(nn * sizeof(float));
        S->Ie = (float*)calloc(nn, sizeof(float));
        S->Ii = (float*)calloc(nn, sizeof(float));
        S->w = (float*)malloc(ss * sizeof(float));
        memcpy(S->V, d->v0, nn * sizeof(float));
        memcpy(S->w, c->w, (size_t)c->s * sizeof(float));
        if (d->acc == ORC_ACC_FIXED) {
            S->qexc = (int32_t*)calloc(ring, sizeof(int32_t));
            S->qinh = (int32_t*)calloc(ring, sizeof(int32_t));
            S->scale_exc = d->scale_exc; S->scale_inh = d->scale_inh;
            S->inv_exc = 1.0f / d->scale_exc; S->inv_inh = 1.0f / d->scale_inh;
        } else {
            S->bexc = (float*)calloc(ring, sizeof(float));
            S->binh = (float*)calloc(ring, sizeof(float));
        }
    }
    return S;
}

void orc_destroy(orc_sim* S) {
    if (!S) return;
    free(S->type); free(S->cset); free(S->rate); free(S->lc); free(S->refr_steps);
    free(S->V); free(S->Ie); free(S->Ii); free(S->bexc); free(S->binh); free(S->w);
    free(S->qexc); free(S->qinh);
    free(S->Vd); free(S->Ied); free(S->Iid); free(S->bexcd); free(S->binhd); free(S->wd);
    free(S->refr); free(S->last); free(S->cnt); free(S->fired); free(S->src_of);
    free(S->mt_col); free(S->mt_dly); free(S->mt_w);
    free(S);
}

/* ---- one step, FP32 state (ORC_ACC_F32 / ORC_ACC_FIXED) -------------------- */
/* Neuron i's update at step s (mgcore.cu:116-234): reads and clears its ring
 * slot, sets fired[i]. */
static inline void neuron_f32(orc_sim* S, int32_t i, int32_t s) {
    const int32_t D = S->D, slot = s % D;
    {
        const float* C = S->lc + 30 * S->cset[i];
        size_t loc = (size_t)i * (size_t)D + (size_t)slot;      /* mgcore.cu:96 */
        float be, bi;
        if (S->acc == ORC_ACC_FIXED) {
            if (S->type[i] == T_LIFB) {
                /* LIFB integrates be + bi only: one fixed-point sum, converted once
                 * (the two scales are equal in this mode) */
                be = (float)(S->qexc[loc] + S->qinh[loc]) * S->inv_exc;
                bi = 0.0f;
            } else {
                be = (float)S->qexc[loc] * S->inv_exc;
                bi = (float)S->qinh[loc] * S->inv_inh;
            }
        } else {
            be = S->bexc[loc];
            bi = S->binh[loc];
        }
        int t = S->type[i];
        int clear = 0;
        if (t == T_POISSON) {                                    /* :121-131 */
            float u = orc_poisson_u(S->seed, (uint32_t)i, (uint32_t)s);
            uint8_t f = S->rate[i] > u;
            S->fired[i] = f;
            if (f) { S->cnt[i]++; S->last[i] = s; }
        } else if (t == T_LIF) {                                 /* :132-163 */
            S->fired[i] = 0;
            if (S->refr[i] > 0) {
                --S->refr[i];
            } else {
                float V = fmaf(S->Ii[i], C[13], fmaf(S->Ie[i], C[12], fmaf(C[17], S->V[i], C[14])));
                S->Ie[i] *= C[15];
                S->Ii[i] *= C[16];
                if (V >= C[4]) {
                    S->fired[i] = 1; S->cnt[i]++; V = C[3]; S->last[i] = s;
                    S->refr[i] = S->refr_steps[S->cset[i]];
                } else {
                    S->Ie[i] += be;
                    S->Ii[i] += bi;
                }
                S->V[i] = V;
            }
            clear = 1;
        } else if (t == T_LIFB) {                                /* :164-196 */
            S->fired[i] = 0;
            if (S->refr[i] > 0) {
                --S->refr[i];
            } else if (S->V[i] >= C[4]) {
                S->fired[i] = 1; S->cnt[i]++; S->V[i] = C[3]; S->last[i] = s;
                S->refr[i] = S->refr_steps[S->cset[i]];
            } else {
                float V = S->V[i];
                S->V[i] = V + fmaf(-0.005f, V, be + bi);
            }
            clear = 1;
        } else if (t == T_LIFE) {                                /* :197-233 */
            S->fired[i] = 0;
            S->Ie[i] = fmaf(S->Ie[i], 0.98f, be);
            S->Ii[i] = fmaf(S->Ii[i], 0.99f, -bi);
            if (S->refr[i] > 0) {
                --S->refr[i];
            } else if (S->V[i] >= C[4]) {
                S->fired[i] = 1; S->cnt[i]++; S->V[i] = C[3]; S->last[i] = s;
                S->refr[i] = S->refr_steps[S->cset[i]];
            } else {
                float V = S->V[i];
                float a = -0.06f - V;
                float b = fmaf(S->Ie[i], 0.0f - V, a);
                float e = fmaf(S->Ii[i], -0.08f - V, b);
                float g = e + 0.02f;
                S->V[i] = fmaf(g, 0.005f, V);
            }
            clear = 1;
        }
        if (clear) {                                             /* clearInput :99-103 */
            if (S->acc == ORC_ACC_FIXED) { S->qexc[loc] = 0; S->qinh[loc] = 0; }
            else { S->bexc[loc] = 0.0f; S->binh[loc] = 0.0f; }
        }
    }
}

static void step_f32(orc_sim* S, int32_t s) {
    const int32_t D = S->D, n = S->n;
    const orc_csr* c = S->csr;
    /* neuron phase: mgcore.cu:116-234 */
    for (int32_t i = 0; i < n; ++i) neuron_f32(S, i, s);
    /* propagation phase: mgcore.cu:271-305, ascending source id, CSR order */
    for (int32_t i = 0; i < n; ++i) {
        if (!S->fired[i]) continue;
        S->se += c->row_ptr[i + 1] - c->row_ptr[i];
        for (int64_t k = c->row_ptr[i]; k < c->row_ptr[i + 1]; ++k) {
            int32_t tar = c->col[k];
            float w = S->w[k];
            size_t loc = (size_t)tar * (size_t)D + (size_t)((s + c->dly[k] + D) % D);   /* :114 */
            if (S->acc == ORC_ACC_FIXED) {
                if (w > 0) S->qexc[loc] += (int32_t)rintf(w * S->scale_exc);
                else S->qinh[loc] += (int32_t)rintf(w * S->scale_inh);
            } else {
                if (w > 0) S->bexc[loc] += w;
                else S->binh[loc] += w;
            }
        }
    }
    /* STDP phase: mgcore.cu:365-412 */
    if (S->stdp) {
        for (int32_t i = 0; i < n; ++i) {
            if (!S->fired[i]) continue;
            for (int64_t k = c->row_ptr[i]; k < c->row_ptr[i + 1]; ++k) {
                if (!c->plastic[k]) continue;
                int32_t dt = S->last[i] - S->last[c->col[k]];
                S->w[k] = clipw(S->w[k] + orc_stdp_dw(S->sc, dt), S->sc);
            }
            for (int64_t p = c->in_ptr[i]; p < c->in_ptr[i + 1]; ++p) {
                int64_t k = c->in_pos[p];
                if (!c->plastic[k]) continue;
                int32_t dt = S->last[S->src_of[k]] - S->last[i];
                S->w[k] = clipw(S->w[k] + orc_stdp_dw(S->sc, dt), S->sc);
            }
        }
    }
}

/* ---- one step, FP64 (accuracy oracle; plain double arithmetic) -------------- */
/* Neuron i's update at step s in FP64 (accuracy oracle): reads and clears its
 * ring slot, sets fired[i]. */
static inline void neuron_f64(orc_sim* S, int32_t i, int32_t s) {
    const int32_t D = S->D, slot = s % D;
    const double k005 = (double)0.005f, k098 = (double)0.98f, k099 = (double)0.99f;
    const double e_l = (double)-0.06f, e_i = (double)-0.08f, i_off = (double)0.02f;
    const float* C = S->lc + 30 * S->cset[i];
    size_t loc = (size_t)i * (size_t)D + (size_t)slot;
    double be = S->bexcd[loc], bi = S->binhd[loc];
    int t = S->type[i];
    int clear = 0;
    if (t == T_POISSON) {
        float u = orc_poisson_u(S->seed, (uint32_t)i, (uint32_t)s);
        uint8_t f = S->rate[i] > u;
        S->fired[i] = f;
        if (f) { S->cnt[i]++; S->last[i] = s; }
    } else if (t == T_LIF) {
        S->fired[i] = 0;
        if (S->refr[i] > 0) {
            --S->refr[i];
        } else {
            double V = (double)C[17] * S->Vd[i] + (double)C[14] + S->Ied[i] * (double)C[12] +
                       S->Iid[i] * (double)C[13];
            S->Ied[i] *= (double)C[15];
            S->Iid[i] *= (double)C[16];
            if (V >= (double)C[4]) {
                S->fired[i] = 1; S->cnt[i]++; V = (double)C[3]; S->last[i] = s;
                S->refr[i] = S->refr_steps[S->cset[i]];
            } else {
                S->Ied[i] += be;
                S->Iid[i] += bi;
            }
            S->Vd[i] = V;
        }
        clear = 1;
    } else if (t == T_LIFB) {
        S->fired[i] = 0;
        if (S->refr[i] > 0) {
            --S->refr[i];
        } else if (S->Vd[i] >= (double)C[4]) {
            S->fired[i] = 1; S->cnt[i]++; S->Vd[i] = (double)C[3]; S->last[i] = s;
            S->refr[i] = S->refr_steps[S->cset[i]];
        } else {
            double V = S->Vd[i];
            S->Vd[i] = V + ((be + bi) - k005 * V);
        }
        clear = 1;
    } else if (t == T_LIFE) {
        S->fired[i] = 0;
        S->Ied[i] = be + S->Ied[i] * k098;
        S->Iid[i] = -bi + S->Iid[i] * k099;
        if (S->refr[i] > 0) {
            --S->refr[i];
        } else if (S->Vd[i] >= (double)C[4]) {
            S->fired[i] = 1; S->cnt[i]++; S->Vd[i] = (double)C[3]; S->last[i] = s;
            S->refr[i] = S->refr_steps[S->cset[i]];
        } else {
            double V = S->Vd[i];
            S->Vd[i] = V + ((e_l - V) + S->Ied[i] * (0.0 - V) + S->Iid[i] * (e_i - V) + i_off) * k005;
        }
        clear = 1;
    }
    if (clear) { S->bexcd[loc] = 0.0; S->binhd[loc] = 0.0; }
}

static void step_f64(orc_sim* S, int32_t s) {
    const int32_t D = S->D, n = S->n;
    const orc_csr* c = S->csr;
    for (int32_t i = 0; i < n; ++i) neuron_f64(S, i, s);
    for (int32_t i = 0; i < n; ++i) {
        if (!S->fired[i]) continue;
        S->se += c->row_ptr[i + 1] - c->row_ptr[i];
        for (int64_t k = c->row_ptr[i]; k < c->row_ptr[i + 1]; ++k) {
            size_t loc = (size_t)c->col[k] * (size_t)D + (size_t)((s + c->dly[k] + D) % D);
            double w = S->wd[k];
            if (w > 0) S->bexcd[loc] += w;
            else S->binhd[loc] += w;
        }
    }
    if (S->stdp) {
        for (int32_t i = 0; i < n; ++i) {
            if (!S->fired[i]) continue;
            for (int64_t k = c->row_ptr[i]; k < c->row_ptr[i + 1]; ++k) {
                if (!c->plastic[k]) continue;
                int32_t dt = S->last[i] - S->last[c->col[k]];
                S->wd[k] = clipwd(S->wd[k] + stdp_dw_d(S->sc, dt), S->sc);
            }
            for (int64_t p = c->in_ptr[i]; p < c->in_ptr[i + 1]; ++p) {
                int64_t k = c->in_pos[p];
                if (!c->plastic[k]) continue;
                int32_t dt = S->last[S->src_of[k]] - S->last[i];
                S->wd[k] = clipwd(S->wd[k] + stdp_dw_d(S->sc, dt), S->sc);
            }
        }
    }
}

int64_t orc_run(orc_sim* S, int32_t nsteps, int32_t* ids, int64_t cap, int64_t* step_off) {
    int64_t w = 0;
    int overflow = 0;
    for (int32_t k = 0; k < nsteps; ++k) {
        int32_t s = (int32_t)S->step;
        if (S->acc == ORC_ACC_F64) step_f64(S, s);
        else step_f32(S, s);
        if (step_off) step_off[k] = w;
        for (int32_t i = 0; i < S->n; ++i) {
            if (!S->fired[i]) continue;
            if (ids) {
                if (w < cap) ids[w] = i;
                else overflow = 1;
            }
            w++;
        }
        S->step++;
    }
    if (step_off) step_off[nsteps] = w;
    return overflow ? -1 : w;
}

/* Multithreaded step loop of a static network (FP32, fixed point or FP64; the
 * CPU timing baseline, and a faster checker for long runs).  The neuron phase
 * runs in parallel over neurons.  For propagation
 * every CSR row is kept sorted by target (a stable sort, so the synapses of one
 * (source, target) pair keep list order) and each thread owns a contiguous
 * target range, which it finds in every fired row by binary search: each target
 * receives its additions in ascending source order, as in the sequential loop,
 * so the results are the sequential oracle's, without atomics. */
static void mt_sort_rows(orc_sim* S, int32_t nthreads) {
    const orc_csr* c = S->csr;
    const size_t ss = (size_t)(c->s ? c->s : 1);
    S->mt_col = (int32_t*)malloc(ss * sizeof(int32_t));
    S->mt_dly = (int32_t*)malloc(ss * sizeof(int32_t));
    S->mt_w = (float*)malloc(ss * sizeof(float));
    int64_t maxlen = 1;
    for (int32_t i = 0; i < S->n; ++i)
        if (c->row_ptr[i + 1] - c->row_ptr[i] > maxlen) maxlen = c->row_ptr[i + 1] - c->row_ptr[i];
    int passes = 1;                                  /* 11-bit digits covering ids < n */
    while (passes < 3 && ((int64_t)1 << (11 * passes)) < (int64_t)S->n) ++passes;
#pragma omp parallel num_threads(nthreads)
    {
        uint32_t* idx = (uint32_t*)malloc((size_t)maxlen * sizeof(uint32_t));
        uint32_t* tmp = (uint32_t*)malloc((size_t)maxlen * sizeof(uint32_t));
        uint32_t cnt[2048];
#pragma omp for schedule(dynamic, 64)
        for (int32_t i = 0; i < S->n; ++i) {
            const int64_t b = c->row_ptr[i], len = c->row_ptr[i + 1] - b;
            for (int64_t j = 0; j < len; ++j) idx[j] = (uint32_t)j;
            for (int p = 0; p < passes; ++p) {       /* LSD radix sort by target: stable */
                const int sh = 11 * p;
                memset(cnt, 0, sizeof(cnt));
                for (int64_t j = 0; j < len; ++j) cnt[((uint32_t)c->col[b + idx[j]] >> sh) & 2047u]++;
                uint32_t acc = 0;
                for (int d = 0; d < 2048; ++d) {
                    const uint32_t x = cnt[d];
                    cnt[d] = acc;
                    acc += x;
                }
                for (int64_t j = 0; j < len; ++j) tmp[cnt[((uint32_t)c->col[b + idx[j]] >> sh) & 2047u]++] = idx[j];
                uint32_t* t = idx;
                idx = tmp;
                tmp = t;
            }
            for (int64_t j = 0; j < len; ++j) {
                const int64_t k = b + idx[j];
                S->mt_col[b + j] = c->col[k];
                S->mt_dly[b + j] = c->dly[k];
                S->mt_w[b + j] = c->w[k];             /* static network: the CSR weights */
            }
        }
        free(idx);
        free(tmp);
    }
}

static int64_t lower_bound32(const int32_t* a, int64_t lo, int64_t hi, int32_t v) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

int64_t orc_run_mt(orc_sim* S, int32_t nsteps, int32_t nthreads, int32_t* ids, int64_t cap, int64_t* step_off) {
    if (S->stdp || nthreads < 1) return -1;
    const int32_t n = S->n, D = S->D, acc = S->acc;
    const orc_csr* c = S->csr;
    if (!S->mt_col) mt_sort_rows(S, nthreads);
    int64_t spikes = 0, se = 0, wr = 0;
    int overflow = 0;
    for (int32_t st = 0; st < nsteps; ++st) {
        const int32_t s = (int32_t)(S->step + st);
        if (step_off) step_off[st] = wr;
#pragma omp parallel num_threads(nthreads) reduction(+ : spikes, se)
        {
#pragma omp for schedule(static, 1024)
            for (int32_t i = 0; i < n; ++i) {
                if (acc == ORC_ACC_F64) neuron_f64(S, i, s);
                else neuron_f32(S, i, s);
            }
            /* the implicit barrier of the loop above: fired[] is complete */
            const int32_t nt = omp_get_num_threads(), me = omp_get_thread_num();
            const int32_t t_lo = (int32_t)((int64_t)n * me / nt), t_hi = (int32_t)((int64_t)n * (me + 1) / nt);
            for (int32_t i = 0; i < n; ++i) {
                if (!S->fired[i]) continue;
                if (me == 0) {
                    ++spikes;
                    se += c->row_ptr[i + 1] - c->row_ptr[i];
                    if (ids) {                        /* ascending id within the step */
                        if (wr < cap) ids[wr] = i;
                        else overflow = 1;
                    }
                    wr++;
                }
                const int64_t b = lower_bound32(S->mt_col, c->row_ptr[i], c->row_ptr[i + 1], t_lo);
                const int64_t e = lower_bound32(S->mt_col, b, c->row_ptr[i + 1], t_hi);
                for (int64_t k = b; k < e; ++k) {
                    const float w = S->mt_w[k];
                    const size_t loc = (size_t)S->mt_col[k] * (size_t)D + (size_t)((s + S->mt_dly[k] + D) % D);
                    if (acc == ORC_ACC_F64) {
                        if (w > 0) S->bexcd[loc] += (double)w;
                        else S->binhd[loc] += (double)w;
                    } else if (acc == ORC_ACC_FIXED) {
                        if (w > 0) S->qexc[loc] += (int32_t)rintf(w * S->scale_exc);
                        else S->qinh[loc] += (int32_t)rintf(w * S->scale_inh);
                    } else {
                        if (w > 0) S->bexc[loc] += w;
                        else S->binh[loc] += w;
                    }
                }
            }
        }
    }
    if (step_off) step_off[nsteps] = wr;
    S->step += nsteps;
    S->se += se;
    return overflow ? -2 : spikes;
}

void orc_get_state(const orc_sim* S, int field, double* out) {
    for (int32_t i = 0; i < S->n; ++i) {
        double v = 0.0;
        int f64 = S->acc == ORC_ACC_F64;
        switch (field) {
            case ORC_F_V: v = f64 ? S->Vd[i] : (double)S->V[i]; break;
            case ORC_F_IEXC: v = f64 ? S->Ied[i] : (double)S->Ie[i]; break;
            case ORC_F_IINH: v = f64 ? S->Iid[i] : (double)S->Ii[i]; break;
            case ORC_F_REFR: v = (double)S->refr[i]; break;
            case ORC_F_LAST: v = (double)S->last[i]; break;
            case ORC_F_CNT: v = (double)S->cnt[i]; break;
            default: break;
        }
        out[i] = v;
    }
}

void orc_get_weights(const orc_sim* S, double* out) {
    for (int64_t k = 0; k < S->s; ++k) out[k] = S->acc == ORC_ACC_F64 ? S->wd[k] : (double)S->w[k];
}

int64_t orc_syn_events(const orc_sim* S) { return S->se; }

/*
 * mgb_oracle.h -- CPU restatement of MGBrain's single-GPU simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path, never part
 * of it: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Each function cites the reference file:line it restates
 * (paths relative to the MGBrain source tree).
 *
 * Pinning (what fixes each function independently of the HIP build):
 *   orc_philox4x32_10   Random123 known-answer vectors (tests/test_oracle_rng.py)
 *   orc_mt_*            C++ standard KAT (10000th output of seed 5489) + numpy's
 *                       legacy MT19937 seeding
 *   orc_glibc_*         glibc rand()/srand() called from the test
 *   orc_gen_net         reference NetGen::gen_net compiled from its own sources
 *                       (oracle/_ref/ref_netgen) on fixtures and live specs, incl.
 *                       quota projections that run short of gen_num; closed-form
 *                       counts (every quota projection emits exactly gen_num)
 *   orc_build_csr       brute-force invariants + delay known answers
 *   orc_sim_* LIFB      closed forms (free decay, single-synapse jump, refractory
 *                       ISI, seq-sum of equal addends), brute force on tiny nets
 *   orc_sim_* LIFE      closed forms in F64 (~1e-12) and F32 (few ulps): free
 *                       relaxation to E_L + I_off and its first spike step,
 *                       conductance decay 0.98^k / 0.99^k (also while
 *                       refractory), the steady-drive fixed point
 *                       V = (E_L + I_off + E_inh I_i) / (1 + I_e + I_i)
 *   orc_sim_* LIF       closed forms in F64 and F32: free decay
 *                       C17^n V0 + C14 (1 - C17^n) / (1 - C17), the response to
 *                       one arrival C12 w sum C17^(k-1-j) C15^j (C13, C16 for w < 0)
 *   STDP                two-neuron scripted pairs against A*expf(-k/tau)
 */
#ifndef MGB_ORACLE_H
#define MGB_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random number generators ------------------------------------------ */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* MT19937, Matsumoto & Nishimura 1998 (std::mt19937). */
typedef struct { uint32_t mt[624]; int mti; } orc_mt19937;
void orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);

/* glibc random(3) TYPE_3 additive-feedback generator (what rand() returns). */
typedef struct { uint32_t r[31]; int idx; } orc_glibc_rand;
void orc_glibc_srand(orc_glibc_rand* g, uint32_t seed);
int32_t orc_glibc_rand_next(orc_glibc_rand* g);

/* ---- network generation (model/netgen.cpp restated) --------------------- */
enum { ORC_CONN_REF = 0, ORC_CONN_BERNOULLI = 1, ORC_CONN_FIXED_INDEGREE = 2, ORC_CONN_REF_SHUFFLE = 3 };

typedef struct {
    int32_t num;
    int32_t type;        /* NeuronType value (macro.h:40-48) */
    int32_t source;
} orc_pop;

typedef struct {
    int32_t src, tar;    /* population indices */
    float wmin, wmax;    /* Projection::wrange (model.h:32) */
    float dmin, dmax;    /* Projection::drange, seconds */
    float ctype;         /* Projection::ctype: 0 one-to-one, 1 full, (0,1) quota */
    int32_t kind;        /* ORC_CONN_* */
    float p;             /* Bernoulli probability */
    int32_t k;           /* fixed in-degree */
    int32_t plastic;     /* per-projection plasticity flag */
} orc_proj;

typedef struct {
    int64_t n_syn;
    int32_t* src;
    int32_t* tar;
    float* weight;
    float* delay;        /* seconds */
    uint8_t* plastic;
} orc_synlist;

/* Generates the synapse list in the reference's generation order.
 * Reference connectors draw from std::mt19937(mt_seed) and glibc rand()
 * seeded with rand_seed (the reference calls neither srand nor a seeded mt;
 * both are exposed so a harness can pin them).  New connectors use Philox
 * keyed by philox_seed.  dw/dt are Config::DW / Config::DT. */
int orc_gen_net(const orc_pop* pops, int32_t n_pops, const orc_proj* pros, int32_t n_pros,
                uint32_t mt_seed, uint32_t rand_seed, uint64_t philox_seed,
                float dw, float dt, orc_synlist* out);
void orc_free_synlist(orc_synlist* s);
/* The fixed in-degree connector's draws for chosen targets alone (target-major, so
 * each target's K draws stand apart from the others'): for each local target
 * tloc[q] of projection j (K draws from sn sources), local source, weight and
 * quantised delay (roundf(delay / step)) of draw k at [q * K + k]. */
void orc_fixk_targets(uint64_t philox_seed, int32_t j, int32_t sn, int32_t K, float wmin, float wmax, float dmin,
                      float dmax, float dw, float dt, float step, const int32_t* tloc, int32_t ntl, int32_t* src,
                      float* w, int32_t* dsteps);

/* ---- CSR build (simulator/mgsim/adj/mgsim.cpp:8-144, npart = 1) ---------- */
typedef struct {
    int32_t n;
    int64_t s;
    int64_t* row_ptr;    /* [n+1] axon_offs */
    int32_t* col;        /* [s]   target, CSR order */
    int32_t* dly;        /* [s]   delay steps */
    float* w;            /* [s]   weight */
    uint8_t* plastic;    /* [s] */
    int64_t* perm;       /* [s]   synapse-list index of each CSR slot (axon_refs) */
    int64_t* in_ptr;     /* [n+1] dend_offs */
    int64_t* in_pos;     /* [s]   CSR slot of each incoming synapse (dend_refs) */
    int32_t max_delay;   /* max quantised delay + 1 == ring length D */
    int32_t min_delay;
} orc_csr;

int orc_build_csr(int32_t n, int64_t s, const int32_t* src, const int32_t* tar,
                  const float* weight, const float* delay_s, const uint8_t* plastic,
                  float step, orc_csr* out);
void orc_free_csr(orc_csr* c);
/* Incoming lists of a CSR given as arrays (e.g. one read back from the device): for
 * every target, the CSR slots of its plastic synapses in slot order (in_ptr [n+1],
 * in_pos [plastic count]); the STDP step uses no other order.  Returns the count. */
int64_t orc_in_csr(int32_t n, int64_t s, const int32_t* col, const uint8_t* plastic, int64_t* in_ptr,
                   int64_t* in_pos);

/* ---- simulation (simulator/mgsim/adj/mgcore.cu:116-412, 535-554) -------- */
enum {
    ORC_ACC_F32 = 0,     /* FP32 state, sequential float ring adds in canonical order */
    ORC_ACC_FIXED = 1,   /* FP32 state, int32 fixed-point ring (LSB 1/scale; LIFB reads exc + inh as one sum) */
    ORC_ACC_F64 = 2      /* FP64 state and accumulation (accuracy oracle) */
};

typedef struct {
    int32_t n;
    const int32_t* type;      /* [n] NeuronType */
    const float* rate;        /* [n] per-step Poisson probability */
    const float* v0;          /* [n] initial membrane potential */
    const int32_t* cset;      /* [n] constant-set index (NULL: all 0) */
    int32_t n_cset;
    const float* lifconst;    /* [n_cset*30] LIFECONSTS */
    const float* stdpconst;   /* [6] STDPCONSTS */
    int32_t stdp;             /* Config::STDP */
    const orc_csr* csr;
    uint64_t seed;            /* Poisson Philox key */
    int32_t acc;              /* ORC_ACC_* */
    float scale_exc, scale_inh;  /* fixed-point scales (powers of two) */
} orc_desc;

typedef struct orc_sim orc_sim;

orc_sim* orc_create(const orc_desc* d);
/* Advances nsteps.  Spike ids of each step are appended in ascending order to
 * ids (capacity cap); step_off[k] is the offset of step k (nsteps+1 entries).
 * Returns the number of spikes written, or -1 if cap was exceeded. */
int64_t orc_run(orc_sim* s, int32_t nsteps, int32_t* ids, int64_t cap, int64_t* step_off);
/* Multithreaded step loop of a static network (OpenMP; every mode, no STDP): the
 * same results as orc_run, bit for bit (each thread owns a target range and adds
 * the fired rows' synapses into it in ascending source order, as orc_run does).
 * ids / cap / step_off as for orc_run (ids may be NULL).  Returns the spike count,
 * -1 with STDP, -2 if cap was exceeded. */
int64_t orc_run_mt(orc_sim* S, int32_t nsteps, int32_t nthreads, int32_t* ids, int64_t cap, int64_t* step_off);
enum { ORC_F_V = 0, ORC_F_IEXC = 1, ORC_F_IINH = 2, ORC_F_REFR = 3, ORC_F_LAST = 4, ORC_F_CNT = 5 };
/* Float fields are written as double. */
void orc_get_state(const orc_sim* s, int field, double* out);
void orc_get_weights(const orc_sim* s, double* out);   /* CSR order */
int64_t orc_syn_events(const orc_sim* s);
void orc_destroy(orc_sim* s);

/* Poisson uniform in (0,1] of neuron id at step s (bit-identical on the GPU). */
float orc_poisson_u(uint64_t seed, uint32_t id, uint32_t step);
/* Initial membrane potential of neuron local_id of a Poisson/network stream,
 * uniform on [lo, hi) with 24 random bits (new feature; the reference starts
 * every neuron at LIFECONSTS[1], mgsim.cpp:97). */
float orc_vinit(uint64_t seed, uint32_t local_id, float lo, float hi);
/* STDP weight change for integer step difference dt (mgcore.cu:346-364). */
float orc_stdp_dw(const float* stdpconst, int32_t dt);

#ifdef __cplusplus
}
#endif
#endif

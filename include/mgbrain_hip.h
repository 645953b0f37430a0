/*
 * mgbrain_hip.h -- C ABI of the MI355X single-GPU simulation step.
 *
 * This is the drop-in boundary under MGBrain's public C++ API.  Each entry
 * point names the reference interface it replaces (paths relative to the
 * MGBrain source tree, simulator/mgsim/adj/ unless stated).  The C++ facade in
 * include/mgbrain/ (class MGSimulator etc.) is written on top of it.
 *
 * Conventions
 *   - Every function returns an mgb_status; on failure mgb_last_error() gives a
 *     thread-local message.  Nothing calls exit() or resets the device (the
 *     reference's CUDACHECK exits, cudamem.cuh:8-26, and ~MGSimulator calls
 *     cudaDeviceReset, mgsim.cpp:405).
 *   - The caller owns every array it passes; the library copies what it keeps.
 *   - One mgb_sim owns one HIP stream on one device.  Calls on one handle are
 *     not thread-safe; distinct handles are independent.
 *   - Neuron ids, CSR order and the synapse-list order are those of the
 *     reference's build_multinet for npart = 1 (mgsim.cpp:8-144).
 */
#ifndef MGBRAIN_HIP_H
#define MGBRAIN_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGB_ABI_VERSION 1

typedef enum mgb_status {
    MGB_OK = 0,
    MGB_ERR_INVALID = 1,      /* bad argument or network (e.g. a delay below one step) */
    MGB_ERR_HIP = 2,          /* HIP runtime error */
    MGB_ERR_OOM = 3,          /* device or host allocation failed */
    MGB_ERR_UNSUPPORTED = 4,  /* valid but outside this library's limits */
    MGB_ERR_CAPACITY = 5      /* spike record capacity exceeded (data kept up to it) */
} mgb_status;

/* NeuronType values, macro.h:40-48. */
enum { MGB_RAW = 0, MGB_LIF0 = 1, MGB_POISSON = 2, MGB_LIF = 3, MGB_LIFE = 4, MGB_LIFB = 5 };

/* Storage of static synaptic weights on the device. */
enum {
    MGB_WF_AUTO = 0,   /* CLASS when every neuron's excitatory and inhibitory inputs each
                          share one weight value, else F32 */
    MGB_WF_CLASS = 1,  /* 4 B/synapse: target|delay|sign; per-class weight, exact integer
                          counts reproduce the reference's FP32 sum of equal addends */
    MGB_WF_F32 = 2,    /* per-synapse FP32 weight, int32 fixed-point accumulation (a LIFB
                          neuron reads its excitatory + inhibitory sums as one) */
    MGB_WF_BF16 = 3,   /* per-synapse BF16 of w / s, s = the source row's largest |w|
                          (plastic synapses stay FP32) */
    MGB_WF_FP8 = 4     /* per-synapse FP8-E4M3 of w / s, s = the source row's largest |w|
                          / 448, or the caller's fp8_scale */
};

typedef struct mgb_opts {
    int32_t device;            /* HIP device ordinal */
    int32_t weight_format;     /* MGB_WF_* */
    int32_t window;            /* steps per neuron launch; 0 = min delay (1 with STDP) */
    int32_t use_graph;         /* replay the step loop from a captured hipGraph */
    int64_t record_capacity;   /* spike records kept on the device; 0 = none */
    int32_t prop_block;        /* propagation workgroup size (64..1024); 0 = tuned default */
    int32_t prop_grid;         /* propagation workgroups; 0 = tuned default */
    int32_t profile;           /* time every kernel with hipEvents (direct launches only) */
    int32_t propagation;       /* MGB_PROP_* */
    int32_t tile_size;         /* TILED: mean target neurons per workgroup (tiles are balanced by
                                  in-degree and never mix neuron types); 0 = default */
    int32_t chunk;             /* TILED: synapses per propagation work item (1, 2, 4, 8, 16; 0 = 8) */
    int32_t persistent;        /* TILED without STDP: MGB_PERSIST_* */
    int32_t spin_limit_us;     /* persistent launches: bound on a workgroup's wait between two
                                  windows before the run fails with MGB_ERR_HIP; 0 = 2 s */
    int32_t reserved[2];
} mgb_opts;

/* Launch form of the fused kernel. */
enum {
    MGB_PERSIST_AUTO = 0,  /* the whole windows of a run (or of a graph block) in one launch when
                              every workgroup fits on the device at once, else one launch per window;
                              16-bit CLASS words in networks of 32768 neurons or more (and at most
                              512 tiles): with list regions (MGB_PERSIST_REGIONS) */
    MGB_PERSIST_OFF = 1,   /* one launch per window */
    MGB_PERSIST_REGIONS = 2, /* as AUTO, with list regions whatever the size (where the words and
                              tiles allow them): in every window of a launch but its last, each
                              workgroup writes its tile's spikes at the tile's own place of the
                              window's list and passes the count with its wait flag */
    MGB_PERSIST_LISTS = 3  /* as AUTO, without list regions: every window's list by a global atomic */
};

/* Spike propagation strategy. */
enum {
    MGB_PROP_AUTO = 0,    /* TILED where the fused kernel takes the network (delays <= 32 steps;
                             with STDP, FP32 or BF16 weights), else DIRECT */
    MGB_PROP_DIRECT = 1,  /* one wave per spike over its CSR row, integer atomics to the ring in HBM */
    MGB_PROP_TILED = 2    /* one fused kernel per window: each workgroup owns a tile of targets,
                             keeps their ring slice in LDS, adds the arriving synapses from a
                             tile-sorted CSR copy and advances the tile's neurons */
};

/* A network as the reference's Network holds it (model/network.h:12-68).
 * Reference: the host->device seam copy_subnet_gpu (mgnet.h:31, mgnet.cpp:92-146)
 * and copy_consts_gpu (mgcore.cuh:84, mgcore.cu:685-748). */
typedef struct mgb_net_desc {
    int32_t n_neurons;
    const int32_t* type;          /* [n] NeuronType */
    const float* rate;            /* [n] Neuron::rate, per-step Poisson probability */
    const float* v_init;          /* [n] initial V; NULL = lifconst[cset][1] (mgsim.cpp:97) */
    const uint8_t* cset;          /* [n] constant-set index; NULL = 0 */
    const uint16_t* stream;       /* [n] Poisson stream (network) index; NULL = 0 */
    int32_t n_csets;              /* >= 1 */
    const float* lifconst;        /* [n_csets][30] LIFECONSTS, mgcore.cu:32-63 */
    const float* stdpconst;       /* [6] STDPCONSTS, mgcore.cu:73-80 */
    int32_t n_streams;            /* >= 1 */
    const int32_t* stream_base;   /* [n_streams] first neuron id of each stream; NULL = 0 */
    const uint64_t* stream_seed;  /* [n_streams] Poisson key of each stream; NULL = seed */
    int64_t n_synapses;
    const int32_t* syn_src;       /* [s] Synapse::src (list order) */
    const int32_t* syn_tar;       /* [s] Synapse::tar */
    const float* syn_weight;      /* [s] Synapse::weight */
    const float* syn_delay;       /* [s] Synapse::delay, seconds */
    const uint8_t* syn_plastic;   /* [s] STDP applies; NULL = all (global Config::STDP) */
    const float* fp8_scale;       /* [n] per-source weight scale for MGB_WF_FP8; NULL = auto */
    float step;                   /* Config::STEP, seconds */
    int32_t stdp;                 /* Config::STDP */
    uint64_t seed;                /* Poisson key when stream_seed is NULL */
} mgb_net_desc;

/* Device-side network generation (the reference generates on the host,
 * model/netgen.cpp:39-229; its placeholder model/netgeng.cu is empty). */
enum { MGB_CONN_BERNOULLI = 1, MGB_CONN_FIXED_INDEGREE = 2, MGB_CONN_ONE_TO_ONE = 3,
       MGB_CONN_FULL = 4 };

typedef struct mgb_gen_pop {
    int32_t num;
    int32_t type;          /* NeuronType */
    float rate;            /* Poisson per-step probability */
    float v_min, v_max;    /* initial V uniform on [v_min, v_max); equal = constant */
    int32_t cset;          /* constant-set index */
    int32_t stream;        /* Poisson stream / network index */
} mgb_gen_pop;

typedef struct mgb_gen_proj {
    int32_t src, tar;      /* population indices */
    int32_t kind;          /* MGB_CONN_* */
    float p;               /* Bernoulli probability */
    int32_t k;             /* fixed in-degree */
    float wmin, wmax;      /* Projection::wrange */
    float dmin, dmax;      /* Projection::drange, seconds */
    int32_t plastic;
    int32_t index;         /* projection index inside its own network (Philox counter) */
    uint64_t seed;         /* Philox key of its network */
} mgb_gen_proj;

typedef struct mgb_gen_desc {
    int32_t n_pops;
    const mgb_gen_pop* pops;
    int32_t n_projs;
    const mgb_gen_proj* projs;
    int32_t n_csets;
    const float* lifconst;        /* [n_csets][30] */
    const float* stdpconst;       /* [6] */
    int32_t n_streams;
    const uint64_t* stream_seed;  /* [n_streams]; NULL = seed */
    float step;                   /* Config::STEP */
    float dt;                     /* Config::DT (delay grid) */
    float dw;                     /* Config::DW (weight grid) */
    int32_t stdp;
    uint64_t seed;                /* Poisson key when stream_seed is NULL */
    uint64_t net_seed;            /* key of the initial-V draws (NetGen's seed) */
} mgb_gen_desc;

typedef struct mgb_info {
    int32_t n_neurons;
    int32_t ring_len;          /* D = max delay + 1 (mgsim.cpp:71) */
    int32_t min_delay;         /* mgsim.cpp:51 */
    int32_t window;            /* steps per neuron launch in use */
    int64_t n_synapses;
    int64_t step;              /* steps simulated so far */
    int64_t syn_events;        /* sum over spikes of out-degree */
    int64_t spikes_recorded;   /* may exceed record_capacity (then data is truncated) */
    int32_t weight_format;     /* MGB_WF_* in use */
    int32_t n_classes;         /* weight classes (CLASS format) */
    float fix_scale_exc;       /* fixed-point scales (F32/BF16/FP8 formats) */
    float fix_scale_inh;
    int64_t device_bytes;      /* device memory held by the handle */
    double last_run_ms;        /* device time of the last mgb_run (hipEvents) */
    int64_t launches;          /* kernel launches issued so far (a graph replay counts once) */
    int64_t spikes;            /* spikes emitted so far */
    /* per-kernel device time of the last mgb_run when opts.profile is set; with
     * use_graph on the fused path a replay's time counts as its windows' fused
     * launches (the graph's one finalize launch included) */
    double neuron_ms, prop_ms, stdp_ms;
    int64_t neuron_launches, prop_launches, stdp_launches;
    int32_t propagation;       /* MGB_PROP_* in use (TILED: the fused kernel's time is prop_ms) */
    int32_t tile_size;         /* the widest tile */
    int32_t tiles;
    int32_t block;             /* workgroup size of the propagation / fused kernel */
    int32_t segments;          /* independent segments (networks of a batch) */
    int32_t word_bytes;        /* bytes per synapse the fused kernel reads (0: direct path; fused
                                  STDP with BF16 static weights: 4, and 4 more for a plastic one) */
    int32_t persistent;        /* 1: runs use the persistent launch (not MGB_PERSIST_OFF, and every
                                  workgroup resident); 2: the persistent launch with list regions */
} mgb_info;

typedef struct mgb_sim mgb_sim;

/* Library version and last error message (thread-local). */
int32_t mgb_abi_version(void);
const char* mgb_last_error(void);

/* Builds a simulator: device CSR (stable counting order by source, i.e. the
 * reference's axon_offs/axon_refs), delay quantisation round(delay/step)
 * (mgsim.cpp:49), ring buffers, state and constants.
 * Replaces MGSimulator::MGSimulator (mgsim.h:53, mgsim.cpp:337-363) with
 * build_multinet (mgsim.cpp:8-144), copy_gnets_to_gpu (:145-174),
 * copy_subnet_gpu (mgnet.cpp:92-146), init_state (mgcore.cu:665-669) and
 * copy_consts_gpu (mgcore.cu:685-748). */
mgb_status mgb_create(const mgb_net_desc* net, const mgb_opts* opts, mgb_sim** out);

/* As mgb_create, with the synapse list generated on the device from
 * population/projection descriptions (new connectors; see mgb_gen_proj). */
mgb_status mgb_create_generated(const mgb_gen_desc* gen, const mgb_opts* opts, mgb_sim** out);

/* Batched form (no counterpart in the reference, which simulates one network
 * per MGSimulator): n_nets independent networks in one simulator, advanced by
 * the same kernel launches.  Network b's neuron ids follow those of network
 * b-1 (read-back is in that order); its constant sets, Poisson streams and
 * seeds are its own and every draw uses ids local to it, so each network
 * evolves exactly as mgb_create_generated(&nets[b]) alone would (CLASS
 * format).  The networks must share the time step and the STDP settings. */
mgb_status mgb_create_batch_generated(const mgb_gen_desc* nets, int32_t n_nets, const mgb_opts* opts,
                                      mgb_sim** out);

/* Advances nsteps time steps.  Replaces MGSimulator::simulate (mgsim.cpp:921-941)
 * -> simulate_sparse (:794-857) -> mgsim_step_sparse (mgcore.cu:535-554). */
mgb_status mgb_run(mgb_sim* sim, int32_t nsteps);

/* Blocks until the handle's stream is idle. */
mgb_status mgb_sync(mgb_sim* sim);

/* Changes the launch shape of the propagation (DIRECT) or fused (TILED) kernel
 * without touching the data layout: workgroup size (64..1024, multiple of 64)
 * and, for TILED, synapses per work item (4, 8, 16).  0 keeps the current
 * value.  Used by the autotuner (mgbrain_amd/tune.py); no reference counterpart. */
mgb_status mgb_set_launch(mgb_sim* sim, int32_t block, int32_t chunk);

mgb_status mgb_get_info(mgb_sim* sim, mgb_info* out);

/* Recorded spikes sorted by (step, neuron id): at most cap entries are
 * written; *n_out receives the number written.  Returns MGB_ERR_CAPACITY if
 * the device record overflowed (the entries kept are still written). */
mgb_status mgb_read_spikes(mgb_sim* sim, int64_t cap, int32_t* steps, int32_t* ids, int64_t* n_out);
/* Drops the recorded spikes (the next record starts empty). */
mgb_status mgb_clear_spikes(mgb_sim* sim);

/* Per-neuron state.  Fields: V_m, I_exc, I_inh (float), Refrac_state,
 * last_fired, fire_cnt (int32) -- NEUBlock, mgstruct.h:8-30. */
enum { MGB_F_V = 0, MGB_F_IEXC = 1, MGB_F_IINH = 2, MGB_F_REFR = 3, MGB_F_LAST = 4, MGB_F_CNT = 5 };
mgb_status mgb_read_state(mgb_sim* sim, int32_t field, void* out);

/* fire_cnt of every neuron; replaces get_subnet_firecnt (mgnet.h:33, mgnet.cpp:148-158). */
mgb_status mgb_fire_counts(mgb_sim* sim, int32_t* out);

/* The fused path's tiles: tile_lo[0..tiles] (mgb_info.tiles + 1 entries), tile g
 * holding neurons [tile_lo[g], tile_lo[g + 1]).  MGB_ERR_UNSUPPORTED on the direct
 * path.  No reference counterpart (introspection of the library's tiling). */
mgb_status mgb_read_tiles(mgb_sim* sim, int32_t* tile_lo);

/* Device CSR in the reference's axon order: row_ptr[n+1] (axon_offs), target,
 * delay in steps and weight of each slot (SYNBlock via axon_refs,
 * mgstruct.h:32-47).  Any output pointer may be NULL. */
mgb_status mgb_read_csr(mgb_sim* sim, int64_t* row_ptr, int32_t* col, int32_t* delay, float* weight);

/* Current weights in CSR order (changed by STDP). */
mgb_status mgb_read_weights(mgb_sim* sim, float* out);

/* As mgb_read_weights, allocating no device memory (no hipMalloc / hipFree):
 * no counterpart in the reference; for callers that cannot allocate (code under
 * graph capture, tight serving loops, harnesses that forbid it).  Its working
 * memory is one fixed device array per GPU (2^24 floats, 64 MiB, held until the
 * process ends): a network needing more -- S slots for the CLASS, BF16 and FP8
 * formats, the padded tile-order length with fused STDP -- gets
 * MGB_ERR_UNSUPPORTED.  Calls on one GPU share that array: calls that overlap on
 * one GPU (from several threads) race. */
mgb_status mgb_read_weights_noalloc(mgb_sim* sim, float* out);

/* Checkpoint of everything a run changes: neuron state, ring buffers, the
 * pending spike lists and their counters, step counters, plastic weights and
 * the recorded spikes (the reference keeps none; its stale network writer is
 * test/networkio.cpp:4-24).  save writes mgb_checkpoint_size() bytes; load
 * restores them into a simulator built from the same network and options,
 * after which runs continue bit-identically. */
mgb_status mgb_checkpoint_size(mgb_sim* sim, int64_t* bytes);
mgb_status mgb_checkpoint_save(mgb_sim* sim, void* buf, int64_t bytes);
mgb_status mgb_checkpoint_load(mgb_sim* sim, const void* buf, int64_t bytes);

/* Frees everything the handle owns.  Replaces ~MGSimulator (mgsim.cpp:364-406)
 * and free_gsubnet(_gpu) (mgnet.h:29-30), without cudaDeviceReset. */
mgb_status mgb_destroy(mgb_sim* sim);

#ifdef __cplusplus
}
#endif
#endif

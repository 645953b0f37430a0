// macro.h -- global configuration and basic types of the MGBrain C++ API.
// Same names and meaning as the reference's src/macro.h:29-171 (Config statics,
// NeuronType, SynapseType, PartType, real, timer); additions are marked "new".
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

namespace MGBrain {

const float CONN_TYPE_O2O = 0;     // macro.h:31-32
const float CONN_TYPE_FULL = 1;
typedef float real;                // macro.h:34-39
typedef unsigned long sid;

enum NeuronType { RAW, LIF0, POISSON, LIF, LIFE, LIFB };   // macro.h:40-48
enum SynapseType { STATIC, STDP };                          // macro.h:49-53
enum PartType { METIS, MODEL, BSIM, SIMPLE };               // macro.h:54-60

struct LIFConsts {
    real P22, P11exc, P11inh, P21exc, P21inh, C_m, V_reset, V_rest, V_thresh, Tau_m, Tau_exc, Tau_inh,
        I_offset, Tau_refrac;
};
struct STDPConsts {
    real A_LTP, A_LTD, TAU_LTP, TAU_LTD, W_max, W_min;
};

// Static configuration (macro.h:89-157, defaults macro.cpp:3-14).
class Config {
   public:
    static real STEP;
    static real DT;
    static real DW;
    static bool STDP;
    static bool SINGLE_GPU;
    static bool PEER_ACCESS;
    static bool SHUFFLE_GEN;
    static bool FIRE_CHECK;
    static bool DENSE_SPIKE;
    static bool STAGGERED;
    static float MIN_GPUMEM;
    static bool SEQUENCE;
    // new: Philox key of the Poisson streams and of seeded network generation
    static uint64_t SEED;
    // new: STDP only on projections created with SynapseType::STDP (the reference
    // applies it to every synapse when STDP is enabled)
    static bool STDP_PER_PROJECTION;
    // new: device weight storage (mgb_opts.weight_format) and spike record capacity
    static int WEIGHT_FORMAT;
    static long long RECORD_CAPACITY;
    static int DEVICE;

    static void set_timestep(real time_step) {
        STEP = time_step;
        DT = time_step;
    }
    static int get_steps(real time) { return (int)std::round(time * 1000 / STEP); }
    static void set_dw(real dw) { DW = dw; }
    static void enable_stdp() { STDP = true; }
    static void enable_single_gpu() {
        SINGLE_GPU = true;
        PEER_ACCESS = false;
    }
    static void enable_shuffle_gen() { SHUFFLE_GEN = true; }
    static void enable_fire_check() { FIRE_CHECK = true; }
    static void enable_dense_spike() { DENSE_SPIKE = true; }
    static void disable_peer_access() { PEER_ACCESS = false; }
    static void enable_staggered() { STAGGERED = true; }
    static void set_min_gpumem(float mem) { MIN_GPUMEM = mem; }
    static void enable_sequence() { SEQUENCE = true; }
    // new
    static void set_seed(uint64_t seed) { SEED = seed; }
    static void enable_stdp_per_projection() { STDP_PER_PROJECTION = true; }
    static void set_weight_format(int wf) { WEIGHT_FORMAT = wf; }
    static void set_record_capacity(long long cap) { RECORD_CAPACITY = cap; }
};

// Wall-clock timer (macro.h:161-171).
class timer {
    std::chrono::time_point<std::chrono::high_resolution_clock> s;

   public:
    timer() { s = std::chrono::high_resolution_clock::now(); }
    double stop() {
        return std::chrono::duration_cast<std::chrono::microseconds>(
                   std::chrono::high_resolution_clock::now() - s)
                   .count() *
               1e-6;
    }
};

}  // namespace MGBrain

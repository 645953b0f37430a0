// mgsim.h -- the simulator class (reference src/simulator/mgsim/adj/mgsim.h:4-60)
// over the MI355X C ABI (include/mgbrain_hip.h).
//
// Same construction, simulate() and get_time() as the reference.  Differences:
//  - errors throw std::runtime_error instead of exit(1) (cudamem.cuh:8-26);
//  - the destructor releases this simulator only (no device reset, mgsim.cpp:405);
//  - nparts > 1 is accepted when Config::SINGLE_GPU is set (all partitions live
//    on device 0 in the reference too, mgsim.cpp:341-352) and simulated as one
//    partition; otherwise it throws;
//  - repeated simulate() calls continue the step count (the reference restarts
//    its step counter at 0 on every call, mgsim.cpp:810);
//  - Poisson inputs are keyed by Config::SEED (the reference seeds from the clock).
#pragma once
#include <stdexcept>
#include <string>

#include "../model/model.h"
#include "../model/network.h"
#include "mgbrain_hip.h"

namespace MGBrain {

class MGSimulator {
    mgb_sim* sim = nullptr;
    double sim_time = 0.0;
    int n = 0;
    void open(const mgb_net_desc& d);

   public:
    MGSimulator(Network& net, std::vector<int>& part, int nparts);
    // new: generate the network on the device from the model's Bernoulli /
    // fixed in-degree projections (reference connectors need the host NetGen)
    MGSimulator(Model& model, uint64_t seed);
    ~MGSimulator();
    MGSimulator(const MGSimulator&) = delete;
    MGSimulator& operator=(const MGSimulator&) = delete;

    void simulate(real time);
    double get_time();
    // reference get_firing_rate (mgsim.cpp:596-615), printing like it
    real get_firing_rate(real time);
    // new read-back
    std::vector<int> fire_counts();
    std::vector<float> membrane_potentials();
    void spikes(std::vector<int>& steps, std::vector<int>& ids);
    long long syn_events();
    // new: checkpoint / resume to and from a file (mgb_checkpoint_save / _load); load
    // into a simulator built from the same network and Config
    void save(const std::string& path);
    void load(const std::string& path);
    mgb_sim* handle() { return sim; }
};

}  // namespace MGBrain

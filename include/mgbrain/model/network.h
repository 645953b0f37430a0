// network.h -- the host network handed to the simulator (reference
// src/model/network.h:12-68).  New: Neuron::v_init, Synapse::plastic.
#pragma once
#include "../macro.h"
#include "model.h"

namespace MGBrain {

struct Pop {
    int id;
    int num;
    bool source;
    NeuronType type;
    std::vector<int> neurons;
    Pop(int _id, int _num, bool _source, NeuronType _type) : id(_id), num(_num), source(_source), type(_type) {
        neurons.resize(_num);
    }
};

struct Neuron {
    int id;
    bool source;
    NeuronType type;
    real rate;                 // per-step Poisson firing probability
    int fire_cnt;
    std::vector<int> nxt;
    std::vector<int> pre;
    std::vector<size_t> syns;  // as the reference: synapse index + 1 (netgen.cpp:108)
    bool has_v_init = false;   // new: per-neuron initial V (else LIFECONSTS[1])
    real v_init = 0;
    Neuron(int _id, bool _source, NeuronType _type) : id(_id), source(_source), type(_type) {
        rate = 0;
        fire_cnt = 0;
    }
};

struct Synapse {
    int src;
    int tar;
    real weight;
    real delay;                // seconds
    bool plastic = false;      // new: STDP applies (per-projection plasticity)
    Synapse(int _src, int _tar, real _weight, real _delay) : src(_src), tar(_tar), weight(_weight), delay(_delay) {}
    Synapse(int _src, int _tar, real _weight, real _delay, bool _plastic)
        : src(_src), tar(_tar), weight(_weight), delay(_delay), plastic(_plastic) {}
};

class Network {
   public:
    std::vector<Pop> pops;
    std::vector<Neuron> neurons;
    std::vector<Synapse> synapses;
    std::array<real, 30> lifconst{};
    std::array<real, 6> stdpconst{};
    bool nlifconst = false;
    bool nstdpconst = false;
};

}  // namespace MGBrain

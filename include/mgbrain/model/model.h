// model.h -- populations, projections and the Model builder (reference
// src/model/model.h:12-63, model.cpp:15-57).  New: Bernoulli and fixed
// in-degree connectors (connect_bernoulli / connect_fixed_indegree).
#pragma once
#include "../macro.h"

namespace MGBrain {

// Connector of a projection.  CONN_REF follows Projection::ctype exactly as the
// reference's NetGen does (0 one-to-one, 1 full, (0,1) quota generator).
enum ConnKind { CONN_REF = 0, CONN_BERNOULLI = 1, CONN_FIXED_INDEGREE = 2 };

struct Population {
    int id;
    int num;
    bool source;
    NeuronType type;
    Population(int _id, int _num, bool _source, NeuronType _type)
        : id(_id), num(_num), source(_source), type(_type) {}
};

struct Projection {
    int src;
    int tar;
    std::array<real, 2> wrange;
    std::array<real, 2> drange;
    float ctype;
    SynapseType stype;
    int conn = CONN_REF;    // new
    float prob = 0.f;       // new: Bernoulli probability
    int indegree = 0;       // new: fixed in-degree
    Projection(int _src, int _tar, std::array<real, 2> _wrange, std::array<real, 2> _drange, float _ctype,
               SynapseType _stype)
        : src(_src), tar(_tar), wrange(_wrange), drange(_drange), ctype(_ctype), stype(_stype) {}
};

class Model {
   public:
    std::vector<Population> pops;
    std::vector<Projection> pros;
    std::array<real, 30> lifconst{};
    bool nlifconst = false;
    std::array<real, 6> stdpconst{};
    bool nstdpconst = false;
    int indexer;
    Model();
    void set_lif_const(std::array<real, 30> lifconst);
    void set_stdp_const(std::array<real, 6> stdpconst);
    Population& create_pop(int num, NeuronType type, bool isSource = false);
    bool connect(Population& src, Population tar, std::array<real, 2> _wrange, std::array<real, 2> _drange,
                 float type, SynapseType _stype);
    bool connect(Population& src, Population tar, real weight, real delay, float type, SynapseType _stype);
    bool connect(Population& src, Population tar, std::array<real, 2> _wrange, std::array<real, 2> _drange,
                 float type);
    bool connect(Population& src, Population tar, real weight, real delay, float type);
    // new connectors (generated identically on the host and on the device)
    bool connect_bernoulli(Population& src, Population tar, std::array<real, 2> wrange,
                           std::array<real, 2> drange, float p, SynapseType stype = STATIC);
    bool connect_bernoulli(Population& src, Population tar, real weight, real delay, float p,
                           SynapseType stype = STATIC);
    bool connect_fixed_indegree(Population& src, Population tar, std::array<real, 2> wrange,
                                std::array<real, 2> drange, int k, SynapseType stype = STATIC);
    bool connect_fixed_indegree(Population& src, Population tar, real weight, real delay, int k,
                                SynapseType stype = STATIC);
    // new: per-population Poisson rate and initial-V range applied by NetGen and by
    // the device generator (the reference sets Neuron::rate by hand after gen_net)
    std::vector<real> pop_rate;
    std::vector<std::array<real, 2>> pop_vrange;
    std::vector<bool> pop_has_vrange;
    void set_pop_rate(const Population& p, real rate);
    void set_pop_vinit(const Population& p, real lo, real hi);
};

}  // namespace MGBrain

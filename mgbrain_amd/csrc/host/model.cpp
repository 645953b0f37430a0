// model.cpp -- Config statics and the Model builder of the C++ facade
// (reference src/macro.cpp:3-14, src/model/model.cpp:3-57).
#include "mgbrain/model/model.h"

namespace MGBrain {

real Config::STEP = 0.001f;
real Config::DT = 0.001f;
real Config::DW = 0.01f;
bool Config::STDP = false;
bool Config::SINGLE_GPU = false;
bool Config::PEER_ACCESS = true;
bool Config::SHUFFLE_GEN = false;
bool Config::FIRE_CHECK = false;
bool Config::DENSE_SPIKE = false;
bool Config::STAGGERED = false;
float Config::MIN_GPUMEM = 3.0f;
bool Config::SEQUENCE = false;
uint64_t Config::SEED = 0;
bool Config::STDP_PER_PROJECTION = false;
int Config::WEIGHT_FORMAT = 0;
long long Config::RECORD_CAPACITY = 0;
int Config::DEVICE = 0;

Model::Model() { indexer = 0; }

void Model::set_lif_const(std::array<real, 30> c) {
    nlifconst = true;
    lifconst = c;
}

void Model::set_stdp_const(std::array<real, 6> c) {
    nstdpconst = true;
    stdpconst = c;
}

Population& Model::create_pop(int num, NeuronType type, bool isSource) {
    pops.emplace_back(indexer++, num, isSource, type);
    pop_rate.push_back(0.0f);
    pop_vrange.push_back({0.0f, 0.0f});
    pop_has_vrange.push_back(false);
    return pops.back();
}

void Model::set_pop_rate(const Population& p, real rate) { pop_rate.at(p.id) = rate; }

void Model::set_pop_vinit(const Population& p, real lo, real hi) {
    pop_vrange.at(p.id) = {lo, hi};
    pop_has_vrange.at(p.id) = true;
}

namespace {
// Adds or overwrites the projection src->tar.  The overwrite test is
// `index > 0` as in the reference (model.cpp:33), so a second connect of the
// first projection appends a duplicate instead of overwriting it.
Projection& upsert(Model& m, const Population& src, const Population& tar, std::array<real, 2> w,
                   std::array<real, 2> d, float ctype, SynapseType stype) {
    int index = -1;
    for (int i = 0; i < (int)m.pros.size(); i++)
        if (m.pros[i].src == src.id && m.pros[i].tar == tar.id) index = i;
    if (index > 0) {
        Projection& p = m.pros[index];
        p.wrange = w;
        p.drange = d;
        p.ctype = ctype;
        p.stype = stype;
        return p;
    }
    m.pros.emplace_back(src.id, tar.id, w, d, ctype, stype);
    return m.pros.back();
}
}  // namespace

bool Model::connect(Population& src, Population tar, std::array<real, 2> w, std::array<real, 2> d, float ctype,
                    SynapseType stype) {
    if (ctype == 0.0 && src.num != tar.num) return false;
    Projection& p = upsert(*this, src, tar, w, d, ctype, stype);
    p.conn = CONN_REF;
    return true;
}

bool Model::connect(Population& src, Population tar, real weight, real delay, float ctype, SynapseType stype) {
    return connect(src, tar, {weight, weight}, {delay, delay}, ctype, stype);
}

bool Model::connect(Population& src, Population tar, std::array<real, 2> w, std::array<real, 2> d, float ctype) {
    return connect(src, tar, w, d, ctype, SynapseType::STATIC);
}

bool Model::connect(Population& src, Population tar, real weight, real delay, float ctype) {
    return connect(src, tar, {weight, weight}, {delay, delay}, ctype, SynapseType::STATIC);
}

bool Model::connect_bernoulli(Population& src, Population tar, std::array<real, 2> w, std::array<real, 2> d,
                              float p, SynapseType stype) {
    if (!(p >= 0.0f && p <= 1.0f)) return false;
    Projection& pr = upsert(*this, src, tar, w, d, p, stype);
    pr.conn = CONN_BERNOULLI;
    pr.prob = p;
    return true;
}

bool Model::connect_bernoulli(Population& src, Population tar, real weight, real delay, float p,
                              SynapseType stype) {
    return connect_bernoulli(src, tar, {weight, weight}, {delay, delay}, p, stype);
}

bool Model::connect_fixed_indegree(Population& src, Population tar, std::array<real, 2> w, std::array<real, 2> d,
                                   int k, SynapseType stype) {
    if (k < 0 || (k > 0 && src.num == 0)) return false;
    Projection& pr = upsert(*this, src, tar, w, d, 0.5f, stype);
    pr.conn = CONN_FIXED_INDEGREE;
    pr.indegree = k;
    return true;
}

bool Model::connect_fixed_indegree(Population& src, Population tar, real weight, real delay, int k,
                                   SynapseType stype) {
    return connect_fixed_indegree(src, tar, {weight, weight}, {delay, delay}, k, stype);
}

}  // namespace MGBrain

// netgen.cpp -- host network generation of the C++ facade (reference
// src/model/netgen.cpp:1-229).  Reference connectors draw from std::mt19937 and
// glibc rand() in the reference's order, so a model generated here with the
// same clock seed and rand() state equals the reference's synapse list bit for
// bit.  The Bernoulli and fixed in-degree connectors draw from Philox streams
// (mgb_rng.h) shared with the device generator.
#include <algorithm>
#include <cstdlib>
#include <random>

#include "mgb_rng.h"
#include "mgbrain/model/netgen.h"

namespace MGBrain {

void NetGen::rand_vector(std::vector<bool>& list, int gen_num, int max_num) {
    list.assign(max_num, false);
    for (int i = 0; i < max_num && i < gen_num; i++) list[i] = true;
    for (int index = max_num - 1; index > 0; --index) {
        const int j = rand() % max_num;
        const bool t = list[index];
        list[index] = list[j];
        list[j] = t;
    }
}

// netgen.cpp:19-31 (note: positive ranges step downwards from `floor`)
real NetGen::rand_weight(real floor, real ceil) {
    if (ceil == floor) return floor;
    const int range = (int)((ceil - floor) / Config::DW) + 1;
    const int k = rand() % range;
    if (floor < 0) return floor + k * Config::DW;
    return floor - k * Config::DW;
}

// netgen.cpp:32-38
real NetGen::rand_delay(real floor, real ceil) {
    if (ceil == floor) return floor;
    const int range = (int)((ceil - floor) / Config::DT) + 1;
    return floor + (rand() % range) * Config::DT;
}

namespace {

struct Builder {
    Network& net;
    void add(int src, int tar, real w, real d, bool plastic) {
        net.synapses.emplace_back(src, tar, w, d, plastic);
        net.neurons[src].nxt.push_back(tar);
        net.neurons[tar].pre.push_back(src);
        net.neurons[src].syns.push_back(net.synapses.size());
    }
};

void generate(Network& net, Model& model, std::mt19937& rng, uint64_t philox_seed) {
    net.nlifconst = model.nlifconst;
    net.nstdpconst = model.nstdpconst;
    net.lifconst = model.lifconst;
    net.stdpconst = model.stdpconst;
    net.pops.reserve(model.pops.size());
    int nindexer = 0;
    for (size_t i = 0; i < model.pops.size(); i++) {                // netgen.cpp:56-68
        const Population& pop = model.pops[i];
        net.pops.emplace_back(pop.id, pop.num, pop.source, pop.type);
        net.neurons.reserve(net.neurons.size() + pop.num);
        for (int j = 0; j < pop.num; j++) {
            const int id = nindexer++;
            net.pops[i].neurons[j] = id;
            net.neurons.emplace_back(id, pop.source, pop.type);
            Neuron& n = net.neurons.back();
            if (i < model.pop_rate.size()) n.rate = model.pop_rate[i];
            if (i < model.pop_has_vrange.size() && model.pop_has_vrange[i]) {
                n.has_v_init = true;
                n.v_init = mgb::vinit_draw(philox_seed, (uint32_t)id, model.pop_vrange[i][0],
                                           model.pop_vrange[i][1]);
            }
        }
    }
    model.indexer = nindexer;
    Builder b{net};
    for (size_t j = 0; j < model.pros.size(); j++) {
        const Projection& pr = model.pros[j];
        const Pop& sp = net.pops[pr.src];
        const Pop& tp = net.pops[pr.tar];
        const int sn = sp.num, tn = tp.num;
        const real minw = pr.wrange[0], maxw = pr.wrange[1];
        const real mind = pr.drange[0], maxd = pr.drange[1];
        const bool plastic = pr.stype == SynapseType::STDP;
        if (pr.conn == CONN_BERNOULLI) {
            const uint64_t thr = (uint64_t)((double)pr.prob * 4294967296.0);
            for (uint64_t m = 0; m < (uint64_t)sn; ++m)
                for (uint64_t n = 0; n < (uint64_t)tn; ++n) {
                    const uint64_t f = m * (uint64_t)tn + n;
                    if ((uint64_t)mgb::philox_pick(philox_seed, f, (uint32_t)j, mgb::kTagBern) >= thr) continue;
                    const real w = mgb::grid_weight(mgb::philox_pick(philox_seed, f, (uint32_t)j, mgb::kTagBernW),
                                                    minw, maxw, Config::DW);
                    const real d = mgb::grid_delay(mgb::philox_pick(philox_seed, f, (uint32_t)j, mgb::kTagBernD),
                                                   mind, maxd, Config::DT);
                    b.add(sp.neurons[m], tp.neurons[n], w, d, plastic);
                }
        } else if (pr.conn == CONN_FIXED_INDEGREE) {
            const uint64_t K = (uint64_t)pr.indegree;
            for (uint64_t n = 0; n < (uint64_t)tn; ++n)
                for (uint64_t k = 0; k < K; ++k) {
                    const uint64_t f = n * K + k;
                    const uint32_t x = mgb::philox_pick(philox_seed, f, (uint32_t)j, mgb::kTagFixK);
                    const uint32_t m = (uint32_t)(((uint64_t)x * (uint64_t)sn) >> 32);
                    const real w = mgb::grid_weight(mgb::philox_pick(philox_seed, f, (uint32_t)j, mgb::kTagFixKW),
                                                    minw, maxw, Config::DW);
                    const real d = mgb::grid_delay(mgb::philox_pick(philox_seed, f, (uint32_t)j, mgb::kTagFixKD),
                                                   mind, maxd, Config::DT);
                    b.add(sp.neurons[m], tp.neurons[n], w, d, plastic);
                }
        } else if (pr.ctype == 1.0f) {                               // full: netgen.cpp:84-111
            net.synapses.reserve(net.synapses.size() + (size_t)sn * tn);
            for (int m = 0; m < sn; m++)
                for (int n = 0; n < tn; n++) {
                    const real w = NetGen::rand_weight(minw, maxw);
                    const real d = NetGen::rand_delay(mind, maxd);
                    b.add(sp.neurons[m], tp.neurons[n], w, d, plastic);
                }
        } else if (pr.ctype == 0.0f) {                               // one-to-one: :112-126
            for (int k = 0; k < sn; k++) {
                const real w = NetGen::rand_weight(minw, maxw);
                const real d = NetGen::rand_delay(mind, maxd);
                b.add(sp.neurons[k], tp.neurons[k], w, d, plastic);
            }
        } else if (pr.ctype < 1.0f && pr.ctype > 0.0f) {
            if (Config::SHUFFLE_GEN) {                               // :127-163
                std::vector<bool> genlist((size_t)sn * tn);
                const size_t gen_num = std::round(genlist.size() * pr.ctype);
                for (size_t k = 0; k < genlist.size(); k++) genlist[k] = k < gen_num;
                // libstdc++'s random_shuffle: swap i with rand() % (i + 1)
                for (size_t i = 1; i < genlist.size(); ++i) {
                    const size_t jj = (size_t)((long)rand() % (long)(i + 1));
                    if (jj != i) {
                        const bool t = genlist[i];
                        genlist[i] = genlist[jj];
                        genlist[jj] = t;
                    }
                }
                for (int m = 0; m < sn; m++)
                    for (int n = 0; n < tn; n++)
                        if (genlist[(size_t)m * tn + n]) {
                            const real w = NetGen::rand_weight(minw, maxw);
                            const real d = NetGen::rand_delay(mind, maxd);
                            b.add(sp.neurons[m], tp.neurons[n], w, d, plastic);
                        }
            } else {                                                 // quota generator :164-227
                const size_t max_num = (size_t)sn * (size_t)tn;
                const size_t gen_num = static_cast<size_t>(std::round(max_num * pr.ctype));
                const double lim = gen_num * 1.1;
                size_t count = 0;
                std::vector<size_t> genlist(gen_num);
                for (size_t m = 0; m < (size_t)sn; m++)
                    for (size_t n = 0; n < (size_t)tn; n++)
                        if (count < gen_num && ((rng() % max_num) < lim)) genlist[count++] = m * tn + n;
                net.synapses.reserve(net.synapses.size() + gen_num);
                // all gen_num entries (:210): unfilled ones are zero, i.e. the projection's
                // first source to its first target, when the 1.1p acceptance ran out early
                for (size_t t = 0; t < genlist.size(); t++) {
                    const int m = (int)(genlist[t] / tn);
                    const int n = (int)(genlist[t] % tn);
                    const real w = NetGen::rand_weight(minw, maxw);
                    const real d = NetGen::rand_delay(mind, maxd);
                    b.add(sp.neurons[m], tp.neurons[n], w, d, plastic);
                }
            }
        }
    }
}

}  // namespace

void NetGen::gen_net(Network& net, Model& model) {
    const auto count = std::chrono::system_clock::now().time_since_epoch().count();
    std::mt19937 rng(count);                                          // netgen.cpp:47
    generate(net, model, rng, (uint64_t)count);
}

void NetGen::gen_net(Network& net, Model& model, uint64_t seed) {
    std::mt19937 rng((std::mt19937::result_type)seed);
    generate(net, model, rng, seed);
}

}  // namespace MGBrain

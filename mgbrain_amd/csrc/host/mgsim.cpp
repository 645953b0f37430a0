// mgsim.cpp -- MGSimulator of the C++ facade over the C ABI (reference
// src/simulator/mgsim/adj/mgsim.cpp:337-363 constructor, :364-406 destructor,
// :596-615 firing rate, :794-857 / :921-941 simulate, :942-945 get_time).
#include "mgbrain/simulator/mgsim.h"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>

namespace MGBrain {

namespace {

// LIFECONSTS initialiser of the reference's __constant__ array (mgcore.cu:32-63):
// in effect when only STDP constants were set (copy_consts_gpu's second overload
// leaves LIFECONSTS untouched, mgcore.cu:735-748).
std::array<real, 30> constant_initialiser() {
    std::array<real, 30> c{};
    c[0] = 0.0001f;
    c[4] = 0.02f;
    c[5] = 0.25f;
    c[6] = 10.0f;
    c[8] = 0.0004f;
    c[10] = 1.0f;
    c[11] = 1.0f;
    c[21] = 0.951229453f;
    c[22] = 0.967216074f;
    c[23] = 0.0799999982f;
    c[24] = 0.0599999987f;
    c[25] = 0.90483743f;
    return c;
}

// copy_consts_gpu(max_delay, dt) defaults (mgcore.cu:685-733): DT = Config::DT, V_THRESH = 15.
std::array<real, 30> default_consts() {
    std::array<real, 30> c = constant_initialiser();
    c[0] = Config::DT;
    c[4] = 15.0f;
    return c;
}

const std::array<real, 6> kStdpDefault = {0.1f, -0.01f, 17, 34, 40, 0};   // mgcore.cu:73-80

void check(mgb_status st) {
    if (st != MGB_OK) throw std::runtime_error(std::string("mgbrain: ") + mgb_last_error());
}

mgb_opts opts_from_config() {
    mgb_opts o;
    std::memset(&o, 0, sizeof(o));
    o.device = Config::DEVICE;
    o.weight_format = Config::WEIGHT_FORMAT;
    o.record_capacity = Config::RECORD_CAPACITY;
    return o;
}

}  // namespace

void MGSimulator::open(const mgb_net_desc& d) {
    mgb_opts o = opts_from_config();
    check(mgb_create(&d, &o, &sim));
}

MGSimulator::MGSimulator(Network& net, std::vector<int>& part, int nparts) {
    if (nparts > 1 && !Config::SINGLE_GPU)
        throw std::invalid_argument("MGSimulator: more than one GPU partition is not supported by this library");
    if (part.size() != net.neurons.size() && nparts > 1)
        throw std::invalid_argument("MGSimulator: partition vector size does not match the network");
    std::cout << (nparts <= 1 ? "run single-net on single-GPU" : "run multi-net on single-GPU") << std::endl;
    const size_t nn = net.neurons.size(), ns = net.synapses.size();
    n = (int)nn;
    std::vector<int32_t> type(nn);
    std::vector<float> rate(nn), v0(nn);
    for (size_t i = 0; i < nn; ++i) {
        net.neurons[i].id = (int)i;                              // mgsim.cpp:26
        type[i] = (int32_t)net.neurons[i].type;
        rate[i] = net.neurons[i].rate;
        v0[i] = net.neurons[i].has_v_init ? net.neurons[i].v_init : net.lifconst[1];   // mgsim.cpp:97
    }
    // constants as copy_gnets_to_gpu selects them (mgsim.cpp:165-172)
    std::array<real, 30> lc;
    std::array<real, 6> sc = kStdpDefault;
    if (net.nlifconst || net.nstdpconst) {
        lc = net.nlifconst ? net.lifconst : constant_initialiser();
        if (net.nstdpconst) sc = net.stdpconst;
    } else {
        lc = default_consts();
    }
    std::vector<int32_t> src(ns), tar(ns);
    std::vector<float> w(ns), d(ns);
    std::vector<uint8_t> pl;
    if (Config::STDP) pl.resize(ns);
    for (size_t k = 0; k < ns; ++k) {
        const Synapse& s = net.synapses[k];
        src[k] = s.src;
        tar[k] = s.tar;
        w[k] = s.weight;
        d[k] = s.delay;
        if (Config::STDP) pl[k] = Config::STDP_PER_PROJECTION ? (s.plastic ? 1 : 0) : 1;
    }
    mgb_net_desc desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.n_neurons = (int32_t)nn;
    desc.type = type.data();
    desc.rate = rate.data();
    desc.v_init = v0.data();
    desc.n_csets = 1;
    desc.lifconst = lc.data();
    desc.stdpconst = sc.data();
    desc.n_streams = 1;
    desc.n_synapses = (int64_t)ns;
    desc.syn_src = src.data();
    desc.syn_tar = tar.data();
    desc.syn_weight = w.data();
    desc.syn_delay = d.data();
    desc.syn_plastic = Config::STDP ? pl.data() : nullptr;
    desc.step = Config::STEP;
    desc.stdp = Config::STDP ? 1 : 0;
    desc.seed = Config::SEED;
    open(desc);
}

MGSimulator::MGSimulator(Model& model, uint64_t seed) {
    std::vector<mgb_gen_pop> pops;
    std::vector<mgb_gen_proj> projs;
    const std::array<real, 30> lc = model.nlifconst ? model.lifconst : default_consts();
    const std::array<real, 6> sc = model.nstdpconst ? model.stdpconst : kStdpDefault;
    n = 0;
    for (size_t i = 0; i < model.pops.size(); ++i) {
        const Population& p = model.pops[i];
        mgb_gen_pop g;
        std::memset(&g, 0, sizeof(g));
        g.num = p.num;
        g.type = (int32_t)p.type;
        g.rate = i < model.pop_rate.size() ? model.pop_rate[i] : 0.0f;
        const bool vr = i < model.pop_has_vrange.size() && model.pop_has_vrange[i];
        g.v_min = vr ? model.pop_vrange[i][0] : model.lifconst[1];
        g.v_max = vr ? model.pop_vrange[i][1] : model.lifconst[1];
        pops.push_back(g);
        n += p.num;
    }
    for (size_t j = 0; j < model.pros.size(); ++j) {
        const Projection& pr = model.pros[j];
        mgb_gen_proj g;
        std::memset(&g, 0, sizeof(g));
        g.src = pr.src;
        g.tar = pr.tar;
        if (pr.conn == CONN_BERNOULLI) g.kind = MGB_CONN_BERNOULLI;
        else if (pr.conn == CONN_FIXED_INDEGREE) g.kind = MGB_CONN_FIXED_INDEGREE;
        else throw std::invalid_argument("MGSimulator(Model&): reference connectors are generated on the host "
                                         "(NetGen::gen_net); use connect_bernoulli / connect_fixed_indegree");
        g.p = pr.prob;
        g.k = pr.indegree;
        g.wmin = pr.wrange[0];
        g.wmax = pr.wrange[1];
        g.dmin = pr.drange[0];
        g.dmax = pr.drange[1];
        g.plastic = Config::STDP && (!Config::STDP_PER_PROJECTION || pr.stype == SynapseType::STDP);
        g.index = (int32_t)j;
        g.seed = seed;
        projs.push_back(g);
    }
    mgb_gen_desc desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.n_pops = (int32_t)pops.size();
    desc.pops = pops.data();
    desc.n_projs = (int32_t)projs.size();
    desc.projs = projs.data();
    desc.n_csets = 1;
    desc.lifconst = lc.data();
    desc.stdpconst = sc.data();
    desc.n_streams = 1;
    desc.step = Config::STEP;
    desc.dt = Config::DT;
    desc.dw = Config::DW;
    desc.stdp = Config::STDP ? 1 : 0;
    desc.seed = Config::SEED;     // Poisson streams, as the host-network constructor
    desc.net_seed = seed;         // initial-V draws, as NetGen::gen_net(net, model, seed)
    mgb_opts o = opts_from_config();
    check(mgb_create_generated(&desc, &o, &sim));
    std::cout << "run single-net on single-GPU" << std::endl;
}

MGSimulator::~MGSimulator() {
    if (sim) mgb_destroy(sim);
}

void MGSimulator::simulate(real time) {
    const int steps = static_cast<int>(std::round(time / Config::STEP));   // mgsim.cpp:801
    timer t;
    check(mgb_run(sim, steps));
    check(mgb_sync(sim));
    sim_time = t.stop();
    if (Config::FIRE_CHECK) get_firing_rate(time);
}

double MGSimulator::get_time() { return sim_time; }

real MGSimulator::get_firing_rate(real time) {
    const std::vector<int> c = fire_counts();
    size_t num = 0;
    for (int v : c) num += (size_t)v;
    const int neusize = (int)c.size();
    std::cout << "zone(0)'fire cnt:" << num << std::endl;
    std::cout << "zone(0)'fire rate:" << (float)num / neusize / time << " hz" << std::endl;
    const float rate = ((float)num / neusize / time);
    std::cout << "fire rate:" << rate << "hz" << std::endl;
    return rate;
}

std::vector<int> MGSimulator::fire_counts() {
    std::vector<int> c((size_t)n);
    check(mgb_fire_counts(sim, c.data()));
    return c;
}

std::vector<float> MGSimulator::membrane_potentials() {
    std::vector<float> v((size_t)n);
    check(mgb_read_state(sim, MGB_F_V, v.data()));
    return v;
}

void MGSimulator::spikes(std::vector<int>& steps, std::vector<int>& ids) {
    mgb_info inf;
    check(mgb_get_info(sim, &inf));
    const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(inf.spikes_recorded, Config::RECORD_CAPACITY));
    steps.resize((size_t)cap);
    ids.resize((size_t)cap);
    int64_t got = 0;
    const mgb_status st = mgb_read_spikes(sim, cap, steps.data(), ids.data(), &got);
    if (st != MGB_OK && st != MGB_ERR_CAPACITY) check(st);
    steps.resize((size_t)got);
    ids.resize((size_t)got);
}

long long MGSimulator::syn_events() {
    mgb_info inf;
    check(mgb_get_info(sim, &inf));
    return inf.syn_events;
}

void MGSimulator::save(const std::string& path) {
    int64_t bytes = 0;
    check(mgb_checkpoint_size(sim, &bytes));
    std::vector<uint8_t> buf((size_t)bytes);
    check(mgb_checkpoint_save(sim, buf.data(), bytes));
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
    if (!f) throw std::runtime_error("cannot write checkpoint " + path);
}

void MGSimulator::load(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read checkpoint " + path);
    std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    check(mgb_checkpoint_load(sim, buf.data(), (int64_t)buf.size()));
}

}  // namespace MGBrain

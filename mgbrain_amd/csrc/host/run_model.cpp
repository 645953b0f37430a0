// run_model.cpp -- drives the C++ facade from a small text spec and writes the
// synapse list, the recorded spikes and the final state as raw binary.  Used by
// the tests to check the facade end to end (Model -> NetGen -> MGSimulator ->
// C ABI -> HIP kernels) against the CPU oracle.
//
// Spec keys (one per line):
//   timestep DT | dw DW | clock SEED (NetGen seed) | srand S | shuffle 0/1
//   seed S (Poisson key) | stdp | stdp_per_projection | stdpconst a b c d e f
//   lifconst2 | lifconst1 | lifconst v_init v_rest v_reset cm tau_m t_ref tse tsi v_th i_off
//   pop NUM TYPE SOURCE | proj S T WMIN WMAX DMIN DMAX CTYPE [STDP]
//   bern S T WMIN WMAX DMIN DMAX P [STDP] | fixk S T WMIN WMAX DMIN DMAX K [STDP]
//   brunel NSYN | vogel NSYN | rate POP R | vinit POP LO HI
//   steps T | record CAP | weight_format W | device_gen 0/1 | nparts N | single_gpu |
//   checkpoint K PATH (save after K steps, finish in a new simulator loaded from PATH)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "mgbrain/mgbrain.h"

using namespace MGBrain;

template <class T>
static void put(FILE* f, const T* p, size_t n) {
    if (n && fwrite(p, sizeof(T), n, f) != n) {
        std::perror("fwrite");
        std::exit(3);
    }
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s SPEC OUT\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    Model m;
    std::vector<Population> pops;
    uint64_t clock_seed = 1;
    unsigned srand_seed = 1;
    int steps = 100, device_gen = 0, nparts = 1, ckpt_at = -1;
    std::string ckpt_path;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string k;
        if (!(ss >> k) || k[0] == '#') continue;
        if (k == "timestep") { float v; ss >> v; Config::set_timestep(v); }
        else if (k == "dw") { float v; ss >> v; Config::set_dw(v); }
        else if (k == "clock") ss >> clock_seed;
        else if (k == "srand") ss >> srand_seed;
        else if (k == "shuffle") { int v; ss >> v; if (v) Config::enable_shuffle_gen(); }
        else if (k == "seed") { uint64_t v; ss >> v; Config::set_seed(v); }
        else if (k == "stdp") Config::enable_stdp();
        else if (k == "stdp_per_projection") Config::enable_stdp_per_projection();
        else if (k == "stdpconst") {
            std::array<real, 6> c{};
            for (auto& x : c) ss >> x;
            m.set_stdp_const(c);
        } else if (k == "lifconst2") m.set_lif_const(ConstGen::gen_life_const2());
        else if (k == "lifconst1") m.set_lif_const(ConstGen::gen_life_const1());
        else if (k == "lifconst") {
            real a[10];
            for (auto& x : a) ss >> x;
            m.set_lif_const(ConstGen::gen_life_const(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9]));
        } else if (k == "pop") {
            int num, type, source;
            ss >> num >> type >> source;
            pops.push_back(m.create_pop(num, (NeuronType)type, source != 0));
        } else if (k == "proj" || k == "bern" || k == "fixk") {
            int s, t, st = 0;
            float wmin, wmax, dmin, dmax, x;
            ss >> s >> t >> wmin >> wmax >> dmin >> dmax >> x;
            if (!(ss >> st)) st = 0;
            const SynapseType stype = st ? SynapseType::STDP : SynapseType::STATIC;
            if (k == "proj") m.connect(pops.at(s), pops.at(t), {wmin, wmax}, {dmin, dmax}, x, stype);
            else if (k == "bern") m.connect_bernoulli(pops.at(s), pops.at(t), {wmin, wmax}, {dmin, dmax}, x, stype);
            else m.connect_fixed_indegree(pops.at(s), pops.at(t), {wmin, wmax}, {dmin, dmax}, (int)x, stype);
        } else if (k == "brunel") {
            int nsyn; ss >> nsyn; ModelMaker::make_brunel(m, nsyn);
            pops = m.pops;
        } else if (k == "vogel") {
            int nsyn; ss >> nsyn; ModelMaker::make_vogel(m, nsyn);
            pops = m.pops;
        } else if (k == "rate") {
            int p; float r; ss >> p >> r; m.set_pop_rate(m.pops.at(p), r);
        } else if (k == "vinit") {
            int p; float lo, hi; ss >> p >> lo >> hi; m.set_pop_vinit(m.pops.at(p), lo, hi);
        } else if (k == "steps") ss >> steps;
        else if (k == "record") { long long c; ss >> c; Config::set_record_capacity(c); }
        else if (k == "weight_format") { int w; ss >> w; Config::set_weight_format(w); }
        else if (k == "device_gen") ss >> device_gen;
        else if (k == "checkpoint") ss >> ckpt_at >> ckpt_path;
        else if (k == "nparts") ss >> nparts;
        else if (k == "single_gpu") Config::enable_single_gpu();
        else {
            std::fprintf(stderr, "unknown spec key %s\n", k.c_str());
            return 2;
        }
    }
    Network net;
    MGSimulator* sim = nullptr;
    try {
        std::vector<int> part;
        if (device_gen) {
            sim = new MGSimulator(m, clock_seed);
        } else {
            std::srand(srand_seed);
            NetGen::gen_net(net, m, clock_seed);
            part.assign(net.neurons.size(), 0);
            sim = new MGSimulator(net, part, nparts);
        }
        if (ckpt_at >= 0) {
            // run ckpt_at steps, save, and finish the run in a new simulator of the
            // same network restored from the file
            sim->simulate(ckpt_at * Config::STEP);
            sim->save(ckpt_path);
            delete sim;
            sim = nullptr;
            sim = device_gen ? new MGSimulator(m, clock_seed) : new MGSimulator(net, part, nparts);
            sim->load(ckpt_path);
            sim->simulate((steps - ckpt_at) * Config::STEP);
        } else {
            sim->simulate(steps * Config::STEP);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        delete sim;
        return 4;
    }
    std::vector<int> st, ids;
    sim->spikes(st, ids);
    std::vector<float> v = sim->membrane_potentials();
    std::vector<int> cnt = sim->fire_counts();
    const long long se = sim->syn_events();
    FILE* f = std::fopen(argv[2], "wb");
    if (!f) return 3;
    int64_t hdr[4] = {(int64_t)v.size(), (int64_t)net.synapses.size(), (int64_t)st.size(), se};
    put(f, hdr, 4);
    std::vector<int32_t> src, tar;
    std::vector<float> w, d;
    std::vector<uint8_t> pl;
    for (const Synapse& s : net.synapses) {
        src.push_back(s.src);
        tar.push_back(s.tar);
        w.push_back(s.weight);
        d.push_back(s.delay);
        pl.push_back(s.plastic ? 1 : 0);
    }
    put(f, src.data(), src.size());
    put(f, tar.data(), tar.size());
    put(f, w.data(), w.size());
    put(f, d.data(), d.size());
    put(f, pl.data(), pl.size());
    put(f, st.data(), st.size());
    put(f, ids.data(), ids.size());
    put(f, v.data(), v.size());
    put(f, cnt.data(), cnt.size());
    std::vector<float> v0;
    std::vector<float> rates;
    for (const Neuron& n : net.neurons) {
        v0.push_back(n.has_v_init ? n.v_init : net.lifconst[1]);
        rates.push_back(n.rate);
    }
    put(f, v0.data(), v0.size());
    put(f, rates.data(), rates.size());
    std::fclose(f);
    delete sim;
    return 0;
}

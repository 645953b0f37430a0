// ref_netgen_main.cpp -- harness around the reference's own host model layer.
//
// TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh compiles this file together
// with the reference's macro.cpp, model/model.cpp and model/netgen.cpp (from
// their own sources, unmodified) into oracle/_ref/ref_netgen.  It reads a small
// text spec, builds a Model through the reference API (Model::create_pop /
// connect, ModelMaker::make_brunel / make_vogel), runs NetGen::gen_net and
// writes the resulting synapse list as raw binary for tests/test_ref_netgen.py.
//
// NetGen::gen_net seeds its std::mt19937 from system_clock::now()
// (model/netgen.cpp:47).  This harness defines system_clock::now() so that the
// clock reads a count given in the spec, which makes the reference generator
// reproducible without touching its source; glibc rand() is seeded with srand.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "macro.h"
#include "model/constgen.h"
#include "model/maker.h"
#include "model/model.h"
#include "model/netgen.h"
#include "model/network.h"

static long g_clock_count = 0;

namespace std {
namespace chrono {
inline namespace _V2 {
system_clock::time_point system_clock::now() noexcept {
    return time_point(duration(g_clock_count));
}
}  // namespace _V2
}  // namespace chrono
}  // namespace std

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
    if (!in) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    Model m;
    std::vector<Population> pops;
    unsigned srand_seed = 1;
    bool have_model = false;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string key;
        if (!(ss >> key) || key[0] == '#') continue;
        if (key == "timestep") {
            float v; ss >> v; Config::set_timestep(v);
        } else if (key == "dw") {
            float v; ss >> v; Config::set_dw(v);
        } else if (key == "clock") {
            ss >> g_clock_count;
        } else if (key == "srand") {
            ss >> srand_seed;
        } else if (key == "shuffle") {
            int v; ss >> v; if (v) Config::enable_shuffle_gen();
        } else if (key == "pop") {
            int num, type, source;
            ss >> num >> type >> source;
            pops.push_back(m.create_pop(num, (NeuronType)type, source != 0));
            have_model = true;
        } else if (key == "proj") {
            int s, t;
            float wmin, wmax, dmin, dmax, ctype;
            ss >> s >> t >> wmin >> wmax >> dmin >> dmax >> ctype;
            m.connect(pops.at(s), pops.at(t), {wmin, wmax}, {dmin, dmax}, ctype);
        } else if (key == "brunel") {
            int nsyn; ss >> nsyn; ModelMaker::make_brunel(m, nsyn); have_model = true;
        } else if (key == "vogel") {
            int nsyn; ss >> nsyn; ModelMaker::make_vogel(m, nsyn); have_model = true;
        } else {
            std::fprintf(stderr, "unknown spec key %s\n", key.c_str());
            return 2;
        }
    }
    if (!have_model) {
        std::fprintf(stderr, "spec has no model\n");
        return 2;
    }
    std::srand(srand_seed);
    Network net;
    NetGen::gen_net(net, m);

    FILE* f = std::fopen(argv[2], "wb");
    if (!f) {
        std::perror("fopen");
        return 3;
    }
    int64_t nn = (int64_t)net.neurons.size(), ns = (int64_t)net.synapses.size();
    put(f, &nn, 1);
    put(f, &ns, 1);
    std::vector<int32_t> types(nn), nxt(nn), pre(nn);
    for (int64_t i = 0; i < nn; ++i) {
        types[i] = (int32_t)net.neurons[i].type;
        nxt[i] = (int32_t)net.neurons[i].nxt.size();
        pre[i] = (int32_t)net.neurons[i].pre.size();
    }
    put(f, types.data(), nn);
    put(f, nxt.data(), nn);
    put(f, pre.data(), nn);
    std::vector<int32_t> src(ns), tar(ns);
    std::vector<float> w(ns), d(ns);
    for (int64_t i = 0; i < ns; ++i) {
        src[i] = net.synapses[i].src;
        tar[i] = net.synapses[i].tar;
        w[i] = net.synapses[i].weight;
        d[i] = net.synapses[i].delay;
    }
    put(f, src.data(), ns);
    put(f, tar.data(), ns);
    put(f, w.data(), ns);
    put(f, d.data(), ns);
    int32_t nl = net.nlifconst ? 1 : 0;
    put(f, &nl, 1);
    put(f, net.lifconst.data(), 30);
    // ConstGen known answers at the current Config::DT (constgen.h:9-159).
    std::array<real, 30> c1 = ConstGen::gen_life_const1();
    std::array<real, 30> c2 = ConstGen::gen_life_const2();
    std::array<real, 30> c3 = ConstGen::gen_life_const3();
    std::array<real, 6> s1 = ConstGen::gen_stdp_const1();
    put(f, c1.data(), 30);
    put(f, c2.data(), 30);
    put(f, c3.data(), 30);
    put(f, s1.data(), 6);
    std::fclose(f);
    return 0;
}

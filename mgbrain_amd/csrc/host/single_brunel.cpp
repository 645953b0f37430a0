// single_brunel.cpp -- the reference's single-GPU Brunel benchmark
// (bench/single_brunel.cpp:11-62) written against this C++ facade, unchanged
// apart from its merge-conflict lines (:17-21 keep enable_single_gpu()).
// Usage: mgb_single_brunel --nsyn=10000000 --rate=0.002 [--seed=N]
#include <ctime>
#include <iostream>

#include "mgbrain/mgbrain.h"

int main(int argc, char* argv[]) {
    parser par(argc, argv);
    clock_t build_s, build_e, part_s, part_e, setup_s, setup_e;
    config::set_timestep(0.0001);
    config::enable_single_gpu();
    config::enable_fire_check();
    if (!par.getStr("--seed").empty()) config::set_seed(par.getNum("--seed"));
    build_s = clock();
    model m;
    size_t nsyn = par.getNum(parser::NSYN);
    float rate = par.getFloat(parser::RATE);
    maker::make_brunel(m, nsyn);
    network net;
    if (!par.getStr("--seed").empty()) netgen::gen_net(net, m, par.getNum("--seed"));
    else netgen::gen_net(net, m);
    build_e = clock();
    std::cout << "build time:" << (build_e - build_s) / 1000 << "ms" << std::endl;
    std::cout << "neu size:" << net.neurons.size() << std::endl;
    std::cout << "syn size:" << net.synapses.size() << std::endl;
    std::cout << "rate:" << rate << std::endl;
    int npart = 1;
    part_s = clock();
    std::vector<int> part = parter::part_network(net, npart, PartType::MODEL);
    part_e = clock();
    std::cout << "part time:" << (float)(part_e - part_s) / 1000 << "ms" << std::endl;
    for (size_t i = 0; i < net.neurons.size(); i++)
        if (net.neurons[i].type == NeuType::POISSON) net.neurons[i].rate = rate;
    setup_s = clock();
    mgsim sim(net, part, npart);
    setup_e = clock();
    std::cout << "steup time:" << (setup_e - setup_s) / 1000 << "ms" << std::endl;
    sim.simulate(1);
    std::cout << "sim time:" << sim.get_time() << std::endl;
    return 0;
}

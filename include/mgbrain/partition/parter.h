// parter.h -- partitioning entry point (reference src/partition/parter.h:35,
// parter.cpp:100-123).  One GPU: every neuron is in partition 0, which is what
// the reference returns for nparts <= 1 (parter.cpp:103-104).  Graph
// partitioning (METIS, MODEL, BSIM, SIMPLE) for several GPUs is not part of
// this build; nparts > 1 throws std::invalid_argument.
#pragma once
#include <stdexcept>

#include "../model/network.h"

namespace MGBrain {

class Partitioner {
   public:
    static std::vector<int> part_network(Network& network, int nparts, PartType type) {
        (void)type;
        if (nparts > 1)
            throw std::invalid_argument("Partitioner: only one partition (one GPU) is supported");
        return std::vector<int>(network.neurons.size(), 0);
    }
};

// Partition statistics printer (reference src/partition/analysis.h) for one
// partition: neuron and synapse counts, zero cut.
class PartitionAnalysis {
    std::string name;
    size_t nn, ns;

   public:
    PartitionAnalysis(std::string _name, std::vector<int>& part, Network& net, int npart)
        : name(std::move(_name)), nn(net.neurons.size()), ns(net.synapses.size()) {
        (void)part;
        (void)npart;
    }
    void printAll() {
        std::cout << name << " part cost:0" << std::endl;
        std::cout << name << " cut-rate:0(0/" << ns << ")" << std::endl;
        std::cout << name << " part zone neu:(0:" << nn << ")\t Standard Deviation:0" << std::endl;
        std::cout << name << " part zone syn:(0:" << ns << ")\t Standard Deviation:0" << std::endl;
    }
};

}  // namespace MGBrain

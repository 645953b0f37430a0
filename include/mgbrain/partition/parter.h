// parter.h -- partitioning entry point (reference src/partition/parter.h:35,
// parter.cpp:100-123).  nparts <= 1: every neuron is in partition 0, as the
// reference returns (parter.cpp:103-104).  nparts > 1 is accepted only with
// Config::SINGLE_GPU, the mode in which the reference places every partition on
// device 0 (mgsim.cpp:341-352) and MGSimulator here simulates them as one
// network: neurons are then split into nparts contiguous id ranges (the
// partition does not change a single-device run).  Graph partitioning (METIS,
// MODEL, BSIM) for several GPUs is not part of this library; nparts > 1 without
// SINGLE_GPU throws std::invalid_argument, as MGSimulator's constructor does.
#pragma once
#include <stdexcept>

#include "../macro.h"
#include "../model/network.h"

namespace MGBrain {

class Partitioner {
   public:
    static std::vector<int> part_network(Network& network, int nparts, PartType type) {
        (void)type;
        const size_t n = network.neurons.size();
        if (nparts <= 1) return std::vector<int>(n, 0);
        if (!Config::SINGLE_GPU)
            throw std::invalid_argument(
                "Partitioner: nparts > 1 needs Config::enable_single_gpu() (one GPU is supported)");
        std::vector<int> part(n);
        for (size_t i = 0; i < n; ++i) part[i] = (int)((i * (size_t)nparts) / n);
        return part;
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

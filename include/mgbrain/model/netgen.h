// netgen.h -- network generation (reference src/model/netgen.h:12-20,
// netgen.cpp:19-229).  gen_net(net, model) seeds std::mt19937 from the clock as
// the reference does; the new overload gen_net(net, model, seed) seeds it with
// `seed` (low 32 bits) and keys the Philox streams of the new connectors with
// `seed`.  Weight/delay ranges of the reference connectors draw glibc rand(),
// exactly like the reference.
#pragma once
#include "../macro.h"
#include "model.h"
#include "network.h"

namespace MGBrain {

class NetGen {
   public:
    static void rand_vector(std::vector<bool>& list, int gen_num, int max_num);
    static real rand_weight(real floor, real ceil);
    static real rand_delay(real floor, real ceil);
    static void gen_net(Network& net, Model& model);
    static void gen_net(Network& net, Model& model, uint64_t seed);   // new
};

}  // namespace MGBrain

// maker.h -- ready-made models (reference src/model/maker.h:5-43).
#pragma once
#include "constgen.h"
#include "model.h"

namespace MGBrain {

class ModelMaker {
   public:
    // Brunel network sized for about nsyn synapses: Poisson:E:I = 5:4:1, quota
    // connector p = 0.1, J = 0.1 mV * 20000/pop, inhibition -5J, 1.6 ms (maker.h:10-27).
    static void make_brunel(Model& m, int nsyn) {
        const float con_prob = 0.1f;
        const int pop_size = static_cast<int>(std::sqrt((float)nsyn / (con_prob * 0.5)));
        Population p = m.create_pop(pop_size * 5 / 10, NeuronType::POISSON, true);
        Population e = m.create_pop(pop_size * 4 / 10, NeuronType::LIFB, false);
        Population i = m.create_pop(pop_size * 1 / 10, NeuronType::LIFB, false);
        const real w1 = 0.0001 * 20000 / pop_size;
        const real w2 = -0.0005 * 20000 / pop_size;
        const real d1 = 0.0016f;
        m.set_lif_const(ConstGen::gen_life_const2());
        m.connect(p, e, w1, d1, con_prob);
        m.connect(p, i, w1, d1, con_prob);
        m.connect(e, e, w1, d1, con_prob);
        m.connect(e, i, w1, d1, con_prob);
        m.connect(i, i, w2, d1, con_prob);
        m.connect(i, e, w2, d1, con_prob);
    }
    // Vogels-Abbott COBA network, p = 0.02, 0.8 ms (maker.h:28-42).
    static void make_vogel(Model& m, int nsyn) {
        const float con_prob = 0.02f;
        const int pop_size = static_cast<int>(std::sqrt((float)nsyn / (con_prob)));
        Population pe = m.create_pop(pop_size * 8 / 10, NeuronType::LIFE, false);
        Population pi = m.create_pop(pop_size * 2 / 10, NeuronType::LIFE, false);
        const real w1 = 0.4 * 16000000 / pop_size / pop_size;
        const real w2 = -5.1 * 16000000 / pop_size / pop_size;
        const real d1 = 0.0008f;
        m.set_lif_const(ConstGen::gen_life_const1());
        m.connect(pe, pe, w1, d1, con_prob);
        m.connect(pe, pi, w1, d1, con_prob);
        m.connect(pi, pe, w2, d1, con_prob);
        m.connect(pi, pi, w2, d1, con_prob);
    }
};

}  // namespace MGBrain

// constgen.h -- LIF constant arrays (reference src/model/constgen.h:6-159).
// The array layout is the reference's LIFECONSTS index map, which the kernels
// read (0 DT, 1 V_INIT, 2 V_REST, 3 V_RESET, 4 V_THRESH, 5 C_M, 6 TAU_M,
// 7 I_OFFSET, 8 TAU_REFRAC, 10/11 TAU_SYN_EXC/INH, 12/13 P_EXC/INH, 14 V_TMP,
// 15/16 C_EXC/INH, 17 CM, 21-25 fixed propagator values); values follow the
// same closed forms so arrays built here equal the reference's bit for bit.
#pragma once
#include "../macro.h"

namespace MGBrain {

class ConstGen {
   public:
    static std::array<real, 30> gen_life_const(real v_init, real v_rest, real v_reset, real cm, real tau_m,
                                               real tau_refrac, real tau_syn_E, real tau_syn_I, real v_thresh,
                                               real i_offset) {
        const real dt = Config::DT;      // snapshot of DT at call time (constgen.h:11)
        real rm = 1.0f;
        if (std::fabs(cm) > 0.00001f) rm = tau_m / cm;
        const real decay_m = tau_m > 0 ? std::exp(-dt / tau_m) : 0.0f;
        const real decay_e = tau_syn_E > 0 ? std::exp(-dt / tau_syn_E) : 0.0f;
        const real decay_i = tau_syn_I > 0 ? std::exp(-dt / tau_syn_I) : 0.0f;
        // the reference scales an offset that it never assigns; every caller passes 0
        const real offset = 0.0f;
        real v_tmp = offset * rm + v_rest;
        v_tmp *= (1 - decay_m);
        real p_exc = rm * tau_syn_E / (tau_syn_E - tau_m);
        real p_inh = rm * tau_syn_I / (tau_syn_I - tau_m);
        p_exc = p_exc * (decay_e - decay_m);
        p_inh = p_inh * (decay_i - decay_m);
        std::array<real, 30> c{};
        c[0] = dt;
        c[1] = v_init;
        c[2] = v_rest;
        c[3] = v_reset;
        c[4] = v_thresh;
        c[5] = cm;
        c[6] = tau_m;
        c[7] = i_offset;
        c[8] = tau_refrac;
        c[10] = tau_syn_E;
        c[11] = tau_syn_I;
        c[12] = p_exc;
        c[13] = p_inh;
        c[14] = v_tmp;
        c[15] = decay_e;
        c[16] = decay_i;
        c[17] = decay_m;
        fixed_propagators(c);
        return c;
    }
    static std::array<real, 30> gen_lif0_const(real v_init, real v_rest, real v_reset, real cm, real tau_m,
                                               real tau_refrac, real tau_syn_E, real tau_syn_I, real v_thresh,
                                               real i_offset) {
        std::array<real, 30> c{};
        c[0] = Config::DT;
        c[1] = v_init;
        c[2] = v_rest;
        c[3] = v_reset;
        c[4] = v_thresh;
        c[5] = cm;
        c[6] = tau_m;
        c[7] = i_offset;
        c[8] = tau_refrac;
        c[10] = tau_syn_E;
        c[11] = tau_syn_I;
        c[12] = 0.9999f;
        c[13] = 0.9999f;
        c[14] = 0.0f;
        c[15] = 0.9999f;
        c[16] = 0.9999f;
        c[17] = 0.0f;
        fixed_propagators(c);
        return c;
    }
    // COBA parameter set used by make_vogel (constgen.h:135-138)
    static std::array<real, 30> gen_life_const1() {
        return gen_life_const(-0.06f, -0.06f, -0.06f, 0, 0, 0.005f, 1, 1, -0.05f, 0);
    }
    // Brunel parameter set used by make_brunel (constgen.h:139-142)
    static std::array<real, 30> gen_life_const2() {
        return gen_life_const(0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.002f, 1.0f, 1.0f, 0.02f, 0.0f);
    }
    static std::array<real, 30> gen_life_const3() {
        return gen_life_const(0.0f, 0.0f, 0.0f, 0.25f, 10.0f, 0.0004f, 1.0f, 1.0f, 15.0f, 0.0f);
    }
    static std::array<real, 6> gen_stdp_const(real a_ltp, real a_ltd, real tau_ltp, real tau_ltd, real w_max,
                                              real w_min) {
        return {a_ltp, a_ltd, tau_ltp, tau_ltd, w_max, w_min};
    }
    static std::array<real, 6> gen_stdp_const1() { return {0.1f, -0.01f, 17, 34, 40, 0}; }
    static std::array<real, 6> gen_stdp_const2() { return {}; }

   private:
    static void fixed_propagators(std::array<real, 30>& c) {
        c[21] = 0.951229453f;
        c[22] = 0.967216074f;
        c[23] = 0.0799999982f;
        c[24] = 0.0599999987f;
        c[25] = 0.90483743f;
    }
};

}  // namespace MGBrain

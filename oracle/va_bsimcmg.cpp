// va_bsimcmg.cpp — CPU evaluation of the BSIM-CMG model for the tests.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  The model code is the header that
// vagen generates from the model's Verilog-A source, compiled here for the
// host by g++ (the product compiles the same header for gfx950 with hipcc).
// This file evaluates the raw six-node model at given node voltages, in
// double or with dual numbers for dI/dV and dQ/dV; the internal-node solve
// and the terminal condensation the product does on the device are
// restated independently in oracle.py (numpy).  Nothing in the reference
// pins BSIM-CMG outputs numerically; the tests pin them by finite
// differences, charge and current conservation, Ids(Vds=0)=0 and
// source/drain exchange symmetry.
#include <cstdint>
#include <cstring>
#include <vector>

#include "va_runtime.h"
#include "bsimcmg.h"

using namespace va_bsimcmg;

namespace {

int set_param(Card& P, Inst& I, const char* name, double value) {
    for (int k = 0; k < n_card; k++)
        if (!std::strcmp(card_table[k].name, name)) {
            char* b = reinterpret_cast<char*>(&P);
            if (card_table[k].is_int) *reinterpret_cast<int*>(b + card_table[k].off) = (int)value;
            else *reinterpret_cast<double*>(b + card_table[k].off) = value;
            b[card_table[k].given_off] = 1;
            return 0;
        }
    for (int k = 0; k < n_inst; k++)
        if (!std::strcmp(inst_table[k].name, name)) {
            char* b = reinterpret_cast<char*>(&I);
            if (inst_table[k].is_int) *reinterpret_cast<int*>(b + inst_table[k].off) = (int)value;
            else *reinterpret_cast<double*>(b + inst_table[k].off) = value;
            b[inst_table[k].given_off] = 1;
            return 0;
        }
    return -1;
}

}  // namespace

extern "C" {

// Six-node raw evaluation: v6[n][6] node voltages (d, g, s, e, di, si);
// outputs I6, Q6 [n][6] and, when derivs, G66, C66 [n][6][6] (row = node
// equation, column = node voltage).  Returns 0 or -1 (unknown parameter).
int or_bsim_raw(int64_t n, const double* L, const double* nfin, const double* v6, double temp_c,
                int nparam, const char** names, const double* values, int derivs, double* I6,
                double* Q6, double* G66, double* C66) {
    static Card P;
    Inst I0;
    std::memset(&P, 0, sizeof(P));
    std::memset(&I0, 0, sizeof(I0));
    for (int k = 0; k < nparam; k++)
        if (set_param(P, I0, names[k], values[k]) != 0) return -1;
    card_defaults(P);
    const double temp = temp_c + 273.15;
    for (int64_t i = 0; i < n; i++) {
        Inst ip = I0;
        ip.L = L[i];
        ip.given_L = 1;
        ip.NFIN = nfin[i];
        ip.given_NFIN = 1;
        inst_defaults(P, ip);
        const double* v = v6 + 6 * i;
        if (derivs) {
            VaDual<6> V[6], Ir[6], Q[6];
            for (int k = 0; k < 6; k++) V[k] = VaDual<6>::var(v[k], k);
            eval<VaDual<6>>(P, ip, temp, V, Ir, Q);
            for (int k = 0; k < 6; k++) {
                I6[6 * i + k] = Ir[k].v;
                Q6[6 * i + k] = Q[k].v;
                for (int j = 0; j < 6; j++) {
                    G66[36 * i + 6 * k + j] = Ir[k].d[j];
                    C66[36 * i + 6 * k + j] = Q[k].d[j];
                }
            }
        } else {
            double Ir[6], Q[6];
            eval<double>(P, ip, temp, v, Ir, Q);
            for (int k = 0; k < 6; k++) {
                I6[6 * i + k] = Ir[k];
                Q6[6 * i + k] = Q[k];
            }
        }
    }
    return 0;
}

}  // extern "C"

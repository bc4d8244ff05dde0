// models/lambda_phi4.hpp -- V = lambda phi^4 / 4 + g2 phi^2 chi^2 / 2.
//
// The massless two-field preheating model (the default model of
// LATTICEEASY, whose model.h layout the reference's model.h keeps).  Terms:
// 0 = lambda phi^4 / 4, 1 = g2 phi^2 chi^2 / 2.  Couplings used: lambda, g2.
#pragma once
#include "../psc_model_api.hpp"

namespace psc_models {

struct LambdaPhi4 {
    static constexpr const char* name = "lambda_phi4";
    static constexpr const char* description = "V = 1/4 lambda PHI^4 + 1/2 g2 PHI^2 CHI^2";

    PSC_HD static constexpr int num_potential_terms(int nflds) { return nflds == 2 ? 2 : 1; }

    template <typename T>
    PSC_HD static T potential_energy(int term, const T* f, const psc::Couplings<T>& c) {
        const T p2 = f[0] * f[0];
        return term == 0 ? (T)0.25 * c.lambda * (p2 * p2) : (T)0.5 * c.g2 * p2 * (f[1] * f[1]);
    }
    template <typename T>
    PSC_HD static T dvdf(int fld, const T* f, const psc::Couplings<T>& c) {
        const T p2 = f[0] * f[0];
        return fld == 0 ? c.lambda * p2 * f[0] + c.g2 * f[0] * (f[1] * f[1]) : c.g2 * p2 * f[1];
    }
    template <typename T>
    PSC_HD static T d2Vdf2(int term, const T* f, const psc::Couplings<T>& c) {
        if (term == 0) return (T)3 * c.lambda * (f[0] * f[0]) + c.g2 * (f[1] * f[1]);
        if (term == 1) return (T)2 * c.g2 * f[0] * f[1];
        return c.g2 * (f[0] * f[0]);
    }
    template <typename T>
    PSC_HD static void effective_mass(T* mass_sq, const T* f, const psc::Couplings<T>& c) {
        mass_sq[0] = (T)3 * c.lambda * (f[0] * f[0]) + c.g2 * (f[1] * f[1]);
        mass_sq[1] = c.g2 * (f[0] * f[0]);
    }
    // (pb + p)^3 - pb^3 = p (3 pb^2 + 3 pb p + p^2); the coupling term as in Preheat.
    template <typename T>
    PSC_HD static void dvdf_delta(const T* mean, const T* d, T* out, const psc::Couplings<T>& c) {
        const T pb = mean[0], qb = mean[1], p = d[0], q = d[1];
        const T chi = qb + q;
        out[0] = c.lambda * p * ((T)3 * pb * pb + p * ((T)3 * pb + p)) +
                 c.g2 * (p * (chi * chi) + pb * q * ((T)2 * qb + q));
        out[1] = c.g2 * (p * ((T)2 * pb + p) * chi + pb * pb * q);
    }
    PSC_HD static constexpr bool has_coupling(int k) { return k == 0 || k == 2; }   // g2, lambda
    template <typename T>
    PSC_HD static T dV_dcoupling(int k, const T* f, const psc::Couplings<T>&) {
        const T p2 = f[0] * f[0];
        return k == 0 ? (T)0.5 * p2 * (f[1] * f[1]) : (T)0.25 * (p2 * p2);
    }
    template <typename T>
    PSC_HD static T d2V_dfdcoupling(int k, int fld, const T* f, const psc::Couplings<T>&) {
        if (k == 0) return fld == 0 ? f[0] * (f[1] * f[1]) : (f[0] * f[0]) * f[1];
        return fld == 0 ? (f[0] * f[0]) * f[0] : (T)0;
    }
};

}  // namespace psc_models

#ifndef PSC_MODEL
#define PSC_MODEL psc_models::LambdaPhi4
#endif

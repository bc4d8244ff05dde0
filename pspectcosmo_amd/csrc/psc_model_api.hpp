// psc_model_api.hpp -- the model interface: the potential of the lattice
// fields as compile-time hooks with the names of the reference's model.h
// (num_potential_terms, potential_energy, dvdf, d2Vdf2, effective_mass;
// model.h:34-104), as templates over the field precision T (float or double)
// that run on the host and on the device.
//
// A model is a struct with static members (see models/preheat.hpp):
//
//   static constexpr const char* name, *description;
//   PSC_HD static constexpr int num_potential_terms(int nflds);   // <= kMaxTerms
//   template <typename T> PSC_HD static T potential_energy(int term, const T* f, const Couplings<T>& c);
//   template <typename T> PSC_HD static T dvdf(int fld, const T* f, const Couplings<T>& c);
//   template <typename T> PSC_HD static T d2Vdf2(int term, const T* f, const Couplings<T>& c);
//   template <typename T> PSC_HD static void effective_mass(T* mass_sq, const T* f, const Couplings<T>& c);
//
// f always holds two values (phi, chi); with one field chi is 0.  d2Vdf2's
// term is the Hessian index fld + fld2 (0: phi phi, 1: phi chi, 2: chi chi),
// as the reference indexes it (evolution.cpp:229).  potential_energy of a
// single point only: the lattice average the reference takes at
// field_values == NULL is formed by the library's reductions.
//
// Optional members:
//   template <typename T> PSC_HD static void dvdf_delta(const T* mean, const T* d, T* out, const Couplings<T>& c);
//     out[fld] = V'_fld(mean + d) - V'_fld(mean), written so that every term
//     carries a factor of the fluctuation d.  The FP32 force pass evaluates
//     V' about the lattice mean this way (the k = 0 part V'(mean) is added in
//     FP64), so that FP32 rounding scales with the fluctuations.  Without it
//     the difference is taken in FP64 (dvdf_delta_default below).
//   template <typename T> PSC_HD static T dV_dcoupling(int k, const T* f, const Couplings<T>& c);
//   template <typename T> PSC_HD static T d2V_dfdcoupling(int k, int fld, const T* f, const Couplings<T>& c);
//   PSC_HD static constexpr bool has_coupling(int k);
//     derivatives of V and of V'_fld with respect to coupling k (the
//     PSC_COUPLING_* index of include/psc.h); needed only by psc_adjoint.
#pragma once

#include <cstddef>
#include <type_traits>

#ifdef __HIPCC__
#define PSC_HD __host__ __device__
#else
#define PSC_HD
#endif

namespace psc {

constexpr int kMaxTerms = 4;   // potential terms a model may have

// The couplings of a member in the field precision: the fields of
// psc_model (include/psc.h) in order.
template <typename T>
struct Couplings {
    T m2, g2, lambda, c[5];
    Couplings() = default;
    template <class P>
    PSC_HD explicit Couplings(const P& p)
        : m2((T)p.m2), g2((T)p.g2), lambda((T)p.lambda), c{(T)p.c[0], (T)p.c[1], (T)p.c[2], (T)p.c[3], (T)p.c[4]} {}
};

template <class M, typename T, class = void>
struct HasDvdfDelta : std::false_type {};
template <class M, typename T>
struct HasDvdfDelta<M, T,
                    std::void_t<decltype(M::dvdf_delta((const T*)nullptr, (const T*)nullptr, (T*)nullptr,
                                                       std::declval<const Couplings<T>&>()))>>
    : std::true_type {};

// V'(mean + d) - V'(mean) per field: the model's own form, or the difference
// of two FP64 evaluations.
template <class M, typename T>
PSC_HD inline void dvdf_delta(const T* mean, const T* d, T* out, const Couplings<T>& c) {
    if constexpr (HasDvdfDelta<M, T>::value) {
        M::dvdf_delta(mean, d, out, c);
    } else {
        const Couplings<double> cd(c);
        const double x[2] = {(double)mean[0] + (double)d[0], (double)mean[1] + (double)d[1]};
        const double m[2] = {(double)mean[0], (double)mean[1]};
        for (int q = 0; q < 2; ++q) out[q] = (T)(M::template dvdf<double>(q, x, cd) - M::template dvdf<double>(q, m, cd));
    }
}

template <class M, class = void>
struct HasCouplingDerivs : std::false_type {};
template <class M>
struct HasCouplingDerivs<M, std::void_t<decltype(M::template dV_dcoupling<double>(
                                0, (const double*)nullptr, std::declval<const Couplings<double>&>()))>>
    : std::true_type {};

// Coupling derivatives for the adjoint: the model's, or none (the adjoint
// then refuses the coupling).
template <class M>
PSC_HD inline bool has_coupling(int k) {
    if constexpr (HasCouplingDerivs<M>::value) return M::has_coupling(k);
    else return false;
}
template <class M, typename T>
PSC_HD inline T dV_dcoupling(int k, const T* f, const Couplings<T>& c) {
    if constexpr (HasCouplingDerivs<M>::value) return M::template dV_dcoupling<T>(k, f, c);
    else return (T)0;
}
template <class M, typename T>
PSC_HD inline T d2V_dfdcoupling(int k, int fld, const T* f, const Couplings<T>& c) {
    if constexpr (HasCouplingDerivs<M>::value) return M::template d2V_dfdcoupling<T>(k, fld, f, c);
    else return (T)0;
}

// Sum of the potential terms at one point.
template <class M, typename T>
PSC_HD inline T potential_sum(int nflds, const T* f, const Couplings<T>& c) {
    const int n = M::num_potential_terms(nflds);
    T v = (T)0;
#pragma unroll
    for (int t = 0; t < kMaxTerms; ++t)   // unrolled: each term's switch folds
        if (t < n) v += M::template potential_energy<T>(t, f, c);
    return v;
}

}  // namespace psc

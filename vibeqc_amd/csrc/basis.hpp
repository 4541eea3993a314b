// Host-side basis handling: normalisation, public AO transforms and the
// Hermite term tables the integral kernels consume.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "generativeqc/generativeqc.h"

namespace vqc {

constexpr int kMaxOrbitalL = 3;    // kernels are instantiated for s..f orbital shells
constexpr int kMaxAuxiliaryL = 3;  // and s..f auxiliary shells
constexpr int kMaxPublicL = 4;     // normalisation/transforms are defined through g

struct Atom {
  int z;
  double xyz[3];
};

struct Shell {
  uint32_t atom;
  uint32_t l;
  std::vector<double> exps;
  std::vector<double> coefs;  // radial factor and contraction scale folded in
};

struct System {
  std::vector<Atom> atoms;
  std::vector<Shell> shells;
  int32_t charge = 0;
  uint32_t multiplicity = 1;
  int32_t representation = GENERATIVEQC_BASIS_CARTESIAN;
  int electrons = 0;

  int functions(uint32_t l) const;
  std::size_t nbf() const;
  double nuclear_repulsion() const;
};

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// CCA Cartesian order: lx descending, then lz ascending.
std::vector<std::array<int, 3>> cartesian_components(int l);

// Validate and normalise a descriptor (semantics of src/molecule/basis.cpp:186-282).
// Returns a status; `detail` receives a message on failure.
generativeqc_status make_system(const generativeqc_system_descriptor& d, System& out, std::string& detail);

// Rows: public AOs of one shell (2l+1 real-spherical in libcint m order, or the
// Cartesian components); columns: raw Cartesian monomials x^i y^j z^k in CCA order.
// The per-component normalisation 1/sqrt((2i-1)!!(2j-1)!!(2k-1)!!) is folded in.
std::vector<double> ao_transform(int l, int representation);

// One term of the Hermite expansion of an auxiliary function:
//   chi(r) = sum_terms coef * (2g)^(-pow) * d^t/dCx d^u/dCy d^v/dCz exp(-g |r-C|^2),
// with the (-1)^(t+u+v) of the Coulomb ket already included in coef.
struct HermiteTerm {
  int32_t t, u, v, pow;
  double coef;
};
// Terms of public function `m` of an auxiliary shell of angular momentum l.
std::vector<HermiteTerm> auxiliary_terms(int l, int representation, int m);

}  // namespace vqc

// Basis normalisation and AO transform tables (host side).
//
// Conventions follow the reference so that AO order and values are
// interchangeable with it:
//   * primitive radial factor (2a/pi)^{3/4} (4a)^{l/2} and contraction
//     renormalisation: src/molecule/basis.cpp:13-22,254-282;
//   * CCA Cartesian order: basis.cpp:33-45;
//   * libcint real-spherical order and coefficients for d and f acting on
//     individually normalised Cartesian AOs: basis.cpp:50-96.
#include "basis.hpp"

#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>
#include <tuple>

#include "common.hpp"

namespace vqc {
namespace {

double odd_double_factorial(int k) {
  double v = 1.0;
  for (int f = 1; f < 2 * k; f += 2) v *= f;
  return v;
}

double component_normalisation(const std::array<int, 3>& c) {
  return 1.0 / std::sqrt(odd_double_factorial(c[0]) * odd_double_factorial(c[1]) *
                         odd_double_factorial(c[2]));
}

int component_index(int l, int lx, int ly, int lz) {
  const auto comps = cartesian_components(l);
  for (std::size_t i = 0; i < comps.size(); ++i)
    if (comps[i][0] == lx && comps[i][1] == ly && comps[i][2] == lz) return static_cast<int>(i);
  throw std::logic_error("Cartesian component not found");
}

// Rows of real-spherical functions over individually normalised Cartesian AOs.
std::vector<double> spherical_over_normalised(int l) {
  const int nc = ncart(l);
  std::vector<double> t(static_cast<std::size_t>(2 * l + 1) * nc, 0.0);
  auto set = [&](int m, int lx, int ly, int lz, double v) {
    t[static_cast<std::size_t>(m) * nc + component_index(l, lx, ly, lz)] = v;
  };
  if (l == 0) {
    set(0, 0, 0, 0, 1.0);
  } else if (l == 1) {
    set(0, 1, 0, 0, 1.0);
    set(1, 0, 1, 0, 1.0);
    set(2, 0, 0, 1, 1.0);
  } else if (l == 2) {
    const double h3 = std::sqrt(3.0) / 2.0;
    set(0, 1, 1, 0, 1.0);
    set(1, 0, 1, 1, 1.0);
    set(2, 2, 0, 0, -0.5);
    set(2, 0, 2, 0, -0.5);
    set(2, 0, 0, 2, 1.0);
    set(3, 1, 0, 1, 1.0);
    set(4, 2, 0, 0, h3);
    set(4, 0, 2, 0, -h3);
  } else if (l == 3) {
    const double a = 3.0 / std::sqrt(8.0), b = std::sqrt(5.0 / 8.0), c = std::sqrt(3.0 / 40.0);
    const double d = std::sqrt(3.0 / 8.0), e = std::sqrt(6.0 / 5.0), f = 3.0 / std::sqrt(20.0);
    const double g = std::sqrt(3.0) / 2.0;
    set(0, 2, 1, 0, a);
    set(0, 0, 3, 0, -b);
    set(1, 1, 1, 1, 1.0);
    set(2, 2, 1, 0, -c);
    set(2, 0, 3, 0, -d);
    set(2, 0, 1, 2, e);
    set(3, 2, 0, 1, -f);
    set(3, 0, 2, 1, -f);
    set(3, 0, 0, 3, 1.0);
    set(4, 3, 0, 0, -d);
    set(4, 1, 2, 0, -c);
    set(4, 1, 0, 2, e);
    set(5, 2, 0, 1, g);
    set(5, 0, 2, 1, -g);
    set(6, 3, 0, 0, b);
    set(6, 1, 2, 0, -a);
  } else {
    throw NotImplemented("real-spherical transforms are implemented for l <= 3");
  }
  return t;
}

}  // namespace

std::vector<std::array<int, 3>> cartesian_components(int l) {
  std::vector<std::array<int, 3>> out;
  for (int lx = l; lx >= 0; --lx)
    for (int lz = 0; lz <= l - lx; ++lz) out.push_back({lx, l - lx - lz, lz});
  return out;
}

int System::functions(uint32_t l) const {
  return representation == GENERATIVEQC_BASIS_SPHERICAL ? static_cast<int>(2 * l + 1)
                                                        : ncart(static_cast<int>(l));
}

std::size_t System::nbf() const {
  std::size_t n = 0;
  for (const auto& s : shells) n += static_cast<std::size_t>(functions(s.l));
  return n;
}

double System::nuclear_repulsion() const {
  double e = 0.0;
  for (std::size_t i = 0; i < atoms.size(); ++i)
    for (std::size_t j = 0; j < i; ++j) {
      const double dx = atoms[i].xyz[0] - atoms[j].xyz[0];
      const double dy = atoms[i].xyz[1] - atoms[j].xyz[1];
      const double dz = atoms[i].xyz[2] - atoms[j].xyz[2];
      e += static_cast<double>(atoms[i].z) * atoms[j].z / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  return e;
}

generativeqc_status make_system(const generativeqc_system_descriptor& d, System& out, std::string& detail) {
  if (!d.atoms || d.atom_count == 0) {
    detail = "a system requires at least one atom";
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  }
  if ((d.shell_count && !d.shells) || (d.primitive_count && !d.primitives)) {
    detail = "shell and primitive arrays must be non-null when counted";
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  }
  System s;
  int64_t nuclear = 0;
  for (uint32_t i = 0; i < d.atom_count; ++i) {
    const auto& a = d.atoms[i];
    if (a.atomic_number <= 0 || a.atomic_number > 118) {
      detail = "atomic numbers must lie in [1, 118]";
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    }
    if (!std::isfinite(a.x) || !std::isfinite(a.y) || !std::isfinite(a.z)) {
      detail = "atom coordinates must be finite Bohr values";
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    }
    s.atoms.push_back({a.atomic_number, {a.x, a.y, a.z}});
    nuclear += a.atomic_number;
  }
  const int64_t electrons = nuclear - static_cast<int64_t>(d.charge);
  if (electrons <= 0 || electrons > std::numeric_limits<int>::max()) {
    detail = "active-electron population is outside the positive native integer range";
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  }
  if (d.multiplicity == 0) {
    detail = "multiplicity must be positive";
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  }
  if (d.basis_representation != GENERATIVEQC_BASIS_CARTESIAN &&
      d.basis_representation != GENERATIVEQC_BASIS_SPHERICAL) {
    detail = "basis representation must be Cartesian or spherical";
    return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
  }
  s.charge = d.charge;
  s.multiplicity = d.multiplicity;
  s.representation = d.basis_representation;
  s.electrons = static_cast<int>(electrons);
  for (uint32_t i = 0; i < d.shell_count; ++i) {
    const auto& sh = d.shells[i];
    if (sh.atom_index >= d.atom_count) {
      detail = "shell atom index is out of range";
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    }
    if (sh.angular_momentum > static_cast<uint32_t>(kMaxPublicL)) {
      detail = "shell angular momentum above g is not supported";
      return GENERATIVEQC_STATUS_NOT_IMPLEMENTED;
    }
    if (sh.primitive_count == 0 ||
        static_cast<uint64_t>(sh.primitive_offset) + sh.primitive_count > d.primitive_count) {
      detail = "each shell requires at least one in-range primitive";
      return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
    }
    Shell out_shell;
    out_shell.atom = sh.atom_index;
    out_shell.l = sh.angular_momentum;
    for (uint32_t k = 0; k < sh.primitive_count; ++k) {
      const auto& p = d.primitives[sh.primitive_offset + k];
      if (!(p.exponent > 0.0) || !std::isfinite(p.exponent) || !std::isfinite(p.coefficient)) {
        detail = "primitive exponents must be positive finite and coefficients finite";
        return GENERATIVEQC_STATUS_INVALID_ARGUMENT;
      }
      out_shell.exps.push_back(p.exponent);
      out_shell.coefs.push_back(p.coefficient);
    }
    // One contraction scale is valid for every AO of the shell because the
    // overlap of two same-centre normalised primitives depends only on l.
    const double l = static_cast<double>(out_shell.l);
    double norm2 = 0.0;
    for (std::size_t a = 0; a < out_shell.exps.size(); ++a)
      for (std::size_t b = 0; b < out_shell.exps.size(); ++b) {
        const double ea = out_shell.exps[a], eb = out_shell.exps[b];
        const double ratio = 2.0 * std::sqrt(ea * eb) / (ea + eb);
        norm2 += out_shell.coefs[a] * out_shell.coefs[b] * std::pow(ratio, l + 1.5);
      }
    if (!(norm2 > 0.0) || !std::isfinite(norm2)) {
      detail = "contracted shell has an invalid normalization";
      return GENERATIVEQC_STATUS_NUMERICAL_FAILURE;
    }
    const double scale = 1.0 / std::sqrt(norm2);
    for (std::size_t k = 0; k < out_shell.exps.size(); ++k) {
      const double e = out_shell.exps[k];
      const bool nonzero = out_shell.coefs[k] != 0.0;
      out_shell.coefs[k] *= scale * std::pow(2.0 * e / M_PI, 0.75) * std::pow(4.0 * e, 0.5 * l);
      if (!std::isfinite(out_shell.coefs[k]) || (nonzero && out_shell.coefs[k] == 0.0)) {
        detail = "normalized primitive coefficient is outside nonzero FP64 range";
        return GENERATIVEQC_STATUS_NUMERICAL_FAILURE;
      }
    }
    s.shells.push_back(std::move(out_shell));
  }
  out = std::move(s);
  return GENERATIVEQC_STATUS_SUCCESS;
}

std::vector<double> ao_transform(int l, int representation) {
  const auto comps = cartesian_components(l);
  const int nc = ncart(l);
  if (representation == GENERATIVEQC_BASIS_CARTESIAN || l < 2) {
    std::vector<double> t(static_cast<std::size_t>(nc) * nc, 0.0);
    for (int i = 0; i < nc; ++i) t[static_cast<std::size_t>(i) * nc + i] = component_normalisation(comps[i]);
    return t;
  }
  auto t = spherical_over_normalised(l);
  for (int m = 0; m < 2 * l + 1; ++m)
    for (int c = 0; c < nc; ++c) t[static_cast<std::size_t>(m) * nc + c] *= component_normalisation(comps[c]);
  return t;
}

std::vector<HermiteTerm> auxiliary_terms(int l, int representation, int m) {
  // 1D Hermite expansion of x^i exp(-g x^2) about its own centre:
  //   x^i e^{-g x^2} = sum_tau h(i,tau) (2g)^{-(i+tau)/2} d^tau/dCx^tau e^{-g x^2},
  // h(0,0) = 1, h(i+1,tau) = h(i,tau-1) + (tau+1) h(i,tau+1).
  constexpr int kMax = 8;
  double h[kMax + 1][kMax + 2] = {};
  h[0][0] = 1.0;
  for (int i = 0; i < kMax; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = 0.0;
      if (t > 0) v += h[i][t - 1];
      if (t + 1 <= i) v += (t + 1) * h[i][t + 1];
      h[i + 1][t] = v;
    }
  const auto comps = cartesian_components(l);
  const auto rows = ao_transform(l, representation);
  const int nc = ncart(l);
  std::map<std::tuple<int, int, int>, double> acc;
  for (int c = 0; c < nc; ++c) {
    const double poly = rows[static_cast<std::size_t>(m) * nc + c];
    if (poly == 0.0) continue;
    const int i = comps[c][0], j = comps[c][1], k = comps[c][2];
    for (int t = i % 2; t <= i; t += 2)
      for (int u = j % 2; u <= j; u += 2)
        for (int v = k % 2; v <= k; v += 2) {
          const double sign = ((t + u + v) & 1) ? -1.0 : 1.0;
          acc[{t, u, v}] += sign * poly * h[i][t] * h[j][u] * h[k][v];
        }
  }
  double largest = 0.0;
  for (const auto& kv : acc) largest = std::max(largest, std::abs(kv.second));
  std::vector<HermiteTerm> out;
  for (const auto& [key, coef] : acc) {
    // Harmonic (real-spherical) polynomials cancel every lower-order Hermite
    // term exactly; drop the rounding residue of that cancellation.
    if (std::abs(coef) <= 1e-13 * largest) continue;
    const auto [t, u, v] = key;
    out.push_back({t, u, v, (l + t + u + v) / 2, coef});
  }
  return out;
}

}  // namespace vqc

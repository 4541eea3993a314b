// Gaussian integral kernels for gfx950 (FP64 VALU).
//
// Mathematics (McMurchie-Davidson), the same integrals the reference evaluates
// in src/integrals/s_integrals.cpp (Boys :126-156, Hermite E :171-211,
// Coulomb R :225-271, primitive ERI :472-524), organised for the GPU:
//
//  * one 64-lane wavefront per (shell pair, 64 auxiliary functions of one
//    angular momentum); each lane owns one auxiliary function, so the packed
//    output row segment A[pair][P0:P0+64] is written with one coalesced store
//    per AO pair;
//  * the bra is evaluated as (e0| for e = la..la+lb and moved to (ab| by the
//    horizontal recurrence (x-B)^b = sum_k C(b,k) (x-A)^k (A-B)^(b-k), so the
//    Hermite E tables stay one-index;
//  * the auxiliary function is expanded in Hermite Gaussians on its own
//    centre; for real-spherical functions only the top degree survives
//    (Hobson's theorem), so the ket costs ncart(lc) multiply-adds per R;
//  * R_{tuv} is built in place, highest degree first, in a compactly indexed
//    register array; every index is a compile-time constant after unrolling;
//  * the Boys function is a tabulated 8-term Taylor step plus downward
//    recursion below T = 36 and the asymptotic form with upward recursion
//    above, branch-free across the recursion.
#include <cmath>

#include "integrals_kernels.hpp"

namespace vqc {
using namespace detail;
namespace {


void dispatch_pair(int la, int lb, int lc, bool cart, const ThreeCenterArgs& args, int ntask, hipStream_t s) {
  switch (la * 10 + lb) {
    case 0: dispatch_3c_00(lc, cart, args, ntask, s); break;
    case 10: dispatch_3c_10(lc, cart, args, ntask, s); break;
    case 11: dispatch_3c_11(lc, cart, args, ntask, s); break;
    case 20: dispatch_3c_20(lc, cart, args, ntask, s); break;
    case 21: dispatch_3c_21(lc, cart, args, ntask, s); break;
    case 22: dispatch_3c_22(lc, cart, args, ntask, s); break;
    case 30: dispatch_3c_30(lc, cart, args, ntask, s); break;
    case 31: dispatch_3c_31(lc, cart, args, ntask, s); break;
    case 32: dispatch_3c_32(lc, cart, args, ntask, s); break;
    case 33: dispatch_3c_33(lc, cart, args, ntask, s); break;
    default: throw NotImplemented("orbital angular momentum above f");
  }
}

void dispatch_metric(int la, int lc, bool cart, const ThreeCenterArgs& args, int ntask, hipStream_t s) {
  switch (la) {
    case 0: dispatch_metric_0(lc, cart, args, ntask, s); break;
    case 1: dispatch_metric_1(lc, cart, args, ntask, s); break;
    case 2: dispatch_metric_2(lc, cart, args, ntask, s); break;
    case 3: dispatch_metric_3(lc, cart, args, ntask, s); break;
    default: throw NotImplemented("auxiliary angular momentum above f");
  }
}

void dispatch_1e(int la, int lb, const OneElectronArgs& args, hipStream_t s) {
  switch (la * 10 + lb) {
    case 0: dispatch_1e_00(args, s); break;
    case 10: dispatch_1e_10(args, s); break;
    case 11: dispatch_1e_11(args, s); break;
    case 20: dispatch_1e_20(args, s); break;
    case 21: dispatch_1e_21(args, s); break;
    case 22: dispatch_1e_22(args, s); break;
    case 30: dispatch_1e_30(args, s); break;
    case 31: dispatch_1e_31(args, s); break;
    case 32: dispatch_1e_32(args, s); break;
    case 33: dispatch_1e_33(args, s); break;
    default: throw NotImplemented("orbital angular momentum above f");
  }
}

long double boys_series(int m, long double t) {
  long double term = 1.0L / (2.0L * m + 1.0L);
  long double sum = term;
  for (int k = 1; k < 4000; ++k) {
    term *= 2.0L * t / (2.0L * m + 2.0L * k + 1.0L);
    sum += term;
    if (term < sum * 1e-21L) break;
  }
  return std::exp(-t) * sum;
}

// Per-pair task lists grouped by the (larger l, smaller l) class, for the
// shell pairs (i, j <= i) with i in [first, last).
std::vector<std::vector<int2>> pair_tasks(const System& sys, int first = 0, int last = -1) {
  std::vector<std::vector<int2>> classes(100);
  const int n = last < 0 ? static_cast<int>(sys.shells.size()) : last;
  for (int i = first; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const int li = sys.shells[i].l, lj = sys.shells[j].l;
      const int2 t = li >= lj ? make_int2(i, j) : make_int2(j, i);
      classes[std::max(li, lj) * 10 + std::min(li, lj)].push_back(t);
    }
  return classes;
}

}  // namespace

const std::vector<double>& boys_table_host() {
  // Every plan uploads the same table; building it costs milliseconds of
  // long-double series per plan, so it is built once (thread-safe static).
  static const std::vector<double> table = [] {
    std::vector<double> t(static_cast<std::size_t>(kBoysRows) * kBoysCols);
    for (int k = 0; k < kBoysRows; ++k)
      for (int m = 0; m < kBoysCols; ++m)
        t[static_cast<std::size_t>(k) * kBoysCols + m] =
            static_cast<double>(boys_series(m, k * static_cast<long double>(kBoysStep)));
    return t;
  }();
  return table;
}

void upload_basis(const System& sys, DeviceBasis& out, bool auxiliary_offsets, hipStream_t s) {
  out.representation = sys.representation;
  out.nshell = static_cast<int>(sys.shells.size());
  out.host_shells.clear();
  std::vector<double> exps, coefs;
  int off = 0;
  for (const auto& sh : sys.shells) {
    DevShell d{};
    d.x = sys.atoms[sh.atom].xyz[0];
    d.y = sys.atoms[sh.atom].xyz[1];
    d.z = sys.atoms[sh.atom].xyz[2];
    d.l = static_cast<int>(sh.l);
    d.prim_off = static_cast<int>(exps.size());
    d.nprim = static_cast<int>(sh.exps.size());
    d.fn_off = off;
    off += sys.functions(sh.l);
    exps.insert(exps.end(), sh.exps.begin(), sh.exps.end());
    coefs.insert(coefs.end(), sh.coefs.begin(), sh.coefs.end());
    out.host_shells.push_back(d);
  }
  out.nfunctions = static_cast<std::size_t>(off);
  if (auxiliary_offsets) {
    // Internal auxiliary order: grouped by l, then shell order, then component.
    int next = 0;
    for (int l = 0; l <= kMaxAuxiliaryL + 1; ++l)
      for (auto& d : out.host_shells)
        if (d.l == l) {
          d.fn_off = next;
          next += sys.functions(l);
        }
  }
  std::vector<double> tr;
  out.transform_off.assign(kMaxPublicL + 1, -1);
  for (int l = 0; l <= kMaxAuxiliaryL; ++l) {
    out.transform_off[l] = static_cast<int>(tr.size());
    const auto rows = ao_transform(l, sys.representation);
    // Pad to ncart(l) rows so kernels can index rows up to ncart(l).
    std::vector<double> padded(static_cast<std::size_t>(ncart(l)) * ncart(l), 0.0);
    std::copy(rows.begin(), rows.end(), padded.begin());
    tr.insert(tr.end(), padded.begin(), padded.end());
  }
  out.shells.upload(out.host_shells, s);
  out.exps.upload(exps, s);
  out.coefs.upload(coefs, s);
  out.transform.upload(tr, s);
  VQC_HIP(hipStreamSynchronize(s));
}

void build_auxiliary_layout(const System& aux, AuxiliaryLayout& out, hipStream_t s) {
  upload_basis(aux, out.basis, true, s);
  const bool cart = aux.representation == GENERATIVEQC_BASIS_CARTESIAN;
  std::vector<double> hermite;
  std::vector<int> coef_off(static_cast<std::size_t>(kMaxAuxiliaryL + 1) * 16, -1);
  for (int l = 0; l <= kMaxAuxiliaryL; ++l) {
    const int nfun = aux.functions(l);
    const int nh = hermite_count(l, cart && l >= 2);
    for (int m = 0; m < nfun; ++m) {
      std::vector<double> dense(nh, 0.0);
      for (const auto& term : auxiliary_terms(l, aux.representation, m)) {
        const int d = term.t + term.u + term.v;
        int k = 0;
        bool found = false;
        for (int dd = l; dd >= 0 && !found; dd -= 2) {
          if (dd == d) {
            k += term.t * (d + 1) - term.t * (term.t - 1) / 2 + term.u;
            found = true;
          } else {
            k += nc(dd);
          }
          if (!(cart && l >= 2) && !found) break;
        }
        if (!found || k >= nh) throw std::logic_error("auxiliary Hermite term outside the kernel table");
        dense[k] = term.coef;
      }
      coef_off[l * 16 + m] = static_cast<int>(hermite.size());
      hermite.insert(hermite.end(), dense.begin(), dense.end());
    }
  }
  std::vector<AuxFunction> fns(out.basis.nfunctions);
  out.public_of_internal.assign(out.basis.nfunctions, 0);
  out.class_begin.assign(kMaxAuxiliaryL + 2, 0);
  int pub = 0;
  for (int sidx = 0; sidx < out.basis.nshell; ++sidx) {
    const auto& d = out.basis.host_shells[sidx];
    if (d.l > kMaxAuxiliaryL) throw NotImplemented("auxiliary angular momentum above f");
    const int nfun = aux.functions(d.l);
    for (int m = 0; m < nfun; ++m) {
      const int idx = d.fn_off + m;
      fns[idx] = AuxFunction{sidx, coef_off[d.l * 16 + m], idx, 0};
      out.public_of_internal[idx] = pub + m;
    }
    pub += nfun;
  }
  // class_begin[l] = first internal index of class l; class_begin[l+1] its end.
  int next = 0;
  for (int l = 0; l <= kMaxAuxiliaryL; ++l) {
    out.class_begin[l] = next;
    for (const auto& d : out.basis.host_shells)
      if (d.l == l) next += aux.functions(l);
  }
  out.class_begin[kMaxAuxiliaryL + 1] = next;
  out.functions.upload(fns, s);
  out.hermite.upload(hermite, s);
  VQC_HIP(hipStreamSynchronize(s));
}

void compute_one_electron(const System& sys, const DeviceBasis& basis, const double* boys, double* S,
                          double* T, double* V, hipStream_t s) {
  std::vector<double4> atoms;
  for (const auto& a : sys.atoms) atoms.push_back(make_double4(a.xyz[0], a.xyz[1], a.xyz[2], a.z));
  DeviceArray<double4> datoms;
  datoms.upload(atoms, s);
  const auto classes = pair_tasks(sys);
  std::vector<DeviceArray<int2>> keep;
  for (int key = 0; key < 100; ++key) {
    if (classes[key].empty()) continue;
    const int la = key / 10, lb = key % 10;
    keep.emplace_back();
    keep.back().upload(classes[key], s);
    OneElectronArgs args{};
    args.shells = basis.shells.data();
    args.exps = basis.exps.data();
    args.coefs = basis.coefs.data();
    args.tasks = keep.back().data();
    args.ntask = static_cast<int>(classes[key].size());
    args.atoms = datoms.data();
    args.natom = static_cast<int>(atoms.size());
    args.ta = basis.transform.data() + basis.transform_off[la];
    args.tb = basis.transform.data() + basis.transform_off[lb];
    args.nfa = sys.functions(la);
    args.nfb = sys.functions(lb);
    args.boys = boys;
    args.S = S;
    args.T = T;
    args.V = V;
    args.n = static_cast<int>(basis.nfunctions);
    args.sub = std::max(2, one_electron_lanes(args.natom, one_electron_pair_bytes(la, lb)));
    dispatch_1e(la, lb, args, s);
  }
  VQC_HIP(hipStreamSynchronize(s));
}

void compute_metric(const AuxiliaryLayout& aux, const double* boys, double* M, hipStream_t s) {
  const auto& b = aux.basis;
  const bool cart = b.representation == GENERATIVEQC_BASIS_CARTESIAN;
  const std::size_t naux = b.nfunctions;
  std::vector<DeviceArray<int2>> keep;
  for (int la = 0; la <= kMaxAuxiliaryL; ++la) {
    std::vector<int2> tasks;
    for (int i = 0; i < b.nshell; ++i)
      if (b.host_shells[i].l == la) tasks.push_back(make_int2(i, i));
    if (tasks.empty()) continue;
    keep.emplace_back();
    keep.back().upload(tasks, s);
    for (int lc = 0; lc <= kMaxAuxiliaryL; ++lc) {
      const int begin = aux.class_begin[lc], end = aux.class_begin[lc + 1];
      if (end <= begin) continue;
      ThreeCenterArgs args{};
      args.bra_shells = b.shells.data();
      args.bra_exps = b.exps.data();
      args.bra_coefs = b.coefs.data();
      args.aux_shells = b.shells.data();
      args.aux_exps = b.exps.data();
      args.aux_coefs = b.coefs.data();
      args.tasks = keep.back().data();
      args.fns = aux.functions.data() + begin;
      args.nfn = end - begin;
      args.hermite = aux.hermite.data();
      args.ta = b.transform.data() + b.transform_off[la];
      args.tb = b.transform.data() + b.transform_off[0];
      args.nfa = b.representation == GENERATIVEQC_BASIS_SPHERICAL ? 2 * la + 1 : ncart(la);
      args.nfb = 1;
      args.boys = boys;
      args.out = M;
      args.ld = naux;
      args.prim_cutoff = 0.0;
      dispatch_metric(la, lc, cart && lc >= 2, args, static_cast<int>(tasks.size()), s);
    }
  }
  VQC_HIP(hipStreamSynchronize(s));
}

void compute_three_center(const System& orb, const DeviceBasis& basis, const AuxiliaryLayout& aux,
                          const double* boys, double* A, std::size_t ld, double screening, hipStream_t s,
                          int shell_begin, int shell_end) {
  const auto& b = aux.basis;
  const bool cart = b.representation == GENERATIVEQC_BASIS_CARTESIAN;
  if (shell_end < 0) shell_end = basis.nshell;
  const std::size_t f0 = shell_begin < basis.nshell ? basis.host_shells[shell_begin].fn_off : basis.nfunctions;
  const std::size_t f1 = shell_end < basis.nshell ? basis.host_shells[shell_end].fn_off : basis.nfunctions;
  const std::size_t row0 = pair_count(f0);
  VQC_HIP(hipMemsetAsync(A, 0, (pair_count(f1) - row0) * ld * sizeof(double), s));
  const auto classes = pair_tasks(orb, shell_begin, shell_end);
  std::vector<DeviceArray<int2>> keep;
  for (int key = 0; key < 100; ++key) {
    if (classes[key].empty()) continue;
    const int la = key / 10, lb = key % 10;
    keep.emplace_back();
    keep.back().upload(classes[key], s);
    for (int lc = 0; lc <= kMaxAuxiliaryL; ++lc) {
      const int begin = aux.class_begin[lc], end = aux.class_begin[lc + 1];
      if (end <= begin) continue;
      ThreeCenterArgs args{};
      args.bra_shells = basis.shells.data();
      args.bra_exps = basis.exps.data();
      args.bra_coefs = basis.coefs.data();
      args.aux_shells = b.shells.data();
      args.aux_exps = b.exps.data();
      args.aux_coefs = b.coefs.data();
      args.tasks = keep.back().data();
      args.fns = aux.functions.data() + begin;
      args.nfn = end - begin;
      args.hermite = aux.hermite.data();
      args.ta = basis.transform.data() + basis.transform_off[la];
      args.tb = basis.transform.data() + basis.transform_off[lb];
      args.nfa = orb.functions(la);
      args.nfb = orb.functions(lb);
      args.boys = boys;
      args.out = A;
      args.ld = ld;
      args.prim_cutoff = screening;
      args.row0 = row0;
      dispatch_pair(la, lb, lc, cart && lc >= 2, args, static_cast<int>(classes[key].size()), s);
    }
  }
  VQC_HIP(hipStreamSynchronize(s));
}

}  // namespace vqc

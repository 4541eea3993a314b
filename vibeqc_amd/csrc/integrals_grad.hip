// Host driver of the nuclear-derivative integral kernels (integrals_grad.hpp):
// launches every shell class with its response weights and reduces the
// per-task / per-function partial derivatives to atoms in a fixed order, so
// the gradient is bitwise reproducible run to run.
#include <algorithm>

#include "integrals_grad.hpp"
#include "integrals_gradient.hpp"

namespace vqc {
using namespace detail;
namespace {

void dispatch_pair_grad(int la, int lb, int lc, bool cart, const ThreeCenterGradArgs& args, int nchunk, hipStream_t s) {
  switch (la * 10 + lb) {
    case 0: dispatch_3c_grad_00(lc, cart, args, nchunk, s); break;
    case 10: dispatch_3c_grad_10(lc, cart, args, nchunk, s); break;
    case 11: dispatch_3c_grad_11(lc, cart, args, nchunk, s); break;
    case 20: dispatch_3c_grad_20(lc, cart, args, nchunk, s); break;
    case 21: dispatch_3c_grad_21(lc, cart, args, nchunk, s); break;
    case 22: dispatch_3c_grad_22(lc, cart, args, nchunk, s); break;
    case 30: dispatch_3c_grad_30(lc, cart, args, nchunk, s); break;
    case 31: dispatch_3c_grad_31(lc, cart, args, nchunk, s); break;
    case 32: dispatch_3c_grad_32(lc, cart, args, nchunk, s); break;
    case 33: dispatch_3c_grad_33(lc, cart, args, nchunk, s); break;
    default: throw NotImplemented("orbital angular momentum above f");
  }
}

void dispatch_metric_grad(int la, int lc, bool cart, const ThreeCenterGradArgs& args, int nchunk, hipStream_t s) {
  switch (la) {
    case 0: dispatch_metric_grad_0(lc, cart, args, nchunk, s); break;
    case 1: dispatch_metric_grad_1(lc, cart, args, nchunk, s); break;
    case 2: dispatch_metric_grad_2(lc, cart, args, nchunk, s); break;
    case 3: dispatch_metric_grad_3(lc, cart, args, nchunk, s); break;
    default: throw NotImplemented("auxiliary angular momentum above f");
  }
}

void dispatch_1e_grad(int la, int lb, const OneElectronGradArgs& args, hipStream_t s) {
  switch (la * 10 + lb) {
    case 0: dispatch_1e_grad_00(args, s); break;
    case 10: dispatch_1e_grad_10(args, s); break;
    case 11: dispatch_1e_grad_11(args, s); break;
    case 20: dispatch_1e_grad_20(args, s); break;
    case 21: dispatch_1e_grad_21(args, s); break;
    case 22: dispatch_1e_grad_22(args, s); break;
    case 30: dispatch_1e_grad_30(args, s); break;
    case 31: dispatch_1e_grad_31(args, s); break;
    case 32: dispatch_1e_grad_32(args, s); break;
    case 33: dispatch_1e_grad_33(args, s); break;
    default: throw NotImplemented("orbital angular momentum above f");
  }
}

constexpr int kReduceThreads = 256;

// grad[atom] += sum over tasks t of the A part of out_ab[t][*] where the
// bra shell sits on `atom`, the B part where the ket shell does, and of
// out_c[*][f] for the functions (or nuclei) f on `atom`.  One block per atom;
// a fixed thread-to-entry map and tree reduction make the sum deterministic.
__global__ __launch_bounds__(kReduceThreads) void reduce_grad_kernel(const double* __restrict__ out_ab, int ntask,
                                                                    int nblk, const int2* __restrict__ task_atoms,
                                                                    const double* __restrict__ out_c, int nrow,
                                                                    int nfn, const int* __restrict__ fn_atoms,
                                                                    double* __restrict__ grad) {
  const int atom = blockIdx.x, tid = threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  for (int t = tid; t < ntask; t += kReduceThreads) {
    const int2 ta = task_atoms[t];
    if (ta.x != atom && ta.y != atom) continue;
    for (int b = 0; b < nblk; ++b) {
      const double* o = out_ab + (static_cast<std::size_t>(t) * nblk + b) * 6;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (ta.x == atom) s[d] += o[d];
        if (ta.y == atom) s[d] += o[3 + d];
      }
    }
  }
  const std::size_t nc_entries = static_cast<std::size_t>(nrow) * nfn;
  for (std::size_t e = tid; e < nc_entries; e += kReduceThreads) {
    if (fn_atoms[e % nfn] != atom) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] += out_c[e * 3 + d];
  }
  __shared__ double red[3][kReduceThreads];
#pragma unroll
  for (int d = 0; d < 3; ++d) red[d][tid] = s[d];
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int d = 0; d < 3; ++d) red[d][tid] += red[d][tid + w];
    __syncthreads();
  }
  if (tid < 3) grad[atom * 3 + tid] += red[tid][0];
}

// Task lists of unordered orbital shell pairs grouped by (larger l, smaller l).
std::vector<std::vector<int2>> grad_pair_tasks(const System& sys) {
  std::vector<std::vector<int2>> classes(100);
  const int n = static_cast<int>(sys.shells.size());
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const int li = sys.shells[i].l, lj = sys.shells[j].l;
      const int2 t = li >= lj ? make_int2(i, j) : make_int2(j, i);
      classes[std::max(li, lj) * 10 + std::min(li, lj)].push_back(t);
    }
  return classes;
}

struct Partials {
  DeviceArray<int2> tasks, task_atoms;
  DeviceArray<int> fn_atoms;
  DeviceArray<double> ab, c;
};

// Enough blocks to fill the GPU while each block reuses its 64 functions'
// coefficients over several tasks.
int tasks_per_block(int ntask, int nfnblk) {
  const long long blocks = static_cast<long long>(ntask) * nfnblk;
  return static_cast<int>(std::max<long long>(1, std::min<long long>(ntask, blocks / 4096)));
}

}  // namespace

void compute_three_center_gradient(const System& orb, const DeviceBasis& basis, const System& aux_sys,
                                   const AuxiliaryLayout& aux, const double* boys, const double* W, std::size_t ld,
                                   double screening, double* grad, hipStream_t s) {
  const auto& b = aux.basis;
  const bool cart = b.representation == GENERATIVEQC_BASIS_CARTESIAN;
  const int natom = static_cast<int>(orb.atoms.size());
  std::vector<int> aux_atom_of_internal(b.nfunctions);
  for (int sh = 0; sh < b.nshell; ++sh)
    for (int m = 0; m < aux_sys.functions(b.host_shells[sh].l); ++m)
      aux_atom_of_internal[b.host_shells[sh].fn_off + m] = static_cast<int>(aux_sys.shells[sh].atom);
  const auto classes = grad_pair_tasks(orb);
  std::vector<Partials> keep;
  for (int key = 0; key < 100; ++key) {
    if (classes[key].empty()) continue;
    const int la = key / 10, lb = key % 10;
    const int ntask = static_cast<int>(classes[key].size());
    std::vector<int2> atoms(ntask);
    for (int t = 0; t < ntask; ++t)
      atoms[t] = make_int2(static_cast<int>(orb.shells[classes[key][t].x].atom),
                           static_cast<int>(orb.shells[classes[key][t].y].atom));
    for (int lc = 0; lc <= kMaxAuxiliaryL; ++lc) {
      const int begin = aux.class_begin[lc], end = aux.class_begin[lc + 1];
      if (end <= begin) continue;
      const int nfn = end - begin, nfnblk = (nfn + 63) / 64;
      const int tpb = tasks_per_block(ntask, nfnblk);
      const int nchunk = (ntask + tpb - 1) / tpb;
      keep.emplace_back();
      Partials& p = keep.back();
      p.tasks.upload(classes[key], s);
      p.task_atoms.upload(atoms, s);
      p.fn_atoms.upload(std::vector<int>(aux_atom_of_internal.begin() + begin, aux_atom_of_internal.begin() + end), s);
      p.ab.resize(static_cast<std::size_t>(ntask) * nfnblk * 6);
      p.c.resize(static_cast<std::size_t>(nchunk) * nfn * 3);
      ThreeCenterGradArgs args{};
      args.bra_shells = basis.shells.data();
      args.bra_exps = basis.exps.data();
      args.bra_coefs = basis.coefs.data();
      args.aux_shells = b.shells.data();
      args.aux_exps = b.exps.data();
      args.aux_coefs = b.coefs.data();
      args.tasks = p.tasks.data();
      args.ntask = ntask;
      args.tasks_per_block = tpb;
      args.fns = aux.functions.data() + begin;
      args.nfn = nfn;
      args.hermite = aux.hermite.data();
      args.ta = basis.transform.data() + basis.transform_off[la];
      args.tb = basis.transform.data() + basis.transform_off[lb];
      args.nfa = orb.functions(la);
      args.nfb = orb.functions(lb);
      args.boys = boys;
      args.w = W;
      args.ld = ld;
      args.prim_cutoff = screening;
      args.out_ab = p.ab.data();
      args.out_c = p.c.data();
      dispatch_pair_grad(la, lb, lc, cart && lc >= 2, args, nchunk, s);
      hipLaunchKernelGGL(reduce_grad_kernel, dim3(natom), dim3(kReduceThreads), 0, s, p.ab.data(), ntask, nfnblk,
                         p.task_atoms.data(), p.c.data(), nchunk, nfn, p.fn_atoms.data(), grad);
      VQC_LAUNCH("reduce_grad_kernel");
    }
  }
  VQC_HIP(hipStreamSynchronize(s));
}

void compute_metric_gradient(const System& aux_sys, const AuxiliaryLayout& aux, const double* boys, const double* W,
                             std::size_t ld, int natom, double* grad, hipStream_t s) {
  const auto& b = aux.basis;
  const bool cart = b.representation == GENERATIVEQC_BASIS_CARTESIAN;
  std::vector<int> aux_atom_of_internal(b.nfunctions);
  for (int sh = 0; sh < b.nshell; ++sh)
    for (int m = 0; m < aux_sys.functions(b.host_shells[sh].l); ++m)
      aux_atom_of_internal[b.host_shells[sh].fn_off + m] = static_cast<int>(aux_sys.shells[sh].atom);
  std::vector<Partials> keep;
  for (int la = 0; la <= kMaxAuxiliaryL; ++la) {
    std::vector<int2> tasks, atoms;
    for (int i = 0; i < b.nshell; ++i)
      if (b.host_shells[i].l == la) {
        tasks.push_back(make_int2(i, i));
        const int at = static_cast<int>(aux_sys.shells[i].atom);
        atoms.push_back(make_int2(at, at));
      }
    if (tasks.empty()) continue;
    const int ntask = static_cast<int>(tasks.size());
    for (int lc = 0; lc <= kMaxAuxiliaryL; ++lc) {
      const int begin = aux.class_begin[lc], end = aux.class_begin[lc + 1];
      if (end <= begin) continue;
      const int nfn = end - begin, nfnblk = (nfn + 63) / 64;
      const int tpb = tasks_per_block(ntask, nfnblk);
      const int nchunk = (ntask + tpb - 1) / tpb;
      keep.emplace_back();
      Partials& p = keep.back();
      p.tasks.upload(tasks, s);
      p.task_atoms.upload(atoms, s);
      p.fn_atoms.upload(std::vector<int>(aux_atom_of_internal.begin() + begin, aux_atom_of_internal.begin() + end), s);
      p.ab.resize(static_cast<std::size_t>(ntask) * nfnblk * 6);
      p.c.resize(static_cast<std::size_t>(nchunk) * nfn * 3);
      ThreeCenterGradArgs args{};
      args.bra_shells = b.shells.data();
      args.bra_exps = b.exps.data();
      args.bra_coefs = b.coefs.data();
      args.aux_shells = b.shells.data();
      args.aux_exps = b.exps.data();
      args.aux_coefs = b.coefs.data();
      args.tasks = p.tasks.data();
      args.ntask = ntask;
      args.tasks_per_block = tpb;
      args.fns = aux.functions.data() + begin;
      args.nfn = nfn;
      args.hermite = aux.hermite.data();
      args.ta = b.transform.data() + b.transform_off[la];
      args.tb = b.transform.data() + b.transform_off[0];
      args.nfa = b.representation == GENERATIVEQC_BASIS_SPHERICAL ? 2 * la + 1 : ncart(la);
      args.nfb = 1;
      args.boys = boys;
      args.w = W;
      args.ld = ld;
      args.prim_cutoff = 0.0;
      args.out_ab = p.ab.data();
      args.out_c = p.c.data();
      dispatch_metric_grad(la, lc, cart && lc >= 2, args, nchunk, s);
      hipLaunchKernelGGL(reduce_grad_kernel, dim3(natom), dim3(kReduceThreads), 0, s, p.ab.data(), ntask, nfnblk,
                         p.task_atoms.data(), p.c.data(), nchunk, nfn, p.fn_atoms.data(), grad);
      VQC_LAUNCH("reduce_grad_kernel");
    }
  }
  VQC_HIP(hipStreamSynchronize(s));
}

void compute_one_electron_gradient(const System& sys, const DeviceBasis& basis, const double* boys,
                                   const double* dens, const double* edens, double* grad, hipStream_t s) {
  std::vector<double4> atoms;
  for (const auto& a : sys.atoms) atoms.push_back(make_double4(a.xyz[0], a.xyz[1], a.xyz[2], a.z));
  const int natom = static_cast<int>(atoms.size());
  DeviceArray<double4> datoms;
  datoms.upload(atoms, s);
  std::vector<int> identity(natom);
  for (int k = 0; k < natom; ++k) identity[k] = k;
  DeviceArray<int> nuclei;
  nuclei.upload(identity, s);
  const auto classes = grad_pair_tasks(sys);
  std::vector<Partials> keep;
  for (int key = 0; key < 100; ++key) {
    if (classes[key].empty()) continue;
    const int la = key / 10, lb = key % 10;
    const int ntask = static_cast<int>(classes[key].size());
    const int sub = one_electron_lanes(natom, one_electron_pair_bytes(la, lb) / 3 * 2);  // D and W slots
    const int nblock = one_electron_grad_blocks(ntask, sub);
    std::vector<int2> ta(ntask);
    for (int t = 0; t < ntask; ++t)
      ta[t] = make_int2(static_cast<int>(sys.shells[classes[key][t].x].atom),
                        static_cast<int>(sys.shells[classes[key][t].y].atom));
    keep.emplace_back();
    Partials& p = keep.back();
    p.tasks.upload(classes[key], s);
    p.task_atoms.upload(ta, s);
    p.ab.resize(static_cast<std::size_t>(ntask) * 6);
    p.c.resize(static_cast<std::size_t>(nblock) * natom * 3);
    OneElectronGradArgs args{};
    args.shells = basis.shells.data();
    args.exps = basis.exps.data();
    args.coefs = basis.coefs.data();
    args.tasks = p.tasks.data();
    args.ntask = ntask;
    args.atoms = datoms.data();
    args.natom = natom;
    args.ta = basis.transform.data() + basis.transform_off[la];
    args.tb = basis.transform.data() + basis.transform_off[lb];
    args.nfa = sys.functions(la);
    args.nfb = sys.functions(lb);
    args.boys = boys;
    args.dens = dens;
    args.edens = edens;
    args.n = static_cast<int>(basis.nfunctions);
    args.out_ab = p.ab.data();
    args.out_c = p.c.data();
    args.sub = sub;
    dispatch_1e_grad(la, lb, args, s);
    hipLaunchKernelGGL(reduce_grad_kernel, dim3(natom), dim3(kReduceThreads), 0, s, p.ab.data(), ntask, 1,
                       p.task_atoms.data(), p.c.data(), nblock, natom, nuclei.data(), grad);
    VQC_LAUNCH("reduce_grad_kernel");
  }
  VQC_HIP(hipStreamSynchronize(s));
}

}  // namespace vqc

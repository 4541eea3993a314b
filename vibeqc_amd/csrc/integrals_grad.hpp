// Nuclear-derivative integral kernels, contracted on the fly with response
// weights, for the DF-RHF gradient.  Shared by the per-class translation
// units integrals_g*.hip and the host driver integrals_grad.hip.
//
// Mathematics: the Gaussian derivative rule d/dA_x g_i(x - A) = 2a g_{i+1} -
// i g_{i-1} per primitive (the same rule the reference's generated derivative
// kernels use, s_integrals.cpp:1220-1304 and :1073-1199), so
//   d/dA_x (ab|P) = 2a (a+1x, b|P) - a_x (a-1x, b|P),
//   d/dB_x (ab|P) = 2b (a, b+1x|P) - b_x (a, b-1x|P),
//   d/dC   (ab|P) = -(d/dA + d/dB)          (translational invariance).
// The bra is evaluated as (e0| for every degree e in [la-1, la+lb+1] with
// three weightings (1, 2a, 2b) accumulated over primitives, and moved to the
// shifted (a'b'| by the same horizontal recurrence as the value kernel.  The
// derivative integrals are contracted at once with the Cartesian weights
// w_c = T_a^T W T_b, so nothing of size (pairs x aux x 9) is ever stored.
#pragma once

#include "integrals_kernels.hpp"

namespace vqc {
namespace detail {

struct ThreeCenterGradArgs {
  const DevShell* bra_shells;
  const double* bra_exps;
  const double* bra_coefs;
  const DevShell* aux_shells;
  const double* aux_exps;
  const double* aux_coefs;
  const int2* tasks;
  int ntask;
  int tasks_per_block;
  const AuxFunction* fns;
  int nfn;
  const double* hermite;
  const double* ta;
  const double* tb;
  int nfa, nfb;
  const double* boys;
  const double* w;  // packed W_A[p][ld] or metric W_M[P][ld] (internal auxiliary columns)
  std::size_t ld;
  double prim_cutoff;
  double* out_ab;  // [ntask][gridDim.y][6]: d/dA xyz, d/dB xyz summed over the block's 64 functions
  double* out_c;   // [gridDim.x][nfn][3]: d/dC of each function summed over the block's tasks
  int accumulate;  // add into out_ab / out_c instead of writing them
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Degree ranges of the (e0| accumulators of one orbital class: raised on A
// (a + 1_d, b), raised on B (a, b + 1_d; none for the metric) and lowered
// (a - 1_d, b and a, b - 1_d), with the entries each holds.
template <int LA, int LB, bool METRIC>
struct GradAccumulators {
  static constexpr int LAB = LA + LB, LHI = LAB + 1;
  static constexpr int UP_A = LA + 1, UP_B = LA;
  static constexpr int DN_LO = LA > 0 ? LA - 1 : 0, DN_HI = LAB - 1;
  static constexpr int NUP_A = ne_offset(UP_A, LHI + 1);
  static constexpr int NUP_B = METRIC ? 0 : ne_offset(UP_B, LHI + 1);
  static constexpr int NDN = DN_HI >= DN_LO ? ne_offset(DN_LO, DN_HI + 1) : 0;
  // Entries held by a launch that keeps the accumulators selected by `mask`
  // (1: raised on A, 2: raised on B, 4: lowered).
  static constexpr int entries(int mask) {
    return ((mask & 1) ? NUP_A : 0) + ((mask & 2) ? NUP_B : 0) + ((mask & 4) ? NDN : 0);
  }
};

// Contracted Cartesian (a b| from one lane's accumulator column in LDS
// (entries 64 doubles apart), whose lowest degree is LO:
// (x-B)^b = sum_k C(b,k) (x-A)^k (A-B)^(b-k).
template <int LO, int AX, int AY, int AZ, int BX, int BY, int BZ>
__device__ __forceinline__ double hrr_column(const double* acc, double ABx, double ABy, double ABz) {
  double v = 0.0;
  static_for<BX + 1>([&](auto KXI) {
    constexpr int kx = VQC_CI(KXI);
    static_for<BY + 1>([&](auto KYI) {
      constexpr int ky = VQC_CI(KYI);
      static_for<BZ + 1>([&](auto KZI) {
        constexpr int kz = VQC_CI(KZI);
        double w = binom(BX, kx) * binom(BY, ky) * binom(BZ, kz);
#pragma unroll
        for (int r = 0; r < BX - kx; ++r) w *= ABx;
#pragma unroll
        for (int r = 0; r < BY - ky; ++r) w *= ABy;
#pragma unroll
        for (int r = 0; r < BZ - kz; ++r) w *= ABz;
        constexpr int le = AX + AY + AZ + kx + ky + kz;
        v += w * acc[(ne_offset(LO, le) + cart_index(AX + kx, AY + ky, AZ + kz)) * 64];
      });
    });
  });
  return v;
}

// d/dA, d/dB and d/dC of (ab|P), contracted with the weights W, for the
// shell pairs of one chunk (blockIdx.x) and 64 auxiliary functions (one per
// lane).  Per bra primitive pair the auxiliary primitives are summed in
// Hermite space (as in the value kernel), contracted with the bra's Hermite
// coefficients once, and added into the (e0| accumulators, which live in LDS
// (one column per lane, so no lane reads another's entries and no barrier is
// needed).  The accumulators selected by MASK are kept; a class whose three
// accumulators exceed the LDS budget runs as several launches with disjoint
// masks that add into the same outputs.  Every register-array index is a
// compile-time constant.
template <int LA, int LB, int LC, bool METRIC, bool CART, int MASK>
__global__ __launch_bounds__(64) void three_center_grad_kernel(const ThreeCenterGradArgs a) {
  using G = GradAccumulators<LA, LB, METRIC>;
  constexpr int LAB = G::LAB, LHI = G::LHI;
  constexpr int L = LHI + LC;
  constexpr int NH = hermite_count(LC, CART);
  constexpr int NA = nc(LA), NB = nc(LB);
  constexpr bool KEEP_A = (MASK & 1) != 0, KEEP_B = !METRIC && (MASK & 2) != 0, KEEP_D = (MASK & 4) != 0 && G::NDN > 0;
  constexpr int OFF_A = 0;
  constexpr int OFF_B = KEEP_A ? G::NUP_A : 0;
  constexpr int OFF_D = OFF_B + (KEEP_B ? G::NUP_B : 0);
  constexpr int NACC = OFF_D + (KEEP_D ? G::NDN : 0);
  constexpr int LO = LA > 0 ? LA - 1 : 0;  // lowest (e0| degree any accumulator needs
  extern __shared__ double grad_acc[];  // [NACC][64]
  double* acc = grad_acc + threadIdx.x;

  const int f = blockIdx.y * 64 + threadIdx.x;
  const bool active = f < a.nfn;
  const AuxFunction fn = a.fns[active ? f : 0];
  const DevShell C = a.aux_shells[fn.shell];
  double hc[NH];
#pragma unroll
  for (int k = 0; k < NH; ++k) hc[k] = a.hermite[fn.coef_off + k];
  double gc[3] = {0.0, 0.0, 0.0};

  const int t0 = blockIdx.x * a.tasks_per_block;
  const int t1 = min(t0 + a.tasks_per_block, a.ntask);
  for (int task_id = t0; task_id < t1; ++task_id) {
    const int2 task = a.tasks[task_id];
    const DevShell A = a.bra_shells[task.x];
    const DevShell B = METRIC ? A : a.bra_shells[task.y];
#pragma unroll 1
    for (int k = 0; k < NACC; ++k) acc[k * 64] = 0.0;
    const double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
    const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
    const int nprim_b = METRIC ? 1 : B.nprim;
    for (int ia = 0; ia < A.nprim; ++ia) {
      const double ea = a.bra_exps[A.prim_off + ia];
      const double ca = a.bra_coefs[A.prim_off + ia];
      for (int ib = 0; ib < nprim_b; ++ib) {
        const double eb = METRIC ? 0.0 : a.bra_exps[B.prim_off + ib];
        const double cb = METRIC ? 1.0 : a.bra_coefs[B.prim_off + ib];
        const double p = ea + eb;
        const double ip = 1.0 / p;
        const double kab = ca * cb * exp(-ea * eb * ip * AB2);
        if (fabs(kab) < a.prim_cutoff) continue;
        const double Px = (ea * A.x + eb * B.x) * ip;
        const double Py = (ea * A.y + eb * B.y) * ip;
        const double Pz = (ea * A.z + eb * B.z) * ip;
        double ws[ntuv(LHI)];
#pragma unroll
        for (int k = 0; k < ntuv(LHI); ++k) ws[k] = 0.0;
        for (int ic = 0; ic < C.nprim; ++ic) {
          const double g = a.aux_exps[C.prim_off + ic];
          const double cg = a.aux_coefs[C.prim_off + ic];
          const double q = p + g;
          const double alpha = p * g / q;
          const double X = Px - C.x, Y = Py - C.y, Z = Pz - C.z;
          double F[L + 1];
          boys_eval<L>(alpha * (X * X + Y * Y + Z * Z), a.boys, F);
          double R[ntuv(L)];
          hermite_r<L>(alpha, X, Y, Z, F, R);
          const double pref = 2.0 * 17.493418327624862 /* pi^{5/2} */ / (p * g * sqrt(q)) * cg;
          ket_contract<LHI, LC, CART, NH, L>(hc, 0.5 / g, pref, R, ws);
        }
        double Ex[LHI + 1][LHI + 1], Ey[LHI + 1][LHI + 1], Ez[LHI + 1][LHI + 1];
        hermite_e0<LHI>(Px - A.x, 0.5 * ip, Ex);
        hermite_e0<LHI>(Py - A.y, 0.5 * ip, Ey);
        hermite_e0<LHI>(Pz - A.z, 0.5 * ip, Ez);
        const double pa = 2.0 * ea * kab, pb = 2.0 * eb * kab;
        bra_contract<LO, LHI, LHI>(Ex, Ey, Ez, ws, [&](auto LI, auto CI, double v) {
          constexpr int le = VQC_CI(LI), c = VQC_CI(CI);
          if constexpr (KEEP_A && le >= G::UP_A) acc[(OFF_A + ne_offset(G::UP_A, le) + c) * 64] += pa * v;
          if constexpr (KEEP_B && le >= G::UP_B) acc[(OFF_B + ne_offset(G::UP_B, le) + c) * 64] += pb * v;
          if constexpr (KEEP_D && le >= G::DN_LO && le <= G::DN_HI)
            acc[(OFF_D + ne_offset(G::DN_LO, le) + c) * 64] += kab * v;
        });
      }
    }

    // Cartesian weights w_c = T_a^T W T_b of this shell pair and function.
    double wc[NA][NB];
    {
      double wp[NA][NB];
#pragma unroll
      for (int ma = 0; ma < NA; ++ma)
#pragma unroll
        for (int mb = 0; mb < NB; ++mb) {
          double w = 0.0;
          if (active && ma < a.nfa && mb < a.nfb) {
            if (METRIC) {
              w = a.w[static_cast<std::size_t>(A.fn_off + ma) * a.ld + fn.index];
            } else {
              const int mu = A.fn_off + ma, nu = B.fn_off + mb;
              const std::size_t hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
              w = a.w[(hi * (hi + 1) / 2 + lo) * a.ld + fn.index] * (task.x != task.y ? 2.0 : 1.0);
            }
          }
          wp[ma][mb] = w;
        }
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
          double v = 0.0;
#pragma unroll
          for (int ma = 0; ma < NA; ++ma)
#pragma unroll
            for (int mb = 0; mb < NB; ++mb) v += a.ta[ma * NA + ia] * a.tb[mb * NB + ib] * wp[ma][mb];
          wc[ia][ib] = v;
        }
    }

    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    static_for<NA>([&](auto IA) {
      constexpr int ia = VQC_CI(IA);
      constexpr int ax = cart_lx(LA, ia), ay = cart_ly(LA, ia), az = cart_lz(LA, ia);
      static_for<NB>([&](auto IB) {
        constexpr int ib = VQC_CI(IB);
        constexpr int bx = cart_lx(LB, ib), by = cart_ly(LB, ib), bz = cart_lz(LB, ib);
        const double w = wc[ia][ib];
        static_for<3>([&](auto DI) {
          constexpr int d = VQC_CI(DI);
          constexpr int dx = d == 0, dy = d == 1, dz = d == 2;
          constexpr int a_d = d == 0 ? ax : (d == 1 ? ay : az);
          constexpr int b_d = d == 0 ? bx : (d == 1 ? by : bz);
          double v = 0.0, vb = 0.0;
          if constexpr (KEEP_A)
            v += hrr_column<G::UP_A, ax + dx, ay + dy, az + dz, bx, by, bz>(acc + OFF_A * 64, ABx, ABy, ABz);
          if constexpr (KEEP_D && a_d > 0)
            v -= a_d * hrr_column<G::DN_LO, ax - dx, ay - dy, az - dz, bx, by, bz>(acc + OFF_D * 64, ABx, ABy, ABz);
          if constexpr (KEEP_B)
            vb += hrr_column<G::UP_B, ax, ay, az, bx + dx, by + dy, bz + dz>(acc + OFF_B * 64, ABx, ABy, ABz);
          if constexpr (!METRIC && KEEP_D && b_d > 0)
            vb -= b_d * hrr_column<G::DN_LO, ax, ay, az, bx - dx, by - dy, bz - dz>(acc + OFF_D * 64, ABx, ABy, ABz);
          gA[d] += w * v;
          gB[d] += w * vb;
        });
      });
    });
#pragma unroll
    for (int d = 0; d < 3; ++d) gc[d] -= gA[d] + gB[d];
    double* o = a.out_ab + (static_cast<std::size_t>(task_id) * gridDim.y + blockIdx.y) * 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double sa = wave_sum(gA[d]);
      const double sb = wave_sum(gB[d]);
      if (threadIdx.x == 0) {
        o[d] = a.accumulate ? o[d] + sa : sa;
        o[3 + d] = a.accumulate ? o[3 + d] + sb : sb;
      }
    }
  }
  if (active) {
    double* o = a.out_c + (static_cast<std::size_t>(blockIdx.x) * a.nfn + f) * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = a.accumulate ? o[d] + gc[d] : gc[d];
  }
}

struct OneElectronGradArgs {
  const DevShell* shells;
  const double* exps;
  const double* coefs;
  const int2* tasks;
  int ntask;
  const double4* atoms;  // x, y, z, Z
  int natom;
  const double* ta;
  const double* tb;
  int nfa, nfb;
  const double* boys;
  const double* dens;   // n x n public AOs: weights of d(T + V)
  const double* edens;  // n x n public AOs: weights of -dS
  int n;
  double* out_ab;  // [ntask][6]
  double* out_c;   // [gridDim.x][natom][3]: Hellmann-Feynman term of each nucleus
  int sub;         // lanes per shell pair (1, 2, 4 or 8): they split the nuclei
};

// One-electron force terms of one shell pair per group of a.sub lanes:
//   sum D d(T + V)/dR - sum W dS/dR,
// the overlap and kinetic parts through d/dB = -d/dA, the nuclear attraction
// through d/dA of each nucleus' term and its Hellmann-Feynman derivative
// d/dC = -(d/dA + d/dB) (s_integrals.cpp:1865-2060; hf_gradient.cpp:8-18).
//
// The pair's Cartesian weights w = TA^T D TB (and the energy-weighted ones)
// are formed once into LDS, one slot per pair of the wave.  Entry loops run
// at run time and pick their Hermite factors by predicates, so no lane holds
// a full table of them: per entry the three axes' value rows E^{ab}_t and
// derivative rows 2a E^{a+1,b}_t - a E^{a-1,b}_t are rebuilt from the
// recurrence.  In round r lane `part` of the group takes nucleus r sub + part,
// so every lane works in every round; for that nucleus it forms dV/dA and the
// Hellmann-Feynman term dV/dC = sum E_tuv (-R_{tuv + e_d}) and takes
// dV/dB = -(dV/dA + dV/dC).  The Hellmann-Feynman term of a nucleus is summed
// over the wave's lanes with the same part, the pair's dA/dB over its group,
// both by fixed butterflies.
template <int LA, int LB>
__global__ __launch_bounds__(64) void one_electron_grad_kernel(const OneElectronGradArgs a) {
  constexpr int NA = nc(LA), NB = nc(LB), L = LA + LB + 1, NP = NA * NB;
  extern __shared__ double pair_weights[];  // [64 / sub][2][NP]: wd, we per pair slot
  const int id = blockIdx.x * 64 + threadIdx.x;
  const int pair = id / a.sub, part = id % a.sub;
  const bool active = pair < a.ntask;
  const int2 task = a.tasks[active ? pair : 0];
  const DevShell A = a.shells[task.x];
  const DevShell B = a.shells[task.y];
  double* w = pair_weights + static_cast<std::size_t>(threadIdx.x / a.sub) * 2 * NP;
  {
    const double fac = active ? (task.x != task.y ? 2.0 : 1.0) : 0.0;
#pragma unroll 1
    for (int e = part; e < NP; e += a.sub) {
      const int i = e / NB, j = e % NB;
      double vd = 0.0, ve = 0.0;
      for (int ma = 0; ma < a.nfa; ++ma)
        for (int mb = 0; mb < a.nfb; ++mb) {
          const double t = a.ta[ma * NA + i] * a.tb[mb * NB + j];
          const std::size_t idx = static_cast<std::size_t>(A.fn_off + ma) * a.n + B.fn_off + mb;
          vd += t * a.dens[idx];
          ve += t * a.edens[idx];
        }
      w[e] = fac * vd;
      w[NP + e] = fac * ve;
    }
  }
  __syncthreads();
  const double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
  const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
  double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
  // Overlap and kinetic: d/dA only, entries split over the group.
  for (int ia = 0; ia < A.nprim; ++ia) {
    const double ea = a.exps[A.prim_off + ia], ca = a.coefs[A.prim_off + ia];
    for (int ib = 0; ib < B.nprim; ++ib) {
      const double eb = a.exps[B.prim_off + ib], cb = a.coefs[B.prim_off + ib];
      const double p = ea + eb, ip = 1.0 / p, h = 0.5 * ip;
      const double kab = ca * cb * exp(-ea * eb * ip * AB2);
      const double P[3] = {(ea * A.x + eb * B.x) * ip, (ea * A.y + eb * B.y) * ip, (ea * A.z + eb * B.z) * ip};
      const double s1 = sqrt(M_PI * ip);
      double ov[3][LA + 2][LB + 3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        overlap_1d<LA + 1, LB + 2>(P[d] - (d == 0 ? A.x : d == 1 ? A.y : A.z), P[d] - (d == 0 ? B.x : d == 1 ? B.y : B.z),
                                   h, ov[d]);
#pragma unroll 1
      for (int e = part; e < NP; e += a.sub) {
        const int i = e / NB, j = e % NB;
        const int ai[3] = {cart_lx(LA, i), cart_ly(LA, i), cart_lz(LA, i)};
        const int bj[3] = {cart_lx(LB, j), cart_ly(LB, j), cart_lz(LB, j)};
        double sv[3], kv[3], sd[3], kd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          // Rows a - 1, a, a + 1 of the 1-D factors at columns b - 2, b, b + 2.
          double f[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
          for (int ii = 0; ii <= LA + 1; ++ii)
#pragma unroll
            for (int jj = 0; jj <= LB + 2; ++jj) {
              const double x = ov[d][ii][jj];
#pragma unroll
              for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                  if (ii == ai[d] + r - 1 && jj == bj[d] + 2 * c - 2) f[r][c] = x;
            }
          const int b = bj[d];
          auto kin = [&](int r) { return -2.0 * eb * eb * f[r][2] + eb * (2 * b + 1) * f[r][1] - 0.5 * b * (b - 1) * f[r][0]; };
          sv[d] = f[1][1] * s1;
          kv[d] = kin(1) * s1;
          sd[d] = (2.0 * ea * f[2][1] - ai[d] * f[0][1]) * s1;
          kd[d] = (2.0 * ea * kin(2) - ai[d] * kin(0)) * s1;
        }
        const double wd = w[e], we = w[NP + e];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int y = (d + 1) % 3, z = (d + 2) % 3;
          const double dS = sd[d] * sv[y] * sv[z];
          const double dT = kd[d] * sv[y] * sv[z] + sd[d] * kv[y] * sv[z] + sd[d] * sv[y] * kv[z];
          gA[d] += kab * (wd * dT - we * dS);
        }
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) gB[d] = -gA[d];
  // Nuclear attraction, one nucleus per lane and round so the Hellmann-Feynman
  // term can be reduced across the wave (every lane runs the same rounds).
  for (int c0 = 0; c0 < a.natom; c0 += a.sub) {
    const int c = c0 + part;
    const bool has = c < a.natom;
    const double4 at = a.atoms[has ? c : 0];
    double dA[3] = {0.0, 0.0, 0.0}, dC[3] = {0.0, 0.0, 0.0};
    for (int ia = 0; ia < A.nprim; ++ia) {
      const double ea = a.exps[A.prim_off + ia], ca = a.coefs[A.prim_off + ia];
      for (int ib = 0; ib < B.nprim; ++ib) {
        const double eb = a.exps[B.prim_off + ib], cb = a.coefs[B.prim_off + ib];
        const double p = ea + eb, ip = 1.0 / p, h = 0.5 * ip;
        const double kab = ca * cb * exp(-ea * eb * ip * AB2);
        const double P[3] = {(ea * A.x + eb * B.x) * ip, (ea * A.y + eb * B.y) * ip, (ea * A.z + eb * B.z) * ip};
        const double PA[3] = {P[0] - A.x, P[1] - A.y, P[2] - A.z};
        const double PB[3] = {P[0] - B.x, P[1] - B.y, P[2] - B.z};
        const double X = P[0] - at.x, Y = P[1] - at.y, Z = P[2] - at.z;
        double F[L + 1];
        boys_eval<L>(p * (X * X + Y * Y + Z * Z), a.boys, F);
        double R[ntuv(L)];
        hermite_r<L>(p, X, Y, Z, F, R);
        const double pref = has ? -at.w * 2.0 * M_PI * ip * kab : 0.0;
#pragma unroll 1
        for (int e = 0; e < NP; ++e) {
          const int i = e / NB, j = e % NB;
          const int ai[3] = {cart_lx(LA, i), cart_ly(LA, i), cart_lz(LA, i)};
          const int bj[3] = {cart_lx(LB, j), cart_ly(LB, j), cart_lz(LB, j)};
          // Value rows E^{ab}_t and A-derivative rows of the three axes.
          double val[3][L + 1], der[3][L + 1];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            double up[L + 1], dn[L + 1];
            hermite_1d<L>(ai[x], bj[x], PA[x], PB[x], h, val[x]);
            hermite_1d<L>(ai[x] + 1, bj[x], PA[x], PB[x], h, up);
            hermite_1d<L>(ai[x] > 0 ? ai[x] - 1 : 0, bj[x], PA[x], PB[x], h, dn);
#pragma unroll
            for (int t = 0; t <= L; ++t) der[x][t] = 2.0 * ea * up[t] - (ai[x] > 0 ? ai[x] * dn[t] : 0.0);
          }
          const int n0 = ai[0] + bj[0], n1 = ai[1] + bj[1], n2 = ai[2] + bj[2];
          const double wt = pref * w[e];
          double sa[3] = {0.0, 0.0, 0.0}, sc[3] = {0.0, 0.0, 0.0};
#pragma unroll
          for (int t = 0; t <= L; ++t)
#pragma unroll
            for (int u = 0; u <= L - t; ++u)
#pragma unroll
              for (int v = 0; v <= L - t - u; ++v) {
                const double r = R[tuv(t, u, v)];
                if (t <= n0 + 1 && u <= n1 && v <= n2) sa[0] += der[0][t] * val[1][u] * val[2][v] * r;
                if (t <= n0 && u <= n1 + 1 && v <= n2) sa[1] += val[0][t] * der[1][u] * val[2][v] * r;
                if (t <= n0 && u <= n1 && v <= n2 + 1) sa[2] += val[0][t] * val[1][u] * der[2][v] * r;
                // dV/dC: the value coefficients against R shifted by one along each axis.
                if (t >= 1 && t - 1 <= n0 && u <= n1 && v <= n2) sc[0] -= val[0][t - 1] * val[1][u] * val[2][v] * r;
                if (u >= 1 && t <= n0 && u - 1 <= n1 && v <= n2) sc[1] -= val[0][t] * val[1][u - 1] * val[2][v] * r;
                if (v >= 1 && t <= n0 && u <= n1 && v - 1 <= n2) sc[2] -= val[0][t] * val[1][u] * val[2][v - 1] * r;
              }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            dA[d] += wt * sa[d];
            dC[d] += wt * sc[d];
          }
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      gA[d] += dA[d];
      gB[d] -= dA[d] + dC[d];
      double hf = dC[d];
      for (int off = 32; off >= a.sub; off >>= 1) hf += __shfl_xor(hf, off);
      if (threadIdx.x < a.sub && has) a.out_c[(static_cast<std::size_t>(blockIdx.x) * a.natom + c) * 3 + d] = hf;
    }
  }
  for (int off = 1; off < a.sub; off <<= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      gA[d] += __shfl_xor(gA[d], off);
      gB[d] += __shfl_xor(gB[d], off);
    }
  if (active && part == 0) {
    double* o = a.out_ab + static_cast<std::size_t>(pair) * 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      o[d] = gA[d];
      o[3 + d] = gB[d];
    }
  }
}

// LDS budget of one gradient block's accumulators: a gfx950 workgroup may
// hold the CU's whole 160 KiB (the f-f class needs 134 KiB, d-d 60 KiB).
constexpr int kGradAccumulatorBytes = 160 << 10;

template <int LA, int LB, int LC, bool METRIC, bool CART, int MASK>
void launch_3c_grad_pass(ThreeCenterGradArgs args, bool accumulate, int nchunk, hipStream_t s) {
  constexpr std::size_t lds = std::size_t{64} * sizeof(double) * GradAccumulators<LA, LB, METRIC>::entries(MASK);
  static_assert(lds <= kGradAccumulatorBytes, "gradient accumulators exceed the LDS budget");
  args.accumulate = accumulate ? 1 : 0;
  dim3 grid(nchunk, (args.nfn + 63) / 64);
  hipLaunchKernelGGL((three_center_grad_kernel<LA, LB, LC, METRIC, CART, MASK>), grid, dim3(64), lds, s, args);
  VQC_LAUNCH("three_center_grad_kernel");
}

// One launch when all three accumulators fit the LDS budget, else two
// (raised-on-A with lowered, then raised-on-B) or three.
template <int LA, int LB, int LC, bool METRIC, bool CART>
void launch_3c_grad(const ThreeCenterGradArgs& args, int nchunk, hipStream_t s) {
  if (args.ntask == 0 || args.nfn == 0) return;
  using G = GradAccumulators<LA, LB, METRIC>;
  constexpr int cap = kGradAccumulatorBytes / (64 * static_cast<int>(sizeof(double)));
  if constexpr (G::entries(7) <= cap) {
    launch_3c_grad_pass<LA, LB, LC, METRIC, CART, 7>(args, false, nchunk, s);
  } else if constexpr (G::entries(5) <= cap && G::entries(2) <= cap) {
    launch_3c_grad_pass<LA, LB, LC, METRIC, CART, 5>(args, false, nchunk, s);
    launch_3c_grad_pass<LA, LB, LC, METRIC, CART, 2>(args, true, nchunk, s);
  } else {
    launch_3c_grad_pass<LA, LB, LC, METRIC, CART, 1>(args, false, nchunk, s);
    launch_3c_grad_pass<LA, LB, LC, METRIC, CART, 2>(args, true, nchunk, s);
    launch_3c_grad_pass<LA, LB, LC, METRIC, CART, 4>(args, true, nchunk, s);
  }
}

template <int LA, int LB, bool METRIC>
void dispatch_lc_grad(int lc, bool cart, const ThreeCenterGradArgs& args, int nchunk, hipStream_t s) {
  switch (lc) {
    case 0: launch_3c_grad<LA, LB, 0, METRIC, false>(args, nchunk, s); break;
    case 1: launch_3c_grad<LA, LB, 1, METRIC, false>(args, nchunk, s); break;
    case 2:
      if (cart) launch_3c_grad<LA, LB, 2, METRIC, true>(args, nchunk, s);
      else launch_3c_grad<LA, LB, 2, METRIC, false>(args, nchunk, s);
      break;
    case 3:
      if (cart) launch_3c_grad<LA, LB, 3, METRIC, true>(args, nchunk, s);
      else launch_3c_grad<LA, LB, 3, METRIC, false>(args, nchunk, s);
      break;
    default: throw NotImplemented("auxiliary angular momentum above f");
  }
}

// Blocks of the one-electron gradient launch for `ntask` pairs at `sub` lanes each.
inline int one_electron_grad_blocks(int ntask, int sub) {
  return static_cast<int>((static_cast<std::size_t>(ntask) * sub + 63) / 64);
}

template <int LA, int LB>
void launch_1e_grad(const OneElectronGradArgs& args, hipStream_t s) {
  if (args.ntask == 0) return;
  if (args.sub < 1 || args.sub > 8 || (args.sub & (args.sub - 1)) != 0)
    throw std::invalid_argument("one-electron lanes per pair must be 1, 2, 4 or 8");
  const std::size_t lds = static_cast<std::size_t>(64 / args.sub) * 2 * nc(LA) * nc(LB) * sizeof(double);
  hipLaunchKernelGGL((one_electron_grad_kernel<LA, LB>), dim3(one_electron_grad_blocks(args.ntask, args.sub)),
                     dim3(64), lds, s, args);
  VQC_LAUNCH("one_electron_grad_kernel");
}

#define VQC_DECLARE_GRAD_CLASS(LA, LB)                                                                       \
  void dispatch_3c_grad_##LA##LB(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s); \
  void dispatch_1e_grad_##LA##LB(const OneElectronGradArgs& a, hipStream_t s);
#define VQC_DEFINE_GRAD_CLASS(LA, LB)                                                                         \
  void dispatch_3c_grad_##LA##LB(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s) { \
    dispatch_lc_grad<LA, LB, false>(lc, cart, a, nchunk, s);                                                   \
  }                                                                                                           \
  void dispatch_1e_grad_##LA##LB(const OneElectronGradArgs& a, hipStream_t s) { launch_1e_grad<LA, LB>(a, s); }
#define VQC_DEFINE_METRIC_GRAD(LA)                                                                              \
  void dispatch_metric_grad_##LA(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s) { \
    dispatch_lc_grad<LA, 0, true>(lc, cart, a, nchunk, s);                                                     \
  }
VQC_DECLARE_GRAD_CLASS(0, 0)
VQC_DECLARE_GRAD_CLASS(1, 0)
VQC_DECLARE_GRAD_CLASS(1, 1)
VQC_DECLARE_GRAD_CLASS(2, 0)
VQC_DECLARE_GRAD_CLASS(2, 1)
VQC_DECLARE_GRAD_CLASS(2, 2)
VQC_DECLARE_GRAD_CLASS(3, 0)
VQC_DECLARE_GRAD_CLASS(3, 1)
VQC_DECLARE_GRAD_CLASS(3, 2)
VQC_DECLARE_GRAD_CLASS(3, 3)
void dispatch_metric_grad_0(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);
void dispatch_metric_grad_1(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);
void dispatch_metric_grad_2(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);
void dispatch_metric_grad_3(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);

}  // namespace detail
}  // namespace vqc

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
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Contracted Cartesian (a b| from the (e0| accumulator whose lowest degree is
// LO: (x-B)^b = sum_k C(b,k) (x-A)^k (A-B)^(b-k).
template <int LO, int NE>
__device__ __forceinline__ double hrr_value(const double (&acc)[NE], int ax, int ay, int az, int bx, int by, int bz,
                                            double ABx, double ABy, double ABz) {
  double v = 0.0;
#pragma unroll
  for (int kx = 0; kx <= bx; ++kx)
#pragma unroll
    for (int ky = 0; ky <= by; ++ky)
#pragma unroll
      for (int kz = 0; kz <= bz; ++kz) {
        double w = binom(bx, kx) * binom(by, ky) * binom(bz, kz);
#pragma unroll
        for (int r = 0; r < bx - kx; ++r) w *= ABx;
#pragma unroll
        for (int r = 0; r < by - ky; ++r) w *= ABy;
#pragma unroll
        for (int r = 0; r < bz - kz; ++r) w *= ABz;
        const int le = ax + ay + az + kx + ky + kz;
        v += w * acc[ne_offset(LO, le) + cart_index(ax + kx, ay + ky, az + kz)];
      }
  return v;
}

template <int LA, int LB, int LC, bool METRIC, bool CART>
__global__ __launch_bounds__(64) void three_center_grad_kernel(const ThreeCenterGradArgs a) {
  constexpr int LAB = LA + LB;
  constexpr int LO = LA > 0 ? LA - 1 : 0;
  constexpr int LHI = LAB + 1;
  constexpr int L = LHI + LC;
  constexpr int NE = ne_offset(LO, LHI + 1);
  constexpr int NH = hermite_count(LC, CART);
  constexpr int NA = nc(LA), NB = nc(LB);
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
    // Cartesian weights of this shell pair and auxiliary function.
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

    double acc0[NE], accA[NE], accB[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) acc0[k] = accA[k] = accB[k] = 0.0;
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
        double Ex[LHI + 1][LHI + 1], Ey[LHI + 1][LHI + 1], Ez[LHI + 1][LHI + 1];
        hermite_e0<LHI>(Px - A.x, 0.5 * ip, Ex);
        hermite_e0<LHI>(Py - A.y, 0.5 * ip, Ey);
        hermite_e0<LHI>(Pz - A.z, 0.5 * ip, Ez);
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
          const double i2g = 0.5 / g;
          double scale[LC + 1];
          scale[0] = 1.0;
#pragma unroll
          for (int k = 1; k <= LC; ++k) scale[k] = scale[k - 1] * i2g;
          double W[ntuv(LHI)];
#pragma unroll
          for (int s = 0; s <= LHI; ++s)
#pragma unroll
            for (int t = 0; t <= s; ++t)
#pragma unroll
              for (int u = 0; u <= s - t; ++u) {
                const int v = s - t - u;
                double w = 0.0;
                int k = 0;
#pragma unroll
                for (int d = LC; d >= 0; d -= 2) {
#pragma unroll
                  for (int tt = 0; tt <= d; ++tt)
#pragma unroll
                    for (int uu = 0; uu <= d - tt; ++uu) {
                      const int vv = d - tt - uu;
                      if (k < NH) w += hc[k] * scale[(LC + d) / 2] * R[tuv(t + tt, u + uu, v + vv)];
                      ++k;
                    }
                  if (!CART) break;
                }
                W[tuv(t, u, v)] = w;
              }
          const double pref = 2.0 * 17.493418327624862 /* pi^{5/2} */ / (p * g * sqrt(q)) * kab * cg;
          const double pa = 2.0 * ea * pref, pb = 2.0 * eb * pref;
#pragma unroll
          for (int le = LO; le <= LHI; ++le)
#pragma unroll
            for (int c = 0; c < nc(le); ++c) {
              const int ex = cart_lx(le, c), ey = cart_ly(le, c), ez = cart_lz(le, c);
              double s = 0.0;
#pragma unroll
              for (int t = 0; t <= ex; ++t) {
                double su = 0.0;
#pragma unroll
                for (int u = 0; u <= ey; ++u) {
                  double sv = 0.0;
#pragma unroll
                  for (int v = 0; v <= ez; ++v) sv += Ez[ez][v] * W[tuv(t, u, v)];
                  su += Ey[ey][u] * sv;
                }
                s += Ex[ex][t] * su;
              }
              const int k = ne_offset(LO, le) + c;
              if (le <= LAB) acc0[k] += pref * s;
              if (le >= LA + 1) accA[k] += pa * s;
              if (!METRIC && le >= LA) accB[k] += pb * s;
            }
        }
      }
    }

    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        const int ax[3] = {cart_lx(LA, ia), cart_ly(LA, ia), cart_lz(LA, ia)};
        const int bx[3] = {cart_lx(LB, ib), cart_ly(LB, ib), cart_lz(LB, ib)};
        const double w = wc[ia][ib];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int up[3] = {ax[0] + (d == 0), ax[1] + (d == 1), ax[2] + (d == 2)};
          double v = hrr_value<LO, NE>(accA, up[0], up[1], up[2], bx[0], bx[1], bx[2], ABx, ABy, ABz);
          if (ax[d] > 0) {
            const int dn[3] = {ax[0] - (d == 0), ax[1] - (d == 1), ax[2] - (d == 2)};
            v -= ax[d] * hrr_value<LO, NE>(acc0, dn[0], dn[1], dn[2], bx[0], bx[1], bx[2], ABx, ABy, ABz);
          }
          gA[d] += w * v;
          if (!METRIC) {
            const int bu[3] = {bx[0] + (d == 0), bx[1] + (d == 1), bx[2] + (d == 2)};
            double vb = hrr_value<LO, NE>(accB, ax[0], ax[1], ax[2], bu[0], bu[1], bu[2], ABx, ABy, ABz);
            if (bx[d] > 0) {
              const int bd[3] = {bx[0] - (d == 0), bx[1] - (d == 1), bx[2] - (d == 2)};
              vb -= bx[d] * hrr_value<LO, NE>(acc0, ax[0], ax[1], ax[2], bd[0], bd[1], bd[2], ABx, ABy, ABz);
            }
            gB[d] += w * vb;
          }
        }
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) gc[d] -= gA[d] + gB[d];
    double* o = a.out_ab + (static_cast<std::size_t>(task_id) * gridDim.y + blockIdx.y) * 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double sa = wave_sum(gA[d]);
      const double sb = wave_sum(gB[d]);
      if (threadIdx.x == 0) {
        o[d] = sa;
        o[3 + d] = sb;
      }
    }
  }
  if (active) {
    double* o = a.out_c + (static_cast<std::size_t>(blockIdx.x) * a.nfn + f) * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = gc[d];
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

template <int LA, int LB, int LC, bool METRIC, bool CART>
void launch_3c_grad(const ThreeCenterGradArgs& args, int nchunk, hipStream_t s) {
  if (args.ntask == 0 || args.nfn == 0) return;
  dim3 grid(nchunk, (args.nfn + 63) / 64);
  hipLaunchKernelGGL((three_center_grad_kernel<LA, LB, LC, METRIC, CART>), grid, dim3(64), 0, s, args);
  VQC_LAUNCH("three_center_grad_kernel");
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
void dispatch_metric_grad_0(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);
void dispatch_metric_grad_1(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);
void dispatch_metric_grad_2(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);
void dispatch_metric_grad_3(int lc, bool cart, const ThreeCenterGradArgs& a, int nchunk, hipStream_t s);

}  // namespace detail
}  // namespace vqc

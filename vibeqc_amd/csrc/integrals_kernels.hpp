// Integral kernel templates shared by the per-class translation units
// integrals_cXY.hip (one per orbital (la, lb) class, compiled in parallel) and
// the host driver integrals.hip.  See integrals.hip for the algorithm notes.
#pragma once

#include <cmath>
#include <type_traits>
#include <utility>

#include "integrals.hpp"

namespace vqc {
namespace detail {

__host__ __device__ constexpr int ntuv(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int tuv(int t, int u, int v) {
  const int s = t + u + v;
  return s * (s + 1) * (s + 2) / 6 + t * (s + 1) - t * (t - 1) / 2 + u;
}
__host__ __device__ constexpr int nc(int l) { return (l + 1) * (l + 2) / 2; }
// Inverse of tuv(): the degree s, t and u of the e-th packed entry.
__host__ __device__ constexpr int tuv_degree(int e) {
  int s = 0;
  while (ntuv(s) <= e) ++s;
  return s;
}
__host__ __device__ constexpr int tuv_t(int e) {
  const int s = tuv_degree(e);
  int r = e - (s == 0 ? 0 : ntuv(s - 1)), t = 0;
  while (r >= s - t + 1) {
    r -= s - t + 1;
    ++t;
  }
  return t;
}
__host__ __device__ constexpr int tuv_u(int e) {
  const int s = tuv_degree(e);
  int r = e - (s == 0 ? 0 : ntuv(s - 1)), t = 0;
  while (r >= s - t + 1) {
    r -= s - t + 1;
    ++t;
  }
  return r;
}
// Position of (t, u, v) among the degree-(t+u+v) entries (tuv() minus the
// entries of lower degree).
__host__ __device__ constexpr int tuv_pos(int t, int u, int v) {
  const int s = t + u + v;
  return t * (s + 1) - t * (t - 1) / 2 + u;
}
__host__ __device__ constexpr int cart_lx(int l, int c) {
  int lx = l;
  while (c > l - lx) {
    c -= l - lx + 1;
    --lx;
  }
  return lx;
}
__host__ __device__ constexpr int cart_lz(int l, int c) {
  int lx = l;
  while (c > l - lx) {
    c -= l - lx + 1;
    --lx;
  }
  return c;
}
__host__ __device__ constexpr int cart_ly(int l, int c) { return l - cart_lx(l, c) - cart_lz(l, c); }
__host__ __device__ constexpr int cart_index(int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  int base = 0;
  for (int x = l; x > lx; --x) base += l - x + 1;
  (void)ly;
  return base + lz;
}
// Number of Cartesian e components with la <= |e| <= la+lb, and the offset of degree l.
__host__ __device__ constexpr int ne_offset(int la, int l) {
  int n = 0;
  for (int k = la; k < l; ++k) n += nc(k);
  return n;
}
// Hermite coefficient entries of an auxiliary function: degree lc only
// (real-spherical) or every degree lc, lc-2, ... (Cartesian).
__host__ __device__ constexpr int hermite_count(int lc, bool cart) {
  int n = 0;
  for (int s = lc; s >= 0; s -= 2) {
    n += nc(s);
    if (!cart) break;
  }
  return n;
}
__host__ __device__ constexpr double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, so
// every register-array index derived from I is a constant expression (a
// run-time index would put the array in per-lane scratch memory).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (N > 0) static_for_impl(f, std::make_integer_sequence<int, N>{});
}
#define VQC_CI(x) decltype(x)::value

template <int M>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ table, double (&F)[M + 1]) {
  const double e = exp(-T);
  if (T < kBoysTmax) {
    const int k = __double2int_rn(T * (1.0 / kBoysStep));
    const double dt = k * kBoysStep - T;
    const double* row = table + k * kBoysCols + M;
    double f = row[7];
    f = row[6] + f * dt * (1.0 / 7.0);
    f = row[5] + f * dt * (1.0 / 6.0);
    f = row[4] + f * dt * (1.0 / 5.0);
    f = row[3] + f * dt * (1.0 / 4.0);
    f = row[2] + f * dt * (1.0 / 3.0);
    f = row[1] + f * dt * (1.0 / 2.0);
    f = row[0] + f * dt;
    F[M] = f;
#pragma unroll
    for (int m = M; m >= 1; --m) F[m - 1] = (2.0 * T * F[m] + e) * (1.0 / (2 * m - 1));
  } else {
    F[0] = 0.5 * sqrt(M_PI / T);
    const double h = 0.5 / T;
#pragma unroll
    for (int m = 1; m <= M; ++m) F[m] = ((2 * m - 1) * F[m - 1] - e) * h;
  }
}

// R_{tuv} = R^{(0)}_{tuv} for t+u+v <= L, in place from the highest degree down:
// at level n the degree-s entries come from the level n+1 entries of degrees
// s-1 and s-2, which are still in place because lower degrees are updated
// later (R_{t+1,u,v} = X R_{t,u,v} + t R_{t-1,u,v}, likewise in u and v).
template <int L>
__device__ __forceinline__ void hermite_r(double alpha, double X, double Y, double Z, const double (&F)[L + 1],
                                          double (&R)[ntuv(L)]) {
  double pw[L + 1];
  pw[0] = 1.0;
#pragma unroll
  for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * alpha);
  R[0] = pw[L] * F[L];
  static_for<L>([&](auto NI) {
    constexpr int n = L - 1 - VQC_CI(NI);
    static_for<L - n>([&](auto SI) {
      constexpr int s = L - n - VQC_CI(SI);
      static_for<s + 1>([&](auto TI) {
        constexpr int t = VQC_CI(TI);
        static_for<s - t + 1>([&](auto UI) {
          constexpr int u = VQC_CI(UI), v = s - t - u;
          double val;
          if constexpr (t > 0) {
            val = X * R[tuv(t - 1, u, v)];
            if constexpr (t > 1) val += (t - 1) * R[tuv(t - 2, u, v)];
          } else if constexpr (u > 0) {
            val = Y * R[tuv(t, u - 1, v)];
            if constexpr (u > 1) val += (u - 1) * R[tuv(t, u - 2, v)];
          } else {
            val = Z * R[tuv(t, u, v - 1)];
            if constexpr (v > 1) val += (v - 1) * R[tuv(t, u, v - 2)];
          }
          R[tuv(t, u, v)] = val;
        });
      });
    });
    R[0] = pw[n] * F[n];
  });
}

// One-index Hermite coefficients E^{i,0}_t (i <= N) of (x-A)^i about P.
template <int N>
__device__ __forceinline__ void hermite_e0(double pa, double h, double (&E)[N + 1][N + 1]) {
#pragma unroll
  for (int i = 0; i <= N; ++i)
#pragma unroll
    for (int t = 0; t <= N; ++t) E[i][t] = 0.0;
  E[0][0] = 1.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int t = 0; t <= i + 1; ++t) {
      double v = 0.0;
      if (t > 0) v += h * E[i][t - 1];
      if (t <= i) v += pa * E[i][t];
      if (t + 1 <= i) v += (t + 1) * E[i][t + 1];
      E[i + 1][t] = v;
    }
}

struct ThreeCenterArgs {
  const DevShell* bra_shells;
  const double* bra_exps;
  const double* bra_coefs;
  const DevShell* aux_shells;
  const double* aux_exps;
  const double* aux_coefs;
  const int2* tasks;
  const AuxFunction* fns;
  int nfn;
  const double* hermite;
  const double* ta;
  const double* tb;
  int nfa, nfb;
  const double* boys;
  double* out;
  std::size_t ld;
  double prim_cutoff;
  std::size_t row0;  // packed pair row stored at out[0] (orbital pairs only)
};

// Ket contraction of one auxiliary primitive into Hermite bra space:
// ws_tuv += pref * sum_k hc_k (2g)^-(LC+|k|)/2 R_{tuv + k}, t+u+v <= LBRA,
// where k runs over the auxiliary function's Hermite terms (top degree only
// for real-spherical functions, every degree LC, LC-2, ... for Cartesian).
template <int LBRA, int LC, bool CART, int NH, int L>
__device__ __forceinline__ void ket_contract(const double (&hc)[NH], double i2g, double pref, const double (&R)[ntuv(L)],
                                             double (&ws)[ntuv(LBRA)]) {
  double scale[LC + 1];
  scale[0] = pref;
#pragma unroll
  for (int k = 1; k <= LC; ++k) scale[k] = scale[k - 1] * i2g;
  static_for<ntuv(LBRA)>([&](auto EI) {
    constexpr int e = VQC_CI(EI);
    constexpr int s = tuv_degree(e), t = tuv_t(e), u = tuv_u(e), v = s - t - u;
    double w = 0.0;
    // Hermite terms of degree d = LC, LC - 2, ... (k counts them in order).
    static_for<LC / 2 + 1>([&](auto DI) {
      constexpr int d = LC - 2 * VQC_CI(DI);
      if constexpr (CART || d == LC) {
        constexpr int k0 = hermite_count(LC, true) - hermite_count(d, true);
        static_for<nc(d)>([&](auto CI) {
          constexpr int c = VQC_CI(CI);
          constexpr int tt = cart_lx(d, c), uu = cart_ly(d, c), vv = cart_lz(d, c);
          constexpr int k = k0 + tuv_pos(tt, uu, vv);
          if constexpr (k < NH) w += hc[k] * scale[(LC + d) / 2] * R[tuv(t + tt, u + uu, v + vv)];
        });
      }
    });
    ws[e] += w;
  });
}

// Bra contraction: out_e (e = Cartesian monomials of degree LO..HI, packed by
// ne_offset) = sum_tuv Ex[ex][t] Ey[ey][u] Ez[ez][v] ws_tuv.
template <int LO, int HI, int LBRA, class Fn>
__device__ __forceinline__ void bra_contract(const double (&Ex)[LBRA + 1][LBRA + 1], const double (&Ey)[LBRA + 1][LBRA + 1],
                                             const double (&Ez)[LBRA + 1][LBRA + 1], const double (&ws)[ntuv(LBRA)],
                                             Fn&& emit) {
  static_for<HI - LO + 1>([&](auto LI) {
    constexpr int le = LO + VQC_CI(LI);
    static_for<nc(le)>([&](auto CI) {
      constexpr int c = VQC_CI(CI);
      constexpr int ex = cart_lx(le, c), ey = cart_ly(le, c), ez = cart_lz(le, c);
      double sum = 0.0;
      static_for<ex + 1>([&](auto TI) {
        constexpr int t = VQC_CI(TI);
        double su = 0.0;
        static_for<ey + 1>([&](auto UI) {
          constexpr int u = VQC_CI(UI);
          double sv = 0.0;
          static_for<ez + 1>([&](auto VI) {
            constexpr int v = VQC_CI(VI);
            sv += Ez[ez][v] * ws[tuv(t, u, v)];
          });
          su += Ey[ey][u] * sv;
        });
        sum += Ex[ex][t] * su;
      });
      emit(std::integral_constant<int, le>{}, std::integral_constant<int, c>{}, sum);
    });
  });
}

// Horizontal recurrence from the (e0| accumulator (degrees LO..) to one
// Cartesian (a b|: (x-B)^b = sum_k C(b,k) (x-A)^k (A-B)^(b-k).
template <int LO, int NE, int AX, int AY, int AZ, int BX, int BY, int BZ>
__device__ __forceinline__ double hrr_static(const double (&acc)[NE], double ABx, double ABy, double ABz) {
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
        v += w * acc[ne_offset(LO, le) + cart_index(AX + kx, AY + ky, AZ + kz)];
      });
    });
  });
  return v;
}

// (ab|P) for one shell pair (blockIdx.x) and 64 auxiliary functions of one
// angular momentum (one per lane).  Per bra primitive pair the auxiliary
// primitives are summed in Hermite space (ws), then contracted with the bra's
// Hermite coefficients once; every register index is a compile-time constant
// (static_for), so nothing is kept in per-lane scratch memory.
template <int LA, int LB, int LC, bool METRIC, bool CART>
__global__ __launch_bounds__(64) void three_center_kernel(const ThreeCenterArgs a) {
  constexpr int LAB = LA + LB;
  constexpr int L = LAB + LC;
  constexpr int NE = ne_offset(LA, LAB + 1);
  constexpr int NH = hermite_count(LC, CART);
  constexpr int NA = nc(LA), NB = nc(LB);
  const int f = blockIdx.y * 64 + threadIdx.x;
  if (f >= a.nfn) return;
  const int2 task = a.tasks[blockIdx.x];
  const DevShell A = a.bra_shells[task.x];
  const DevShell B = METRIC ? A : a.bra_shells[task.y];
  const AuxFunction fn = a.fns[f];
  const DevShell C = a.aux_shells[fn.shell];

  double hc[NH];
#pragma unroll
  for (int k = 0; k < NH; ++k) hc[k] = a.hermite[fn.coef_off + k];

  double acc[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) acc[k] = 0.0;

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
      double ws[ntuv(LAB)];
#pragma unroll
      for (int k = 0; k < ntuv(LAB); ++k) ws[k] = 0.0;
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
        ket_contract<LAB, LC, CART, NH, L>(hc, 0.5 / g, pref, R, ws);
      }
      double Ex[LAB + 1][LAB + 1], Ey[LAB + 1][LAB + 1], Ez[LAB + 1][LAB + 1];
      hermite_e0<LAB>(Px - A.x, 0.5 * ip, Ex);
      hermite_e0<LAB>(Py - A.y, 0.5 * ip, Ey);
      hermite_e0<LAB>(Pz - A.z, 0.5 * ip, Ez);
      bra_contract<LA, LAB, LAB>(Ex, Ey, Ez, ws, [&](auto LI, auto CI, double v) {
        acc[ne_offset(LA, VQC_CI(LI)) + VQC_CI(CI)] += kab * v;
      });
    }
  }

  // Horizontal recurrence to Cartesian (ab|, then public AOs.  The transform
  // rows past nfa / nfb are zero padding, so the loops have constant trip
  // counts; only the stores test the public counts.
  double cart[NA][NB];
  static_for<NA>([&](auto IA) {
    constexpr int ia = VQC_CI(IA);
    static_for<NB>([&](auto IB) {
      constexpr int ib = VQC_CI(IB);
      cart[ia][ib] = hrr_static<LA, NE, cart_lx(LA, ia), cart_ly(LA, ia), cart_lz(LA, ia), cart_lx(LB, ib),
                                cart_ly(LB, ib), cart_lz(LB, ib)>(acc, ABx, ABy, ABz);
    });
  });
  double half[NA][NB];  // [public a][cartesian b]
#pragma unroll
  for (int ma = 0; ma < NA; ++ma)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      double v = 0.0;
#pragma unroll
      for (int ia = 0; ia < NA; ++ia) v += a.ta[ma * NA + ia] * cart[ia][ib];
      half[ma][ib] = v;
    }
#pragma unroll
  for (int ma = 0; ma < NA; ++ma) {
#pragma unroll
    for (int mb = 0; mb < NB; ++mb) {
      if (ma >= a.nfa || mb >= a.nfb) continue;
      double v = 0.0;
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) v += a.tb[mb * NB + ib] * half[ma][ib];
      if (METRIC) {
        a.out[static_cast<std::size_t>(A.fn_off + ma) * a.ld + fn.index] = v;
      } else {
        const int mu = A.fn_off + ma, nu = B.fn_off + mb;
        if (task.x == task.y && nu > mu) continue;
        const std::size_t hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
        a.out[(hi * (hi + 1) / 2 + lo - a.row0) * a.ld + fn.index] = v;
      }
    }
  }
}

struct OneElectronArgs {
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
  double* S;
  double* T;
  double* V;
  int n;
  int sub;  // lanes per shell pair (1, 2, 4 or 8): they split the nuclei of V
};

// Lanes per shell pair for the one-electron kernels: enough that each lane
// takes about four nuclei of the V sum, at most eight, and at least enough
// that the 64-lane block's LDS slots (`pair_bytes` per shell pair, 64 / sub
// pairs per block) stay within 64 KiB.
constexpr std::size_t kOneElectronLds = std::size_t{64} << 10;
inline int one_electron_lanes(int natom, std::size_t pair_bytes) {
  int sub = 1;
  while (sub < 8 && (sub * 4 < natom || (64 / sub) * pair_bytes > kOneElectronLds)) sub *= 2;
  return sub;
}
// LDS bytes per shell pair of the value kernel (S, T, V slots).
inline std::size_t one_electron_pair_bytes(int la, int lb) {
  return std::size_t{3} * ((la + 1) * (la + 2) / 2) * ((lb + 1) * (lb + 2) / 2) * sizeof(double);
}

// Overlap factors E^{ij}_0 along one axis for i <= I, j <= J by the
// Obara-Saika recurrences S(i+1,j) = PA S(i,j) + h (i S(i-1,j) + j S(i,j-1))
// and S(0,j+1) = PB S(0,j) + h j S(0,j-1), h = 1/(2p), S(0,0) = 1.
template <int I, int J>
__device__ __forceinline__ void overlap_1d(double pa, double pb, double h, double (&s)[I + 1][J + 1]) {
#pragma unroll
  for (int i = 0; i <= I; ++i)
#pragma unroll
    for (int j = 0; j <= J; ++j) {
      double v = 1.0;
      if (i > 0) {
        v = pa * s[i - 1][j];
        if (i > 1) v += h * (i - 1) * s[i - 2][j];
        if (j > 0) v += h * j * s[i - 1][j - 1];
      } else if (j > 0) {
        v = pb * s[0][j - 1];
        if (j > 1) v += h * (j - 1) * s[0][j - 2];
      }
      s[i][j] = v;
    }
}

// Hermite expansion coefficients E^{ij}_t (t <= i + j) of one axis of a
// Gaussian product, for exponents i, j that are compile-time constants after
// unrolling: E^{i+1,j}_t = h E^{ij}_{t-1} + PA E^{ij}_t + (t+1) E^{ij}_{t+1}
// from E^{00}_0 = 1, then the same transfer in j with PB.
template <int T>
__device__ __forceinline__ void hermite_1d(int i, int j, double pa, double pb, double h, double (&e)[T + 1]) {
#pragma unroll
  for (int t = 0; t <= T; ++t) e[t] = t == 0 ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < T; ++k) {
    if (k >= i + j) break;
    const int deg = k;  // degree of the current row
    const double f = k < i ? pa : pb;
    double nxt[T + 1];
#pragma unroll
    for (int t = 0; t <= T; ++t) {
      double v = 0.0;
      if (t <= deg + 1) {
        if (t > 0) v += h * e[t - 1];
        if (t <= deg) v += f * e[t];
        if (t + 1 <= deg) v += (t + 1) * e[t + 1];
      }
      nxt[t] = v;
    }
#pragma unroll
    for (int t = 0; t <= T; ++t) e[t] = nxt[t];
  }
}

// Overlap, kinetic and nuclear-attraction integrals of one shell pair per
// group of a.sub consecutive lanes (s_integrals.cpp:1865-2060).
//
// Per primitive pair the group's lanes split the Cartesian entries (i, j)
// of S and T (entry e belongs to lane e % sub) and the nuclei of V.  V is
// contracted once per primitive pair: every lane sums the Coulomb Hermite
// integrals of its nuclei, R~_tuv = sum_C -Z_C R_tuv(P - C), the group adds
// its partial sums by a fixed butterfly, and the owner of each entry forms
// sum_tuv E^x_t E^y_u E^z_v R~_tuv.  The Cartesian accumulators live in LDS,
// one slot per pair of the wave, so no lane keeps the whole S/T/V block in
// registers.  Lanes past the last pair compute pair 0 into their own slot and
// write nothing, so every lane reaches the shuffles.
template <int LA, int LB>
__global__ __launch_bounds__(64) void one_electron_kernel(const OneElectronArgs a) {
  constexpr int NA = nc(LA), NB = nc(LB), L = LA + LB, NP = NA * NB;
  extern __shared__ double cart_acc[];  // [64 / sub][3][NP]: S, T, V per pair slot
  const int id = blockIdx.x * 64 + threadIdx.x;
  const int pair = id / a.sub, part = id % a.sub;
  const bool active = pair < a.ntask;
  const int2 task = a.tasks[active ? pair : 0];
  const DevShell A = a.shells[task.x];
  const DevShell B = a.shells[task.y];
  double* slot = cart_acc + static_cast<std::size_t>(threadIdx.x / a.sub) * 3 * NP;
#pragma unroll
  for (int e = 0; e < NP; ++e)
    if (e % a.sub == part) slot[e] = slot[NP + e] = slot[2 * NP + e] = 0.0;
  const double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
  const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
  for (int ia = 0; ia < A.nprim; ++ia) {
    const double ea = a.exps[A.prim_off + ia], ca = a.coefs[A.prim_off + ia];
    for (int ib = 0; ib < B.nprim; ++ib) {
      const double eb = a.exps[B.prim_off + ib], cb = a.coefs[B.prim_off + ib];
      const double p = ea + eb, ip = 1.0 / p, h = 0.5 * ip;
      const double kab = ca * cb * exp(-ea * eb * ip * AB2);
      const double P[3] = {(ea * A.x + eb * B.x) * ip, (ea * A.y + eb * B.y) * ip, (ea * A.z + eb * B.z) * ip};
      const double PA[3] = {P[0] - A.x, P[1] - A.y, P[2] - A.z};
      const double PB[3] = {P[0] - B.x, P[1] - B.y, P[2] - B.z};
      const double s1 = sqrt(M_PI * ip);
      {
        // S and T: separable, from 1-D overlap factors up to j + 2.  The
        // entry loop runs at run time (its exponents select factors by
        // predicates, so every register index stays a constant).
        double ov[3][LA + 1][LB + 3];
#pragma unroll
        for (int d = 0; d < 3; ++d) overlap_1d<LA, LB + 2>(PA[d], PB[d], h, ov[d]);
#pragma unroll 1
        for (int e = part; e < NP; e += a.sub) {
          const int i = e / NB, j = e % NB;
          const int ai[3] = {cart_lx(LA, i), cart_ly(LA, i), cart_lz(LA, i)};
          const int bj[3] = {cart_lx(LB, j), cart_ly(LB, j), cart_lz(LB, j)};
          double sd[3], kd[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double s0 = 0.0, sp = 0.0, sm = 0.0;
#pragma unroll
            for (int ii = 0; ii <= LA; ++ii)
#pragma unroll
              for (int jj = 0; jj <= LB + 2; ++jj) {
                const double x = ov[d][ii][jj];
                if (ii == ai[d]) {
                  s0 = jj == bj[d] ? x : s0;
                  sp = jj == bj[d] + 2 ? x : sp;
                  sm = jj + 2 == bj[d] ? x : sm;
                }
              }
            sd[d] = s0 * s1;
            kd[d] = (-2.0 * eb * eb * sp + eb * (2 * bj[d] + 1) * s0 - 0.5 * bj[d] * (bj[d] - 1) * sm) * s1;
          }
          slot[e] += kab * sd[0] * sd[1] * sd[2];
          slot[NP + e] += kab * (kd[0] * sd[1] * sd[2] + sd[0] * kd[1] * sd[2] + sd[0] * sd[1] * kd[2]);
        }
      }
      // V: nuclear Hermite sums of this lane's nuclei, then the group's total.
      double rs[ntuv(L)];
#pragma unroll
      for (int k = 0; k < ntuv(L); ++k) rs[k] = 0.0;
      for (int c = part; c < a.natom; c += a.sub) {
        const double4 at = a.atoms[c];
        const double X = P[0] - at.x, Y = P[1] - at.y, Z = P[2] - at.z;
        double F[L + 1];
        boys_eval<L>(p * (X * X + Y * Y + Z * Z), a.boys, F);
        double R[ntuv(L)];
        hermite_r<L>(p, X, Y, Z, F, R);
#pragma unroll
        for (int k = 0; k < ntuv(L); ++k) rs[k] -= at.w * R[k];
      }
      for (int off = 1; off < a.sub; off <<= 1)
#pragma unroll
        for (int k = 0; k < ntuv(L); ++k) rs[k] += __shfl_xor(rs[k], off);
      const double pref = 2.0 * M_PI * ip * kab;
#pragma unroll 1
      for (int e = part; e < NP; e += a.sub) {
        const int i = e / NB, j = e % NB;
        const int ai[3] = {cart_lx(LA, i), cart_ly(LA, i), cart_lz(LA, i)};
        const int bj[3] = {cart_lx(LB, j), cart_ly(LB, j), cart_lz(LB, j)};
        double ex[L + 1], ey[L + 1], ez[L + 1];
        hermite_1d<L>(ai[0], bj[0], PA[0], PB[0], h, ex);
        hermite_1d<L>(ai[1], bj[1], PA[1], PB[1], h, ey);
        hermite_1d<L>(ai[2], bj[2], PA[2], PB[2], h, ez);
        const int nx = ai[0] + bj[0], ny = ai[1] + bj[1], nz = ai[2] + bj[2];
        double val = 0.0;
#pragma unroll
        for (int t = 0; t <= L; ++t)
#pragma unroll
          for (int u = 0; u <= L - t; ++u)
#pragma unroll
            for (int v = 0; v <= L - t - u; ++v)
              if (t <= nx && u <= ny && v <= nz) val += ex[t] * ey[u] * ez[v] * rs[tuv(t, u, v)];
        slot[2 * NP + e] += pref * val;
      }
    }
  }
  // The group's entries were written by its own lanes (one wave); publish
  // them to the whole group before the transform.
  __syncthreads();
  if (!active) return;
  // Public AOs: out = TA x TB^T, written to both triangles; the group's lanes
  // split the public entries.
  for (int m = part; m < a.nfa * a.nfb; m += a.sub) {
    const int ma = m / a.nfb, mb = m % a.nfb;
    double os = 0.0, ok = 0.0, ov = 0.0;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const double w = a.ta[ma * NA + i] * a.tb[mb * NB + j];
        os += w * slot[i * NB + j];
        ok += w * slot[NP + i * NB + j];
        ov += w * slot[2 * NP + i * NB + j];
      }
    const std::size_t mu = A.fn_off + ma, nu = B.fn_off + mb;
    a.S[mu * a.n + nu] = os;
    a.S[nu * a.n + mu] = os;
    a.T[mu * a.n + nu] = ok;
    a.T[nu * a.n + mu] = ok;
    a.V[mu * a.n + nu] = ov;
    a.V[nu * a.n + mu] = ov;
  }
}

template <int LA, int LB, int LC, bool METRIC, bool CART>
void launch_3c(const ThreeCenterArgs& args, int ntask, hipStream_t s) {
  if (ntask == 0 || args.nfn == 0) return;
  dim3 grid(ntask, (args.nfn + 63) / 64);
  hipLaunchKernelGGL((three_center_kernel<LA, LB, LC, METRIC, CART>), grid, dim3(64), 0, s, args);
  VQC_LAUNCH("three_center_kernel");
}

template <int LA, int LB, bool METRIC>
void dispatch_lc(int lc, bool cart, const ThreeCenterArgs& args, int ntask, hipStream_t s) {
  switch (lc) {
    case 0: launch_3c<LA, LB, 0, METRIC, false>(args, ntask, s); break;
    case 1: launch_3c<LA, LB, 1, METRIC, false>(args, ntask, s); break;
    case 2:
      if (cart) launch_3c<LA, LB, 2, METRIC, true>(args, ntask, s);
      else launch_3c<LA, LB, 2, METRIC, false>(args, ntask, s);
      break;
    case 3:
      if (cart) launch_3c<LA, LB, 3, METRIC, true>(args, ntask, s);
      else launch_3c<LA, LB, 3, METRIC, false>(args, ntask, s);
      break;
    default: throw NotImplemented("auxiliary angular momentum above f");
  }
}

template <int LA, int LB>
void launch_1e(const OneElectronArgs& args, hipStream_t s) {
  if (args.ntask == 0) return;
  if (args.sub < 1 || args.sub > 8 || (args.sub & (args.sub - 1)) != 0)
    throw std::invalid_argument("one-electron lanes per pair must be 1, 2, 4 or 8");
  const std::size_t lanes = static_cast<std::size_t>(args.ntask) * args.sub;
  const std::size_t lds = static_cast<std::size_t>(64 / args.sub) * 3 * nc(LA) * nc(LB) * sizeof(double);
  hipLaunchKernelGGL((one_electron_kernel<LA, LB>), dim3(static_cast<unsigned>((lanes + 63) / 64)), dim3(64), lds, s,
                     args);
  VQC_LAUNCH("one_electron_kernel");
}


// Per-class entry points, defined in integrals_cXY.hip.
#define VQC_DECLARE_CLASS(LA, LB)                                                                  \
  void dispatch_3c_##LA##LB(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s); \
  void dispatch_1e_##LA##LB(const OneElectronArgs& a, hipStream_t s);
#define VQC_DEFINE_CLASS(LA, LB)                                                                    \
  void dispatch_3c_##LA##LB(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s) { \
    dispatch_lc<LA, LB, false>(lc, cart, a, ntask, s);                                               \
  }                                                                                                 \
  void dispatch_1e_##LA##LB(const OneElectronArgs& a, hipStream_t s) { launch_1e<LA, LB>(a, s); }
#define VQC_DEFINE_METRIC(LA)                                                                        \
  void dispatch_metric_##LA(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s) { \
    dispatch_lc<LA, 0, true>(lc, cart, a, ntask, s);                                                 \
  }
VQC_DECLARE_CLASS(0, 0)
VQC_DECLARE_CLASS(1, 0)
VQC_DECLARE_CLASS(1, 1)
VQC_DECLARE_CLASS(2, 0)
VQC_DECLARE_CLASS(2, 1)
VQC_DECLARE_CLASS(2, 2)
VQC_DECLARE_CLASS(3, 0)
VQC_DECLARE_CLASS(3, 1)
VQC_DECLARE_CLASS(3, 2)
VQC_DECLARE_CLASS(3, 3)
void dispatch_metric_0(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s);
void dispatch_metric_1(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s);
void dispatch_metric_2(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s);
void dispatch_metric_3(int lc, bool cart, const ThreeCenterArgs& a, int ntask, hipStream_t s);

}  // namespace detail
}  // namespace vqc

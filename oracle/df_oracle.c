/*
 * df_oracle.c -- FP64 CPU oracle for the Gaussian integrals on the DF-JK path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls it.
 *
 * It restates, in plain C, the McMurchie-Davidson (Hermite) evaluation that
 * the reference uses as its general integral fallback:
 *   - Boys function: positive series below T = m_max + 16, erf plus upward
 *     recurrence above (src/integrals/s_integrals.cpp:126-148);
 *   - Hermite expansion coefficients E^{ij}_t (s_integrals.cpp:171-211);
 *   - Coulomb Hermite integrals R_{tuv} (s_integrals.cpp:225-271);
 *   - the primitive ERI assembly (s_integrals.cpp:472-524) specialised to
 *     three-centre (mu nu|P) and two-centre (P|Q) Coulomb integrals;
 *   - one-electron overlap, kinetic and nuclear-attraction integrals used for
 *     Hcore (s_integrals.cpp:1865-2060);
 *   - nuclear derivatives of all of the above, contracted with caller weights
 *     (the DF force response, density_fitting.cpp:898-917 and
 *     s_integrals.cpp:1073-1199; one-electron terms of hf_gradient.cpp:8-18),
 *     from the Gaussian derivative rule d/dA_x g_i = 2a g_{i+1} - i g_{i-1}
 *     applied per primitive and translational invariance for the remaining
 *     centre.
 * Outputs are CARTESIAN, with every Cartesian AO individually normalised
 * (primitive radial factor and contraction scale already folded into the
 * coefficients by the caller, component factor 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)
 * applied here, src/molecule/basis.cpp:179-184).  The caller applies the
 * real-spherical transforms.
 *
 * Cartesian order is CCA (src/molecule/basis.cpp:33-45): lx descending, then
 * lz ascending.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_MAX_L 4
#define ORACLE_MAX_LTOT (4 * ORACLE_MAX_L + 2)

typedef struct {
  const double *centers;   /* [nshell][3] */
  const int *l;            /* [nshell] */
  const int *prim_offset;  /* [nshell] */
  const int *prim_count;   /* [nshell] */
  const double *exps;      /* [nprim] */
  const double *coefs;     /* [nprim] normalised primitive coefficients */
  int nshell;
} oracle_basis;

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void cart_components(int l, int comps[][3]) {
  int n = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int lz = 0; lz <= l - lx; ++lz) {
      comps[n][0] = lx;
      comps[n][1] = l - lx - lz;
      comps[n][2] = lz;
      ++n;
    }
}

static double odd_double_factorial(int k) {
  double v = 1.0;
  for (int f = 1; f < 2 * k; f += 2) v *= (double)f;
  return v;
}

static double component_norm(const int c[3]) {
  return 1.0 / sqrt(odd_double_factorial(c[0]) * odd_double_factorial(c[1]) *
                    odd_double_factorial(c[2]));
}

/* Boys F_m(T), m = 0..mmax (s_integrals.cpp:126-148). */
static void boys(int mmax, double t, double *f) {
  const double ex = exp(-t);
  if (t < (double)mmax + 16.0) {
    for (int n = 0; n <= mmax; ++n) {
      double term = 1.0 / (2.0 * n + 1.0);
      double sum = term;
      for (int k = 1; k < 512; ++k) {
        term *= 2.0 * t / (2.0 * n + 2.0 * k + 1.0);
        sum += term;
        if (term <= sum * 2.0e-16) break;
      }
      f[n] = ex * sum;
    }
  } else {
    f[0] = 0.5 * sqrt(M_PI / t) * erf(sqrt(t));
    for (int n = 1; n <= mmax; ++n) f[n] = ((2.0 * n - 1.0) * f[n - 1] - ex) / (2.0 * t);
  }
}

/* 1D Hermite coefficients E^{ij}_t, i<=la, j<=lb, stored E[(i*(lb+1)+j)*(la+lb+1)+t]. */
static void hermite_1d(int la, int lb, double p, double xpa, double xpb, double *e) {
  const int tdim = la + lb + 1;
  const int jdim = lb + 1;
  memset(e, 0, sizeof(double) * (size_t)(la + 1) * jdim * tdim);
#define E(i, j, t) e[((i) * jdim + (j)) * tdim + (t)]
  const double h = 0.5 / p;
  E(0, 0, 0) = 1.0;
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = xpa * (t <= i ? E(i, 0, t) : 0.0);
      if (t > 0) v += h * E(i, 0, t - 1);
      if (t + 1 <= i) v += (t + 1) * E(i, 0, t + 1);
      E(i + 1, 0, t) = v;
    }
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j < lb; ++j)
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = xpb * (t <= i + j ? E(i, j, t) : 0.0);
        if (t > 0) v += h * E(i, j, t - 1);
        if (t + 1 <= i + j) v += (t + 1) * E(i, j, t + 1);
        E(i, j + 1, t) = v;
      }
#undef E
}

/* Hermite Coulomb integrals R_{tuv} = R^{(0)}_{tuv}, t+u+v <= L, stored R[(t*(L+1)+u)*(L+1)+v]. */
static void hermite_r(int L, double alpha, double x, double y, double z, double *r) {
  const int d = L + 1;
  const size_t per = (size_t)d * d * d;
  double f[ORACLE_MAX_LTOT + 2];
  double *work = (double *)calloc(per * (size_t)(L + 1), sizeof(double));
  boys(L, alpha * (x * x + y * y + z * z), f);
#define W(n, t, u, v) work[(size_t)(n) * per + ((size_t)(t) * d + (u)) * d + (v)]
  double scale = 1.0;
  for (int n = 0; n <= L; ++n) {
    W(n, 0, 0, 0) = scale * f[n];
    scale *= -2.0 * alpha;
  }
  for (int n = L - 1; n >= 0; --n) {
    const int top = L - n;
    for (int s = 1; s <= top; ++s)
      for (int t = 0; t <= s; ++t)
        for (int u = 0; u <= s - t; ++u) {
          const int v = s - t - u;
          double val;
          if (t > 0) {
            val = x * W(n + 1, t - 1, u, v);
            if (t > 1) val += (t - 1) * W(n + 1, t - 2, u, v);
          } else if (u > 0) {
            val = y * W(n + 1, t, u - 1, v);
            if (u > 1) val += (u - 1) * W(n + 1, t, u - 2, v);
          } else {
            val = z * W(n + 1, t, u, v - 1);
            if (v > 1) val += (v - 1) * W(n + 1, t, u, v - 2);
          }
          W(n, t, u, v) = val;
        }
  }
  memcpy(r, work, sizeof(double) * per);
#undef W
  free(work);
}

int oracle_cart_count(const oracle_basis *b) {
  int n = 0;
  for (int s = 0; s < b->nshell; ++s) n += ncart(b->l[s]);
  return n;
}

static void cart_offsets(const oracle_basis *b, int *off) {
  int n = 0;
  for (int s = 0; s < b->nshell; ++s) {
    off[s] = n;
    n += ncart(b->l[s]);
  }
}

/*
 * Three-centre Cartesian integrals (mu nu | P):
 *   out[(mu * nc + nu) * naux_c + P]
 * (ab|c) = 2 pi^{5/2} / (p g sqrt(p+g)) K_ab
 *          sum_{tuv} E^{ab}_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi} E^{c}_{tau nu phi}
 *          R_{t+tau,u+nu,v+phi}(p g/(p+g), P - C)
 * A two-centre metric is the same expression with a unit s function
 * (exponent zero) as the second bra function.
 */
static void three_center_block(const oracle_basis *ob, int sa, int sb, const oracle_basis *ab,
                               int sc, int use_unit_b, double *block) {
  const int la = ob->l[sa], lb = use_unit_b ? 0 : ob->l[sb], lc = ab->l[sc];
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc);
  int ca[15][3], cb[15][3], cc[15][3];
  cart_components(la, ca);
  cart_components(lb, cb);
  cart_components(lc, cc);
  const double *A = &ob->centers[3 * sa];
  const double *B = use_unit_b ? A : &ob->centers[3 * sb];
  const double *C = &ab->centers[3 * sc];
  const int lab = la + lb, L = la + lb + lc;
  const int d = L + 1;
  const int tdim = lab + 1;
  double ex[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 1) * (2 * ORACLE_MAX_L + 1)];
  double ey[sizeof(ex) / sizeof(double)], ez[sizeof(ex) / sizeof(double)];
  double gx[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 1)], gy[sizeof(gx) / sizeof(double)],
      gz[sizeof(gx) / sizeof(double)];
  double *r = (double *)malloc(sizeof(double) * (size_t)d * d * d);
  double *w = (double *)malloc(sizeof(double) * (size_t)nc * tdim * tdim * tdim);
  memset(block, 0, sizeof(double) * (size_t)na * nb * nc);
  const int npa = ob->prim_count[sa], npb = use_unit_b ? 1 : ob->prim_count[sb];
  const int npc = ab->prim_count[sc];
  for (int ia = 0; ia < npa; ++ia) {
    const double a = ob->exps[ob->prim_offset[sa] + ia];
    const double ka = ob->coefs[ob->prim_offset[sa] + ia];
    for (int ib = 0; ib < npb; ++ib) {
      const double b = use_unit_b ? 0.0 : ob->exps[ob->prim_offset[sb] + ib];
      const double kb = use_unit_b ? 1.0 : ob->coefs[ob->prim_offset[sb] + ib];
      const double p = a + b;
      double P[3], pa[3], pb[3], ab2 = 0.0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A[x] + b * B[x]) / p;
        pa[x] = P[x] - A[x];
        pb[x] = P[x] - B[x];
        ab2 += (A[x] - B[x]) * (A[x] - B[x]);
      }
      const double kab = ka * kb * exp(-a * b / p * ab2);
      hermite_1d(la, lb, p, pa[0], pb[0], ex);
      hermite_1d(la, lb, p, pa[1], pb[1], ey);
      hermite_1d(la, lb, p, pa[2], pb[2], ez);
      for (int ic = 0; ic < npc; ++ic) {
        const double g = ab->exps[ab->prim_offset[sc] + ic];
        const double kc = ab->coefs[ab->prim_offset[sc] + ic];
        const double alpha = p * g / (p + g);
        const double pref = 2.0 * pow(M_PI, 2.5) / (p * g * sqrt(p + g)) * kab * kc;
        hermite_r(L, alpha, P[0] - C[0], P[1] - C[1], P[2] - C[2], r);
        hermite_1d(lc, 0, g, 0.0, 0.0, gx);
        memcpy(gy, gx, sizeof(gx));
        memcpy(gz, gx, sizeof(gx));
        /* ket contraction: w[k][t][u][v] = sum (-1)^{tau+nu+phi} e_tau e_nu e_phi R */
        for (int k = 0; k < nc; ++k)
          for (int t = 0; t <= lab; ++t)
            for (int u = 0; u <= lab - t; ++u)
              for (int v = 0; v <= lab - t - u; ++v) {
                double s = 0.0;
                for (int tau = 0; tau <= cc[k][0]; ++tau) {
                  const double e1 = gx[cc[k][0] * (lc + 1) + tau];
                  if (e1 == 0.0) continue;
                  for (int nu = 0; nu <= cc[k][1]; ++nu) {
                    const double e2 = gy[cc[k][1] * (lc + 1) + nu];
                    if (e2 == 0.0) continue;
                    for (int phi = 0; phi <= cc[k][2]; ++phi) {
                      const double e3 = gz[cc[k][2] * (lc + 1) + phi];
                      if (e3 == 0.0) continue;
                      const double sign = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                      s += sign * e1 * e2 * e3 * r[((t + tau) * d + (u + nu)) * d + (v + phi)];
                    }
                  }
                }
                w[((k * tdim + t) * tdim + u) * tdim + v] = s;
              }
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            const double *exij = &ex[(ca[i][0] * (lb + 1) + cb[j][0]) * tdim];
            const double *eyij = &ey[(ca[i][1] * (lb + 1) + cb[j][1]) * tdim];
            const double *ezij = &ez[(ca[i][2] * (lb + 1) + cb[j][2]) * tdim];
            const int tx = ca[i][0] + cb[j][0], ty = ca[i][1] + cb[j][1], tz = ca[i][2] + cb[j][2];
            for (int k = 0; k < nc; ++k) {
              double s = 0.0;
              for (int t = 0; t <= tx; ++t)
                for (int u = 0; u <= ty; ++u)
                  for (int v = 0; v <= tz; ++v)
                    s += exij[t] * eyij[u] * ezij[v] * w[((k * tdim + t) * tdim + u) * tdim + v];
              block[(i * nb + j) * nc + k] += pref * s;
            }
          }
      }
    }
  }
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j)
      for (int k = 0; k < nc; ++k)
        block[(i * nb + j) * nc + k] *=
            component_norm(ca[i]) * component_norm(cb[j]) * component_norm(cc[k]);
  free(r);
  free(w);
}

int oracle_three_center(const oracle_basis *ob, const oracle_basis *ab, double *out) {
  for (int s = 0; s < ob->nshell; ++s)
    if (ob->l[s] > ORACLE_MAX_L) return 1;
  for (int s = 0; s < ab->nshell; ++s)
    if (ab->l[s] > ORACLE_MAX_L) return 1;
  const int n = oracle_cart_count(ob), naux = oracle_cart_count(ab);
  int *oo = (int *)malloc(sizeof(int) * ob->nshell);
  int *ao = (int *)malloc(sizeof(int) * ab->nshell);
  cart_offsets(ob, oo);
  cart_offsets(ab, ao);
  const int npair = ob->nshell * (ob->nshell + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
  for (int pair = 0; pair < npair; ++pair) {
    int sa = (int)((sqrt(8.0 * pair + 1.0) - 1.0) / 2.0);
    while (sa * (sa + 1) / 2 > pair) --sa;
    while ((sa + 1) * (sa + 2) / 2 <= pair) ++sa;
    const int sb = pair - sa * (sa + 1) / 2;
    const int na = ncart(ob->l[sa]), nb = ncart(ob->l[sb]);
    double *block = (double *)malloc(sizeof(double) * 15 * 15 * 15);
    for (int sc = 0; sc < ab->nshell; ++sc) {
      const int nc = ncart(ab->l[sc]);
      three_center_block(ob, sa, sb, ab, sc, 0, block);
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j)
          for (int k = 0; k < nc; ++k) {
            const double v = block[(i * nb + j) * nc + k];
            const size_t mu = (size_t)oo[sa] + i, nu = (size_t)oo[sb] + j, P = (size_t)ao[sc] + k;
            out[(mu * n + nu) * naux + P] = v;
            out[(nu * n + mu) * naux + P] = v;
          }
    }
    free(block);
  }
  free(oo);
  free(ao);
  return 0;
}

int oracle_metric(const oracle_basis *ab, double *out) {
  for (int s = 0; s < ab->nshell; ++s)
    if (ab->l[s] > ORACLE_MAX_L) return 1;
  const int naux = oracle_cart_count(ab);
  int *ao = (int *)malloc(sizeof(int) * ab->nshell);
  cart_offsets(ab, ao);
#pragma omp parallel for schedule(dynamic, 1)
  for (int sa = 0; sa < ab->nshell; ++sa) {
    double *block = (double *)malloc(sizeof(double) * 15 * 15);
    const int na = ncart(ab->l[sa]);
    for (int sc = 0; sc < ab->nshell; ++sc) {
      const int nc = ncart(ab->l[sc]);
      three_center_block(ab, sa, sa, ab, sc, 1, block);
      for (int i = 0; i < na; ++i)
        for (int k = 0; k < nc; ++k) out[(size_t)(ao[sa] + i) * naux + ao[sc] + k] = block[i * nc + k];
    }
    free(block);
  }
  free(ao);
  return 0;
}

/*
 * One-electron Cartesian integrals: overlap S, kinetic T and nuclear attraction
 * V = -sum_C Z_C <a| 1/|r-C| |b>, all nc x nc row-major.
 */
int oracle_one_electron(const oracle_basis *ob, int natom, const double *charges,
                        const double *atom_xyz, double *S, double *T, double *V) {
  for (int s = 0; s < ob->nshell; ++s)
    if (ob->l[s] > ORACLE_MAX_L) return 1;
  const int n = oracle_cart_count(ob);
  int *oo = (int *)malloc(sizeof(int) * ob->nshell);
  cart_offsets(ob, oo);
  const int npair = ob->nshell * ob->nshell;
#pragma omp parallel for schedule(dynamic, 1)
  for (int pair = 0; pair < npair; ++pair) {
    const int sa = pair / ob->nshell, sb = pair % ob->nshell;
    const int la = ob->l[sa], lb = ob->l[sb];
    const int na = ncart(la), nb = ncart(lb);
    int ca[15][3], cb[15][3];
    cart_components(la, ca);
    cart_components(lb, cb);
    const double *A = &ob->centers[3 * sa], *B = &ob->centers[3 * sb];
    const int lbk = lb + 2; /* kinetic needs j+2 */
    const int tdim = la + lbk + 1;
    double ex[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 3) * (2 * ORACLE_MAX_L + 3)];
    double ey[sizeof(ex) / sizeof(double)], ez[sizeof(ex) / sizeof(double)];
    const int L = la + lb, d = L + 1;
    double *r = (double *)malloc(sizeof(double) * (size_t)d * d * d);
    double sblk[225], tblk[225], vblk[225];
    memset(sblk, 0, sizeof(sblk));
    memset(tblk, 0, sizeof(tblk));
    memset(vblk, 0, sizeof(vblk));
    for (int ia = 0; ia < ob->prim_count[sa]; ++ia) {
      const double a = ob->exps[ob->prim_offset[sa] + ia];
      const double ka = ob->coefs[ob->prim_offset[sa] + ia];
      for (int ib = 0; ib < ob->prim_count[sb]; ++ib) {
        const double b = ob->exps[ob->prim_offset[sb] + ib];
        const double kb = ob->coefs[ob->prim_offset[sb] + ib];
        const double p = a + b;
        double P[3], pa[3], pb[3], ab2 = 0.0;
        for (int x = 0; x < 3; ++x) {
          P[x] = (a * A[x] + b * B[x]) / p;
          pa[x] = P[x] - A[x];
          pb[x] = P[x] - B[x];
          ab2 += (A[x] - B[x]) * (A[x] - B[x]);
        }
        const double kab = ka * kb * exp(-a * b / p * ab2);
        hermite_1d(la, lbk, p, pa[0], pb[0], ex);
        hermite_1d(la, lbk, p, pa[1], pb[1], ey);
        hermite_1d(la, lbk, p, pa[2], pb[2], ez);
        const double s1 = sqrt(M_PI / p);
#define EX(i, j, t) ex[((i) * (lbk + 1) + (j)) * tdim + (t)]
#define EY(i, j, t) ey[((i) * (lbk + 1) + (j)) * tdim + (t)]
#define EZ(i, j, t) ez[((i) * (lbk + 1) + (j)) * tdim + (t)]
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            const int ix = ca[i][0], iy = ca[i][1], iz = ca[i][2];
            const int jx = cb[j][0], jy = cb[j][1], jz = cb[j][2];
            const double sx = EX(ix, jx, 0) * s1, sy = EY(iy, jy, 0) * s1, sz = EZ(iz, jz, 0) * s1;
            /* 1D kinetic: -2b^2 S_{i,j+2} + b(2j+1) S_{ij} - j(j-1)/2 S_{i,j-2} */
            double kx = -2.0 * b * b * EX(ix, jx + 2, 0) * s1 + b * (2 * jx + 1) * sx;
            if (jx >= 2) kx -= 0.5 * jx * (jx - 1) * EX(ix, jx - 2, 0) * s1;
            double ky = -2.0 * b * b * EY(iy, jy + 2, 0) * s1 + b * (2 * jy + 1) * sy;
            if (jy >= 2) ky -= 0.5 * jy * (jy - 1) * EY(iy, jy - 2, 0) * s1;
            double kz = -2.0 * b * b * EZ(iz, jz + 2, 0) * s1 + b * (2 * jz + 1) * sz;
            if (jz >= 2) kz -= 0.5 * jz * (jz - 1) * EZ(iz, jz - 2, 0) * s1;
            sblk[i * nb + j] += kab * sx * sy * sz;
            tblk[i * nb + j] += kab * (kx * sy * sz + sx * ky * sz + sx * sy * kz);
          }
        for (int c = 0; c < natom; ++c) {
          const double *Cc = &atom_xyz[3 * c];
          hermite_r(L, p, P[0] - Cc[0], P[1] - Cc[1], P[2] - Cc[2], r);
          const double pref = -charges[c] * 2.0 * M_PI / p * kab;
          for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
              const int ix = ca[i][0], iy = ca[i][1], iz = ca[i][2];
              const int jx = cb[j][0], jy = cb[j][1], jz = cb[j][2];
              double s = 0.0;
              for (int t = 0; t <= ix + jx; ++t)
                for (int u = 0; u <= iy + jy; ++u)
                  for (int v = 0; v <= iz + jz; ++v)
                    s += EX(ix, jx, t) * EY(iy, jy, u) * EZ(iz, jz, v) * r[(t * d + u) * d + v];
              vblk[i * nb + j] += pref * s;
            }
        }
#undef EX
#undef EY
#undef EZ
      }
    }
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < nb; ++j) {
        const double f = component_norm(ca[i]) * component_norm(cb[j]);
        const size_t idx = (size_t)(oo[sa] + i) * n + oo[sb] + j;
        S[idx] = f * sblk[i * nb + j];
        T[idx] = f * tblk[i * nb + j];
        V[idx] = f * vblk[i * nb + j];
      }
    free(r);
  }
  free(oo);
  return 0;
}

/* ------------------------------------------------------------------------
 * Nuclear derivatives.  For a primitive pair, D^A_{ij,t} = 2a E_{i+1,j,t} -
 * i E_{i-1,j,t} and D^B_{ij,t} = 2b E_{i,j+1,t} - j E_{i,j-1,t} are the
 * Hermite coefficients of d/dA_x and d/dB_x of the product along x.
 * ------------------------------------------------------------------------ */

#define DEXE(e, jd, td, i, j, t) ((i) < 0 || (j) < 0 ? 0.0 : (e)[((i) * (jd) + (j)) * (td) + (t)])

static double dcoef(const double *e, int jd, int td, int i, int j, int t, int which, double a, double b) {
  if (which == 0) return 2.0 * a * DEXE(e, jd, td, i + 1, j, t) - (double)i * DEXE(e, jd, td, i - 1, j, t);
  return 2.0 * b * DEXE(e, jd, td, i, j + 1, t) - (double)j * DEXE(e, jd, td, i, j - 1, t);
}

/* d[0..5][(i*nb+j)*nc+k]: d/dA_{x,y,z} and d/dB_{x,y,z} of the Cartesian block
 * (ab|c) (or (a|c) with use_unit_b).  d/dC = -(d/dA + d/dB). */
static void three_center_deriv_block(const oracle_basis *ob, int sa, int sb, const oracle_basis *ab, int sc,
                                     int use_unit_b, double *d) {
  const int la = ob->l[sa], lb = use_unit_b ? 0 : ob->l[sb], lc = ab->l[sc];
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc);
  int ca[15][3], cb[15][3], cc[15][3];
  cart_components(la, ca);
  cart_components(lb, cb);
  cart_components(lc, cc);
  const double *A = &ob->centers[3 * sa];
  const double *B = use_unit_b ? A : &ob->centers[3 * sb];
  const double *C = &ab->centers[3 * sc];
  const int lab1 = la + lb + 1, L = lab1 + lc;
  const int dr = L + 1, tw = lab1 + 1;
  const int jd = lb + 2, td = la + lb + 3; /* tables for i <= la+1, j <= lb+1 */
  double *ex = (double *)malloc(sizeof(double) * (size_t)(la + 2) * jd * td * 3);
  double *ey = ex + (size_t)(la + 2) * jd * td, *ez = ey + (size_t)(la + 2) * jd * td;
  double gx[(ORACLE_MAX_L + 1) * (ORACLE_MAX_L + 1)];
  double *r = (double *)malloc(sizeof(double) * (size_t)dr * dr * dr);
  double *w = (double *)malloc(sizeof(double) * (size_t)nc * tw * tw * tw);
  const size_t blk = (size_t)na * nb * nc;
  memset(d, 0, sizeof(double) * 6 * blk);
  const int npa = ob->prim_count[sa], npb = use_unit_b ? 1 : ob->prim_count[sb];
  const int npc = ab->prim_count[sc];
  for (int ia = 0; ia < npa; ++ia) {
    const double a = ob->exps[ob->prim_offset[sa] + ia];
    const double ka = ob->coefs[ob->prim_offset[sa] + ia];
    for (int ib = 0; ib < npb; ++ib) {
      const double b = use_unit_b ? 0.0 : ob->exps[ob->prim_offset[sb] + ib];
      const double kb = use_unit_b ? 1.0 : ob->coefs[ob->prim_offset[sb] + ib];
      const double p = a + b;
      double P[3], pa[3], pb[3], ab2 = 0.0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A[x] + b * B[x]) / p;
        pa[x] = P[x] - A[x];
        pb[x] = P[x] - B[x];
        ab2 += (A[x] - B[x]) * (A[x] - B[x]);
      }
      const double kab = ka * kb * exp(-a * b / p * ab2);
      hermite_1d(la + 1, lb + 1, p, pa[0], pb[0], ex);
      hermite_1d(la + 1, lb + 1, p, pa[1], pb[1], ey);
      hermite_1d(la + 1, lb + 1, p, pa[2], pb[2], ez);
      const double *E[3] = {ex, ey, ez};
      for (int ic = 0; ic < npc; ++ic) {
        const double g = ab->exps[ab->prim_offset[sc] + ic];
        const double kc = ab->coefs[ab->prim_offset[sc] + ic];
        const double alpha = p * g / (p + g);
        const double pref = 2.0 * pow(M_PI, 2.5) / (p * g * sqrt(p + g)) * kab * kc;
        hermite_r(L, alpha, P[0] - C[0], P[1] - C[1], P[2] - C[2], r);
        hermite_1d(lc, 0, g, 0.0, 0.0, gx);
        for (int k = 0; k < nc; ++k)
          for (int t = 0; t <= lab1; ++t)
            for (int u = 0; u <= lab1 - t; ++u)
              for (int v = 0; v <= lab1 - t - u; ++v) {
                double sum = 0.0;
                for (int tau = 0; tau <= cc[k][0]; ++tau)
                  for (int nu = 0; nu <= cc[k][1]; ++nu)
                    for (int phi = 0; phi <= cc[k][2]; ++phi) {
                      const double e = gx[cc[k][0] * (lc + 1) + tau] * gx[cc[k][1] * (lc + 1) + nu] *
                                       gx[cc[k][2] * (lc + 1) + phi];
                      if (e == 0.0) continue;
                      const double sign = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                      sum += sign * e * r[((t + tau) * dr + (u + nu)) * dr + (v + phi)];
                    }
                w[((k * tw + t) * tw + u) * tw + v] = sum;
              }
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            const int ii[3] = {ca[i][0], ca[i][1], ca[i][2]}, jj[3] = {cb[j][0], cb[j][1], cb[j][2]};
            for (int which = 0; which < (use_unit_b ? 1 : 2); ++which)
              for (int dir = 0; dir < 3; ++dir) {
                /* Coefficients along each axis: derivative along dir, plain elsewhere. */
                int tmax[3];
                for (int x = 0; x < 3; ++x) tmax[x] = ii[x] + jj[x] + (x == dir ? 1 : 0);
                for (int k = 0; k < nc; ++k) {
                  double sum = 0.0;
                  for (int t = 0; t <= tmax[0]; ++t) {
                    const double c0 = dir == 0 ? dcoef(E[0], jd, td, ii[0], jj[0], t, which, a, b)
                                               : DEXE(E[0], jd, td, ii[0], jj[0], t);
                    if (c0 == 0.0) continue;
                    for (int u = 0; u <= tmax[1]; ++u) {
                      const double c1 = dir == 1 ? dcoef(E[1], jd, td, ii[1], jj[1], u, which, a, b)
                                                 : DEXE(E[1], jd, td, ii[1], jj[1], u);
                      if (c1 == 0.0) continue;
                      for (int v = 0; v <= tmax[2]; ++v) {
                        const double c2 = dir == 2 ? dcoef(E[2], jd, td, ii[2], jj[2], v, which, a, b)
                                                   : DEXE(E[2], jd, td, ii[2], jj[2], v);
                        sum += c0 * c1 * c2 * w[((k * tw + t) * tw + u) * tw + v];
                      }
                    }
                  }
                  d[(size_t)(which * 3 + dir) * blk + (i * nb + j) * nc + k] += pref * sum;
                }
              }
          }
      }
    }
  }
  for (int q = 0; q < 6; ++q)
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nc; ++k)
          d[(size_t)q * blk + (i * nb + j) * nc + k] *=
              component_norm(ca[i]) * component_norm(cb[j]) * component_norm(cc[k]);
  free(ex);
  free(r);
  free(w);
}

static void add_grad(double *grad, int atom, const double g[3], double sign) {
  for (int x = 0; x < 3; ++x) grad[3 * atom + x] += sign * g[x];
}

/*
 * grad[natom*3] += sum_{mu nu P} w[(mu*n+nu)*naux+P] d(mu nu|P)/dR over the
 * full Cartesian square (mu, nu) (w need not be symmetric).
 */
int oracle_grad_three_center(const oracle_basis *ob, const oracle_basis *ab, const int *atom_o, const int *atom_a,
                             const double *w, int natom, double *grad) {
  for (int s = 0; s < ob->nshell; ++s)
    if (ob->l[s] + 1 > ORACLE_MAX_L) return 1;
  for (int s = 0; s < ab->nshell; ++s)
    if (ab->l[s] > ORACLE_MAX_L) return 1;
  const int n = oracle_cart_count(ob), naux = oracle_cart_count(ab);
  int *oo = (int *)malloc(sizeof(int) * ob->nshell);
  int *ao = (int *)malloc(sizeof(int) * ab->nshell);
  cart_offsets(ob, oo);
  cart_offsets(ab, ao);
  const int npair = ob->nshell * (ob->nshell + 1) / 2;
#pragma omp parallel
  {
    double *g = (double *)calloc((size_t)natom * 3, sizeof(double));
    double *d = (double *)malloc(sizeof(double) * 6 * 15 * 15 * 15);
#pragma omp for schedule(dynamic, 1)
    for (int pair = 0; pair < npair; ++pair) {
      int sa = (int)((sqrt(8.0 * pair + 1.0) - 1.0) / 2.0);
      while (sa * (sa + 1) / 2 > pair) --sa;
      while ((sa + 1) * (sa + 2) / 2 <= pair) ++sa;
      const int sb = pair - sa * (sa + 1) / 2;
      const int na = ncart(ob->l[sa]), nb = ncart(ob->l[sb]);
      for (int sc = 0; sc < ab->nshell; ++sc) {
        const int nc = ncart(ab->l[sc]);
        const size_t blk = (size_t)na * nb * nc;
        three_center_deriv_block(ob, sa, sb, ab, sc, 0, d);
        double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j)
            for (int k = 0; k < nc; ++k) {
              const size_t mu = (size_t)oo[sa] + i, nu = (size_t)oo[sb] + j, P = (size_t)ao[sc] + k;
              double wt = w[(mu * n + nu) * naux + P];
              if (sa != sb) wt += w[(nu * n + mu) * naux + P];
              for (int x = 0; x < 3; ++x) {
                gA[x] += wt * d[(size_t)x * blk + (i * nb + j) * nc + k];
                gB[x] += wt * d[(size_t)(3 + x) * blk + (i * nb + j) * nc + k];
              }
            }
        add_grad(g, atom_o[sa], gA, 1.0);
        add_grad(g, atom_o[sb], gB, 1.0);
        add_grad(g, atom_a[sc], gA, -1.0);
        add_grad(g, atom_a[sc], gB, -1.0);
      }
    }
#pragma omp critical
    for (int k = 0; k < natom * 3; ++k) grad[k] += g[k];
    free(g);
    free(d);
  }
  free(oo);
  free(ao);
  return 0;
}

/* grad += sum_{PQ} w[P*naux+Q] d(P|Q)/dR over the full Cartesian square. */
int oracle_grad_metric(const oracle_basis *ab, const int *atom_a, const double *w, int natom, double *grad) {
  for (int s = 0; s < ab->nshell; ++s)
    if (ab->l[s] + 1 > ORACLE_MAX_L) return 1;
  const int naux = oracle_cart_count(ab);
  int *ao = (int *)malloc(sizeof(int) * ab->nshell);
  cart_offsets(ab, ao);
#pragma omp parallel
  {
    double *g = (double *)calloc((size_t)natom * 3, sizeof(double));
    double *d = (double *)malloc(sizeof(double) * 6 * 15 * 15);
#pragma omp for schedule(dynamic, 1)
    for (int sa = 0; sa < ab->nshell; ++sa) {
      const int na = ncart(ab->l[sa]);
      for (int sc = 0; sc < ab->nshell; ++sc) {
        const int nc = ncart(ab->l[sc]);
        three_center_deriv_block(ab, sa, sa, ab, sc, 1, d);
        double gA[3] = {0, 0, 0};
        for (int i = 0; i < na; ++i)
          for (int k = 0; k < nc; ++k) {
            const double wt = w[(size_t)(ao[sa] + i) * naux + ao[sc] + k];
            for (int x = 0; x < 3; ++x) gA[x] += wt * d[(size_t)x * na * nc + i * nc + k];
          }
        add_grad(g, atom_a[sa], gA, 1.0);
        add_grad(g, atom_a[sc], gA, -1.0);
      }
    }
#pragma omp critical
    for (int k = 0; k < natom * 3; ++k) grad[k] += g[k];
    free(g);
    free(d);
  }
  free(ao);
  return 0;
}

/*
 * One-electron force terms: grad += sum_{ab} dens[a][b] d(T + V)_{ab}/dR
 * - edens[a][b] dS_{ab}/dR over the full Cartesian square, including the
 * Hellmann-Feynman derivative of V with respect to each nucleus.
 */
int oracle_grad_one_electron(const oracle_basis *ob, int natom, const double *charges, const double *atom_xyz,
                             const int *atom_o, const double *dens, const double *edens, double *grad) {
  for (int s = 0; s < ob->nshell; ++s)
    if (ob->l[s] + 1 > ORACLE_MAX_L) return 1;
  const int n = oracle_cart_count(ob);
  int *oo = (int *)malloc(sizeof(int) * ob->nshell);
  cart_offsets(ob, oo);
  const int npair = ob->nshell * ob->nshell;
#pragma omp parallel
  {
    double *g = (double *)calloc((size_t)natom * 3, sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int pair = 0; pair < npair; ++pair) {
      const int sa = pair / ob->nshell, sb = pair % ob->nshell;
      const int la = ob->l[sa], lb = ob->l[sb];
      const int na = ncart(la), nb = ncart(lb);
      int ca[15][3], cb[15][3];
      cart_components(la, ca);
      cart_components(lb, cb);
      const double *A = &ob->centers[3 * sa], *B = &ob->centers[3 * sb];
      const int jd = lb + 3, td = la + lb + 4; /* i <= la+1, j <= lb+2 */
      double *ex = (double *)malloc(sizeof(double) * (size_t)(la + 2) * jd * td * 3);
      double *ey = ex + (size_t)(la + 2) * jd * td, *ez = ey + (size_t)(la + 2) * jd * td;
      const int L = la + lb + 1, dr = L + 1;
      double *r = (double *)malloc(sizeof(double) * (size_t)dr * dr * dr);
      double wts[225], ewts[225];
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
          const double f = component_norm(ca[i]) * component_norm(cb[j]);
          const size_t idx = (size_t)(oo[sa] + i) * n + oo[sb] + j;
          wts[i * nb + j] = f * dens[idx];
          ewts[i * nb + j] = f * edens[idx];
        }
      double gA[3] = {0, 0, 0};  /* S and T: d/dB = -d/dA */
      double *gV = (double *)calloc((size_t)natom * 3, sizeof(double));
      for (int ia = 0; ia < ob->prim_count[sa]; ++ia) {
        const double a = ob->exps[ob->prim_offset[sa] + ia];
        const double ka = ob->coefs[ob->prim_offset[sa] + ia];
        for (int ib = 0; ib < ob->prim_count[sb]; ++ib) {
          const double b = ob->exps[ob->prim_offset[sb] + ib];
          const double kb = ob->coefs[ob->prim_offset[sb] + ib];
          const double p = a + b;
          double P[3], pa[3], pb[3], ab2 = 0.0;
          for (int x = 0; x < 3; ++x) {
            P[x] = (a * A[x] + b * B[x]) / p;
            pa[x] = P[x] - A[x];
            pb[x] = P[x] - B[x];
            ab2 += (A[x] - B[x]) * (A[x] - B[x]);
          }
          const double kab = ka * kb * exp(-a * b / p * ab2);
          hermite_1d(la + 1, lb + 2, p, pa[0], pb[0], ex);
          hermite_1d(la + 1, lb + 2, p, pa[1], pb[1], ey);
          hermite_1d(la + 1, lb + 2, p, pa[2], pb[2], ez);
          const double *E[3] = {ex, ey, ez};
          const double s1 = sqrt(M_PI / p);
          for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
              const int ii[3] = {ca[i][0], ca[i][1], ca[i][2]}, jj[3] = {cb[j][0], cb[j][1], cb[j][2]};
              double sv[3], kv[3], sd[3], kd[3];
              for (int x = 0; x < 3; ++x) {
                const double *e = E[x];
                const int i0 = ii[x], j0 = jj[x];
                sv[x] = DEXE(e, jd, td, i0, j0, 0) * s1;
                kv[x] = (-2.0 * b * b * DEXE(e, jd, td, i0, j0 + 2, 0) + b * (2 * j0 + 1) * DEXE(e, jd, td, i0, j0, 0) -
                         (j0 >= 2 ? 0.5 * j0 * (j0 - 1) * DEXE(e, jd, td, i0, j0 - 2, 0) : 0.0)) *
                        s1;
                /* d/dA_x applied to the bra: 2a (i+1) - i (i-1) in the same formulas. */
                sd[x] = (2.0 * a * DEXE(e, jd, td, i0 + 1, j0, 0) - i0 * DEXE(e, jd, td, i0 - 1, j0, 0)) * s1;
                const double kp = -2.0 * b * b * DEXE(e, jd, td, i0 + 1, j0 + 2, 0) +
                                  b * (2 * j0 + 1) * DEXE(e, jd, td, i0 + 1, j0, 0) -
                                  (j0 >= 2 ? 0.5 * j0 * (j0 - 1) * DEXE(e, jd, td, i0 + 1, j0 - 2, 0) : 0.0);
                const double km = -2.0 * b * b * DEXE(e, jd, td, i0 - 1, j0 + 2, 0) +
                                  b * (2 * j0 + 1) * DEXE(e, jd, td, i0 - 1, j0, 0) -
                                  (j0 >= 2 ? 0.5 * j0 * (j0 - 1) * DEXE(e, jd, td, i0 - 1, j0 - 2, 0) : 0.0);
                kd[x] = (2.0 * a * kp - i0 * km) * s1;
              }
              const double wd = wts[i * nb + j], we = ewts[i * nb + j];
              for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                const double dS = sd[x] * sv[y] * sv[z];
                const double dT = kd[x] * sv[y] * sv[z] + sd[x] * kv[y] * sv[z] + sd[x] * sv[y] * kv[z];
                gA[x] += kab * (wd * dT - we * dS);
              }
            }
          for (int c = 0; c < natom; ++c) {
            const double *Cc = &atom_xyz[3 * c];
            hermite_r(L, p, P[0] - Cc[0], P[1] - Cc[1], P[2] - Cc[2], r);
            const double pref = -charges[c] * 2.0 * M_PI / p * kab;
            for (int i = 0; i < na; ++i)
              for (int j = 0; j < nb; ++j) {
                const double wd = wts[i * nb + j];
                if (wd == 0.0) continue;
                const int ii[3] = {ca[i][0], ca[i][1], ca[i][2]}, jj[3] = {cb[j][0], cb[j][1], cb[j][2]};
                for (int which = 0; which < 2; ++which)
                  for (int dir = 0; dir < 3; ++dir) {
                    double sum = 0.0;
                    for (int t = 0; t <= ii[0] + jj[0] + (dir == 0); ++t) {
                      const double c0 = dir == 0 ? dcoef(E[0], jd, td, ii[0], jj[0], t, which, a, b)
                                                 : DEXE(E[0], jd, td, ii[0], jj[0], t);
                      for (int u = 0; u <= ii[1] + jj[1] + (dir == 1); ++u) {
                        const double c1 = dir == 1 ? dcoef(E[1], jd, td, ii[1], jj[1], u, which, a, b)
                                                   : DEXE(E[1], jd, td, ii[1], jj[1], u);
                        for (int v = 0; v <= ii[2] + jj[2] + (dir == 2); ++v) {
                          const double c2 = dir == 2 ? dcoef(E[2], jd, td, ii[2], jj[2], v, which, a, b)
                                                     : DEXE(E[2], jd, td, ii[2], jj[2], v);
                          sum += c0 * c1 * c2 * r[(t * dr + u) * dr + v];
                        }
                      }
                    }
                    const double val = pref * wd * sum;
                    const int atom = which == 0 ? atom_o[sa] : atom_o[sb];
                    g[3 * atom + dir] += val;
                    gV[3 * c + dir] -= val; /* Hellmann-Feynman: d/dC = -(d/dA + d/dB) */
                  }
              }
          }
        }
      }
      add_grad(g, atom_o[sa], gA, 1.0);
      add_grad(g, atom_o[sb], gA, -1.0);
      for (int k = 0; k < natom * 3; ++k) g[k] += gV[k];
      free(gV);
      free(ex);
      free(r);
    }
#pragma omp critical
    for (int k = 0; k < natom * 3; ++k) grad[k] += g[k];
    free(g);
  }
  free(oo);
  return 0;
}

/* Boys values exposed for tests: f[0..mmax]. */
void oracle_boys(int mmax, double t, double *f) { boys(mmax, t, f); }

/*
 * pfdd_oracle.c -- FP64 CPU restatement of the PFDD GL step.
 *
 * TEST INFRASTRUCTURE ONLY (see pfdd_oracle.h): the checker for the HIP path.
 *
 * Reference: PFDD_Final_GPU.c (G.c).  Each function names the lines it follows.
 * What pins each function (tests/ names in brackets):
 *   orc_fft3d            brute-force DFT, numpy.fft golden vectors, delta/plane-wave
 *                        closed forms, and the compiled reference fourn (G.c:2758)
 *                        [test_oracle_fft.py]
 *   orc_frec1            compiled reference frec at 32x2x32 [test_oracle_setup.py]
 *   orc_bmatrix_*        literal loop nest vs factored form (1e-13), isotropic closed
 *                        form, symmetry/evenness/homogeneity, and the compiled reference
 *                        Bmatrix at 32x2x32 [test_oracle_elastic.py]
 *   orc_ddsum_at,        compiled reference Fmatrix DD / FF at 32x2x32
 *   orc_fmatrix_at       [test_oracle_elastic.py]
 *   orc_plasticevolv     compiled reference plasticevolv on inputs exact in FP32,
 *                        uniform-xi scalar recurrence [test_oracle_gl.py]
 *   orc_prologue,        the GL loop driven through the compiled reference's own
 *   orc_gl_steps         fourn/Bmatrix/Fmatrix/virtualevolv/plasticevolv at 32x2x32,
 *                        10 steps, FP32 tolerance; uniform-xi closed form
 *                        [test_oracle_gl.py]
 *   orc_strain_stress    compiled reference strain()/stress() at 32x2x32, and
 *                        <sigma> = sigma_applied*mu for periodic cells
 *                        [test_oracle_strain.py]
 *   orc_line_points,     compiled reference avestrain (energy_stressTimePf,
 *   orc_line_energy,     energy_in4) and ResidualEnergy at 32x2x32 on random
 *   orc_residual_energy  xi, stresses and D coefficients; a brute-force
 *                        restatement of the line rule [test_oracle_energy.py]
 */
#include "pfdd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * FFT: iterative Stockham autosort, mixed radix 2/3/4/5, double twiddles.
 * Conventions of G.c:209-273: unnormalised, sign -1 forward.
 * ====================================================================== */

typedef struct {
  int n, nrad, rad[32];
  double *tw; /* tw[2m] = cos(2 pi m/n), tw[2m+1] = sin(2 pi m/n) */
} orc_plan1;

static int plan1_init(orc_plan1 *pl, int n)
{
  static const long double TWO_PI = 6.283185307179586476925286766559005768L;
  int m = n;
  pl->n = n;
  pl->nrad = 0;
  while (m % 4 == 0) { pl->rad[pl->nrad++] = 4; m /= 4; }
  while (m % 2 == 0) { pl->rad[pl->nrad++] = 2; m /= 2; }
  while (m % 3 == 0) { pl->rad[pl->nrad++] = 3; m /= 3; }
  while (m % 5 == 0) { pl->rad[pl->nrad++] = 5; m /= 5; }
  if (m != 1 || n < 1) return -1;
  pl->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    long double a = TWO_PI * (long double)k / (long double)n;
    pl->tw[2 * k] = (double)cosl(a);
    pl->tw[2 * k + 1] = (double)sinl(a);
  }
  return 0;
}

static void plan1_free(orc_plan1 *pl) { free(pl->tw); pl->tw = NULL; }

/* y[q] = sum_r v[r] exp(sign 2 pi i r q / R) for R in {2,3,4,5}. */
static void dft_small(int R, double *vr, double *vi, int sign)
{
  double s = (double)sign;
  if (R == 2) {
    double ar = vr[0], ai = vi[0];
    vr[0] = ar + vr[1]; vi[0] = ai + vi[1];
    vr[1] = ar - vr[1]; vi[1] = ai - vi[1];
  } else if (R == 4) {
    double t0r = vr[0] + vr[2], t0i = vi[0] + vi[2];
    double t1r = vr[0] - vr[2], t1i = vi[0] - vi[2];
    double t2r = vr[1] + vr[3], t2i = vi[1] + vi[3];
    double t3r = vr[1] - vr[3], t3i = vi[1] - vi[3];
    /* (s i) * t3 = (-s t3i, s t3r) */
    vr[0] = t0r + t2r; vi[0] = t0i + t2i;
    vr[2] = t0r - t2r; vi[2] = t0i - t2i;
    vr[1] = t1r - s * t3i; vi[1] = t1i + s * t3r;
    vr[3] = t1r + s * t3i; vi[3] = t1i - s * t3r;
  } else if (R == 3) {
    const double h = 0.86602540378443864676; /* sqrt(3)/2 */
    double tr = vr[1] + vr[2], ti = vi[1] + vi[2];
    double ur = vr[1] - vr[2], ui = vi[1] - vi[2];
    double mr = vr[0] - 0.5 * tr, mi = vi[0] - 0.5 * ti;
    vr[0] += tr; vi[0] += ti;
    /* s i h u = (-s h ui, s h ur) */
    vr[1] = mr - s * h * ui; vi[1] = mi + s * h * ur;
    vr[2] = mr + s * h * ui; vi[2] = mi - s * h * ur;
  } else { /* R == 5, direct sum */
    static const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    static const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    double cr[5] = {1.0, c1, c2, c2, c1};
    double si[5] = {0.0, s1, s2, -s2, -s1};
    double yr[5], yi[5];
    for (int q = 0; q < 5; q++) {
      double ar = 0, ai = 0;
      for (int r = 0; r < 5; r++) {
        int m = (r * q) % 5;
        double wr = cr[m], wi = s * si[m];
        ar += vr[r] * wr - vi[r] * wi;
        ai += vr[r] * wi + vi[r] * wr;
      }
      yr[q] = ar; yi[q] = ai;
    }
    memcpy(vr, yr, sizeof yr);
    memcpy(vi, yi, sizeof yi);
  }
}

/* In-place 1D transform of x[2n] (interleaved) using scratch w[2n]. */
static void fft1d(double *x, double *w, const orc_plan1 *pl, int sign)
{
  const int n = pl->n;
  double *src = x, *dst = w;
  int Ns = 1;
  if (n == 1) return;
  for (int p = 0; p < pl->nrad; p++) {
    const int R = pl->rad[p], m = n / R, step = n / (Ns * R);
    for (int j = 0; j < m; j++) {
      const int k = j % Ns;
      double vr[5] = {0}, vi[5] = {0};
      for (int r = 0; r < R; r++) {
        const int idx = k * r * step;
        const double c = pl->tw[2 * idx], s = sign * pl->tw[2 * idx + 1];
        const double a = src[2 * (j + r * m)], b = src[2 * (j + r * m) + 1];
        vr[r] = a * c - b * s;
        vi[r] = a * s + b * c;
      }
      dft_small(R, vr, vi, sign);
      const int d = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) {
        dst[2 * (d + r * Ns)] = vr[r];
        dst[2 * (d + r * Ns) + 1] = vi[r];
      }
    }
    double *t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

/* Transform along one axis of a [outer][n][inner] complex array. */
static void fft_axis(double *data, long outer, int n, long inner, const orc_plan1 *pl, int sign)
{
  const int PB = 8; /* pencils gathered together for strided axes */
  long nblk = (inner + PB - 1) / PB;
  long total = outer * nblk;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n * PB);
    double *wrk = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
    for (long t = 0; t < total; t++) {
      long o = t / nblk, i0 = (t % nblk) * PB;
      int np = (int)((inner - i0) < PB ? (inner - i0) : PB);
      double *base = data + 2 * (o * (long)n * inner + i0);
      if (inner == 1) {
        fft1d(base, wrk, pl, sign);
        continue;
      }
      for (int q = 0; q < n; q++)
        for (int p = 0; p < np; p++) {
          buf[2 * ((long)p * n + q)] = base[2 * ((long)q * inner + p)];
          buf[2 * ((long)p * n + q) + 1] = base[2 * ((long)q * inner + p) + 1];
        }
      for (int p = 0; p < np; p++) fft1d(buf + 2 * (long)p * n, wrk, pl, sign);
      for (int q = 0; q < n; q++)
        for (int p = 0; p < np; p++) {
          base[2 * ((long)q * inner + p)] = buf[2 * ((long)p * n + q)];
          base[2 * ((long)q * inner + p) + 1] = buf[2 * ((long)p * n + q) + 1];
        }
    }
    free(buf);
    free(wrk);
  }
}

int orc_fft3d(double *data, int n1, int n2, int n3, int batch, int sign)
{
  orc_plan1 p1, p2, p3;
  if (plan1_init(&p1, n1) || plan1_init(&p2, n2) || plan1_init(&p3, n3)) return -1;
  long P = (long)n1 * n2 * n3;
  for (int b = 0; b < batch; b++) {
    double *d = data + 2 * P * b;
    fft_axis(d, (long)n1 * n2, n3, 1, &p3, sign);
    fft_axis(d, n1, n2, n3, &p2, sign);
    fft_axis(d, 1, n1, (long)n2 * n3, &p1, sign);
  }
  plan1_free(&p1); plan1_free(&p2); plan1_free(&p3);
  return 0;
}

/* ======================================================================
 * k-grid and elastic kernels
 * ====================================================================== */

/* frec, G.c:2698-2741: i/(N d) below N/2, (i-N)/(N d) from N/2 up, 0 at i=0. */
double orc_frec1(int i, int n, double d)
{
  if (i == 0) return 0.0;
  if (i < n / 2) return (double)i / (double)n / d;
  return ((double)i - (double)n) / (double)n / d;
}

/* Cijkl of Bmatrix G.c:348-358 (mu = C44, lambda = C12, mu' = C11-C12-2C44). */
void orc_cijkl(double C[81], double C11, double C12, double C44)
{
  double mu = C44, ll = C12, mup = C11 - C12 - 2 * C44;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int m = 0; m < 3; m++) {
          double dik = i == k, djm = j == m, dim = i == m, djk = j == k, dij = i == j, dkm = k == m;
          double d4 = (i == j && j == k && k == m);
          C[((i * 3 + j) * 3 + k) * 3 + m] = mu * (dik * djm + dim * djk) + ll * dij * dkm + mup * d4;
        }
}

/* Green's function of G.c:386 at one k; returns 0 when |k| = 0. */
static int green(const double fk[3], double mu, double ll, double mup, double G[9])
{
  double fk2 = fk[0] * fk[0] + fk[1] * fk[1] + fk[2] * fk[2];
  if (!(fk2 > 0)) return 0;
  double z2[3] = {fk[0] * fk[0] / fk2, fk[1] * fk[1] / fk2, fk[2] * fk[2] / fk2};
  double kk = 1.0 + (mu + ll) * (1 / (mu + mup * z2[0]) * z2[0] + 1 / (mu + mup * z2[1]) * z2[1] +
                                 1 / (mu + mup * z2[2]) * z2[2]);
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 3; i++)
      G[k * 3 + i] = 1 / fk2 * ((i == k ? 1.0 : 0.0) / (mu + mup * z2[k]) -
                                fk[i] * fk[k] / fk2 / (mu + mup * z2[i]) / (mu + mup * z2[k]) * (mu + ll) / kk);
  return 1;
}

/* The literal seven-deep sum for A_mnuv (G.c:383-394), then B (G.c:396-402). */
void orc_bmatrix_literal_at(const double *fk, int npts, const double *eps_a, int na,
                            const double *eps_b, int nb, double C11, double C12, double C44,
                            double *out)
{
  double C[81];
  double mu = C44, ll = C12, mup = C11 - C12 - 2 * C44;
  orc_cijkl(C, C11, C12, C44);
  for (int q = 0; q < npts; q++) {
    const double *f = fk + 3 * q;
    double G[9], A[81];
    double *o = out + (size_t)q * na * nb;
    for (int t = 0; t < na * nb; t++) o[t] = 0.0;
    if (!green(f, mu, ll, mup, G)) continue;
    for (int m = 0; m < 3; m++)
      for (int n = 0; n < 3; n++)
        for (int u = 0; u < 3; u++)
          for (int v = 0; v < 3; v++) {
            double a = 0.0;
            for (int i = 0; i < 3; i++)
              for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++)
                  for (int l = 0; l < 3; l++)
                    a -= C[((k * 3 + l) * 3 + u) * 3 + v] * C[((i * 3 + j) * 3 + m) * 3 + n] *
                         G[k * 3 + i] * f[j] * f[l];
            a += C[((m * 3 + n) * 3 + u) * 3 + v];
            A[((m * 3 + n) * 3 + u) * 3 + v] = a;
          }
    for (int x = 0; x < na; x++)
      for (int y = 0; y < nb; y++) {
        double s = 0.0;
        for (int mn = 0; mn < 9; mn++)
          for (int uv = 0; uv < 9; uv++) s += A[mn * 9 + uv] * eps_a[9 * x + mn] * eps_b[9 * y + uv];
        o[x * nb + y] = s / mu;
      }
  }
}

/* s = C:e (3x3) */
static void c_colon(const double C[81], const double *e, double s[9])
{
  for (int ij = 0; ij < 9; ij++) {
    double a = 0.0;
    for (int mn = 0; mn < 9; mn++) a += C[ij * 9 + mn] * e[mn];
    s[ij] = a;
  }
}

void orc_bmatrix_at(const double *fk, int npts, const double *eps_a, int na, const double *eps_b,
                    int nb, double C11, double C12, double C44, double *out)
{
  double C[81];
  double mu = C44, ll = C12, mup = C11 - C12 - 2 * C44;
  orc_cijkl(C, C11, C12, C44);
  double *sa = (double *)malloc(sizeof(double) * 9 * (na + nb));
  double *sb = sa + 9 * na;
  for (int x = 0; x < na; x++) c_colon(C, eps_a + 9 * x, sa + 9 * x);
  for (int y = 0; y < nb; y++) c_colon(C, eps_b + 9 * y, sb + 9 * y);
  for (int q = 0; q < npts; q++) {
    const double *f = fk + 3 * q;
    double G[9];
    double *o = out + (size_t)q * na * nb;
    if (!green(f, mu, ll, mup, G)) {
      for (int t = 0; t < na * nb; t++) o[t] = 0.0;
      continue;
    }
    for (int x = 0; x < na; x++) {
      double ua[3], gua[3];
      for (int i = 0; i < 3; i++) ua[i] = sa[9 * x + 3 * i] * f[0] + sa[9 * x + 3 * i + 1] * f[1] + sa[9 * x + 3 * i + 2] * f[2];
      for (int k = 0; k < 3; k++) gua[k] = G[k * 3 + 0] * ua[0] + G[k * 3 + 1] * ua[1] + G[k * 3 + 2] * ua[2];
      for (int y = 0; y < nb; y++) {
        double ub[3], se = 0.0;
        for (int i = 0; i < 3; i++) ub[i] = sb[9 * y + 3 * i] * f[0] + sb[9 * y + 3 * i + 1] * f[1] + sb[9 * y + 3 * i + 2] * f[2];
        for (int mn = 0; mn < 9; mn++) se += sa[9 * x + mn] * eps_b[9 * y + mn];
        o[x * nb + y] = (se - (gua[0] * ub[0] + gua[1] * ub[1] + gua[2] * ub[2])) / mu;
      }
    }
  }
  free(sa);
}

/* sum_ij D_b,ij / mu, D_b = A:e_b (Fmatrix G.c:580-591, DD = D/mu G.c:588). */
void orc_ddsum_at(const double *fk, int npts, const double *eps_b, int nb, double C11, double C12,
                  double C44, double *out)
{
  double C[81], T[9];
  double mu = C44, ll = C12, mup = C11 - C12 - 2 * C44;
  orc_cijkl(C, C11, C12, C44);
  for (int pq = 0; pq < 9; pq++) {
    double a = 0.0;
    for (int ij = 0; ij < 9; ij++) a += C[pq * 9 + ij];
    T[pq] = a; /* T_pq = sum_ij C_pqij */
  }
  for (int q = 0; q < npts; q++) {
    const double *f = fk + 3 * q;
    double G[9];
    int nz = green(f, mu, ll, mup, G);
    for (int y = 0; y < nb; y++) {
      if (!nz) { out[(size_t)q * nb + y] = 0.0; continue; }
      double s[9], u[3], gu[3], tot = 0.0;
      c_colon(C, eps_b + 9 * y, s);
      for (int i = 0; i < 3; i++) u[i] = s[3 * i] * f[0] + s[3 * i + 1] * f[1] + s[3 * i + 2] * f[2];
      for (int k = 0; k < 3; k++) gu[k] = G[k * 3] * u[0] + G[k * 3 + 1] * u[1] + G[k * 3 + 2] * u[2];
      for (int ij = 0; ij < 9; ij++) tot += s[ij];
      for (int p = 0; p < 3; p++)
        for (int qq = 0; qq < 3; qq++) tot -= T[p * 3 + qq] * f[qq] * gu[p];
      out[(size_t)q * nb + y] = tot / mu;
    }
  }
}

/* FF_a,ij = f_i (G u_a)_j with u_a = (C:e_a) f  (Fmatrix G.c:565-571). */
void orc_fmatrix_at(const double *fk, int npts, const double *eps_a, int na, double C11, double C12,
                    double C44, double *out)
{
  double C[81];
  double mu = C44, ll = C12, mup = C11 - C12 - 2 * C44;
  orc_cijkl(C, C11, C12, C44);
  for (int q = 0; q < npts; q++) {
    const double *f = fk + 3 * q;
    double G[9];
    int nz = green(f, mu, ll, mup, G);
    for (int x = 0; x < na; x++) {
      double *o = out + ((size_t)q * na + x) * 9;
      if (!nz) { for (int t = 0; t < 9; t++) o[t] = 0.0; continue; }
      double s[9], u[3], gu[3];
      c_colon(C, eps_a + 9 * x, s);
      for (int i = 0; i < 3; i++) u[i] = s[3 * i] * f[0] + s[3 * i + 1] * f[1] + s[3 * i + 2] * f[2];
      for (int j = 0; j < 3; j++) gu[j] = G[j * 3] * u[0] + G[j * 3 + 1] * u[1] + G[j * 3 + 2] * u[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[3 * i + j] = f[i] * gu[j];
    }
  }
}

/* ======================================================================
 * GL loop
 * ====================================================================== */

struct orc_state {
  orc_params p;
  long P;
  int ntri;
  double *xi, *bc, *tau, *data, *data2, *Bsym, *dds;
  int tau_pp;
  double tau_rate[64];  /* uniform-tau ramp: tau_a(s) = tau_a + tau_rate_a * s */
  long ramp_step;
};

static void k_of(const orc_state *s, long pt, double f[3])
{
  int n2 = s->p.n2, n3 = s->p.n3;
  int i = (int)(pt / ((long)n2 * n3)), j = (int)((pt / n3) % n2), k = (int)(pt % n3);
  f[0] = orc_frec1(i, s->p.n1, s->p.d1);
  f[1] = orc_frec1(j, n2, s->p.d2);
  f[2] = orc_frec1(k, n3, s->p.d3);
}

orc_state *orc_create(const orc_params *p)
{
  orc_state *s = (orc_state *)calloc(1, sizeof(*s));
  int S = p->ns;
  s->p = *p;
  s->P = (long)p->n1 * p->n2 * p->n3;
  s->ntri = S * (S + 1) / 2;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  s->xi = (double *)calloc((size_t)2 * S * s->P, sizeof(double));
  s->data = (double *)calloc((size_t)2 * S * s->P, sizeof(double));
  s->data2 = (double *)calloc((size_t)2 * S * s->P, sizeof(double));
  s->tau = (double *)calloc((size_t)S, sizeof(double));
  s->Bsym = (double *)malloc(sizeof(double) * s->ntri * (size_t)s->P);
  s->dds = (double *)malloc(sizeof(double) * S * (size_t)s->P);
  /* B(k) for every k (Bmatrix G.c:328-412), kept as the upper triangle. */
#pragma omp parallel
  {
    double *bfull = (double *)malloc(sizeof(double) * S * S);
#pragma omp for schedule(static)
    for (long pt = 0; pt < s->P; pt++) {
      double f[3];
      k_of(s, pt, f);
      orc_bmatrix_at(f, 1, &p->eps[0][0], S, &p->eps[0][0], S, p->C11, p->C12, p->C44, bfull);
      int t = 0;
      for (int a = 0; a < S; a++)
        for (int b = a; b < S; b++) s->Bsym[pt * s->ntri + t++] = bfull[a * S + b];
      orc_ddsum_at(f, 1, &p->eps[0][0], S, p->C11, p->C12, p->C44, s->dds + pt * S);
    }
    free(bfull);
  }
  return s;
}

void orc_destroy(orc_state *s)
{
  if (!s) return;
  free(s->xi); free(s->bc); free(s->tau); free(s->data); free(s->data2);
  free(s->Bsym); free(s->dds);
  free(s);
}

void orc_set_xi(orc_state *s, const double *xi, const double *xi_bc)
{
  long n = 2 * (long)s->p.ns * s->P;
  memcpy(s->xi, xi, sizeof(double) * n);
  free(s->bc);
  s->bc = NULL;
  if (xi_bc) {
    s->bc = (double *)malloc(sizeof(double) * s->p.ns * (size_t)s->P);
    for (long q = 0; q < s->p.ns * s->P; q++) s->bc[q] = xi_bc[2 * q];
  }
}

void orc_get_xi(const orc_state *s, double *xi)
{
  memcpy(xi, s->xi, sizeof(double) * 2 * (size_t)s->p.ns * s->P);
}

void orc_set_tau(orc_state *s, const double *tau, int per_point)
{
  long n = per_point ? (long)s->p.ns * s->P : s->p.ns;
  free(s->tau);
  s->tau = (double *)malloc(sizeof(double) * n);
  memcpy(s->tau, tau, sizeof(double) * n);
  s->tau_pp = per_point;
  memset(s->tau_rate, 0, sizeof(s->tau_rate));
  s->ramp_step = 0;
}

void orc_set_tau_ramp(orc_state *s, const double *tau0, const double *rate)
{
  orc_set_tau(s, tau0, 0);
  for (int a = 0; a < s->p.ns; a++) s->tau_rate[a] = rate[a];
}

/* data2_a = sum_b B_ab data_b, b ascending (G.c:4609-4629); quirk adds
 * data_a * sum_v DD_a,v (virtualevolv G.c:1426-1441). */
static void contract(orc_state *s, int quirk)
{
  const int S = s->p.ns, nt = s->ntri;
  const long P = s->P;
#pragma omp parallel for schedule(static)
  for (long pt = 0; pt < P; pt++) {
    const double *B = s->Bsym + pt * nt;
    for (int a = 0; a < S; a++) {
      double re = 0.0, im = 0.0;
      for (int b = 0; b < S; b++) {
        int lo = a < b ? a : b, hi = a < b ? b : a;
        double bab = B[lo * S - lo * (lo - 1) / 2 + (hi - lo)];
        re += s->data[2 * (b * P + pt)] * bab;
        im += s->data[2 * (b * P + pt) + 1] * bab;
      }
      if (quirk) {
        double dd = s->dds[pt * S + a];
        re += s->data[2 * (a * P + pt)] * dd;
        im += s->data[2 * (a * P + pt) + 1] * dd;
      }
      s->data2[2 * (a * P + pt)] = re;
      s->data2[2 * (a * P + pt) + 1] = im;
    }
  }
}

static void elastic_solve(orc_state *s, int quirk)
{
  const int S = s->p.ns;
  memcpy(s->data, s->xi, sizeof(double) * 2 * (size_t)S * s->P);       /* G.c:4581-4590 */
  orc_fft3d(s->data, s->p.n1, s->p.n2, s->p.n3, S, -1);                  /* G.c:4599 */
  contract(s, quirk);                                                    /* G.c:4601-4629 */
  orc_fft3d(s->data2, s->p.n1, s->p.n2, s->p.n3, S, +1);                 /* G.c:4638 */
}

void orc_prologue(orc_state *s, int quirk) { elastic_solve(s, quirk); }

double orc_plasticevolv(int ns, long P, double *xi, const double *xi_bc, const double *d2,
                        const double *cd2, const double *asf2, const double *dslip,
                        const double *tau, int tau_per_point, double uq2, double pi_const,
                        double *gamma_a)
{
  /* Update of G.c:3487-3489; gamma sums of G.c:3492-3493 taken over fixed
   * chunks and combined in chunk order so the result does not depend on the
   * thread count. */
  const long CH = 4096;
  const long nch = (P + CH - 1) / CH;
  const double nsize = (double)P;
  double gamma = 0.0;
  double *part = (double *)malloc(sizeof(double) * nch);
  for (int a = 0; a < ns; a++) {
    double *x = xi + 2 * (long)a * P;
    const double *d = d2 + 2 * (long)a * P;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nch; c++) {
      long lo = c * CH, hi = lo + CH < P ? lo + CH : P;
      double acc = 0.0;
      for (long pt = lo; pt < hi; pt++) {
        int evolve = xi_bc ? (xi_bc[2 * ((long)a * P + pt)] == 0.0) : 1;
        if (evolve) {
          double t = tau_per_point ? tau[pt * ns + a] : tau[a];
          x[2 * pt] -= cd2[a] * (uq2 * d[2 * pt] / nsize + asf2[a] * pi_const * sin(2.0 * pi_const * x[2 * pt]) -
                                 t / dslip[a]);
          x[2 * pt + 1] -= uq2 * cd2[a] * (d[2 * pt + 1] / nsize);
        }
        acc += x[2 * pt] / nsize;
      }
      part[c] = acc;
    }
    double ga = 0.0;
    for (long c = 0; c < nch; c++) ga += part[c];
    if (gamma_a) gamma_a[a] = ga;
    gamma += ga;
  }
  free(part);
  return gamma;
}

void orc_gl_steps(orc_state *s, int nsteps, double *gamma, double *gamma_a)
{
  const int S = s->p.ns;
  double *bc2 = NULL;
  if (s->bc) { /* re-expand the mask to the interleaved shape plasticevolv reads */
    bc2 = (double *)calloc((size_t)2 * S * s->P, sizeof(double));
    for (long q = 0; q < S * s->P; q++) bc2[2 * q] = s->bc[q];
  }
  double *tau_eff = (double *)malloc(sizeof(double) * (s->tau_pp ? (size_t)S * s->P : (size_t)S));
  for (int it = 0; it < nsteps; it++) {
    double ga[64];
    const long n = s->tau_pp ? (long)S * s->P : S;
    for (long q = 0; q < n; q++) tau_eff[q] = s->tau[q] + s->tau_rate[q % S] * (double)s->ramp_step;
    s->ramp_step++;
    double g = orc_plasticevolv(S, s->P, s->xi, bc2, s->data2, s->p.cd2, s->p.asf2, s->p.dslip,
                                tau_eff, s->tau_pp, s->p.uq2, s->p.pi_const, ga);
    if (gamma) gamma[it] = g;
    if (gamma_a) memcpy(gamma_a + (size_t)it * S, ga, sizeof(double) * S);
    elastic_solve(s, 0);
  }
  free(tau_eff);
  free(bc2);
}

/* ======================================================================
 * strain / stress (G.c:752-972, 4058-4070)
 * ====================================================================== */
void orc_strain_stress(orc_state *s, const double *avepst_uniform, double *aveps, double *avesigma,
                       double *eps_field, double *sig_field)
{
  const int S = s->p.ns;
  const long P = s->P;
  double C[81], avepst[9];
  orc_cijkl(C, s->p.C11, s->p.C12, s->p.C44);
  double *beta = (double *)calloc((size_t)2 * 9 * P, sizeof(double));
  memcpy(s->data, s->xi, sizeof(double) * 2 * (size_t)S * P);
  orc_fft3d(s->data, s->p.n1, s->p.n2, s->p.n3, S, -1);
  /* beta_ij(k) = sum_a FF_a,ij(k) xi_a(k), field slot i + 3j (G.c:768-781) */
#pragma omp parallel
  {
    double *ff = (double *)malloc(sizeof(double) * 9 * S);
#pragma omp for schedule(static)
    for (long pt = 0; pt < P; pt++) {
      double f[3];
      k_of(s, pt, f);
      orc_fmatrix_at(f, 1, &s->p.eps[0][0], S, s->p.C11, s->p.C12, s->p.C44, ff);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double re = 0.0, im = 0.0;
          for (int a = 0; a < S; a++) {
            re += s->data[2 * (a * P + pt)] * ff[9 * a + 3 * i + j];
            im += s->data[2 * (a * P + pt) + 1] * ff[9 * a + 3 * i + j];
          }
          beta[2 * ((i + 3 * j) * P + pt)] = re;
          beta[2 * ((i + 3 * j) * P + pt) + 1] = im;
        }
    }
    free(ff);
  }
  orc_fft3d(beta, s->p.n1, s->p.n2, s->p.n3, 9, +1);
  /* avepsd = sum_a e_a <Re xi_a> (avestrain G.c:670-686) */
  double mean[64];
  for (int a = 0; a < S; a++) {
    double acc = 0.0;
    for (long pt = 0; pt < P; pt++) acc += s->xi[2 * (a * P + pt)];
    mean[a] = acc / (double)P;
  }
  for (int ij = 0; ij < 9; ij++) {
    double v = avepst_uniform[ij];
    for (int a = 0; a < S; a++) v += s->p.eps[a][ij] * mean[a];
    avepst[ij] = v;
  }
  double se[9] = {0}, ss[9] = {0};
#pragma omp parallel
  {
    double lse[9] = {0}, lss[9] = {0};
#pragma omp for schedule(static)
    for (long pt = 0; pt < P; pt++) {
      double b[9], e[9], estar[9], sg[9];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) b[3 * i + j] = beta[2 * ((i + 3 * j) * P + pt)] / (double)P + avepst[3 * i + j];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) e[3 * i + j] = (b[3 * i + j] + b[3 * j + i]) / 2.;
      for (int kl = 0; kl < 9; kl++) {
        double v = 0.0;
        for (int a = 0; a < S; a++) v += s->p.eps[a][kl] * s->xi[2 * (a * P + pt)];
        estar[kl] = v;
      }
      for (int ij = 0; ij < 9; ij++) {
        double v = 0.0;
        for (int kl = 0; kl < 9; kl++) v += C[ij * 9 + kl] * (e[kl] - estar[kl]);
        sg[ij] = v;
        lse[ij] += e[ij];
        lss[ij] += sg[ij];
        if (eps_field) eps_field[ij * P + pt] = e[ij];
        if (sig_field) sig_field[ij * P + pt] = sg[ij];
      }
    }
#pragma omp critical
    for (int ij = 0; ij < 9; ij++) { se[ij] += lse[ij]; ss[ij] += lss[ij]; }
  }
  for (int ij = 0; ij < 9; ij++) {
    aveps[ij] = se[ij] / (double)P;
    avesigma[ij] = ss[ij] / (double)P;
  }
  free(beta);
}

/* ======================================================================
 * Energy diagnostics on the interface line (avestrain, ResidualEnergy).
 * ====================================================================== */

int orc_line_points(int n1, int n2, int n3, const double n[3], const int pp[3], long *pts, int cap)
{
  int cnt = 0;
  const int k1 = pp[0];
  if (k1 < 0 || k1 >= n1) return 0;
  for (int k2 = 0; k2 < n2; k2++)
    for (int k3 = 0; k3 < n3; k3++)
      if (fabs(n[0] * (k1 - pp[0]) + n[1] * (k2 - pp[1]) + n[2] * (k3 - pp[2])) <= 1.0) {
        if (cnt < cap) pts[cnt] = ((long)k1 * n2 + k2) * n3 + k3;
        cnt++;
      }
  return cnt;
}

/* S_ijkl of avestrain (G.c:656-666) */
static double s_ijkl(int i, int j, int k, int l, double S11, double S12, double S44)
{
  const double dij = i == j, dkl = k == l, dik = i == k, djl = j == l, dil = i == l, djk = j == k;
  const double d4 = (i == j && j == k && k == l) ? 1.0 : 0.0;
  return (S11 + 2 * S12) / 3.0 * dij * dkl + S44 / 4.0 * (dik * djl + dil * djk) - S44 / 2.0 * d4 +
         (S11 - S12) * (d4 - dij * dkl / 3.0);
}

double orc_line_energy(const orc_state *s, const double sigma[9], double S11, double S12, double S44,
                       double mu, const long *pts, int npts, double *e4_inc)
{
  const int S = s->p.ns;
  double e3 = 0.0, e4 = 0.0;
  double ssig = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) ssig += s_ijkl(i, j, k, l, S11, S12, S44) * sigma[3 * k + l] * sigma[3 * i + j];
  for (int q = 0; q < npts; q++) {
    const long pt = pts[q];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double sp = 0.0;
        for (int a = 0; a < S; a++) sp += s->xi[2 * ((long)a * s->P + pt)] * s->p.eps[a][3 * i + j];
        e3 += -sp * sigma[3 * i + j];
      }
    e4 += -ssig * 0.5 * mu;
  }
  if (e4_inc) *e4_inc = e4;
  return e3;
}

double orc_residual_energy(const orc_state *s, const int pp[3], int ns1, double D00, double D01, double D10,
                           double D11)
{
  const int S = s->p.ns, n2 = s->p.n2, n3 = s->p.n3;
  double en = 0.0;
  for (int j = 0; j < n2; j++) {
    const long pt = ((long)pp[0] * n2 + j) * n3 + pp[2];
    double x0 = 0.0, x1 = 0.0;
    for (int a = 0; a < S; a++) {
      const int m = a / ns1;
      if (m == 0) x0 += s->xi[2 * ((long)a * s->P + pt)];
      if (m == 1) x1 += s->xi[2 * ((long)a * s->P + pt)];
    }
    en += fabs(-D11 * x1 * x1 + (D10 + D01) * x1 * x0 - D00 * x0 * x0);
  }
  return en;
}


/*
 * pfdd_setup.c -- host-side setup helpers exported by libpfdd_gpu.so.
 * They restate the reference's setup routines (PFDD_Final_GPU.c) with runtime
 * sizes, plus the seeded loop generator the benchmark configurations use.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>

#include "pfdd_gpu.h"

/* setfcc, G.c:1812-1909.  sqrt(3) is held in a float there (s3), so the
 * normals are 1/(float)sqrt(3), not 1/sqrt(3). */
void pfdd_setfcc(double xn[12][3], double xb[12][3])
{
  static const int nsg[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};
  static const int bv[12][3] = {{-1, 1, 0}, {0, -1, 1}, {1, 0, -1}, {-1, -1, 0}, {1, 0, 1}, {0, -1, 1},
                                {-1, -1, 0}, {1, 0, -1}, {0, -1, -1}, {-1, 1, 0}, {1, 0, 1}, {0, -1, -1}};
  float s3 = (float)sqrt(3.0);
  for (int a = 0; a < 12; a++)
    for (int i = 0; i < 3; i++) {
      xn[a][i] = nsg[a / 3][i] * (1.0 / s3);
      xb[a][i] = bv[a][i] / 2.0;
    }
}

/* setorient normalisation, G.c:2036-2055: each axis row is divided by its
 * squared length. */
void pfdd_setorient(const double rows[3][3], double theta[3][3])
{
  for (int i = 0; i < 3; i++) {
    double r = 0.0;
    for (int j = 0; j < 3; j++) r += rows[i][j] * rows[i][j];
    for (int j = 0; j < 3; j++) theta[i][j] = rows[i][j] / r;
  }
}

/* seteps, G.c:1950-1978: epst = b (x) n / dslip, eps = theta epst theta^T. */
void pfdd_seteps(int ns, const double (*xn)[3], const double (*xb)[3], const double *dslip,
                 const double theta[3][3], double (*eps)[3][3])
{
  for (int a = 0; a < ns; a++) {
    double et[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) et[i][j] = xb[a][i] * xn[a][j] / dslip[a];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double v = 0.0;
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) v += theta[i][k] * theta[j][l] * et[k][l];
        eps[a][i][j] = v;
      }
  }
}

/* resolSS, G.c:1607-1622, for a spatially uniform sigma. */
void pfdd_resolve_tau(int ns, const double sigma[3][3], const double (*eps)[3][3], double *tau)
{
  for (int a = 0; a < ns; a++) {
    double t = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) t = t + sigma[i][j] * eps[a][i][j];
    tau[a] = t;
  }
}

/* SplitMix64 (Steele, Lea & Flood, OOPSLA 2014; constants as published). */
unsigned long long pfdd_splitmix64(unsigned long long *state)
{
  unsigned long long z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double pfdd_uniform01(unsigned long long *state)
{
  return (double)(pfdd_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

static double min_image(double d, int n)
{
  return d - n * floor(d / n + 0.5);
}

long pfdd_add_loop(double *xi, int n1, int n2, int n3, int sys, const double normal[3],
                   const double center[3], double radius)
{
  double len = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
  double nh[3] = {normal[0] / len, normal[1] / len, normal[2] / len};
  double h = fabs(nh[0]);
  if (fabs(nh[1]) > h) h = fabs(nh[1]);
  if (fabs(nh[2]) > h) h = fabs(nh[2]);
  long P = (long)n1 * n2 * n3, count = 0;
  double *x = xi + 2 * (long)sys * P;
  double r2 = radius * radius;
  for (int i = 0; i < n1; i++) {
    double dx = min_image(i - center[0], n1);
    for (int j = 0; j < n2; j++) {
      double dy = min_image(j - center[1], n2);
      for (int k = 0; k < n3; k++) {
        double dz = min_image(k - center[2], n3);
        double hn = nh[0] * dx + nh[1] * dy + nh[2] * dz;
        if (!(fabs(hn) < 0.5 * h)) continue;
        double px = dx - hn * nh[0], py = dy - hn * nh[1], pz = dz - hn * nh[2];
        if (px * px + py * py + pz * pz <= r2) {
          x[2 * (((long)i * n2 + j) * n3 + k)] = 1.0;
          count++;
        }
      }
    }
  }
  return count;
}

int pfdd_seed_loops(double *xi, int n1, int n2, int n3, int ns, const double (*normals)[3], int nloops,
                    double rmin, double rmax, unsigned long long *rng_state, double *loops_out)
{
  if (!xi || !normals || !rng_state || ns < 1 || nloops < 0) return PFDD_E_SHAPE;
  for (int l = 0; l < nloops; l++) {
    double r = rmin + (rmax - rmin) * pfdd_uniform01(rng_state);
    int a = (int)(pfdd_uniform01(rng_state) * ns);
    if (a > ns - 1) a = ns - 1;
    double c[3];
    c[0] = pfdd_uniform01(rng_state) * n1;
    c[1] = pfdd_uniform01(rng_state) * n2;
    c[2] = pfdd_uniform01(rng_state) * n3;
    long cnt = pfdd_add_loop(xi, n1, n2, n3, a, normals[a], c, r);
    if (loops_out) {
      double *o = loops_out + 6 * l;
      o[0] = a; o[1] = c[0]; o[2] = c[1]; o[3] = c[2]; o[4] = r; o[5] = (double)cnt;
    }
  }
  return PFDD_OK;
}

void pfdd_seed_sigma(double amp, unsigned long long *rng_state, double sigma[3][3])
{
  static const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  for (int t = 0; t < 6; t++) {
    double v = -amp + 2 * amp * pfdd_uniform01(rng_state);
    sigma[ij[t][0]][ij[t][1]] = v;
    sigma[ij[t][1]][ij[t][0]] = v;
  }
}

int pfdd_line_points(int n1, int n2, int n3, const double normal[3], const int ppoint[3], long long *pts, int cap)
{
  /* The predicate of avestrain and Imatrix (G.c:707, 714, 2596), evaluated
   * in the same order of double operations so that points on the |.| = 1
   * boundary fall on the same side. */
  const int k1 = ppoint[0];
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || k1 < 0 || k1 >= n1) return 0;
  int cnt = 0;
  for (int k2 = 0; k2 < n2; k2++) {
    const double d12 = normal[0] * (k1 - ppoint[0]) + normal[1] * (k2 - ppoint[1]);
    for (int k3 = 0; k3 < n3; k3++) {
      if (fabs(d12 + normal[2] * (k3 - ppoint[2])) > 1.0) continue;
      if (pts && cnt < cap) pts[cnt] = ((long long)k1 * n2 + k2) * n3 + k3;
      cnt++;
    }
  }
  return cnt;
}

/* plasticconverge, G.c:3732-3785.  A step "hits" when the relative change of
 * gamma is below 1e-8 (or |gamma| < 1e-12; for gamma == 0: |change| < 1e-15
 * or |gamma| < 1e-14).  From step 29 on, a hit extends the run of hits when
 * the last 30 steps all hit and resets it otherwise.  The log line is printed
 * after that update, as in the reference, and the sixth consecutive run ends
 * the loop (return 0). */
int pfdd_plastic_converge(double gamma, double gammalast, int it_plastic, int *testplastic, int *countgamma)
{
  const int nonzero = gamma != 0;
  const double d = nonzero ? (gamma - gammalast) / gamma : gamma - gammalast;
  const int hit = nonzero ? (fabs(d) < 1e-8 || fabs(gamma) < 1e-12) : (fabs(d) < 1e-15 || fabs(gamma) < 1e-14);
  if (hit) {
    testplastic[it_plastic] = 1;
    if (it_plastic >= 29) {
      int run = 0;
      for (int i = 0; i < 30; i++) run += testplastic[it_plastic - i];
      *countgamma = run == 30 ? *countgamma + 1 : 0;
    }
    if (nonzero)
      printf("(gamma-gammalast)/gamma = %e     %d\n", d, *countgamma);
    else
      printf("gamma-gammalast = %e     %d\n", d, *countgamma);
  }
  if (*countgamma == 5) {
    (*countgamma)++;
    return 0;
  }
  return 1;
}

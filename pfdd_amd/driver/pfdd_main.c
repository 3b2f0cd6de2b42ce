/*
 * pfdd -- driver of the PFDD FFT-elasticity + Ginzburg-Landau loop on one
 * MI355X through libpfdd_gpu.so.
 *
 * It restates the control flow of main (PFDD_Final_GPU.c:4087-4847) for the
 * GL path with runtime parameters instead of compile-time #defines: outer
 * steps it = 0..NT-1, strain/stress at the steps where calc_strain_stress
 * runs (G.c:4044), the outer-step prologue with the virtualevolv data2 term
 * (G.c:4450-4452, 1426-1447), and the GL loop (G.c:4565-4691) in fixed-step
 * mode or with plasticconverge (G.c:3732-3785).  interfacevolv and the
 * virtual-strain update are off (single material, xi_v = 0).  Log lines and
 * the files stress-strain.dat, phasefield.dat, pfout.dat, strain1D.dat,
 * stress1D.dat, strain.txt, stress.txt, xi.dat and xi_bc.dat keep the
 * reference formats.
 *
 * With energy=1 (the default) it also writes checkEnergy.dat per outer step
 * and checkEbarrier.dat per GL step from the line sums of Re xi on the
 * interface line (avestrain G.c:697-718).
 *
 *   pfdd [preset] [key=value ...]
 * presets: c0 (32^3 x 3, one loop, 500 steps), c1 (128^3 x 12, 16 loops,
 * 1000 steps), c2 (256^3 x 12, 64 loops, 2000 steps, mixed precision).
 * keys: n1 n2 n3 systems(plane111|fcc) nloops rmin rmax seed sigma_amp tau0
 *       steps nt converge plastic_max precision(fp64|mixed|mixed_lean|ref32) bmode(recompute|stored)
 *       quirk pfout_every fields(1|0) energy(1|0) plan autotune restart(1|0) xi_in outdir
 *
 * Checkpoints (G.c:4361-4382, 4540-4547, 4777-4788): the run ends by writing
 * xi.dat and xi_bc.dat, and xi1.dat / xi_bc1.dat (the itp == 0 copies), as
 * raw doubles [NS+9][N1][N2][N3][2]; the last outer step writes FinalXi_w.dat.
 * restart=1 starts from xi.dat / xi_bc.dat in outdir instead of the seeded
 * loops (the reference's choice == 2; the applied stress is still drawn from
 * the seeded stream); xi_in=FILE reads xi alone from FILE at it == 0, as
 * input_xi_from_file does with FinalXi.dat.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "pfdd_gpu.h"

#define NV 9 /* virtual-strain fields kept in the binary dumps (G.c:88) */

static double wtime(void)
{
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec / 1e6;
}

typedef struct {
  int n1, n2, n3;
  char systems[16];     /* plane111: FCC systems 0-2 rotated so (111) = z; fcc: all 12 */
  int nloops;           /* 0: one loop of radius rmin at the grid centre on system 0 */
  double rmin, rmax;
  unsigned long long seed;
  double sigma_amp;     /* random symmetric applied stress amplitude (fcc) */
  double tau0;          /* resolved shear on system 0 (plane111) */
  int steps;            /* GL steps per plastic outer step (fixed-step mode) */
  int nt;               /* outer steps NT (G.c:90); GL loops run for 0 < it < NT-1 */
  int converge;         /* 1: stop on plasticconverge, at most plastic_max steps */
  int plastic_max;
  int precision, bmode, quirk;
  int plan;             /* kernel plan word (PFDD_PLAN_*), 0 = AUTO */
  int autotune;         /* > 0: pfdd_autotune over that many GL steps before the run */
  int pfout_every;      /* 0: no pfout.dat */
  int fields;           /* 1: strain1D.dat / stress1D.dat per strain call, strain.txt / stress.txt at the end */
  int energy;           /* 1: checkEnergy.dat per outer step, checkEbarrier.dat per GL step */
  int restart;          /* 1: read xi / xi_bc from xi.dat / xi_bc.dat in outdir (choice == 2, G.c:4366-4382) */
  char xi_in[512];      /* non-empty: read xi from this file at it == 0 (input_xi_from_file, G.c:3966-3987) */
  char outdir[512];
} config;

static void preset(config *c, const char *name)
{
  memset(c, 0, sizeof(*c));
  c->nt = 3;
  c->plastic_max = 3000;
  c->quirk = 1;
  c->fields = 1;
  c->energy = 1;
  c->precision = PFDD_FP64;
  c->bmode = PFDD_B_RECOMPUTE;
  c->seed = 20261202ULL;
  strcpy(c->outdir, ".");
  if (!strcmp(name, "c0")) {
    c->n1 = c->n2 = c->n3 = 32;
    strcpy(c->systems, "plane111");
    c->nloops = 0;
    c->rmin = c->rmax = 8.0;
    c->tau0 = 0.01;
    c->steps = 500;
  } else if (!strcmp(name, "c1") || !strcmp(name, "c2")) {
    int big = !strcmp(name, "c2");
    c->n1 = c->n2 = c->n3 = big ? 256 : 128;
    strcpy(c->systems, "fcc");
    c->nloops = big ? 64 : 16;
    c->rmin = 5.0;
    c->rmax = 20.0;
    c->seed += big ? 2 : 1;
    c->sigma_amp = 0.02;
    c->steps = big ? 2000 : 1000;
    if (big) c->precision = PFDD_MIXED;
  } else {
    fprintf(stderr, "pfdd: unknown preset %s\n", name);
    exit(2);
  }
}

static void set_key(config *c, const char *kv)
{
  char key[64];
  const char *eq = strchr(kv, '=');
  if (!eq || eq - kv >= (long)sizeof(key)) {
    fprintf(stderr, "pfdd: expected key=value, got %s\n", kv);
    exit(2);
  }
  memcpy(key, kv, eq - kv);
  key[eq - kv] = 0;
  const char *v = eq + 1;
  if (!strcmp(key, "n1")) c->n1 = atoi(v);
  else if (!strcmp(key, "n2")) c->n2 = atoi(v);
  else if (!strcmp(key, "n3")) c->n3 = atoi(v);
  else if (!strcmp(key, "systems")) snprintf(c->systems, sizeof(c->systems), "%s", v);
  else if (!strcmp(key, "nloops")) c->nloops = atoi(v);
  else if (!strcmp(key, "rmin")) c->rmin = atof(v);
  else if (!strcmp(key, "rmax")) c->rmax = atof(v);
  else if (!strcmp(key, "seed")) c->seed = strtoull(v, NULL, 10);
  else if (!strcmp(key, "sigma_amp")) c->sigma_amp = atof(v);
  else if (!strcmp(key, "tau0")) c->tau0 = atof(v);
  else if (!strcmp(key, "steps")) c->steps = atoi(v);
  else if (!strcmp(key, "nt")) c->nt = atoi(v);
  else if (!strcmp(key, "converge")) c->converge = atoi(v);
  else if (!strcmp(key, "plastic_max")) c->plastic_max = atoi(v);
  else if (!strcmp(key, "precision"))
    c->precision = !strcmp(v, "mixed")        ? PFDD_MIXED
                   : !strcmp(v, "mixed_lean") ? PFDD_MIXED_LEAN
                   : !strcmp(v, "ref32")      ? PFDD_REF32
                                              : PFDD_FP64;
  else if (!strcmp(key, "bmode")) c->bmode = !strcmp(v, "stored") ? PFDD_B_STORED : PFDD_B_RECOMPUTE;
  else if (!strcmp(key, "quirk")) c->quirk = atoi(v);
  else if (!strcmp(key, "plan")) c->plan = (int)strtol(v, NULL, 0);
  else if (!strcmp(key, "autotune")) c->autotune = atoi(v);
  else if (!strcmp(key, "pfout_every")) c->pfout_every = atoi(v);
  else if (!strcmp(key, "fields")) c->fields = atoi(v);
  else if (!strcmp(key, "energy")) c->energy = atoi(v);
  else if (!strcmp(key, "restart")) c->restart = atoi(v);
  else if (!strcmp(key, "xi_in")) snprintf(c->xi_in, sizeof(c->xi_in), "%s", v);
  else if (!strcmp(key, "outdir")) snprintf(c->outdir, sizeof(c->outdir), "%s", v);
  else {
    fprintf(stderr, "pfdd: unknown key %s\n", key);
    exit(2);
  }
}

/* Field outputs of calc_strain_stress (G.c:4039-4085).  e and s are the
 * strain and stress fields [3][3][N1][N2][N3] of pfdd_strain_stress; mu
 * divides the stress for the sigma columns. */
static double fld(const double *f, int i, int j, long pt, long P) { return f[(long)(3 * i + j) * P + pt]; }

/* strain1D.dat / stress1D.dat: the 33 component along N1 at (N2/2, N3/2)
 * (G.c:4074-4083; the zone header names N3 as the reference's does). */
static void write_1d(FILE *f, const double *x, double scale, int n1, int n2, int n3)
{
  const long P = (long)n1 * n2 * n3;
  fprintf(f, "zone   I = %d \n", n3);
  for (int k1 = 0; k1 < n1; k1++)
    fprintf(f, "%d %lf \n", k1, fld(x, 2, 2, ((long)k1 * n2 + n2 / 2) * n3 + n3 / 2, P) * scale);
}

/* strain.txt (G.c:818-831): every point, the nine components; the header is
 * the reference's (eight names for nine columns). */
static void write_strain_txt(FILE *f, const double *e, int n1, int n2, int n3)
{
  const long P = (long)n1 * n2 * n3;
  fprintf(f, "x,y,z,eps11,eps13,eps21,eps22,eps23,eps31,eps32,eps33\n");
  for (int k1 = 0; k1 < n1; k1++)
    for (int k2 = 0; k2 < n2; k2++)
      for (int k3 = 0; k3 < n3; k3++) {
        const long pt = ((long)k1 * n2 + k2) * n3 + k3;
        fprintf(f, "%d,%d,%d,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf \n", k1, k2, k3, fld(e, 0, 0, pt, P),
                fld(e, 0, 1, pt, P), fld(e, 0, 2, pt, P), fld(e, 1, 0, pt, P), fld(e, 1, 1, pt, P),
                fld(e, 1, 2, pt, P), fld(e, 2, 0, pt, P), fld(e, 2, 1, pt, P), fld(e, 2, 2, pt, P));
      }
}

/* stress.txt (G.c:927-944): the k2 = 0 plane.  The reference's first six
 * columns are the virtual-field stress sigmav = sigma_applied - IFFT(sum_b
 * B_bv xi_b)/P of virtualevolv (G.c:1455-1464); with the virtual eigenstrain
 * e_u (x) e_v, sum_b B_bv xi_b(k) = -sigma_uv(k)/mu for k != 0, so that
 * field is the stress C:(eps - eps*) over mu, which is what is written; the
 * last six columns are C:(eps - eps*) itself (datasigma). */
static void write_stress_txt(FILE *f, const double *s, double mu, int n1, int n2, int n3)
{
  const long P = (long)n1 * n2 * n3;
  const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
  fprintf(f, "x,y,z,sigma11,sigma12,sigma13,sigma22,sigma23,sigma33,datasigma11,datasigma12,datasigma13,"
             "datasigma22,datasigma23,datasigma33 \n");
  for (int k1 = 0; k1 < n1; k1++)
    for (int k3 = 0; k3 < n3; k3++) {
      const long pt = (long)k1 * n2 * n3 + k3;  /* k2 = 0 */
      fprintf(f, "%d,%d,%d", k1, 0, k3);
      for (int t = 0; t < 6; t++) fprintf(f, ",%lf", fld(s, ii[t], jj[t], pt, P) / mu);
      for (int t = 0; t < 6; t++) fprintf(f, ",%lf", fld(s, ii[t], jj[t], pt, P));
      fprintf(f, "\n");
    }
}

static FILE *open_out(const config *c, const char *name, const char *mode)
{
  char path[1024];
  snprintf(path, sizeof(path), "%s/%s", c->outdir, name);
  FILE *f = fopen(path, mode);
  if (!f) {
    perror(path);
    exit(1);
  }
  return f;
}

static void die(int rc, const char *what)
{
  if (rc != PFDD_OK) {
    fprintf(stderr, "pfdd: %s failed: %s (%d)\n", what, pfdd_strerror(rc), rc);
    exit(1);
  }
}

/* Reads n doubles of a binary dump ([NSV][N1][N2][N3][2], G.c:4366-4369). */
static void read_dump(const char *path, double *x, size_t n)
{
  FILE *f = fopen(path, "rb");
  if (!f) {
    perror(path);
    exit(1);
  }
  const size_t got = fread(x, sizeof(double), n, f);
  fclose(f);
  if (got != n) {
    fprintf(stderr, "pfdd: %s holds %zu of the %zu doubles of [NS+9][N1][N2][N3][2]\n", path, got, n);
    exit(1);
  }
}

static void write_dump(const config *c, const char *name, const double *x, size_t n)
{
  FILE *f = open_out(c, name, "wb");
  if (fwrite(x, sizeof(double), n, f) != n) {
    fprintf(stderr, "pfdd: short write of %s\n", name);
    exit(1);
  }
  fclose(f);
}

/* The compliance S_ijkl of avestrain (G.c:652-666), cubic form. */
static double s_ijkl(const pfdd_params *p, int i, int j, int k, int l)
{
  double C11 = p->C11, C12 = p->C12, C44 = p->C44;
  double S11 = 1. / 3. * (1 / (C11 + 2 * C12) + 2 / (C11 - C12));
  double S12 = 1. / 3. * (1 / (C11 + 2 * C12) - 1 / (C11 - C12));
  double S44 = 1 / C44;
  double dij = i == j, dkl = k == l, dik = i == k, djl = j == l, dil = i == l, djk = j == k;
  double d4 = (i == j && j == k && k == l);
  return (S11 + 2 * S12) / 3.0 * dij * dkl + S44 / 4.0 * (dik * djl + dil * djk) - S44 / 2.0 * d4 +
         (S11 - S12) * (d4 - dij * dkl / 3.0);
}

/* S_ijkl:sigma*mu, the xi-independent part of avepst (avestrain G.c:720-729). */
static void avepst_uniform(const pfdd_params *p, const double sigma[3][3], double out[9])
{
  const double mu = p->C44;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0.0;
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) v += s_ijkl(p, i, j, k, l) * sigma[k][l] * mu;
      out[3 * i + j] = v;
    }
}

/* Energy diagnostics on the interface line (avestrain G.c:697-718 with a
 * uniform applied stress and no virtual strain): energy_in3 =
 * -sum_a (e_a : sigma) L_a with L_a the line sum of Re xi_a, and the increment
 * of energy_in4 per avestrain call, -mu/2 npts sigma:S:sigma.  energy_in
 * (Energy_calculation) and energy_in2 (Imatrix) read only the virtual fields
 * and ResidualEnergy needs two materials; all three are 0 here. */
typedef struct {
  int npts;
  double e_sig[PFDD_MAX_NS]; /* e_a : sigma */
  double e4_inc;
} line_energy;

static void line_energy_init(line_energy *le, const pfdd_params *p, const double sigma[3][3], int npts)
{
  const double mu = p->C44;
  double ssig = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) ssig += s_ijkl(p, i, j, k, l) * sigma[k][l] * sigma[i][j];
  le->npts = npts;
  le->e4_inc = -ssig * 0.5 * mu * npts;
  for (int a = 0; a < p->ns; a++) {
    double v = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) v += p->eps[a][i][j] * sigma[i][j];
    le->e_sig[a] = v;
  }
}

static double line_e3(const line_energy *le, int ns, const double *lsum)
{
  double e3 = 0.0;
  for (int a = 0; a < ns; a++) e3 += -le->e_sig[a] * lsum[a];
  return e3;
}

int main(int argc, char **argv)
{
  config cfg;
  int argi = 1;
  if (argc > 1 && !strchr(argv[1], '=')) preset(&cfg, argv[argi++]);
  else preset(&cfg, "c0");
  for (; argi < argc; argi++) set_key(&cfg, argv[argi]);

  double tmstart = wtime();
  const double nu = 0.33, mu = 1.0, lam = 2 * nu / (1 - 2 * nu);
  pfdd_params p;
  memset(&p, 0, sizeof(p));
  p.n1 = cfg.n1; p.n2 = cfg.n2; p.n3 = cfg.n3;
  p.precision = cfg.precision;
  p.bmode = cfg.bmode;
  p.plan = cfg.plan;
  p.C44 = mu; p.C12 = lam;
  p.C11 = 2. * p.C44 + p.C12; /* MT == 1: isotropic (G.c:4243-4245) */
  p.d1 = p.d2 = p.d3 = 1.0;
  p.uq2 = 1.0;
  p.pi_const = 3.141592654;
  printf("This is an isotropic material.\n");
  printf("Material constants for material 0\n");
  printf("C11 = %e, C12 = %e, C44 = %e\n", p.C11, p.C12, p.C44);

  /* slip systems and eigenstrains (setfcc G.c:1812, setorient G.c:1986, seteps G.c:1912) */
  double xn12[12][3], xb12[12][3], theta[3][3], normals[PFDD_MAX_NS][3], dslip[PFDD_MAX_NS];
  pfdd_setfcc(xn12, xb12);
  int nsys = 0;
  double rows[3][3];
  if (!strcmp(cfg.systems, "plane111")) {
    const double a = 1 / sqrt(2.), b = 1 / sqrt(6.), t = 1 / sqrt(3.);
    double r[3][3] = {{-a, a, 0.0}, {-b, -b, 2 * b}, {t, t, t}};
    memcpy(rows, r, sizeof(rows));
    nsys = 3;
  } else if (!strcmp(cfg.systems, "fcc")) {
    double r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    memcpy(rows, r, sizeof(rows));
    nsys = 12;
  } else {
    fprintf(stderr, "pfdd: unknown systems %s\n", cfg.systems);
    return 2;
  }
  pfdd_setorient(rows, theta);
  p.ns = nsys;
  for (int a = 0; a < nsys; a++) {
    dslip[a] = 1.0;
    p.cd2[a] = 0.1;                    /* G.c:4290 */
    p.asf2[a] = 0.0262;                /* G.c:4279-4282 */
    p.dslip[a] = 1.0;
    for (int i = 0; i < 3; i++) {
      normals[a][i] = 0.0;
      for (int k = 0; k < 3; k++) normals[a][i] += theta[i][k] * xn12[a][k];
    }
  }
  for (int a = 0; a < nsys && a < 2; a++) p.asf2[a] = 0.0064659518668288441;
  pfdd_seteps(nsys, (const double(*)[3])xn12, (const double(*)[3])xb12, dslip, theta, p.eps);
  printf("Slip systems %d\n", nsys);

  /* initial xi (seeded loops) and applied stress */
  const long P = (long)p.n1 * p.n2 * p.n3;
  const int nsv = nsys + NV;
  double *xi = calloc((size_t)2 * nsv * P, sizeof(double));
  double *xi_bc = calloc((size_t)2 * nsv * P, sizeof(double));
  if (!xi || !xi_bc) {
    fprintf(stderr, "pfdd: host allocation failed\n");
    return 1;
  }
  for (long q = (long)nsys * P; q < (long)nsv * P; q++) xi_bc[2 * q] = 1.0; /* virtual fields frozen */
  if (!cfg.restart) printf(" Generating initial data\n"); /* choice == 1 (G.c:4347-4348) */
  double sigma[3][3] = {{0}};
  unsigned long long rng = cfg.seed;
  if (cfg.nloops == 0) {
    double c[3] = {p.n1 / 2.0, p.n2 / 2.0, p.n3 / 2.0};
    pfdd_add_loop(xi, p.n1, p.n2, p.n3, 0, normals[0], c, cfg.rmin);
  } else {
    die(pfdd_seed_loops(xi, p.n1, p.n2, p.n3, nsys, (const double(*)[3])normals, cfg.nloops, cfg.rmin, cfg.rmax,
                        &rng, NULL), "pfdd_seed_loops");
  }
  if (!strcmp(cfg.systems, "plane111")) {
    double se[3][3], se_e = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) se[i][j] = 0.5 * (p.eps[0][i][j] + p.eps[0][j][i]);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) se_e += se[i][j] * p.eps[0][i][j];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) sigma[i][j] = cfg.tau0 * se[i][j] / se_e;
  } else {
    pfdd_seed_sigma(cfg.sigma_amp, &rng, sigma);
  }
  if (cfg.restart) { /* choice == 2: xi and xi_bc from the dumps of an earlier run (G.c:4366-4369) */
    char path[1100];
    snprintf(path, sizeof(path), "%s/xi.dat", cfg.outdir);
    read_dump(path, xi, (size_t)2 * nsv * P);
    snprintf(path, sizeof(path), "%s/xi_bc.dat", cfg.outdir);
    read_dump(path, xi_bc, (size_t)2 * nsv * P);
  }
  double tau[PFDD_MAX_NS];
  pfdd_resolve_tau(nsys, (const double(*)[3])sigma, (const double(*)[3][3])p.eps, tau);
  printf("Resolved shear stresses\n");

  printf("Frequencies\n");
  printf("Set interaction matrix\n");
  pfdd_ctx *ctx = NULL;
  die(pfdd_create(&p, &ctx), "pfdd_create");
  die(pfdd_set_tau(ctx, tau, 0), "pfdd_set_tau");
  die(pfdd_upload_xi(ctx, xi, xi_bc, nsv), "pfdd_upload_xi");
  if (cfg.autotune > 0) {
    int plan = 0;
    float ms = 0.f;
    die(pfdd_autotune(ctx, cfg.autotune, &plan, &ms), "pfdd_autotune");
    printf("pfdd: autotuned plan 0x%x, %.4f ms per GL step\n", plan, ms);
  }

  FILE *of2 = open_out(&cfg, "stress-strain.dat", "w");
  FILE *ofpfv = open_out(&cfg, "phasefield.dat", "w");
  FILE *ofpfout = cfg.pfout_every > 0 ? open_out(&cfg, "pfout.dat", "w") : NULL;
  FILE *of4 = cfg.fields ? open_out(&cfg, "strain1D.dat", "w") : NULL;
  FILE *of6 = cfg.fields ? open_out(&cfg, "stress1D.dat", "w") : NULL;
  double *efield = cfg.fields ? malloc(sizeof(double) * 9 * (size_t)P) : NULL;
  double *sfield = cfg.fields ? malloc(sizeof(double) * 9 * (size_t)P) : NULL;
  if (cfg.fields && (!efield || !sfield)) {
    fprintf(stderr, "pfdd: out of host memory for the strain/stress fields\n");
    return 1;
  }
  double ap_u[9];
  avepst_uniform(&p, sigma, ap_u);
  int maxsteps = cfg.converge ? cfg.plastic_max : cfg.steps;
  double *gam = malloc(sizeof(double) * (maxsteps > 0 ? maxsteps : 1));
  double *gam_a = malloc(sizeof(double) * nsys * (maxsteps > 0 ? maxsteps : 1));
  int *testplastic = calloc(cfg.plastic_max > 0 ? cfg.plastic_max : 1, sizeof(int));
  double gamma = 0.0, gammalast = 0.0, gamma1[PFDD_MAX_NS] = {0};
  long total_steps = 0;
  double gl_time = 0.0;

  /* interface line of the energy diagnostics: interface_n and ppoint of main (G.c:4200-4203, 4234-4236) */
  line_energy le;
  memset(&le, 0, sizeof(le));
  FILE *ofE = NULL, *ofB = NULL;
  double *lsum = NULL, e4 = 0.0;
  int it_checkEbarrier = 0;
  if (cfg.energy) {
    const double inorm[3] = {sqrt(3.) / 3., 0.0, sqrt(6.) / 3.};
    const int pp[3] = {p.n1 / 2, 0, p.n3 * 3 / 4};
    const int np = pfdd_line_points(p.n1, p.n2, p.n3, inorm, pp, NULL, 0);
    long long *pts = malloc(sizeof(long long) * (np > 0 ? np : 1));
    lsum = malloc(sizeof(double) * nsys * (maxsteps > 0 ? maxsteps : 1));
    if (!pts || !lsum) {
      fprintf(stderr, "pfdd: host allocation failed\n");
      return 1;
    }
    pfdd_line_points(p.n1, p.n2, p.n3, inorm, pp, pts, np);
    die(pfdd_set_line_points(ctx, pts, np), "pfdd_set_line_points");
    free(pts);
    line_energy_init(&le, &p, (const double(*)[3])sigma, np);
    ofE = open_out(&cfg, "checkEnergy.dat", "w");
    ofB = open_out(&cfg, "checkEbarrier.dat", "w");
  }

  printf("start time step \n");
  for (int it = 0; it < cfg.nt; it++) {
    const int itp = 0, itp2 = 0;
    printf("time step, %d %d %d\n", it, itp, itp2);
    printf("xi = %e\n", 0.0); /* xi[2*NS*N1*N2*N3]: the first virtual field */
    /* calc_strain_stress (G.c:4044) */
    if (it == cfg.nt - 1 || (it != 0 && it != cfg.nt - 1)) {
      double aveps[9], avesig[9];
      die(pfdd_strain_stress(ctx, ap_u, aveps, avesig, efield, sfield), "pfdd_strain_stress");
      if (cfg.fields && it == cfg.nt - 1) {
        printf("printing strains %d %d\n", it, itp);
        FILE *of3 = open_out(&cfg, "strain.txt", "w");
        write_strain_txt(of3, efield, p.n1, p.n2, p.n3);
        fclose(of3);
        printf("printing stress %d %d\n", it, itp);
        FILE *of5 = open_out(&cfg, "stress.txt", "w");
        write_stress_txt(of5, sfield, mu, p.n1, p.n2, p.n3);
        fclose(of5);
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) fprintf(of2, "%e %e ", aveps[3 * i + j], avesig[3 * i + j] / mu);
      if (cfg.fields) {
        write_1d(of4, efield, 1.0, p.n1, p.n2, p.n3);
        write_1d(of6, sfield, 1.0 / mu, p.n1, p.n2, p.n3);
      }
    }
    if (it == 0 && cfg.xi_in[0]) { /* input_xi_from_file (G.c:3966-3987): xi only, xi_bc kept */
      read_dump(cfg.xi_in, xi, (size_t)2 * nsv * P);
      die(pfdd_upload_xi(ctx, xi, xi_bc, nsv), "pfdd_upload_xi");
    }
    /* fft_forward + calc_data2_datag + virtualevolv data2 term */
    die(pfdd_begin_outer(ctx, cfg.quirk), "pfdd_begin_outer");
    if (cfg.energy) { /* avestrain with energy_in4 reset (G.c:4439-4442), checkEnergy line (G.c:4530-4536) */
      double ls[PFDD_MAX_NS];
      die(pfdd_line_sums(ctx, ls), "pfdd_line_sums");
      const double e3 = line_e3(&le, nsys, ls);
      e4 = le.e4_inc;
      fprintf(ofE, "%d    %lf   %lf   %lf   %lf   %lf   %lf\n", it, 0.0, 0.0, e3, e4, 0.0, e3 + e4);
    }
    if (it == cfg.nt - 1) { /* FinalXi_w.dat (G.c:4540-4547) */
      die(pfdd_download_xi(ctx, xi, nsv), "pfdd_download_xi");
      write_dump(&cfg, "FinalXi_w.dat", xi, (size_t)2 * nsv * P);
    }
    printf("vflag %d, time %d \n", 0, it);
    if (it != 0 && it != cfg.nt - 1) {
      printf("here plastic!\n");
      printf("evolve plastic then evolve interface\n");
      int countgamma = 0, it_plastic = 0, check = 1;
      memset(testplastic, 0, sizeof(int) * (cfg.plastic_max > 0 ? cfg.plastic_max : 1));
      double t0 = wtime();
      if (!cfg.converge) {
        if (cfg.energy) die(pfdd_gl_steps_lines(ctx, cfg.steps, gam, gam_a, lsum), "pfdd_gl_steps_lines");
        else die(pfdd_gl_steps(ctx, cfg.steps, gam, gam_a), "pfdd_gl_steps");
        gl_time += wtime() - t0;
        for (int s = 0; s < cfg.steps; s++) {
          gamma = gam[s];
          printf("in evolve plastic:    %d    %d\n", s, 1);
          if (cfg.energy) { /* G.c:4646-4650 after every GL step */
            const double e3 = line_e3(&le, nsys, lsum + (long)s * nsys);
            e4 += le.e4_inc;
            fprintf(ofB, "%d   %lf   %lf\n", it_checkEbarrier++, 0.0, fabs(e3 + e4));
          }
          fprintf(ofpfv, "%d       %d         %e           %e           %e\n", it, s + 1, gam_a[(long)s * nsys],
                  nsys > 1 ? gam_a[(long)s * nsys + 1] : 0.0, gamma);
        }
        for (int a = 0; a < nsys; a++) gamma1[a] = cfg.steps > 0 ? gam_a[(long)(cfg.steps - 1) * nsys + a] : 0.0;
        total_steps += cfg.steps;
      } else {
        do { /* G.c:4574-4691, one GL step per call */
          gammalast = gamma;
          if (cfg.energy) die(pfdd_gl_steps_lines(ctx, 1, gam, gam_a, lsum), "pfdd_gl_steps_lines");
          else die(pfdd_gl_steps(ctx, 1, gam, gam_a), "pfdd_gl_steps");
          gamma = gam[0];
          for (int a = 0; a < nsys; a++) gamma1[a] = gam_a[a];
          check = pfdd_plastic_converge(gamma, gammalast, it_plastic, testplastic, &countgamma);
          printf("in evolve plastic:    %d    %d\n", it_plastic, check);
          if (cfg.energy && check) { /* only while plasticconverge says go on (G.c:4589, 4646-4650) */
            const double e3 = line_e3(&le, nsys, lsum);
            e4 += le.e4_inc;
            fprintf(ofB, "%d   %lf   %lf\n", it_checkEbarrier++, 0.0, fabs(e3 + e4));
          }
          it_plastic++;
          total_steps++;
          fprintf(ofpfv, "%d       %d         %e           %e           %e\n", it, it_plastic, gamma1[0],
                  nsys > 1 ? gamma1[1] : 0.0, gamma);
          if (ofpfout && (it_plastic % cfg.pfout_every == 0 || it_plastic == cfg.plastic_max || !check)) {
            die(pfdd_download_xi(ctx, xi, nsv), "pfdd_download_xi");
            fprintf(ofpfout, "zone   I = %d K = %d\n", p.n3, p.n1);
            for (int i = 0; i < p.n1; i++)
              for (int k = 0; k < p.n3; k++) fprintf(ofpfout, "%d   %d   %lf\n", k, i, xi[2 * ((long)i * p.n2 * p.n3 + k)]);
          }
        } while (check && it_plastic < cfg.plastic_max);
        gl_time += wtime() - t0;
      }
    }
    if (it == cfg.nt - 1) {
      for (int a = 0; a < nsys; a++) fprintf(of2, "\t%e", gamma1[a]);
      fprintf(of2, " %e \n", gamma);
    }
  }
  /* binary dumps xi.dat / xi_bc.dat and, for itp == 0, xi1.dat / xi_bc1.dat,
   * [NSV][N1][N2][N3][2] doubles (G.c:4777-4788) */
  die(pfdd_download_xi(ctx, xi, nsv), "pfdd_download_xi");
  write_dump(&cfg, "xi.dat", xi, (size_t)2 * nsv * P);
  write_dump(&cfg, "xi_bc.dat", xi_bc, (size_t)2 * nsv * P);
  write_dump(&cfg, "xi1.dat", xi, (size_t)2 * nsv * P);
  write_dump(&cfg, "xi_bc1.dat", xi_bc, (size_t)2 * nsv * P);
  fclose(of2);
  fclose(ofpfv);
  if (ofpfout) fclose(ofpfout);
  if (of4) fclose(of4);
  if (of6) fclose(of6);
  if (ofE) fclose(ofE);
  if (ofB) fclose(ofB);
  free(lsum);
  free(efield);
  free(sfield);
  pfdd_destroy(ctx);
  if (gl_time > 0)
    printf("pfdd: %ld GL steps in %.3f s (%.1f GL steps/s)\n", total_steps, gl_time, total_steps / gl_time);
  printf("wall time %fs\n", wtime() - tmstart);
  free(xi);
  free(xi_bc);
  free(gam);
  free(gam_a);
  free(testplastic);
  return 0;
}

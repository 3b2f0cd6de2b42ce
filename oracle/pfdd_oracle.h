/*
 * pfdd_oracle.h -- CPU restatement of the PFDD FFT-elasticity + Ginzburg-Landau
 * step (reference: PFDD_Final_GPU.c, abbreviated G.c below).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * pfdd_amd/: it is loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and by nothing else.  The product never links,
 * loads or falls back to it.
 *
 * Everything is FP64.  Complex arrays are interleaved (re, im) doubles in the
 * reference layout [field][N1][N2][N3][2] (G.c:38-43), N3 fastest.
 *
 * Where each function is pinned (which test checks it against what) is listed
 * in the header comment of pfdd_oracle.c.
 */
#ifndef PFDD_ORACLE_H
#define PFDD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Batched 3D complex DFT, in place, unnormalised (G.c:209-273).
 * sign = -1: forward, exp(-2 pi i k.x/N); sign = +1: backward.
 * Sizes must factor into 2, 3 and 5.  Returns 0 or -1 on a bad size. */
int orc_fft3d(double *data, int n1, int n2, int n3, int batch, int sign);

/* 1D frequency of index i on an n-point axis with spacing d (frec, G.c:2683-2754). */
double orc_frec1(int i, int n, double d);

/* Isotropic/cubic stiffness C_ijkl, row-major [3][3][3][3] (Bmatrix G.c:348-358). */
void orc_cijkl(double C[81], double C11, double C12, double C44);

/* B_ab(k) evaluated by the literal loop nest of Bmatrix (G.c:361-410) at npts
 * explicit k-vectors fk[npts][3]; out[npts][na][nb] = sum A_mnuv ea_mn eb_uv / mu. */
void orc_bmatrix_literal_at(const double *fk, int npts,
                            const double *eps_a, int na,
                            const double *eps_b, int nb,
                            double C11, double C12, double C44, double *out);

/* Same quantity in the factored form B_ab = (s_a:e_b - u_a.G.u_b)/mu with
 * s_a = C:e_a, u_a = s_a f; used by the GL loop below. */
void orc_bmatrix_at(const double *fk, int npts,
                    const double *eps_a, int na,
                    const double *eps_b, int nb,
                    double C11, double C12, double C44, double *out);

/* sum_{ij} D_b,ij(k) / mu with D_b = A(k):e_b (Fmatrix G.c:580-591), the
 * per-field factor that virtualevolv adds to data2 (G.c:1426-1441). */
void orc_ddsum_at(const double *fk, int npts, const double *eps_b, int nb,
                  double C11, double C12, double C44, double *out /*[npts][nb]*/);

/* FF_a,ij(k) = sum C_klmn G_jk e_a,mn f_i f_l  (Fmatrix G.c:565-571). */
void orc_fmatrix_at(const double *fk, int npts, const double *eps_a, int na,
                    double C11, double C12, double C44, double *out /*[npts][na][3][3]*/);

/* ---- GL loop state ------------------------------------------------------ */
typedef struct {
  int n1, n2, n3, ns;
  double d1, d2, d3;
  double C11, C12, C44;
  double cd2[64], asf2[64], dslip[64];   /* per slip system */
  double uq2;                            /* 1.0 (G.c:4195) */
  double pi_const;                       /* 3.141592654 (G.c:95) */
  double eps[64][9];                     /* eigenstrain of each system, row-major 3x3 */
  int nthreads;                          /* OpenMP threads, 0 = default */
} orc_params;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);

/* Load the complex order parameter xi[ns][P][2] and the evolve mask
 * (xi_bc[ns][P][2], only the real slot is read, G.c:3480; NULL = all evolve)
 * and the resolved shear: tau[ns] (uniform) or tau[P][ns] (per_point=1). */
void orc_set_xi(orc_state *s, const double *xi, const double *xi_bc);
void orc_get_xi(const orc_state *s, double *xi);
void orc_set_tau(orc_state *s, const double *tau, int per_point);
/* Uniform-tau ramp: the s-th GL step after the call uses tau0_a + rate_a * s. */
void orc_set_tau_ramp(orc_state *s, const double *tau0, const double *rate);

/* Outer-step prologue (G.c:4450-4452 + virtualevolv G.c:1426-1447):
 * data = FFT(xi); data2 = B.data (+ data_b * sum DD_b when quirk);
 * data2 = IFFT(data2).  The next GL step reads data2. */
void orc_prologue(orc_state *s, int quirk);

/* nsteps iterations of the GL loop body (G.c:4574-4644) in fixed-step mode:
 * plasticevolv, copy xi -> data, FFT, contraction, IFFT.
 * gamma[nsteps] and gamma_a[nsteps][ns] may be NULL. */
void orc_gl_steps(orc_state *s, int nsteps, double *gamma, double *gamma_a);

/* One plasticevolv (G.c:3470-3501) on explicit arrays; used by the tests that
 * compare against the compiled reference.  d2 is the unnormalised inverse
 * transform [ns][P][2]; returns gamma, fills gamma_a[ns]. */
double orc_plasticevolv(int ns, long P, double *xi, const double *xi_bc, const double *d2,
                        const double *cd2, const double *asf2, const double *dslip,
                        const double *tau, int tau_per_point, double uq2, double pi_const,
                        double *gamma_a);

/* strain + stress + averages (strain G.c:752-835, stress G.c:837-972,
 * calc_strain_stress G.c:4058-4070) for the current xi.
 * avepst_uniform[9] = S:sigma_applied*mu (the part of avestrain G.c:714-729
 * that does not depend on xi); the xi-dependent avepsd term is added here.
 * Outputs aveps[9], avesigma[9] (sigma, not divided by mu); eps_field and
 * sig_field ([9][P] real parts) may be NULL. */
void orc_strain_stress(orc_state *s, const double *avepst_uniform,
                       double *aveps, double *avesigma, double *eps_field, double *sig_field);

/* ---- energy diagnostics on the interface line ---------------------------
 * Points with k1 == pp[0] and |n.(k - pp)| <= 1 (avestrain G.c:707/714,
 * Energy_calculation G.c:3440), in k1, k2, k3 order as point indices
 * (k1 N2 + k2) N3 + k3; at most cap are written; returns the count. */
int orc_line_points(int n1, int n2, int n3, const double n[3], const int pp[3], long *pts, int cap);

/* avestrain's energies (G.c:641-749) for a uniform applied stress sigma[9]
 * (units of mu) and zero virtual strains: returns energy_stressTimePf =
 * -sum_{line} sum_ij strain_p,ij sigma_ij, strain_p = sum_a Re xi_a e_a;
 * *e4_inc = -sum_{line} sum_ijkl S_ijkl sigma_kl sigma_ij mu / 2, what each
 * call adds to energy_in4. */
double orc_line_energy(const orc_state *s, const double sigma[9], double S11, double S12, double S44,
                       double mu, const long *pts, int npts, double *e4_inc);

/* ResidualEnergy (G.c:3920-3947): sum over k2 of
 * |-D11 x1^2 + (D10 + D01) x1 x0 - D00 x0^2| at (pp[0], k2, pp[2]), x_m the
 * sum of Re xi_a over the systems of material m (a / ns1 == m). */
double orc_residual_energy(const orc_state *s, const int pp[3], int ns1, double D00, double D01, double D10,
                           double D11);

#ifdef __cplusplus
}
#endif
#endif

/*
 * pfdd_gpu.h -- C ABI of libpfdd_gpu.so: the PFDD FFT-elasticity +
 * Ginzburg-Landau step on one MI355X (gfx950), hand-written HIP kernels.
 *
 * The reference (PFDD_Final_GPU.c, "G.c") has no library boundary; its two
 * seams are the FFT dispatch (fft_forward/fft_backward, G.c:209-273) and the
 * GL loop body inside main (G.c:4574-4691).  Each entry point below names the
 * reference code it replaces.
 *
 * Conventions
 *  - Return codes: 0 on success, a negative PFDD_E_* code otherwise; nothing
 *    asserts or exits inside the library.  pfdd_strerror() names the code.
 *  - Host arrays use the reference layouts: complex fields are interleaved
 *    doubles [field][N1][N2][N3][2] (G.c:38-43), N3 fastest.
 *  - One context owns its device memory and one HIP stream.  Calls on a
 *    context are asynchronous on that stream unless they return host data.
 *    Contexts share no state; use one context per host thread.
 *  - Only the NS slip-system fields evolve.  The reference's NV=9 virtual
 *    strain fields are identically zero on the GL path (xi_v = 0, xi_bc = 1);
 *    pfdd_upload_xi rejects non-zero virtual fields with PFDD_E_UNSUPPORTED.
 */
#ifndef PFDD_GPU_H
#define PFDD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PFDD_MAX_NS 16

enum {
  PFDD_OK = 0,
  PFDD_E_HIP = -1,          /* a HIP runtime call failed */
  PFDD_E_ALLOC = -2,        /* device or host allocation failed */
  PFDD_E_SHAPE = -3,        /* bad grid / NS / batch / pointer argument */
  PFDD_E_RADIX = -4,        /* a grid edge is not a supported FFT length */
  PFDD_E_NAN = -5,          /* non-finite value in xi or in a parameter */
  PFDD_E_UNSUPPORTED = -6,  /* input outside what the GL path supports */
  PFDD_E_STATE = -7,        /* call order (e.g. steps before upload) */
  PFDD_E_LIMIT = -8         /* input larger than the fixed staging array of pfdd_fft_host_noalloc */
};

/* Precision of FFT + contraction.  PFDD_MIXED: FP32 FFTs and contraction,
 * FP64 update on a compensated FP32 master xi (12 B a point, Re xi to ~48
 * bits).  PFDD_MIXED_LEAN: the same with xi kept as a plain FP32 complex
 * (8 B a point), rounded to FP32 after every update, and the update's sine
 * taken in FP32 after an FP64 reduction of its argument mod 2 pi.
 * PFDD_REF32: the reference's own float arithmetic, for checks against the
 * compiled reference: float data/data2 arrays (G.c:61-63), the FOURN
 * transform bit for bit (G.c:2756-2827; the axes N3, N2, N1 in both
 * directions), data2 accumulated as float sums of double products
 * (G.c:4609-4629, virtualevolv's terms G.c:1426-1447), the FP64 master xi
 * (G.c:4146) and the float uq2*data2/nsize of plasticevolv (G.c:3486-3489).
 * Power-of-two edges only (PFDD_E_RADIX otherwise), the unfused
 * field-major kernels only (PFDD_E_SHAPE for other plan flags), B(k) formed
 * in FP64 at each k-point in either bmode; no ensembles and no autotuner
 * (PFDD_E_UNSUPPORTED). */
enum { PFDD_FP64 = 0, PFDD_MIXED = 1, PFDD_MIXED_LEAN = 2, PFDD_REF32 = 3 };
enum { PFDD_B_STORED = 0, PFDD_B_RECOMPUTE = 1 }; /* interaction matrix B(k) */
enum { PFDD_FORWARD = -1, PFDD_BACKWARD = 1 };    /* FFT sign (G.c:106, FFTW_FORWARD) */

/* Runtime replacement for the compile-time constants of G.c:81-95 and the
 * constants set in main (G.c:4186-4300). */
typedef struct {
  int n1, n2, n3;              /* grid; each edge 2^a 3^b, 1 < n <= 1024 */
  int ns;                      /* active slip systems, 1..PFDD_MAX_NS */
  int precision;               /* PFDD_FP64 | PFDD_MIXED | PFDD_MIXED_LEAN | PFDD_REF32 */
  int bmode;                   /* PFDD_B_STORED | PFDD_B_RECOMPUTE */
  double C11, C12, C44;        /* elastic constants; mu = C44 (G.c:337) */
  double d1, d2, d3;           /* grid spacing of frec (G.c:4266-4268) */
  double cd2[PFDD_MAX_NS];     /* CD2: dt * kinetic coefficient (G.c:4290) */
  double asf2[PFDD_MAX_NS];    /* Asf2: gamma-surface amplitude (G.c:4278-4285) */
  double dslip[PFDD_MAX_NS];   /* dslip2 (setMat G.c:2480) */
  double uq2;                  /* 1.0 (G.c:4195) */
  double pi_const;             /* 3.141592654 (G.c:95) */
  double eps[PFDD_MAX_NS][3][3]; /* eigenstrains e_a (seteps G.c:1912-1980) */
  int plan;                    /* kernel plan: PFDD_PLAN_AUTO (0) or PFDD_PLAN_* flags and tiles */
} pfdd_params;

/* Kernel plan word (no reference counterpart; results are the same to
 * rounding for every plan).  AUTO picks the plan measured fastest at
 * 128^3 x 12; pfdd_autotune measures the candidates on the actual problem.
 * Bits 0-4 are flags; bits 8-11 and 12-15 optionally fix the tile of the
 * fused row / column kernel as log2(tile) + 1 (0 = the largest that fits);
 * a tile that does not fit is rejected with PFDD_E_SHAPE; bits 16-18 pick
 * the radix sequence of the standalone passes per axis. */
enum {
  PFDD_PLAN_AUTO = 0,
  PFDD_PLAN_UNFUSED_ROWS = 1,  /* plain N3 passes + pointwise update instead of the fused row kernel */
  PFDD_PLAN_UNFUSED_COLS = 2,  /* plain N1 passes + pointwise contraction instead of the fused column kernel */
  PFDD_PLAN_FUSED_COLS = 4,    /* fused column kernel whenever its tile fits in LDS */
  PFDD_PLAN_NO_PLANE = 8,      /* no LDS-resident N2 x N3 plane kernel (AUTO uses it for N2 = N3 <= 96) */
  PFDD_PLAN_PLANE = 16,        /* the plane kernel also at 128^2 (mixed precision only) */
  PFDD_PLAN_ROW_TILE_SHIFT = 8,  /* rows along N3 per block of the fused row kernel */
  PFDD_PLAN_COL_TILE_SHIFT = 12, /* (i2,i3) columns per block of the fused column kernel */
  /* the standalone FFT passes along an axis take radix-4 butterflies instead
   * of radix-8 first (more passes, twice the threads per pencil) */
  PFDD_PLAN_RADIX4_AXIS1 = 1 << 16,
  PFDD_PLAN_RADIX4_AXIS2 = 1 << 17,
  PFDD_PLAN_RADIX4_AXIS3 = 1 << 18,
  /* layout of the FFT work array.  INTERLEAVED keeps the ns fields of a grid
   * point together ([instance][N1][N2][N3][ns]), so that the fused row and
   * column kernels read and write all fields of their rows / columns as one
   * contiguous stretch; it needs the fused row kernel and excludes the plane
   * kernel, and the row / column tile bits then count rows / columns of all
   * fields.  FIELD_MAJOR ([instance][ns][N1][N2][N3]) forces the other. */
  PFDD_PLAN_INTERLEAVED = 1 << 19,
  PFDD_PLAN_FIELD_MAJOR = 1 << 20
};
#define PFDD_PLAN_ROW_TILE_LOG2(l) (((l) + 1) << PFDD_PLAN_ROW_TILE_SHIFT)
#define PFDD_PLAN_COL_TILE_LOG2(l) (((l) + 1) << PFDD_PLAN_COL_TILE_SHIFT)

typedef struct pfdd_ctx pfdd_ctx;

/* Allocate device state and build what B(k) needs (replaces the setup calls
 * frec + Bmatrix of main, G.c:4322-4325; B is built on the device in FP64 and
 * stored in the working precision when bmode == PFDD_B_STORED). */
int pfdd_create(const pfdd_params *p, pfdd_ctx **out);
void pfdd_destroy(pfdd_ctx *c);

/* Resolved shear tau_a (resolSS G.c:1607-1622): tau[ns] when per_point == 0,
 * or the reference array tau[N1][N2][N3][NS] when per_point == 1. */
int pfdd_set_tau(pfdd_ctx *c, const double *tau, int per_point);

/* Linear ramp of the uniform resolved shear: the s-th GL step after this
 * call (s = 0, 1, ...) uses tau_a(s) = tau0_a + rate_a * s, as if the
 * applied stress of setstress/resolSS (G.c:3820, 1607) were raised between
 * steps.  pfdd_set_tau ends the ramp. */
int pfdd_set_tau_ramp(pfdd_ctx *c, const double *tau0, const double *rate);

/* Copy the order parameter in.  xi and xi_bc are the reference arrays
 * [nsv][N1][N2][N3][2] with nsv >= ns (G.c:4146-4147); fields ns..nsv-1 are
 * the virtual strains and must be zero.  Only the real slot of xi_bc is read
 * (G.c:3480); xi_bc may be NULL (everything evolves). */
int pfdd_upload_xi(pfdd_ctx *c, const double *xi, const double *xi_bc, int nsv);

/* Copy the order parameter out into xi[nsv][N1][N2][N3][2]; virtual fields
 * (ns..nsv-1) are written as zero.  Synchronous. */
int pfdd_download_xi(pfdd_ctx *c, double *xi, int nsv);

/* Outer-step prologue: data = FFT(xi); data2 = B.data, plus, when quirk != 0,
 * the term data_b * sum_v DD_b,v that virtualevolv adds before its inverse
 * transform (G.c:4450-4452, G.c:1426-1447); data2 = IFFT(data2).  The first
 * GL step after it reads that data2. */
int pfdd_begin_outer(pfdd_ctx *c, int quirk);

/* nsteps iterations of the GL loop body in fixed-step mode (G.c:4574-4644
 * with interfacevolv off): plasticevolv, FFT(xi), contraction with B(k),
 * inverse FFT.  gamma[nsteps] and gamma_a[nsteps][ns] (may be NULL) receive
 * the sums of plasticevolv (G.c:3492-3493).  Synchronous when gamma or
 * gamma_a is non-NULL, asynchronous otherwise. */
int pfdd_gl_steps(pfdd_ctx *c, int nsteps, double *gamma, double *gamma_a);

/* Interface-line diagnostics.  avestrain (G.c:697-718) sums its
 * stress x plastic-strain energy (energy_in3) and applied-stress energy
 * (energy_in4) over the grid points with k1 == ppoint[0] and
 * |interface_n . (k - ppoint)| <= 1; with a uniform applied stress and no
 * virtual strain, energy_in3 = -sum_a (e_a : sigma) sum_line Re xi_a, so the
 * device only needs the line sums of Re xi_a.
 *
 * pfdd_set_line_points sets the point list of the context: n flat indices
 * (k1 N2 + k2) N3 + k3 in [0, N1 N2 N3) (pfdd_line_points builds the
 * reference's list); n = 0 clears it.  While a list is set every GL step also
 * writes the line sums of that step.  pfdd_line_sums returns sums[ns] for the
 * current xi (synchronous); pfdd_gl_steps_lines is pfdd_gl_steps that also
 * returns line_sums[nsteps][ns], the sums after each step (synchronous,
 * PFDD_E_STATE without a list). */
int pfdd_set_line_points(pfdd_ctx *c, const long long *pts, int n);
int pfdd_line_sums(pfdd_ctx *c, double *sums);
int pfdd_gl_steps_lines(pfdd_ctx *c, int nsteps, double *gamma, double *gamma_a, double *line_sums);

/* Strain and stress of the current xi (strain G.c:752-835, stress
 * G.c:837-972, averages G.c:4058-4070).  avepst_uniform[3][3] is the part
 * of avepst that does not depend on xi, S:sigma_applied*mu (avestrain
 * G.c:714-729); the xi-dependent part sum_a e_a <xi_a> is added on the
 * device.  Fills aveps[3][3] and avesigma[3][3] (not divided by mu);
 * eps_field / sig_field ([3][3][N1][N2][N3] doubles, real parts) may be
 * NULL.  Synchronous.  It reuses the FFT work array, so (as in main, where
 * calc_strain_stress precedes the GL loop, G.c:4461) pfdd_begin_outer must
 * run again before the next pfdd_gl_steps (PFDD_E_STATE otherwise). */
int pfdd_strain_stress(pfdd_ctx *c, const double *avepst_uniform, double *aveps,
                       double *avesigma, double *eps_field, double *sig_field);

/* Raw FFT seam (fft_forward / fft_backward, G.c:209-273): batched, in place,
 * unnormalised 3D C2C on a device buffer of batch contiguous N1*N2*N3 complex
 * blocks.  dev_buf is double2 when prec == PFDD_FP64, float2 when
 * prec == PFDD_MIXED (PFDD_E_SHAPE for other values), or float2 transformed
 * as fourn does it, bit for bit, when prec == PFDD_REF32 on a PFDD_REF32
 * context (PFDD_E_STATE on others).
 * dir = PFDD_FORWARD (-1) or PFDD_BACKWARD (+1).  Asynchronous. */
int pfdd_fft(pfdd_ctx *c, void *dev_buf, int batch, int dir, int prec);

/* Kernel plan (no reference counterpart): apply a plan word, or read the
 * applied one.  Changing the plan needs a new pfdd_begin_outer. */
int pfdd_set_plan(pfdd_ctx *c, int plan);
int pfdd_get_plan(const pfdd_ctx *c);

/* Autotuner (no reference counterpart): times every candidate plan (fusion
 * choices x tiles of the fused kernels, then radix sequences of the
 * standalone passes of the best one) over `steps` GL steps of the loaded
 * problem and applies the fastest; *plan and *ms_per_step (may be NULL)
 * receive it and its time.  Needs xi uploaded.  xi, tau and the ramp clock
 * are restored afterwards; pfdd_begin_outer must run again.  Synchronous. */
int pfdd_autotune(pfdd_ctx *c, int steps, int *plan, float *ms_per_step);

/* Wait for the context's stream. */
int pfdd_sync(pfdd_ctx *c);

/* The context's hipStream_t (as void*), for timing with HIP events. */
void *pfdd_stream(pfdd_ctx *c);

/* Device bytes held by the context. */
long long pfdd_device_bytes(const pfdd_ctx *c);

/* Profiling aid, no reference counterpart: enqueue `reps` launches of one
 * kernel stage of the GL step on the context stream, so that HIP events or
 * rocprofv3 can time it in isolation.  It overwrites the xi / work-array
 * state (upload xi again afterwards).  Asynchronous. */
enum {
  PFDD_STAGE_UPDATE = 0,    /* plasticevolv + copy + per-block gamma sums (fused: + N3 passes) */
  PFDD_STAGE_FFT_AXIS3 = 1, /* one pass along N3 (contiguous) over all fields */
  PFDD_STAGE_FFT_AXIS2 = 2, /* one pass along N2 */
  PFDD_STAGE_FFT_AXIS1 = 3, /* one pass along N1 */
  PFDD_STAGE_CONTRACT = 4,  /* B(k) contraction (fused: + N1 passes) */
  PFDD_STAGE_STEP = 5,      /* one whole GL step */
  PFDD_STAGE_COPY = 6,      /* plain copy xi -> work array (bandwidth reference) */
  PFDD_STAGE_FINISH = 7     /* the one-block pass that turns block sums into gamma */
};
int pfdd_enqueue_stage(pfdd_ctx *c, int stage, int reps);

/* Profiling aid, no reference counterpart: runs nsteps eager GL steps (as
 * pfdd_gl_steps, advancing the state) with a HIP event recorded after every
 * kernel launch, and returns per stage (index PFDD_STAGE_*) the time per
 * step (ms[8], the sum over that stage's launches in one step) and the
 * launches per step (launches[8]).  Each launch is timed inside the real
 * sequence of a step, between the events of its neighbours.  Needs
 * pfdd_begin_outer.  Synchronous. */
int pfdd_time_step_kernels(pfdd_ctx *c, int nsteps, float *ms, int *launches);

/* Launches of each stage per GL step for this context (launches[8]):
 * PFDD_STAGE_UPDATE and PFDD_STAGE_CONTRACT name the fused row and column
 * kernels when the grid allows them, and the passes they absorb count 0. */
int pfdd_step_plan(const pfdd_ctx *c, int *launches);

const char *pfdd_strerror(int rc);

/* ---- batched ensembles ----------------------------------------------------
 * Many independent PFDD instances (the reference runs one per process).
 * params[count] holds one parameter set per instance; instances whose
 * parameters agree in everything (grid, material, slip systems, precision,
 * B mode, plan) form a class that runs as one batched launch per kernel, and
 * all classes share one stream and one captured graph per GRAPH_STEPS steps.
 * Instance indices are 0..count-1 in params order.  The per-instance calls
 * have the meaning of their single-context counterparts above. */
typedef struct pfdd_ens pfdd_ens;
int pfdd_create_ensemble(const pfdd_params *params, int count, pfdd_ens **out);
void pfdd_destroy_ensemble(pfdd_ens *e);
/* Number of classes, the class of each instance (class_of[count]) and each
 * class's plan (plans[nclasses]); the arrays may be NULL. */
int pfdd_ensemble_info(const pfdd_ens *e, int *nclasses, int *class_of, int *plans);
/* tau[ns] (uniform) or tau[N1][N2][N3][ns] (per_point) of one instance. */
int pfdd_ensemble_set_tau(pfdd_ens *e, int inst, const double *tau, int per_point);
int pfdd_ensemble_upload_xi(pfdd_ens *e, int inst, const double *xi, const double *xi_bc, int nsv);
int pfdd_ensemble_download_xi(pfdd_ens *e, int inst, double *xi, int nsv);
/* Prologue of every instance (see pfdd_begin_outer); needs every xi uploaded. */
int pfdd_ensemble_begin_outer(pfdd_ens *e, int quirk);
/* nsteps GL steps of every instance; gamma[nsteps][count] (may be NULL)
 * receives each instance's gamma.  Synchronous when gamma != NULL. */
int pfdd_ensemble_gl_steps(pfdd_ens *e, int nsteps, double *gamma);
/* pfdd_autotune for each class, then the choice of mode (below);
 * plans[nclasses] and ms_per_step[nclasses] may be NULL.  xi is restored;
 * pfdd_ensemble_begin_outer must run again afterwards. */
int pfdd_ensemble_autotune(pfdd_ens *e, int steps, int *plans, float *ms_per_step);
int pfdd_ensemble_set_plan(pfdd_ens *e, int cls, int plan);
/* How a GL step of the ensemble runs (no reference counterpart):
 * PFDD_ENS_PHASED: every class's kernels; PFDD_ENS_PERSISTENT: one
 * persistent launch per step and precision whose blocks draw work items from
 * a counter: the plane items of every class, then each class's gamma finish
 * and fused column items, each of which waits until its class's plane items
 * have completed (a per-class completion counter, no grid-wide barrier).  In
 * both modes GRAPH_STEPS steps of the whole ensemble are one captured
 * hipGraph.  Persistent mode needs every class cubic with N in
 * {32,48,64,96}, ns <= 4, recomputed B, no frozen points and uniform shear
 * (PFDD_E_UNSUPPORTED otherwise); it sets each class's plan to
 * PFDD_PLAN_PLANE | PFDD_PLAN_FUSED_COLS.  Both modes give the same bits.
 * A completion wait is bounded (about 2 s); one that runs out skips its item
 * and makes the next synchronous call (pfdd_ensemble_gl_steps with gamma,
 * pfdd_ensemble_sync) return PFDD_E_HIP.
 * pfdd_ensemble_begin_outer must run again after a change.
 * pfdd_ensemble_autotune tunes the class plans, then times both modes and
 * keeps the faster; pfdd_ensemble_get_mode returns the mode and the two
 * times per step of that comparison (0 when not measured). */
enum { PFDD_ENS_PHASED = 0, PFDD_ENS_PERSISTENT = 1 };
int pfdd_ensemble_set_mode(pfdd_ens *e, int mode);
int pfdd_ensemble_get_mode(const pfdd_ens *e, float *ms_phased, float *ms_persistent);
int pfdd_ensemble_sync(pfdd_ens *e);
void *pfdd_ensemble_stream(pfdd_ens *e);
long long pfdd_ensemble_device_bytes(const pfdd_ens *e);

/* ---- host-side setup helpers (pure host code) ---------------------------
 * These restate the reference's setup routines with runtime sizes so that a
 * driver can build inputs without its compile-time constants. */

/* The 12 FCC slip systems of setfcc (G.c:1812-1909): n = {111}/sqrt3 with a
 * float sqrt3, b = <110>/2. */
void pfdd_setfcc(double xn[12][3], double xb[12][3]);

/* Orientation rows as setorient normalises them (G.c:2036-2055: each row
 * divided by its squared length). */
void pfdd_setorient(const double rows[3][3], double theta[3][3]);

/* e_a = theta (b_a x n_a / dslip_a) theta^T, unsymmetrised (seteps G.c:1950-1978). */
void pfdd_seteps(int ns, const double (*xn)[3], const double (*xb)[3], const double *dslip,
                 const double theta[3][3], double (*eps)[3][3]);

/* tau_a = sum_ij sigma_ij e_a,ij for a uniform sigma (resolSS G.c:1607-1622). */
void pfdd_resolve_tau(int ns, const double sigma[3][3], const double (*eps)[3][3], double *tau);

/* SplitMix64 stream (Steele, Lea & Flood 2014) and a uniform double in [0,1). */
unsigned long long pfdd_splitmix64(unsigned long long *state);
double pfdd_uniform01(unsigned long long *state);

/* Voxelised circular loop: sets xi[sys] = 1 (real part) at the grid points x
 * with |n.(x-c)| < h/2 and |(x-c) - (n.(x-c)) n| <= r, where x-c is the
 * periodic minimum image, n = normal/|normal| and h = max_i |n_i| is the
 * spacing of lattice planes with that normal.  xi is [>=ns][N1][N2][N3][2].
 * Returns the number of points set. */
long pfdd_add_loop(double *xi, int n1, int n2, int n3, int sys, const double normal[3],
                   const double center[3], double radius);

/* Seeded loops of the benchmark workloads: for each of nloops loops, draw
 * (in this order from the SplitMix64 stream) the radius r ~ U[rmin, rmax),
 * the system a = floor(U ns), and the centre c_i ~ U[0, N_i); then
 * pfdd_add_loop(xi, ..., a, normals[a], c, r).  loops_out (NULL or
 * [nloops][6]) receives a, c, r and the number of points set. */
int pfdd_seed_loops(double *xi, int n1, int n2, int n3, int ns, const double (*normals)[3], int nloops,
                    double rmin, double rmax, unsigned long long *rng_state, double *loops_out);

/* The interface line of avestrain (G.c:707, 714): the points with
 * k1 == ppoint[0] and |normal . (k - ppoint)| <= 1, in k2, k3 order, as flat
 * indices (k1 N2 + k2) N3 + k3.  Writes at most cap of them to pts (may be
 * NULL) and returns the count. */
int pfdd_line_points(int n1, int n2, int n3, const double normal[3], const int ppoint[3], long long *pts, int cap);

/* Symmetric applied stress with the six independent components (11, 12, 13,
 * 22, 23, 33) drawn in that order from U[-amp, amp). */
void pfdd_seed_sigma(double amp, unsigned long long *rng_state, double sigma[3][3]);

/* plasticconverge (G.c:3732-3785), the convergence test of the GL loop:
 * records step it_plastic in testplastic[] (zeroed by the caller before the
 * loop), updates the run counter *countgamma, prints the reference's log line
 * ("(gamma-gammalast)/gamma = ..." or "gamma-gammalast = ...") on stdout
 * after that update, and returns 0 when the loop has converged, 1 otherwise. */
int pfdd_plastic_converge(double gamma, double gammalast, int it_plastic, int *testplastic, int *countgamma);

/* ---- the reference's FFT seam on host arrays ----------------------------
 * fft_forward / fft_backward have the exact signature and semantics of
 * G.c:209-273 (in place, unnormalised, batch contiguous N1*N2*N3 complex
 * float blocks, forward sign -1) and run on the GPU: each call copies the
 * batch to the device, transforms it in FP32 and copies it back.  When every
 * edge is a power of two the transform is fourn's (G.c:2756-2827) bit for
 * bit, the result of the reference's -DUSE_FOURN build; edges with a factor
 * 3 (which fourn does not take) get the FP32 Stockham passes.
 * pfdd_fft_host_init fixes the grid (the reference's N1..N3); it must be
 * called first and again to change the grid.  Returns 0 or a PFDD_E_* code;
 * the seam functions themselves return nothing, as in the reference, and
 * pfdd_fft_host_status() reports the last error. */
int pfdd_fft_host_init(int n1, int n2, int n3);
void fft_forward(float *data, int batch);
void fft_backward(float *data, int batch);
int pfdd_fft_host_status(void);
void pfdd_fft_host_release(void);

/* Allocation-free form of the host seam.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  fft_forward /
 * fft_backward stage the batch in a device buffer that they allocate and
 * grow; this call stages it in a fixed device array of
 * PFDD_NOALLOC_MAX_POINTS complex floats (16 MiB), which each GPU holds from
 * its first pfdd_fft_host_init until the process ends, and it calls no
 * allocating routine.  dir = PFDD_FORWARD or PFDD_BACKWARD; the result is
 * the same as that of fft_forward / fft_backward.  Returns 0, PFDD_E_LIMIT
 * when batch * N1*N2*N3 > PFDD_NOALLOC_MAX_POINTS, PFDD_E_STATE before
 * pfdd_fft_host_init, or another PFDD_E_* code.  Synchronous.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU race. */
#define PFDD_NOALLOC_MAX_POINTS (1 << 21)
int pfdd_fft_host_noalloc(float *data, int batch, int dir);

#ifdef __cplusplus
}
#endif
#endif

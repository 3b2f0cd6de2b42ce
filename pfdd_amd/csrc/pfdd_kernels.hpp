// pfdd_kernels.hpp -- HIP kernels of the GL step (gfx950).
//
// Data layout in HBM (one context of B instances of the same grid, S slip
// systems each; field f = b * S + a; B = 1 for a single run):
//   xi    master order parameter [B*S][N1][N2][N3] (complex, as G.c:64):
//         FP64: double2; mixed: Re as float2 (hi, lo) followed by Im as float
//         (struct Master)
//   W     cx<T>   [B*S][N1][N2][N3]   FFT work array: data / data2 of the reference
//   Bsym  T       [S(S+1)/2][P]       upper triangle of B(k), stored mode only (shared by the instances)
//   mask  u8      [B*S][P]            1 = frozen (xi_bc != 0), only when some point is frozen
//   taupp double  [B][P][S]           per-point resolved shear, only when not uniform
//   tau   double2 [B*S]               uniform resolved shear and its ramp rate per field
// T = double (PFDD_FP64) or float (PFDD_MIXED).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "pfdd_fft.hpp"

namespace pfdd {

constexpr int MAXS = 16;

// The master order parameter.  FP64 contexts keep xi as double2.  Mixed
// contexts keep a compensated form, 12 bytes per point instead of 16: Re xi
// as the unevaluated sum hi + lo of two floats (48 significant bits; the
// update itself runs in FP64 on hi + lo and splits the result again), and
// Im xi as a float.  Im xi is the small component that the non-Hermitian
// B(k) of the Nyquist planes feeds (|Im xi| ~ 1e-3); the FP32 transforms
// around it carry errors of the same relative size.
template <typename T> struct Master;
// The master is read once and written once per GL step and is larger than
// the caches: its accesses are non-temporal (measured at 128^3 x 12: FP64
// step 0.989 -> 0.968 ms, mixed 0.556 -> 0.537 ms; the other passes of the
// step keep more of L2.  Non-temporal stores with ordinary loads: 0.978 ms).
template <typename V> __device__ __forceinline__ V xld(const V* p) { return __builtin_nontemporal_load(p); }
template <typename V> __device__ __forceinline__ void xst(V* p, V v) { __builtin_nontemporal_store(v, p); }
// raw(i) loads a point in its stored form and widen() turns that into the
// FP64 (Re, Im) the update works on: a kernel that loads its points long
// before the update holds the narrower raw form in registers meanwhile.
template <> struct Master<double> {
  double2* p;
  using Raw = double2;
  __device__ __forceinline__ Raw raw(long i) const {
    const VecOf<double>::t v = xld(reinterpret_cast<const VecOf<double>::t*>(p + i));
    return make_double2(v.x, v.y);
  }
  static __device__ __forceinline__ double2 widen(Raw v) { return v; }
  __device__ __forceinline__ double2 load(long i) const { return raw(i); }
  __device__ __forceinline__ void store(long i, double2 v) const {
    VecOf<double>::t w;
    w.x = v.x;
    w.y = v.y;
    xst(reinterpret_cast<VecOf<double>::t*>(p + i), w);
  }
  __device__ __forceinline__ double re(long i) const { return p[i].x; }
};
// The two FP32 forms as separate types, for the per-step kernels, which are
// instantiated per form (a per-access branch on the form measured 0.155 ->
// 0.183 ms for the lean FP32 row kernel at 128^3 x 12).
struct MasterF32C {  // compensated: (hi, lo) of Re xi, and Im xi
  float2* r;
  float* im;
  struct Raw {
    VecOf<float>::t h;
    float im;
  };
  __device__ __forceinline__ Raw raw(long i) const {
    return Raw{xld(reinterpret_cast<const VecOf<float>::t*>(r + i)), xld(im + i)};
  }
  static __device__ __forceinline__ double2 widen(Raw v) {
    return make_double2((double)v.h.x + (double)v.h.y, (double)v.im);
  }
  __device__ __forceinline__ double2 load(long i) const { return widen(raw(i)); }
  __device__ __forceinline__ void store(long i, double2 v) const {
    const float hi = (float)v.x;
    VecOf<float>::t w;
    w.x = hi;
    w.y = (float)(v.x - (double)hi);
    xst(reinterpret_cast<VecOf<float>::t*>(r + i), w);
    xst(im + i, (float)v.y);
  }
  __device__ __forceinline__ double re(long i) const {
    const float2 h = r[i];
    return (double)h.x + (double)h.y;
  }
};
struct MasterF32L {  // lean: (Re xi, Im xi) as a plain FP32 complex
  float2* r;
  using Raw = VecOf<float>::t;
  __device__ __forceinline__ Raw raw(long i) const { return xld(reinterpret_cast<const VecOf<float>::t*>(r + i)); }
  static __device__ __forceinline__ double2 widen(Raw h) { return make_double2((double)h.x, (double)h.y); }
  __device__ __forceinline__ double2 load(long i) const { return widen(raw(i)); }
  __device__ __forceinline__ void store(long i, double2 v) const {
    VecOf<float>::t w;
    w.x = (float)v.x;
    w.y = (float)v.y;
    xst(reinterpret_cast<VecOf<float>::t*>(r + i), w);
  }
  __device__ __forceinline__ double re(long i) const { return (double)r[i].x; }
};
// Either form, chosen per access (the kernels that run once per outer
// iteration or only in non-default plans).
// A third form serves PFDD_REF32 (FP32 work array, FP64 master as in the
// reference, G.c:4146): d points to the double2 master and r, im are null.
template <> struct Master<float> {
  float2* r;  // (hi, lo) of Re xi; (Re xi, Im xi) in the lean form
  float* im;  // Im xi; null in the lean form
  double2* d = nullptr;  // PFDD_REF32: the FP64 master
  using Raw = double2;
  __device__ __forceinline__ double2 load(long i) const {
    if (d) return Master<double>{d}.load(i);
    return im ? MasterF32C{r, im}.load(i) : MasterF32L{r}.load(i);
  }
  __device__ __forceinline__ Raw raw(long i) const { return load(i); }
  static __device__ __forceinline__ double2 widen(Raw v) { return v; }
  __device__ __forceinline__ void store(long i, double2 v) const {
    if (d) Master<double>{d}.store(i, v);
    else if (im) MasterF32C{r, im}.store(i, v);
    else MasterF32L{r}.store(i, v);
  }
  __device__ __forceinline__ double re(long i) const {
    if (d) return d[i].x;
    return im ? MasterF32C{r, im}.re(i) : MasterF32L{r}.re(i);
  }
};
// Bytes per point of the master, and the master over a base allocation of n
// points.  The lean form (PFDD_MIXED_LEAN) keeps xi as a plain FP32 complex,
// 8 bytes a point: xi is rounded to FP32 after every update (the update
// itself still runs in FP64).
template <typename T> constexpr size_t master_bytes(bool lean) { return sizeof(T) == 8 ? 16 : lean ? 8 : 12; }
template <typename T> __host__ __device__ inline Master<T> make_master(void* base, long n, bool lean);
template <> __host__ __device__ inline Master<double> make_master<double>(void* base, long, bool) {
  return Master<double>{reinterpret_cast<double2*>(base)};
}
template <> __host__ __device__ inline Master<float> make_master<float>(void* base, long n, bool lean) {
  float2* r = reinterpret_cast<float2*>(base);
  return Master<float>{r, lean ? nullptr : reinterpret_cast<float*>(r + n)};
}

// Constants of the elastic kernel, shared by every k-point (read through the
// scalar cache: every lane of a wave reads the same address).
struct ElasticConsts {
  double mu, ll, mup, inv_mu;      // C44, C12, C11-C12-2C44 (Bmatrix G.c:345-347)
  double sig[MAXS][9];             // s_a = C:e_a
  double SE[MAXS][MAXS];           // s_a:e_b
  double Tsum[MAXS];               // sum_ij s_a,ij
  double T[9];                     // T_pq = sum_ij C_pqij
  double C[81];                    // C_ijkl
  double eps[MAXS][9];             // e_a
  // the stress-spectrum form of the contraction (contract_point): s_b = C:e_b
  // and w_t sym(e_a)_t / mu over the six symmetric components
  // t = (00, 01, 02, 11, 12, 22), w_t = 1 on the diagonal and 2 off it
  double sig6[MAXS][6];
  double e6m[MAXS][6];
  float sig6f[MAXS][6];
  float e6mf[MAXS][6];
  double d[3];                     // grid spacing
  const double* freq[3];           // per-axis frec tables (device), freq[a][i] = frec1(i, n[a], d[a])
  double g_iso;                    // (mu+ll) / (mu^2 kk) for mup == 0
  int iso;                         // mup == 0: isotropic Green's function
  int n[3];
  int ns;
};

// frec, G.c:2698-2741
__device__ __forceinline__ double frec1(int i, int n, double d) {
  if (i == 0) return 0.0;
  if (i < n / 2) return (double)i / (double)n / d;
  return ((double)i - (double)n) / (double)n / d;
}

// Green's function of Bmatrix (G.c:386) as a symmetric 3x3 (g00 g01 g02 g11 g12 g22).
// Returns false at k = 0.
__device__ __forceinline__ bool green6(const double f[3], const ElasticConsts* __restrict__ ec, double g[6]) {
  const double fk2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
  if (!(fk2 > 0.0)) return false;
  if (ec->iso) {  // mup = 0: kk = 1 + (mu+ll)/mu, G = (delta/mu - f f (mu+ll)/(mu^2 kk |f|^2)) / |f|^2
    const double inv = 1 / fk2, c = ec->g_iso * inv, im = ec->inv_mu;
    g[0] = inv * (im - f[0] * f[0] * c);
    g[1] = inv * (-f[0] * f[1] * c);
    g[2] = inv * (-f[0] * f[2] * c);
    g[3] = inv * (im - f[1] * f[1] * c);
    g[4] = inv * (-f[1] * f[2] * c);
    g[5] = inv * (im - f[2] * f[2] * c);
    return true;
  }
  const double mu = ec->mu, ll = ec->ll, mup = ec->mup;
  const double z0 = f[0] * f[0] / fk2, z1 = f[1] * f[1] / fk2, z2 = f[2] * f[2] / fk2;
  const double m0 = mu + mup * z0, m1 = mu + mup * z1, m2 = mu + mup * z2;
  const double kk = 1.0 + (mu + ll) * (1 / m0 * z0 + 1 / m1 * z1 + 1 / m2 * z2);
  const double inv = 1 / fk2, c = (mu + ll) / kk;
  const double m[3] = {m0, m1, m2};
  int t = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = i; k < 3; k++)
      g[t++] = inv * ((i == k ? 1.0 : 0.0) / m[k] - f[i] * f[k] / fk2 / m[i] / m[k] * c);
  return true;
}

__device__ __forceinline__ void gmul(const double g[6], const double u[3], double o[3]) {
  o[0] = g[0] * u[0] + g[1] * u[1] + g[2] * u[2];
  o[1] = g[1] * u[0] + g[3] * u[1] + g[4] * u[2];
  o[2] = g[2] * u[0] + g[4] * u[1] + g[5] * u[2];
}

// k-vector of flat point index pt (< 2^31: every edge is at most 1024)
__device__ __forceinline__ void kvec(long pt, const ElasticConsts* __restrict__ ec, double f[3]) {
  const unsigned n2 = (unsigned)ec->n[1], n3 = (unsigned)ec->n[2];
  const unsigned p = (unsigned)pt;
  const unsigned r = p / n3, k = p - r * n3;
  const unsigned i = r / n2, j = r - i * n2;
  f[0] = ec->freq[0][i];
  f[1] = ec->freq[1][j];
  f[2] = ec->freq[2][k];
}

__device__ __forceinline__ int tri_index(int a, int b, int S) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo * S - lo * (lo - 1) / 2 + (hi - lo);
}

// ---------------------------------------------------------------- reductions
// Block sum (blockDim.x a multiple of 64, at most 1024).
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; i++) s += red[i];
  return s;  // valid in thread 0
}

// ---------------------------------------------------------------- FFT kernels
constexpr int pow2_floor(int v) {
  int t = 1;
  while (t * 2 <= v) t *= 2;
  return t;
}
// Rows per block of the contiguous-axis kernels: a power of two, so that the
// block size TILE*TPP is a multiple of 64 whenever it reaches 64.
template <typename T, int N, int RL = 8> constexpr int tile_rows() {
  constexpr int tpp = make_plan(N, RL).tpp;
  int t = pow2_floor(256 / tpp > 0 ? 256 / tpp : 1);
  const int lim = 32768 / (pitch<N>() * (int)sizeof(cx<T>));
  while (t > lim && t > 1) t /= 2;
  return t;
}
// Columns per block of the strided-axis kernel: 512 bytes of each row read
// together in FP64 (32 columns) and 256 in FP32 (32 columns), within 1024
// threads and 80 KB of (dynamic) LDS, so that two blocks still fit on a CU.
// Measured at 128^3 x 12: FP64 0.956 -> 0.936 ms per step with 512-byte rows;
// FP32 0.508 ms with 256-byte rows, 0.529 with 512.
// WIDE FP64 tiles, for long lengths along an axis whose rows lie far apart
// (the N1 pass at N1 >= 384), may take all 160 KB of LDS, one block per CU:
// 16 columns at N1 = 512 instead of 8 (c4, 512^3 x 3 FP64: 18.2 -> 16.9 ms
// per step; the N2 passes of the same length were slower so).  In FP32 the
// tile is held to 16 columns there by the 1024 threads of a block; 32
// columns at two pencils per thread measured slower (4.1 vs 3.6 ms for the
// two N1 passes of c4).
constexpr int COLS_LDS_MAX = 80 * 1024, COLS_LDS_WIDE = 160 * 1024;
template <typename T, int N, int RL = 8, bool WIDE = false> constexpr int tile_cols() {
  constexpr int tpp = make_plan(N, RL).tpp;
  constexpr int row_bytes = sizeof(T) == 8 ? 512 : 256;
  int t = pow2_floor(256 / tpp > 0 ? 256 / tpp : 1);
  if (t < row_bytes / (int)sizeof(cx<T>)) t = row_bytes / (int)sizeof(cx<T>);
  const int lim = (WIDE ? COLS_LDS_WIDE : COLS_LDS_MAX) / (pitch<N>() * (int)sizeof(cx<T>));
  while (t > lim && t > 1) t /= 2;
  while (t * tpp > 1024 && t > 1) t /= 2;
  return t;
}
template <typename T, int N, int RL = 8, bool WIDE = false> constexpr int cols_threads() {
  return tile_cols<T, N, RL, WIDE>() * make_plan(N, RL).tpp;
}
template <typename T, int N, int RL = 8, bool WIDE = false> constexpr size_t cols_lds_bytes() {
  return (size_t)tile_cols<T, N, RL, WIDE>() * pitch<N>() * sizeof(cx<T>);
}
// whether a length has a wide tile that differs from the ordinary one (at
// N = 256 the wide FP64 tile measured no faster: c2, 256^3 x 12)
template <typename T, int N, int RL> constexpr bool has_wide_cols() {
  return sizeof(T) == 8 && N >= 384 && tile_cols<T, N, RL, true>() != tile_cols<T, N, RL, false>();
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Kernels whose neighbouring tiles share cache lines (the strided-axis
// passes, when a tile row is narrower than a line) give each XCD a
// contiguous range of tiles instead: block b -> tile xcd_tile(b, nb).
constexpr int NUM_XCD = 8;
__device__ __forceinline__ long xcd_tile(long b, long nb) {
  const long q = nb / NUM_XCD, r = nb % NUM_XCD;
  const long x = b % NUM_XCD, i = b / NUM_XCD;
  return x * q + (x < r ? x : r) + i;
}

// Transform along the contiguous axis: rows of N consecutive elements.  The
// first pass reads the rows from HBM, the last writes them back.
template <typename T, int N, int RL = 8>
__global__ void __launch_bounds__((tile_rows<T, N, RL>() * make_plan(N, RL).tpp))
k_fft_rows(cx<T>* __restrict__ data, long nrows, T sign, const cx<T>* __restrict__ tw) {
  constexpr int TPP = make_plan(N, RL).tpp, TILE = tile_rows<T, N, RL>(), LD = pitch<N>();
  __shared__ cx<T> sm[TILE * LD];
  __shared__ cx<T> twl[N];
  // the table is stored to LDS after the first pass's HBM reads (which use no twiddles) are issued
  TwStage<T, N, TILE * TPP> ts;
  ts.load(tw);
  const long row0 = (long)blockIdx.x * TILE;
  const int nr = (int)((nrows - row0) < TILE ? (nrows - row0) : TILE);
  const RowPencils<T> g{data + row0 * N, N};
  const LdsPencils<T> l{sm, LD};
  fft_pencils<T, N, 1, RL>(g, l, g, (int)threadIdx.x / TPP, 1 << 30, nr, (int)threadIdx.x % TPP, sign, twl,
                           [&] { ts.store(twl); });
}

// Transform along a strided axis of an [outer][N][inner] array; each block
// takes `tile` consecutive inner indices (tile divides inner, tile <= TILE).
// Threads run pencil-fastest, so a wave's loads of one row of the tile are
// contiguous.
template <typename T, int N, int RL = 8, bool WIDE = false>
__global__ void __launch_bounds__((cols_threads<T, N, RL, WIDE>()))
k_fft_cols(cx<T>* __restrict__ data, long inner, int tile, T sign, const cx<T>* __restrict__ tw) {
  constexpr int TILE = tile_cols<T, N, RL, WIDE>(), LD = pitch<N>();
  extern __shared__ unsigned char lds_raw[];  // cols_lds_bytes<T, N, RL, WIDE>()
  cx<T>* sm = reinterpret_cast<cx<T>*>(lds_raw);
  __shared__ cx<T> twl[N];
  // the table is stored to LDS after the first pass's HBM reads (which use no twiddles) are issued
  TwStage<T, N, cols_threads<T, N, RL, WIDE>()> ts;
  ts.load(tw);
  const long tiles_per_outer = inner / tile;
  const long bt = xcd_tile(blockIdx.x, gridDim.x);
  const long o = bt / tiles_per_outer;
  const long i0 = (bt % tiles_per_outer) * tile;
  const ColPencils<T> g{data + o * N * inner + i0, inner};
  const LdsPencils<T> l{sm, LD};
  fft_pencils<T, N, 1, RL>(g, l, g, (int)threadIdx.x % TILE, 1 << 30, tile, (int)threadIdx.x / TILE, sign, twl,
                           [&] { ts.store(twl); });
}

// ---------------------------------------------------------------- GL kernels
struct StepArgs {
  long P;
  int ns;
  int nblk;             // blocks per field of k_update
  double uq2, pi;
  float uq2f, nsizef;   // uq2 and nsize as the reference's float arithmetic sees them (PFDD_REF32)
  double cd2[MAXS], asf2[MAXS], dslip[MAXS];  // per slip system
  // per field f = instance * ns + system: (tau, rate); the s-th GL step
  // after the ramp was set uses tau + rate * s
  const double2* tau;
};

// Per-field coefficients of plasticevolv (G.c:3470-3501) for one block.
struct FieldCoef {
  double cd2, asf2, dsl, tau, rate;
  long tbase;  // per-point resolved shear of point pt: taupp[tbase + pt * ns] ([instance][P][ns])
};
__device__ __forceinline__ FieldCoef field_coef(const StepArgs& a, int f, const int* __restrict__ step_ctr) {
  const int s = f % a.ns;
  FieldCoef c;
  c.cd2 = a.cd2[s];
  c.asf2 = a.asf2[s];
  c.dsl = a.dslip[s];
  const double2 t = a.tau[f];
  c.rate = t.y * (double)step_ctr[1];
  c.tau = t.x + c.rate;
  c.tbase = (long)(f / a.ns) * a.P * a.ns + s;
  return c;
}
// The update of one point (G.c:3485-3489): d is the unnormalised inverse
// transform of data2.
// inv_n = 1 / P (exact for the power-of-two grids; within an ulp of the
// reference's division by nsize otherwise).  F32SIN (the lean FP32 master
// only, whose xi carries 24 bits): the sine is taken in FP32 after reducing
// its FP64 argument mod 2 pi in FP64 (Cody-Waite, two-part 2 pi), so the
// argument keeps the reference's pi constant and the error stays that of one
// FP32 sine (measured at 128^3 x 12: 171 -> 157 us for the lean row kernel).
template <bool TAUPP, bool F32SIN = false>
__device__ __forceinline__ void gl_point(double2& x, double dre, double dim, const FieldCoef& c, const StepArgs& a,
                                         const double* __restrict__ taupp, long pt, double inv_n) {
  const double t = TAUPP ? taupp[c.tbase + pt * a.ns] + c.rate : c.tau;
  const double arg = 2.0 * a.pi * x.x;
  double sn;
  if constexpr (F32SIN) {
    constexpr double TWO_PI_HI = 6.283185307179586, TWO_PI_LO = 2.4492935982947064e-16;
    const double k = rint(arg * (1.0 / TWO_PI_HI));
    const double r = fma(-k, TWO_PI_LO, fma(-k, TWO_PI_HI, arg));
    sn = (double)sinf((float)r);
  } else {
    sn = sin(arg);
  }
  x.x -= c.cd2 * (a.uq2 * dre * inv_n + c.asf2 * a.pi * sn - t / c.dsl);
  x.y -= a.uq2 * c.cd2 * (dim * inv_n);
}
// plasticevolv with the reference's own arithmetic on a float data2
// (PFDD_REF32): uq2*DATA2/nsize and DATA2_im/nsize are float operations
// (float uq2, float data2, int nsize converted to float), the rest is double,
// evaluated left to right with every product and sum rounded on its own
// (no fused multiply-adds), as the compiled reference does (G.c:3486-3489).
template <bool TAUPP>
__device__ __forceinline__ void gl_point_r32(double2& x, float dre, float dim, const FieldCoef& c, const StepArgs& a,
                                             const double* __restrict__ taupp, long pt) {
#pragma clang fp contract(off)
  const double t = TAUPP ? taupp[c.tbase + pt * a.ns] + c.rate : c.tau;
  const float qre = (a.uq2f * dre) / a.nsizef;
  const float qim = dim / a.nsizef;
  const double sn = sin(2.0 * a.pi * x.x);
  x.x = x.x - c.cd2 * (((double)qre + c.asf2 * a.pi * sn) - t / c.dsl);
  x.y = x.y - ((double)a.uq2f * c.cd2) * (double)qim;
}
template <typename M> constexpr bool lean_master() { return __is_same(M, MasterF32L); }

constexpr int UPD_THREADS = 256, UPD_EPT = 4;

// Second pass of the per-field sums, one block: each field's partials are
// added in a fixed order (in up to 16 chunks, each lane-strided and then a
// butterfly over a wave, then the chunks in order); then each instance's
// fields (ns consecutive ones) are added in order.
// Output row: [instance][ns + 1] = the ns field sums and their total.  With
// step_ctr the row is out + step_ctr[0] * row length and both counters
// advance (step_ctr[0]: history row, step_ctr[1]: GL steps under the current
// tau ramp).  nfields <= FIN_MAX_FIELDS, a multiple of ns.
constexpr int FIN_THREADS = 1024, FIN_MAX_FIELDS = 4096;
// The block's work (FIN_THREADS threads); fsum: nfields doubles of LDS.
__device__ __forceinline__ void finish_body(const double* __restrict__ partials, int nfields, int nblk, int ns,
                                            double* out, int* step_ctr, double* fsum) {
  const int lane = threadIdx.x & 63;
  const int ninst = nfields / ns;
  const long row = step_ctr ? (long)step_ctr[0] : 0;
  // a field's partials are cut into nch chunks (by nblk alone, so a field
  // is added the same way whatever the number of fields); the 16 waves take
  // (field, chunk) pairs, 16 / nch fields at a time
  __shared__ double red[FIN_THREADS / 64];
  constexpr int NW = FIN_THREADS / 64;
  const int w = threadIdx.x >> 6;
  const int nch = nblk / 2048 < 1 ? 1 : nblk / 2048 > NW ? NW : nblk / 2048;
  const int g = NW / nch;
  const int len = (nblk + nch - 1) / nch;
  for (int f0 = 0; f0 < nfields; f0 += g) {
    const int f = f0 + w / nch, ch = w % nch;
    if (w < g * nch && f < nfields) {
      const int b0 = ch * len, b1 = b0 + len < nblk ? b0 + len : nblk;
      double v = 0.0;
      for (int b = b0 + lane; b < b1; b += 64) v += partials[(long)f * nblk + b];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[w] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < g && f0 + (int)threadIdx.x < nfields) {
      double t = 0.0;
      for (int c = 0; c < nch; c++) t += red[threadIdx.x * nch + c];
      fsum[f0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
  double* dst = out + row * (long)ninst * (ns + 1);
  for (int b = threadIdx.x; b < ninst; b += FIN_THREADS) {
    double tot = 0.0;
    for (int a = 0; a < ns; a++) {
      const double v = fsum[b * ns + a];
      dst[b * (ns + 1) + a] = v;
      tot += v;
    }
    dst[b * (ns + 1) + ns] = tot;
  }
  if (step_ctr && threadIdx.x == 0) {
    step_ctr[0] = step_ctr[0] + 1;  // history row (read by every thread before the barrier above)
    step_ctr[1] = step_ctr[1] + 1;  // GL steps since the tau ramp was set
  }
}

// First stage of the per-field sums when a field has many block sums (the
// row kernels write one per row block): block (j, f) adds the
// j-th of FOLD consecutive ranges of field f's block sums in a fixed order,
// giving FOLD sums per field for k_finish_sums.
constexpr int FOLD = 16;
__global__ void __launch_bounds__(256)
k_fold_partials(const double* __restrict__ in, int nblk, double* __restrict__ out) {
  __shared__ double red[256 / 64];
  const int f = blockIdx.y, j = blockIdx.x;
  const int per = (nblk + FOLD - 1) / FOLD;
  const int b0 = j * per, b1 = b0 + per < nblk ? b0 + per : nblk;
  double acc = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += 256) acc += in[(long)f * nblk + b];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[(long)f * FOLD + j] = s;
}

__global__ void __launch_bounds__(FIN_THREADS)
k_finish_sums(const double* __restrict__ partials, int nfields, int nblk, int ns, double* out, int* step_ctr) {
  __shared__ double fsum[FIN_MAX_FIELDS];
  finish_body(partials, nfields, nblk, ns, out, step_ctr, fsum);
}

// plasticevolv (G.c:3470-3501) fused with the copy xi -> data (G.c:4581-4590)
// and the gamma sums.  grid = (nblk, B*S).
template <typename T, bool MASK, bool TAUPP>
__global__ void __launch_bounds__(UPD_THREADS)
k_update(Master<T> xi, cx<T>* __restrict__ W, const uint8_t* __restrict__ mask,
         const double* __restrict__ taupp, StepArgs a, double* __restrict__ partials,
         const int* __restrict__ step_ctr) {
  __shared__ double red[UPD_THREADS / 64];
  const int f = blockIdx.y;
  const long P = a.P;
  const double inv_n = 1.0 / (double)P;
  const long off = (long)f * P;
  const long base = (long)blockIdx.x * (UPD_THREADS * UPD_EPT) + threadIdx.x;
  const FieldCoef c = field_coef(a, f, step_ctr);
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < UPD_EPT; e++) {
    const long pt = base + (long)e * UPD_THREADS;
    if (pt < P) {
      double2 x = xi.load(off + pt);
      const cx<T> d = W[off + pt];
      if (!MASK || mask[off + pt] == 0) {
        if constexpr (sizeof(T) == 4) {
          if (xi.d) gl_point_r32<TAUPP>(x, d.x, d.y, c, a, taupp, pt);
          else if (!xi.im) gl_point<TAUPP, true>(x, (double)d.x, (double)d.y, c, a, taupp, pt, inv_n);
          else gl_point<TAUPP>(x, (double)d.x, (double)d.y, c, a, taupp, pt, inv_n);
        } else {
          gl_point<TAUPP>(x, (double)d.x, (double)d.y, c, a, taupp, pt, inv_n);
        }
        xi.store(off + pt, x);
      }
      W[off + pt] = {(T)x.x, (T)x.y};
      acc += x.x;  // (the block sum is scaled by 1/P once)
    }
  }
  const double s = block_sum(acc, red) * inv_n;
  if (threadIdx.x == 0) partials[(long)f * a.nblk + blockIdx.x] = s;
}

// Sum of Re xi per field (avestrain's xi means, G.c:670-686).  grid = (nblk, S).
template <typename T>
__global__ void __launch_bounds__(UPD_THREADS)
k_xi_sums(Master<T> xi, long P, int nblk, double* __restrict__ partials) {
  __shared__ double red[UPD_THREADS / 64];
  const int f = blockIdx.y;
  const long base = (long)blockIdx.x * (UPD_THREADS * UPD_EPT) + threadIdx.x;
  double acc = 0.0;
  for (int e = 0; e < UPD_EPT; e++) {
    const long pt = base + (long)e * UPD_THREADS;
    if (pt < P) acc += xi.re((long)f * P + pt);
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) partials[(long)f * nblk + blockIdx.x] = s;
}

// Sum of Re xi over a point list per field (the interface line of avestrain,
// G.c:697-718).  grid = F blocks.  The result goes to row step_ctr[0] - 1 of
// out[rows][F] (the step that has just finished: the step's last reduction
// advances step_ctr[0]) or, without step_ctr, to row 0.
template <typename T>
__global__ void __launch_bounds__(UPD_THREADS)
k_line_sums(Master<T> xi, long P, const long* __restrict__ pts, int npts, int F, double* __restrict__ out,
            const int* __restrict__ step_ctr) {
  __shared__ double red[UPD_THREADS / 64];
  const int f = blockIdx.x;
  double acc = 0.0;
  for (int q = threadIdx.x; q < npts; q += UPD_THREADS) acc += xi.re((long)f * P + pts[q]);
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[(long)(step_ctr ? step_ctr[0] - 1 : 0) * F + f] = s;
}

template <typename T>
__global__ void k_xi_to_w(Master<T> xi, cx<T>* __restrict__ W, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double2 x = xi.load(i);
    W[i] = {(T)x.x, (T)x.y};
  }
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Contraction at one k-point (G.c:4609-4629): out_a = sum_b B_ab(k) v_b,
// written to out[a * ostride] (v stays in registers, so out may alias the
// locations v was read from).  Stored mode reads B from Bsym; recompute mode
// never forms B and takes the stress-spectrum form (contract_stress_form).
// quirk != 0 adds v_a * sum_ij DD_a,ij (virtualevolv G.c:1426-1441) in a
// second pass over a.  UNROLL_A unrolls the stored mode's loop over outputs
// (more registers); rolled, v is still only indexed by unrolled loops, so it
// stays in registers.
__device__ __forceinline__ void sig_f(const ElasticConsts* __restrict__ ec, int a, const double f[3], double u[3]) {
  const double* s = ec->sig[a];
  u[0] = s[0] * f[0] + s[1] * f[1] + s[2] * f[2];
  u[1] = s[3] * f[0] + s[4] * f[1] + s[5] * f[2];
  u[2] = s[6] * f[0] + s[7] * f[1] + s[8] * f[2];
}

template <typename T> __device__ __forceinline__ const T* sig6_of(const ElasticConsts* __restrict__ ec, int b);
template <> __device__ __forceinline__ const double* sig6_of<double>(const ElasticConsts* __restrict__ ec, int b) {
  return ec->sig6[b];
}
template <> __device__ __forceinline__ const float* sig6_of<float>(const ElasticConsts* __restrict__ ec, int b) {
  return ec->sig6f[b];
}
template <typename T> __device__ __forceinline__ const T* e6m_of(const ElasticConsts* __restrict__ ec, int a);
template <> __device__ __forceinline__ const double* e6m_of<double>(const ElasticConsts* __restrict__ ec, int a) {
  return ec->e6m[a];
}
template <> __device__ __forceinline__ const float* e6m_of<float>(const ElasticConsts* __restrict__ ec, int a) {
  return ec->e6mf[a];
}

// Recomputed B(k) in the stress-spectrum form.  With s = sum_b s_b v_b
// (s_b = C:e_b symmetric, so s_a:e_b = e_a:s_b), t = s f and g = G t,
//   sum_b B_ab v_b = (e_a : s - u_a . g) / mu = e_a : r / mu,
//   r = s - C : sym(g (x) f),
// the B_ab = (s_a:e_b - u_a.G.u_b)/mu of Bmatrix (G.c:386-402) with u_a =
// s_a f and the sums regrouped: 6 S + 6 S complex-by-real products per
// k-point instead of S^2 + 6 S.  For the cubic C of cijkl,
// C:X = lambda tr(X) I + 2 mu X + mu' diag(X).  f and G are evaluated in
// FP64 and rounded to T; the sums run in T.  Writes out[a * ostride] (v
// stays in registers, so out may alias where v was read from).
template <typename T, int SM, bool NTS>
__device__ __forceinline__ void contract_stress_form(const cx<T> (&v)[SM], const ElasticConsts* __restrict__ ec,
                                                     const double f[3], const double g[6], int S, cx<T>* out,
                                                     long ostride) {
  cx<T> sv[6];
#pragma unroll
  for (int t = 0; t < 6; t++) sv[t] = {(T)0, (T)0};
#pragma unroll
  for (int b = 0; b < SM; b++) {
    if (b >= S) continue;
    const T* sb = sig6_of<T>(ec, b);
#pragma unroll
    for (int t = 0; t < 6; t++) {
      sv[t].x += sb[t] * v[b].x;
      sv[t].y += sb[t] * v[b].y;
    }
  }
  const T f0 = (T)f[0], f1 = (T)f[1], f2 = (T)f[2];
  // t = s f
  cx<T> tv[3];
  tv[0] = {sv[0].x * f0 + sv[1].x * f1 + sv[2].x * f2, sv[0].y * f0 + sv[1].y * f1 + sv[2].y * f2};
  tv[1] = {sv[1].x * f0 + sv[3].x * f1 + sv[4].x * f2, sv[1].y * f0 + sv[3].y * f1 + sv[4].y * f2};
  tv[2] = {sv[2].x * f0 + sv[4].x * f1 + sv[5].x * f2, sv[2].y * f0 + sv[4].y * f1 + sv[5].y * f2};
  // g = G t
  const T g0 = (T)g[0], g1 = (T)g[1], g2 = (T)g[2], g3 = (T)g[3], g4 = (T)g[4], g5 = (T)g[5];
  cx<T> gv[3];
  gv[0] = {g0 * tv[0].x + g1 * tv[1].x + g2 * tv[2].x, g0 * tv[0].y + g1 * tv[1].y + g2 * tv[2].y};
  gv[1] = {g1 * tv[0].x + g3 * tv[1].x + g4 * tv[2].x, g1 * tv[0].y + g3 * tv[1].y + g4 * tv[2].y};
  gv[2] = {g2 * tv[0].x + g4 * tv[1].x + g5 * tv[2].x, g2 * tv[0].y + g4 * tv[1].y + g5 * tv[2].y};
  // r = s - C : Y, Y = sym(g (x) f): diagonal lambda tr(Y) + (2 mu + mu') Y_ii, off-diagonal mu (g_i f_j + g_j f_i)
  const T lam = (T)ec->ll, mu = (T)ec->mu, dg = (T)(2 * ec->mu + ec->mup);
  const cx<T> y00 = {gv[0].x * f0, gv[0].y * f0}, y11 = {gv[1].x * f1, gv[1].y * f1},
              y22 = {gv[2].x * f2, gv[2].y * f2};
  const cx<T> tr = {lam * (y00.x + y11.x + y22.x), lam * (y00.y + y11.y + y22.y)};
  cx<T> rv[6];
  rv[0] = {sv[0].x - (tr.x + dg * y00.x), sv[0].y - (tr.y + dg * y00.y)};
  rv[3] = {sv[3].x - (tr.x + dg * y11.x), sv[3].y - (tr.y + dg * y11.y)};
  rv[5] = {sv[5].x - (tr.x + dg * y22.x), sv[5].y - (tr.y + dg * y22.y)};
  rv[1] = {sv[1].x - mu * (gv[0].x * f1 + gv[1].x * f0), sv[1].y - mu * (gv[0].y * f1 + gv[1].y * f0)};
  rv[2] = {sv[2].x - mu * (gv[0].x * f2 + gv[2].x * f0), sv[2].y - mu * (gv[0].y * f2 + gv[2].y * f0)};
  rv[4] = {sv[4].x - mu * (gv[1].x * f2 + gv[2].x * f1), sv[4].y - mu * (gv[1].y * f2 + gv[2].y * f1)};
#pragma unroll
  for (int a = 0; a < SM; a++) {
    if (a >= S) continue;
    const T* ea = e6m_of<T>(ec, a);
    cx<T> o = {(T)0, (T)0};
#pragma unroll
    for (int t = 0; t < 6; t++) {
      o.x += ea[t] * rv[t].x;
      o.y += ea[t] * rv[t].y;
    }
    if constexpr (NTS) gstore<T>(out + a * ostride, o);
    else out[a * ostride] = o;
  }
}

// contract_stress_form for v read from (and the result written back to)
// LDS, io[b * stride]: the v_b are accumulated into s one at a time and the
// outputs written once all of them are read, so only s, t, g and r stay in
// registers.  No virtualevolv term (the GL steps).
template <typename T, int S>
__device__ __forceinline__ void contract_stress_form_lds(cx<T>* io, int stride, const ElasticConsts* __restrict__ ec,
                                                         long pt) {
  double f[3], g[6];
  kvec(pt, ec, f);
  if (!green6(f, ec, g)) {  // B(0) = 0
#pragma unroll
    for (int a = 0; a < S; a++) io[a * stride] = {(T)0, (T)0};
    return;
  }
  cx<T> sv[6];
#pragma unroll
  for (int t = 0; t < 6; t++) sv[t] = {(T)0, (T)0};
#pragma unroll
  for (int b = 0; b < S; b++) {
    const cx<T> vb = io[b * stride];
    const T* sb = sig6_of<T>(ec, b);
#pragma unroll
    for (int t = 0; t < 6; t++) {
      sv[t].x += sb[t] * vb.x;
      sv[t].y += sb[t] * vb.y;
    }
  }
  const T f0 = (T)f[0], f1 = (T)f[1], f2 = (T)f[2];
  cx<T> tv[3];
  tv[0] = {sv[0].x * f0 + sv[1].x * f1 + sv[2].x * f2, sv[0].y * f0 + sv[1].y * f1 + sv[2].y * f2};
  tv[1] = {sv[1].x * f0 + sv[3].x * f1 + sv[4].x * f2, sv[1].y * f0 + sv[3].y * f1 + sv[4].y * f2};
  tv[2] = {sv[2].x * f0 + sv[4].x * f1 + sv[5].x * f2, sv[2].y * f0 + sv[4].y * f1 + sv[5].y * f2};
  const T g0 = (T)g[0], g1 = (T)g[1], g2 = (T)g[2], g3 = (T)g[3], g4 = (T)g[4], g5 = (T)g[5];
  cx<T> gv[3];
  gv[0] = {g0 * tv[0].x + g1 * tv[1].x + g2 * tv[2].x, g0 * tv[0].y + g1 * tv[1].y + g2 * tv[2].y};
  gv[1] = {g1 * tv[0].x + g3 * tv[1].x + g4 * tv[2].x, g1 * tv[0].y + g3 * tv[1].y + g4 * tv[2].y};
  gv[2] = {g2 * tv[0].x + g4 * tv[1].x + g5 * tv[2].x, g2 * tv[0].y + g4 * tv[1].y + g5 * tv[2].y};
  const T lam = (T)ec->ll, mu = (T)ec->mu, dg = (T)(2 * ec->mu + ec->mup);
  const cx<T> y00 = {gv[0].x * f0, gv[0].y * f0}, y11 = {gv[1].x * f1, gv[1].y * f1},
              y22 = {gv[2].x * f2, gv[2].y * f2};
  const cx<T> tr = {lam * (y00.x + y11.x + y22.x), lam * (y00.y + y11.y + y22.y)};
  cx<T> rv[6];
  rv[0] = {sv[0].x - (tr.x + dg * y00.x), sv[0].y - (tr.y + dg * y00.y)};
  rv[3] = {sv[3].x - (tr.x + dg * y11.x), sv[3].y - (tr.y + dg * y11.y)};
  rv[5] = {sv[5].x - (tr.x + dg * y22.x), sv[5].y - (tr.y + dg * y22.y)};
  rv[1] = {sv[1].x - mu * (gv[0].x * f1 + gv[1].x * f0), sv[1].y - mu * (gv[0].y * f1 + gv[1].y * f0)};
  rv[2] = {sv[2].x - mu * (gv[0].x * f2 + gv[2].x * f0), sv[2].y - mu * (gv[0].y * f2 + gv[2].y * f0)};
  rv[4] = {sv[4].x - mu * (gv[1].x * f2 + gv[2].x * f1), sv[4].y - mu * (gv[1].y * f2 + gv[2].y * f1)};
#pragma unroll
  for (int a = 0; a < S; a++) {
    const T* ea = e6m_of<T>(ec, a);
    cx<T> o = {(T)0, (T)0};
#pragma unroll
    for (int t = 0; t < 6; t++) {
      o.x += ea[t] * rv[t].x;
      o.y += ea[t] * rv[t].y;
    }
    io[a * stride] = o;
  }
}

template <typename T, int SM, bool UNROLL_A, bool NTS = false>
__device__ __forceinline__ void contract_point(const cx<T> (&v)[SM], const ElasticConsts* __restrict__ ec, long pt,
                                               int S, int quirk, bool stored, const T* __restrict__ Bsym, long P,
                                               cx<T>* out, long ostride) {
  auto put = [](cx<T>* q, cx<T> x) {
    if constexpr (NTS) gstore<T>(q, x);
    else *q = x;
  };
  double f[3], g[6];
  kvec(pt, ec, f);
  if (!green6(f, ec, g)) {  // B(0) = 0 and DD(0) = 0
#pragma unroll
    for (int a = 0; a < SM; a++)
      if (a < S) put(out + a * ostride, {(T)0, (T)0});
    return;
  }
  if (stored) {
    auto row = [&](int a) -> cx<T> {
      cx<T> acc = {(T)0, (T)0};
#pragma unroll
      for (int b = 0; b < SM; b++) {
        if (b >= S) continue;
        const T bab = Bsym[(long)tri_index(a, b, S) * P + pt];
        acc.x += v[b].x * bab;
        acc.y += v[b].y * bab;
      }
      return acc;
    };
    if (UNROLL_A) {
#pragma unroll
      for (int a = 0; a < SM; a++)
        if (a < S) put(out + a * ostride, row(a));
    } else {
#pragma unroll 1
      for (int a = 0; a < S; a++) put(out + a * ostride, row(a));
    }
  } else {
    contract_stress_form<T, SM, NTS>(v, ec, f, g, S, out, ostride);
  }
  if (quirk) {
    double tf[3];
    const double* t9 = ec->T;
    tf[0] = t9[0] * f[0] + t9[1] * f[1] + t9[2] * f[2];
    tf[1] = t9[3] * f[0] + t9[4] * f[1] + t9[5] * f[2];
    tf[2] = t9[6] * f[0] + t9[7] * f[1] + t9[8] * f[2];
#pragma unroll
    for (int a = 0; a < SM; a++) {
      if (a >= S) continue;
      double u[3], gu[3];
      sig_f(ec, a, f, u);
      gmul(g, u, gu);
      const T dd = (T)((ec->Tsum[a] - (gu[0] * tf[0] + gu[1] * tf[1] + gu[2] * tf[2])) * ec->inv_mu);
      cx<T> o = out[a * ostride];
      o.x += v[a].x * dd;
      o.y += v[a].y * dd;
      out[a * ostride] = o;
    }
  }
}

// Pointwise contraction in place on W (used when the fused column kernel
// does not fit).
// IL: the interleaved layout (fields of a point together).
// Threads per block: 128 in FP32 (0.076 -> 0.072 ms at 128^3 x 12), 256 in
// FP64 (128 measured 0.135 -> 0.143 ms).
template <typename T> constexpr int contract_threads() { return sizeof(T) == 4 ? 128 : 256; }
template <typename T, int SM, bool STORED, bool IL>
__global__ void __launch_bounds__(contract_threads<T>())
k_contract(cx<T>* __restrict__ W, const T* __restrict__ Bsym, const ElasticConsts* __restrict__ ec,
           long P, int quirk) {
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= P) return;
  const int S = ec->ns;
  W += (long)blockIdx.y * S * P;  // instance
  cx<T> v[SM];
#pragma unroll
  for (int b = 0; b < SM; b++)
    if (b < S) v[b] = IL ? W[pt * S + b] : W[(long)b * P + pt];
  // non-temporal stores in FP64 (0.142 -> 0.137 ms at 128^3 x 12); the FP32
  // array (201 MB there) stays in the 256 MB last-level cache for the next
  // pass, which then measured faster with ordinary stores
  contract_point<T, SM, true, sizeof(T) == 8>(v, ec, pt, S, quirk, STORED, Bsym, P, IL ? W + pt * S : W + pt,
                                              IL ? 1 : P);
}

// ---------------------------------------------------------------- PFDD_REF32
// The reference's float arithmetic, for checks against the compiled
// reference itself: its FOURN transform bit for bit, its float data2
// accumulation and its float uq2*data2/nsize.
//
// One axis of fourn (G.c:2758-2827) on lines of n = 2^logn points of an
// [outer][n][inner] float array: the bit-reversal permutation, then the
// radix-2 decimation-in-time stages, each butterfly in float with the
// twiddle (float)wr, (float)wi of the reference's double recurrence
// (tw: n - 1 entries, the half-span-h stage at tw[h - 1 + j]), every product
// and sum rounded on its own.  The stages of one line are sequential and its
// butterflies within a stage independent, so the order of work does not
// change a bit.  fourn runs the axes N3, N2, N1 in both directions.
// One block per line.
constexpr int FOURN_THREADS = 256;
__global__ void __launch_bounds__(FOURN_THREADS)
k_fourn_axis(cx<float>* __restrict__ data, long inner, int n, int logn, const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  __shared__ cx<float> sm[1024];
  const long line = blockIdx.x;
  const long o = line / inner, i = line - o * inner;
  cx<float>* base = data + o * n * inner + i;
  for (int k = threadIdx.x; k < n; k += blockDim.x)
    sm[k] = base[(long)(__brev((unsigned)k) >> (32 - logn)) * inner];
  __syncthreads();
  for (int h = 1; h < n; h <<= 1) {
    for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const int j = b & (h - 1), k1 = ((b - j) << 1) + j, k2 = k1 + h;
      const float2 w = tw[h - 1 + j];
      const cx<float> d1 = sm[k1], d2 = sm[k2];
      const float tr = w.x * d2.x - w.y * d2.y;
      const float ti = w.x * d2.y + w.y * d2.x;
      sm[k2] = {d1.x - tr, d1.y - ti};
      sm[k1] = {d1.x + tr, d1.y + ti};
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) base[(long)k * inner] = sm[k];
}

// Fmatrix's D_a (G.c:512-576) at one k-point, in Fmatrix's own arithmetic:
// its G, A and D are float arrays (G.c:424-429), so every assignment to them
// rounds to float, while the products and sums on the right are double.  The
// loops keep Fmatrix's order (A_mnuv over i, j, k, l; D_a,ij over k, l), which
// fixes the rounding; G_ki, which Fmatrix recomputes inside the loop nest with
// the same inputs, is computed once.  C is the unrotated C_ijkl (Fmatrix's
// C_rotate for the identity orientation).  D[a][i*3+j] for a < S.
template <int SM>
__device__ void fmatrix_d_r32(const double fk[3], const ElasticConsts* __restrict__ ec, int S, float (&D)[SM][9]) {
#pragma clang fp contract(off)
  for (int a = 0; a < SM; a++)
    for (int t = 0; t < 9; t++) D[a][t] = 0.0f;
  const double fk2 = fk[0] * fk[0] + fk[1] * fk[1] + fk[2] * fk[2];
  if (!(fk2 > 0)) return;
  const double mu = ec->mu, ll = ec->ll, mup = ec->mup;
  double z2[3];
  for (int i = 0; i < 3; i++) z2[i] = fk[i] * fk[i] / fk2;
  const double kk =
      1.0 + (mu + ll) * (1 / (mu + mup * z2[0]) * z2[0] + 1 / (mu + mup * z2[1]) * z2[1] + 1 / (mu + mup * z2[2]) * z2[2]);
  float G[3][3];  // G[k][i]
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 3; i++)
      G[k][i] = (float)(1 / fk2 *
                        ((i == k ? 1 : 0) / (mu + mup * z2[k]) -
                         fk[i] * fk[k] / fk2 / (mu + mup * z2[i]) / (mu + mup * z2[k]) * (mu + ll) / kk));
  const double* C = ec->C;  // C[((i*3+j)*3+k)*3+l]
  for (int mn = 0; mn < 9; mn++) {
    float Arow[9];  // A[m][n][u][v]
    for (int uv = 0; uv < 9; uv++) {
      float A = 0.0f;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
              A = (float)((double)A - C[(k * 3 + l) * 9 + uv] * C[(i * 3 + j) * 9 + mn] * (double)G[k][i] * fk[j] * fk[l]);
      Arow[uv] = (float)((double)A + C[mn * 9 + uv]);
    }
    for (int a = 0; a < SM; a++) {
      if (a >= S) continue;
      float d = 0.0f;
      for (int kl = 0; kl < 9; kl++) d = (float)((double)d + (double)Arow[kl] * ec->eps[a][kl]);
      D[a][mn] = d;
    }
  }
}

// The contraction in the reference's arithmetic (G.c:4609-4629): data2_a
// starts at 0 and takes data_b * BB_ab for b ascending, each product and sum
// in double and the sum rounded to float on every store.  B_ab is formed in
// FP64 at the k-point and scaled by 1/mu (the reference's stored BB, G.c:395-401).  With
// quirk, virtualevolv's nine terms data_a * DD (G.c:1426-1447) follow, in
// the reference's order, with DD_a(v) = D_a[v % 3][v / 3] / mu (the slot its
// index I5V(..., a, v) lands on, G.c:38, 56) and D_a from fmatrix_d_r32.
template <int SM>
__global__ void __launch_bounds__(256)
k_contract_r32(cx<float>* __restrict__ W, const ElasticConsts* __restrict__ ec, long P, int quirk) {
#pragma clang fp contract(off)
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= P) return;
  const int S = ec->ns;
  cx<float> v[SM];
#pragma unroll
  for (int b = 0; b < SM; b++)
    if (b < S) v[b] = W[(long)b * P + pt];
  double f[3], g[6];
  kvec(pt, ec, f);
  const bool nz = green6(f, ec, g);
  double u[SM][3];
#pragma unroll
  for (int b = 0; b < SM; b++)
    if (b < S) sig_f(ec, b, f, u[b]);
  float Dq[SM][9];  // Fmatrix's D (prologue only)
  if (quirk) fmatrix_d_r32<SM>(f, ec, S, Dq);
#pragma unroll
  for (int a = 0; a < SM; a++) {
    if (a >= S) continue;
    double gu[3];
    gmul(g, u[a], gu);
    float ar = 0.0f, ai = 0.0f;
#pragma unroll
    for (int b = 0; b < SM; b++) {
      if (b >= S) continue;
      const double bab = nz ? (ec->SE[a][b] - (gu[0] * u[b][0] + gu[1] * u[b][1] + gu[2] * u[b][2])) * ec->inv_mu : 0.0;
      ar = (float)((double)ar + (double)v[b].x * bab);
      ai = (float)((double)ai + (double)v[b].y * bab);
    }
    if (quirk) {
      for (int vv = 0; vv < 9; vv++) {
        const double dd = (double)Dq[a][(vv % 3) * 3 + vv / 3] / ec->mu;
        ar = (float)((double)ar + (double)v[a].x * dd);
        ai = (float)((double)ai + (double)v[a].y * dd);
      }
    }
    W[(long)a * P + pt] = {ar, ai};
  }
}

// ---------------------------------------------------------------- fused kernels
// K1: one block owns `tile` rows (along N3) of one field.  Inverse transform
// along N3 finishes data2 in real space (G.c:4638), then plasticevolv
// (G.c:3470-3501) updates xi, the new xi is copied into the rows
// (G.c:4581-4590) and transformed forward along N3 (first pass of
// G.c:4599).  The block's sum of Re xi / P goes to partials[f][block].
template <typename T, int N, bool MASK, bool TAUPP, typename M = Master<T>>
__global__ void __launch_bounds__((tile_rows<T, N>() * make_plan(N).tpp))
k_row_step(cx<T>* __restrict__ W, M xi, const uint8_t* __restrict__ mask,
           const double* __restrict__ taupp, StepArgs a, int tile, const cx<T>* __restrict__ tw,
           double* __restrict__ partials, const int* __restrict__ step_ctr) {
  constexpr int TPP = make_plan(N).tpp, TILE = tile_rows<T, N>(), NT = TILE * TPP, LD = pitch<N>();
  static_assert(NT % 64 == 0, "block must be whole waves");
  __shared__ cx<T> sm[TILE * LD];
  __shared__ double red[NT / 64];
  __shared__ cx<T> twl[N];
  TwStage<T, N, NT> ts;  // stored to LDS after the first pass's HBM reads are issued
  ts.load(tw);
  const long P = a.P;
  const long rows_per_field = P / N;
  const long row0 = (long)blockIdx.x * tile;
  const int f = (int)(row0 / rows_per_field);
  const long pt0 = (row0 - (long)f * rows_per_field) * N;
  const long off = (long)f * P + pt0;
  cx<T>* wb = W + off;
  const RowPencils<T> g{wb, N};
  const LdsPencils<T> l{sm, LD};
  const int pen = (int)threadIdx.x / TPP, q = (int)threadIdx.x % TPP;
  // In FP64 the thread's xi values are loaded before the inverse transform,
  // so their latency overlaps it (0.283 -> 0.277 ms at 128^3 x 12); in FP32
  // that measured slower (0.175 -> 0.205 ms) and they are loaded in the
  // update loop, unrolled (0.184 -> 0.175 ms against a rolled loop).
  constexpr bool EARLY = sizeof(T) == 8;
  constexpr int EPT = (TILE * N + NT - 1) / NT;
  double2 xv[EARLY ? EPT : 1];
  if constexpr (EARLY) {
#pragma unroll
    for (int k = 0; k < EPT; k++) {
      const int e = threadIdx.x + k * NT;
      if (e < tile * N) xv[k] = xi.load(off + e);
    }
  }
  fft_pencils<T, N, 1>(g, l, l, pen, 1 << 30, tile, q, (T)1, twl, [&] { ts.store(twl); });
  const double inv_n = 1.0 / (double)P;
  const FieldCoef c = field_coef(a, f, step_ctr);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < EPT; k++) {
    const int e = threadIdx.x + k * NT;
    if (e >= tile * N) break;
    const int li = (e / N) * LD + (e % N);
    const cx<T> d = sm[li];
    double2 x = EARLY ? xv[EARLY ? k : 0] : xi.load(off + e);
    if (!MASK || mask[off + e] == 0) {
      gl_point<TAUPP, lean_master<M>()>(x, (double)d.x, (double)d.y, c, a, taupp, pt0 + e, inv_n);
      xi.store(off + e, x);
    }
    sm[li] = {(T)x.x, (T)x.y};
    acc += x.x;  // (the block sum is scaled by 1/P once)
  }
  __syncthreads();
  fft_pencils<T, N, 1>(l, l, g, pen, 1 << 30, tile, q, (T)-1, twl);
  const double sum = block_sum(acc, red) * inv_n;
  if (threadIdx.x == 0) partials[(long)f * (rows_per_field / tile) + (pt0 / N) / tile] = sum;
}

// ---------------------------------------------------------------- interleaved layout
// W[instance][point][ns]: the fields of a grid point sit together, so the
// rows (or columns) of every field that one block owns are a single
// contiguous stretch of W.  The block stages that stretch into LDS pencils
// with coalesced loads, transforms there and writes it back the same way.
// One TPP-lane group per pencil; blockDim.x >= pencils * TPP.

// Walks e = (hi * M + mid) * S + lo in steps of `step` with carries instead
// of divisions (S and M are run-time).
struct Walk3 {
  int lo, mid, hi, dlo, dmid, dhi, S, M;
  __device__ __forceinline__ Walk3(int e, int step, int S_, int M_) : S(S_), M(M_) {
    lo = e % S;
    mid = (e / S) % M;
    hi = e / S / M;
    dlo = step % S;
    dmid = (step / S) % M;
    dhi = step / S / M;
  }
  __device__ __forceinline__ void next() {
    lo += dlo;
    mid += dmid;
    hi += dhi;
    if (lo >= S) {
      lo -= S;
      mid++;
    }
    if (mid >= M) {
      mid -= M;
      hi++;
    }
  }
};

// Staging of a block's stretch of W (chunk = npen * N elements, blockDim.x >=
// npen * TPP threads, so at most EPT = N / TPP per thread).  Element
// e = (hi * M + mid) * S + lo sits at g[goff(w)] and at sm[lidx(w)].  With
// BATCH the thread's elements go in batches (4 of 16 bytes or 8 of 8 bytes)
// whose loads are all issued before the batch's first store; measured at
// 128^3 x 12 this helps the FP32 row kernel (0.31 -> 0.22 ms) and slows the
// FP64 kernels, whose extra registers spill.
template <typename T, int EPT, bool IN, bool BATCH, typename GOff, typename LIdx>
__device__ __forceinline__ void il_stage(cx<T>* g, cx<T>* sm, int chunk, int S, int M, GOff goff, LIdx lidx) {
  if constexpr (!BATCH) {
    Walk3 w(threadIdx.x, blockDim.x, S, M);
    for (int e = threadIdx.x; e < chunk; e += blockDim.x, w.next()) {
      if (IN)
        sm[lidx(w)] = g[goff(w)];
      else
        g[goff(w)] = sm[lidx(w)];
    }
  } else {
    constexpr int BB = (sizeof(T) == 8 ? 4 : 8) < EPT ? (sizeof(T) == 8 ? 4 : 8) : EPT;
    Walk3 wr(threadIdx.x, blockDim.x, S, M), ww = wr;
#pragma unroll 1
    for (int k0 = 0; k0 < EPT; k0 += BB) {
      cx<T> buf[BB];
#pragma unroll
      for (int k = 0; k < BB; k++, wr.next())
        if (k0 + k < EPT && (int)threadIdx.x + (k0 + k) * (int)blockDim.x < chunk)
          buf[k] = IN ? g[goff(wr)] : sm[lidx(wr)];
#pragma unroll
      for (int k = 0; k < BB; k++, ww.next())
        if (k0 + k < EPT && (int)threadIdx.x + (k0 + k) * (int)blockDim.x < chunk) {
          if (IN)
            sm[lidx(ww)] = buf[k];
          else
            g[goff(ww)] = buf[k];
        }
    }
  }
}

// prologue copy xi -> W in the interleaved layout; n = B * S * P
template <typename T>
__global__ void k_xi_to_w_il(Master<T> xi, cx<T>* __restrict__ W, long P, int S, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const long ip = e / S;  // instance * P + point
  const int a = (int)(e - ip * S);
  const long b = ip / P, pt = ip - b * P;
  const double2 x = xi.load((b * S + a) * P + pt);
  W[e] = {(T)x.x, (T)x.y};
}

// Row kernel, interleaved layout: block (blk, inst) owns rows blk*R ..
// blk*R+R-1 (along N3) of every field: pencil a*R + r.  Inverse N3
// transform (G.c:4638), plasticevolv (G.c:3470-3501) with the copy into the
// FFT array (G.c:4581-4590), forward N3 transform (first pass of G.c:4599).
// Each field's sum of Re xi / P over the block goes to
// partials[inst*S + a][blk]; it adds the field's threads in thread order.
// Dynamic LDS: S*R pencils of pitch N+1, then blockDim.x doubles.
template <typename T, int N, bool MASK, bool TAUPP, typename M = Master<T>>
__global__ void __launch_bounds__(1024)
k_row_step_il(cx<T>* __restrict__ W, M xi, const uint8_t* __restrict__ mask,
              const double* __restrict__ taupp, StepArgs a, int R, const cx<T>* __restrict__ tw,
              double* __restrict__ partials, const int* __restrict__ step_ctr) {
  constexpr int TPP = make_plan(N).tpp, LD = pitch<N>();
  extern __shared__ unsigned char lds_raw[];
  cx<T>* sm = reinterpret_cast<cx<T>*>(lds_raw);
  const int S = a.ns, npen = S * R, NT = blockDim.x;
  double* red = reinterpret_cast<double*>(sm + (long)npen * LD);
  __shared__ cx<T> twl[N];
  stage_twiddles<T, N>(twl, tw);
  const long P = a.P;
  const long nblk = P / ((long)N * R);
  const long blk = blockIdx.x;
  const int inst = blockIdx.y;
  const long pt0 = blk * R * N;
  cx<T>* wb = W + ((long)inst * P + pt0) * S;
  const int chunk = R * N * S;
  constexpr int EPT = N / TPP;  // elements per pencil lane
  constexpr bool BATCH = sizeof(T) == 4;
  // element (r * N + i) * S + a: pencil a * R + r, position i
  auto goff = [](const Walk3& w) { return (long)(w.hi * w.M + w.mid) * w.S + w.lo; };
  auto lidx = [R](const Walk3& w) { return (w.lo * R + w.hi) * LD + w.mid; };
  il_stage<T, EPT, true, BATCH>(wb, sm, chunk, S, N, goff, lidx);
  __syncthreads();
  const int p = threadIdx.x / TPP, q = threadIdx.x % TPP;
  fft_lds_multi<T, N, 1, LD>(sm, p, 1 << 30, npen, q, (T)1, twl);
  const double inv_n = 1.0 / (double)P;
  double acc = 0.0;
  if (p < npen) {
    const int fa = p / R, r = p - fa * R, f = inst * S + fa;
    const FieldCoef c = field_coef(a, f, step_ctr);
    const long rpt = pt0 + (long)r * N;  // first point of the row
    const long off = (long)f * P + rpt;
    constexpr int XB = !BATCH ? 1 : EPT < 4 ? EPT : 4;  // xi loads in flight per batch
#pragma unroll
    for (int j0 = 0; j0 < EPT; j0 += XB) {
      double2 xv[XB];
#pragma unroll
      for (int j = 0; j < XB; j++)
        if (j0 + j < EPT) xv[j] = xi.load(off + q + (j0 + j) * TPP);
#pragma unroll
      for (int j = 0; j < XB; j++) {
        if (j0 + j >= EPT) continue;
        const int i = q + (j0 + j) * TPP, li = p * LD + i;
        const cx<T> d = sm[li];
        double2 x = xv[j];
        if (!MASK || mask[off + i] == 0) {
          gl_point<TAUPP, lean_master<M>()>(x, (double)d.x, (double)d.y, c, a, taupp, rpt + i, inv_n);
          xi.store(off + i, x);
        }
        sm[li] = {(T)x.x, (T)x.y};
        acc += x.x;  // (scaled by 1/P once per field)
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < S) {
    const int g = R * TPP;
    double s = 0.0;
    for (int k = 0; k < g; k++) s += red[threadIdx.x * g + k];
    partials[((long)inst * S + threadIdx.x) * nblk + blk] = s * inv_n;
  }
  fft_lds_multi<T, N, 1, LD>(sm, p, 1 << 30, npen, q, (T)-1, twl);
  il_stage<T, EPT, false, BATCH>(wb, sm, chunk, S, N, goff, lidx);
}

// Row kernel, interleaved layout, with S and the rows per block (R) fixed at
// compile time, no frozen points and a uniform shear.  One TPP-lane group per
// pencil (field f = p / R, row r = p % R), so each thread transforms one
// pencil and updates the EPT = N / TPP points of it that it holds; its master
// values are loaded before the inverse transform, so that they arrive while
// it runs, and its share of the block's contiguous stretch of W is loaded
// (and later stored) in one batch.  Per point and per pencil the same
// arithmetic as k_row_step_il; a field's block sum adds its pencils' lanes in
// a fixed butterfly.  grid = (row blocks, instances).
template <typename T, int N, int S, int R> struct RowIlFixed {
  static constexpr int TPP = make_plan(N).tpp, SH = 3, LD = padded_pitch<N, SH>(), NPEN = S * R;
  static constexpr int NT = NPEN * TPP;             // one TPP-lane group per pencil
  static constexpr int EPT = N / TPP;               // points per lane
  static constexpr int EL = R * N * S / NT;         // stretch elements per thread (= EPT)
  static constexpr size_t LDS = (size_t)NPEN * LD * sizeof(cx<T>);
};
template <typename T, int N, int S, int R, typename M>
__global__ void __launch_bounds__((RowIlFixed<T, N, S, R>::NT))
k_row_il_fixed(cx<T>* __restrict__ W, M xi, StepArgs a, const cx<T>* __restrict__ tw,
               double* __restrict__ partials, const int* __restrict__ step_ctr) {
  using K = RowIlFixed<T, N, S, R>;
  constexpr int NT = K::NT, LD = K::LD, TPP = K::TPP, EPT = K::EPT, EL = K::EL, SH = K::SH;
  static_assert(NT % 64 == 0 && NT <= 1024 && K::LDS <= 64 * 1024 && (R * N * S) % NT == 0, "tile shape");
  static_assert(TPP <= 64 && 64 % TPP == 0, "a pencil's lanes lie in one wave");
  __shared__ cx<T> sm[K::NPEN * LD];
  __shared__ cx<T> twl[N];  // the twiddle table, read by every pass
  const long P = a.P;
  const long nblk = P / ((long)N * R);
  const long blk = blockIdx.x;
  const int inst = blockIdx.y, t = threadIdx.x;
  const long pt0 = blk * R * N;
  cx<T>* wb = W + ((long)inst * P + pt0) * S;  // the block's R*N*S contiguous elements
  const int p = t / TPP, q = t % TPP, f = p / R, r = p % R;
  const long xoff = ((long)inst * S + f) * P + pt0 + (long)r * N;  // this pencil's row of the master
  // Loads in flight together: the stretch of W, the twiddle table, and the
  // thread's master points, which the update uses after the inverse
  // transform (the LDS stores below wait only for the first two).
  TwStage<T, N, NT> ts;
  typename M::Raw xv[EPT];
  // Element t + k*NT of the stretch is point po0 + k*KP of field fe: with
  // NT a multiple of S, the thread's field is the same for every k, and
  // with KP a multiple of 2^SH the padded LDS slot moves by a constant.
  constexpr int KP = NT / S;
  static_assert(NT % S == 0 && KP % (1 << SH) == 0 && N % KP == 0, "stretch walk");
  const int po0 = t / S, fe = t % S;
  cx<T>* lrow = sm + (fe * R + po0 / N) * LD + lds_pos<SH>(po0 % N);
  auto lslot = [](int k) { return (k * KP) / N * LD + (k * KP) % N + ((k * KP) % N >> SH); };
  {
    cx<T> buf[EL];
#pragma unroll
    for (int k = 0; k < EL; k++) buf[k] = wb[t + k * NT];
    ts.load(tw);
#pragma unroll
    for (int j = 0; j < EPT; j++) xv[j] = xi.raw(xoff + q + j * TPP);
#pragma unroll
    for (int k = 0; k < EL; k++) lrow[lslot(k)] = buf[k];
    ts.store(twl);
  }
  __syncthreads();
  fft_lds_multi<T, N, 1, LD, SH>(sm, p, 1 << 30, K::NPEN, q, (T)1, twl);
  const double inv_n = 1.0 / (double)P;
  const FieldCoef c = field_coef(a, inst * S + f, step_ctr);
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < EPT; j++) {
    const int i = q + j * TPP;
    cx<T>* cell = sm + p * LD + lds_pos<SH>(i);
    const cx<T> d = *cell;
    double2 x = M::widen(xv[j]);
    gl_point<false, lean_master<M>()>(x, (double)d.x, (double)d.y, c, a, nullptr, pt0 + (long)r * N + i, inv_n);
    xi.store(xoff + i, x);
    *cell = {(T)x.x, (T)x.y};
    acc += x.x;  // (scaled by 1/P once per field)
  }
  __syncthreads();
  fft_lds_multi<T, N, 1, LD, SH>(sm, p, 1 << 30, K::NPEN, q, (T)-1, twl);
  {
    cx<T> buf[EL];
#pragma unroll
    for (int k = 0; k < EL; k++) buf[k] = lrow[lslot(k)];
#pragma unroll
    for (int k = 0; k < EL; k++) gstore<T>(wb + t + k * NT, buf[k]);
  }
  // the field's R pencils are R * TPP consecutive lanes of one wave (TPP divides 64; R * TPP <= 64)
  static_assert(R * TPP <= 64 && 64 % (R * TPP) == 0, "a field's lanes lie in one wave");
#pragma unroll
  for (int o = R * TPP / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((t % (R * TPP)) == 0) partials[((long)inst * S + f) * nblk + blk] = acc * inv_n;
}

template <typename T, int N> constexpr size_t row_il_lds_bytes(int S, int R, int nt) {
  return (size_t)S * R * pitch<N>() * sizeof(cx<T>) + (size_t)nt * sizeof(double);
}

// Column kernel, interleaved layout: block (tile bt, inst) owns the `tile`
// consecutive (i2,i3) columns from bt*tile on, every field: pencil
// a*tile + p.  Forward N1 transform (last pass of G.c:4599), contraction
// with the recomputed B(k) (G.c:4609-4629), inverse N1 transform (first
// pass of G.c:4638).  Dynamic LDS: S*tile pencils of pitch N+1.
template <typename T, int N, int SM>
__global__ void __launch_bounds__(1024)
k_col_contract_il(cx<T>* __restrict__ W, const ElasticConsts* __restrict__ ec, long P, int tile, int quirk,
                  const cx<T>* __restrict__ tw) {
  constexpr int TPP = make_plan(N).tpp, LD = pitch<N>();
  extern __shared__ unsigned char lds_raw[];
  cx<T>* sm = reinterpret_cast<cx<T>*>(lds_raw);
  __shared__ cx<T> twl[N];
  stage_twiddles<T, N>(twl, tw);
  const int S = ec->ns, npen = S * tile, NT = blockDim.x;
  const long inner = P / N;
  const long m0 = xcd_tile(blockIdx.x, gridDim.x) * tile;
  cx<T>* wb = W + ((long)blockIdx.y * P + m0) * S;  // row n of the tile at wb + n * inner * S
  const int chunk = N * tile * S;
  constexpr int EPT = N / TPP;
  // element (n * tile + c) * S + a: pencil a * tile + c, position n
  const long rs = inner * S;
  auto goff = [rs](const Walk3& w) { return (long)w.hi * rs + w.mid * w.S + w.lo; };
  auto lidx = [tile](const Walk3& w) { return (w.lo * tile + w.mid) * LD + w.hi; };
  il_stage<T, EPT, true, false>(wb, sm, chunk, S, tile, goff, lidx);
  __syncthreads();
  const int p = threadIdx.x / TPP, q = threadIdx.x % TPP;
  fft_lds_multi<T, N, 1, LD>(sm, p, 1 << 30, npen, q, (T)-1, twl);
  for (int e = threadIdx.x; e < N * tile; e += NT) {
    const int c = e / N, n = e - c * N;  // consecutive threads read consecutive LDS words
    cx<T> v[SM];
#pragma unroll
    for (int b = 0; b < SM; b++)
      if (b < S) v[b] = sm[(b * tile + c) * LD + n];
    contract_point<T, SM, false>(v, ec, (long)n * inner + m0 + c, S, quirk, false, (const T*)nullptr, P,
                                 sm + c * LD + n, (long)tile * LD);
  }
  __syncthreads();
  fft_lds_multi<T, N, 1, LD>(sm, p, 1 << 30, npen, q, (T)1, twl);
  il_stage<T, EPT, false, false>(wb, sm, chunk, S, tile, goff, lidx);
}

// Column kernel, interleaved layout, with S and the tile (TC columns) fixed at
// compile time, for the GL steps (no virtualevolv term): one TPP-lane group
// per pencil (field a = p / TC, column c = p % TC); a thread's share of the
// tile is loaded in one batch before its LDS stores, and stored in one batch
// after its LDS reads; the k-points are contracted in the stress-spectrum
// form.  The stores are ordinary (not streaming): a tile's row segments
// (S * TC complex) share cache lines with the neighbouring tiles, and
// non-temporal stores of them wrote 12% more bytes than the data.  Same
// arithmetic per k-point and per pencil as k_col_contract_il.
// grid = (column tiles, instances).
template <typename T, int N, int S, int TC> struct ColIlFixed {
  // Pencil pitch LD and one pad slot per 2^SH positions, chosen by the LDS
  // bank rules measured on gfx950 (64 banks of 4 bytes; a 16-byte read
  // serves 16 lanes a cycle, in the lane groups {0-3, 12-15, 20-27} and
  // {4-11, 16-19, 28-31} of each half-wave; a 16-byte write serves 8
  // consecutive lanes a cycle across 32 banks; an 8-byte read serves 32
  // lanes).  At N = 128 the FP64 kernel takes one pad per 8 points and pitch
  // 147 (the fewest conflict cycles of the pitches and paddings that fit),
  // the FP32 kernel one pad per 16 points (its pencils' 16 lanes then read
  // and write 16 consecutive positions on distinct banks); elsewhere one pad
  // per 8 points.
  static constexpr int TPP = make_plan(N).tpp, SH = N == 128 && sizeof(T) == 4 ? 4 : 3,
                       LD = padded_pitch<N, SH>() + (N == 128 && sizeof(T) == 8 ? 2 : 0), NPEN = S * TC, ROW = S * TC;
  // Minimum waves per SIMD asked of the register allocator: the FP32 form
  // fits in the 72 registers that seven waves allow without spilling (at
  // N = 128 four blocks per CU instead of three: 0.124 -> 0.102 ms at
  // 128^3 x 12); the FP64 form is bound by LDS.
  static constexpr int MINW = sizeof(T) == 4 ? 6 : 1;
  static constexpr int NT = NPEN * TPP;             // one TPP-lane group per pencil
  static constexpr int EL = N * ROW / NT;           // tile elements per thread
  static constexpr size_t LDS = (size_t)NPEN * LD * sizeof(cx<T>);
};
template <typename T, int N, int S, int TC>
__global__ void __launch_bounds__((ColIlFixed<T, N, S, TC>::NT), (ColIlFixed<T, N, S, TC>::MINW))
k_col_il_fixed(cx<T>* __restrict__ W, const ElasticConsts* __restrict__ ec, long P, const cx<T>* __restrict__ tw) {
  using K = ColIlFixed<T, N, S, TC>;
  constexpr int NT = K::NT, LD = K::LD, ROW = K::ROW, EL = K::EL, TPP = K::TPP, SH = K::SH;
  static_assert(NT % 64 == 0 && NT <= 1024 && K::LDS <= 64 * 1024 && (N * ROW) % NT == 0, "tile shape");
  __shared__ cx<T> sm[K::NPEN * LD];
  __shared__ cx<T> twl[N];  // the twiddle table, read by every pass
  const long M = P / N;  // (i2, i3) columns
  const long m0 = xcd_tile(blockIdx.x, gridDim.x) * TC;
  cx<T>* wb = W + ((long)blockIdx.y * P + m0) * S;  // row n of the tile at wb + n * M * S
  const long rs = M * S;
  const int t = threadIdx.x;
  // Tile element t + k*NT is row n0 + k*KN, offset r in the row: with NT a
  // multiple of ROW the thread's offset r (column c = r / S, field a = r % S)
  // is the same for every k.  Row offsets stay below N * rs = P * S < 2^31.
  constexpr int KN = NT / ROW;
  static_assert(NT % ROW == 0 && KN % (1 << SH) == 0, "tile walk");
  const int n0 = t / ROW, r = t % ROW;
  const cx<T>* gsrc = wb + (unsigned)(n0 * (int)rs + r);
  cx<T>* lcol = sm + ((r % S) * TC + r / S) * LD + lds_pos<SH>(n0);  // pencil a * TC + c
  {
    cx<T> buf[EL];
#pragma unroll
    for (int k = 0; k < EL; k++) buf[k] = gsrc[(unsigned)(k * KN * (int)rs)];
    // (the twiddle loads after the tile's, stored with it)
    std::conditional_t<N == 128, Shfl128Tw<T, NT>, TwStage<T, N, NT>> ts;
    ts.load(tw);
#pragma unroll
    for (int k = 0; k < EL; k++) lcol[k * (KN + (KN >> SH))] = buf[k];
    ts.store(twl);
  }
  __syncthreads();
  const int p = t / TPP, q = t % TPP;
  auto fft = [&](T sign) {
    if constexpr (N == 128) fft128_lds_shfl<T, LD, SH>(sm, p, K::NPEN, q, sign, twl);
    else fft_lds_multi<T, N, 1, LD, SH>(sm, p, 1 << 30, K::NPEN, q, sign, twl);
  };
  fft((T)-1);
  for (int e = t; e < N * TC; e += NT) {
    const int c = e / N, n = e % N;  // consecutive threads read consecutive LDS words
    contract_stress_form_lds<T, S>(sm + c * LD + lds_pos<SH>(n), TC * LD, ec, (long)n * M + m0 + c);
  }
  __syncthreads();
  fft((T)1);
  {
    cx<T> buf[EL];
#pragma unroll
    for (int k = 0; k < EL; k++) buf[k] = lcol[k * (KN + (KN >> SH))];
    cx<T>* gdst = wb + (unsigned)(n0 * (int)rs + r);
#pragma unroll
    for (int k = 0; k < EL; k++) gdst[(unsigned)(k * KN * (int)rs)] = buf[k];
  }
}

template <typename T, int N> constexpr size_t col_il_lds_bytes(int S, int tile) {
  return (size_t)S * tile * pitch<N>() * sizeof(cx<T>);
}

// Plane kernel: one block owns the (i2, i3) plane of one field at one i1 and
// keeps it in LDS (row-major, pitch N3+1).  It finishes the inverse
// transform along N2 and N3 (the first N2 pass reads HBM), applies
// plasticevolv (G.c:3470-3501) and the copy to the FFT array, then runs the
// forward N3 and N2 passes (the last one writes HBM): with the N1 passes and
// the contraction this is the whole GL step in two HBM round trips of W.
// Block sum of Re xi / P -> partials[f][i1].
template <typename T, int N2, int N3>
constexpr size_t plane_lds_bytes() { return (size_t)N2 * (N3 + 1) * sizeof(cx<T>); }
// 1024 threads, or 512 for the 128^2 FP32 plane, whose passes would not fit
// in the 128 registers per thread that a 1024-thread block allows.
template <typename T, int N2, int N3>
constexpr int plane_threads() { return (sizeof(T) == 4 && N2 * N3 >= 16384) ? 512 : 1024; }

// The block's work for plane `blk` (= field * N1 + i1); sm is the LDS plane.
template <typename T, int N2, int N3, bool MASK, bool TAUPP, typename M>
__device__ __forceinline__ void plane_body(cx<T>* __restrict__ W, M xi,
                                           const uint8_t* __restrict__ mask, const double* __restrict__ taupp,
                                           const StepArgs& a, const cx<T>* __restrict__ tw2,
                                           const cx<T>* __restrict__ tw3, double* __restrict__ partials,
                                           const int* __restrict__ step_ctr, long blk, cx<T>* sm) {
  constexpr int LD = N3 + 1, PLANE = N2 * N3, PLANE_THREADS = plane_threads<T, N2, N3>();
  constexpr int TPP2 = make_plan(N2).tpp, TPP3 = make_plan(N3).tpp;
  constexpr int PPI2 = PLANE_THREADS / TPP2, PPI3 = PLANE_THREADS / TPP3;
  constexpr int G2 = (N3 + PPI2 - 1) / PPI2, G3 = (N2 + PPI3 - 1) / PPI3;  // N3 column pencils, N2 row pencils
  __shared__ double red[PLANE_THREADS / 64];
  const long P = a.P;
  const int n1 = (int)(P / PLANE);
  const int f = (int)(blk / n1), i1 = (int)(blk - (long)f * n1);
  const long off = (long)f * P + (long)i1 * PLANE;
  const ColPencils<T> gcol{W + off, N3};
  const LdsColPencils<T> lcol{sm, LD};
  const LdsPencils<T> lrow{sm, LD};
  const int t = threadIdx.x;
  // threads past PPI*TPP (when TPP does not divide the block) sit out
  const int c_p0 = t < PPI2 * TPP2 ? t % PPI2 : (1 << 30), c_q = t / PPI2;
  const int r_p0 = t < PPI3 * TPP3 ? t / TPP3 : (1 << 30), r_q = t % TPP3;
  // All of this thread's xi values are loaded in one batch, so
  // their latency is paid once (one block per CU leaves no other block to
  // hide it).  Up to 4 points per thread the batch is issued before the
  // inverse passes and lands while they run; larger batches would not fit
  // in registers beside the passes and are issued after them.
  // (The masked 96^2 FP64 variant has no registers to spare for the batch.)
  constexpr int EPT = (PLANE + PLANE_THREADS - 1) / PLANE_THREADS;
  constexpr bool EARLY = EPT <= 4;
  constexpr bool BATCH = EARLY || !MASK || sizeof(T) == 4;
  double2 xv[EPT];
  auto fetch = [&]() {
#pragma unroll
    for (int k = 0; k < EPT; k++) {
      const int e = t + k * PLANE_THREADS;
      if (BATCH && e < PLANE) xv[k] = xi.load(off + e);
    }
  };
  if (EARLY) fetch();
  // inverse along N2 (columns: pencil i3, pencil-fastest threads), HBM -> LDS
  fft_pencils<T, N2, G2>(gcol, lcol, lcol, c_p0, PPI2, N3, c_q, (T)1, tw2);
  // inverse along N3 (rows)
  fft_pencils<T, N3, G3>(lrow, lrow, lrow, r_p0, PPI3, N2, r_q, (T)1, tw3);
  if (!EARLY) fetch();
  const double inv_n = 1.0 / (double)P;
  const FieldCoef c = field_coef(a, f, step_ctr);
  double acc = 0.0;
  auto point = [&](int e, double2 x) {
    const int li = (e / N3) * LD + (e % N3);
    const cx<T> d = sm[li];
    if (!MASK || mask[off + e] == 0) {
      gl_point<TAUPP, lean_master<M>()>(x, (double)d.x, (double)d.y, c, a, taupp, (long)i1 * PLANE + e, inv_n);
      xi.store(off + e, x);
    }
    sm[li] = {(T)x.x, (T)x.y};
    acc += x.x;  // (the block sum is scaled by 1/P once)
  };
  if constexpr (BATCH) {
#pragma unroll
    for (int k = 0; k < EPT; k++) {
      const int e = t + k * PLANE_THREADS;
      if (e < PLANE) point(e, xv[k]);
    }
  } else {
    for (int e = t; e < PLANE; e += PLANE_THREADS) point(e, xi.load(off + e));
  }
  __syncthreads();
  fft_pencils<T, N3, G3>(lrow, lrow, lrow, r_p0, PPI3, N2, r_q, (T)-1, tw3);
  fft_pencils<T, N2, G2>(lcol, lcol, gcol, c_p0, PPI2, N3, c_q, (T)-1, tw2);
  const double sum = block_sum(acc, red) * inv_n;
  if (t == 0) partials[(long)f * n1 + i1] = sum;
}

template <typename T, int N2, int N3, bool MASK, bool TAUPP, typename M = Master<T>>
__global__ void __launch_bounds__((plane_threads<T, N2, N3>()))
k_plane_step(cx<T>* __restrict__ W, M xi, const uint8_t* __restrict__ mask,
             const double* __restrict__ taupp, StepArgs a, const cx<T>* __restrict__ tw2,
             const cx<T>* __restrict__ tw3, double* __restrict__ partials, const int* __restrict__ step_ctr) {
  extern __shared__ unsigned char lds_raw[];
  // The two axes' twiddle tables in LDS, read there by the passes (stored
  // before the first pass, whose barrier orders them before the second).
  __shared__ cx<T> twl2[N2], twl3[N3];
  stage_twiddles<T, N2>(twl2, tw2);
  stage_twiddles<T, N3>(twl3, tw3);
  plane_body<T, N2, N3, MASK, TAUPP, M>(W, xi, mask, taupp, a, twl2, twl3, partials, step_ctr, blockIdx.x,
                                     reinterpret_cast<cx<T>*>(lds_raw));
}

// K3: one block owns `tile` consecutive (i2,i3) columns of every field.
// Forward transform along N1 (last pass of G.c:4599), contraction with B(k)
// at every k of the tile (G.c:4609-4629), inverse transform along N1 (first
// pass of G.c:4638).  LDS holds S*tile pencils (dynamic, S*tile*LD complex).
// B(k) is recomputed per k (the stored-B mode takes the unfused passes).
constexpr int C3_THREADS = 1024;
// Pencils one block can hold: one per TPP-thread group for FP64 (register
// budget of 128 VGPRs at 1024 threads), two for FP32.
template <typename T, int N> constexpr int c3_max_pencils() {
  return (sizeof(T) == 8 ? 1 : 2) * (C3_THREADS / make_plan(N).tpp);
}

// The block's work for column tile `bt` of instance `inst`; sm holds the
// S * tile pencils.
template <typename T, int N, int SM>
__device__ __forceinline__ void col_contract_body(cx<T>* __restrict__ W, const ElasticConsts* __restrict__ ec, long P,
                                                  int tile, int quirk, const cx<T>* __restrict__ tw, long bt, int inst,
                                                  cx<T>* sm) {
  constexpr int TPP = make_plan(N).tpp, LD = pitch<N>(), PPI = C3_THREADS / TPP;
  constexpr int G = (c3_max_pencils<T, N>() + PPI - 1) / PPI;
  const int S = ec->ns;
  W += (long)inst * S * P;
  const long inner = P / N;
  const long m0 = bt * tile;
  const int npen = S * tile;
  const int per_field = N * tile;
  for (int e = threadIdx.x; e < S * per_field; e += C3_THREADS) {
    const int fld = e / per_field, r = e - fld * per_field, n = r / tile, p = r - n * tile;
    sm[(fld * tile + p) * LD + n] = W[(long)fld * P + (long)n * inner + m0 + p];
  }
  __syncthreads();
  const int q = threadIdx.x % TPP, p0 = threadIdx.x / TPP;
  fft_lds_multi<T, N, G, LD>(sm, p0, PPI, npen, q, (T)-1, tw);
  for (int e = threadIdx.x; e < per_field; e += C3_THREADS) {
    const int p = e / N, n = e - p * N;  // consecutive threads read consecutive LDS words
    cx<T> v[SM];
#pragma unroll
    for (int b = 0; b < SM; b++)
      if (b < S) v[b] = sm[(b * tile + p) * LD + n];
    contract_point<T, SM, false>(v, ec, (long)n * inner + m0 + p, S, quirk, false, (const T*)nullptr, P, sm + p * LD + n,
                          (long)tile * LD);
  }
  __syncthreads();
  fft_lds_multi<T, N, G, LD>(sm, p0, PPI, npen, q, (T)1, tw);
  for (int e = threadIdx.x; e < S * per_field; e += C3_THREADS) {
    const int fld = e / per_field, r = e - fld * per_field, n = r / tile, p = r - n * tile;
    W[(long)fld * P + (long)n * inner + m0 + p] = sm[(fld * tile + p) * LD + n];
  }
}

// grid = (column tiles, instances)
template <typename T, int N, int SM>
__global__ void __launch_bounds__(C3_THREADS)
k_col_contract(cx<T>* __restrict__ W, const ElasticConsts* __restrict__ ec, long P, int tile, int quirk,
               const cx<T>* __restrict__ tw) {
  extern __shared__ unsigned char lds_raw[];
  col_contract_body<T, N, SM>(W, ec, P, tile, quirk, tw, xcd_tile(blockIdx.x, gridDim.x), blockIdx.y,
                              reinterpret_cast<cx<T>*>(lds_raw));
}

// Stored B(k): upper triangle in the working precision, built in FP64.
template <typename T, int SM>
__global__ void __launch_bounds__(256)
k_build_B(T* __restrict__ Bsym, const ElasticConsts* __restrict__ ec, long P) {
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= P) return;
  const int S = ec->ns;
  double f[3], g[6];
  kvec(pt, ec, f);
  const bool nz = green6(f, ec, g);
  double u[SM][3];
#pragma unroll
  for (int b = 0; b < SM; b++)
    if (b < S) {
      const double* s = ec->sig[b];
      u[b][0] = s[0] * f[0] + s[1] * f[1] + s[2] * f[2];
      u[b][1] = s[3] * f[0] + s[4] * f[1] + s[5] * f[2];
      u[b][2] = s[6] * f[0] + s[7] * f[1] + s[8] * f[2];
    }
#pragma unroll
  for (int a = 0; a < SM; a++) {
    if (a >= S) continue;
    double gu[3];
    gmul(g, u[a], gu);
#pragma unroll
    for (int b = 0; b < SM; b++) {
      if (b < a || b >= S) continue;
      const double bab = nz ? (ec->SE[a][b] - (gu[0] * u[b][0] + gu[1] * u[b][1] + gu[2] * u[b][2])) * ec->inv_mu : 0.0;
      Bsym[(long)tri_index(a, b, S) * P + pt] = (T)bab;
    }
  }
}

// Strain spectrum: beta_ij(k) = sum_a FF_a,ij xi_a = f_i (G v)_j with
// v = sum_a u_a xi_a (Fmatrix G.c:565-571, strain G.c:768-781); only its
// symmetric part is kept (strain G.c:806-817): E[t], t over (00,01,02,11,12,22).
template <typename T, int SM>
__global__ void __launch_bounds__(256)
k_strain_spectrum(const cx<T>* __restrict__ W, cx<T>* __restrict__ E, const ElasticConsts* __restrict__ ec, long P) {
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= P) return;
  const int S = ec->ns;
  double f[3], g[6];
  kvec(pt, ec, f);
  const bool nz = green6(f, ec, g);
  double vr[3] = {0, 0, 0}, vi[3] = {0, 0, 0};
  if (nz) {
#pragma unroll
    for (int b = 0; b < SM; b++) {
      if (b >= S) continue;
      const double* s = ec->sig[b];
      const double u0 = s[0] * f[0] + s[1] * f[1] + s[2] * f[2];
      const double u1 = s[3] * f[0] + s[4] * f[1] + s[5] * f[2];
      const double u2 = s[6] * f[0] + s[7] * f[1] + s[8] * f[2];
      const cx<T> x = W[(long)b * P + pt];
      vr[0] += u0 * (double)x.x; vr[1] += u1 * (double)x.x; vr[2] += u2 * (double)x.x;
      vi[0] += u0 * (double)x.y; vi[1] += u1 * (double)x.y; vi[2] += u2 * (double)x.y;
    }
  }
  double gr[3], gi[3];
  gmul(g, vr, gr);
  gmul(g, vi, gi);
  const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int t = 0; t < 6; t++) {
    const int i = ii[t], j = jj[t];
    cx<T> o = {(T)0, (T)0};
    if (nz) {
      o.x = (T)((f[i] * gr[j] + f[j] * gr[i]) / 2.);
      o.y = (T)((f[i] * gi[j] + f[j] * gi[i]) / 2.);
    }
    E[(long)t * P + pt] = o;
  }
}

// Per point: eps = Re E / P + sym(avepst); eps* = sum_a e_a Re xi_a;
// sigma = C:(eps - eps*) (stress G.c:897-919); 18 block sums -> partials.
template <typename T, int SM>
__global__ void __launch_bounds__(256)
k_strain_stress(const cx<T>* __restrict__ E, Master<T> xi, const ElasticConsts* __restrict__ ec,
                const double* __restrict__ avepst, long P, double* __restrict__ partials, int nblk,
                double* eps_field, double* sig_field) {
  __shared__ double red[4];
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int S = ec->ns;
  double e[9] = {0}, sg[9] = {0};
  if (pt < P) {
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int t = 0; t < 6; t++) {
      const int i = ii[t], j = jj[t];
      const double v = (double)E[(long)t * P + pt].x / (double)P + (avepst[3 * i + j] + avepst[3 * j + i]) / 2.;
      e[3 * i + j] = v;
      e[3 * j + i] = v;
    }
    double es[9] = {0};
#pragma unroll
    for (int a = 0; a < SM; a++) {
      if (a >= S) continue;
      const double x = xi.re((long)a * P + pt);
#pragma unroll
      for (int kl = 0; kl < 9; kl++) es[kl] += ec->eps[a][kl] * x;
    }
#pragma unroll
    for (int ij = 0; ij < 9; ij++) {
      double v = 0.0;
#pragma unroll
      for (int kl = 0; kl < 9; kl++) v += ec->C[ij * 9 + kl] * (e[kl] - es[kl]);
      sg[ij] = v;
    }
    if (eps_field)
      for (int ij = 0; ij < 9; ij++) eps_field[(long)ij * P + pt] = e[ij];
    if (sig_field)
      for (int ij = 0; ij < 9; ij++) sig_field[(long)ij * P + pt] = sg[ij];
  }
  for (int c = 0; c < 18; c++) {
    const double s = block_sum(c < 9 ? e[c] : sg[c - 9], red);
    if (threadIdx.x == 0) partials[(long)c * nblk + blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------- persistent ensemble step
// One ordinary launch runs a whole GL step of every class of an ensemble (of
// one precision).  Blocks draw items from one work counter: first the plane
// items (plane_body) of every class, larger grids first; then the gamma
// finish of each class; then the fused column items (col_contract_body).  An
// item of the second kind waits until every plane item of its class has
// completed (a per-class completion counter): since all plane items are drawn
// before any other item, the blocks it waits for are already running and
// never wait themselves, so no grid-wide barrier or co-residency guarantee is
// needed.  Each wait is bounded; a wait that runs out records an error and
// the item is skipped (the host reports it).  The items run exactly the code
// of the plane / fused column / finish kernels, so a step gives the same bits
// as those kernels launched per class.  Supported classes: cubic grids N in
// {32, 48, 64, 96}, S <= 4, recomputed B, no frozen points, uniform shear.
constexpr int ENS_THREADS = 1024, ENS_MAX_CLASSES = 8;
// q: [0] work counter, [1] error flag (sticky: set by a wait that ran out,
// cleared by the host), [2 + k] completed plane items of class k
constexpr int ENS_Q_INTS = 2 + ENS_MAX_CLASSES;
// spins of at most ~64 clocks each before a wait gives up (~2 s at 2.4 GHz)
constexpr long ENS_SPIN_LIMIT = 1L << 26;
struct EnsClassDev {
  void* W;                 // cx<T> [B*S][P]
  void* xi;                // the master (Master<T> over B*S*P points)
  const ElasticConsts* ec;
  const void* tw[3];       // cx<T> twiddles of the N1, N2, N3 axes
  StepArgs sa;
  double* partials;
  double* hist;
  int* step_ctr;
  long P;
  int n, F, B, S, c3_tile;
  int lean;                // master in the lean FP32 form
  long plane_first, col_first;  // first item of the class among the plane / column items
  long n_planes;           // plane items of the class (F * n)
};
struct EnsDev {
  int* q;                  // ENS_Q_INTS counters; all but the error flag are zeroed before every launch
  int nc;
  long n_plane, n_items;   // plane items; all items (planes, nc finishes, column items)
  EnsClassDev cls[ENS_MAX_CLASSES];
};

__global__ void k_ens_reset(int* q) {
  if (threadIdx.x < ENS_Q_INTS && threadIdx.x != 1) q[threadIdx.x] = 0;
}

// (Bodies are not inlined into the persistent kernel, so its registers are
// those of the largest body rather than of all of them together.)
template <typename T, int N, typename M>
__device__ __forceinline__ void ens_plane_m(const EnsClassDev* __restrict__ c, long local, cx<T>* sm, M xi) {
  plane_body<T, N, N, false, false, M>(reinterpret_cast<cx<T>*>(c->W), xi, nullptr, nullptr, c->sa,
                                       reinterpret_cast<const cx<T>*>(c->tw[1]),
                                       reinterpret_cast<const cx<T>*>(c->tw[2]), c->partials, c->step_ctr, local, sm);
}
template <typename T, int N>
__device__ __noinline__ void ens_plane(const EnsClassDev* __restrict__ c, long local, cx<T>* sm) {
  const Master<T> m = make_master<T>(c->xi, (long)c->F * c->P, c->lean != 0);
  if constexpr (sizeof(T) == 8) {
    ens_plane_m<T, N>(c, local, sm, m);
  } else if (m.im) {
    ens_plane_m<T, N>(c, local, sm, MasterF32C{m.r, m.im});
  } else {
    ens_plane_m<T, N>(c, local, sm, MasterF32L{m.r});
  }
}
template <typename T, int N>
__device__ __noinline__ void ens_cols(const EnsClassDev* __restrict__ c, long local, cx<T>* sm) {
  const long tiles = (long)N * N / c->c3_tile;
  col_contract_body<T, N, 4>(reinterpret_cast<cx<T>*>(c->W), c->ec, c->P, c->c3_tile, 0,
                             reinterpret_cast<const cx<T>*>(c->tw[0]), local % tiles, (int)(local / tiles), sm);
}
__device__ __noinline__ void ens_finish(const EnsClassDev* __restrict__ c, double* fsum) {
  finish_body(c->partials, c->F, c->n, c->S, c->hist, c->step_ctr, fsum);
}

template <typename T>
__global__ void __launch_bounds__(ENS_THREADS) k_ens_step(const EnsDev* __restrict__ d) {
  extern __shared__ unsigned char lds_raw[];
  cx<T>* sm = reinterpret_cast<cx<T>*>(lds_raw);
  __shared__ int slot;
  __shared__ int ready;        // bit k: this block has seen class k's planes complete
  __shared__ int go;
  int* const q = d->q;
  if (threadIdx.x == 0) ready = 0;
  for (;;) {
    __syncthreads();  // the previous item is done with LDS and `slot`
    if (threadIdx.x == 0) slot = atomicAdd(&q[0], 1);
    __syncthreads();
    const long it = slot;
    if (it >= d->n_items) break;
    if (it < d->n_plane) {
      int k = 0;
      while (k + 1 < d->nc && it >= d->cls[k + 1].plane_first) k++;
      const EnsClassDev* c = &d->cls[k];
      const long local = it - c->plane_first;
      switch (c->n) {
        case 32: ens_plane<T, 32>(c, local, sm); break;
        case 48: ens_plane<T, 48>(c, local, sm); break;
        case 64: ens_plane<T, 64>(c, local, sm); break;
        case 96: ens_plane<T, 96>(c, local, sm); break;
        default: break;
      }
      __syncthreads();  // every wave's stores of the plane are done
      if (threadIdx.x == 0) {
        __threadfence();  // ... and visible device-wide before the count
        atomicAdd(&q[2 + k], 1);
      }
      continue;
    }
    // a finish or column item of class k: wait for the class's plane items
    const long r = it - d->n_plane;
    int k = 0;
    if (r < d->nc) {
      k = (int)r;
    } else {
      while (k + 1 < d->nc && r >= d->cls[k + 1].col_first + d->nc) k++;
    }
    const EnsClassDev* c = &d->cls[k];
    if (threadIdx.x == 0) {
      int ok = 1;
      if (!(ready & (1 << k))) {
        long spins = 0;
        while (__hip_atomic_load(&q[2 + k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < (int)c->n_planes) {
          if (++spins > ENS_SPIN_LIMIT) {
            ok = 0;
            atomicOr(&q[1], 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
        if (ok) ready |= 1 << k;
      }
      go = ok;
    }
    __syncthreads();
    if (!go) continue;
    if (r < d->nc) {
      ens_finish(c, reinterpret_cast<double*>(lds_raw));
      continue;
    }
    const long local = r - d->nc - c->col_first;
    switch (c->n) {
      case 32: ens_cols<T, 32>(c, local, sm); break;
      case 48: ens_cols<T, 48>(c, local, sm); break;
      case 64: ens_cols<T, 64>(c, local, sm); break;
      case 96: ens_cols<T, 96>(c, local, sm); break;
      default: break;
    }
  }
}

}  // namespace pfdd

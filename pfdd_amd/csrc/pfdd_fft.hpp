// pfdd_fft.hpp -- device building blocks of the batched 3D C2C FFT.
//
// A 1D transform of length N runs on pencils staged in LDS: each pass of the
// Stockham autosort algorithm (Govindaraju et al., SC08, "High performance
// discrete Fourier transforms on graphics processors") reads R points per
// butterfly from LDS into registers, applies the twiddles and a radix-R
// butterfly, and writes the R results back in natural order.  Radices are
// 8, 4, 2 and 3; TPP threads share one pencil.  Unnormalised; sign -1 is the
// forward transform exp(-2 pi i k x / N) (G.c:209-273).
#pragma once
#include <hip/hip_runtime.h>

namespace pfdd {

template <typename T> struct V2;
template <> struct V2<double> { using t = double2; };
template <> struct V2<float> { using t = float2; };
template <typename T> using cx = typename V2<T>::t;

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  cx<T> r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
// (s i) * a for s = +-1
template <typename T> __device__ __forceinline__ cx<T> mul_si(cx<T> a, T s) { return {-s * a.y, s * a.x}; }

// HBM accesses of the passes that read or write W directly (first and last
// pass).  In FP64 the stores are non-temporal, and so are the loads of the
// contiguous rows: W (403 MB at 128^3 x 12) is larger than the 256 MB
// last-level cache and is streamed; measured 0.99 vs 1.01 ms per step
// (non-temporal loads of the strided columns too: 1.02 ms).  The FP32 W
// (201 MB there) fits in that cache once the master's accesses bypass it, and
// ordinary accesses let the next pass find it there: 0.509 vs 0.533 ms per
// mixed step.
template <typename T> struct VecOf;
template <> struct VecOf<double> { typedef double t __attribute__((ext_vector_type(2))); };
template <> struct VecOf<float> { typedef float t __attribute__((ext_vector_type(2))); };
template <typename T, bool NT> __device__ __forceinline__ cx<T> gload(const cx<T>* p) {
  if constexpr (NT && sizeof(T) == 8) {
    const typename VecOf<T>::t v = __builtin_nontemporal_load(reinterpret_cast<const typename VecOf<T>::t*>(p));
    return {v.x, v.y};
  } else {
    return *p;
  }
}
template <typename T> __device__ __forceinline__ void gstore(cx<T>* p, cx<T> v) {
  if constexpr (sizeof(T) == 8) {
    typename VecOf<T>::t w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<typename VecOf<T>::t*>(p));
  } else {
    *p = v;
  }
}

// ---------------------------------------------------------------- plans
struct Plan {
  int np;      // number of passes
  int r[12];   // radix of each pass
  int tpp;     // threads per pencil = N / max radix
};

// Radix sequence of an n-point transform: with rlim = 8, radix-8 passes
// first, then at most one radix-4 and one radix-2 pass, then radix-3 passes;
// with rlim = 4 the power of two is taken in radix-4 passes (and at most
// one radix-2 pass): more passes, but twice the threads per pencil and half
// the registers per thread.
constexpr Plan make_plan(int n, int rlim = 8) {
  Plan p{};
  int m = n, rmax = 1;
  while (rlim >= 8 && m % 8 == 0) { p.r[p.np++] = 8; m /= 8; }
  if (rlim >= 8) {
    if (m % 4 == 0) { p.r[p.np++] = 4; m /= 4; }
  } else {
    while (m % 4 == 0) { p.r[p.np++] = 4; m /= 4; }
  }
  if (m % 2 == 0) { p.r[p.np++] = 2; m /= 2; }
  while (m % 3 == 0) { p.r[p.np++] = 3; m /= 3; }
  for (int i = 0; i < p.np; i++) rmax = p.r[i] > rmax ? p.r[i] : rmax;
  p.tpp = n / rmax;
  return p;
}
// whether the radix-4 sequence of n differs from the default one
constexpr bool has_r4_variant(int n) { return make_plan(n, 4).np != make_plan(n, 8).np; }
// product of the radices before pass k (the Stockham "Ns")
constexpr int ns_before(int n, int k, int rlim = 8) {
  Plan p = make_plan(n, rlim);
  int s = 1;
  for (int i = 0; i < k; i++) s *= p.r[i];
  return s;
}
constexpr bool supported_len(int n) {
  if (n < 2) return false;
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  return m == 1;
}

// ---------------------------------------------------------------- butterflies
template <typename T, int R> struct Bfly;

template <typename T> struct Bfly<T, 2> {
  __device__ __forceinline__ static void run(cx<T>* v, T) {
    cx<T> a = v[0];
    v[0] = cadd<T>(a, v[1]);
    v[1] = csub<T>(a, v[1]);
  }
};

template <typename T> struct Bfly<T, 4> {
  __device__ __forceinline__ static void run(cx<T>* v, T s) {
    cx<T> t0 = cadd<T>(v[0], v[2]), t1 = csub<T>(v[0], v[2]);
    cx<T> t2 = cadd<T>(v[1], v[3]), t3 = mul_si<T>(csub<T>(v[1], v[3]), s);
    v[0] = cadd<T>(t0, t2);
    v[2] = csub<T>(t0, t2);
    v[1] = cadd<T>(t1, t3);
    v[3] = csub<T>(t1, t3);
  }
};

template <typename T> struct Bfly<T, 3> {
  __device__ __forceinline__ static void run(cx<T>* v, T s) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    cx<T> t = cadd<T>(v[1], v[2]), u = csub<T>(v[1], v[2]);
    cx<T> m = {v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y};
    cx<T> hu = mul_si<T>({h * u.x, h * u.y}, s);
    v[0] = cadd<T>(v[0], t);
    v[1] = cadd<T>(m, hu);
    v[2] = csub<T>(m, hu);
  }
};

// 8 = 2 x 4: two radix-4 on the even and odd samples, then w8 twiddles.
template <typename T> struct Bfly<T, 8> {
  __device__ __forceinline__ static void run(cx<T>* v, T s) {
    const T c = (T)0.70710678118654752440;  // sqrt(1/2)
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<T, 4>::run(e, s);
    Bfly<T, 4>::run(o, s);
    // o[q] *= w8^(s q): q=1: c(1 + s i), q=2: s i, q=3: c(-1 + s i)
    cx<T> o1 = {c * (o[1].x - s * o[1].y), c * (o[1].y + s * o[1].x)};
    cx<T> o2 = mul_si<T>(o[2], s);
    cx<T> o3 = {c * (-o[3].x - s * o[3].y), c * (-o[3].y + s * o[3].x)};
    v[0] = cadd<T>(e[0], o[0]);
    v[4] = csub<T>(e[0], o[0]);
    v[1] = cadd<T>(e[1], o1);
    v[5] = csub<T>(e[1], o1);
    v[2] = cadd<T>(e[2], o2);
    v[6] = csub<T>(e[2], o2);
    v[3] = cadd<T>(e[3], o3);
    v[7] = csub<T>(e[3], o3);
  }
};

// ---------------------------------------------------------------- pencil access
// A transform reads its input and writes its output through accessors, so
// the first Stockham pass can load straight from HBM into registers and the
// last pass can store straight back; only the passes in between go through
// LDS.  Element i of pencil p:
// Position i of a pencil sits at i + (i >> SH) when SH > 0: one pad slot per
// 2^SH points, so that the first radix-8 pass's writes (stride 8 within a
// pencil) fall on distinct LDS banks.
template <int SH> __device__ __forceinline__ int lds_pos(int i) { return SH ? i + (i >> SH) : i; }
template <int N, int SH> constexpr int padded_pitch() { return SH ? N + (N >> SH) + 1 : N + 1; }

template <typename T, int SH = 0> struct LdsPencils {     // sm[p * ld + lds_pos<SH>(i)]
  cx<T>* sm;
  int ld;
  static constexpr bool in_lds = true;
  __device__ __forceinline__ cx<T> load(int p, int i) const { return sm[p * ld + lds_pos<SH>(i)]; }
  __device__ __forceinline__ void store(int p, int i, cx<T> v) const { sm[p * ld + lds_pos<SH>(i)] = v; }
};
template <typename T> struct RowPencils {      // contiguous rows: base[p * n + i]
  cx<T>* base;
  int n;
  static constexpr bool in_lds = false;
  __device__ __forceinline__ cx<T> load(int p, int i) const { return gload<T, true>(base + (long)p * n + i); }
  __device__ __forceinline__ void store(int p, int i, cx<T> v) const { gstore<T>(base + (long)p * n + i, v); }
};
template <typename T> struct ColPencils {      // strided columns: base[i * inner + p]
  cx<T>* base;
  long inner;
  static constexpr bool in_lds = false;
  __device__ __forceinline__ cx<T> load(int p, int i) const { return gload<T, false>(base + (long)i * inner + p); }
  __device__ __forceinline__ void store(int p, int i, cx<T> v) const { gstore<T>(base + (long)i * inner + p, v); }
};

template <typename T> struct LdsColPencils {  // columns of a row-major LDS plane: sm[i * ld + p]
  cx<T>* sm;
  int ld;
  static constexpr bool in_lds = true;
  __device__ __forceinline__ cx<T> load(int p, int i) const { return sm[i * ld + p]; }
  __device__ __forceinline__ void store(int p, int i, cx<T> v) const { sm[i * ld + p] = v; }
};

// ---------------------------------------------------------------- passes
// One Stockham pass on G pencils per thread, reading through `in` and
// writing through `out`.  Pencil g of this thread is p0 + g * pstride and is
// processed when < npen; q in [0, TPP) is the thread's lane within the
// pencil.  tw[m] = exp(+2 pi i m / N); the sign s flips its imaginary part.
// Every thread of the block calls it: there is a barrier between the read
// and the write phase, and one after the write phase when out is in LDS.
// `hook` runs after the read phase, before the barrier (the kernels that
// read the first pass from HBM store their twiddle table there; see TwStage).
struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};
template <typename T, int N, int R, int NSB, int TPP, int G, typename In, typename Out, typename Hook = NoHook>
__device__ __forceinline__ void stockham_pass(const In& in, const Out& out, int p0, int pstride, int npen, int q,
                                              T s, const cx<T>* __restrict__ tw, const Hook& hook = Hook{}) {
  constexpr int M = N / R;
  constexpr int J = (M + TPP - 1) / TPP;
  constexpr int STEP = N / (NSB * R);
  cx<T> v[G][J][R];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int p = p0 + g * pstride;
    if (p >= npen) continue;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const int j = q + jj * TPP;
      if ((M % TPP == 0) || j < M) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[g][jj][r] = in.load(p, j + r * M);
        if (NSB > 1) {
          const int k = j % NSB;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            cx<T> w = tw[k * r * STEP];
            w.y *= s;
            v[g][jj][r] = cmul<T>(v[g][jj][r], w);
          }
        }
        Bfly<T, R>::run(v[g][jj], s);
      }
    }
  }
  hook();
  if (In::in_lds || Out::in_lds) __syncthreads();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int p = p0 + g * pstride;
    if (p >= npen) continue;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const int j = q + jj * TPP;
      if ((M % TPP == 0) || j < M) {
        const int d = (j / NSB) * NSB * R + (j % NSB);
#pragma unroll
        for (int r = 0; r < R; ++r) out.store(p, d + r * NSB, v[g][jj][r]);
      }
    }
  }
  if (Out::in_lds) __syncthreads();
}

template <typename T, int N, int G, int K, int RL, typename In, typename Mid, typename Out, typename Hook = NoHook>
__device__ __forceinline__ void fft_from(const In& in, const Mid& lds, const Out& out, int p0, int pstride,
                                         int npen, int q, T s, const cx<T>* __restrict__ tw,
                                         const Hook& hook = Hook{}) {
  constexpr Plan pl = make_plan(N, RL);
  if constexpr (K < pl.np) {
    constexpr int R = pl.r[K], NSB = ns_before(N, K, RL), TPP = pl.tpp;
    if constexpr (pl.np == 1) {
      stockham_pass<T, N, R, NSB, TPP, G>(in, out, p0, pstride, npen, q, s, tw, hook);
    } else if constexpr (K == 0) {
      stockham_pass<T, N, R, NSB, TPP, G>(in, lds, p0, pstride, npen, q, s, tw, hook);
    } else if constexpr (K == pl.np - 1) {
      stockham_pass<T, N, R, NSB, TPP, G>(lds, out, p0, pstride, npen, q, s, tw);
    } else {
      stockham_pass<T, N, R, NSB, TPP, G>(lds, lds, p0, pstride, npen, q, s, tw);
    }
    fft_from<T, N, G, K + 1, RL>(in, lds, out, p0, pstride, npen, q, s, tw);
  }
}

// N-point transforms of up to G pencils per thread: the first pass reads
// through `in`, the last writes through `out`, the passes in between use
// `lds` (which needs npen pencils of pitch >= N).  When the input is already
// in LDS the caller synchronises first; the function ends synchronised when
// the output is in LDS.
template <typename T, int N, int G, int RL = 8, typename In, typename Mid, typename Out, typename Hook = NoHook>
__device__ __forceinline__ void fft_pencils(const In& in, const Mid& lds, const Out& out, int p0,
                                            int pstride, int npen, int q, T s, const cx<T>* __restrict__ tw,
                                            const Hook& hook = Hook{}) {
  fft_from<T, N, G, 0, RL>(in, lds, out, p0, pstride, npen, q, s, tw, hook);
}

// All-LDS transforms of up to G pencils per thread (pencil p at sm + p*LD,
// position i at lds_pos<SH>(i)).
template <typename T, int N, int G, int LD, int SH = 0>
__device__ __forceinline__ void fft_lds_multi(cx<T>* sm, int p0, int pstride, int npen, int q, T s,
                                              const cx<T>* __restrict__ tw) {
  const LdsPencils<T, SH> l{sm, LD};
  fft_pencils<T, N, G>(l, l, l, p0, pstride, npen, q, s, tw);
}

// The 128-point transform of the compile-time column kernel (plan 8, 8, 2;
// 16 lanes per pencil, one radix-8 butterfly per lane and pass), with its
// last two passes fused through a wave shuffle instead of an LDS round
// trip.  After the second pass lane j of a pencil holds the outputs for
// positions (j / 8) * 64 + j % 8 + 8 r (r < 8), and the radix-2 pass pairs
// position p < 64 with p + 64: those two sit in lanes j and j ^ 8.  One
// swap of four values with the partner lane (__shfl_xor over 16 lanes)
// gives each lane four complete pairs -- the lower lane (j < 8) the pairs
// p = j + 8 r for even r, the upper one for odd r -- whose butterflies it
// finishes and writes to LDS.  This drops the second pass's LDS stores, the
// third pass's LDS loads and one barrier.  The first pass is the ordinary
// one.  Splitting the pairs by the parity of r makes the 16 lanes of a
// pencil store 16 consecutive positions at a time.
//
// The twiddles come from a 120-entry table in the order the lanes read them
// (Shfl128Tw), so that the 8 distinct factors a pass reads at once sit in 8
// consecutive entries: entry (r - 1) * 8 + k = w^(2 k r) for the second
// pass (k = j % 8, r = 1..7), entry 56 + m = w^m for the radix-2 pass
// (m < 64), w = exp(+2 pi i / 128); the same factors as the plain table, so
// the results are the same bits.  Ends synchronised (the result is in LDS).
template <typename T, int LD, int SH>
__device__ __forceinline__ void fft128_lds_shfl(cx<T>* sm, int p, int npen, int q, T s, const cx<T>* __restrict__ tw) {
  constexpr int N = 128;
  static_assert(make_plan(N).np == 3 && make_plan(N).r[0] == 8 && make_plan(N).r[1] == 8 &&
                    make_plan(N).r[2] == 2 && make_plan(N).tpp == 16, "plan 8, 8, 2 with 16 lanes");
  const LdsPencils<T, SH> l{sm, LD};
  stockham_pass<T, N, 8, 1, 16, 1>(l, l, p, 1 << 30, npen, q, s, tw);  // no twiddles; ends with a barrier
  cx<T> v[8];
  const bool on = p < npen;
  const int j8 = q % 8;
  if (on) {
    // second pass: NSB = 8, M = 16 butterflies (one per lane), twiddles w^(2 (j % 8) r)
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = l.load(p, q + r * 16);
#pragma unroll
    for (int r = 1; r < 8; r++) {
      cx<T> w = tw[(r - 1) * 8 + j8];
      w.y *= s;
      v[r] = cmul<T>(v[r], w);
    }
    Bfly<T, 8>::run(v, s);
  }
  // exchange with lane j ^ 8: the lower lane keeps v[2i] (position j + 16 i)
  // and takes the partner's v[2i] (position 64 + j + 16 i); the upper lane
  // keeps v[2i + 1] (position 64 + j % 8 + 8 (2i + 1)) and takes the
  // partner's v[2i + 1]
  const bool lower = q < 8;
  cx<T> a[4], b[4];  // pair (position pp, pp + 64): a at pp, b at pp + 64
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const cx<T> send = lower ? v[2 * i + 1] : v[2 * i];
    cx<T> got;
    got.x = __shfl_xor(send.x, 8, 16);
    got.y = __shfl_xor(send.y, 8, 16);
    a[i] = lower ? v[2 * i] : got;
    b[i] = lower ? got : v[2 * i + 1];
  }
  __syncthreads();  // every lane's second-pass inputs are read before the stores below
  if (on) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int pp = j8 + 8 * (2 * i + (lower ? 0 : 1));  // third pass: NSB = 64, twiddle w^pp
      cx<T> w = tw[56 + pp];
      w.y *= s;
      cx<T> pr[2] = {a[i], cmul<T>(b[i], w)};
      Bfly<T, 2>::run(pr, s);
      l.store(p, pp, pr[0]);
      l.store(p, pp + 64, pr[1]);
    }
  }
  __syncthreads();
}

// The twiddle table of fft128_lds_shfl, staged into LDS like TwStage: entry
// e < 56 is tw[2 (e % 8) (e / 8 + 1)], entry 56 + m is tw[m], from the
// plain 128-entry table tw[m] = w^m.
template <typename T, int NT> struct Shfl128Tw {
  static constexpr int E = 120, K = (E + NT - 1) / NT;
  cx<T> v[K];
  __device__ __forceinline__ void load(const cx<T>* __restrict__ tw) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = (int)threadIdx.x + k * NT < E ? (int)threadIdx.x + k * NT : E - 1;  // (as in TwStage)
      v[k] = tw[e < 56 ? 2 * (e % 8) * (e / 8 + 1) : e - 56];
    }
  }
  __device__ __forceinline__ void store(cx<T>* twl) const {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = (int)threadIdx.x + k * NT;
      if (e < E) twl[e] = v[k];
    }
  }
};

// One pencil per TPP-thread group: pencil p = threadIdx.x / TPP at sm + p*LD.
template <typename T, int N, int LD>
__device__ __forceinline__ void fft_lds(cx<T>* sm, int npen, T s, const cx<T>* __restrict__ tw) {
  constexpr int TPP = make_plan(N).tpp;
  fft_lds_multi<T, N, 1, LD>(sm, (int)threadIdx.x / TPP, 1 << 30, npen, (int)threadIdx.x % TPP, s, tw);
}

// The kernels copy the N-entry twiddle table of their axis into LDS at the
// start (twl), so that the passes read it there rather than from global
// memory (at 128^3 x 12 the fused column kernel went from 0.217 to 0.189 ms
// and the interleaved FP32 row kernel from 0.196 to 0.174 ms).  The first
// barrier of the transform precedes the first pass that reads it.
template <typename T, int N>
__device__ __forceinline__ void stage_twiddles(cx<T>* twl, const cx<T>* __restrict__ tw) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) twl[k] = tw[k];
}

// The same copy split in two, for kernels that load a tile into LDS at the
// start: load() issues this thread's global loads of the table, store()
// writes them to LDS.  Issued after the tile's loads and stored with them,
// the table's latency overlaps the tile's instead of preceding it (the
// store of stage_twiddles waits for its load, and the in-order load counter
// makes that wait cover every load issued before it).
template <typename T, int N, int NT> struct TwStage {
  static constexpr int K = (N + NT - 1) / NT;
  cx<T> v[K];
  __device__ __forceinline__ void load(const cx<T>* __restrict__ tw) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = (int)threadIdx.x + k * NT;
      // every thread loads (an entry in range): loads under a branch were
      // kept in scratch memory, each waited for before the next was issued
      v[k] = tw[N % NT == 0 || i < N ? i : N - 1];
    }
  }
  __device__ __forceinline__ void store(cx<T>* twl) const {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = (int)threadIdx.x + k * NT;
      if (N % NT == 0 || i < N) twl[i] = v[k];
    }
  }
};

// LDS row pitch (in complex elements) for a pencil of N points.
template <int N> constexpr int pitch() { return N + 1; }

}  // namespace pfdd

// pfdd_api.hip -- the C ABI of libpfdd_gpu.so (include/pfdd_gpu.h).
//
// One pfdd_ctx owns its device memory and one HIP stream (default priority).
// A GL step is: k_update (plasticevolv + copy to the FFT array + gamma sums),
// forward 3D FFT, contraction with B(k), inverse 3D FFT.  Chunks of steps
// are captured once into a hipGraph and replayed.  A context may hold B
// independent instances of one grid (the classes of an ensemble): every
// kernel then runs over the B*S fields, and only the contraction couples the
// S fields of one instance.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <functional>
#include <type_traits>
#include <new>
#include <vector>

#include "pfdd_gpu.h"
#include "pfdd_kernels.hpp"

using namespace pfdd;

#define HIPCK(x)                                   \
  do {                                             \
    hipError_t e_ = (x);                           \
    if (e_ != hipSuccess) return PFDD_E_HIP;       \
  } while (0)

namespace {

constexpr int GRAPH_STEPS = 8;      // GL steps per captured graph
constexpr int HIST_CAP = 1024;      // gamma history rows per chunk

struct Twiddles {
  int n = 0;
  double2* d = nullptr;
  float2* f = nullptr;
};

}  // namespace

struct pfdd_ctx {
  pfdd_params p{};
  long P = 0;
  int S = 0;
  int B = 1;                 // instances of this grid
  int F = 0;                 // fields, B * S
  int ntri = 0;
  int fp64 = 1;
  int lean = 0;              // PFDD_MIXED_LEAN: plain FP32 master
  int ref32 = 0;             // PFDD_REF32: FP32 work array, FP64 master, the reference's float arithmetic
  int stored = 0;
  hipStream_t stream = nullptr;
  int owns_stream = 1;       // 0: the stream belongs to an ensemble
  // device buffers
  void* xi = nullptr;         // the master: Master<T> over F*P points (16 or 12 bytes each)
  size_t xi_bytes = 0;
  void* W = nullptr;         // cx<T> [S][P]
  void* Bsym = nullptr;      // T [ntri][P]
  uint8_t* mask = nullptr;
  double* taupp = nullptr;   // [B][P][S], when some instance has a per-point shear
  double2* taud = nullptr;   // [F] (tau, ramp rate)
  double* partials = nullptr;
  size_t partials_len = 0;
  double* hist = nullptr;    // [HIST_CAP][B][S+1]
  int* step_ctr = nullptr;
  double* small = nullptr;   // reduction outputs (<= 32 doubles)
  double* avepst_d = nullptr;
  ElasticConsts* ec = nullptr;
  double* freq = nullptr;    // frec tables of the three axes, n1 + n2 + n3 doubles
  void* E6 = nullptr;        // strain spectrum [6][P], allocated on first use
  double* efield = nullptr;  // [9][P] strain/stress fields, on demand
  double* sfield = nullptr;
  long* line_pts = nullptr;  // interface-line point list (pfdd_set_line_points)
  int nline = 0;
  double* line_out = nullptr; // [HIST_CAP][F] line sums per step; row 0 also serves pfdd_line_sums
  Twiddles tw[3];
  float2* fourn_tw[3][2] = {};  // PFDD_REF32: fourn's twiddles per axis, [0] forward, [1] backward
  StepArgs sa{};
  int upd_nblk = 0;
  int row_tile = 0;          // rows per block of the fused row kernel (0: not fused)
  int row_nblk = 0;          // its blocks per field
  int c3_tile = 0;           // columns per block of the fused column kernel (0: not fused)
  int plane = 0;             // 1: the LDS-resident plane kernel does the N2/N3 passes and the update
  int il = 0;                // 1: W interleaved, [B][P][S] (fields of a point together); else [B][S][P]
  int il_fixed = 0;          // 1: the interleaved column kernel with compile-time S and tile (k_col_il_fixed)
  int il_fixed_row = 0;      // 1: the interleaved row kernel with compile-time S and rows (k_row_il_fixed)
  size_t c3_lds = 0;         // its dynamic LDS bytes
  int plan = 0;              // the applied PFDD_PLAN_* word
  int r4 = 0;                // bit ax: radix-4 sequence for the standalone passes along axis ax
  int has_xi = 0;
  int prologue_done = 0;
  hipGraphExec_t graph = nullptr;
  long long bytes = 0;
  // kernel timing: when set, an event is recorded after every launch of a
  // step, tagged with its stage (pfdd_time_step_kernels)
  std::vector<std::pair<int, hipEvent_t>>* tap = nullptr;
  std::vector<hipEvent_t>* tap_pool = nullptr;
  // per instance
  std::vector<double2> tauh; // host copy of taud
  std::vector<uint8_t> inst_pp, inst_frozen, inst_up;
};

namespace {

constexpr int kLens[] = {2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};

bool len_ok(int n) {
  for (int v : kLens)
    if (v == n) return true;
  return false;
}

// Records the timing event of a launch of `stage` (see pfdd_ctx::tap).
void tap(pfdd_ctx* c, int stage) {
  if (!c->tap || c->tap->size() >= c->tap_pool->size()) return;
  hipEvent_t e = (*c->tap_pool)[c->tap->size()];
  (void)hipEventRecord(e, c->stream);
  c->tap->push_back({stage, e});
}

template <typename T> const cx<T>* tw_of(const Twiddles& t);
template <> const double2* tw_of<double>(const Twiddles& t) { return t.d; }
template <> const float2* tw_of<float>(const Twiddles& t) { return t.f; }

template <typename T> Master<T> xm(const pfdd_ctx* c) { return make_master<T>(c->xi, (long)c->F * c->P, c->lean != 0); }
template <> Master<float> xm<float>(const pfdd_ctx* c) {
  if (c->ref32) return Master<float>{nullptr, nullptr, (double2*)c->xi};
  return make_master<float>(c->xi, (long)c->F * c->P, c->lean != 0);
}
// whether the master is the double2 array of the reference layout (FP64 and PFDD_REF32)
bool master64(const pfdd_ctx* c) { return c->fp64 || c->ref32; }
// Calls f with the context's master as the static form the per-step kernels
// are instantiated for (Master<double>, MasterF32C or MasterF32L).
template <typename T, typename Fn> void with_master(const pfdd_ctx* c, Fn&& f) {
  const Master<T> m = xm<T>(c);
  if constexpr (sizeof(T) == 8) {
    f(m);
  } else if (m.im) {
    f(MasterF32C{m.r, m.im});
  } else {
    f(MasterF32L{m.r});
  }
}

// radix limit of the standalone passes along axis ax (0: N1, 1: N2, 2: N3)
int rl_of(const pfdd_ctx* c, int ax) { return (c->r4 >> ax) & 1 ? 4 : 8; }

template <typename T, int N, int RL>
void launch_axis_n(cx<T>* data, long outer, long inner, int sign, const cx<T>* tw, hipStream_t st) {
  constexpr int TPP = make_plan(N, RL).tpp;
  if (inner == 1) {
    constexpr int TILE = tile_rows<T, N, RL>();
    const long nrows = outer;
    const long nb = (nrows + TILE - 1) / TILE;
    hipLaunchKernelGGL((k_fft_rows<T, N, RL>), dim3((unsigned)nb), dim3(TILE * TPP), 0, st, data, nrows, (T)sign, tw);
  } else {
    // rows at least 1 MB apart (the N1 pass of a large grid) take the wide tile where it differs
    auto go = [&](auto wide_tag) {
      constexpr bool WIDE = decltype(wide_tag)::value;
      constexpr int TILE = tile_cols<T, N, RL, WIDE>();
      int tile = 1;
      while (tile * 2 <= TILE && inner % (tile * 2) == 0) tile *= 2;
      const long nb = outer * (inner / tile);
      hipLaunchKernelGGL((k_fft_cols<T, N, RL, WIDE>), dim3((unsigned)nb), dim3(cols_threads<T, N, RL, WIDE>()),
                         (cols_lds_bytes<T, N, RL, WIDE>()), st, data, inner, tile, (T)sign, tw);
    };
    if constexpr (has_wide_cols<T, N, RL>()) {
      if (inner * (long)sizeof(cx<T>) >= (1L << 20)) {
        go(std::true_type{});
        return;
      }
    }
    go(std::false_type{});
  }
}

// rl = 4 takes the radix-4 sequence where the length has one
template <typename T, int N>
void launch_axis_r(cx<T>* data, long outer, long inner, int sign, const cx<T>* tw, hipStream_t st, int rl) {
  if constexpr (has_r4_variant(N)) {
    if (rl == 4) {
      launch_axis_n<T, N, 4>(data, outer, inner, sign, tw, st);
      return;
    }
  }
  launch_axis_n<T, N, 8>(data, outer, inner, sign, tw, st);
}

// The column kernels of length N may take more than 64 KB of dynamic LDS;
// the attribute is set once per process for every instantiation (in
// ctx_create, outside any graph capture).
template <typename T, int N>
int cols_attr_n() {
  HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_cols<T, N, 8>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)cols_lds_bytes<T, N, 8>()));
  if constexpr (has_wide_cols<T, N, 8>())
    HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_cols<T, N, 8, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)cols_lds_bytes<T, N, 8, true>()));
  if constexpr (has_r4_variant(N)) {
    HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_cols<T, N, 4>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)cols_lds_bytes<T, N, 4>()));
    if constexpr (has_wide_cols<T, N, 4>())
      HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_cols<T, N, 4, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)cols_lds_bytes<T, N, 4, true>()));
  }
  return PFDD_OK;
}

int cols_attr(int n) {
  int rc = PFDD_OK;
  switch (n) {
#define PFDD_LEN(NN) case NN: if (!(rc = cols_attr_n<double, NN>())) rc = cols_attr_n<float, NN>(); break;
    PFDD_LEN(2) PFDD_LEN(3) PFDD_LEN(4) PFDD_LEN(6) PFDD_LEN(8) PFDD_LEN(12) PFDD_LEN(16) PFDD_LEN(24)
    PFDD_LEN(32) PFDD_LEN(48) PFDD_LEN(64) PFDD_LEN(96) PFDD_LEN(128) PFDD_LEN(192) PFDD_LEN(256)
    PFDD_LEN(384) PFDD_LEN(512)
#undef PFDD_LEN
    default: return PFDD_E_RADIX;
  }
  return rc;
}

// Transform along one axis of an [outer][n][inner] complex array.
template <typename T>
int launch_axis(cx<T>* data, long outer, int n, long inner, int sign, const cx<T>* tw, hipStream_t st, int rl = 8) {
  switch (n) {
#define PFDD_LEN(NN) case NN: launch_axis_r<T, NN>(data, outer, inner, sign, tw, st, rl); break;
    PFDD_LEN(2) PFDD_LEN(3) PFDD_LEN(4) PFDD_LEN(6) PFDD_LEN(8) PFDD_LEN(12) PFDD_LEN(16) PFDD_LEN(24)
    PFDD_LEN(32) PFDD_LEN(48) PFDD_LEN(64) PFDD_LEN(96) PFDD_LEN(128) PFDD_LEN(192) PFDD_LEN(256)
    PFDD_LEN(384) PFDD_LEN(512)
#undef PFDD_LEN
    default: return PFDD_E_RADIX;
  }
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

template <typename T>
int fft3d(pfdd_ctx* c, cx<T>* data, int batch, int sign) {
  const int n1 = c->p.n1, n2 = c->p.n2, n3 = c->p.n3;
  int rc;
  if ((rc = launch_axis<T>(data, (long)batch * n1 * n2, n3, 1, sign,
                           tw_of<T>(c->tw[2]), c->stream, rl_of(c, 2)))) return rc;
  if ((rc = launch_axis<T>(data, (long)batch * n1, n2, n3, sign,
                           tw_of<T>(c->tw[1]), c->stream, rl_of(c, 1)))) return rc;
  if ((rc = launch_axis<T>(data, (long)batch, n1, (long)n2 * n3, sign,
                           tw_of<T>(c->tw[0]), c->stream, rl_of(c, 0)))) return rc;
  return PFDD_OK;
}

// The contraction in the reference's float arithmetic (PFDD_REF32).
int contract_r32(pfdd_ctx* c, int quirk) {
  const dim3 grid((unsigned)((c->P + 255) / 256));
  cx<float>* W = (cx<float>*)c->W;
#define PFDD_CR(SM) hipLaunchKernelGGL((k_contract_r32<SM>), grid, dim3(256), 0, c->stream, W, c->ec, c->P, quirk)
  if (c->S <= 4) PFDD_CR(4); else if (c->S <= 8) PFDD_CR(8); else if (c->S <= 12) PFDD_CR(12); else PFDD_CR(16);
#undef PFDD_CR
  tap(c, PFDD_STAGE_CONTRACT);
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

template <typename T>
int contract(pfdd_ctx* c, int quirk) {
  if (sizeof(T) == 4 && c->ref32) return contract_r32(c, quirk);
  constexpr int bs = contract_threads<T>();
  const dim3 grid((unsigned)((c->P + bs - 1) / bs), (unsigned)c->B);
  cx<T>* W = (cx<T>*)c->W;
  const T* B = (const T*)c->Bsym;
#define PFDD_CT1(SM, IL)                                                                                       \
  if (c->stored)                                                                                               \
    hipLaunchKernelGGL((k_contract<T, SM, true, IL>), grid, dim3(bs), 0, c->stream, W, B, c->ec, c->P, quirk); \
  else                                                                                                         \
    hipLaunchKernelGGL((k_contract<T, SM, false, IL>), grid, dim3(bs), 0, c->stream, W, B, c->ec, c->P, quirk);
#define PFDD_CT(SM)     \
  if (c->il) {          \
    PFDD_CT1(SM, true)  \
  } else {              \
    PFDD_CT1(SM, false) \
  }
  if (c->S <= 4) { PFDD_CT(4) } else if (c->S <= 8) { PFDD_CT(8) } else if (c->S <= 12) { PFDD_CT(12) } else { PFDD_CT(16) }
#undef PFDD_CT
#undef PFDD_CT1
  tap(c, PFDD_STAGE_CONTRACT);
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

template <typename T>
int update(pfdd_ctx* c, bool finish = true) {
  dim3 grid((unsigned)c->upd_nblk, (unsigned)c->F);
  cx<T>* W = (cx<T>*)c->W;
  const bool m = c->mask != nullptr, tp = c->taupp != nullptr;
#define PFDD_UP(MM, TT)                                                                                  \
  hipLaunchKernelGGL((k_update<T, MM, TT>), grid, dim3(UPD_THREADS), 0, c->stream, xm<T>(c), W, c->mask, c->taupp, \
                     c->sa, c->partials, c->step_ctr)
  if (m && tp) PFDD_UP(true, true);
  else if (m) PFDD_UP(true, false);
  else if (tp) PFDD_UP(false, true);
  else PFDD_UP(false, false);
#undef PFDD_UP
  tap(c, PFDD_STAGE_UPDATE);
  if (finish) {
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(FIN_THREADS), 0, c->stream, c->partials, c->F, c->upd_nblk, c->S,
                       c->hist, c->step_ctr);
    tap(c, PFDD_STAGE_FINISH);
  }
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

constexpr int kFusedLens[] = {32, 48, 64, 96, 128, 192, 256, 384, 512};
constexpr size_t C3_LDS_BUDGET = 104 * 1024;
// The fused column kernel holds S*tile pencils in LDS; at NS=12 that allows
// only 4 (FP64) or 8 (FP32) columns and one block per CU, and it measures
// slower than the three unfused passes (0.71 vs 0.61 ms FP64 at 128^3x12).
// It is used only when it gets at least this many columns per block.
constexpr int C3_MIN_TILE = 16;

bool fused_len(int n) {
  for (int v : kFusedLens)
    if (v == n) return true;
  return false;
}

int row_tile_max(int n, bool fp64) {
  switch (n) {
#define PFDD_L(NN) case NN: return fp64 ? tile_rows<double, NN>() : tile_rows<float, NN>();
    PFDD_L(32) PFDD_L(48) PFDD_L(64) PFDD_L(96) PFDD_L(128) PFDD_L(192) PFDD_L(256) PFDD_L(384) PFDD_L(512)
#undef PFDD_L
    default: return 0;
  }
}

template <typename T, int N>
void k1_launch_n(pfdd_ctx* c) {
  cx<T>* W = (cx<T>*)c->W;
  const long nb = (long)c->F * c->row_nblk;
  constexpr int NT = tile_rows<T, N>() * make_plan(N).tpp;
  const cx<T>* tw = tw_of<T>(c->tw[2]);
  with_master<T>(c, [&](auto xi) {
    using M = decltype(xi);
#define PFDD_K1(MM, TT)                                                                                          \
  hipLaunchKernelGGL((k_row_step<T, N, MM, TT, M>), dim3((unsigned)nb), dim3(NT), 0, c->stream, W, xi, c->mask, \
                     c->taupp, c->sa, c->row_tile, tw, c->partials, c->step_ctr)
    const bool m = c->mask != nullptr, tp = c->taupp != nullptr;
    if (m && tp) PFDD_K1(true, true);
    else if (m) PFDD_K1(true, false);
    else if (tp) PFDD_K1(false, true);
    else PFDD_K1(false, false);
#undef PFDD_K1
  });
}

// Threads of an interleaved-layout block: one TPP-lane group per pencil, in
// whole waves.
template <int N> int il_threads(int npen) { return ((npen * make_plan(N).tpp + 63) / 64) * 64; }

// Row kernel, interleaved layout (c->row_tile rows of every field per block);
// attr_only sets its dynamic-LDS limit instead of launching.
template <typename T, int N>
int k1il_launch_n(pfdd_ctx* c, bool attr_only) {
  const int nt = il_threads<N>(c->S * c->row_tile);
  const size_t lds = row_il_lds_bytes<T, N>(c->S, c->row_tile, nt);
  cx<T>* W = (cx<T>*)c->W;
  const cx<T>* tw = tw_of<T>(c->tw[2]);
  int rc = PFDD_OK;
  with_master<T>(c, [&](auto xi) {
    using M = decltype(xi);
#define PFDD_K1(MM, TT)                                                                                              \
  if (attr_only) {                                                                                                   \
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_row_step_il<T, N, MM, TT, M>),                          \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)                    \
      rc = PFDD_E_HIP;                                                                                               \
  } else {                                                                                                           \
    hipLaunchKernelGGL((k_row_step_il<T, N, MM, TT, M>), dim3((unsigned)c->row_nblk, (unsigned)c->B), dim3(nt), lds, \
                       c->stream, W, xi, c->mask, c->taupp, c->sa, c->row_tile, tw, c->partials, c->step_ctr);     \
  }
    if (attr_only) {
      PFDD_K1(true, true) PFDD_K1(true, false) PFDD_K1(false, true) PFDD_K1(false, false)
    } else {
      const bool m = c->mask != nullptr, tp = c->taupp != nullptr;
      if (m && tp) { PFDD_K1(true, true) }
      else if (m) { PFDD_K1(true, false) }
      else if (tp) { PFDD_K1(false, true) }
      else { PFDD_K1(false, false) }
    }
#undef PFDD_K1
  });
  return rc;
}

// The number of fields the compile-time interleaved row and column kernels
// are instantiated for at each length (0: none): S = 12 at 128 and 256 (the
// 12 FCC systems), S = 3 at 512 (three systems on a 512^3 grid).
constexpr int il_fixed_S(int n) { return n == 128 || n == 256 ? 12 : n == 512 ? 3 : 0; }
// Rows per block of the compile-time interleaved row kernel (k_row_il_fixed)
// for ns fields along n3, or 0 where it is not instantiated.
int il_fixed_rows(int S, int n3) { return S == il_fixed_S(n3) ? 1 : 0; }
template <typename T, int N>
bool k1il_fixed(pfdd_ctx* c) {
  if constexpr (il_fixed_S(N) > 0) {
    constexpr int SF = il_fixed_S(N);
    if (!c->il_fixed_row || c->S != SF || c->mask || c->taupp) return false;
    with_master<T>(c, [&](auto xi) {
      using M = decltype(xi);
      hipLaunchKernelGGL((k_row_il_fixed<T, N, SF, 1, M>), dim3((unsigned)c->row_nblk, (unsigned)c->B),
                         dim3(RowIlFixed<T, N, SF, 1>::NT), 0, c->stream, (cx<T>*)c->W, xi, c->sa,
                         tw_of<T>(c->tw[2]), c->partials, c->step_ctr);
    });
    return true;
  }
  return false;
}

template <typename T>
int k1il_attr(pfdd_ctx* c) {
  switch (c->p.n3) {
#define PFDD_L(NN) case NN: return k1il_launch_n<T, NN>(c, true);
    PFDD_L(32) PFDD_L(48) PFDD_L(64) PFDD_L(96) PFDD_L(128) PFDD_L(192) PFDD_L(256) PFDD_L(384) PFDD_L(512)
#undef PFDD_L
    default: return PFDD_E_RADIX;
  }
}

// The row kernels write one block sum per field and row block; above this
// many per field they are folded to FOLD in parallel before the one-block
// finish (128^3 x 12 FP64, 2048 per field: 0.926 -> 0.922 ms per step).
constexpr int FOLD_ABOVE = 512;
bool folds(const pfdd_ctx* c) { return c->row_nblk > FOLD_ABOVE; }

// The gamma sums of the fused row kernel's block sums (k_finish_sums, with
// step_ctr or without it).
void finish_rows(pfdd_ctx* c, int* step_ctr) {
  const double* src = c->partials;
  int nblk = c->row_nblk;
  if (folds(c)) {
    double* dst = c->partials + c->partials_len - (size_t)c->F * FOLD;
    hipLaunchKernelGGL(k_fold_partials, dim3(FOLD, (unsigned)c->F), dim3(256), 0, c->stream, c->partials,
                       c->row_nblk, dst);
    tap(c, PFDD_STAGE_FINISH);
    src = dst;
    nblk = FOLD;
  }
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(FIN_THREADS), 0, c->stream, src, c->F, nblk, c->S, c->hist,
                     step_ctr);
  tap(c, PFDD_STAGE_FINISH);
}

template <typename T>
int k1_launch(pfdd_ctx* c, bool finish = true) {
  switch (c->p.n3) {
#define PFDD_L(NN)                                        \
  case NN:                                                \
    if (c->il) {                                          \
      if (!k1il_fixed<T, NN>(c)) k1il_launch_n<T, NN>(c, false); \
    } else {                                              \
      k1_launch_n<T, NN>(c);                              \
    }                                                     \
    break;
    PFDD_L(32) PFDD_L(48) PFDD_L(64) PFDD_L(96) PFDD_L(128) PFDD_L(192) PFDD_L(256) PFDD_L(384) PFDD_L(512)
#undef PFDD_L
    default: return PFDD_E_RADIX;
  }
  tap(c, PFDD_STAGE_UPDATE);
  if (finish) finish_rows(c, c->step_ctr);
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

template <typename T, int N, int SM>
const void* k3_symbol() {
  return reinterpret_cast<const void*>(&k_col_contract<T, N, SM>);
}

template <typename T, int N>
int k3_launch_n(pfdd_ctx* c, int quirk, bool attr_only) {
  const long nb = ((long)c->p.n2 * c->p.n3) / c->c3_tile;
  cx<T>* W = (cx<T>*)c->W;
  const cx<T>* tw = tw_of<T>(c->tw[0]);
#define PFDD_K3(SM)                                                                                             \
  if (attr_only) {                                                                                              \
    HIPCK(hipFuncSetAttribute(k3_symbol<T, N, SM>(), hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->c3_lds)); \
  } else {                                                                                                      \
    hipLaunchKernelGGL((k_col_contract<T, N, SM>), dim3((unsigned)nb, (unsigned)c->B), dim3(C3_THREADS), c->c3_lds, \
                       c->stream, W, c->ec, c->P, c->c3_tile, quirk, tw);                                       \
  }
  if (c->S <= 4) { PFDD_K3(4) } else if (c->S <= 8) { PFDD_K3(8) } else if (c->S <= 12) { PFDD_K3(12) } else { PFDD_K3(16) }
#undef PFDD_K3
  if (!attr_only) tap(c, PFDD_STAGE_CONTRACT);
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

// Tile of the compile-time interleaved column kernel (k_col_il_fixed) for
// ns fields along n1, or 0 where it is not instantiated.
// Columns per block: one in FP64; two in FP32, four at S = 3 (N = 512),
// where a column's row segment is only 3 fields long: 96-byte segments
// instead of 48 took the FP32 kernel from 3.9 to 3.1 ms at 512^3 x 3 (two
// columns in FP64, 63 KB of LDS and two blocks per CU, went from 7.1 to
// 8.2 ms).
template <typename T, int N> constexpr int il_fixed_tc() {
  return sizeof(T) == 8 ? 1 : N == 512 ? 4 : 2;
}
int il_fixed_tile(int S, int n1, bool fp64) {
  if (S != il_fixed_S(n1)) return 0;
  return fp64 ? 1 : n1 == 512 ? 4 : 2;
}
template <typename T, int N>
void k3il_fixed(pfdd_ctx* c) {
  constexpr int TC = il_fixed_tc<T, N>(), SF = il_fixed_S(N);
  const long nb = ((long)c->p.n2 * c->p.n3) / TC;
  hipLaunchKernelGGL((k_col_il_fixed<T, N, SF, TC>), dim3((unsigned)nb, (unsigned)c->B),
                     dim3(ColIlFixed<T, N, SF, TC>::NT), 0, c->stream, (cx<T>*)c->W, c->ec, c->P,
                     tw_of<T>(c->tw[0]));
}

// Column kernel, interleaved layout (c->c3_tile columns of every field per block).
template <typename T, int N>
int k3il_launch_n(pfdd_ctx* c, int quirk, bool attr_only) {
  if constexpr (il_fixed_S(N) > 0) {
    if (!attr_only && c->il_fixed && !quirk) {  // (the prologue's virtualevolv term takes the generic kernel)
      k3il_fixed<T, N>(c);
      tap(c, PFDD_STAGE_CONTRACT);
      return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
    }
  }
  const long nb = ((long)c->p.n2 * c->p.n3) / c->c3_tile;
  const int nt = il_threads<N>(c->S * c->c3_tile);
  const size_t lds = col_il_lds_bytes<T, N>(c->S, c->c3_tile);
  cx<T>* W = (cx<T>*)c->W;
  const cx<T>* tw = tw_of<T>(c->tw[0]);
#define PFDD_K3(SM)                                                                                            \
  if (attr_only) {                                                                                             \
    HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_col_contract_il<T, N, SM>),                     \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                         \
  } else {                                                                                                     \
    hipLaunchKernelGGL((k_col_contract_il<T, N, SM>), dim3((unsigned)nb, (unsigned)c->B), dim3(nt), lds,        \
                       c->stream, W, c->ec, c->P, c->c3_tile, quirk, tw);                                      \
  }
  if (c->S <= 4) { PFDD_K3(4) } else if (c->S <= 8) { PFDD_K3(8) } else if (c->S <= 12) { PFDD_K3(12) } else { PFDD_K3(16) }
#undef PFDD_K3
  if (!attr_only) tap(c, PFDD_STAGE_CONTRACT);
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

template <typename T>
int k3_launch(pfdd_ctx* c, int quirk, bool attr_only = false) {
  switch (c->p.n1) {
#define PFDD_L(NN) \
  case NN: return c->il ? k3il_launch_n<T, NN>(c, quirk, attr_only) : k3_launch_n<T, NN>(c, quirk, attr_only);
    PFDD_L(32) PFDD_L(48) PFDD_L(64) PFDD_L(96) PFDD_L(128) PFDD_L(192) PFDD_L(256) PFDD_L(384) PFDD_L(512)
#undef PFDD_L
    default: return PFDD_E_RADIX;
  }
}

// The plane kernel exists for square N2 = N3 planes that fit in LDS.
bool plane_len(int n, bool fp64) {
  if (fp64) return n == 32 || n == 48 || n == 64 || n == 96;
  return n == 32 || n == 48 || n == 64 || n == 96 || n == 128;
}

template <typename T, int N>
int plane_launch_n(pfdd_ctx* c, bool attr_only) {
  constexpr size_t lds = plane_lds_bytes<T, N, N>();
  const cx<T>* tw2 = tw_of<T>(c->tw[1]);
  const cx<T>* tw3 = tw_of<T>(c->tw[2]);
  const unsigned nb = (unsigned)(c->F * c->p.n1);
  cx<T>* W = (cx<T>*)c->W;
  int rc = PFDD_OK;
  with_master<T>(c, [&](auto xi) {
    using M = decltype(xi);
#define PFDD_PL(MM, TT)                                                                                             \
  if (attr_only) {                                                                                                  \
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_step<T, N, N, MM, TT, M>),                       \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)                   \
      rc = PFDD_E_HIP;                                                                                              \
  } else {                                                                                                          \
    hipLaunchKernelGGL((k_plane_step<T, N, N, MM, TT, M>), dim3(nb), dim3(plane_threads<T, N, N>()), lds, c->stream, \
                       W, xi, c->mask, c->taupp, c->sa, tw2, tw3, c->partials, c->step_ctr);                       \
  }
    if (attr_only) {
      PFDD_PL(true, true) PFDD_PL(true, false) PFDD_PL(false, true) PFDD_PL(false, false)
    } else {
      const bool m = c->mask != nullptr, tp = c->taupp != nullptr;
      if (m && tp) { PFDD_PL(true, true) }
      else if (m) { PFDD_PL(true, false) }
      else if (tp) { PFDD_PL(false, true) }
      else { PFDD_PL(false, false) }
    }
#undef PFDD_PL
  });
  if (rc) return rc;
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

template <typename T>
int plane_launch(pfdd_ctx* c, bool finish = true, bool attr_only = false) {
  int rc;
  switch (c->p.n3) {
#define PFDD_L(NN) case NN: rc = plane_launch_n<T, NN>(c, attr_only); break;
    PFDD_L(32) PFDD_L(48) PFDD_L(64) PFDD_L(96)
#undef PFDD_L
    case 128:
      if (sizeof(T) == 4) {
        rc = plane_launch_n<float, 128>(c, attr_only);
        break;
      }
      return PFDD_E_RADIX;
    default:
      return PFDD_E_RADIX;
  }
  if (rc || attr_only) return rc;
  tap(c, PFDD_STAGE_UPDATE);
  if (finish) {
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(FIN_THREADS), 0, c->stream, c->partials, c->F, c->p.n1, c->S,
                       c->hist, c->step_ctr);
    tap(c, PFDD_STAGE_FINISH);
  }
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

// The N1 part of the elastic solve: forward N1, contraction, inverse N1
// (fused into one kernel when it fits).
// One FFT pass of W along axis ax (0: N1, 1: N2, 2: N3) for the context's
// layout: W is [outer][n][inner] with the fields in `outer` (field-major)
// or innermost (interleaved).
// One axis of fourn (PFDD_REF32) over lines of n points of an
// [outer][n][inner] float array; the table is the axis's for `sign`.
int fourn_axis(cx<float>* data, long outer, int n, long inner, const float2* tw, hipStream_t st) {
  int logn = 0;
  while ((1 << logn) < n) logn++;
  const int nt = n / 2 >= FOURN_THREADS ? FOURN_THREADS : n / 2 > 64 ? n / 2 : 64;
  hipLaunchKernelGGL(k_fourn_axis, dim3((unsigned)(outer * inner)), dim3(nt), 0, st, data, inner, n, logn, tw);
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}
const float2* fourn_tw_of(const pfdd_ctx* c, int ax, int sign) { return c->fourn_tw[ax][sign == PFDD_FORWARD ? 0 : 1]; }
// fourn on batch contiguous N1*N2*N3 blocks: the axes N3, N2, N1 (G.c:2768-2826)
int fourn3d(pfdd_ctx* c, float2* buf, int batch, int sign) {
  const long n1 = c->p.n1, n2 = c->p.n2, n3 = c->p.n3;
  cx<float>* d = (cx<float>*)buf;
  int rc;
  if ((rc = fourn_axis(d, batch * n1 * n2, (int)n3, 1, fourn_tw_of(c, 2, sign), c->stream))) return rc;
  if ((rc = fourn_axis(d, batch * n1, (int)n2, n3, fourn_tw_of(c, 1, sign), c->stream))) return rc;
  return fourn_axis(d, batch, (int)n1, n2 * n3, fourn_tw_of(c, 0, sign), c->stream);
}

template <typename T>
int pass(pfdd_ctx* c, int ax, int sign) {
  const long n1 = c->p.n1, n2 = c->p.n2, n3 = c->p.n3;
  const long fo = c->il ? c->B : c->F, fi = c->il ? c->S : 1;
  long outer, inner;
  int n;
  if (ax == 2) {
    outer = fo * n1 * n2, n = (int)n3, inner = fi;
  } else if (ax == 1) {
    outer = fo * n1, n = (int)n2, inner = n3 * fi;
  } else {
    outer = fo, n = (int)n1, inner = n2 * n3 * fi;
  }
  int rc;
  if (sizeof(T) == 4 && c->ref32)
    rc = fourn_axis((cx<float>*)c->W, outer, n, inner, fourn_tw_of(c, ax, sign), c->stream);
  else
    rc = launch_axis<T>((cx<T>*)c->W, outer, n, inner, sign, tw_of<T>(c->tw[ax]), c->stream, rl_of(c, ax));
  tap(c, ax == 0 ? PFDD_STAGE_FFT_AXIS1 : ax == 1 ? PFDD_STAGE_FFT_AXIS2 : PFDD_STAGE_FFT_AXIS3);
  return rc;
}

template <typename T>
int n1_part(pfdd_ctx* c, int quirk) {
  int rc;
  if (c->c3_tile > 0) return k3_launch<T>(c, quirk);
  if ((rc = pass<T>(c, 0, PFDD_FORWARD))) return rc;
  if ((rc = contract<T>(c, quirk))) return rc;
  return pass<T>(c, 0, PFDD_BACKWARD);
}

// Passes between the forward N3 pass and the inverse N3 pass: forward N2,
// then (forward N1, contraction, inverse N1) -- fused when it fits -- then
// inverse N2.
template <typename T>
int middle(pfdd_ctx* c, int quirk) {
  int rc;
  if ((rc = pass<T>(c, 1, PFDD_FORWARD))) return rc;
  if ((rc = n1_part<T>(c, quirk))) return rc;
  return pass<T>(c, 1, PFDD_BACKWARD);
}

template <typename T>
int axis3(pfdd_ctx* c, int sign) {
  return pass<T>(c, 2, sign);
}

// Between GL steps W holds data2 with the inverse transform done along N1
// and N2; the N3 pass is finished by the next step's row kernel (or, when
// the row kernel is not fused, W holds data2 fully in real space).
template <typename T>
int one_step(pfdd_ctx* c) {
  int rc;
  if (c->plane) {
    if ((rc = plane_launch<T>(c))) return rc;
    return n1_part<T>(c, 0);
  }
  if (c->row_tile > 0) {
    if ((rc = k1_launch<T>(c))) return rc;
  } else {
    if ((rc = update<T>(c))) return rc;
    if ((rc = axis3<T>(c, PFDD_FORWARD))) return rc;
  }
  if ((rc = middle<T>(c, 0))) return rc;
  if (c->row_tile == 0) return axis3<T>(c, PFDD_BACKWARD);
  return PFDD_OK;
}

template <typename T>
int prologue(pfdd_ctx* c, int quirk) {
  int rc;
  const long n = (long)c->F * c->P;
  if (c->il)
    hipLaunchKernelGGL((k_xi_to_w_il<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, xm<T>(c),
                       (cx<T>*)c->W, c->P, c->S, n);
  else
    hipLaunchKernelGGL((k_xi_to_w<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, xm<T>(c),
                       (cx<T>*)c->W, n);
  if ((rc = axis3<T>(c, PFDD_FORWARD))) return rc;
  if (c->plane) {  // the next step's plane kernel finishes the N2 and N3 inverse passes
    if ((rc = pass<T>(c, 1, PFDD_FORWARD))) return rc;
    return n1_part<T>(c, quirk);
  }
  if ((rc = middle<T>(c, quirk))) return rc;
  if (c->row_tile == 0) return axis3<T>(c, PFDD_BACKWARD);
  return PFDD_OK;
}

// PFDD_REF32: the step and the prologue in the reference's order (G.c:4574-4644,
// 4450-4452 + 1426-1447): fourn runs the axes N3, N2, N1 in both directions.
int ref32_solve(pfdd_ctx* c, int quirk) {
  int rc;
  for (int ax = 2; ax >= 0; ax--)
    if ((rc = pass<float>(c, ax, PFDD_FORWARD))) return rc;
  if ((rc = contract<float>(c, quirk))) return rc;
  for (int ax = 2; ax >= 0; ax--)
    if ((rc = pass<float>(c, ax, PFDD_BACKWARD))) return rc;
  return PFDD_OK;
}
int ref32_step(pfdd_ctx* c) {
  const int rc = update<float>(c);
  return rc ? rc : ref32_solve(c, 0);
}
int ref32_prologue(pfdd_ctx* c, int quirk) {
  const long n = (long)c->F * c->P;
  hipLaunchKernelGGL((k_xi_to_w<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, xm<float>(c),
                     (cx<float>*)c->W, n);
  return ref32_solve(c, quirk);
}

template <typename T>
void line_sums(pfdd_ctx* c, const int* step_ctr) {
  hipLaunchKernelGGL((k_line_sums<T>), dim3((unsigned)c->F), dim3(UPD_THREADS), 0, c->stream, xm<T>(c), c->P,
                     (const long*)c->line_pts, c->nline, c->F, c->line_out, step_ctr);
}

int step(pfdd_ctx* c) {
  const int rc = c->ref32 ? ref32_step(c) : c->fp64 ? one_step<double>(c) : one_step<float>(c);
  if (rc || c->nline == 0) return rc;
  if (c->fp64) line_sums<double>(c, c->step_ctr);
  else line_sums<float>(c, c->step_ctr);
  tap(c, PFDD_STAGE_FINISH);
  return PFDD_OK;
}

int make_twiddles(Twiddles& t, int n, long long& bytes, hipStream_t st) {
  static const long double TWO_PI = 6.283185307179586476925286766559005768L;
  std::vector<double2> hd(n);
  std::vector<float2> hf(n);
  for (int k = 0; k < n; k++) {
    const long double a = TWO_PI * (long double)k / (long double)n;
    hd[k] = make_double2((double)cosl(a), (double)sinl(a));
    hf[k] = make_float2((float)hd[k].x, (float)hd[k].y);
  }
  t.n = n;
  HIPCK(hipMalloc(&t.d, sizeof(double2) * n));
  HIPCK(hipMalloc(&t.f, sizeof(float2) * n));
  HIPCK(hipMemcpyAsync(t.d, hd.data(), sizeof(double2) * n, hipMemcpyHostToDevice, st));
  HIPCK(hipMemcpyAsync(t.f, hf.data(), sizeof(float2) * n, hipMemcpyHostToDevice, st));
  HIPCK(hipStreamSynchronize(st));
  bytes += (long long)n * 24;
  return PFDD_OK;
}

// fourn's twiddles (G.c:2803-2823) for one axis of n = 2^m points and one
// sign.  The stage of half-span h (ifp1/ip1 in fourn) uses theta = isign *
// 6.28318530717959 / (2h) and the values (float)wr, (float)wi of the double
// recurrence from wr = 1, wi = 0, one per j < h: stored at tw[h - 1 + j].
// Evaluated in the reference's operation order without fused multiply-adds,
// so the table holds the floats fourn multiplies by.
int make_fourn_twiddles(float2*& out, int n, int isign, long long& bytes, hipStream_t st) {
#pragma clang fp contract(off)
  std::vector<float2> h((size_t)n - 1);
  for (int half = 1; half < n; half <<= 1) {
    const double theta = isign * 6.28318530717959 / (double)(2 * half);
    const double wtemp = std::sin(0.5 * theta);
    const double wpr = -2.0 * wtemp * wtemp, wpi = std::sin(theta);
    double wr = 1.0, wi = 0.0;
    for (int j = 0; j < half; j++) {
      h[(size_t)half - 1 + j] = make_float2((float)wr, (float)wi);
      const double wt = wr;
      wr = wt * wpr - wi * wpi + wr;
      wi = wi * wpr + wt * wpi + wi;
    }
  }
  HIPCK(hipMalloc(&out, sizeof(float2) * h.size()));
  HIPCK(hipMemcpyAsync(out, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, st));
  HIPCK(hipStreamSynchronize(st));
  bytes += (long long)sizeof(float2) * (long long)h.size();
  return PFDD_OK;
}

void cijkl(double C[81], double C11, double C12, double C44) {
  const double mu = C44, ll = C12, mup = C11 - C12 - 2 * C44;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int m = 0; m < 3; m++) {
          const double dik = i == k, djm = j == m, dim = i == m, djk = j == k, dij = i == j, dkm = k == m;
          const double d4 = (i == j && j == k && k == m);
          C[((i * 3 + j) * 3 + k) * 3 + m] = mu * (dik * djm + dim * djk) + ll * dij * dkm + mup * d4;
        }
}

template <typename T>
void* dev_alloc(size_t bytes, long long& total) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  total += (long long)bytes;
  return p;
}

int ensure_strain_buffers(pfdd_ctx* c, bool fields) {
  const size_t esz = c->fp64 ? sizeof(double2) : sizeof(float2);
  if (!c->E6) {
    c->E6 = dev_alloc<char>(esz * 6 * (size_t)c->P, c->bytes);
    if (!c->E6) return PFDD_E_ALLOC;
  }
  if (fields && !c->efield) {
    c->efield = (double*)dev_alloc<char>(sizeof(double) * 9 * (size_t)c->P, c->bytes);
    c->sfield = (double*)dev_alloc<char>(sizeof(double) * 9 * (size_t)c->P, c->bytes);
    if (!c->efield || !c->sfield) return PFDD_E_ALLOC;
  }
  return PFDD_OK;
}

template <typename T>
int strain_stress_impl(pfdd_ctx* c, const double* avepst_uniform, double* aveps, double* avesigma,
                       double* eps_field, double* sig_field) {
  int rc;
  const bool fields = eps_field || sig_field;
  if ((rc = ensure_strain_buffers(c, fields))) return rc;
  cx<T>* W = (cx<T>*)c->W;
  cx<T>* E = (cx<T>*)c->E6;
  const long P = c->P;
  const int S = c->S;
  // 1. <xi_a> (avestrain G.c:670-686)
  {
    dim3 grid((unsigned)c->upd_nblk, (unsigned)S);
    hipLaunchKernelGGL((k_xi_sums<T>), grid, dim3(UPD_THREADS), 0, c->stream, xm<T>(c), P, c->upd_nblk, c->partials);
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(FIN_THREADS), 0, c->stream, c->partials, S, c->upd_nblk, S,
                       c->small, (int*)nullptr);
    HIPCK(hipPeekAtLastError());
  }
  double sums[MAXS + 1];
  HIPCK(hipMemcpyAsync(sums, c->small, sizeof(double) * (S + 1), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  double avepst[9];
  for (int ij = 0; ij < 9; ij++) {
    double v = avepst_uniform[ij];
    for (int a = 0; a < S; a++) v += c->p.eps[a][ij / 3][ij % 3] * (sums[a] / (double)P);
    avepst[ij] = v;
  }
  HIPCK(hipMemcpyAsync(c->avepst_d, avepst, sizeof(avepst), hipMemcpyHostToDevice, c->stream));
  // 2. spectrum of xi, strain spectrum, 6 inverse transforms (the work array
  //    is rebuilt from xi afterwards, so the GL state is left untouched)
  const unsigned nb = (unsigned)((P + 255) / 256);
  hipLaunchKernelGGL((k_xi_to_w<T>), dim3((unsigned)((S * P + 255) / 256)), dim3(256), 0, c->stream, xm<T>(c), W, S * P);
  if ((rc = fft3d<T>(c, W, S, PFDD_FORWARD))) return rc;
#define PFDD_SS(SM) hipLaunchKernelGGL((k_strain_spectrum<T, SM>), dim3(nb), dim3(256), 0, c->stream, W, E, c->ec, P)
  if (S <= 4) PFDD_SS(4); else if (S <= 8) PFDD_SS(8); else PFDD_SS(16);
#undef PFDD_SS
  HIPCK(hipPeekAtLastError());
  if ((rc = fft3d<T>(c, E, 6, PFDD_BACKWARD))) return rc;
  // 3. per-point strain and stress, 18 device sums
  double* ef = fields ? c->efield : nullptr;
  double* sf = fields ? c->sfield : nullptr;
#define PFDD_SR(SM)                                                                                       \
  hipLaunchKernelGGL((k_strain_stress<T, SM>), dim3(nb), dim3(256), 0, c->stream, E, xm<T>(c), c->ec, c->avepst_d, P, \
                     c->partials, (int)nb, ef, sf)
  if (S <= 4) PFDD_SR(4); else if (S <= 8) PFDD_SR(8); else PFDD_SR(16);
#undef PFDD_SR
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(FIN_THREADS), 0, c->stream, c->partials, 18, (int)nb, 18, c->small,
                     (int*)nullptr);
  HIPCK(hipPeekAtLastError());
  double out[19];
  HIPCK(hipMemcpyAsync(out, c->small, sizeof(double) * 19, hipMemcpyDeviceToHost, c->stream));
  if (eps_field) HIPCK(hipMemcpyAsync(eps_field, ef, sizeof(double) * 9 * P, hipMemcpyDeviceToHost, c->stream));
  if (sig_field) HIPCK(hipMemcpyAsync(sig_field, sf, sizeof(double) * 9 * P, hipMemcpyDeviceToHost, c->stream));
  // restore the GL state: W must again hold the inverse-transformed data2
  // (the prologue has to be rerun by the caller if it had been done)
  HIPCK(hipStreamSynchronize(c->stream));
  for (int ij = 0; ij < 9; ij++) {
    aveps[ij] = out[ij] / (double)P;
    avesigma[ij] = out[9 + ij] / (double)P;
  }
  return PFDD_OK;
}

}  // namespace

extern "C" {

const char* pfdd_strerror(int rc) {
  switch (rc) {
    case PFDD_OK: return "ok";
    case PFDD_E_HIP: return "HIP runtime error";
    case PFDD_E_ALLOC: return "allocation failed";
    case PFDD_E_SHAPE: return "bad shape or argument";
    case PFDD_E_RADIX: return "unsupported FFT length";
    case PFDD_E_NAN: return "non-finite value";
    case PFDD_E_UNSUPPORTED: return "unsupported input";
    case PFDD_E_STATE: return "call out of order";
    case PFDD_E_LIMIT: return "exceeds the fixed staging array of an allocation-free call";
    default: return "unknown error";
  }
}

void pfdd_destroy(pfdd_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  void* bufs[] = {c->xi,   c->W,        c->Bsym, c->mask, c->taupp, c->taud,   c->partials, c->hist,
                  c->step_ctr, c->small, c->avepst_d, c->ec, c->E6,  c->efield, c->sfield,   c->freq,
                  c->line_pts, c->line_out};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto& t : c->tw) {
    if (t.d) (void)hipFree(t.d);
    if (t.f) (void)hipFree(t.f);
  }
  for (auto& ax : c->fourn_tw)
    for (float2* t : ax)
      if (t) (void)hipFree(t);
  if (c->stream && c->owns_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

namespace {

constexpr int PLAN_FLAGS = PFDD_PLAN_UNFUSED_ROWS | PFDD_PLAN_UNFUSED_COLS | PFDD_PLAN_FUSED_COLS |
                           PFDD_PLAN_NO_PLANE | PFDD_PLAN_PLANE | PFDD_PLAN_INTERLEAVED | PFDD_PLAN_FIELD_MAJOR;
// LDS of one workgroup on gfx950
constexpr size_t LDS_MAX = 160 * 1024;
// AUTO tiles of the interleaved kernels stay within this much LDS (two or
// more blocks per CU)
constexpr size_t IL_AUTO_LDS = 64 * 1024;

// Rows (per block) of the interleaved row kernel and columns of the
// interleaved column kernel: the requested ones (log2 + 1 in the plan word,
// PFDD_E_SHAPE when they do not fit) or, for AUTO, one row, and the widest
// column tile whose rows of all fields stay within 256 bytes.
int il_tiles(pfdd_ctx* c, int plan, int& row_tile, int& c3_tile, size_t& c3_lds) {
  const pfdd_params& p = c->p;
  const int S = c->S;
  const size_t csz = c->fp64 ? sizeof(double2) : sizeof(float2);
  const int rreq = (plan >> PFDD_PLAN_ROW_TILE_SHIFT) & 15, creq = (plan >> PFDD_PLAN_COL_TILE_SHIFT) & 15;
  const long rows = (long)p.n1 * p.n2, inner = (long)p.n2 * p.n3;
  const int tpp3 = make_plan(p.n3).tpp, tpp1 = make_plan(p.n1).tpp;
  auto nt = [](int npen, int tpp) { return ((npen * tpp + 63) / 64) * 64; };
  auto row_lds = [&](int r) { return (size_t)S * r * (p.n3 + 1) * csz + (size_t)nt(S * r, tpp3) * sizeof(double); };
  // (the kernels also hold their axis's twiddle table in LDS: n complex)
  auto row_fits = [&](int r) {
    return nt(S * r, tpp3) <= 1024 && rows % r == 0 && row_lds(r) + p.n3 * csz <= LDS_MAX;
  };
  row_tile = rreq ? 1 << (rreq - 1) : 1;
  if (!row_fits(row_tile)) return PFDD_E_SHAPE;
  c3_tile = 0;
  c3_lds = 0;
  if (fused_len(p.n1) && !c->stored && !(plan & PFDD_PLAN_UNFUSED_COLS)) {
    auto col_lds = [&](int t) { return (size_t)S * t * (p.n1 + 1) * csz; };
    auto col_fits = [&](int t) {
      return nt(S * t, tpp1) <= 1024 && inner % t == 0 && col_lds(t) + p.n1 * csz <= LDS_MAX;
    };
    int t = 1;
    if (creq) {
      t = 1 << (creq - 1);
      if (!col_fits(t)) return PFDD_E_SHAPE;
    } else if (const int ft = il_fixed_tile(S, p.n1, c->fp64 != 0); ft > 0 && inner % ft == 0) {
      t = ft;
    } else {
      while (col_fits(t * 2) && col_lds(t * 2) <= IL_AUTO_LDS && (size_t)t * 2 * S * csz <= 256) t *= 2;
      if (!col_fits(t)) t = 0;
    }
    if (t > 0) {
      c3_tile = t;
      c3_lds = col_lds(t);
    }
  } else if (creq) {
    return PFDD_E_SHAPE;
  }
  return PFDD_OK;
}
constexpr int PLAN_TILES = (15 << PFDD_PLAN_ROW_TILE_SHIFT) | (15 << PFDD_PLAN_COL_TILE_SHIFT) |
                           PFDD_PLAN_RADIX4_AXIS1 | PFDD_PLAN_RADIX4_AXIS2 | PFDD_PLAN_RADIX4_AXIS3;

void drop_graph(pfdd_ctx* c) {
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
}

// Kernel choice for a plan word (PFDD_E_SHAPE when a requested tile does not
// fit).  The state of W between steps depends on the kernels, so a new
// prologue is needed afterwards.
int apply_plan(pfdd_ctx* c, int plan) {
  const pfdd_params& p = c->p;
  if (plan & ~(PLAN_FLAGS | PLAN_TILES)) return PFDD_E_SHAPE;
  const int S = c->S;
  const size_t csz = c->fp64 ? sizeof(double2) : sizeof(float2);
  const int rreq = (plan >> PFDD_PLAN_ROW_TILE_SHIFT) & 15, creq = (plan >> PFDD_PLAN_COL_TILE_SHIFT) & 15;
  int row_tile = 0, row_nblk = 0, c3_tile = 0, plane = 0, il = 0;
  size_t c3_lds = 0;
  if ((plan & PFDD_PLAN_INTERLEAVED) && (plan & PFDD_PLAN_FIELD_MAJOR)) return PFDD_E_SHAPE;
  if (c->ref32) {  // the unfused field-major kernels only: k_update, fourn passes, k_contract_r32
    if (plan & ~(PFDD_PLAN_FIELD_MAJOR | PFDD_PLAN_UNFUSED_ROWS | PFDD_PLAN_UNFUSED_COLS | PFDD_PLAN_NO_PLANE))
      return PFDD_E_SHAPE;
    c->r4 = 0;
    c->row_tile = c->row_nblk = c->c3_tile = 0;
    c->c3_lds = 0;
    c->plane = c->il = c->il_fixed = c->il_fixed_row = 0;
    drop_graph(c);
    c->plan = plan;
    c->prologue_done = 0;
    return PFDD_OK;
  }
  // AUTO takes the interleaved layout where its compile-time row and column
  // kernels exist (S = 12 at N1 = N3 = 128 or 256, S = 3 at 512; recomputed
  // B) and measured faster than the field-major plan: at 128^3 x 12 0.686
  // against 0.909 ms per FP64 step and 0.400 against 0.466 ms per mixed
  // step; at 256^3 x 12 3.82 against 4.52 ms (mixed) and 3.43 against 3.80
  // ms (lean mixed), but 8.60 against 7.54 ms in FP64, where the row
  // kernel's 60 KB of LDS leave two blocks per CU; at 512^3 x 3 8.34
  // against 9.47 ms (mixed) and 7.71 against 8.84 ms (lean mixed), but
  // 17.2 against 16.9 ms in FP64, where the column kernel's 48-byte row
  // segments (three fields of one column) run at 1.8 TB/s.  So FP64 takes
  // it at 128 only.  Without the compile-time kernels the field-major plans
  // measured faster.
  const bool auto_il = !(plan & (PFDD_PLAN_FIELD_MAJOR | PFDD_PLAN_UNFUSED_ROWS | PFDD_PLAN_UNFUSED_COLS |
                                 PFDD_PLAN_FUSED_COLS | PFDD_PLAN_PLANE | PFDD_PLAN_NO_PLANE)) &&
                       !c->stored && il_fixed_tile(S, p.n1, c->fp64 != 0) > 0 && il_fixed_rows(S, p.n3) > 0 &&
                       (!c->fp64 || (p.n1 == 128 && p.n3 == 128));
  if ((plan & PFDD_PLAN_INTERLEAVED) || auto_il) {
    // interleaved layout: the interleaved row kernel, no plane kernel
    if (!fused_len(p.n3) || (plan & (PFDD_PLAN_UNFUSED_ROWS | PFDD_PLAN_PLANE))) return PFDD_E_SHAPE;
    const int rc = il_tiles(c, plan, row_tile, c3_tile, c3_lds);
    if (rc) return rc;
    row_nblk = (int)((long)p.n1 * p.n2 / row_tile);
    il = 1;
  } else if (fused_len(p.n3) && !(plan & PFDD_PLAN_UNFUSED_ROWS)) {
    // fused row kernel: rows per block = the largest power of two <= its
    // compile-time tile that divides the N1*N2 rows of a field (or the one asked for)
    const long rows = (long)p.n1 * p.n2;
    const int tmax = row_tile_max(p.n3, c->fp64);
    int t = 1;
    if (rreq) {
      t = 1 << (rreq - 1);
      if (t > tmax || rows % t) return PFDD_E_SHAPE;
    } else {
      while (t * 2 <= tmax && rows % (t * 2) == 0) t *= 2;
    }
    row_tile = t;
    row_nblk = (int)(rows / t);
  } else if (rreq) {
    return PFDD_E_SHAPE;
  }
  // fused column kernel: S*tile pencils of N1 points in LDS
  if (il) {
    // (tiles chosen above)
  } else if (fused_len(p.n1) && !c->stored && !(plan & PFDD_PLAN_UNFUSED_COLS)) {
    const size_t per_col = (size_t)S * (p.n1 + 1) * csz;
    const int maxpen = (c->fp64 ? 1 : 2) * (C3_THREADS / make_plan(p.n1).tpp);
    const long inner = (long)p.n2 * p.n3;
    auto fits = [&](int t) { return t * S <= maxpen && t * per_col <= C3_LDS_BUDGET && inner % t == 0; };
    int t = 0;
    if (creq) {
      t = 1 << (creq - 1);
      if (!fits(t)) return PFDD_E_SHAPE;
    } else {
      for (int cand = 1; fits(cand); cand *= 2) t = cand;
      if (t < ((plan & PFDD_PLAN_FUSED_COLS) ? 1 : C3_MIN_TILE)) t = 0;
    }
    if (t > 0) {
      c3_tile = t;
      c3_lds = t * per_col;
    }
  } else if (creq) {
    return PFDD_E_SHAPE;
  }
  // the plane kernel holds one block per CU; at 128^2 (FP32) it measured no faster than the row kernel +
  // two N2 passes (0.36 vs 0.35 ms at 128^3 x 12), so AUTO takes it up to 96^2 and PFDD_PLAN_PLANE forces it
  if (!il && !(plan & PFDD_PLAN_NO_PLANE) && p.n2 == p.n3 && plane_len(p.n3, c->fp64) &&
      (p.n3 <= 96 || (plan & PFDD_PLAN_PLANE)))
    plane = 1;
  c->r4 = ((plan & PFDD_PLAN_RADIX4_AXIS1) ? 1 : 0) | ((plan & PFDD_PLAN_RADIX4_AXIS2) ? 2 : 0) |
          ((plan & PFDD_PLAN_RADIX4_AXIS3) ? 4 : 0);
  c->row_tile = row_tile;
  c->row_nblk = row_nblk;
  c->c3_tile = c3_tile;
  c->c3_lds = c3_lds;
  c->plane = plane;
  c->il = il;
  c->il_fixed = il && c3_tile > 0 && c3_tile == il_fixed_tile(S, p.n1, c->fp64 != 0);
  c->il_fixed_row = il && row_tile == il_fixed_rows(S, p.n3);
  int rc = PFDD_OK;
  if (il && (rc = c->fp64 ? k1il_attr<double>(c) : k1il_attr<float>(c))) return rc;
  if (c3_tile > 0 && (rc = c->fp64 ? k3_launch<double>(c, 0, true) : k3_launch<float>(c, 0, true))) return rc;
  if (plane && (rc = c->fp64 ? plane_launch<double>(c, false, true) : plane_launch<float>(c, false, true))) return rc;
  drop_graph(c);
  c->plan = plan;
  c->prologue_done = 0;
  return PFDD_OK;
}

int validate_params(const pfdd_params* p) {
  if (p->ns < 1 || p->ns > PFDD_MAX_NS) return PFDD_E_SHAPE;
  if (p->precision != PFDD_FP64 && p->precision != PFDD_MIXED && p->precision != PFDD_MIXED_LEAN &&
      p->precision != PFDD_REF32)
    return PFDD_E_SHAPE;
  if (p->bmode != PFDD_B_STORED && p->bmode != PFDD_B_RECOMPUTE) return PFDD_E_SHAPE;
  const int ns3[3] = {p->n1, p->n2, p->n3};
  for (int i = 0; i < 3; i++)
    if (!len_ok(ns3[i]) || (p->precision == PFDD_REF32 && (ns3[i] & (ns3[i] - 1)))) return PFDD_E_RADIX;
  // PFDD_REF32 runs the unfused field-major kernels only
  if (p->precision == PFDD_REF32 &&
      (p->plan & ~(PFDD_PLAN_FIELD_MAJOR | PFDD_PLAN_UNFUSED_ROWS | PFDD_PLAN_UNFUSED_COLS | PFDD_PLAN_NO_PLANE)))
    return PFDD_E_SHAPE;
  const double* chk[] = {&p->C11, &p->C12, &p->C44, &p->d1, &p->d2, &p->d3, &p->uq2, &p->pi_const};
  for (const double* v : chk)
    if (!std::isfinite(*v)) return PFDD_E_NAN;
  if (!(p->C44 > 0) || !(p->d1 > 0) || !(p->d2 > 0) || !(p->d3 > 0)) return PFDD_E_SHAPE;
  for (int a = 0; a < p->ns; a++) {
    if (!std::isfinite(p->cd2[a]) || !std::isfinite(p->asf2[a]) || !std::isfinite(p->dslip[a]) || p->dslip[a] == 0)
      return PFDD_E_NAN;
    for (int i = 0; i < 9; i++)
      if (!std::isfinite(p->eps[a][i / 3][i % 3])) return PFDD_E_NAN;
  }
  if (p->plan & ~(PLAN_FLAGS | PLAN_TILES)) return PFDD_E_SHAPE;
  return PFDD_OK;
}

int push_tau(pfdd_ctx* c) {
  HIPCK(hipMemcpyAsync(c->taud, c->tauh.data(), sizeof(double2) * c->F, hipMemcpyHostToDevice, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return PFDD_OK;
}

// A context of `batch` instances of the grid in *p.  With `shared` the
// context runs on that stream (an ensemble's) and does not own it.
int ctx_create(const pfdd_params* p, int batch, hipStream_t shared, pfdd_ctx** out) {
  *out = nullptr;
  int rc = validate_params(p);
  if (rc) return rc;
  if (batch < 1 || (long)batch * p->ns > FIN_MAX_FIELDS) return PFDD_E_SHAPE;
  pfdd_ctx* c = new (std::nothrow) pfdd_ctx();
  if (!c) return PFDD_E_ALLOC;
  c->p = *p;
  c->S = p->ns;
  c->B = batch;
  c->F = batch * p->ns;
  c->P = (long)p->n1 * p->n2 * p->n3;
  c->ntri = c->S * (c->S + 1) / 2;
  c->fp64 = p->precision == PFDD_FP64;
  c->lean = p->precision == PFDD_MIXED_LEAN;
  c->ref32 = p->precision == PFDD_REF32;
  c->stored = p->bmode == PFDD_B_STORED && !c->ref32;  // PFDD_REF32 forms B in FP64 at each k-point
  c->tauh.assign((size_t)c->F, make_double2(0.0, 0.0));
  c->inst_pp.assign((size_t)batch, 0);
  c->inst_frozen.assign((size_t)batch, 0);
  c->inst_up.assign((size_t)batch, 0);
#define FAIL(code) do { rc = (code); goto fail; } while (0)
  {
    if (shared) {
      c->stream = shared;
      c->owns_stream = 0;
    } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      FAIL(PFDD_E_HIP);
    }
    const size_t csz = c->fp64 ? sizeof(double2) : sizeof(float2);
    const size_t rsz = c->fp64 ? sizeof(double) : sizeof(float);
    const long P = c->P;
    const int S = c->S, F = c->F;
    const int ns3[3] = {p->n1, p->n2, p->n3};
    c->upd_nblk = (int)((P + UPD_THREADS * UPD_EPT - 1) / (UPD_THREADS * UPD_EPT));
    const long strain_nblk = (P + 255) / 256;
    // block sums of every plan (the row kernel's tile may go down to one row)
    // (+ the folded sums of the interleaved row kernel, at the end)
    c->partials_len = (size_t)std::max<long>(std::max<long>((long)F * c->upd_nblk, 18 * strain_nblk),
                                             std::max<long>((long)F * p->n1 * p->n2, (long)F * p->n1)) +
                      (size_t)F * FOLD;
    c->xi_bytes = (master64(c) ? master_bytes<double>(false) : master_bytes<float>(c->lean != 0)) * F * P;
    if (!(c->xi = dev_alloc<char>(c->xi_bytes, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->W = dev_alloc<char>(csz * F * P, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->partials = (double*)dev_alloc<char>(sizeof(double) * c->partials_len, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->hist = (double*)dev_alloc<char>(sizeof(double) * HIST_CAP * c->B * (S + 1), c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->step_ctr = (int*)dev_alloc<char>(sizeof(int) * 4, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->small = (double*)dev_alloc<char>(sizeof(double) * 32, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->avepst_d = (double*)dev_alloc<char>(sizeof(double) * 9, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->ec = (ElasticConsts*)dev_alloc<char>(sizeof(ElasticConsts), c->bytes))) FAIL(PFDD_E_ALLOC);
    if (!(c->taud = (double2*)dev_alloc<char>(sizeof(double2) * F, c->bytes))) FAIL(PFDD_E_ALLOC);
    if (hipMemsetAsync(c->step_ctr, 0, sizeof(int) * 4, c->stream) != hipSuccess) FAIL(PFDD_E_HIP);
    if (hipMemsetAsync(c->xi, 0, c->xi_bytes, c->stream) != hipSuccess) FAIL(PFDD_E_HIP);
    if (hipMemsetAsync(c->W, 0, csz * F * P, c->stream) != hipSuccess) FAIL(PFDD_E_HIP);
    if ((rc = push_tau(c))) goto fail;
    // twiddle tables, one per axis; LDS attributes of the axis kernels
    for (int i = 0; i < 3; i++) {
      if ((rc = make_twiddles(c->tw[i], ns3[i], c->bytes, c->stream))) goto fail;
      if ((rc = cols_attr(ns3[i]))) goto fail;
      if (c->ref32)
        for (int d = 0; d < 2; d++)
          if ((rc = make_fourn_twiddles(c->fourn_tw[i][d], ns3[i], d ? PFDD_BACKWARD : PFDD_FORWARD, c->bytes,
                                        c->stream)))
            goto fail;
    }
    // elastic constants (Bmatrix G.c:340-358; s_a = C:e_a etc.)
    ElasticConsts h{};
    h.mu = p->C44;
    h.ll = p->C12;
    h.mup = p->C11 - p->C12 - 2 * p->C44;
    h.inv_mu = 1.0 / p->C44;
    cijkl(h.C, p->C11, p->C12, p->C44);
    for (int pq = 0; pq < 9; pq++) {
      double t = 0.0;
      for (int ij = 0; ij < 9; ij++) t += h.C[pq * 9 + ij];
      h.T[pq] = t;
    }
    for (int a = 0; a < S; a++) {
      for (int ij = 0; ij < 9; ij++) h.eps[a][ij] = p->eps[a][ij / 3][ij % 3];
      double ts = 0.0;
      for (int ij = 0; ij < 9; ij++) {
        double s = 0.0;
        for (int mn = 0; mn < 9; mn++) s += h.C[ij * 9 + mn] * h.eps[a][mn];
        h.sig[a][ij] = s;
        ts += s;
      }
      h.Tsum[a] = ts;
    }
    {  // six symmetric components (00, 01, 02, 11, 12, 22) of s_b and of w_t sym(e_a)_t / mu
      const int ti[6] = {0, 0, 0, 1, 1, 2}, tj[6] = {0, 1, 2, 1, 2, 2};
      for (int a = 0; a < S; a++)
        for (int t = 0; t < 6; t++) {
          const int i = ti[t], j = tj[t];
          h.sig6[a][t] = 0.5 * (h.sig[a][3 * i + j] + h.sig[a][3 * j + i]);
          h.e6m[a][t] = (i == j ? 1.0 : 2.0) * 0.5 * (h.eps[a][3 * i + j] + h.eps[a][3 * j + i]) * h.inv_mu;
          h.sig6f[a][t] = (float)h.sig6[a][t];
          h.e6mf[a][t] = (float)h.e6m[a][t];
        }
    }
    for (int a = 0; a < S; a++)
      for (int b = 0; b < S; b++) {
        double se = 0.0;
        for (int mn = 0; mn < 9; mn++) se += h.sig[a][mn] * h.eps[b][mn];
        h.SE[a][b] = se;
      }
    h.d[0] = p->d1; h.d[1] = p->d2; h.d[2] = p->d3;
    // frec (G.c:2698-2741) per axis, evaluated with the same arithmetic as the reference
    std::vector<double> fr((size_t)p->n1 + p->n2 + p->n3);
    {
      size_t o = 0;
      for (int ax = 0; ax < 3; ax++)
        for (int i = 0; i < ns3[ax]; i++, o++) {
          const double dd = h.d[ax];
          fr[o] = i == 0 ? 0.0 : i < ns3[ax] / 2 ? (double)i / (double)ns3[ax] / dd
                                                  : ((double)i - (double)ns3[ax]) / (double)ns3[ax] / dd;
        }
    }
    if (!(c->freq = (double*)dev_alloc<char>(sizeof(double) * fr.size(), c->bytes))) FAIL(PFDD_E_ALLOC);
    if (hipMemcpyAsync(c->freq, fr.data(), sizeof(double) * fr.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
      FAIL(PFDD_E_HIP);
    h.freq[0] = c->freq;
    h.freq[1] = c->freq + p->n1;
    h.freq[2] = c->freq + p->n1 + p->n2;
    h.iso = h.mup == 0.0;
    h.g_iso = (h.mu + h.ll) / (h.mu * h.mu * (1.0 + (h.mu + h.ll) / h.mu));
    h.n[0] = p->n1; h.n[1] = p->n2; h.n[2] = p->n3;
    h.ns = S;
    if (hipMemcpyAsync(c->ec, &h, sizeof(h), hipMemcpyHostToDevice, c->stream) != hipSuccess) FAIL(PFDD_E_HIP);
    if (hipStreamSynchronize(c->stream) != hipSuccess) FAIL(PFDD_E_HIP);
    // step arguments
    c->sa.P = P;
    c->sa.ns = S;
    c->sa.nblk = c->upd_nblk;
    c->sa.uq2 = p->uq2;
    c->sa.pi = p->pi_const;
    c->sa.uq2f = (float)p->uq2;
    c->sa.nsizef = (float)P;
    for (int a = 0; a < S; a++) {
      c->sa.cd2[a] = p->cd2[a];
      c->sa.asf2[a] = p->asf2[a];
      c->sa.dslip[a] = p->dslip[a];
    }
    c->sa.tau = c->taud;
    if (c->stored) {
      if (!(c->Bsym = dev_alloc<char>(rsz * c->ntri * P, c->bytes))) FAIL(PFDD_E_ALLOC);
      const unsigned nb = (unsigned)((P + 255) / 256);
#define PFDD_BB(T, SM) hipLaunchKernelGGL((k_build_B<T, SM>), dim3(nb), dim3(256), 0, c->stream, (T*)c->Bsym, c->ec, P)
      if (c->fp64) {
        if (S <= 4) PFDD_BB(double, 4); else if (S <= 8) PFDD_BB(double, 8); else PFDD_BB(double, 16);
      } else {
        if (S <= 4) PFDD_BB(float, 4); else if (S <= 8) PFDD_BB(float, 8); else PFDD_BB(float, 16);
      }
#undef PFDD_BB
      if (hipPeekAtLastError() != hipSuccess) FAIL(PFDD_E_HIP);
    }
    if ((rc = apply_plan(c, p->plan))) goto fail;
    if (hipStreamSynchronize(c->stream) != hipSuccess) FAIL(PFDD_E_HIP);
  }
#undef FAIL
  *out = c;
  return PFDD_OK;
fail:
  pfdd_destroy(c);
  return rc;
}

// Resolved shear of instance b: tau[S] (uniform) or tau[P][S] (per point).
// Rates are reset to zero.  While any instance has a per-point shear the
// kernels read taupp for every instance, so uniform ones are expanded there.
int set_tau_inst(pfdd_ctx* c, int b, const double* tau, int per_point) {
  const int S = c->S;
  const long P = c->P;
  const size_t n = (size_t)P * S;
  if (!per_point) {
    for (int a = 0; a < S; a++)
      if (!std::isfinite(tau[a])) return PFDD_E_NAN;
  } else {
    for (size_t i = 0; i < n; i++)
      if (!std::isfinite(tau[i])) return PFDD_E_NAN;
  }
  HIPCK(hipStreamSynchronize(c->stream));
  c->inst_pp[b] = per_point ? 1 : 0;
  for (int a = 0; a < S; a++) c->tauh[(size_t)b * S + a] = make_double2(per_point ? 0.0 : tau[a], 0.0);
  bool any_pp = false;
  for (uint8_t v : c->inst_pp) any_pp = any_pp || v;
  if (!any_pp) {
    if (c->taupp) {
      (void)hipFree(c->taupp);
      c->taupp = nullptr;
    }
  } else {
    std::vector<double> h;
    auto expand = [&](int j) -> int {  // uniform shear of instance j -> its taupp slice
      h.resize(n);
      for (long pt = 0; pt < P; pt++)
        for (int a = 0; a < S; a++) h[(size_t)pt * S + a] = c->tauh[(size_t)j * S + a].x;
      HIPCK(hipMemcpyAsync(c->taupp + (size_t)j * n, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
      HIPCK(hipStreamSynchronize(c->stream));
      return PFDD_OK;
    };
    int rc;
    if (!c->taupp) {
      if (!(c->taupp = (double*)dev_alloc<char>(sizeof(double) * n * c->B, c->bytes))) return PFDD_E_ALLOC;
      for (int j = 0; j < c->B; j++)
        if (j != b && !c->inst_pp[j] && (rc = expand(j))) return rc;
    }
    if (per_point) {
      HIPCK(hipMemcpyAsync(c->taupp + (size_t)b * n, tau, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
      HIPCK(hipStreamSynchronize(c->stream));
    } else if ((rc = expand(b))) {
      return rc;
    }
  }
  drop_graph(c);  // the captured graph holds the old kernel arguments
  return push_tau(c);
}

int upload_inst(pfdd_ctx* c, int b, const double* xi, const double* xi_bc, int nsv) {
  if (!xi || nsv < c->S) return PFDD_E_SHAPE;
  const long P = c->P;
  const int S = c->S;
  const size_t n = (size_t)S * P, off = (size_t)b * n;
  for (size_t i = 0; i < 2 * n; i++)
    if (!std::isfinite(xi[i])) return PFDD_E_NAN;
  for (size_t i = 2 * n; i < 2 * (size_t)nsv * P; i++)
    if (xi[i] != 0.0) return PFDD_E_UNSUPPORTED;
  bool frozen = false;
  std::vector<uint8_t> m;
  if (xi_bc) {
    for (size_t i = 0; i < n && !frozen; i++) frozen = xi_bc[2 * i] != 0.0;
    if (frozen) {
      m.resize(n);
      for (size_t i = 0; i < n; i++) m[i] = xi_bc[2 * i] != 0.0;
    }
  }
  HIPCK(hipStreamSynchronize(c->stream));
  if (master64(c)) {
    HIPCK(hipMemcpyAsync((double2*)c->xi + off, xi, sizeof(double2) * n, hipMemcpyHostToDevice, c->stream));
  } else {  // split into the compensated form of Master<float> (or round, in the lean form)
    std::vector<float2> r(n);
    std::vector<float> im(c->lean ? 0 : n);
    for (size_t i = 0; i < n; i++) {
      const double re = xi[2 * i];
      const float hi = (float)re;
      if (c->lean) {
        r[i] = make_float2(hi, (float)xi[2 * i + 1]);
      } else {
        r[i] = make_float2(hi, (float)(re - (double)hi));
        im[i] = (float)xi[2 * i + 1];
      }
    }
    const Master<float> m = xm<float>(c);
    HIPCK(hipMemcpyAsync(m.r + off, r.data(), sizeof(float2) * n, hipMemcpyHostToDevice, c->stream));
    if (m.im) HIPCK(hipMemcpyAsync(m.im + off, im.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  c->inst_frozen[b] = frozen ? 1 : 0;
  bool any = false;
  for (uint8_t v : c->inst_frozen) any = any || v;
  if (frozen) {
    if (!c->mask) {
      if (!(c->mask = (uint8_t*)dev_alloc<char>(n * c->B, c->bytes))) return PFDD_E_ALLOC;
      HIPCK(hipMemsetAsync(c->mask, 0, n * c->B, c->stream));
    }
    HIPCK(hipMemcpyAsync(c->mask + off, m.data(), n, hipMemcpyHostToDevice, c->stream));
  } else if (c->mask) {
    if (!any) {
      HIPCK(hipStreamSynchronize(c->stream));
      (void)hipFree(c->mask);
      c->mask = nullptr;
    } else {
      HIPCK(hipMemsetAsync(c->mask + off, 0, n, c->stream));
    }
  }
  HIPCK(hipStreamSynchronize(c->stream));
  drop_graph(c);
  c->inst_up[b] = 1;
  bool all = true;
  for (uint8_t v : c->inst_up) all = all && v;
  c->has_xi = all ? 1 : 0;
  c->prologue_done = 0;
  return PFDD_OK;
}

int download_inst(pfdd_ctx* c, int b, double* xi, int nsv) {
  if (!xi || nsv < c->S) return PFDD_E_SHAPE;
  const long P = c->P;
  const size_t n = (size_t)c->S * P;
  HIPCK(hipStreamSynchronize(c->stream));
  if (master64(c)) {
    HIPCK(hipMemcpyAsync(xi, (double2*)c->xi + (size_t)b * n, sizeof(double2) * n, hipMemcpyDeviceToHost,
                         c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float2> r(n);
    const Master<float> m = xm<float>(c);
    std::vector<float> im(m.im ? n : 0);
    HIPCK(hipMemcpyAsync(r.data(), m.r + (size_t)b * n, sizeof(float2) * n, hipMemcpyDeviceToHost, c->stream));
    if (m.im)
      HIPCK(hipMemcpyAsync(im.data(), m.im + (size_t)b * n, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) {
      if (m.im) {
        xi[2 * i] = (double)r[i].x + (double)r[i].y;
        xi[2 * i + 1] = (double)im[i];
      } else {
        xi[2 * i] = (double)r[i].x;
        xi[2 * i + 1] = (double)r[i].y;
      }
    }
  }
  if (nsv > c->S) memset(xi + 2 * n, 0, sizeof(double) * 2 * (size_t)(nsv - c->S) * P);
  return PFDD_OK;
}

int begin_outer(pfdd_ctx* c, int quirk) {
  if (!c->has_xi) return PFDD_E_STATE;
  const int rc = c->ref32  ? ref32_prologue(c, quirk ? 1 : 0)
                 : c->fp64 ? prologue<double>(c, quirk ? 1 : 0)
                           : prologue<float>(c, quirk ? 1 : 0);
  if (rc) return rc;
  c->prologue_done = 1;
  return PFDD_OK;
}

// Runs nsteps GL steps of every context in cs (all on one stream), in chunks
// of up to HIST_CAP steps; GRAPH_STEPS steps of all of them are captured once
// into *graph and replayed.  step_all (when given) enqueues one step of all
// of them instead of step() per context.  After each chunk, sink(ctx index,
// first step, steps, history rows) receives the gamma history of each context.
template <typename Sink>
int run_steps(pfdd_ctx* const* cs, int nc, hipStream_t st, hipGraphExec_t* graph, int nsteps, bool want_hist,
              Sink&& sink, const std::function<int()>* step_all = nullptr) {
  auto one = [&]() -> int {
    if (step_all) return (*step_all)();
    for (int i = 0; i < nc; i++) {
      const int rc = step(cs[i]);
      if (rc) return rc;
    }
    return PFDD_OK;
  };
  int done = 0;
  std::vector<double> h;
  while (done < nsteps) {
    const int chunk = std::min(HIST_CAP, nsteps - done);
    for (int i = 0; i < nc; i++)
      hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, cs[i]->step_ctr, 0);
    HIPCK(hipPeekAtLastError());
    int left = chunk;
    if (left >= GRAPH_STEPS && !*graph) {
      hipGraph_t g;
      HIPCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = PFDD_OK;
      for (int s = 0; s < GRAPH_STEPS && rc == PFDD_OK; s++) rc = one();
      hipError_t ec = hipStreamEndCapture(st, &g);
      if (rc) {
        if (ec == hipSuccess) (void)hipGraphDestroy(g);
        return rc;
      }
      HIPCK(ec);
      ec = hipGraphInstantiate(graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCK(ec);
    }
    while (left >= GRAPH_STEPS && *graph) {
      HIPCK(hipGraphLaunch(*graph, st));
      left -= GRAPH_STEPS;
    }
    for (; left > 0; left--) {
      const int rc = one();
      if (rc) return rc;
    }
    if (want_hist) {
      for (int i = 0; i < nc; i++) {
        const size_t row = (size_t)cs[i]->B * (cs[i]->S + 1);
        h.resize((size_t)chunk * row);
        HIPCK(hipMemcpyAsync(h.data(), cs[i]->hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        sink(i, done, chunk, h.data());
      }
    }
    done += chunk;
  }
  return PFDD_OK;
}

int gl_steps_ctx(pfdd_ctx* c, int nsteps, double* gamma, double* gamma_a) {
  const int S = c->S;
  return run_steps(&c, 1, c->stream, &c->graph, nsteps, gamma || gamma_a,
                   [&](int, int first, int n, const double* h) {
                     for (int s = 0; s < n; s++) {
                       if (gamma) gamma[first + s] = h[(size_t)s * (S + 1) + S];
                       if (gamma_a)
                         for (int a = 0; a < S; a++) gamma_a[(size_t)(first + s) * S + a] = h[(size_t)s * (S + 1) + a];
                     }
                   });
}

// Candidate plan words of the autotuner: the kernel-fusion choices, and for
// the fused row and column kernels their largest tile and two smaller ones.
std::vector<int> plan_candidates(pfdd_ctx* c) {
  const int base[] = {PFDD_PLAN_AUTO,
                      PFDD_PLAN_NO_PLANE,
                      PFDD_PLAN_NO_PLANE | PFDD_PLAN_UNFUSED_ROWS,
                      PFDD_PLAN_NO_PLANE | PFDD_PLAN_UNFUSED_COLS,
                      PFDD_PLAN_NO_PLANE | PFDD_PLAN_UNFUSED_ROWS | PFDD_PLAN_UNFUSED_COLS,
                      PFDD_PLAN_NO_PLANE | PFDD_PLAN_FUSED_COLS,
                      PFDD_PLAN_PLANE,
                      PFDD_PLAN_PLANE | PFDD_PLAN_UNFUSED_COLS,
                      PFDD_PLAN_PLANE | PFDD_PLAN_FUSED_COLS};
  std::vector<int> out;
  std::vector<long> seen;
  const int saved = c->plan;
  auto add = [&](int plan) {
    if (apply_plan(c, plan) != PFDD_OK) return;
    const long sig = ((long)c->row_tile << 32) | ((long)c->c3_tile << 8) | (c->il << 1) | c->plane;
    for (long v : seen)
      if (v == sig) return;
    seen.push_back(sig);
    out.push_back(plan);
  };
  for (int b : base) {
    add(b);
    if (apply_plan(c, b) != PFDD_OK) continue;
    const int rt = c->row_tile, ct = c->c3_tile, pl = c->plane;
    int lg = 0;
    if (rt > 0 && !pl)
      for (lg = 0; (1 << lg) < rt; lg++) {}
    for (int d = 1; d <= 2 && rt > 0 && !pl && lg - d >= 0; d++) add(b | ((lg - d + 1) << PFDD_PLAN_ROW_TILE_SHIFT));
    if (ct > 0)
      for (lg = 0; (1 << lg) < ct; lg++) {}
    for (int d = 1; d <= 2 && ct > 0 && lg - d >= 0; d++) add(b | ((lg - d + 1) << PFDD_PLAN_COL_TILE_SHIFT));
  }
  // interleaved layout: 1, 2 or 4 rows per row block; the AUTO column tile,
  // half and twice it, and the unfused N1 passes
  if (apply_plan(c, PFDD_PLAN_INTERLEAVED) == PFDD_OK) {
    const int ct = c->c3_tile;
    int lg = 0;
    while (ct > 0 && (1 << lg) < ct) lg++;
    for (int r = 0; r <= 2; r++) {
      const int rb = (r + 1) << PFDD_PLAN_ROW_TILE_SHIFT;
      add(PFDD_PLAN_INTERLEAVED | rb | PFDD_PLAN_UNFUSED_COLS);
      if (ct == 0) continue;
      for (int d = -1; d <= 1; d++)
        if (lg + d >= 0) add(PFDD_PLAN_INTERLEAVED | rb | ((lg + d + 1) << PFDD_PLAN_COL_TILE_SHIFT));
    }
  }
  (void)apply_plan(c, saved);
  return out;
}

// Times every candidate plan over `steps` GL steps (after GRAPH_STEPS
// warm-up steps that also capture the graph) and applies the fastest.  xi and
// the step counters are saved before and restored after each candidate, so
// the context's state is as before except that a new prologue is needed.
int autotune_ctx(pfdd_ctx* c, int steps, int* best_plan, float* best_ms) {
  if (c->ref32) return PFDD_E_UNSUPPORTED;  // one plan only
  if (!c->has_xi) return PFDD_E_STATE;
  if (steps < 1) return PFDD_E_SHAPE;
  const size_t xbytes = c->xi_bytes;
  void* xsave = nullptr;
  int* csave = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PFDD_OK;
  const std::vector<int> cands = plan_candidates(c);
  int best = c->plan;
  float best_t = 0.f;
  if (hipMalloc(&xsave, xbytes) != hipSuccess) return PFDD_E_ALLOC;
  if (hipMalloc(&csave, sizeof(int) * 4) != hipSuccess) {
    (void)hipFree(xsave);
    return PFDD_E_ALLOC;
  }
  auto restore = [&]() -> int {
    HIPCK(hipMemcpyAsync(c->xi, xsave, xbytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->step_ctr, csave, sizeof(int) * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return PFDD_OK;
  };
  const int timed = ((steps + GRAPH_STEPS - 1) / GRAPH_STEPS) * GRAPH_STEPS;
  bool saved = false;
  do {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { rc = PFDD_E_HIP; break; }
    if (hipMemcpyAsync(xsave, c->xi, xbytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(csave, c->step_ctr, sizeof(int) * 4, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = PFDD_E_HIP;
      break;
    }
    saved = true;
    // Two rounds over a candidate list; each candidate keeps its faster
    // round (the first also absorbs clock ramp-up and first-launch costs).
    auto tune = [&](const std::vector<int>& list) -> int {
      std::vector<float> tmin(list.size(), 0.f);
      for (int round = 0; round < 2; round++)
        for (size_t ci = 0; ci < list.size(); ci++) {
          int r;
          if ((r = apply_plan(c, list[ci]))) return r;
          if ((r = begin_outer(c, 0))) return r;
          if ((r = gl_steps_ctx(c, GRAPH_STEPS, nullptr, nullptr))) return r;
          HIPCK(hipEventRecord(e0, c->stream));
          if ((r = gl_steps_ctx(c, timed, nullptr, nullptr))) return r;
          HIPCK(hipEventRecord(e1, c->stream));
          HIPCK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCK(hipEventElapsedTime(&ms, e0, e1));
          ms /= (float)timed;
          if (tmin[ci] == 0.f || ms < tmin[ci]) tmin[ci] = ms;
          if ((r = restore())) return r;
        }
      best_t = 0.f;
      for (size_t ci = 0; ci < list.size(); ci++)
        if (best_t == 0.f || tmin[ci] < best_t) {
          best_t = tmin[ci];
          best = list[ci];
        }
      return PFDD_OK;
    };
    // stage 1: kernel fusion and tiles
    if ((rc = tune(cands))) break;
    // stage 2: radix-4 sequences for the standalone passes the best plan uses
    if ((rc = apply_plan(c, best))) break;
    const int n3[3] = {c->p.n1, c->p.n2, c->p.n3};
    const bool standalone[3] = {c->c3_tile == 0, !c->plane, !c->plane && c->row_tile == 0};
    const int bit[3] = {PFDD_PLAN_RADIX4_AXIS1, PFDD_PLAN_RADIX4_AXIS2, PFDD_PLAN_RADIX4_AXIS3};
    int axes = 0;
    for (int ax = 0; ax < 3; ax++)
      if (standalone[ax] && has_r4_variant(n3[ax])) axes |= bit[ax];
    if (axes) {
      std::vector<int> radix = {best};
      for (int sub = axes; sub > 0; sub = (sub - 1) & axes) radix.push_back(best | sub);
      if ((rc = tune(radix))) break;
    }
  } while (0);
  if (saved) {
    const int rc1 = restore();
    if (rc == PFDD_OK) rc = rc1;
  }
  const int rc2 = apply_plan(c, best);
  if (rc == PFDD_OK) rc = rc2;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(xsave);
  (void)hipFree(csave);
  if (rc == PFDD_OK) {
    if (best_plan) *best_plan = best;
    if (best_ms) *best_ms = best_t;
  }
  return rc;
}

}  // namespace

extern "C" {

int pfdd_create(const pfdd_params* p, pfdd_ctx** out) {
  if (!p || !out) return PFDD_E_SHAPE;
  return ctx_create(p, 1, nullptr, out);
}

int pfdd_set_plan(pfdd_ctx* c, int plan) {
  if (!c) return PFDD_E_SHAPE;
  HIPCK(hipStreamSynchronize(c->stream));
  return apply_plan(c, plan);
}

int pfdd_get_plan(const pfdd_ctx* c) { return c ? c->plan : PFDD_E_SHAPE; }

int pfdd_autotune(pfdd_ctx* c, int steps, int* plan, float* ms_per_step) {
  if (!c) return PFDD_E_SHAPE;
  return autotune_ctx(c, steps, plan, ms_per_step);
}

int pfdd_set_tau_ramp(pfdd_ctx* c, const double* tau0, const double* rate) {
  if (!c || !tau0 || !rate) return PFDD_E_SHAPE;
  for (int a = 0; a < c->S; a++)
    if (!std::isfinite(tau0[a]) || !std::isfinite(rate[a])) return PFDD_E_NAN;
  int rc = pfdd_set_tau(c, tau0, 0);
  if (rc) return rc;
  for (int a = 0; a < c->S; a++) c->tauh[a].y = rate[a];
  return push_tau(c);
}

int pfdd_set_tau(pfdd_ctx* c, const double* tau, int per_point) {
  if (!c || !tau) return PFDD_E_SHAPE;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->step_ctr + 1, 0);  // restart the ramp clock
  HIPCK(hipPeekAtLastError());
  return set_tau_inst(c, 0, tau, per_point);
}

int pfdd_upload_xi(pfdd_ctx* c, const double* xi, const double* xi_bc, int nsv) {
  if (!c) return PFDD_E_SHAPE;
  return upload_inst(c, 0, xi, xi_bc, nsv);
}

int pfdd_download_xi(pfdd_ctx* c, double* xi, int nsv) {
  if (!c) return PFDD_E_SHAPE;
  return download_inst(c, 0, xi, nsv);
}

int pfdd_begin_outer(pfdd_ctx* c, int quirk) {
  if (!c) return PFDD_E_SHAPE;
  return begin_outer(c, quirk);
}

int pfdd_gl_steps(pfdd_ctx* c, int nsteps, double* gamma, double* gamma_a) {
  if (!c || nsteps < 0) return PFDD_E_SHAPE;
  if (!c->prologue_done) return PFDD_E_STATE;
  return gl_steps_ctx(c, nsteps, gamma, gamma_a);
}

int pfdd_set_line_points(pfdd_ctx* c, const long long* pts, int n) {
  if (!c || n < 0 || (long)n > c->P || (n > 0 && !pts)) return PFDD_E_SHAPE;
  for (int i = 0; i < n; i++)
    if (pts[i] < 0 || pts[i] >= c->P) return PFDD_E_SHAPE;
  HIPCK(hipStreamSynchronize(c->stream));
  drop_graph(c);  // the captured steps hold the old list (or none)
  if (c->line_pts) {
    (void)hipFree(c->line_pts);
    (void)hipFree(c->line_out);
    c->bytes -= (long long)(sizeof(long) * c->nline + sizeof(double) * HIST_CAP * c->F);
    c->line_pts = nullptr;
    c->line_out = nullptr;
    c->nline = 0;
  }
  if (n == 0) return PFDD_OK;
  static_assert(sizeof(long) == sizeof(long long), "64-bit point indices");
  c->line_pts = (long*)dev_alloc<char>(sizeof(long) * n, c->bytes);
  c->line_out = (double*)dev_alloc<char>(sizeof(double) * HIST_CAP * c->F, c->bytes);
  if (!c->line_pts || !c->line_out) {
    (void)hipFree(c->line_pts);
    (void)hipFree(c->line_out);
    c->line_pts = nullptr;
    c->line_out = nullptr;
    return PFDD_E_ALLOC;
  }
  HIPCK(hipMemcpy(c->line_pts, pts, sizeof(long) * n, hipMemcpyHostToDevice));
  c->nline = n;
  return PFDD_OK;
}

int pfdd_line_sums(pfdd_ctx* c, double* sums) {
  if (!c || !sums) return PFDD_E_SHAPE;
  if (!c->has_xi || c->nline == 0) return PFDD_E_STATE;
  if (c->fp64) line_sums<double>(c, nullptr);
  else line_sums<float>(c, nullptr);
  HIPCK(hipPeekAtLastError());
  HIPCK(hipMemcpyAsync(sums, c->line_out, sizeof(double) * c->F, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return PFDD_OK;
}

int pfdd_gl_steps_lines(pfdd_ctx* c, int nsteps, double* gamma, double* gamma_a, double* line_sums) {
  if (!c || nsteps < 0 || !line_sums) return PFDD_E_SHAPE;
  if (!c->prologue_done || c->nline == 0) return PFDD_E_STATE;
  const int S = c->S, F = c->F;
  int copy_rc = PFDD_OK;
  std::vector<double> rows;
  const int rc = run_steps(&c, 1, c->stream, &c->graph, nsteps, true, [&](int, int first, int n, const double* h) {
    for (int s = 0; s < n; s++) {
      if (gamma) gamma[first + s] = h[(size_t)s * (S + 1) + S];
      if (gamma_a)
        for (int a = 0; a < S; a++) gamma_a[(size_t)(first + s) * S + a] = h[(size_t)s * (S + 1) + a];
    }
    rows.resize((size_t)n * F);
    if (copy_rc == PFDD_OK &&
        (hipMemcpyAsync(rows.data(), c->line_out, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, c->stream) !=
             hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
      copy_rc = PFDD_E_HIP;
    if (copy_rc == PFDD_OK) std::copy(rows.begin(), rows.end(), line_sums + (size_t)first * F);
  });
  return rc ? rc : copy_rc;
}

int pfdd_strain_stress(pfdd_ctx* c, const double* avepst_uniform, double* aveps, double* avesigma,
                       double* eps_field, double* sig_field) {
  if (!c || !avepst_uniform || !aveps || !avesigma) return PFDD_E_SHAPE;
  if (!c->has_xi) return PFDD_E_STATE;
  int rc = c->fp64 ? strain_stress_impl<double>(c, avepst_uniform, aveps, avesigma, eps_field, sig_field)
                   : strain_stress_impl<float>(c, avepst_uniform, aveps, avesigma, eps_field, sig_field);
  // the work array was reused; the next GL step needs a new prologue
  c->prologue_done = 0;
  return rc;
}

int pfdd_fft(pfdd_ctx* c, void* dev_buf, int batch, int dir, int prec) {
  if (!c || !dev_buf || batch < 1) return PFDD_E_SHAPE;
  if (dir != PFDD_FORWARD && dir != PFDD_BACKWARD) return PFDD_E_SHAPE;
  if (prec == PFDD_FP64) return fft3d<double>(c, (double2*)dev_buf, batch, dir);
  if (prec == PFDD_MIXED) return fft3d<float>(c, (float2*)dev_buf, batch, dir);
  if (prec == PFDD_REF32) return c->ref32 ? fourn3d(c, (float2*)dev_buf, batch, dir) : PFDD_E_STATE;
  return PFDD_E_SHAPE;
}

int pfdd_sync(pfdd_ctx* c) {
  if (!c) return PFDD_E_SHAPE;
  HIPCK(hipStreamSynchronize(c->stream));
  return PFDD_OK;
}

void* pfdd_stream(pfdd_ctx* c) { return c ? (void*)c->stream : nullptr; }

long long pfdd_device_bytes(const pfdd_ctx* c) { return c ? c->bytes : 0; }

}  // extern "C"

namespace {

template <typename T>
int enqueue_stage(pfdd_ctx* c, int stage, int reps) {
  cx<T>* W = (cx<T>*)c->W;
  const long n = (long)c->F * c->P;
  int rc = PFDD_OK;
  for (int r = 0; r < reps && rc == PFDD_OK; r++) {
    const int sign = (r & 1) ? PFDD_BACKWARD : PFDD_FORWARD;
    switch (stage) {
      case PFDD_STAGE_UPDATE:
        rc = c->plane ? plane_launch<T>(c, false) : c->row_tile > 0 ? k1_launch<T>(c, false) : update<T>(c, false);
        break;
      case PFDD_STAGE_FINISH:
        if (!c->plane && c->row_tile > 0)
          finish_rows(c, nullptr);
        else
          hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(FIN_THREADS), 0, c->stream, c->partials, c->F,
                             c->plane ? c->p.n1 : c->upd_nblk, c->S, c->hist, (int*)nullptr);
        rc = hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
        break;
      case PFDD_STAGE_FFT_AXIS3:
        rc = pass<T>(c, 2, sign);
        break;
      case PFDD_STAGE_FFT_AXIS2:
        rc = pass<T>(c, 1, sign);
        break;
      case PFDD_STAGE_FFT_AXIS1:
        rc = pass<T>(c, 0, sign);
        break;
      case PFDD_STAGE_CONTRACT:
        rc = c->c3_tile > 0 ? k3_launch<T>(c, 0) : contract<T>(c, 0);
        break;
      case PFDD_STAGE_STEP:
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->step_ctr, 0);
        rc = c->ref32 ? ref32_step(c) : one_step<T>(c);
        break;
      case PFDD_STAGE_COPY:
        hipLaunchKernelGGL((k_xi_to_w<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, xm<T>(c), W, n);
        rc = hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
        break;
      default:
        return PFDD_E_SHAPE;
    }
  }
  return rc;
}

}  // namespace

extern "C" int pfdd_step_plan(const pfdd_ctx* c, int* launches) {
  if (!c || !launches) return PFDD_E_SHAPE;
  const bool r = c->row_tile > 0 || c->plane, k3 = c->c3_tile > 0;
  launches[PFDD_STAGE_UPDATE] = 1;
  launches[PFDD_STAGE_FFT_AXIS3] = r ? 0 : 2;
  launches[PFDD_STAGE_FFT_AXIS2] = c->plane ? 0 : 2;
  launches[PFDD_STAGE_FFT_AXIS1] = k3 ? 0 : 2;
  launches[PFDD_STAGE_CONTRACT] = 1;
  launches[PFDD_STAGE_STEP] = 0;
  launches[PFDD_STAGE_COPY] = 0;
  // + the fold of the row kernel's block sums when there are many, + the line sums while a list is set
  launches[PFDD_STAGE_FINISH] = 1 + (folds(c) ? 1 : 0) + (c->nline > 0 ? 1 : 0);
  return PFDD_OK;
}

extern "C" int pfdd_time_step_kernels(pfdd_ctx* c, int nsteps, float* ms, int* launches) {
  if (!c || nsteps < 1 || nsteps > HIST_CAP || !ms || !launches) return PFDD_E_SHAPE;
  if (!c->prologue_done) return PFDD_E_STATE;
  const int per_step_max = 16;
  std::vector<hipEvent_t> pool((size_t)nsteps * per_step_max + 1, nullptr);
  std::vector<std::pair<int, hipEvent_t>> taps;
  int rc = PFDD_OK;
  // timing-only events: no system-scope fence (cache writeback) at each record,
  // which would otherwise add to the time of the launch before it
  for (auto& e : pool)
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) rc = PFDD_E_HIP;
  if (rc == PFDD_OK) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->step_ctr, 0);
    taps.reserve(pool.size());
    c->tap_pool = &pool;
    c->tap = &taps;
    tap(c, -1);  // start mark
    for (int s = 0; s < nsteps && rc == PFDD_OK; s++) rc = step(c);
    c->tap = nullptr;
    c->tap_pool = nullptr;
    if (rc == PFDD_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = PFDD_E_HIP;
  }
  if (rc == PFDD_OK) {
    for (int k = 0; k < 8; k++) {
      ms[k] = 0.f;
      launches[k] = 0;
    }
    for (size_t i = 1; i < taps.size(); i++) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, taps[i - 1].second, taps[i].second) != hipSuccess) {
        rc = PFDD_E_HIP;
        break;
      }
      const int st = taps[i].first;
      if (st >= 0 && st < 8) {
        ms[st] += t / (float)nsteps;
        launches[st]++;
      }
    }
    for (int k = 0; k < 8; k++) launches[k] /= nsteps;
  }
  for (auto& e : pool)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

extern "C" int pfdd_enqueue_stage(pfdd_ctx* c, int stage, int reps) {
  if (!c || reps < 0 || stage < PFDD_STAGE_UPDATE || stage > PFDD_STAGE_FINISH) return PFDD_E_SHAPE;
  c->prologue_done = 0;
  return c->fp64 ? enqueue_stage<double>(c, stage, reps) : enqueue_stage<float>(c, stage, reps);
}

// ---------------------------------------------------------------- ensembles
// Instances with the same grid and material (every pfdd_params field but the
// applied shear, which is set per instance) form a class; each class is one
// batched context, and every class runs on the ensemble's stream, so that
// GRAPH_STEPS steps of the whole ensemble are one captured graph.
namespace {
// The persistent step of the classes of one precision (k_ens_step).
struct PersGroup {
  EnsDev* dev = nullptr;
  int* q = nullptr;
  int nc = 0;
  int grid = 0;
  size_t lds = 0;
};
}  // namespace

struct pfdd_ens {
  int count = 0;
  hipStream_t stream = nullptr;
  std::vector<pfdd_ctx*> cls;
  std::vector<int> cls_of, slot_of;  // per instance
  hipGraphExec_t graph = nullptr;
  int prologue_done = 0;
  int mode = PFDD_ENS_PHASED;
  PersGroup pg[2];                   // [0] FP32 classes, [1] FP64 classes
  int pers_built = 0;
  float ms_phased = 0.f, ms_persistent = 0.f;  // last autotune measurement of each mode
};

namespace {

bool same_class(const pfdd_params& a, const pfdd_params& b) {
  if (a.n1 != b.n1 || a.n2 != b.n2 || a.n3 != b.n3 || a.ns != b.ns || a.precision != b.precision ||
      a.bmode != b.bmode || a.plan != b.plan)
    return false;
  const double x[] = {a.C11, a.C12, a.C44, a.d1, a.d2, a.d3, a.uq2, a.pi_const};
  const double y[] = {b.C11, b.C12, b.C44, b.d1, b.d2, b.d3, b.uq2, b.pi_const};
  for (int i = 0; i < 8; i++)
    if (x[i] != y[i]) return false;
  for (int s = 0; s < a.ns; s++) {
    if (a.cd2[s] != b.cd2[s] || a.asf2[s] != b.asf2[s] || a.dslip[s] != b.dslip[s]) return false;
    for (int i = 0; i < 9; i++)
      if (a.eps[s][i / 3][i % 3] != b.eps[s][i / 3][i % 3]) return false;
  }
  return true;
}

void ens_drop_graph(pfdd_ens* e) {
  if (e->graph) {
    (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
}

bool ens_inst_ok(const pfdd_ens* e, int inst) { return e && inst >= 0 && inst < e->count; }

bool pers_len(int n) { return n == 32 || n == 48 || n == 64 || n == 96; }

// A class the persistent step can run (see k_ens_step).
bool pers_eligible(const pfdd_ctx* c) {
  return c->S <= 4 && !c->stored && !c->mask && !c->taupp && c->p.n1 == c->p.n2 && c->p.n2 == c->p.n3 &&
         pers_len(c->p.n1);
}

size_t plane_lds_of(int n, bool fp64) {
  switch (n) {
#define PFDD_L(NN) case NN: return fp64 ? plane_lds_bytes<double, NN, NN>() : plane_lds_bytes<float, NN, NN>();
    PFDD_L(32) PFDD_L(48) PFDD_L(64) PFDD_L(96)
#undef PFDD_L
    default: return 0;
  }
}

template <typename T>
const void* ens_kernel() {
  return reinterpret_cast<const void*>(&k_ens_step<T>);
}

// Work lists and launch shape of the persistent step, per precision.
int pers_build(pfdd_ens* e) {
  int dev = 0, ncu = 0;
  HIPCK(hipGetDevice(&dev));
  HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  for (int fp = 0; fp < 2; fp++) {
    std::vector<pfdd_ctx*> cl;
    for (pfdd_ctx* c : e->cls)
      if (c->fp64 == fp) cl.push_back(c);
    PersGroup& g = e->pg[fp];
    g.nc = (int)cl.size();
    if (cl.empty()) continue;
    if ((int)cl.size() > ENS_MAX_CLASSES) return PFDD_E_UNSUPPORTED;
    // larger grids first, so the long items start early
    std::stable_sort(cl.begin(), cl.end(), [](const pfdd_ctx* a, const pfdd_ctx* b) { return a->p.n1 > b->p.n1; });
    EnsDev h{};
    h.nc = (int)cl.size();
    long plane = 0, col = 0;
    size_t lds = 0;
    for (int k = 0; k < h.nc; k++) {
      const pfdd_ctx* c = cl[k];
      if (!pers_eligible(c) || !c->plane || c->c3_tile <= 0) return PFDD_E_UNSUPPORTED;
      EnsClassDev& d = h.cls[k];
      d.W = c->W;
      d.xi = c->xi;
      d.lean = c->lean;
      d.ec = c->ec;
      for (int ax = 0; ax < 3; ax++) d.tw[ax] = fp ? (const void*)c->tw[ax].d : (const void*)c->tw[ax].f;
      d.sa = c->sa;
      d.partials = c->partials;
      d.hist = c->hist;
      d.step_ctr = c->step_ctr;
      d.P = c->P;
      d.n = c->p.n1;
      d.F = c->F;
      d.B = c->B;
      d.S = c->S;
      d.c3_tile = c->c3_tile;
      d.plane_first = plane;
      d.col_first = col;
      d.n_planes = (long)c->F * c->p.n1;
      plane += d.n_planes;
      col += (long)c->B * ((long)c->p.n2 * c->p.n3 / c->c3_tile);
      lds = std::max(lds, std::max(plane_lds_of(c->p.n1, fp), std::max(c->c3_lds, sizeof(double) * c->F)));
    }
    if (plane + h.nc + col >= (1L << 31)) return PFDD_E_UNSUPPORTED;  // items are drawn as int
    h.n_plane = plane;
    h.n_items = plane + h.nc + col;
    if (!g.q) {
      if (hipMalloc(&g.q, sizeof(int) * ENS_Q_INTS) != hipSuccess) return PFDD_E_ALLOC;
      HIPCK(hipMemsetAsync(g.q, 0, sizeof(int) * ENS_Q_INTS, e->stream));
    }
    h.q = g.q;
    if (!g.dev && hipMalloc(&g.dev, sizeof(EnsDev)) != hipSuccess) return PFDD_E_ALLOC;
    HIPCK(hipMemcpyAsync(g.dev, &h, sizeof(EnsDev), hipMemcpyHostToDevice, e->stream));
    HIPCK(hipStreamSynchronize(e->stream));
    const void* kern = fp ? ens_kernel<double>() : ens_kernel<float>();
    HIPCK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, ENS_THREADS, lds));
    if (per_cu < 1) return PFDD_E_UNSUPPORTED;
    g.grid = per_cu * ncu;
    g.lds = lds;
  }
  e->pers_built = 1;
  return PFDD_OK;
}

// One GL step of every class: per precision, the counter reset and one launch
// of k_ens_step (both ordinary kernels, so the step can be captured).
int pers_step(pfdd_ens* e) {
  for (int fp = 0; fp < 2; fp++) {
    PersGroup& g = e->pg[fp];
    if (!g.nc) continue;
    hipLaunchKernelGGL(k_ens_reset, dim3(1), dim3(64), 0, e->stream, g.q);
    if (fp)
      hipLaunchKernelGGL(k_ens_step<double>, dim3((unsigned)g.grid), dim3(ENS_THREADS), g.lds, e->stream, g.dev);
    else
      hipLaunchKernelGGL(k_ens_step<float>, dim3((unsigned)g.grid), dim3(ENS_THREADS), g.lds, e->stream, g.dev);
  }
  return hipPeekAtLastError() == hipSuccess ? PFDD_OK : PFDD_E_HIP;
}

// Error flags of the persistent step (a completion wait that ran out); the
// flags are cleared.  Synchronous.
int pers_check(pfdd_ens* e) {
  int bad = 0;
  for (PersGroup& g : e->pg) {
    if (!g.nc || !g.q) continue;
    int flag = 0;
    HIPCK(hipMemcpyAsync(&flag, g.q + 1, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCK(hipStreamSynchronize(e->stream));
    if (flag) {
      bad = 1;
      hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, e->stream, g.q + 1, 0);
      HIPCK(hipStreamSynchronize(e->stream));
    }
  }
  return bad ? PFDD_E_HIP : PFDD_OK;
}

// Plans the persistent step needs (plane kernel + fused columns).
int pers_plans(pfdd_ens* e) {
  for (pfdd_ctx* c : e->cls) {
    if (!pers_eligible(c)) return PFDD_E_UNSUPPORTED;
    const int rc = apply_plan(c, PFDD_PLAN_PLANE | PFDD_PLAN_FUSED_COLS);
    if (rc) return rc;
    if (!c->plane || c->c3_tile <= 0) return PFDD_E_UNSUPPORTED;
  }
  return PFDD_OK;
}

}  // namespace

extern "C" {

void pfdd_destroy_ensemble(pfdd_ens* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  for (PersGroup& g : e->pg) {
    if (g.dev) (void)hipFree(g.dev);
    if (g.q) (void)hipFree(g.q);
  }
  for (pfdd_ctx* c : e->cls) pfdd_destroy(c);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int pfdd_create_ensemble(const pfdd_params* params, int count, pfdd_ens** out) {
  if (!params || !out || count < 1) return PFDD_E_SHAPE;
  *out = nullptr;
  for (int i = 0; i < count; i++) {
    const int rc = validate_params(&params[i]);
    if (rc) return rc;
    if (params[i].precision == PFDD_REF32) return PFDD_E_UNSUPPORTED;
  }
  pfdd_ens* e = new (std::nothrow) pfdd_ens();
  if (!e) return PFDD_E_ALLOC;
  e->count = count;
  e->cls_of.assign((size_t)count, -1);
  e->slot_of.assign((size_t)count, 0);
  std::vector<int> first;  // first instance of each class
  std::vector<int> size;
  for (int i = 0; i < count; i++) {
    int k = 0;
    while (k < (int)first.size() && !same_class(params[first[k]], params[i])) k++;
    if (k == (int)first.size()) {
      first.push_back(i);
      size.push_back(0);
    }
    e->cls_of[i] = k;
    e->slot_of[i] = size[k]++;
  }
  int rc = PFDD_OK;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    rc = PFDD_E_HIP;
  } else {
    for (size_t k = 0; k < first.size() && rc == PFDD_OK; k++) {
      pfdd_ctx* c = nullptr;
      rc = ctx_create(&params[first[k]], size[k], e->stream, &c);
      if (rc == PFDD_OK) e->cls.push_back(c);
    }
  }
  if (rc) {
    pfdd_destroy_ensemble(e);
    return rc;
  }
  *out = e;
  return PFDD_OK;
}

int pfdd_ensemble_info(const pfdd_ens* e, int* nclasses, int* class_of, int* plans) {
  if (!e) return PFDD_E_SHAPE;
  if (nclasses) *nclasses = (int)e->cls.size();
  if (class_of)
    for (int i = 0; i < e->count; i++) class_of[i] = e->cls_of[i];
  if (plans)
    for (size_t k = 0; k < e->cls.size(); k++) plans[k] = e->cls[k]->plan;
  return PFDD_OK;
}

int pfdd_ensemble_set_tau(pfdd_ens* e, int inst, const double* tau, int per_point) {
  if (!ens_inst_ok(e, inst) || !tau) return PFDD_E_SHAPE;
  ens_drop_graph(e);
  return set_tau_inst(e->cls[e->cls_of[inst]], e->slot_of[inst], tau, per_point);
}

int pfdd_ensemble_upload_xi(pfdd_ens* e, int inst, const double* xi, const double* xi_bc, int nsv) {
  if (!ens_inst_ok(e, inst)) return PFDD_E_SHAPE;
  ens_drop_graph(e);
  e->prologue_done = 0;
  return upload_inst(e->cls[e->cls_of[inst]], e->slot_of[inst], xi, xi_bc, nsv);
}

int pfdd_ensemble_download_xi(pfdd_ens* e, int inst, double* xi, int nsv) {
  if (!ens_inst_ok(e, inst)) return PFDD_E_SHAPE;
  return download_inst(e->cls[e->cls_of[inst]], e->slot_of[inst], xi, nsv);
}

int pfdd_ensemble_begin_outer(pfdd_ens* e, int quirk) {
  if (!e) return PFDD_E_SHAPE;
  for (pfdd_ctx* c : e->cls)
    if (!c->has_xi) return PFDD_E_STATE;
  for (pfdd_ctx* c : e->cls) {
    const int rc = begin_outer(c, quirk);
    if (rc) return rc;
  }
  e->prologue_done = 1;
  return PFDD_OK;
}

int pfdd_ensemble_gl_steps(pfdd_ens* e, int nsteps, double* gamma) {
  if (!e || nsteps < 0) return PFDD_E_SHAPE;
  if (!e->prologue_done) return PFDD_E_STATE;
  // instances of each class in slot order
  std::vector<std::vector<int>> members(e->cls.size());
  for (int i = 0; i < e->count; i++) members[e->cls_of[i]].push_back(i);
  const int count = e->count;
  const bool pers = e->mode == PFDD_ENS_PERSISTENT;
  std::function<int()> pstep;
  if (pers) {
    for (pfdd_ctx* c : e->cls)
      if (!pers_eligible(c)) return PFDD_E_UNSUPPORTED;  // e.g. a per-point shear set since
    if (!e->pers_built) {
      const int rc = pers_build(e);
      if (rc) return rc;
    }
    pstep = [e]() { return pers_step(e); };
  }
  int rc = run_steps(e->cls.data(), (int)e->cls.size(), e->stream, &e->graph, nsteps, gamma != nullptr,
                     [&](int k, int first, int n, const double* h) {
                       const int S = e->cls[k]->S;
                       const size_t row = (size_t)e->cls[k]->B * (S + 1);
                       for (int s = 0; s < n; s++)
                         for (size_t b = 0; b < members[k].size(); b++)
                           gamma[(size_t)(first + s) * count + members[k][b]] = h[(size_t)s * row + b * (S + 1) + S];
                     },
                     pers ? &pstep : nullptr);
  // a completion wait of the persistent step that ran out shows at the next
  // synchronous call (with gamma, at once)
  if (rc == PFDD_OK && pers && gamma) rc = pers_check(e);
  return rc;
}

int pfdd_ensemble_set_mode(pfdd_ens* e, int mode) {
  if (!e || (mode != PFDD_ENS_PHASED && mode != PFDD_ENS_PERSISTENT)) return PFDD_E_SHAPE;
  HIPCK(hipStreamSynchronize(e->stream));
  ens_drop_graph(e);
  e->prologue_done = 0;
  e->pers_built = 0;
  if (mode == PFDD_ENS_PERSISTENT) {
    const int rc = pers_plans(e);
    if (rc) {
      for (pfdd_ctx* c : e->cls) (void)apply_plan(c, c->p.plan);
      e->mode = PFDD_ENS_PHASED;
      return rc;
    }
  }
  e->mode = mode;
  return PFDD_OK;
}

int pfdd_ensemble_get_mode(const pfdd_ens* e, float* ms_phased, float* ms_persistent) {
  if (!e) return PFDD_E_SHAPE;
  if (ms_phased) *ms_phased = e->ms_phased;
  if (ms_persistent) *ms_persistent = e->ms_persistent;
  return e->mode;
}

int pfdd_ensemble_autotune(pfdd_ens* e, int steps, int* plans, float* ms_per_step) {
  if (!e) return PFDD_E_SHAPE;
  int rc = pfdd_ensemble_set_mode(e, PFDD_ENS_PHASED);
  if (rc) return rc;
  std::vector<int> tuned(e->cls.size());
  for (size_t k = 0; k < e->cls.size(); k++) {
    float ms = 0.f;
    if ((rc = autotune_ctx(e->cls[k], steps, &tuned[k], &ms))) return rc;
    if (plans) plans[k] = tuned[k];
    if (ms_per_step) ms_per_step[k] = ms;
  }
  // the mode: the tuned phased plans against the persistent step
  e->ms_phased = e->ms_persistent = 0.f;
  bool elig = true;
  for (pfdd_ctx* c : e->cls) elig = elig && pers_eligible(c);
  if (!elig) return PFDD_OK;
  std::vector<void*> xsave(e->cls.size(), nullptr);
  std::vector<int*> csave(e->cls.size(), nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int timed = ((std::max(steps, 1) + GRAPH_STEPS - 1) / GRAPH_STEPS) * GRAPH_STEPS;
  auto restore = [&]() -> int {
    for (size_t k = 0; k < e->cls.size(); k++) {
      pfdd_ctx* c = e->cls[k];
      HIPCK(hipMemcpyAsync(c->xi, xsave[k], c->xi_bytes, hipMemcpyDeviceToDevice, e->stream));
      HIPCK(hipMemcpyAsync(c->step_ctr, csave[k], sizeof(int) * 4, hipMemcpyDeviceToDevice, e->stream));
    }
    HIPCK(hipStreamSynchronize(e->stream));
    return PFDD_OK;
  };
  auto time_mode = [&](int mode, float* best) -> int {
    int r;
    if (mode == PFDD_ENS_PHASED) {
      if ((r = pfdd_ensemble_set_mode(e, PFDD_ENS_PHASED))) return r;
      for (size_t k = 0; k < e->cls.size(); k++)
        if ((r = apply_plan(e->cls[k], tuned[k]))) return r;
    } else if ((r = pfdd_ensemble_set_mode(e, PFDD_ENS_PERSISTENT))) {
      return r;
    }
    if ((r = pfdd_ensemble_begin_outer(e, 0))) return r;
    if ((r = pfdd_ensemble_gl_steps(e, GRAPH_STEPS, nullptr))) return r;
    HIPCK(hipEventRecord(e0, e->stream));
    if ((r = pfdd_ensemble_gl_steps(e, timed, nullptr))) return r;
    HIPCK(hipEventRecord(e1, e->stream));
    HIPCK(hipEventSynchronize(e1));
    if (mode == PFDD_ENS_PERSISTENT && (r = pers_check(e))) return r;
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, e0, e1));
    ms /= (float)timed;
    if (*best == 0.f || ms < *best) *best = ms;
    return restore();
  };
  bool saved = false;
  do {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { rc = PFDD_E_HIP; break; }
    for (size_t k = 0; k < e->cls.size() && rc == PFDD_OK; k++) {
      pfdd_ctx* c = e->cls[k];
      const size_t xb = c->xi_bytes;
      if (hipMalloc(&xsave[k], xb) != hipSuccess || hipMalloc(&csave[k], sizeof(int) * 4) != hipSuccess) {
        rc = PFDD_E_ALLOC;
        break;
      }
      if (hipMemcpyAsync(xsave[k], c->xi, xb, hipMemcpyDeviceToDevice, e->stream) != hipSuccess ||
          hipMemcpyAsync(csave[k], c->step_ctr, sizeof(int) * 4, hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
        rc = PFDD_E_HIP;
    }
    if (rc || hipStreamSynchronize(e->stream) != hipSuccess) {
      if (!rc) rc = PFDD_E_HIP;
      break;
    }
    saved = true;
    for (int round = 0; round < 2 && rc == PFDD_OK; round++) {
      if ((rc = time_mode(PFDD_ENS_PHASED, &e->ms_phased))) break;
      rc = time_mode(PFDD_ENS_PERSISTENT, &e->ms_persistent);
    }
  } while (0);
  if (saved) {
    const int r = restore();
    if (rc == PFDD_OK) rc = r;
  }
  // keep the faster mode (the phased one on any failure of the persistent step)
  if (rc == PFDD_OK && e->ms_persistent > 0.f && e->ms_persistent < e->ms_phased) {
    rc = pfdd_ensemble_set_mode(e, PFDD_ENS_PERSISTENT);
  } else {
    const int r = pfdd_ensemble_set_mode(e, PFDD_ENS_PHASED);
    for (size_t k = 0; k < e->cls.size(); k++) (void)apply_plan(e->cls[k], tuned[k]);
    if (rc == PFDD_OK) rc = r;
  }
  for (void* x : xsave)
    if (x) (void)hipFree(x);
  for (int* x : csave)
    if (x) (void)hipFree(x);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  e->prologue_done = 0;
  return rc;
}

int pfdd_ensemble_set_plan(pfdd_ens* e, int cls, int plan) {
  if (!e || cls < 0 || cls >= (int)e->cls.size()) return PFDD_E_SHAPE;
  ens_drop_graph(e);
  e->prologue_done = 0;
  HIPCK(hipStreamSynchronize(e->stream));
  return apply_plan(e->cls[cls], plan);
}

int pfdd_ensemble_sync(pfdd_ens* e) {
  if (!e) return PFDD_E_SHAPE;
  HIPCK(hipStreamSynchronize(e->stream));
  return e->pers_built ? pers_check(e) : PFDD_OK;
}

void* pfdd_ensemble_stream(pfdd_ens* e) { return e ? (void*)e->stream : nullptr; }

long long pfdd_ensemble_device_bytes(const pfdd_ens* e) {
  long long b = 0;
  if (e)
    for (const pfdd_ctx* c : e->cls) b += c->bytes;
  return b;
}

}  // extern "C"

// ---------------------------------------------------------------- FFT seam on host arrays
// Process-wide, like the static plans of G.c:124,170; not thread-safe.
namespace {
struct HostSeam {
  pfdd_ctx* ctx = nullptr;
  float2* buf = nullptr;
  size_t cap = 0;  // complex elements
  int status = PFDD_E_STATE;
  int dev = -1;    // device of ctx (current at pfdd_fft_host_init)
};
HostSeam g_seam;
// The seam's transform: fourn bit for bit when every edge is a power of two
// (the seam's context is then a PFDD_REF32 one), the FP32 Stockham passes
// otherwise (edges with a factor 3, which fourn does not take).
int seam_fft(pfdd_ctx* c, float2* buf, int batch, int dir) {
  return c->ref32 ? fourn3d(c, buf, batch, dir) : fft3d<float>(c, buf, batch, dir);
}

// Staging buffer of the allocation-free seam call (pfdd_fft_host_noalloc):
// a fixed device array, one per GPU for the life of the process.  In host
// code `scratch` is only a symbol; its device address on each GPU comes from
// hipGetSymbolAddress (taken once per device in pfdd_fft_host_init, with that
// device current) and is the only form passed to copies and kernels.
constexpr size_t NOALLOC_BYTES = (size_t)PFDD_NOALLOC_MAX_POINTS * sizeof(float2);
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int seam_run(float* data, int batch, int dir) {
  if (!g_seam.ctx) return PFDD_E_STATE;
  if (!data || batch < 1) return PFDD_E_SHAPE;
  pfdd_ctx* c = g_seam.ctx;
  const size_t n = (size_t)batch * c->P;
  if (n > g_seam.cap) {
    if (g_seam.buf) (void)hipFree(g_seam.buf);
    g_seam.buf = nullptr;
    g_seam.cap = 0;
    if (hipMalloc(&g_seam.buf, n * sizeof(float2)) != hipSuccess) return PFDD_E_ALLOC;
    g_seam.cap = n;
  }
  HIPCK(hipMemcpyAsync(g_seam.buf, data, n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  const int rc = seam_fft(c, g_seam.buf, batch, dir);
  if (rc) return rc;
  HIPCK(hipMemcpyAsync(data, g_seam.buf, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return PFDD_OK;
}
}  // namespace

extern "C" {

void pfdd_fft_host_release(void) {
  if (g_seam.buf) (void)hipFree(g_seam.buf);
  g_seam.buf = nullptr;
  g_seam.cap = 0;
  pfdd_destroy(g_seam.ctx);
  g_seam.ctx = nullptr;
  g_seam.status = PFDD_E_STATE;
}

int pfdd_fft_host_init(int n1, int n2, int n3) {
  pfdd_fft_host_release();
  pfdd_params p{};
  p.n1 = n1;
  p.n2 = n2;
  p.n3 = n3;
  p.ns = 1;
  const auto pow2 = [](int v) { return v > 1 && (v & (v - 1)) == 0; };
  p.precision = pow2(n1) && pow2(n2) && pow2(n3) ? PFDD_REF32 : PFDD_MIXED;
  p.bmode = PFDD_B_RECOMPUTE;
  p.C11 = 3.0;
  p.C12 = 1.0;
  p.C44 = 1.0;
  p.d1 = p.d2 = p.d3 = 1.0;
  p.cd2[0] = 0.1;
  p.dslip[0] = 1.0;
  p.uq2 = 1.0;
  p.pi_const = 3.141592654;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return g_seam.status = PFDD_E_HIP;
  // the allocation-free call's staging array on this device: its address, once per device
  if (dev >= 0 && dev < NOALLOC_MAX_DEVICES && !g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return g_seam.status = PFDD_E_HIP;
    g_scratch_ptr[dev] = ptr;
  }
  g_seam.status = pfdd_create(&p, &g_seam.ctx);
  g_seam.dev = g_seam.status == PFDD_OK ? dev : -1;
  return g_seam.status;
}

int pfdd_fft_host_noalloc(float* data, int batch, int dir) {
  if (!g_seam.ctx || g_seam.dev < 0 || g_seam.dev >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[g_seam.dev])
    return PFDD_E_STATE;
  if (!data || batch < 1) return PFDD_E_SHAPE;
  if (dir != PFDD_FORWARD && dir != PFDD_BACKWARD) return PFDD_E_SHAPE;
  pfdd_ctx* c = g_seam.ctx;
  if ((long long)batch * c->P > (long long)PFDD_NOALLOC_MAX_POINTS) return PFDD_E_LIMIT;
  const size_t n = (size_t)batch * c->P;
  // every element the transform reads is copied in first: nothing has to start at zero
  float2* buf = (float2*)g_scratch_ptr[g_seam.dev];
  HIPCK(hipMemcpyAsync(buf, data, n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  const int rc = seam_fft(c, buf, batch, dir);
  if (rc) return rc;
  HIPCK(hipMemcpyAsync(data, buf, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return PFDD_OK;
}

void fft_forward(float* data, int batch) { g_seam.status = seam_run(data, batch, PFDD_FORWARD); }
void fft_backward(float* data, int batch) { g_seam.status = seam_run(data, batch, PFDD_BACKWARD); }
int pfdd_fft_host_status(void) { return g_seam.status; }

}  // extern "C"

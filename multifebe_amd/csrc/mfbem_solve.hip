// Dense LU solve on MI355X (gfx950): the replacement of solve_lse_r / solve_lse_c
// (src/solve_lse_r.f90:25-232, src/solve_lse_c.f90:25-219), i.e. LAPACK ?getrf + ?getrs.
//
// Factorisation: blocked right-looking LU with partial pivoting, column-major, outer block NB.
//   * panel (m x NB) factored as narrow blocks of IB columns; a narrow block takes one kernel
//     launch per column: every workgroup scales its rows of the pivot column, applies the rank-1
//     update to the remaining columns of the narrow block and reports its local pivot candidate;
//     the last workgroup to finish (atomic counter) picks the global pivot (largest |re|+|im|,
//     lowest index on ties: LAPACK i?amax) and swaps the two rows of the narrow block.
//   * row interchanges: a run of swaps becomes its net list of row moves (k_swap_plan), applied
//     as a parallel gather over columns (k_apply_moves): per narrow block on the other panel
//     columns, per panel on every column outside the panel.
//   * the unit-lower and upper diagonal blocks are inverted (IB x IB leaves in LDS, then doubling
//     with batched GEMMs), so every triangular solve becomes a GEMM; the trailing update
//     A22 -= A21 A12 is the one large contraction and runs on MFMA through rocBLAS GEMM.
// Mixed precision (MFB_PREC_FP32_REFINE): the FP64 A is kept, an FP32 (complex FP32) copy is
// factored, and FP64 iterative refinement x += LU32 \ (b - A x) runs until the LAPACK zcgesv
// criterion ||r|| <= ||x|| ||A|| eps sqrt(n) holds; if it does not within 30 sweeps, or the
// residual stops decreasing, the FP64 LU solves instead.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <type_traits>
#include <utility>
#include <vector>

#include "mfbem.h"
#include "mfbem_internal.h"

namespace mfb {

struct cf {
  float re, im;
};
struct cz {
  double re, im;
};

template <typename T>
struct Tr;
template <>
struct Tr<double> {
  typedef double R;
  static constexpr bool cplx = false;
};
template <>
struct Tr<float> {
  typedef float R;
  static constexpr bool cplx = false;
};
template <>
struct Tr<cz> {
  typedef double R;
  static constexpr bool cplx = true;
};
template <>
struct Tr<cf> {
  typedef float R;
  static constexpr bool cplx = true;
};

__host__ __device__ inline double tabs1(double a) { return fabs(a); }
__host__ __device__ inline float tabs1(float a) { return fabsf(a); }
__host__ __device__ inline double tabs1(cz a) { return fabs(a.re) + fabs(a.im); }
__host__ __device__ inline float tabs1(cf a) { return fabsf(a.re) + fabsf(a.im); }
__host__ __device__ inline double tmod(double a) { return fabs(a); }
__host__ __device__ inline float tmod(float a) { return fabsf(a); }
__host__ __device__ inline double tmod(cz a) { return hypot(a.re, a.im); }
__host__ __device__ inline float tmod(cf a) { return hypotf(a.re, a.im); }
__host__ __device__ inline double tmul(double a, double b) { return a * b; }
__host__ __device__ inline float tmul(float a, float b) { return a * b; }
__host__ __device__ inline cz tmul(cz a, cz b) { return cz{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline cf tmul(cf a, cf b) { return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline double tsub(double a, double b) { return a - b; }
__host__ __device__ inline float tsub(float a, float b) { return a - b; }
__host__ __device__ inline cz tsub(cz a, cz b) { return cz{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cf tsub(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline double tadd(double a, double b) { return a + b; }
__host__ __device__ inline float tadd(float a, float b) { return a + b; }
__host__ __device__ inline cz tadd(cz a, cz b) { return cz{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cf tadd(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
// reciprocal (Smith's algorithm for the complex case)
__host__ __device__ inline double trcp(double a) { return 1.0 / a; }
__host__ __device__ inline float trcp(float a) { return 1.0f / a; }
template <typename C, typename R>
__host__ __device__ inline C crcp(C a) {
  if (fabs((double)a.re) >= fabs((double)a.im)) {
    R r = a.im / a.re, d = a.re + a.im * r;
    return C{(R)1 / d, -r / d};
  }
  R r = a.re / a.im, d = a.re * r + a.im;
  return C{r / d, (R)-1 / d};
}
__host__ __device__ inline cz trcp(cz a) { return crcp<cz, double>(a); }
__host__ __device__ inline cf trcp(cf a) { return crcp<cf, float>(a); }
template <typename T>
__host__ __device__ inline T tzero();
template <> __host__ __device__ inline double tzero<double>() { return 0.0; }
template <> __host__ __device__ inline float tzero<float>() { return 0.0f; }
template <> __host__ __device__ inline cz tzero<cz>() { return cz{0.0, 0.0}; }
template <> __host__ __device__ inline cf tzero<cf>() { return cf{0.0f, 0.0f}; }
template <typename T>
__host__ __device__ inline T tone();
template <> __host__ __device__ inline double tone<double>() { return 1.0; }
template <> __host__ __device__ inline float tone<float>() { return 1.0f; }
template <> __host__ __device__ inline cz tone<cz>() { return cz{1.0, 0.0}; }
template <> __host__ __device__ inline cf tone<cf>() { return cf{1.0f, 0.0f}; }
__host__ __device__ inline bool tiszero(double a) { return a == 0.0; }
__host__ __device__ inline bool tiszero(float a) { return a == 0.0f; }
__host__ __device__ inline bool tiszero(cz a) { return a.re == 0.0 && a.im == 0.0; }
__host__ __device__ inline bool tiszero(cf a) { return a.re == 0.0f && a.im == 0.0f; }

constexpr int NB = 256;     // outer block: width of the trailing GEMM contraction
constexpr int IB = 32;      // narrow block of the panel
constexpr int ROWS = 256;   // rows per workgroup in the column kernel

// ------------------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------------------
struct PivotScratch {
  double* val;   // per-workgroup candidate |a|
  int* idx;      // per-workgroup candidate row
  unsigned* counter;
  int* info;     // first zero pivot (global 1-based column), 0 if none
};

// pivot candidate order: larger |a|_1 wins, the lower row on ties (LAPACK i?amax)
__device__ inline void cand_max(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__device__ inline void wave_cand_max(double& v, int& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double v2 = __shfl_xor(v, off);
    const int i2 = __shfl_xor(i, off);
    cand_max(v, i, v2, i2);
  }
}

// One step of the narrow-block factorisation.  B: narrow block (m x w, ld), rows relative to its
// top-left corner; one thread per row.  Step j >= 0: scale column j below the pivot and update
// columns j+1..w-1 (pivot row staged in LDS, the thread's row held in registers).  Then
// (j + 1 < w): choose the pivot of column j+1 and swap it into row j+1 (last workgroup).
// ipiv: 0-based pivot rows relative to the block; gcol0: global index of column 0 (for info).
template <typename T>
__global__ void __launch_bounds__(ROWS) k_col_step(T* B, long long ld, int m, int w, int j,
                                                   int* ipiv, PivotScratch ps, int gcol0) {
  constexpr int NW = ROWS / 64;
  const int tid = threadIdx.x;
  const int r = blockIdx.x * ROWS + tid;
  const int jn = j + 1;
  __shared__ T prow[IB];
  __shared__ double wv[NW];
  __shared__ int wi[NW];
  __shared__ bool last;
  __shared__ int psel;
  if (j >= 0) {
    if (tid < w) prow[tid] = B[j + ld * tid];
    __syncthreads();
  }
  T x[IB];
  const bool act = r < m && r > j;   // rows below the current pivot row
  const int c0 = j > 0 ? j : 0;
  if (act) {
#pragma unroll
    for (int c = 0; c < IB; c++)
      if (c >= c0 && c < w) x[c] = B[r + ld * c];
  }
  if (j >= 0 && act) {
    const T piv = prow[j];
    if (!tiszero(piv)) {
      const T inv = trcp(piv);
      T l = tzero<T>();
#pragma unroll
      for (int c = 0; c < IB; c++)
        if (c == j) l = tmul(x[c], inv);
#pragma unroll
      for (int c = 0; c < IB; c++) {
        if (c == j) x[c] = l;
        else if (c > j && c < w) x[c] = tsub(x[c], tmul(l, prow[c]));
      }
#pragma unroll
      for (int c = 0; c < IB; c++)
        if (c >= j && c < w) B[r + ld * c] = x[c];
    }
  }
  if (jn >= w) return;
  double v = -1.0;
  int vi = 0x7fffffff;
  if (act) {
#pragma unroll
    for (int c = 0; c < IB; c++)
      if (c == jn) v = (double)tabs1(x[c]);
    vi = r;
  }
  wave_cand_max(v, vi);
  if ((tid & 63) == 0) {
    wv[tid >> 6] = v;
    wi[tid >> 6] = vi;
  }
  __syncthreads();
  if (tid == 0) {
    double bv = wv[0];
    int bi = wi[0];
    for (int q = 1; q < NW; q++) cand_max(bv, bi, wv[q], wi[q]);
    ps.val[blockIdx.x] = bv;
    ps.idx[blockIdx.x] = bi;
    __threadfence();
    const unsigned done = atomicAdd(ps.counter, 1u);
    last = (done == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  v = -1.0;
  vi = 0x7fffffff;
  for (int g = tid; g < (int)gridDim.x; g += ROWS)
    cand_max(v, vi, __hip_atomic_load(&ps.val[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
             __hip_atomic_load(&ps.idx[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  wave_cand_max(v, vi);
  __syncthreads();
  if ((tid & 63) == 0) {
    wv[tid >> 6] = v;
    wi[tid >> 6] = vi;
  }
  __syncthreads();
  if (tid == 0) {
    double bv = wv[0];
    int bi = wi[0];
    for (int q = 1; q < NW; q++) cand_max(bv, bi, wv[q], wi[q]);
    if (bi == 0x7fffffff) bi = jn;
    ipiv[jn] = bi;
    psel = bi;
    if (bv == 0.0 && *ps.info == 0) *ps.info = gcol0 + jn + 1;
    *ps.counter = 0;
  }
  __syncthreads();
  const int p = psel;
  if (p != jn && tid < w) {
    const T a = B[jn + ld * tid];
    B[jn + ld * tid] = B[p + ld * tid];
    B[p + ld * tid] = a;
  }
}

// Persistent narrow-block factorisation: the whole m x w block (w <= IB) in one launch (grid
// within the co-resident capacity), one thread per row holding its row in registers for all w columns.  Column j: every
// workgroup posts its pivot candidate (|a|_1, row) with the candidate's whole row, and workgroup
// 0 posts row j; every workgroup then waits for all posts of column j, picks the same winner
// (LAPACK i?amax order), the owner of the pivot row takes row j's content, row j takes the pivot
// row, and the rows below are scaled and updated.  The arithmetic is that of k_col_step.
//
// Synchronisation without a shared counter: a workgroup's post is a record (tag, |a|_1) plus the
// candidate row, written with relaxed agent-scope atomics (coherent across XCDs without cache
// write-back or invalidation); the data is written first, the workgroup waits until it is
// performed (s_waitcnt), and then writes the tag = (launch epoch, column, row).  Readers poll the
// tags (one thread per workgroup slot) and read the data after seeing the tag.  Posts alternate
// between two halves by column parity: a workgroup posts into half (j & 1) again only at column
// j + 2, after it has seen every post of column j + 1, and each workgroup posts column j + 1 only
// after its reads of column j.  Every spin-wait is bounded: on expiry the thread sets *err and
// goes on, so the grid always drains (the host reports the error).
constexpr int PANEL_GMAX = 256;   // workgroups of the persistent panel (rows <= 65,536)

// Every workgroup's record sits on its own 256-byte line: posts and polls of different
// workgroups never touch the same line.
constexpr int REC = 32;   // record stride in 8-byte words: tag, |a|_1, padding

struct PanelSync {
  unsigned long long* rec;   // [2][PANEL_GMAX][REC]: word 0 = epoch << 32 | column << 16 | row,
                             // word 1 = candidate |a|_1 (double bits)
  void* crow;                // [2][PANEL_GMAX][IB] candidate rows (T; FP32 types: stamped words)
  void* rowj;                // [2][IB] row j (T; FP32 types: stamped words)
  unsigned* err;             // time-out flag
  unsigned spin;             // spin limit of a wait (MFB_SPIN_LIMIT; expiry sets *err)
};

// Agent-coherent stores and loads of the posted values (relaxed atomics at agent scope).
__device__ inline void post_val(float* p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void post_val(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void post_val(cf* p, cf v) { post_val(&p->re, v.re); post_val(&p->im, v.im); }
__device__ inline void post_val(cz* p, cz v) { post_val(&p->re, v.re); post_val(&p->im, v.im); }
__device__ inline float fetch_val(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double fetch_val(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline cf fetch_val(const cf* p) { return cf{fetch_val(&p->re), fetch_val(&p->im)}; }
__device__ inline cz fetch_val(const cz* p) { return cz{fetch_val(&p->re), fetch_val(&p->im)}; }

// a matrix entry read once (non-temporal: the residual GEMV streams A past the caches; the
// same loads made the FP64 -> FP32 conversion 13 % slower and the triangular solves no faster)
__device__ inline double stream_ld(const double* p) { return __builtin_nontemporal_load(p); }
__device__ inline float stream_ld(const float* p) { return __builtin_nontemporal_load(p); }
__device__ inline cz stream_ld(const cz* p) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 w = __builtin_nontemporal_load((const d2*)p);
  return cz{w.x, w.y};
}
__device__ inline cf stream_ld(const cf* p) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 w = __builtin_nontemporal_load((const f2*)p);
  return cf{w.x, w.y};
}

// candidate order of the panel in the arithmetic's own real type (|a|_1 of a complex float is a
// float: comparing floats picks the same pivot as comparing their exact double values)
template <typename R>
__device__ inline void cand_max_t(R& v, int& i, R v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}
// Wave arg-max in pivot order through the wave-reduction library (DPP, no LDS round trips):
// the largest |a|_1 (non-negative, so its IEEE bits order like the values), then the lowest row
// among the lanes that hold it.  v < 0 marks a lane without a candidate; a wave without any
// candidate returns v = -1, i = INT_MAX.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(
    unsigned long long);
extern "C" __device__ __attribute__((const)) int __ockl_wfred_min_i32(int);
__device__ inline unsigned long long key_bits(float v) { return __float_as_uint(v); }
__device__ inline unsigned long long key_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ inline void from_key(unsigned long long k, float& v) { v = __uint_as_float((unsigned)k); }
__device__ inline void from_key(unsigned long long k, double& v) { v = __longlong_as_double((long long)k); }

template <typename R>
__device__ inline void wave_cand_max_t(R& v, int& i) {
  const bool has = v >= (R)0;
  const unsigned long long key = has ? key_bits(v) : 0ull;
  const unsigned long long kmax = __ockl_wfred_max_u64(key);
  const int imin = __ockl_wfred_min_i32(has && key == kmax ? i : 0x7fffffff);
  if (imin == 0x7fffffff) {
    v = (R)-1;
  } else {
    from_key(kmax, v);
  }
  i = imin;
}

template <typename T>
struct PanelCtx {
  typedef typename Tr<T>::R R;
  int tid, g, G, r, m, w, gcol0;
  unsigned epoch;
  int* ipiv;
  int* info;
  PanelSync ps;
  T* prow;    // LDS [IB]
  T* srowj;   // LDS [IB]
  R* wv;      // LDS [ROWS / 64]
  int* wi;    // LDS [ROWS / 64]
  T* stage;   // LDS [2 * IB]: candidate row, row J (posted one element per lane)
};

// Column J of the persistent panel for FP32 / complex FP32 entries: every posted 32-bit value
// travels in one 64-bit word with a 32-bit stamp (launch epoch x 32 + J; a slot set is reused two
// columns later, and only after every reader has left it), so a reader that sees the stamp has the
// value: no store has to be performed before another (no s_waitcnt round trip between the data
// and a tag).  Records: word 0 = stamp | |a|_1 bits, word 1 = stamp | candidate row; candidate
// rows and row J as stamped floats.  Pivot order and arithmetic are those of the FP64 protocol.
template <typename T, int J>
__device__ __attribute__((always_inline)) inline void panel_col_s(T (&x)[IB], const PanelCtx<T>& k) {
  constexpr int NW = ROWS / 64;
  constexpr int par = J & 1;
  constexpr int NF = sizeof(T) / sizeof(float);   // floats per entry
  constexpr int RW = IB * NF;                      // stamped words per row
  static_assert(64 + RW <= 128 && RW <= 64, "posting lanes");
  const int tid = k.tid, g = k.g, r = k.r, m = k.m;
  // bit 31 set: never equal to the upper word of the FP64 protocol's tags (the epoch itself)
  const unsigned long long st =
      (unsigned long long)(0x80000000u | ((k.epoch * 32u + (unsigned)J) & 0x7fffffffu)) << 32;
  unsigned long long* crow = (unsigned long long*)k.ps.crow;   // [2][PANEL_GMAX][RW]
  unsigned long long* rowj = (unsigned long long*)k.ps.rowj;   // [2][RW]
  auto put = [](unsigned long long* a, unsigned long long v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  // a stamped word of this column (bounded spin; on expiry the flag is set and the word used)
  auto take = [&](const unsigned long long* a) {
    unsigned long long w = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while ((w & 0xffffffff00000000ull) != st) {
      if (++spins > k.ps.spin) {
        atomicOr(k.ps.err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      w = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return w;
  };
  // local candidate of column J among rows J .. m-1
  float v = -1.0f;
  int vi = 0x7fffffff;
  if (r < m && r >= J) {
    v = tabs1(x[J]);
    vi = r;
  }
  wave_cand_max_t(v, vi);
  if ((tid & 63) == 0) {
    k.wv[tid >> 6] = v;
    k.wi[tid >> 6] = vi;
  }
  __syncthreads();
  float bv = k.wv[0];
  int bi = k.wi[0];
#pragma unroll
  for (int q = 1; q < NW; q++) cand_max_t(bv, bi, k.wv[q], k.wi[q]);
  if (r == bi) {
#pragma unroll
    for (int c = 0; c < IB; c++) k.stage[c] = x[c];
  }
  if (r == J) {   // workgroup 0
#pragma unroll
    for (int c = 0; c < IB; c++) k.stage[IB + c] = x[c];
  }
  __syncthreads();
  const float* sf = (const float*)k.stage;
  if (tid < RW)
    put(crow + ((size_t)par * PANEL_GMAX + g) * RW + tid, st | __float_as_uint(sf[tid]));
  if (g == 0 && tid >= 64 && tid < 64 + RW)
    put(rowj + par * RW + (tid - 64), st | __float_as_uint(sf[RW + tid - 64]));
  unsigned long long* myrec = k.ps.rec + ((size_t)par * PANEL_GMAX + g) * REC;
  if (tid == 128) {
    put(myrec, st | __float_as_uint(bv));
    put(myrec + 1, st | (unsigned)bi);
  }
  // every workgroup's record of column J (thread t reads slot t), then the winner
  v = -1.0f;
  vi = 0x7fffffff;
  if (tid < k.G) {
    const unsigned long long* rt = k.ps.rec + ((size_t)par * PANEL_GMAX + tid) * REC;
    const unsigned long long a = take(rt), b = take(rt + 1);
    v = __uint_as_float((unsigned)a);
    vi = (int)(unsigned)b;
    if (v < 0.0f) vi = 0x7fffffff;   // a workgroup without rows at or below J
  }
  wave_cand_max_t(v, vi);
  if ((tid & 63) == 0) {
    k.wv[tid >> 6] = v;
    k.wi[tid >> 6] = vi;
  }
  __syncthreads();
  float pv = k.wv[0];
  int p = k.wi[0];
#pragma unroll
  for (int q = 1; q < NW; q++) cand_max_t(pv, p, k.wv[q], k.wi[q]);
  if (p < J || p >= m) p = J;   // no candidate (or data of a timed-out wait)
  if (g == 0 && tid == 0) {
    k.ipiv[J] = p;
    if (pv == 0.0f && *k.info == 0) *k.info = k.gcol0 + J + 1;
  }
  // the pivot row and row J, one stamped word per lane
  if (tid < RW) {
    float* pf = (float*)k.prow;
    pf[tid] = __uint_as_float((unsigned)take(crow + ((size_t)par * PANEL_GMAX + p / ROWS) * RW + tid));
    if (p / ROWS == g && p != J)
      ((float*)k.srowj)[tid] = __uint_as_float((unsigned)take(rowj + par * RW + tid));
  }
  __syncthreads();
  if (r == p && p != J) {
#pragma unroll
    for (int c = 0; c < IB; c++) x[c] = k.srowj[c];
  }
  if (r == J) {
#pragma unroll
    for (int c = 0; c < IB; c++) x[c] = k.prow[c];
  }
  if (r < m && r > J) {
    const T piv = k.prow[J];
    if (!tiszero(piv)) {
      const T l = tmul(x[J], trcp(piv));
      x[J] = l;
#pragma unroll
      for (int c = J + 1; c < IB; c++) x[c] = tsub(x[c], tmul(l, k.prow[c]));
    }
  }
}

// Column J of the persistent panel (J a compile-time constant, so every register index of the
// thread's row x is one too).
template <typename T, int J>
__device__ __attribute__((always_inline)) inline void panel_col(T (&x)[IB], const PanelCtx<T>& k) {
  typedef typename Tr<T>::R R;
  constexpr int NW = ROWS / 64;
  constexpr int par = J & 1;
  if (J >= k.w) return;   // uniform: the whole grid skips the same columns
  if constexpr (std::is_same<R, float>::value) {
    panel_col_s<T, J>(x, k);
    return;
  }
  const int tid = k.tid, g = k.g, r = k.r, m = k.m;
  const unsigned long long stamp = ((unsigned long long)k.epoch << 32) | ((unsigned)J << 16);
  T* crow = (T*)k.ps.crow;
  T* rowj = (T*)k.ps.rowj;
  // local candidate of column J among rows J .. m-1
  R v = (R)-1;
  int vi = 0x7fffffff;
  if (r < m && r >= J) {
    v = tabs1(x[J]);
    vi = r;
  }
  wave_cand_max_t(v, vi);
  if ((tid & 63) == 0) {
    k.wv[tid >> 6] = v;
    k.wi[tid >> 6] = vi;
  }
  __syncthreads();
  // every thread combines the waves' candidates (no second barrier)
  R bv = k.wv[0];
  int bi = k.wi[0];
#pragma unroll
  for (int q = 1; q < NW; q++) cand_max_t(bv, bi, k.wv[q], k.wi[q]);
  // post the data (whole rows: columns beyond w are zeros), staged through LDS so that every
  // lane stores one element; wait until it is performed, then the tag
  if (r == bi) {
#pragma unroll
    for (int c = 0; c < IB; c++) k.stage[c] = x[c];
  }
  if (r == J) {   // workgroup 0
#pragma unroll
    for (int c = 0; c < IB; c++) k.stage[IB + c] = x[c];
  }
  __syncthreads();
  if (tid < IB && bi != 0x7fffffff)
    post_val(crow + ((size_t)par * PANEL_GMAX + g) * IB + tid, k.stage[tid]);
  if (g == 0 && tid >= 64 && tid < 64 + IB)
    post_val(rowj + par * IB + (tid - 64), k.stage[IB + tid - 64]);
  unsigned long long* myrec = k.ps.rec + ((size_t)par * PANEL_GMAX + g) * REC;
  if (tid == 128) post_val((double*)(myrec + 1), (double)bv);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0)
    __hip_atomic_store(myrec, stamp | (unsigned)(bi & 0xffff), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // wait for every workgroup's post of column J (thread t watches slot t), then the winner
  v = (R)-1;
  vi = 0x7fffffff;
  if (tid < k.G) {
    const unsigned long long* rt = k.ps.rec + ((size_t)par * PANEL_GMAX + tid) * REC;
    unsigned long long tg = 0;
    unsigned spins = 0;
    while (((tg = __hip_atomic_load(rt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 16) !=
           (stamp >> 16)) {
      if (++spins > k.ps.spin) {
        atomicOr(k.ps.err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    v = (R)fetch_val((const double*)(rt + 1));
    vi = (int)(tg & 0xffff);
    if (v < (R)0) vi = 0x7fffffff;   // a workgroup without rows at or below J
  }
  wave_cand_max_t(v, vi);
  if ((tid & 63) == 0) {
    k.wv[tid >> 6] = v;
    k.wi[tid >> 6] = vi;
  }
  __syncthreads();
  // every thread picks the same winner (no second barrier)
  R pv = k.wv[0];
  int p = k.wi[0];
#pragma unroll
  for (int q = 1; q < NW; q++) cand_max_t(pv, p, k.wv[q], k.wi[q]);
  if (p < J || p >= m) p = J;   // no candidate (or data of a timed-out wait)
  if (g == 0 && tid == 0) {
    k.ipiv[J] = p;
    if (pv == (R)0 && *k.info == 0) *k.info = k.gcol0 + J + 1;
  }
  // the pivot row and row J, one element per lane (both loads in flight together)
  if (tid < IB) {
    k.prow[tid] = fetch_val(crow + ((size_t)par * PANEL_GMAX + p / ROWS) * IB + tid);
    if (p / ROWS == g && p != J) k.srowj[tid] = fetch_val(rowj + par * IB + tid);
  }
  __syncthreads();
  if (r == p && p != J) {
#pragma unroll
    for (int c = 0; c < IB; c++) x[c] = k.srowj[c];
  }
  if (r == J) {
#pragma unroll
    for (int c = 0; c < IB; c++) x[c] = k.prow[c];
  }
  if (r < m && r > J) {
    const T piv = k.prow[J];
    if (!tiszero(piv)) {
      const T l = tmul(x[J], trcp(piv));
      x[J] = l;
#pragma unroll
      for (int c = J + 1; c < IB; c++) x[c] = tsub(x[c], tmul(l, k.prow[c]));
    }
  }
}

template <typename T, int... Js>
__device__ __attribute__((always_inline)) inline void panel_cols(T (&x)[IB], const PanelCtx<T>& k,
                                                                 std::integer_sequence<int, Js...>) {
  (panel_col<T, Js>(x, k), ...);
}

template <typename T>
__global__ void __launch_bounds__(ROWS) k_panel(T* B, long long ld, int m, int w, int* ipiv,
                                                PanelSync ps, int* info, int gcol0,
                                                unsigned epoch) {
  typedef typename Tr<T>::R R;
  __shared__ T prow[IB], srowj[IB], stage[2 * IB];
  __shared__ R wv[ROWS / 64];
  __shared__ int wi[ROWS / 64];
  PanelCtx<T> k;
  k.tid = threadIdx.x;
  k.g = blockIdx.x;
  k.G = gridDim.x;
  k.r = k.g * ROWS + k.tid;
  k.m = m;
  k.w = w;
  k.gcol0 = gcol0;
  k.epoch = epoch;
  k.ipiv = ipiv;
  k.info = info;
  k.ps = ps;
  k.prow = prow;
  k.srowj = srowj;
  k.wv = wv;
  k.wi = wi;
  k.stage = stage;
  const int r = k.r;
  // the thread's row, zero beyond column w (the updates run unmasked over IB columns)
  T x[IB];
#pragma unroll
  for (int c = 0; c < IB; c++) x[c] = (r < m && c < w) ? B[r + ld * c] : tzero<T>();
  panel_cols<T>(x, k, std::make_integer_sequence<int, IB>{});
  if (r < m) {
#pragma unroll
    for (int c = 0; c < IB; c++)
      if (c < w) B[r + ld * c] = x[c];
  }
}

// Row interchanges.  A run of LAPACK-order swaps (row R <-> row piv(R), R = r0 .. r0+cnt-1, in
// increasing order; piv(R) = floor(R / IB) * IB + ipiv[R] because each narrow block stores its
// pivots relative to its first row) is turned into the net row moves it produces, dst <- src.
// Rows of the run are slots 0..cnt-1; a pivot row below the run gets slot cnt + (index of the
// first swap that names it).  One workgroup; cnt <= NB.
template <int MAXS>
__global__ void __launch_bounds__(256) k_swap_plan(const int* ipiv, int r0, int cnt, int* moves,
                                                   int* nmoves) {
  __shared__ int slot[MAXS], rowof[2 * MAXS], content[2 * MAXS], cnt_s;
  const int t = threadIdx.x;
  for (int i = t; i < 2 * cnt; i += blockDim.x) { content[i] = i; rowof[i] = -1; }
  if (t == 0) cnt_s = 0;
  __syncthreads();
  for (int i = t; i < cnt; i += blockDim.x) {
    const int R = r0 + i, g = R / IB * IB + ipiv[R];
    rowof[i] = R;
    if (g < r0 + cnt) {
      slot[i] = g - r0;
    } else {
      int f = i;
      for (int j = 0; j < i; j++)
        if (ipiv[r0 + j] + (r0 + j) / IB * IB == g) { f = j; break; }
      slot[i] = cnt + f;
      if (f == i) rowof[cnt + i] = g;
    }
  }
  __syncthreads();
  if (t == 0)
    for (int i = 0; i < cnt; i++) {
      const int a = content[i];
      content[i] = content[slot[i]];
      content[slot[i]] = a;
    }
  __syncthreads();
  for (int sl = t; sl < 2 * cnt; sl += blockDim.x) {
    if (content[sl] != sl) {
      const int q = atomicAdd(&cnt_s, 1);
      moves[q] = rowof[sl];
      moves[2 * MAXS + q] = rowof[content[sl]];
    }
  }
  __syncthreads();
  if (t == 0) *nmoves = cnt_s;
}

// Apply the row moves of k_swap_plan to the columns [c0, c1) and [c2, c3): every workgroup owns
// CPC consecutive columns, reads all moved elements of them into registers, then writes them.
template <typename T, int MAXS, int CPC>
__global__ void __launch_bounds__(256) k_apply_moves(T* A, long long lda, const int* moves,
                                                     const int* nmoves, int c0, int c1, int c2,
                                                     int c3) {
  constexpr int PER = (2 * MAXS * CPC + 255) / 256;
  const int nm = *nmoves;
  const int n1 = c1 - c0, ntot = n1 + (c3 - c2);
  const int cb = blockIdx.x * CPC;
  T v[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = threadIdx.x + 256 * q, cc = e / (2 * MAXS), mi = e % (2 * MAXS);
    if (mi < nm && cb + cc < ntot) {
      const int c = cb + cc < n1 ? c0 + cb + cc : c2 + (cb + cc - n1);
      v[q] = A[moves[2 * MAXS + mi] + lda * (long long)c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = threadIdx.x + 256 * q, cc = e / (2 * MAXS), mi = e % (2 * MAXS);
    if (mi < nm && cb + cc < ntot) {
      const int c = cb + cc < n1 ? c0 + cb + cc : c2 + (cb + cc - n1);
      A[moves[mi] + lda * (long long)c] = v[q];
    }
  }
}

// Shuffle within a 32-lane half-wave.
template <typename T>
__device__ inline T shfl32(T v, int src);
template <> __device__ inline float shfl32<float>(float v, int src) { return __shfl(v, src, 32); }
template <> __device__ inline double shfl32<double>(double v, int src) { return __shfl(v, src, 32); }
template <> __device__ inline cf shfl32<cf>(cf v, int src) {
  return cf{__shfl(v.re, src, 32), __shfl(v.im, src, 32)};
}
template <> __device__ inline cz shfl32<cz>(cz v, int src) {
  return cz{__shfl(v.re, src, 32), __shfl(v.im, src, 32)};
}

// Interchanges of one narrow block, then the block's unit-lower solve, in one launch: the run of
// w <= IB swaps at rows r0 .. r0+w-1 (pivots relative to r0, r0 a multiple of IB; the rules of
// k_swap_plan) is applied to the panel columns [c0, c1) and [c2, c3), and on [c2, c3) the top w
// rows then become L^{-1} X (L = the narrow block's unit lower w x w factor at L, ld ldl).  Every
// workgroup builds the move list in LDS.  One 32-lane half-wave per column: all moved elements and
// the column's top w rows (as they are after the moves) are read before a barrier, then the
// moves below the top rows are written and the top rows solved by forward substitution (lane r
// holds x_r and row r of L; step t broadcasts the final x_t: x_r -= L[r][t] x_t for r > t) and
// written.
template <typename T>
__global__ void __launch_bounds__(256) k_swap_trsm(T* A, long long lda, const int* ipiv, int r0,
                                                   int w, const T* L, long long ldl, int c0, int c1,
                                                   int c2, int c3) {
  static_assert(IB == 32, "one half-wave per column");
  __shared__ int slot[IB], rowof[2 * IB], content[2 * IB], mdst[2 * IB], msrc[2 * IB], top[IB];
  __shared__ int nm_s;
  const int t = threadIdx.x;
  if (t < 2 * w) {
    content[t] = t;
    rowof[t] = -1;
  }
  if (t < IB) top[t] = r0 + t;
  if (t == 0) nm_s = 0;
  __syncthreads();
  if (t < w) {
    const int R = r0 + t, g = r0 + ipiv[R];
    rowof[t] = R;
    if (g < r0 + w) {
      slot[t] = g - r0;
    } else {
      int f = t;
      for (int j = 0; j < t; j++)
        if (r0 + ipiv[r0 + j] == g) { f = j; break; }
      slot[t] = w + f;
      if (f == t) rowof[w + t] = g;
    }
  }
  __syncthreads();
  if (t == 0)
    for (int i = 0; i < w; i++) {
      const int a = content[i];
      content[i] = content[slot[i]];
      content[slot[i]] = a;
    }
  __syncthreads();
  if (t < 2 * w && content[t] != t) {
    const int q = atomicAdd(&nm_s, 1);
    mdst[q] = rowof[t];
    msrc[q] = rowof[content[t]];
    if (t < w) top[t] = rowof[content[t]];   // slot t < w is row r0 + t
  }
  __syncthreads();
  const int nm = nm_s;
  const int n1 = c1 - c0, ntot = n1 + (c3 - c2);
  const int cc = blockIdx.x * 8 + (t >> 5), r = t & 31;
  if (cc >= ntot) return;   // whole half-waves leave together; no barrier follows
  const bool solve = cc >= n1;
  T* col = A + lda * (long long)(solve ? c2 + (cc - n1) : c0 + cc);
  T v0 = tzero<T>(), v1 = tzero<T>(), x = tzero<T>();
  if (r < nm) v0 = col[msrc[r]];
  if (r + 32 < nm) v1 = col[msrc[r + 32]];
  if (solve && r < w) x = col[top[r]];
  // all reads of this column (this half-wave's lanes only) are complete before its writes
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const int lim = solve ? r0 + w : -1;   // solved top rows are written below, not moved
  if (r < nm && mdst[r] >= lim) col[mdst[r]] = v0;
  if (r + 32 < nm && mdst[r + 32] >= lim) col[mdst[r + 32]] = v1;
  if (!solve) return;
  T Lr[IB];
#pragma unroll
  for (int q = 0; q < IB; q++) Lr[q] = (r < w && q < r) ? L[r + ldl * q] : tzero<T>();
#pragma unroll
  for (int q = 0; q < IB - 1; q++) {
    const T xq = shfl32(x, q);
    if (r > q) x = tsub(x, tmul(Lr[q], xq));
  }
  if (r < w) col[r0 + r] = x;
}

// Inverses of the IB x IB diagonal sub-blocks of the nb x nb triangular block at A (unit lower
// when UPPER is false, upper otherwise), written into the diagonal of Inv (ld ldi).  Two
// workgroups of 16 half-waves per sub-block: half-wave c computes column c of the inverse by substitution on
// e_c, lane r holding x_r and row r of the sub-block in registers, the solved x_t broadcast by a
// shuffle (lower: forward, x_r -= L_rt x_t for r > t; upper: backward, x_t /= U_tt, then
// x_r -= U_rt x_t for r < t).  blockIdx.z selects one of several such blocks, at A + z sa and
// Inv + z si.
template <typename T, bool UPPER>
__global__ void __launch_bounds__(IB * 16) k_inv_leaf(const T* A, long long lda, int nb, T* Inv,
                                                      long long ldi, long long sa, long long si) {
  static_assert(IB == 32, "one half-wave per column");
  A += sa * blockIdx.z;
  Inv += si * blockIdx.z;
  const int o = blockIdx.x * IB, w = min(IB, nb - o);   // blockIdx.y: which 16 columns
  const int c = blockIdx.y * 16 + (threadIdx.x >> 5), r = threadIdx.x & 31;
  T a[IB];
#pragma unroll
  for (int t = 0; t < IB; t++) {
    const bool in = r < w && t < w && (UPPER ? t > r : t < r);
    a[t] = in ? A[o + r + lda * (long long)(o + t)] : tzero<T>();
  }
  T x = (r == c && r < w) ? tone<T>() : tzero<T>();
  if (!UPPER) {
#pragma unroll
    for (int t = 0; t < IB - 1; t++) {
      const T xt = shfl32(x, t);
      if (r > t) x = tsub(x, tmul(a[t], xt));
    }
  } else {
    const T dinv = (r < w) ? trcp(A[o + r + lda * (long long)(o + r)]) : tzero<T>();
#pragma unroll
    for (int t = IB - 1; t >= 0; t--) {
      if (r == t) x = tmul(x, dinv);
      const T xt = shfl32(x, t);
      if (r < t) x = tsub(x, tmul(a[t], xt));
    }
  }
  if (r < w && c < w) Inv[o + r + ldi * (long long)(o + c)] = x;
}

// dst (m x n, ld) = 0
template <typename T>
__global__ void k_zero_mat(T* dst, long long ld, int m, int n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < (long long)m * n; t += (long long)gridDim.x * blockDim.x)
    dst[t % m + ld * (t / m)] = tzero<T>();
}

template <typename Ts, typename Td>
__device__ inline Td conv(Ts a);
template <> __device__ inline float conv<double, float>(double a) { return (float)a; }
template <> __device__ inline double conv<float, double>(float a) { return (double)a; }
template <> __device__ inline cf conv<cz, cf>(cz a) { return cf{(float)a.re, (float)a.im}; }
template <> __device__ inline cz conv<cf, cz>(cf a) { return cz{(double)a.re, (double)a.im}; }
template <> __device__ inline double conv<double, double>(double a) { return a; }
template <> __device__ inline cz conv<cz, cz>(cz a) { return a; }

// dst(m x n, ldd) = conv(src(m x n, lds))
template <typename Ts, typename Td>
__global__ void k_convert(const Ts* src, long long lds, Td* dst, long long ldd, int m, int n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)m * n;
  for (; t < tot; t += (long long)gridDim.x * blockDim.x) {
    long long c = t / m, r = t % m;
    dst[r + ldd * c] = conv<Ts, Td>(src[r + lds * c]);
  }
}

// y[i] = x[perm[i]] for every right-hand side
template <typename T>
__global__ void k_permute(const T* x, long long ldx, T* y, long long ldy, const int* perm, int n,
                          int nrhs) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * nrhs) return;
  int c = (int)(t / n), i = (int)(t % n);
  y[i + ldy * c] = x[perm[i] + ldx * c];
}

// x += conv(d)
template <typename Tl, typename Th>
__global__ void k_axpy_conv(const Tl* d, Th* x, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = tadd(x[t], conv<Tl, Th>(d[t]));
}

// out[0] = max |v_i| (modulus) with a block reduction + atomic max on the bit pattern (v >= 0)
template <typename T>
__global__ void k_maxabs(const T* v, long long n, unsigned long long* out) {
  __shared__ double s[256];
  double m = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    double a = (double)tmod(v[t]);
    if (!(a <= m)) m = a;  // NaN propagates
  }
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      double a = s[threadIdx.x], b = s[threadIdx.x + k];
      if (!(b <= a)) s[threadIdx.x] = b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double r = s[0];
    if (r != r) r = INFINITY;
    atomicMax(out, (unsigned long long)__double_as_longlong(r));
  }
}

// power-of-two column scale c_j ~ 1 / max_i |a_ij| (exact in floating point); one block per column
template <typename T>
__global__ void k_colscale(const T* A, long long lda, int n, double* cs) {
  __shared__ double s[256];
  int c = blockIdx.x;
  double m = 0;
  for (int r = threadIdx.x; r < n; r += 256) {
    double a = (double)tmod(A[r + lda * (long long)c]);
    if (a > m) m = a;
  }
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k && s[threadIdx.x + k] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) cs[c] = s[0] > 0 ? exp2(-round(log2(s[0]))) : 1.0;
}

template <typename T>
__device__ inline T tscale(T a, double f);
template <> __device__ inline double tscale<double>(double a, double f) { return a * f; }
template <> __device__ inline cz tscale<cz>(cz a, double f) { return cz{a.re * f, a.im * f}; }

// dst(n x n) = conv(A diag(cs)), and rs[r] += sum over the block's columns of |a_rc| cs_c (rs
// zeroed before; ||A diag(cs)||_inf is its largest entry).  Grid (row blocks of 256, chunks of
// CV_COLS columns): reads and writes coalesced over rows, one atomic add per row and chunk.
constexpr int CV_COLS = 128;
// |a| without hypot's scaling (the entries of an assembled matrix are far from overflow)
__device__ inline double tmod_fast(double a) { return fabs(a); }
__device__ inline double tmod_fast(cz a) { return sqrt(a.re * a.re + a.im * a.im); }
template <typename Ts, typename Td>
__global__ void __launch_bounds__(256) k_convert_scaled(const Ts* src, long long lds, Td* dst,
                                                        long long ldd, int n, const double* cs,
                                                        double* rs) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * CV_COLS, c1 = min(n, c0 + CV_COLS);
  if (r >= n) return;
  double sum = 0;
  const Ts* sp = src + r + lds * (long long)c0;
  Td* dp = dst + r + ldd * (long long)c0;
  int c = c0;
  // eight columns at a time: eight loads in flight per thread
  for (; c + 8 <= c1; c += 8, sp += 8 * lds, dp += 8 * ldd) {
    Ts a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = sp[lds * u];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const Ts b = tscale(a[u], cs[c + u]);
      sum += tmod_fast(b);
      dp[ldd * u] = conv<Ts, Td>(b);
    }
  }
  for (; c < c1; c++, sp += lds, dp += ldd) {
    const Ts b = tscale(*sp, cs[c]);
    sum += tmod_fast(b);
    *dp = conv<Ts, Td>(b);
  }
  atomicAdd(&rs[r], sum);
}

// x += diag(cs) conv(d), column-major n x nrhs
template <typename Tl, typename Th>
__global__ void k_axpy_scaled(const Tl* d, Th* x, int n, long long tot, const double* cs) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < tot) x[t] = tadd(x[t], tscale(conv<Tl, Th>(d[t]), cs[t % n]));
}

// y = diag(1/cs) x (scaled unknowns, for the convergence test)
template <typename T>
__global__ void k_unscale(const T* x, T* y, int n, long long tot, const double* cs) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < tot) y[t] = tscale(x[t], 1.0 / cs[t % n]);
}

// Few right-hand sides (nrhs <= SMALL_RHS): triangular sweeps and residuals without library GEMMs.
constexpr int SMALL_RHS = 4;
constexpr int GV_COLS = 512;   // columns per partial sum of the residual GEMV

// y = Inv t for one NB block (Inv: jb x jb inside an NB x NB array); grid.y = right-hand side
template <typename T>
__global__ void __launch_bounds__(NB) k_blk_matvec(const T* Inv, int jb, const T* t, long long ldt,
                                                   T* y, long long ldy) {
  __shared__ T st[NB];
  const int q = blockIdx.y, i = threadIdx.x;
  if (i < jb) st[i] = t[i + ldt * q];
  __syncthreads();
  if (i >= jb) return;
  T acc = tzero<T>();
  for (int j = 0; j < jb; j++) acc = tadd(acc, tmul(Inv[i + (long long)NB * j], st[j]));
  y[i + ldy * q] = acc;
}

// t[0:m) -= A[0:m, 0:kk) y[0:kk), kk <= NB; one thread per row, grid (row blocks, rhs)
template <typename T>
__global__ void __launch_bounds__(256) k_gemv_sub(const T* A, long long lda, int m, int kk,
                                                  const T* y, long long ldy, T* t, long long ldt) {
  __shared__ T sy[NB];
  const int q = blockIdx.y;
  for (int j = threadIdx.x; j < kk; j += 256) sy[j] = y[j + ldy * q];
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m) return;
  T acc = tzero<T>();
  for (int j = 0; j < kk; j++) acc = tadd(acc, tmul(A[r + lda * (long long)j], sy[j]));
  t[r + ldt * q] = tsub(t[r + ldt * q], acc);
}

// Blocked triangular solve with one right-hand side in one persistent launch (grid within the
// co-resident capacity; workgroups are numbered in solve order, so each waits only on
// lower-indexed, earlier dispatched, ones): one workgroup owns
// row block i (NB rows; TRSV_CG threads per row, each over a quarter of a block's columns, their
// partial sums added in a fixed order).  It subtracts the contributions of the blocks it
// depends on (lower: k = 0 .. i-1, upper: k = nb-1 .. i+1, i.e. the order of the launch-per-
// block sweep) as each block's solution is published, then applies the inverse of its diagonal
// block and publishes its own solution: the values with agent-coherent stores, then (after
// s_waitcnt) a tag = launch epoch.  Spin-waits are bounded (on expiry *err is set and the
// workgroup goes on).
struct SolveSync {
  unsigned long long* rec;   // [PANEL_GMAX][REC] word 0: epoch of the block's published solution
  unsigned* err;
  unsigned spin;             // spin limit of a wait (MFB_SPIN_LIMIT; expiry sets *err)
};

constexpr int TRSV_CG = 4;                  // column groups per row (memory parallelism)
constexpr int TRSV_THREADS = NB * TRSV_CG;  // one thread per (row, column group)

template <typename T, bool UPPER>
__global__ void __launch_bounds__(TRSV_THREADS) k_trsv_blocked(const T* A, long long lda, int n,
                                                               const T* Inv, const T* M,
                                                               const T* in, T* out, SolveSync ss,
                                                               unsigned epoch) {
  constexpr int CW = NB / TRSV_CG;   // columns per group
  __shared__ T sy[NB];
  __shared__ T part[TRSV_CG][NB];
  // the upper solve runs from the last block: workgroup 0 takes it, so that in both triangles a
  // workgroup waits only on lower-indexed (earlier dispatched) workgroups
  const int nb = gridDim.x, i = UPPER ? nb - 1 - (int)blockIdx.x : (int)blockIdx.x;
  const int lr = threadIdx.x % NB, q = threadIdx.x / NB;
  const int r = i * NB + lr;
  const bool valid = r < n;
  T t = (valid && q == 0) ? in[r] : tzero<T>();   // the running right-hand side (group 0)
  // row lr of the block at a0 (leading dimension ld) times the columns [q CW, (q+1) CW) of sy
  auto partial = [&](const T* a0, long long ld, int jb) {
    T acc = tzero<T>();
    const int j0 = q * CW, j1 = min(jb, j0 + CW);
    const T* a = a0 + ld * j0;
#pragma unroll 16
    for (int j = j0; j < j1; j++, a += ld) acc = tadd(acc, tmul(*a, sy[j]));
    return acc;
  };
  auto combine = [&]() {
    T acc = part[0][lr];
#pragma unroll
    for (int g = 1; g < TRSV_CG; g++) acc = tadd(acc, part[g][lr]);
    return acc;
  };
  // wait for block k's published solution and stage it in sy
  auto fetch_block = [&](int k) {
    if (threadIdx.x == 0) {
      const unsigned long long* tg = ss.rec + (size_t)k * REC;
      unsigned spins = 0;
      while (__hip_atomic_load(tg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++spins > ss.spin) {
          atomicOr(ss.err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    const int c0 = k * NB, jb = min(NB, n - c0);
    if (threadIdx.x < jb) sy[threadIdx.x] = fetch_val(out + c0 + threadIdx.x);
    __syncthreads();
    return jb;
  };
  const int cnt = UPPER ? nb - 1 - i : i;   // blocks this one depends on
  // all but the nearest block: t -= A_ik y_k
  for (int st = 0; st + 1 < cnt; st++) {
    const int k = UPPER ? nb - 1 - st : st;
    const int jb = fetch_block(k);
    part[q][lr] = valid ? partial(A + r + lda * (long long)(k * NB), lda, jb) : tzero<T>();
    __syncthreads();
    if (q == 0 && valid) t = tsub(t, combine());
  }
  // z = Inv_ii t (before the nearest block is known), then y_i = z - M_i y_near with
  // M_i = Inv_ii A_i,near: one matrix-vector product on the chain from block to block
  const int jbi = min(NB, n - i * NB);
  if (q == 0) sy[lr] = valid ? t : tzero<T>();
  __syncthreads();
  part[q][lr] = valid ? partial(Inv + (size_t)NB * NB * i + lr, (long long)NB, jbi) : tzero<T>();
  __syncthreads();
  T z = tzero<T>();
  if (q == 0 && valid) z = combine();
  if (cnt > 0) {
    __syncthreads();   // part and sy are reused
    const int k = UPPER ? i + 1 : i - 1;
    const int jb = fetch_block(k);
    part[q][lr] = valid ? partial(M + (size_t)NB * NB * i + lr, (long long)NB, jb) : tzero<T>();
    __syncthreads();
    if (q == 0 && valid) z = tsub(z, combine());
  }
  if (q == 0 && valid) post_val(out + r, z);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(ss.rec + (size_t)i * REC, (unsigned long long)epoch, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// The same solve with each row block split over SPLIT (4 or 2) workgroups of NB / SPLIT rows
// (SPLIT^2 column groups of NB / SPLIT^2 columns per row): every matrix-vector product of a
// block step reads 1 / SPLIT of its block per workgroup, SPLIT CUs' bandwidth per row block
// instead of one.  The right-hand side of the diagonal-inverse product needs the whole block's
// reduced vector, so the parts exchange it once per block (posted values, then a tag per part)
// before y_{i-1} is needed; y_i is published per part (tag slot SPLIT i + part), reduced-vector
// tags in slots PANEL_GMAX + SPLIT i + part.  Grid = SPLIT x (row blocks) workgroups.
template <typename T, bool UPPER, int TRSV_SPLIT>
__global__ void __launch_bounds__(TRSV_THREADS) k_trsv_split(const T* A, long long lda, int n,
                                                             const T* Inv, const T* M,
                                                             const T* in, T* out, T* rbuf,
                                                             SolveSync ss, unsigned epoch) {
  constexpr int TRSV_QR = NB / TRSV_SPLIT;               // rows per workgroup
  constexpr int TRSV_SCG = TRSV_THREADS / TRSV_QR;       // column groups per row
  constexpr int CW = NB / TRSV_SCG;   // columns per group
  __shared__ T sy[NB];
  __shared__ T part[TRSV_SCG][TRSV_QR];
  const int nb = gridDim.x / TRSV_SPLIT, qd = blockIdx.x % TRSV_SPLIT;
  const int i = UPPER ? nb - 1 - (int)(blockIdx.x / TRSV_SPLIT) : (int)(blockIdx.x / TRSV_SPLIT);
  const int lr = threadIdx.x % TRSV_QR, q = threadIdx.x / TRSV_QR;
  const int rb = qd * TRSV_QR + lr;   // row within the block
  const int r = i * NB + rb;
  const bool valid = r < n;
  T t = (valid && q == 0) ? in[r] : tzero<T>();
  auto partial = [&](const T* a0, long long ld, int jb) {
    T acc = tzero<T>();
    const int j0 = q * CW, j1 = min(jb, j0 + CW);
    const T* a = a0 + ld * j0;
#pragma unroll 16
    for (int j = j0; j < j1; j++, a += ld) acc = tadd(acc, tmul(*a, sy[j]));
    return acc;
  };
  auto combine = [&]() {
    T acc = part[0][lr];
#pragma unroll
    for (int g = 1; g < TRSV_SCG; g++) acc = tadd(acc, part[g][lr]);
    return acc;
  };
  // wait for the TRSV_SPLIT tags at slots base + TRSV_SPLIT k + (0 .. TRSV_SPLIT-1), then stage
  // jb values of src (block k) in sy
  auto fetch = [&](int base, int k, const T* src) {
    if (threadIdx.x < TRSV_SPLIT) {
      const unsigned long long* tg = ss.rec + (size_t)(base + TRSV_SPLIT * k + threadIdx.x) * REC;
      unsigned spins = 0;
      while (__hip_atomic_load(tg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++spins > ss.spin) {
          atomicOr(ss.err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    const int c0 = k * NB, jb = min(NB, n - c0);
    if (threadIdx.x < jb) sy[threadIdx.x] = fetch_val(src + c0 + threadIdx.x);
    __syncthreads();
    return jb;
  };
  // values first, s_waitcnt, then the tag
  auto publish = [&](T* dst, T v, int slot) {
    if (q == 0 && valid) post_val(dst, v);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(ss.rec + (size_t)slot * REC, (unsigned long long)epoch, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  };
  const int cnt = UPPER ? nb - 1 - i : i;
  for (int st = 0; st + 1 < cnt; st++) {
    const int k = UPPER ? nb - 1 - st : st;
    const int jb = fetch(0, k, out);
    part[q][lr] = valid ? partial(A + r + lda * (long long)(k * NB), lda, jb) : tzero<T>();
    __syncthreads();
    if (q == 0 && valid) t = tsub(t, combine());
  }
  // exchange the reduced right-hand side of block i between its parts
  publish(rbuf + r, t, PANEL_GMAX + TRSV_SPLIT * i + qd);
  const int jbi = fetch(PANEL_GMAX, i, rbuf);
  part[q][lr] = valid ? partial(Inv + (size_t)NB * NB * i + rb, (long long)NB, jbi) : tzero<T>();
  __syncthreads();
  T z = tzero<T>();
  if (q == 0 && valid) z = combine();
  if (cnt > 0) {
    const int k = UPPER ? i + 1 : i - 1;
    const int jb = fetch(0, k, out);   // its first barrier orders the reads of part and sy above
    part[q][lr] = valid ? partial(M + (size_t)NB * NB * i + rb, (long long)NB, jb) : tzero<T>();
    __syncthreads();
    if (q == 0 && valid) z = tsub(z, combine());
  }
  publish(out + r, z, TRSV_SPLIT * i + qd);
}

// component c of an FP32 / complex FP32 value
__device__ inline float fpart(float v, int) { return v; }
__device__ inline float fpart(cf v, int c) { return c ? v.im : v.re; }

// k_trsv_split for FP32 / complex FP32 entries with stamped posts and operands read ahead.
// Every posted 32-bit value (the solution y, the reduced right-hand side exchanged between the
// parts of a block) travels in one 64-bit word whose upper half is the launch epoch: a reader that
// sees the epoch has the value, so a post is a plain store (no s_waitcnt, no barrier, no tag) and
// a wait is one load per word.  ys: [n][NF] words of y, rs: [NB x blocks][NF] words of the reduced
// right-hand sides.  Each matrix block's part of a workgroup is loaded into registers before the
// wait for the vector it multiplies, so the loads of M_i (the product on the chain from block to
// block) are in flight, or done, while y_{i -/+ 1} is still being solved.  The sums run in the
// order of k_trsv_split.  Grid = SPLIT x (row blocks) workgroups of THREADS, all co-resident;
// every wave covers 64 rows of one column group.
template <typename T, bool UPPER, int TRSV_SPLIT, int THREADS>
__global__ void __launch_bounds__(THREADS) k_trsv_stamped(const T* A, long long lda, int n,
                                                               const T* Inv, const T* M,
                                                               const T* in, T* out,
                                                               unsigned long long* ys,
                                                               unsigned long long* rs,
                                                               SolveSync ss, unsigned epoch) {
  constexpr int NF = sizeof(T) / sizeof(float);          // floats per entry
  constexpr int TRSV_QR = NB / TRSV_SPLIT;               // rows per workgroup
  constexpr int TRSV_SCG = THREADS / TRSV_QR;            // column groups per row
  constexpr int CW = NB / TRSV_SCG;                      // columns per group
  static_assert(NB * NF <= THREADS && TRSV_QR % 64 == 0, "one posted word per thread");
  __shared__ T sy[NB];
  __shared__ T part[TRSV_SCG][TRSV_QR];
  const int nb = gridDim.x / TRSV_SPLIT, qd = blockIdx.x % TRSV_SPLIT;
  const int i = UPPER ? nb - 1 - (int)(blockIdx.x / TRSV_SPLIT) : (int)(blockIdx.x / TRSV_SPLIT);
  const int lr = threadIdx.x % TRSV_QR, q = threadIdx.x / TRSV_QR;
  const int rb = qd * TRSV_QR + lr;   // row within the block
  const int r = i * NB + rb;
  const bool valid = r < n;
  const unsigned long long st = (unsigned long long)epoch << 32;
  T t = (valid && q == 0) ? in[r] : tzero<T>();
  T a[CW];
  // row `off` of the block whose column 0 is at a0 (leading dimension ld), columns
  // [q CW, (q+1) CW) below jb; q is one value per wave, so every column address is scalar
  auto load = [&](const T* a0, long long ld, int off, int jb) {
    const int j0 = __builtin_amdgcn_readfirstlane(q) * CW;
    const T* p = a0 + ld * j0;
#pragma unroll
    for (int jj = 0; jj < CW; jj++)
      a[jj] = (valid && j0 + jj < jb) ? p[ld * jj + off] : tzero<T>();
  };
  auto dot = [&]() {
    T acc = tzero<T>();
#pragma unroll
    for (int jj = 0; jj < CW; jj++) acc = tadd(acc, tmul(a[jj], sy[q * CW + jj]));
    return acc;
  };
  auto combine = [&]() {
    T acc = part[0][lr];
#pragma unroll
    for (int g = 1; g < TRSV_SCG; g++) acc = tadd(acc, part[g][lr]);
    return acc;
  };
  // block k of a stamped vector into sy (zeros past its end), then one barrier; sy and part are
  // free here: every read of them precedes the barrier ahead of the caller's combine
  auto fetch = [&](const unsigned long long* src, int k) {
    const int c0 = k * NB, jb = min(NB, n - c0);
    const int w = threadIdx.x;
    if (w < NB * NF) {
      float v = 0.0f;
      if (w < jb * NF) {
        const unsigned long long* pw = src + (size_t)c0 * NF + w;
        unsigned long long x = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while ((x & 0xffffffff00000000ull) != st) {
          if (++spins > ss.spin) {
            atomicOr(ss.err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          x = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        v = __uint_as_float((unsigned)x);
      }
      ((float*)sy)[w] = v;
    }
    __syncthreads();
    return jb;
  };
  auto post = [&](unsigned long long* dst, T v) {
#pragma unroll
    for (int c = 0; c < NF; c++)
      __hip_atomic_store(dst + (size_t)r * NF + c, st | __float_as_uint(fpart(v, c)),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  const int cnt = UPPER ? nb - 1 - i : i;
  for (int s_ = 0; s_ + 1 < cnt; s_++) {
    const int k = UPPER ? nb - 1 - s_ : s_;
    load(A + lda * (long long)(k * NB), lda, r, min(NB, n - k * NB));
    fetch(ys, k);
    part[q][lr] = dot();
    __syncthreads();
    if (q == 0 && valid) t = tsub(t, combine());
  }
  // the reduced right-hand side of block i from all its parts, then z = Inv_ii t
  if (q == 0 && valid) post(rs, t);
  const int jbi = min(NB, n - i * NB);
  load(Inv + (size_t)NB * NB * i, (long long)NB, rb, jbi);
  fetch(rs, i);
  part[q][lr] = dot();
  __syncthreads();
  T z = tzero<T>();
  if (q == 0 && valid) z = combine();
  if (cnt > 0) {
    const int k = UPPER ? i + 1 : i - 1;
    load(M + (size_t)NB * NB * i, (long long)NB, rb, min(NB, n - k * NB));
    fetch(ys, k);
    part[q][lr] = dot();
    __syncthreads();
    if (q == 0 && valid) z = tsub(z, combine());
  }
  if (q == 0 && valid) {
    out[r] = z;
    post(ys, z);
  }
}

// part[(q * nch + ch) * n + r] = sum over the columns of chunk ch of A[r, c] x[c, q]
template <typename T>
__global__ void __launch_bounds__(256) k_gemv_partial(const T* A, long long lda, int n, const T* x,
                                                      long long ldx, T* part) {
  __shared__ T sx[GV_COLS];
  const int ch = blockIdx.y, q = blockIdx.z, nch = gridDim.y;
  const int c0 = ch * GV_COLS, nc = min(GV_COLS, n - c0);
  for (int j = threadIdx.x; j < nc; j += 256) sx[j] = x[c0 + j + ldx * q];
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const T* a = A + r + lda * (long long)c0;
  T acc = tzero<T>();
  for (int j = 0; j < nc; j++) acc = tadd(acc, tmul(stream_ld(a + lda * (long long)j), sx[j]));
  part[((long long)q * nch + ch) * n + r] = acc;
}

// res[r, q] = b[r, q] - sum_ch part (fixed order: repeatable to the last bit)
template <typename T>
__global__ void k_gemv_reduce(const T* part, int n, int nch, const T* b, long long ldb, T* res,
                              long long ldr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
  if (r >= n) return;
  T acc = tzero<T>();
  for (int ch = 0; ch < nch; ch++) acc = tadd(acc, part[((long long)q * nch + ch) * n + r]);
  res[r + ldr * q] = tsub(b[r + ldb * q], acc);
}

// ------------------------------------------------------------------------------------------------
// rocBLAS GEMM dispatch (column-major, no transposes): C = alpha A B + beta C
// ------------------------------------------------------------------------------------------------
static rocblas_status gemm(rocblas_handle h, int m, int n, int k, double alpha, const double* A,
                           long long lda, const double* B, long long ldb, double beta, double* C,
                           long long ldc) {
  return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda,
                       B, ldb, &beta, C, ldc);
}
static rocblas_status gemm(rocblas_handle h, int m, int n, int k, double alpha, const float* A,
                           long long lda, const float* B, long long ldb, double beta, float* C,
                           long long ldc) {
  float a = (float)alpha, b = (float)beta;
  return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &a, A, lda, B,
                       ldb, &b, C, ldc);
}
static rocblas_status gemm(rocblas_handle h, int m, int n, int k, double alpha, const cz* A,
                           long long lda, const cz* B, long long ldb, double beta, cz* C,
                           long long ldc) {
  rocblas_double_complex a(alpha, 0.0), b(beta, 0.0);
  return rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &a,
                       (const rocblas_double_complex*)A, lda, (const rocblas_double_complex*)B, ldb,
                       &b, (rocblas_double_complex*)C, ldc);
}
static rocblas_status gemm(rocblas_handle h, int m, int n, int k, double alpha, const cf* A,
                           long long lda, const cf* B, long long ldb, double beta, cf* C,
                           long long ldc) {
  rocblas_float_complex a((float)alpha, 0.0f), b((float)beta, 0.0f);
  return rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &a,
                       (const rocblas_float_complex*)A, lda, (const rocblas_float_complex*)B, ldb,
                       &b, (rocblas_float_complex*)C, ldc);
}

// strided-batched variants (used for the triangular inverses)
static rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, double alpha, const double* A,
                              long long lda, long long sa, const double* B, long long ldb,
                              long long sb, double beta, double* C, long long ldc, long long sc,
                              int cnt) {
  return rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k,
                                       &alpha, A, lda, sa, B, ldb, sb, &beta, C, ldc, sc, cnt);
}
static rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, double alpha, const float* A,
                              long long lda, long long sa, const float* B, long long ldb,
                              long long sb, double beta, float* C, long long ldc, long long sc,
                              int cnt) {
  float a = (float)alpha, b = (float)beta;
  return rocblas_sgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k,
                                       &a, A, lda, sa, B, ldb, sb, &b, C, ldc, sc, cnt);
}
static rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, double alpha, const cz* A,
                              long long lda, long long sa, const cz* B, long long ldb,
                              long long sb, double beta, cz* C, long long ldc, long long sc,
                              int cnt) {
  rocblas_double_complex a(alpha, 0.0), b(beta, 0.0);
  return rocblas_zgemm_strided_batched(
      h, rocblas_operation_none, rocblas_operation_none, m, n, k, &a,
      (const rocblas_double_complex*)A, lda, sa, (const rocblas_double_complex*)B, ldb, sb, &b,
      (rocblas_double_complex*)C, ldc, sc, cnt);
}
static rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, double alpha, const cf* A,
                              long long lda, long long sa, const cf* B, long long ldb,
                              long long sb, double beta, cf* C, long long ldc, long long sc,
                              int cnt) {
  rocblas_float_complex a((float)alpha, 0.0f), b((float)beta, 0.0f);
  return rocblas_cgemm_strided_batched(
      h, rocblas_operation_none, rocblas_operation_none, m, n, k, &a,
      (const rocblas_float_complex*)A, lda, sa, (const rocblas_float_complex*)B, ldb, sb, &b,
      (rocblas_float_complex*)C, ldc, sc, cnt);
}

#define MFB_BLAS(call)                                                       \
  do {                                                                       \
    rocblas_status st_ = (call);                                             \
    if (st_ != rocblas_status_success) {                                     \
      set_error("%s failed: rocblas status %d (%s:%d)", #call, (int)st_,    \
                __FILE__, __LINE__);                                         \
      return MFB_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

// ------------------------------------------------------------------------------------------------
// Split FP16 trailing update of the FP32 and complex FP32 factorisations:
//   A22 -= A21 A12  as  A22 -= [Lh Lh Ll] [Rh; Rl; Rh]
// with x = hi + lo, hi = fp16(x), lo = fp16(x - hi) for every entry of A21 (L) and A12 (R): one
// FP16 GEMM with FP32 accumulation on MFMA over three times the panel width.  The dropped lo*lo
// product and the rounding of lo leave about 2^-21 |L| |R|, close to an FP32 product; the FP64
// refinement that follows every mixed factorisation absorbs it.  A complex operand becomes the
// real matrix acting on interleaved (re, im) rows: L entry -> [[ar, -ai], [ai, ar]], R entry ->
// [br; bi], so the complex FP32 A22 is updated in place as a real matrix of 2 nr rows (ld 2 lda).
// L21 entries are bounded by 1 (partial pivoting); U12 entries by the growth of the column-scaled
// matrix (columns scaled to a largest entry in [0.5, 1)): both lie far inside FP16's range, and
// entries below its normal range (6e-5) keep an absolute accuracy of 2^-24.
// ------------------------------------------------------------------------------------------------
template <typename T> struct RealW;
template <> struct RealW<float> { static constexpr int R = 1; };
template <> struct RealW<cf> { static constexpr int R = 2; };
template <typename T>
constexpr bool kSplitType = std::is_same<T, float>::value || std::is_same<T, cf>::value;

__device__ inline void split_h(float x, __half& h, __half& l) {
  h = __float2half_rn(x);
  l = __float2half_rn(x - __half2float(h));
}

// left operand: A21 (m x kb, lda) -> H (R m x 3 R kb, ld ldh), column blocks [hi | hi | lo]
__global__ void k_split_left(const float* A, long long lda, int m, int kb, __half* H,
                             long long ldh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
  if (i >= m) return;
  __half h, lo;
  split_h(A[i + lda * l], h, lo);
  H[i + ldh * l] = h;
  H[i + ldh * (kb + l)] = h;
  H[i + ldh * (2LL * kb + l)] = lo;
}
__global__ void k_split_left(const cf* A, long long lda, int m, int kb, __half* H, long long ldh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
  if (i >= m) return;
  const cf a = A[i + lda * l];
  __half rh, rl, ih, il;
  split_h(a.re, rh, rl);
  split_h(a.im, ih, il);
  for (int blk = 0; blk < 3; blk++) {
    const __half r = blk < 2 ? rh : rl, im = blk < 2 ? ih : il;
    __half* c0 = H + ldh * (2LL * kb * blk + 2LL * l) + 2LL * i;
    __half* c1 = c0 + ldh;
    c0[0] = r;
    c0[1] = im;
    c1[0] = __hneg(im);
    c1[1] = r;
  }
}
// C (m x n, ldc, FP32) := alpha A B + beta C, A and B FP16 (FP32 accumulation)
static rocblas_status gemm_f16(rocblas_handle h, int m, int n, int k, float alpha, const __half* A,
                               long long lda, const __half* B, long long ldb, float beta, float* C,
                               long long ldc) {
  return rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A,
                         rocblas_datatype_f16_r, (rocblas_int)lda, B, rocblas_datatype_f16_r,
                         (rocblas_int)ldb, &beta, C, rocblas_datatype_f32_r, (rocblas_int)ldc, C,
                         rocblas_datatype_f32_r, (rocblas_int)ldc, rocblas_datatype_f32_r,
                         rocblas_gemm_algo_standard, 0, 0);
}

// The split of B (kb x nc, ldb), row blocks of pw rows (the panels whose left operands are
// [hi | hi | lo] column blocks of Hl) each as [hi; lo; hi] in H (ld ldH), B also stored into S (ld
// lds): the U rows of a super-panel go from the work buffer they were computed in to the factor
// and to the FP16 operand in one pass
__global__ void k_split_right_store(const float* B, long long ldb, int kb, int pw, int nc,
                                    __half* H, long long ldH, float* S, long long lds) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (l >= kb || j >= nc) return;
  const int i = l / pw, li = l - i * pw, w = min(pw, kb - i * pw);
  const float b = B[l + ldb * j];
  S[l + lds * j] = b;
  __half h, lo;
  split_h(b, h, lo);
  __half* c = H + ldH * j + 3LL * pw * i + li;
  c[0] = h;
  c[w] = lo;
  c[2 * w] = h;
}
__global__ void k_split_right_store(const cf* B, long long ldb, int kb, int pw, int nc, __half* H,
                                    long long ldH, cf* S, long long lds) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (l >= kb || j >= nc) return;
  const int i = l / pw, li = l - i * pw, w = min(pw, kb - i * pw);
  const cf b = B[l + ldb * j];
  S[l + lds * j] = b;
  __half rh, rl, ih, il;
  split_h(b.re, rh, rl);
  split_h(b.im, ih, il);
  __half* c = H + ldH * j + 6LL * pw * i + 2 * li;
  c[0] = rh;
  c[1] = ih;
  c[2 * w] = rl;
  c[2 * w + 1] = il;
  c[4 * w] = rh;
  c[4 * w + 1] = ih;
}

// ------------------------------------------------------------------------------------------------
// factorisation object
// ------------------------------------------------------------------------------------------------
template <typename T>
struct LU {
  int n = 0;
  T* A = nullptr;          // factors (in place)
  long long lda = 0;
  int* ipiv = nullptr;     // device, 0-based global
  T* Linv = nullptr;       // inverses of the diagonal blocks, NB x NB each
  T* Uinv = nullptr;
  T* Lm = nullptr;         // Linv_i L_i,i-1 (NB x NB each; block 0 unused)
  T* Um = nullptr;         // Uinv_i U_i,i+1 (NB x NB each; the last block unused)
  T* work = nullptr;       // NB x n work panel
  std::vector<int> perm;   // row permutation P (b_perm[i] = b[perm[i]])
  int* dperm = nullptr;
  int info = 0;
  bool split16 = false;    // FP32 / complex FP32 factors: trailing updates as three FP16 products
};

struct Ctx {
  rocblas_handle blas = nullptr;
  PivotScratch ps{};
  int ps_cap = 0;
  int device = -1;
  PanelSync pan{};          // persistent-panel posting buffers
  unsigned pan_epoch = 0;   // launch counter of the persistent panel (tags of its posts)
  SolveSync sol{};          // published-block tags of the persistent triangular solves
  unsigned sol_epoch = 0;
  int trsv_grid = 0;        // largest co-resident grid of k_trsv_blocked (all instantiations)
  int trsv_split_grid[2] = {};  // the same for k_trsv_split with 4 and with 2 parts per block
  int trsv_stamp_grid[2] = {};  // the same for k_trsv_stamped (FP32 types; no tag slots, no cap)
  int panel_grid[2] = {};   // largest co-resident grid of k_panel<cf/float>, k_panel<cz/double>
  int panel_mode = -1;      // 1: persistent panels, 0: one launch per column
  // a persistent kernel's bounded wait expired (workgroups not co-resident): the panel of getrf
  // (the caller redoes the factorisation with one launch per column where it can) and the number
  // of such time-outs recovered so far (reported as fell_back bit 2)
  bool panel_timeout = false;
  long long recovered = 0;
  int* moves = nullptr;   // row moves of k_swap_plan: dst[2 NB], src[2 NB]
  int* nmoves = nullptr;
  // look-ahead: the trailing update beyond the next panel runs on a second stream (own rocBLAS
  // handle) while the next panel is factored; events order the two
  int lookahead = -1;       // 1: on (persistent panels only), 0: off (MFB_LOOKAHEAD=0)
  int la_main = 0;    // trailing-update columns on the caller's stream (MFB_LA_MAIN; 0: SB)
  int la_chunk = 0;   // columns per trailing-update GEMM on s2 (MFB_LA_CHUNK; 0: 768 for the
                     // split FP16 updates of depth 6 R NB, 2048 otherwise)
  hipStream_t s2 = nullptr;
  rocblas_handle blas2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // grow-only device workspaces of the factorisations and solves (no allocation, and no device-
  // wide synchronisation from hipFree, once a size has been seen)
  void* ws[27] = {};
  size_t ws_cap[27] = {};
};

enum WsSlot {
  WS_IPIV, WS_LINV, WS_UINV, WS_WORK, WS_DPERM, WS_TMP, WS_A32, WS_CS,
  WS_X, WS_Y, WS_R, WS_R32, WS_T32, WS_RS, WS_SCR, WS_PART, WS_LM, WS_UM, WS_WORK2, WS_RX,
  WS_HL, WS_HR, WS_HL2, WS_HR2, WS_PB, WS_YS, WS_RSS, WS_COUNT
};
static_assert(WS_COUNT == sizeof(Ctx::ws) / sizeof(void*), "one workspace per slot");

static Ctx& ctx() {
  static Ctx c;
  return c;
}

// Caller-chosen schedule of the dense solves (mfb_solver_set_schedule; -1: the environment's or
// the built-in choice): look-ahead on/off, columns per side-stream GEMM, most parts per row block
// of the persistent triangular solves.  Process-wide, read at the start of each factorisation or
// solve.
struct Schedule {
  int lookahead = -1;
  int la_chunk = -1;
  int trsv_parts = -1;
};
static Schedule& schedule() {
  static Schedule g;
  return g;
}

static int ensure_ctx(hipStream_t s, int n) {
  Ctx& c = ctx();
  int dev = 0;
  MFB_HIP(hipGetDevice(&dev));
  if (c.blas && c.device != dev) {
    rocblas_destroy_handle(c.blas);
    c.blas = nullptr;
    c.ps = PivotScratch{};
    c.ps_cap = 0;
    c.moves = nullptr;
    c.nmoves = nullptr;
    c.pan = PanelSync{};
    c.pan_epoch = 0;
    c.sol = SolveSync{};
    c.sol_epoch = 0;
    c.panel_mode = -1;
    // the second stream, handle and events belonged to the previous device
    if (c.blas2) rocblas_destroy_handle(c.blas2);
    c.blas2 = nullptr;
    c.s2 = nullptr;
    c.ev_fork = c.ev_join = nullptr;
    c.lookahead = -1;
    for (int k = 0; k < WS_COUNT; k++) {
      c.ws[k] = nullptr;
      c.ws_cap[k] = 0;
    }
  }
  if (!c.blas) {
    MFB_BLAS(rocblas_create_handle(&c.blas));
    c.device = dev;
  }
  MFB_BLAS(rocblas_set_stream(c.blas, s));
  int need = (n + ROWS - 1) / ROWS + 1;
  if (need > c.ps_cap) {
    if (c.ps.val) { (void)hipFree(c.ps.val); (void)hipFree(c.ps.idx); (void)hipFree(c.ps.counter); (void)hipFree(c.ps.info); }
    MFB_HIP(hipMalloc(&c.ps.val, sizeof(double) * need));
    MFB_HIP(hipMalloc(&c.ps.idx, sizeof(int) * need));
    MFB_HIP(hipMalloc(&c.ps.counter, sizeof(unsigned)));
    MFB_HIP(hipMalloc(&c.ps.info, sizeof(int)));
    MFB_HIP(hipMemset(c.ps.counter, 0, sizeof(unsigned)));
    c.ps_cap = need;
  }
  if (!c.moves) {
    MFB_HIP(hipMalloc(&c.moves, sizeof(int) * 4 * NB));
    MFB_HIP(hipMalloc(&c.nmoves, sizeof(int)));
  }
  if (!c.pan.rec) {
    const size_t rb = sizeof(unsigned long long) * 2 * PANEL_GMAX * REC;
    MFB_HIP(hipMalloc(&c.pan.rec, rb));
    MFB_HIP(hipMalloc(&c.pan.err, sizeof(unsigned)));
    MFB_HIP(hipMalloc(&c.pan.crow, sizeof(cz) * 2 * PANEL_GMAX * IB));
    MFB_HIP(hipMalloc(&c.pan.rowj, sizeof(cz) * 2 * IB));
    MFB_HIP(hipMemset(c.pan.crow, 0, sizeof(cz) * 2 * PANEL_GMAX * IB));
    MFB_HIP(hipMemset(c.pan.rowj, 0, sizeof(cz) * 2 * IB));
    MFB_HIP(hipMemset(c.pan.rec, 0, rb));
    MFB_HIP(hipMemset(c.pan.err, 0, sizeof(unsigned)));
    MFB_HIP(hipDeviceSynchronize());
    c.pan_epoch = 0;
  }
  if (!c.sol.rec) {
    // block tags (PANEL_GMAX slots) and the split solve's exchange tags (PANEL_GMAX more)
    const size_t rb = sizeof(unsigned long long) * 2 * PANEL_GMAX * REC;
    MFB_HIP(hipMalloc(&c.sol.rec, rb));
    MFB_HIP(hipMalloc(&c.sol.err, sizeof(unsigned)));
    MFB_HIP(hipMemset(c.sol.rec, 0, rb));
    MFB_HIP(hipMemset(c.sol.err, 0, sizeof(unsigned)));
    MFB_HIP(hipDeviceSynchronize());
    c.sol_epoch = 0;
  }
  {
    // spin limit of the persistent kernels' waits (MFB_SPIN_LIMIT=<spins>: forces time-outs in
    // tests of the recovery; default 2^22, ~0.2 s)
    const char* sl = getenv("MFB_SPIN_LIMIT");
    const unsigned spin = sl ? (unsigned)strtoul(sl, nullptr, 10) : (1u << 22);
    c.pan.spin = spin;
    c.sol.spin = spin;
  }
  if (c.panel_mode < 0) {
    // persistent kernels (panels, one-right-hand-side solves) are ordinary launches whose grid
    // fits co-resident (the occupancy bounds below) and whose waits are bounded.  Not cooperative
    // launches: on ROCm those wait for every other stream's workgroups to leave the CUs, and a
    // second process on the GPU could not make one while the first held the device.
    // MFB_PANEL=column selects the one-launch-per-column kernel.
    int cus = 0;
    MFB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const char* env = getenv("MFB_PANEL");
    c.panel_mode = !(env && strcmp(env, "column") == 0) ? 1 : 0;
    int b0 = 0, b1 = 0;
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, k_panel<cf>, ROWS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, k_panel<cz>, ROWS, 0));
    int f0 = 0, f1 = 0;
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&f0, k_panel<float>, ROWS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&f1, k_panel<double>, ROWS, 0));
    c.panel_grid[0] = std::min(PANEL_GMAX, std::min(b0, f0) * cus);
    c.panel_grid[1] = std::min(PANEL_GMAX, std::min(b1, f1) * cus);
    int t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t0, k_trsv_blocked<cz, false>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t1, k_trsv_blocked<cz, true>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t2, k_trsv_blocked<double, false>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t3, k_trsv_blocked<double, true>, TRSV_THREADS, 0));
    c.trsv_grid = std::min(PANEL_GMAX, std::min(std::min(t0, t1), std::min(t2, t3)) * cus);
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t0, k_trsv_split<cz, false, 4>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t1, k_trsv_split<cz, true, 4>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t2, k_trsv_split<double, false, 4>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t3, k_trsv_split<double, true, 4>, TRSV_THREADS, 0));
    c.trsv_split_grid[0] = std::min(PANEL_GMAX, std::min(std::min(t0, t1), std::min(t2, t3)) * cus);
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t0, k_trsv_split<cz, false, 2>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t1, k_trsv_split<cz, true, 2>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t2, k_trsv_split<double, false, 2>, TRSV_THREADS, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t3, k_trsv_split<double, true, 2>, TRSV_THREADS, 0));
    c.trsv_split_grid[1] = std::min(PANEL_GMAX, std::min(std::min(t0, t1), std::min(t2, t3)) * cus);
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t0, k_trsv_stamped<cf, false, 4, 512>, 512, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t1, k_trsv_stamped<cf, true, 4, 512>, 512, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t2, k_trsv_stamped<float, false, 4, 512>, 512, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t3, k_trsv_stamped<float, true, 4, 512>, 512, 0));
    c.trsv_stamp_grid[0] = std::min(std::min(t0, t1), std::min(t2, t3)) * cus;
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t0, k_trsv_stamped<cf, false, 2, 1024>, 1024, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t1, k_trsv_stamped<cf, true, 2, 1024>, 1024, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t2, k_trsv_stamped<float, false, 2, 1024>, 1024, 0));
    MFB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&t3, k_trsv_stamped<float, true, 2, 1024>, 1024, 0));
    c.trsv_stamp_grid[1] = std::min(std::min(t0, t1), std::min(t2, t3)) * cus;
  }
  if (c.lookahead < 0) {
    const char* env = getenv("MFB_LOOKAHEAD");
    c.lookahead = (c.panel_mode == 1 && !(env && strcmp(env, "0") == 0)) ? 1 : 0;
    const char* ch = getenv("MFB_LA_CHUNK");
    c.la_chunk = ch && atoi(ch) > 0 ? std::max(NB, atoi(ch) / NB * NB) : 0;   // 0: by type
    const char* lm = getenv("MFB_LA_MAIN");
    c.la_main = lm && atoi(lm) > 0 ? std::max(NB, atoi(lm) / NB * NB) : 0;   // 0: SB
    if (c.panel_mode == 1) {   // also with MFB_LOOKAHEAD=0: a schedule may switch it on
      MFB_HIP(hipStreamCreateWithFlags(&c.s2, hipStreamNonBlocking));
      MFB_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
      MFB_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
      MFB_BLAS(rocblas_create_handle(&c.blas2));
      MFB_BLAS(rocblas_set_stream(c.blas2, c.s2));
    }
  }
  return MFB_OK;
}

// device workspace `slot` of at least `bytes` (grown, never shrunk)
template <typename P>
static int ws_get(Ctx& c, int slot, size_t bytes, P** out) {
  if (c.ws_cap[slot] < bytes) {
    if (c.ws[slot]) (void)hipFree(c.ws[slot]);
    c.ws[slot] = nullptr;
    c.ws_cap[slot] = 0;
    MFB_HIP(hipMalloc(&c.ws[slot], bytes));
    c.ws_cap[slot] = bytes;
  }
  *out = reinterpret_cast<P*>(c.ws[slot]);
  return MFB_OK;
}

// the factor's arrays live in the context's workspaces: releasing it frees nothing
template <typename T>
static void lu_free(LU<T>& f) {
  f = LU<T>();
}

// Inverse of the jb x jb triangular diagonal block at P (unit lower, or upper) into Inv (NB x NB,
// ld NB): IB x IB leaf inverses, then doubling with strided-batched GEMMs,
//   lower: X21 = -inv(B22) B21 inv(B11);   upper: X12 = -inv(B11) B12 inv(B22).
// tmp holds NB * NB / 2 elements.
template <typename T>
static int tri_inverse(Ctx& c, const T* P, long long lda, int jb, T* Inv, bool upper, T* tmp,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_zero_mat<T>, dim3(64), dim3(256), 0, s, Inv, (long long)NB, NB, NB);
  if (upper)
    hipLaunchKernelGGL((k_inv_leaf<T, true>), dim3((jb + IB - 1) / IB, 2), dim3(IB * 16), 0, s,
                       P, lda, jb, Inv, (long long)NB, 0LL, 0LL);
  else
    hipLaunchKernelGGL((k_inv_leaf<T, false>), dim3((jb + IB - 1) / IB, 2), dim3(IB * 16), 0, s,
                       P, lda, jb, Inv, (long long)NB, 0LL, 0LL);
  for (int h = IB; h < jb; h *= 2) {
    const int full = jb / (2 * h);           // pairs with both halves of size h
    const int o_last = full * 2 * h;         // a trailing pair with a short second half
    const int h2 = std::min(h, jb - o_last - h);
    const long long sA = 2LL * h * (1 + lda), sI = 2LL * h * (1 + NB), sT = (long long)h * h;
    if (!upper) {
      if (full > 0) {
        MFB_BLAS(gemm_sb(c.blas, h, h, h, 1.0, P + h, lda, sA, Inv, NB, sI, 0.0, tmp, h, sT, full));
        MFB_BLAS(gemm_sb(c.blas, h, h, h, -1.0, Inv + h + (long long)NB * h, NB, sI, tmp, h, sT,
                         0.0, Inv + h, NB, sI, full));
      }
      if (h2 > 0) {
        const long long o = o_last;
        MFB_BLAS(gemm(c.blas, h2, h, h, 1.0, P + (o + h) + lda * o, lda, Inv + o + NB * o, NB,
                      0.0, tmp, h2));
        MFB_BLAS(gemm(c.blas, h2, h, h2, -1.0, Inv + (o + h) + NB * (o + h), NB, tmp, h2, 0.0,
                      Inv + (o + h) + NB * o, NB));
      }
    } else {
      if (full > 0) {
        MFB_BLAS(gemm_sb(c.blas, h, h, h, 1.0, P + lda * h, lda, sA, Inv + h + (long long)NB * h,
                         NB, sI, 0.0, tmp, h, sT, full));
        MFB_BLAS(gemm_sb(c.blas, h, h, h, -1.0, Inv, NB, sI, tmp, h, sT, 0.0,
                         Inv + (long long)NB * h, NB, sI, full));
      }
      if (h2 > 0) {
        const long long o = o_last;
        MFB_BLAS(gemm(c.blas, h, h2, h2, 1.0, P + o + lda * (o + h), lda,
                      Inv + (o + h) + NB * (o + h), NB, 0.0, tmp, h));
        MFB_BLAS(gemm(c.blas, h, h2, h, -1.0, Inv + o + NB * o, NB, tmp, h, 0.0,
                      Inv + o + NB * (o + h), NB));
      }
    }
  }
  return MFB_OK;
}

// Inverses of the first nfull NB x NB upper-triangular diagonal blocks of the factored A into
// Uinv (block b at Uinv + b NB^2, ld NB), all blocks at once: one leaf launch over every block,
// then per doubling level and pair position two strided-batched GEMMs over the blocks
// (X12 = -inv(B11) B12 inv(B22)).  tmp holds nfull NB^2 / 4 elements.  The U blocks are final
// once their panel is factored (later interchanges touch only rows below them), so this runs
// once after the factorisation instead of once per panel.
template <typename T>
static int tri_inverse_upper_blocks(Ctx& c, const T* A, long long lda, int nfull, T* Uinv, T* tmp,
                                    hipStream_t s) {
  if (nfull <= 0) return MFB_OK;
  hipLaunchKernelGGL(k_zero_mat<T>, dim3(1024), dim3(256), 0, s, Uinv, (long long)NB, NB,
                     NB * nfull);
  const long long sd = (long long)NB * (1 + lda), si = (long long)NB * NB;
  hipLaunchKernelGGL((k_inv_leaf<T, true>), dim3(NB / IB, 2, nfull), dim3(IB * 16), 0, s, A, lda,
                     NB, Uinv, (long long)NB, sd, si);
  for (int h = IB; h < NB; h *= 2) {
    const long long sT = (long long)h * h;
    for (int o = 0; o < NB; o += 2 * h) {
      MFB_BLAS(gemm_sb(c.blas, h, h, h, 1.0, A + o + lda * (long long)(o + h), lda, sd,
                       Uinv + (o + h) + (long long)NB * (o + h), NB, si, 0.0, tmp, h, sT, nfull));
      MFB_BLAS(gemm_sb(c.blas, h, h, h, -1.0, Uinv + o + (long long)NB * o, NB, si, tmp, h, sT,
                       0.0, Uinv + o + (long long)NB * (o + h), NB, si, nfull));
    }
  }
  return MFB_OK;
}

// Row moves of the swaps at rows [r0, r0 + cnt) applied to columns [c0, c1) and [c2, c3).
template <typename T>
static void apply_swaps(Ctx& c, T* A, long long lda, const int* ipiv, int r0, int cnt, int c0,
                        int c1, int c2, int c3, hipStream_t s) {
  const int ncols = (c1 - c0) + (c3 - c2);
  if (ncols <= 0 || cnt <= 0) return;
  if (cnt <= IB) {
    hipLaunchKernelGGL(k_swap_plan<IB>, dim3(1), dim3(256), 0, s, ipiv, r0, cnt, c.moves,
                       c.nmoves);
    constexpr int CPC = 16;
    hipLaunchKernelGGL((k_apply_moves<T, IB, CPC>), dim3((ncols + CPC - 1) / CPC), dim3(256), 0,
                       s, A, lda, c.moves, c.nmoves, c0, c1, c2, c3);
  } else {
    hipLaunchKernelGGL(k_swap_plan<NB>, dim3(1), dim3(256), 0, s, ipiv, r0, cnt, c.moves,
                       c.nmoves);
    constexpr int CPC = 2;
    hipLaunchKernelGGL((k_apply_moves<T, NB, CPC>), dim3((ncols + CPC - 1) / CPC), dim3(256), 0,
                       s, A, lda, c.moves, c.nmoves, c0, c1, c2, c3);
  }
}

// LU of the n x n matrix A (lda) in place.  f.A = A.  Returns MFB_OK; f.info > 0 on a zero pivot.
template <typename T>
static int getrf(LU<T>& f, T* A, long long lda, int n, hipStream_t s) {
  Ctx& c = ctx();
  int rc = ensure_ctx(s, n);
  if (rc) return rc;
  c.panel_timeout = false;
  f.n = n;
  f.A = A;
  f.lda = lda;
  const int nblk = (n + NB - 1) / NB;
  rc = ws_get(c, WS_IPIV, sizeof(int) * n, &f.ipiv);
  if (!rc) rc = ws_get(c, WS_LINV, sizeof(T) * (size_t)NB * NB * nblk, &f.Linv);
  if (!rc) rc = ws_get(c, WS_UINV, sizeof(T) * (size_t)NB * NB * nblk, &f.Uinv);
  // work buffers: U rows of a super-panel (up to 2 NB rows) for every column
  if (!rc) rc = ws_get(c, WS_WORK, sizeof(T) * (size_t)2 * NB * std::max(n, NB), &f.work);
  if (!rc) rc = ws_get(c, WS_DPERM, sizeof(int) * n, &f.dperm);
  // look-ahead (persistent panels): after panel k, the next panel's columns are updated on s and
  // the rest of the trailing matrix on s2, which runs beside the factorisation of panel k + 1
  // (that touches only its own columns); s waits for s2 before the interchanges of panel k + 1
  // reach the other columns.  The panel's workgroups start as s2's GEMM workgroups leave the CUs.
  const int la_req = schedule().lookahead >= 0 ? schedule().lookahead : c.lookahead;
  const bool la = la_req == 1 && c.panel_mode == 1 && c.s2;
  T* work2 = nullptr;
  if (!rc && la) rc = ws_get(c, WS_WORK2, sizeof(T) * (size_t)2 * NB * std::max(n, NB), &work2);
  if (rc) return rc;
  bool pending = false;
  MFB_HIP(hipMemsetAsync(c.ps.info, 0, sizeof(int), s));
  // Split FP16 factorisations take the panels two at a time: a super-panel of SB = 2 NB columns is
  // factored panel by panel (the second panel's columns updated from the first inside it), and
  // the columns beyond it get one trailing update of depth 2 NB (one FP16 GEMM of depth 6 R NB,
  // which runs ~30 % faster per flop than two of depth 3 R NB and reads and writes the FP32
  // trailing matrix half as often).  Other factorisations: SB = NB, one panel per super-panel.
  int SB = NB;
  if constexpr (kSplitType<T>)
    if (f.split16) SB = 2 * NB;
  int sp = 0;   // super-panel count (the split left operand alternates between two buffers)
  for (int k = 0; k < n; k += SB, sp++) {
    const int kend = std::min(n, k + SB);
    const int g0 = k + std::min(NB, n - k);   // first row of the split left operand
    __half* Hl = nullptr;
    long long ldh = 0;
    if constexpr (kSplitType<T>) {
      if (f.split16 && g0 < n) {
        constexpr int R = RealW<T>::R;
        // the buffer that the far update of super-panel sp - 1, still running on s2, does not read
        ldh = ((long long)R * (n - g0) + 63) / 64 * 64;
        rc = ws_get(c, (sp & 1) ? WS_HL2 : WS_HL, sizeof(__half) * (size_t)ldh * 3 * R * (kend - k),
                    &Hl);
        if (rc) return rc;
      }
    }
    // left split operand of panel k1 (width jb) from its rows r0 .. n (r0 >= k1 + jb): Hl's
    // column block of the panel, Hl's rows from r0
    auto split_left = [&](int k1, int jb, int r0) {
      if constexpr (kSplitType<T>) {
        constexpr int R = RealW<T>::R;
        if (Hl && r0 < n)
          hipLaunchKernelGGL(k_split_left, dim3((n - r0 + 255) / 256, jb), dim3(256), 0, s,
                             (const T*)(A + r0 + lda * (long long)k1), lda, n - r0, jb,
                             Hl + (long long)R * (r0 - g0) + ldh * 3LL * R * (k1 - k), ldh);
      }
    };
    for (int k1 = k; k1 < kend; k1 += NB) {
      const int jb = std::min(NB, n - k1);
      const int m = n - k1;
      T* P = A + k1 + lda * (long long)k1;  // panel top-left
      for (int jj = 0; jj < jb; jj += IB) {
        const int w = std::min(IB, jb - jj);
        const int mm = m - jj;
        T* Bk = P + jj + lda * (long long)jj;
        int* piv = f.ipiv + k1 + jj;
        const int grid = (mm + ROWS - 1) / ROWS;
        if (c.panel_mode == 1 && grid <= c.panel_grid[sizeof(typename Tr<T>::R) == 8 ? 1 : 0]) {
          long long ld_ = lda;
          int mm_ = mm, w_ = w, g0_ = k1 + jj;
          int* info_ = c.ps.info;
          PanelSync pan = c.pan;
          // epochs start at 1 (tags are zero-initialised); 2^32 launches before a repeat
          unsigned ep = ++c.pan_epoch;
          if (ep == 0) ep = ++c.pan_epoch;
          hipLaunchKernelGGL(k_panel<T>, dim3(grid), dim3(ROWS), 0, s, Bk, ld_, mm_, w_, piv, pan,
                             info_, g0_, ep);
        } else {
          for (int j = -1; j < w; j++)
            hipLaunchKernelGGL(k_col_step<T>, dim3(grid), dim3(ROWS), 0, s, Bk, lda, mm, w, j, piv,
                               c.ps, k1 + jj);
        }
        // interchanges of this narrow block on the other columns of the panel, and the solve
        // U12 := L11^{-1} U12 on the columns to its right
        const int nside = jj + (jb - jj - w);
        if (nside > 0)
          hipLaunchKernelGGL(k_swap_trsm<T>, dim3((nside + 7) / 8), dim3(256), 0, s, A, lda,
                             f.ipiv, k1 + jj, w, (const T*)Bk, lda, k1, k1 + jj, k1 + jj + w,
                             k1 + jb);
        if (jj + w < jb) {
          const int nr = jb - jj - w;
          T* U12 = Bk + lda * (long long)w;  // w x nr
          if (mm - w > 0)
            MFB_BLAS(gemm(c.blas, mm - w, nr, w, -1.0, Bk + w, lda, U12, lda, 1.0, U12 + w, lda));
        }
      }
      const int blk = k1 / NB;
      T* Li = f.Linv + (size_t)NB * NB * blk;
      T* Ui = f.Uinv + (size_t)NB * NB * blk;
      // the diagonal block's inverses read only this panel (final now): before waiting for s2
      rc = tri_inverse(c, P, lda, jb, Li, false, f.work, s);
      if (rc) return rc;
      if (jb < NB) {   // a short last block; the full ones are inverted together below
        rc = tri_inverse(c, P, lda, jb, Ui, true, f.work, s);
        if (rc) return rc;
      }
      // interchanges of this panel on the super-panel's other columns; the super-panel's columns
      // right of the panel: U rows, then rows k1 + jb .. n updated (all on s: the next panel of
      // the super-panel waits for them)
      apply_swaps(c, A, lda, f.ipiv, k1, jb, k, k1, k1 + jb, kend, s);
      if (k1 + jb < kend) {
        const int cw = kend - k1 - jb, nr = n - k1 - jb;
        T* A12 = P + lda * (long long)jb;
        MFB_BLAS(gemm(c.blas, jb, cw, jb, 1.0, Li, NB, A12, lda, 0.0, f.work, NB));
        bool done = false;
        if constexpr (kSplitType<T>) {
          if (Hl) {
            // the U rows go from the work panel to the factor with their FP16 operand
            constexpr int R = RealW<T>::R;
            __half* Hr2 = nullptr;
            rc = ws_get(c, WS_HR2, sizeof(__half) * (size_t)3 * R * jb * cw, &Hr2);
            if (rc) return rc;
            split_left(k1, jb, k1 + jb);
            hipLaunchKernelGGL(k_split_right_store, dim3((jb + 255) / 256, cw), dim3(256), 0, s,
                               (const T*)f.work, (long long)NB, jb, jb, cw, Hr2, 3LL * R * jb,
                               A12, lda);
            MFB_BLAS(gemm_f16(c.blas, R * nr, cw, 3 * R * jb, -1.0f,
                              Hl + (long long)R * (k1 + jb - g0) + ldh * 3LL * R * (k1 - k), ldh,
                              Hr2, 3LL * R * jb, 1.0f, (float*)(A12 + jb), R * lda));
            done = true;
          }
        }
        if (!done) {
          MFB_HIP(hipMemcpy2DAsync(A12, lda * sizeof(T), f.work, NB * sizeof(T), jb * sizeof(T),
                                   cw, hipMemcpyDeviceToDevice, s));
          MFB_BLAS(gemm(c.blas, nr, cw, jb, -1.0, P + jb, lda, A12, lda, 1.0, A12 + jb, lda));
        }
      }
    }
    // split left operands of the super-panel's panels from rows kend .. n (a panel's rows are
    // final once the later panels' interchanges have reached its columns): Hl is not read by s2
    for (int k1 = k; k1 < kend; k1 += NB) split_left(k1, std::min(NB, n - k1), kend);
    if (pending) {   // s2's update of the other columns (and its reads of the previous L panels)
      MFB_HIP(hipStreamWaitEvent(s, c.ev_join, 0));
      pending = false;
    }
    // interchanges of the super-panel's panels, in order, on the columns outside it
    for (int k1 = k; k1 < kend; k1 += NB)
      apply_swaps(c, A, lda, f.ipiv, k1, std::min(NB, n - k1), 0, k, kend, n, s);
    if (kend < n) {
      const int nr = n - kend, kb = kend - k;
      __half* Hr = nullptr;
      if constexpr (kSplitType<T>) {
        if (Hl) {
          rc = ws_get(c, WS_HR, sizeof(__half) * (size_t)3 * RealW<T>::R * kb * nr, &Hr);
          if (rc) return rc;
        }
      }
      // U rows k .. kend of the columns [col0, col0 + cw) (the super-panel's panels in turn: a
      // panel's rows first lose the earlier panels' contributions, then L11^{-1}), through `wk`
      // With the split operands the U rows stay in wk (ld 2 NB) and one kernel stores them into
      // the factor while it builds the FP16 right operand of their columns (Hr); the chunked
      // updates then only run the GEMM.
      auto urows = [&](rocblas_handle h, hipStream_t st, T* wk, int col0, int cw) -> int {
        const bool keep = Hr != nullptr;
        const long long ldw = keep ? 2LL * NB : (long long)NB;
        for (int k1 = k; k1 < kend; k1 += NB) {
          const int jb = std::min(NB, n - k1);
          T* Ak = A + k1 + lda * (long long)col0;
          T* wk1 = keep ? wk + (k1 - k) : wk;
          if (k1 > k) {
            if (keep)
              MFB_BLAS(gemm(h, jb, cw, k1 - k, -1.0, A + k1 + lda * (long long)k, lda, wk, ldw,
                            1.0, Ak, lda));
            else
              MFB_BLAS(gemm(h, jb, cw, k1 - k, -1.0, A + k1 + lda * (long long)k, lda,
                            A + k + lda * (long long)col0, lda, 1.0, Ak, lda));
          }
          MFB_BLAS(gemm(h, jb, cw, jb, 1.0, f.Linv + (size_t)NB * NB * (k1 / NB), NB, Ak, lda,
                        0.0, wk1, ldw));
          if (!keep)
            MFB_HIP(hipMemcpy2DAsync(Ak, lda * sizeof(T), wk, NB * sizeof(T), jb * sizeof(T), cw,
                                     hipMemcpyDeviceToDevice, st));
        }
        if constexpr (kSplitType<T>) {
          if (keep) {
            constexpr int R = RealW<T>::R;
            hipLaunchKernelGGL(k_split_right_store, dim3((kb + 255) / 256, cw), dim3(256), 0, st,
                               (const T*)wk, ldw, kb, NB, cw, Hr + 3LL * R * kb * (col0 - kend),
                               3LL * R * kb, A + k + lda * (long long)col0, lda);
          }
        }
        return MFB_OK;
      };
      // A22 (rows kend .. n, cw columns from kend + c0) -= A21 A12
      auto update = [&](rocblas_handle h, hipStream_t st, int c0, int cw) -> int {
        T* B = A + k + lda * (long long)(kend + c0);
        if constexpr (kSplitType<T>) {
          if (Hl) {   // the right operand was built with the U rows (urows)
            constexpr int R = RealW<T>::R;
            MFB_BLAS(gemm_f16(h, R * nr, cw, 3 * R * kb, -1.0f,
                              Hl + (long long)R * (kend - g0), ldh,
                              Hr + 3LL * R * kb * c0, 3LL * R * kb, 1.0f, (float*)(B + kb),
                              R * lda));
            return MFB_OK;
          }
        }
        MFB_BLAS(gemm(h, nr, cw, kb, -1.0, A + kend + lda * (long long)k, lda, B, lda, 1.0,
                      B + kb, lda));
        return MFB_OK;
      };
      // with look-ahead only the next super-panel's columns on s
      const int j2 = la ? std::min(std::max(SB, c.la_main), nr) : nr;
      if (j2 < nr) {   // the rest on s2, started first (it reads only this super-panel and Linv)
        const int nr2 = nr - j2;
        MFB_HIP(hipEventRecord(c.ev_fork, s));
        MFB_HIP(hipStreamWaitEvent(c.s2, c.ev_fork, 0));
        rc = urows(c.blas2, c.s2, work2, kend + j2, nr2);
        if (rc) return rc;
        // in column chunks: between two GEMM launches the dispatcher can start the next panel's
        // workgroups (one large GEMM keeps the CUs until its last workgroups are dispatched)
        const int chunk_req = schedule().la_chunk >= 0 ? schedule().la_chunk : c.la_chunk;
        const int chunk = chunk_req > 0 ? chunk_req : (SB > NB ? 768 : 2048);
        for (int c0 = 0; c0 < nr2; c0 += chunk) {
          rc = update(c.blas2, c.s2, j2 + c0, std::min(chunk, nr2 - c0));
          if (rc) return rc;
        }
        MFB_HIP(hipEventRecord(c.ev_join, c.s2));
        pending = true;
      }
      rc = urows(c.blas, s, f.work, kend, j2);
      if (rc) return rc;
      rc = update(c.blas, s, 0, j2);
      if (rc) return rc;
    }
  }
  if (pending) MFB_HIP(hipStreamWaitEvent(s, c.ev_join, 0));
  rc = tri_inverse_upper_blocks(c, (const T*)A, lda, n / NB, f.Uinv, f.work, s);
  if (rc) return rc;
  // products of the diagonal-block inverses with their nearest off-diagonal blocks, for the
  // one-right-hand-side triangular solves: Lm_i = Linv_i L_i,i-1, Um_i = Uinv_i U_i,i+1
  if (nblk > 1) {
    rc = ws_get(c, WS_LM, sizeof(T) * (size_t)NB * NB * nblk, &f.Lm);
    if (!rc) rc = ws_get(c, WS_UM, sizeof(T) * (size_t)NB * NB * nblk, &f.Um);
    if (rc) return rc;
    const int jl = n - (nblk - 1) * NB;   // rows of the last block
    const long long sd = (long long)NB * (1 + lda), si = (long long)NB * NB;
    if (nblk > 2) {
      MFB_BLAS(gemm_sb(c.blas, NB, NB, NB, 1.0, f.Linv + si, NB, si, A + NB, lda, sd, 0.0,
                       f.Lm + si, NB, si, nblk - 2));
      MFB_BLAS(gemm_sb(c.blas, NB, NB, NB, 1.0, f.Uinv, NB, si, A + lda * (long long)NB, lda, sd,
                       0.0, f.Um, NB, si, nblk - 2));
    }
    const long long last = (long long)(nblk - 1) * NB;
    MFB_BLAS(gemm(c.blas, jl, NB, jl, 1.0, f.Linv + si * (nblk - 1), NB,
                  A + last + lda * (last - NB), lda, 0.0, f.Lm + si * (nblk - 1), NB));
    MFB_BLAS(gemm(c.blas, NB, jl, NB, 1.0, f.Uinv + si * (nblk - 2), NB,
                  A + (last - NB) + lda * last, lda, 0.0, f.Um + si * (nblk - 2), NB));
  } else {
    f.Lm = f.Um = f.Linv;   // one block: no neighbour, never read
  }
  MFB_HIP(hipGetLastError());
  // pivots are relative to each narrow block's first row: make them global 0-based, then build
  // the row permutation
  std::vector<int> hp(n);
  MFB_HIP(hipMemcpyAsync(hp.data(), f.ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  int info = 0;
  unsigned perr = 0;
  MFB_HIP(hipMemcpyAsync(&info, c.ps.info, sizeof(int), hipMemcpyDeviceToHost, s));
  MFB_HIP(hipMemcpyAsync(&perr, c.pan.err, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  if (perr) {
    const unsigned zero = 0;
    (void)hipMemcpy(c.pan.err, &zero, sizeof(unsigned), hipMemcpyHostToDevice);
    set_error("getrf: panel grid barrier timed out (workgroups not co-resident)");
    c.panel_timeout = true;
    return MFB_ERR_HIP;
  }
  for (int k = 0; k < n; k += IB)
    for (int i = k; i < std::min(n, k + IB); i++) hp[i] += k;
  f.perm.resize(n);
  for (int i = 0; i < n; i++) f.perm[i] = i;
  for (int i = 0; i < n; i++) std::swap(f.perm[i], f.perm[hp[i]]);
  MFB_HIP(hipMemcpyAsync(f.ipiv, hp.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  MFB_HIP(hipMemcpyAsync(f.dperm, f.perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  MFB_HIP(hipStreamSynchronize(s));
  f.info = info;
  return MFB_OK;
}

// X (n x nrhs, ldx) := A^{-1} X with the factors of f; tmp: n x nrhs scratch (ld n)
template <typename T>
static int getrs(const LU<T>& f, T* X, long long ldx, int nrhs, T* tmp, hipStream_t s) {
  Ctx& c = ctx();
  const int n = f.n;
  const long long tot = (long long)n * nrhs;
  hipLaunchKernelGGL(k_permute<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, X, ldx,
                     tmp, n, f.dperm, n, nrhs);
  const int nblk_all = (n + NB - 1) / NB;
  if (nrhs == 1 && c.panel_mode == 1 && nblk_all <= c.trsv_grid && f.Lm) {
    // one persistent launch per triangle: L y = P b (tmp -> X), then U x = y (X -> tmp); row
    // blocks split over 4 workgroups when that grid fits co-resident, else over 2 when that one
    // does (MFB_TRSV_SPLIT=0: never, =2: at most 2); FP32 types post stamped words
    constexpr bool stamped = std::is_same<typename Tr<T>::R, float>::value;
    const int* sg = stamped ? c.trsv_stamp_grid : c.trsv_split_grid;
    const char* sp_env = getenv("MFB_TRSV_SPLIT");
    const int maxp = schedule().trsv_parts > 0 ? schedule().trsv_parts
                     : !sp_env                 ? 4
                     : strcmp(sp_env, "0") == 0 ? 1
                     : strcmp(sp_env, "2") == 0 ? 2 : 4;
    const int parts = maxp >= 4 && 4 * nblk_all <= sg[0]   ? 4
                      : maxp >= 2 && 2 * nblk_all <= sg[1] ? 2 : 1;
    const bool split = parts > 1;
    T* rx = nullptr;
    unsigned long long *ys = nullptr, *rss = nullptr;
    if (split && !stamped) {
      const int rc = ws_get(c, WS_RX, sizeof(T) * (size_t)NB * nblk_all, &rx);
      if (rc) return rc;
    }
    if (split && stamped) {
      // posted words start at zero (epoch 0 is never used), so no stale word can carry an epoch
      const size_t words = (size_t)NB * nblk_all * (sizeof(T) / sizeof(float));
      for (int slot : {WS_YS, WS_RSS}) {
        const bool fresh = c.ws_cap[slot] < sizeof(unsigned long long) * words;
        unsigned long long* p = nullptr;
        const int rc = ws_get(c, slot, sizeof(unsigned long long) * words, &p);
        if (rc) return rc;
        if (fresh) MFB_HIP(hipMemsetAsync(p, 0, sizeof(unsigned long long) * words, s));
        (slot == WS_YS ? ys : rss) = p;
      }
    }
    // P b kept for the launch-per-block solve, should a persistent launch's wait expire
    T* pb = nullptr;
    {
      const int rc = ws_get(c, WS_PB, sizeof(T) * (size_t)n, &pb);
      if (rc) return rc;
    }
    MFB_HIP(hipMemcpyAsync(pb, tmp, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    for (int up = 0; up < 2; up++) {
      const T* A_ = f.A;
      long long lda_ = f.lda;
      int n_ = n;
      const T* inv = up ? f.Uinv : f.Linv;
      const T* mm = up ? f.Um : f.Lm;
      const T* in = up ? X : tmp;
      T* out = up ? tmp : X;
      SolveSync ss = c.sol;
      unsigned ep = ++c.sol_epoch;
      if (ep == 0) ep = ++c.sol_epoch;
      if (split && stamped) {
        if constexpr (stamped) {
          const dim3 g(parts * nblk_all);
          if (parts == 4 && up)
            hipLaunchKernelGGL((k_trsv_stamped<T, true, 4, 512>), g, dim3(512), 0, s, A_, lda_,
                               n_, inv, mm, in, out, ys, rss, ss, ep);
          else if (parts == 4)
            hipLaunchKernelGGL((k_trsv_stamped<T, false, 4, 512>), g, dim3(512), 0, s, A_, lda_,
                               n_, inv, mm, in, out, ys, rss, ss, ep);
          else if (up)
            hipLaunchKernelGGL((k_trsv_stamped<T, true, 2, 1024>), g, dim3(1024), 0, s, A_, lda_,
                               n_, inv, mm, in, out, ys, rss, ss, ep);
          else
            hipLaunchKernelGGL((k_trsv_stamped<T, false, 2, 1024>), g, dim3(1024), 0, s, A_, lda_,
                               n_, inv, mm, in, out, ys, rss, ss, ep);
        }
      } else if (split) {
        const dim3 g(parts * nblk_all);
        if (parts == 4 && up)
          hipLaunchKernelGGL((k_trsv_split<T, true, 4>), g, dim3(TRSV_THREADS), 0, s, A_, lda_,
                             n_, inv, mm, in, out, rx, ss, ep);
        else if (parts == 4)
          hipLaunchKernelGGL((k_trsv_split<T, false, 4>), g, dim3(TRSV_THREADS), 0, s, A_, lda_,
                             n_, inv, mm, in, out, rx, ss, ep);
        else if (up)
          hipLaunchKernelGGL((k_trsv_split<T, true, 2>), g, dim3(TRSV_THREADS), 0, s, A_, lda_,
                             n_, inv, mm, in, out, rx, ss, ep);
        else
          hipLaunchKernelGGL((k_trsv_split<T, false, 2>), g, dim3(TRSV_THREADS), 0, s, A_, lda_,
                             n_, inv, mm, in, out, rx, ss, ep);
      } else if (up) {
        hipLaunchKernelGGL((k_trsv_blocked<T, true>), dim3(nblk_all), dim3(TRSV_THREADS), 0, s, A_,
                           lda_, n_, inv, mm, in, out, ss, ep);
      } else {
        hipLaunchKernelGGL((k_trsv_blocked<T, false>), dim3(nblk_all), dim3(TRSV_THREADS), 0, s,
                           A_, lda_, n_, inv, mm, in, out, ss, ep);
      }
    }
    MFB_HIP(hipMemcpyAsync(X, tmp, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    unsigned herr = 0;
    MFB_HIP(hipMemcpyAsync(&herr, c.sol.err, sizeof herr, hipMemcpyDeviceToHost, s));
    MFB_HIP(hipStreamSynchronize(s));
    if (!herr) return MFB_OK;
    // a wait expired (workgroups not co-resident): solve again with one launch per block step
    const unsigned zero = 0;
    MFB_HIP(hipMemcpy(c.sol.err, &zero, sizeof zero, hipMemcpyHostToDevice));
    MFB_HIP(hipMemcpyAsync(tmp, pb, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    c.recovered++;
  }
  if (nrhs <= SMALL_RHS) {
    const int nblk = (n + NB - 1) / NB;
    for (int k = 0; k < n; k += NB) {
      const int jb = std::min(NB, n - k);
      hipLaunchKernelGGL(k_blk_matvec<T>, dim3(1, nrhs), dim3(NB), 0, s,
                         f.Linv + (size_t)NB * NB * (k / NB), jb, tmp + k, (long long)n, X + k, ldx);
      const int m = n - k - jb;
      if (m > 0)
        hipLaunchKernelGGL(k_gemv_sub<T>, dim3((m + 255) / 256, nrhs), dim3(256), 0, s,
                           f.A + (k + jb) + f.lda * (long long)k, f.lda, m, jb, X + k, ldx,
                           tmp + k + jb, (long long)n);
    }
    MFB_HIP(hipMemcpy2DAsync(tmp, n * sizeof(T), X, ldx * sizeof(T), n * sizeof(T), nrhs,
                             hipMemcpyDeviceToDevice, s));
    for (int b = nblk - 1; b >= 0; b--) {
      const int k = b * NB, jb = std::min(NB, n - k);
      hipLaunchKernelGGL(k_blk_matvec<T>, dim3(1, nrhs), dim3(NB), 0, s,
                         f.Uinv + (size_t)NB * NB * b, jb, tmp + k, (long long)n, X + k, ldx);
      if (k > 0)
        hipLaunchKernelGGL(k_gemv_sub<T>, dim3((k + 255) / 256, nrhs), dim3(256), 0, s,
                           f.A + f.lda * (long long)k, f.lda, k, jb, X + k, ldx, tmp, (long long)n);
    }
    return MFB_OK;
  }
  // forward: y_k = Linv_kk b_k ; b_{k+1:} -= L_{k+1:,k} y_k   (y written back into X)
  for (int k = 0; k < n; k += NB) {
    const int jb = std::min(NB, n - k);
    const T* Li = f.Linv + (size_t)NB * NB * (k / NB);
    MFB_BLAS(gemm(c.blas, jb, nrhs, jb, 1.0, Li, NB, tmp + k, n, 0.0, X + k, ldx));
    if (k + jb < n)
      MFB_BLAS(gemm(c.blas, n - k - jb, nrhs, jb, -1.0, f.A + (k + jb) + f.lda * (long long)k,
                    f.lda, X + k, ldx, 1.0, tmp + k + jb, n));
  }
  // backward: x_k = Uinv_kk y_k ; y_{:k} -= U_{:k,k} x_k
  MFB_HIP(hipMemcpy2DAsync(tmp, n * sizeof(T), X, ldx * sizeof(T), n * sizeof(T), nrhs,
                           hipMemcpyDeviceToDevice, s));
  const int nblk = (n + NB - 1) / NB;
  for (int b = nblk - 1; b >= 0; b--) {
    const int k = b * NB, jb = std::min(NB, n - k);
    const T* Ui = f.Uinv + (size_t)NB * NB * b;
    MFB_BLAS(gemm(c.blas, jb, nrhs, jb, 1.0, Ui, NB, tmp + k, n, 0.0, X + k, ldx));
    if (k > 0)
      MFB_BLAS(gemm(c.blas, k, nrhs, jb, -1.0, f.A + f.lda * (long long)k, f.lda, X + k, ldx, 1.0,
                    tmp, n));
  }
  return MFB_OK;
}

template <typename T>
static double maxabs(const T* v, long long n, unsigned long long* dscr, hipStream_t s) {
  (void)hipMemsetAsync(dscr, 0, sizeof(unsigned long long), s);
  int blocks = (int)std::min<long long>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_maxabs<T>, dim3(std::max(blocks, 1)), dim3(256), 0, s, v, n, dscr);
  unsigned long long h = 0;
  (void)hipMemcpyAsync(&h, dscr, sizeof h, hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  double d;
  std::memcpy(&d, &h, sizeof d);
  return d;
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// FP64 path: factor A in place and solve
template <typename T>
static int gesv64(int n, int nrhs, T* A, long long lda, int* ipiv_out, T* B, long long ldb,
                  mfb_solve_report* rep, hipStream_t s) {
  LU<T> f;
  const long long rec0 = ctx().recovered;
  auto t0 = std::chrono::steady_clock::now();
  int rc = getrf(f, A, lda, n, s);
  if (rc) { lu_free(f); return rc; }
  if (rep) rep->factor_ms = ms_since(t0);
  if (f.info) {
    if (rep) rep->info = f.info;
    set_error("U(%d,%d) is exactly zero: the matrix is singular", f.info, f.info);
    lu_free(f);
    return MFB_ERR_SINGULAR;
  }
  auto t1 = std::chrono::steady_clock::now();
  T* tmp = nullptr;
  rc = ws_get(ctx(), WS_TMP, sizeof(T) * (size_t)n * nrhs, &tmp);
  if (!rc) rc = getrs(f, B, ldb, nrhs, tmp, s);
  if (ipiv_out && !rc) {
    // LAPACK format: 1-based
    std::vector<int> hp(n);
    MFB_HIP(hipMemcpyAsync(hp.data(), f.ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    MFB_HIP(hipStreamSynchronize(s));
    for (auto& v : hp) v += 1;
    MFB_HIP(hipMemcpyAsync(ipiv_out, hp.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  }
  (void)hipStreamSynchronize(s);
  if (rep) {
    rep->solve_ms = ms_since(t1);
    rep->precision_mode = MFB_PREC_FP64;
    rep->info = 0;
    if (ctx().recovered > rec0) rep->fell_back |= 2;
  }
  lu_free(f);
  return rc;
}

// Mixed precision: Tl is the FP32 type of Th.
// BEM columns mix h (O(1)) and g (O(1/mu)) entries: the FP32 factorisation works on the
// column-equilibrated matrix A C (C = diag of powers of two, as LAPACK ?geequ would choose them up
// to rounding), and convergence is judged on the equilibrated unknowns y = C^{-1} x.
// the FP32-factor modes: plain FP32 trailing updates, or split FP16 ones
inline bool is_mixed(int mode) {
  return mode == MFB_PREC_FP32_REFINE || mode == MFB_PREC_SPLIT16_REFINE;
}

template <typename Th, typename Tl>
struct MixedLU {
  LU<Tl> f;               // FP32 factors of A C
  const Th* A = nullptr;  // the FP64 matrix (residuals b - A x)
  long long lda = 0;
  int n = 0;
  double* cs = nullptr;   // C
  double anorm = 0;       // ||A C||_inf
};

// FP32 factorisation of A C into the context's workspaces (WS_A32, WS_CS and the LU slots)
template <typename Th, typename Tl>
static int mixed_factor(MixedLU<Th, Tl>& mx, const Th* A, long long lda, int n, hipStream_t s,
                        bool split16 = false) {
  Ctx& c = ctx();
  int rc = ensure_ctx(s, n);
  if (rc) return rc;
  Tl* A32 = nullptr;
  double *cs = nullptr, *rs = nullptr;
  unsigned long long* scr = nullptr;
  rc = ws_get(c, WS_A32, sizeof(Tl) * (size_t)n * n, &A32);
  if (!rc) rc = ws_get(c, WS_CS, sizeof(double) * n, &cs);
  if (!rc) rc = ws_get(c, WS_RS, sizeof(double) * n, &rs);
  if (!rc) rc = ws_get(c, WS_SCR, sizeof(unsigned long long), &scr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_colscale<Th>, dim3(n), dim3(256), 0, s, A, lda, n, cs);
  // the FP32 copy of A C and the row sums of |A C| in one pass over A
  hipLaunchKernelGGL(k_zero_mat<double>, dim3((n + 255) / 256), dim3(256), 0, s, rs, (long long)n,
                     n, 1);
  hipLaunchKernelGGL((k_convert_scaled<Th, Tl>), dim3((n + 255) / 256, (n + CV_COLS - 1) / CV_COLS),
                     dim3(256), 0, s, A, lda, A32, n, n, (const double*)cs, rs);
  mx.A = A;
  mx.lda = lda;
  mx.n = n;
  mx.cs = cs;
  mx.f.split16 = split16;
  rc = getrf(mx.f, A32, n, n, s);
  if (rc == MFB_ERR_HIP && c.panel_timeout) {
    // the persistent panel's wait expired: the FP32 copy is rebuilt from A (intact) and factored
    // with one launch per column
    c.panel_timeout = false;
    hipLaunchKernelGGL(k_zero_mat<double>, dim3((n + 255) / 256), dim3(256), 0, s, rs,
                       (long long)n, n, 1);
    hipLaunchKernelGGL((k_convert_scaled<Th, Tl>),
                       dim3((n + 255) / 256, (n + CV_COLS - 1) / CV_COLS), dim3(256), 0, s, A, lda,
                       A32, n, n, (const double*)cs, rs);
    const int pm = c.panel_mode;
    c.panel_mode = 0;
    rc = getrf(mx.f, A32, n, n, s);
    c.panel_mode = pm;
    if (!rc) c.recovered++;
  }
  if (rc) return rc;
  // ||A C||_inf (the row sums came with the conversion)
  mx.anorm = maxabs(rs, n, scr, s);
  return MFB_OK;
}

// FP64 iterative refinement with the FP32 factors: B (n x nrhs, ldb) is overwritten with the
// solution only when the LAPACK zcgesv criterion is met (*ok); otherwise B is left unchanged.
template <typename Th, typename Tl>
static int mixed_refine(const MixedLU<Th, Tl>& mx, Th* B, long long ldb, int nrhs, bool* ok,
                        int* iters, double* berr_out, hipStream_t s) {
  Ctx& c = ctx();
  const int n = mx.n;
  const long long tot = (long long)n * nrhs;
  Th *x = nullptr, *r = nullptr, *y = nullptr;
  Tl *r32 = nullptr, *t32 = nullptr;
  unsigned long long* scr = nullptr;
  int rc = ws_get(c, WS_X, sizeof(Th) * tot, &x);
  if (!rc) rc = ws_get(c, WS_Y, sizeof(Th) * tot, &y);
  if (!rc) rc = ws_get(c, WS_R, sizeof(Th) * tot, &r);
  if (!rc) rc = ws_get(c, WS_R32, sizeof(Tl) * tot, &r32);
  if (!rc) rc = ws_get(c, WS_T32, sizeof(Tl) * tot, &t32);
  if (!rc) rc = ws_get(c, WS_SCR, sizeof(unsigned long long), &scr);
  Th* part = nullptr;  // partial sums of the residual GEMV (few right-hand sides)
  if (!rc && nrhs <= SMALL_RHS)
    rc = ws_get(c, WS_PART, sizeof(Th) * (size_t)((n + GV_COLS - 1) / GV_COLS) * n * nrhs, &part);
  if (rc) return rc;
  const Th* A = mx.A;
  const long long lda = mx.lda;
  const double eps = 1.1102230246251565e-16;  // dlamch('Epsilon')
  const double cte = mx.anorm * eps * std::sqrt((double)n);
  const unsigned eg = (unsigned)std::min<long long>((tot + 255) / 256, 65536);
  *ok = false;
  int it = 0;
  double berr = 0, prev_rn = INFINITY;
  int stalls = 0;
  MFB_HIP(hipMemsetAsync(x, 0, sizeof(Th) * tot, s));
  MFB_HIP(hipMemcpy2DAsync(r, n * sizeof(Th), B, ldb * sizeof(Th), n * sizeof(Th), nrhs,
                           hipMemcpyDeviceToDevice, s));
  const double bnorm = maxabs(r, tot, scr, s);
  for (it = 0; it < 30; it++) {
    hipLaunchKernelGGL((k_convert<Th, Tl>), dim3(eg), dim3(256), 0, s, r, n, r32, n, n, nrhs);
    rc = getrs(mx.f, r32, n, nrhs, t32, s);
    if (rc) return rc;
    hipLaunchKernelGGL((k_axpy_scaled<Tl, Th>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       s, r32, x, n, tot, (const double*)mx.cs);
    // r = b - A x in FP64
    if (part) {
      const int nch = (n + GV_COLS - 1) / GV_COLS;
      hipLaunchKernelGGL(k_gemv_partial<Th>, dim3((n + 255) / 256, nch, nrhs), dim3(256), 0, s, A,
                         lda, n, x, (long long)n, part);
      hipLaunchKernelGGL(k_gemv_reduce<Th>, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, part, n,
                         nch, B, ldb, r, (long long)n);
    } else {
      MFB_HIP(hipMemcpy2DAsync(r, n * sizeof(Th), B, ldb * sizeof(Th), n * sizeof(Th), nrhs,
                               hipMemcpyDeviceToDevice, s));
      MFB_BLAS(gemm(c.blas, n, nrhs, n, -1.0, A, lda, x, n, 1.0, r, n));
    }
    hipLaunchKernelGGL(k_unscale<Th>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, x, y,
                       n, tot, (const double*)mx.cs);
    const double rn = maxabs(r, tot, scr, s), yn = maxabs(y, tot, scr, s);
    berr = rn / (mx.anorm * yn + bnorm);
    if (rn <= yn * cte) { *ok = true; it++; break; }
    if (!(rn < 0.5 * prev_rn)) stalls++;
    else stalls = 0;
    if (stalls >= 3 || !(rn == rn)) break;
    prev_rn = rn;
  }
  if (*ok) {
    MFB_HIP(hipMemcpy2DAsync(B, ldb * sizeof(Th), x, n * sizeof(Th), n * sizeof(Th), nrhs,
                             hipMemcpyDeviceToDevice, s));
    MFB_HIP(hipStreamSynchronize(s));
  }
  *iters = it;
  *berr_out = berr;
  return MFB_OK;
}

template <typename Th, typename Tl>
static int gesv_mixed(int n, int nrhs, Th* A, long long lda, int* ipiv_out, Th* B, long long ldb,
                      mfb_solve_report* rep, hipStream_t s, int mode) {
  auto t0 = std::chrono::steady_clock::now();
  const long long rec0 = ctx().recovered;
  MixedLU<Th, Tl> mx;
  int rc = mixed_factor(mx, A, lda, n, s, mode == MFB_PREC_SPLIT16_REFINE);
  if (rc) { lu_free(mx.f); return rc; }
  if (rep) rep->factor_ms = ms_since(t0);
  auto t1 = std::chrono::steady_clock::now();
  bool ok = false;
  int it = 0;
  double berr = 0;
  if (mx.f.info == 0) {
    rc = mixed_refine(mx, B, ldb, nrhs, &ok, &it, &berr, s);
    if (rc) { lu_free(mx.f); return rc; }
  }
  lu_free(mx.f);
  if (ok) {
    if (rep) {
      rep->precision_mode = mode;
      rep->refinement_steps = it;
      rep->fell_back = ctx().recovered > rec0 ? 2 : 0;
      rep->backward_error = berr;
      rep->solve_ms = ms_since(t1);
      rep->info = 0;
    }
    return MFB_OK;
  }
  // fall back to the FP64 LU (on a copy: A stays intact for the caller in this mode)
  Th* A64 = nullptr;
  MFB_HIP(hipMalloc(&A64, sizeof(Th) * (size_t)n * n));
  MFB_HIP(hipMemcpy2DAsync(A64, n * sizeof(Th), A, lda * sizeof(Th), n * sizeof(Th), n,
                           hipMemcpyDeviceToDevice, s));
  rc = gesv64<Th>(n, nrhs, A64, n, ipiv_out, B, ldb, rep, s);
  (void)hipFree(A64);
  if (rep) {
    rep->fell_back = 1 | (ctx().recovered > rec0 ? 2 : 0);
    rep->refinement_steps = it;
    rep->precision_mode = MFB_PREC_FP64;
  }
  return rc;
}

}  // namespace mfb

using namespace mfb;

extern "C" int mfb_dgesv_device(int n, int nrhs, double* A, long long lda, int* ipiv, double* B,
                                long long ldb, int precision_mode, mfb_solve_report* report,
                                void* stream) {
  MFB_ARG(n > 0 && nrhs > 0 && A && B && lda >= n && ldb >= n, "invalid dgesv arguments");
  if (report) std::memset(report, 0, sizeof *report);
  hipStream_t s = (hipStream_t)stream;
  if (precision_mode == MFB_PREC_FP64) return gesv64<double>(n, nrhs, A, lda, ipiv, B, ldb, report, s);
  if (is_mixed(precision_mode))
    return gesv_mixed<double, float>(n, nrhs, A, lda, ipiv, B, ldb, report, s, precision_mode);
  set_error("unknown precision_mode %d", precision_mode);
  return MFB_ERR_ARG;
}

extern "C" int mfb_zgesv_device(int n, int nrhs, double* A, long long lda, int* ipiv, double* B,
                                long long ldb, int precision_mode, mfb_solve_report* report,
                                void* stream) {
  MFB_ARG(n > 0 && nrhs > 0 && A && B && lda >= n && ldb >= n, "invalid zgesv arguments");
  if (report) std::memset(report, 0, sizeof *report);
  hipStream_t s = (hipStream_t)stream;
  if (precision_mode == MFB_PREC_FP64)
    return gesv64<cz>(n, nrhs, (cz*)A, lda, ipiv, (cz*)B, ldb, report, s);
  if (is_mixed(precision_mode))
    return gesv_mixed<cz, cf>(n, nrhs, (cz*)A, lda, ipiv, (cz*)B, ldb, report, s, precision_mode);
  set_error("unknown precision_mode %d", precision_mode);
  return MFB_ERR_ARG;
}

// ------------------------------------------------------------------------------------------------
// host-array drop-ins for solve_lse_r / solve_lse_c
// ------------------------------------------------------------------------------------------------
namespace mfb {

// ?geequ / ?laqge (LAPACK): row and column scale factors of A and their application
template <typename T>
__global__ void __launch_bounds__(256) k_row_absmax(const T* A, long long lda, int n, double* r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int j = 0; j < n; j++) v = fmax(v, (double)tabs1(A[i + lda * (long long)j]));
  r[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(256) k_col_absmax(const T* A, long long lda, int n,
                                                    const double* r, double* c) {
  const int j = blockIdx.x;
  __shared__ double red[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v = fmax(v, (double)tabs1(A[i + lda * (long long)j]) * r[i]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) c[j] = red[0];
}

// A(i, j) := (c_j r_i) A(i, j); r or c null: that factor is 1
template <typename T>
__global__ void k_scale_rc(T* A, long long lda, int m, int ncol, const double* r, const double* c) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)m * ncol) return;
  const int j = (int)(t / m), i = (int)(t % m);
  double f;
  if (r && c) f = c[j] * r[i];
  else if (r) f = r[i];
  else f = c[j];
  T& a = A[i + lda * (long long)j];
  a = tscale(a, f);
}

// FP32 factors of A C written out as FP64 LAPACK-format factors of A: L as is, U C^{-1}
template <typename Tl, typename Th>
__global__ void k_factor_out(const Tl* F, int n, const double* cs, Th* out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * n) return;
  const int j = (int)(t / n), i = (int)(t % n);
  const Th v = conv<Tl, Th>(F[t]);
  out[t] = i <= j ? tscale(v, 1.0 / cs[j]) : v;
}

template <typename T>
struct LowOf;
template <>
struct LowOf<double> {
  typedef float T;
};
template <>
struct LowOf<cz> {
  typedef cf T;
};

// A factor kept on the device between solve_lse calls (the reference re-solves with
// factorize = .false., src/multifebe.f90:119): it owns its arrays, taken over from the context's
// workspaces after the factorisation, and is keyed by the caller's host matrix, n and mode.
template <typename T>
struct Kept {
  typedef typename LowOf<T>::T Tl;
  const void* key = nullptr;
  int n = 0;
  int mode = -1;        // precision mode the caller asked for
  int eff = -1;         // mode of the kept factor (FP64 after a mixed-mode fallback)
  T* dA = nullptr;      // FP64 LU factors (eff FP64) or the FP64 matrix (eff mixed)
  T* dAo = nullptr;     // the unfactored (scaled) matrix, kept for the condition / refine stages
  LU<T> f;
  MixedLU<T, Tl> mx;
  std::vector<void*> owned;
};
static Kept<double> kept_r;
static Kept<cz> kept_c;

static void* ws_take(Ctx& c, int slot) {
  void* p = c.ws[slot];
  c.ws[slot] = nullptr;
  c.ws_cap[slot] = 0;
  return p;
}

// the arrays of a factor just computed in the context's workspaces move to `owned`
template <typename T>
static void take_lu(Ctx& c, const LU<T>& f, std::vector<void*>& owned) {
  for (int slot : {WS_IPIV, WS_LINV, WS_UINV, WS_DPERM}) owned.push_back(ws_take(c, slot));
  if (f.Lm != f.Linv) {
    owned.push_back(ws_take(c, WS_LM));
    owned.push_back(ws_take(c, WS_UM));
  }
}

template <typename T>
static void kept_release(Kept<T>& k) {
  for (void* p : k.owned)
    if (p) (void)hipFree(p);
  if (k.dA) (void)hipFree(k.dA);
  if (k.dAo) (void)hipFree(k.dAo);
  k = Kept<T>();
}

// ?geequ + ?laqge on the device matrix dA: r, c (host, n) and equed as LAPACK returns them
template <typename T>
static int equilibrate(T* dA, int n, double* r, double* c, char* equed, hipStream_t s) {
  const double smlnum = 2.2250738585072014e-308;  // dlamch('S')
  const double bignum = 1.0 / smlnum;
  const double prec = 2.220446049250313e-16;      // dlamch('P')
  double *dr = nullptr, *dc = nullptr;
  MFB_HIP(hipMalloc(&dr, sizeof(double) * n));
  MFB_HIP(hipMalloc(&dc, sizeof(double) * n));
  struct Free {
    double *a, *b;
    ~Free() { (void)hipFree(a); (void)hipFree(b); }
  } fr{dr, dc};
  hipLaunchKernelGGL(k_row_absmax<T>, dim3((n + 255) / 256), dim3(256), 0, s, dA, (long long)n, n, dr);
  MFB_HIP(hipMemcpyAsync(r, dr, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  double rcmin = bignum, rcmax = 0.0;
  for (int i = 0; i < n; i++) { rcmax = std::max(rcmax, r[i]); rcmin = std::min(rcmin, r[i]); }
  const double amax = rcmax;
  if (rcmin == 0.0) {
    for (int i = 0; i < n; i++)
      if (r[i] == 0.0) {
        set_error("scaling has failed: row %d of A is zero (?geequ info %d)", i + 1, i + 1);
        return MFB_ERR_SINGULAR;
      }
  }
  for (int i = 0; i < n; i++) r[i] = 1.0 / std::min(std::max(r[i], smlnum), bignum);
  const double rowcnd = std::max(rcmin, smlnum) / std::min(rcmax, bignum);
  MFB_HIP(hipMemcpyAsync(dr, r, sizeof(double) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_col_absmax<T>, dim3(n), dim3(256), 0, s, (const T*)dA, (long long)n, n,
                     (const double*)dr, dc);
  MFB_HIP(hipMemcpyAsync(c, dc, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  rcmin = bignum;
  rcmax = 0.0;
  for (int j = 0; j < n; j++) { rcmin = std::min(rcmin, c[j]); rcmax = std::max(rcmax, c[j]); }
  if (rcmin == 0.0) {
    for (int j = 0; j < n; j++)
      if (c[j] == 0.0) {
        set_error("scaling has failed: column %d of A is zero (?geequ info %d)", j + 1, n + j + 1);
        return MFB_ERR_SINGULAR;
      }
  }
  for (int j = 0; j < n; j++) c[j] = 1.0 / std::min(std::max(c[j], smlnum), bignum);
  const double colcnd = std::max(rcmin, smlnum) / std::min(rcmax, bignum);
  // ?laqge
  const double thresh = 0.1, small = smlnum / prec, large = 1.0 / small;
  const bool rows = !(rowcnd >= thresh && amax >= small && amax <= large);
  const bool cols = !(colcnd >= thresh);
  *equed = rows ? (cols ? 'B' : 'R') : (cols ? 'C' : 'N');
  if (rows || cols) {
    MFB_HIP(hipMemcpyAsync(dc, c, sizeof(double) * n, hipMemcpyHostToDevice, s));
    const long long tot = (long long)n * n;
    hipLaunchKernelGGL(k_scale_rc<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dA,
                       (long long)n, n, n, rows ? (const double*)dr : nullptr,
                       cols ? (const double*)dc : nullptr);
    MFB_HIP(hipStreamSynchronize(s));
  }
  return MFB_OK;
}

// rows of the device right-hand sides times the host vector v
template <typename T>
static int scale_rows(T* dB, int n, int nrhs, const double* v, hipStream_t s) {
  double* dv = nullptr;
  MFB_HIP(hipMalloc(&dv, sizeof(double) * n));
  hipError_t he = hipMemcpyAsync(dv, v, sizeof(double) * n, hipMemcpyHostToDevice, s);
  if (he == hipSuccess) {
    const long long tot = (long long)n * nrhs;
    hipLaunchKernelGGL(k_scale_rc<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dB,
                       (long long)n, n, nrhs, (const double*)dv, (const double*)nullptr);
    he = hipStreamSynchronize(s);
  }
  (void)hipFree(dv);
  if (he != hipSuccess) {
    set_error("scale_rows: %s", hipGetErrorString(he));
    return MFB_ERR_HIP;
  }
  return MFB_OK;
}

// Factorise the kept matrix k.dA in the requested mode; the factor's arrays become k's.
template <typename T>
static int kept_factor(Kept<T>& k, int mode, hipStream_t s) {
  Ctx& c = ctx();
  const int n = k.n;
  if (is_mixed(mode)) {
    int rc = mixed_factor(k.mx, (const T*)k.dA, n, n, s, mode == MFB_PREC_SPLIT16_REFINE);
    if (rc) return rc;
    take_lu(c, k.mx.f, k.owned);
    k.owned.push_back(ws_take(c, WS_A32));
    k.owned.push_back(ws_take(c, WS_CS));
    if (k.mx.f.info == 0) {
      k.eff = mode;
      return MFB_OK;
    }
    // a zero pivot in FP32: the FP64 factorisation decides
  }
  int rc = getrf(k.f, k.dA, n, n, s);
  if (rc) return rc;
  take_lu(c, k.f, k.owned);
  k.eff = MFB_PREC_FP64;
  if (k.f.info) {
    set_error("U(%d,%d) is exactly zero: the matrix is singular (?getrf info %d)", k.f.info,
              k.f.info, k.f.info);
    return MFB_ERR_SINGULAR;
  }
  return MFB_OK;
}

// ------------------------------------------------------------------------------------------------
// Condition estimate and refinement stages of solve_lse (src/solve_lse_c.f90:140-165, 191-206):
// LAPACK ?gecon and ?gerfs on the kept FP64 factors, their steps as LAPACK 3 takes them (the
// Hager-Higham 1-norm estimator ?lacn2 in reverse-communication order, ?gerfs' stopping rules and
// error bound); the triangular solves run on the GPU (rocBLAS trsv on the factors), the estimator's
// vector steps on the host.  ?latrs' overflow scaling is not reproduced (scale 1: the
// factors of a solvable system do not overflow a triangular solve).
// ------------------------------------------------------------------------------------------------
inline double cabs1(double a) { return std::fabs(a); }
inline double cabs1(cz a) { return std::fabs(a.re) + std::fabs(a.im); }
inline double hmod(double a) { return std::fabs(a); }
inline double hmod(cz a) { return std::hypot(a.re, a.im); }
inline double tscale_h(double a, double f) { return a * f; }
inline cz tscale_h(cz a, double f) { return cz{a.re * f, a.im * f}; }

// y[perm[i]] = x[i]
template <typename T>
__global__ void k_unpermute(const T* x, T* y, const int* perm, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[perm[i]] = x[i];
}
// out[j] = sum_i |A(i, j)| (modulus)
template <typename T>
__global__ void k_colabs(const T* A, long long lda, int n, double* out) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const T a = A[i + lda * j];
    if constexpr (std::is_same<T, cz>::value) acc += hypot(a.re, a.im);
    else acc += fabs(a);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}
// r = b - A x and w = cabs1(b) + sum_k cabs1(A(i, k)) cabs1(x(k)), one thread per row
__device__ inline double dcabs1(double a) { return fabs(a); }
__device__ inline double dcabs1(cz a) { return fabs(a.re) + fabs(a.im); }
template <typename T>
__global__ void k_resid_abs(const T* A, long long lda, int n, const T* x, const T* b, T* r,
                            double* w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T acc = tzero<T>();
  double wa = 0.0;
  for (int k = 0; k < n; k++) {
    const T a = A[i + lda * k];
    acc = tadd(acc, tmul(a, x[k]));
    wa += dcabs1(a) * dcabs1(x[k]);
  }
  r[i] = tsub(b[i], acc);
  w[i] = dcabs1(b[i]) + wa;
}

static rocblas_status trsv(rocblas_handle h, rocblas_fill uplo, rocblas_operation op,
                           rocblas_diagonal dg, int n, const double* A, long long lda, double* x) {
  return rocblas_dtrsv(h, uplo, op == rocblas_operation_conjugate_transpose
                                    ? rocblas_operation_transpose : op,
                       dg, n, A, (rocblas_int)lda, x, 1);
}
static rocblas_status trsv(rocblas_handle h, rocblas_fill uplo, rocblas_operation op,
                           rocblas_diagonal dg, int n, const cz* A, long long lda, cz* x) {
  return rocblas_ztrsv(h, uplo, op, dg, n, (const rocblas_double_complex*)A, (rocblas_int)lda,
                       (rocblas_double_complex*)x, 1);
}

// x := M x on the device with the FP64 factors P A = L U of f:
//   0: (L U)^{-1}   1: (L U)^{-H}   2: A^{-1}   3: A^{-H}
template <typename T>
static int lu_apply(const LU<T>& f, T* x, T* tmp, int op, hipStream_t s) {
  Ctx& c = ctx();
  const int n = f.n;
  const rocblas_operation N = rocblas_operation_none, H = rocblas_operation_conjugate_transpose;
  const rocblas_fill lo = rocblas_fill_lower, up = rocblas_fill_upper;
  const rocblas_diagonal unit = rocblas_diagonal_unit, nonunit = rocblas_diagonal_non_unit;
  MFB_BLAS(rocblas_set_stream(c.blas, s));
  if (op == 2) {
    hipLaunchKernelGGL(k_permute<T>, dim3((n + 255) / 256), dim3(256), 0, s, x, (long long)n, tmp,
                       (long long)n, f.dperm, n, 1);
    MFB_HIP(hipMemcpyAsync(x, tmp, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
  }
  if (op == 0 || op == 2) {
    MFB_BLAS(trsv(c.blas, lo, N, unit, n, f.A, f.lda, x));
    MFB_BLAS(trsv(c.blas, up, N, nonunit, n, f.A, f.lda, x));
  } else {
    MFB_BLAS(trsv(c.blas, up, H, nonunit, n, f.A, f.lda, x));
    MFB_BLAS(trsv(c.blas, lo, H, unit, n, f.A, f.lda, x));
  }
  if (op == 3) {
    hipLaunchKernelGGL(k_unpermute<T>, dim3((n + 255) / 256), dim3(256), 0, s, x, tmp, f.dperm, n);
    MFB_HIP(hipMemcpyAsync(x, tmp, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
  }
  MFB_HIP(hipGetLastError());
  return MFB_OK;
}

// The 1-norm of an implicit matrix M (LAPACK ?lacn2, Higham's modification of Hager's method):
// apply(1, x) computes x := M x, apply(2, x) x := M^H x (host vectors).  Returns the estimate.
static int lacn2(int n, const std::function<int(int, std::vector<double>&)>& apply, double* est) {
  std::vector<double> x(n, 1.0 / n), v;
  std::vector<int> isgn(n);
  int rc = apply(1, x);
  if (rc) return rc;
  if (n == 1) { *est = std::fabs(x[0]); return MFB_OK; }
  auto asum = [&](const std::vector<double>& y) { double t = 0; for (double a : y) t += std::fabs(a); return t; };
  auto imax = [&](const std::vector<double>& y) {
    int j = 0;
    for (int i = 1; i < n; i++) if (std::fabs(y[i]) > std::fabs(y[j])) j = i;
    return j;
  };
  double e = asum(x);
  for (int i = 0; i < n; i++) { x[i] = x[i] >= 0 ? 1.0 : -1.0; isgn[i] = (int)x[i]; }
  if ((rc = apply(2, x))) return rc;
  int j = imax(x), iter = 2;
  while (true) {
    std::fill(x.begin(), x.end(), 0.0);
    x[j] = 1.0;
    if ((rc = apply(1, x))) return rc;
    v = x;
    const double estold = e;
    e = asum(v);
    bool same = true;
    for (int i = 0; i < n && same; i++) same = (x[i] >= 0 ? 1 : -1) == isgn[i];
    if (same || e <= estold) break;   // repeated sign vector, or no increase
    for (int i = 0; i < n; i++) { x[i] = x[i] >= 0 ? 1.0 : -1.0; isgn[i] = (int)x[i]; }
    if ((rc = apply(2, x))) return rc;
    const int jlast = j;
    j = imax(x);
    if (!(x[jlast] != std::fabs(x[j]) && iter < 5)) break;
    iter++;
  }
  double alt = 1.0;
  for (int i = 0; i < n; i++) { x[i] = alt * (1.0 + (double)i / (double)(n - 1)); alt = -alt; }
  if ((rc = apply(1, x))) return rc;
  const double temp = 2.0 * (asum(x) / (3.0 * n));
  *est = temp > e ? temp : e;
  return MFB_OK;
}
static int lacn2(int n, const std::function<int(int, std::vector<cz>&)>& apply, double* est) {
  const double safmin = 2.2250738585072014e-308;
  std::vector<cz> x(n, cz{1.0 / n, 0.0});
  int rc = apply(1, x);
  if (rc) return rc;
  if (n == 1) { *est = hmod(x[0]); return MFB_OK; }
  auto asum = [&](const std::vector<cz>& y) { double t = 0; for (const cz& a : y) t += hmod(a); return t; };
  auto imax = [&](const std::vector<cz>& y) {
    int j = 0;
    for (int i = 1; i < n; i++) if (hmod(y[i]) > hmod(y[j])) j = i;
    return j;
  };
  auto unit_phase = [&](std::vector<cz>& y) {
    for (cz& a : y) {
      const double m = hmod(a);
      a = m > safmin ? cz{a.re / m, a.im / m} : cz{1.0, 0.0};
    }
  };
  double e = asum(x);
  unit_phase(x);
  if ((rc = apply(2, x))) return rc;
  int j = imax(x), iter = 2;
  while (true) {
    std::fill(x.begin(), x.end(), cz{0.0, 0.0});
    x[j] = cz{1.0, 0.0};
    if ((rc = apply(1, x))) return rc;
    const double estold = e;
    e = asum(x);
    if (e <= estold) break;
    unit_phase(x);
    if ((rc = apply(2, x))) return rc;
    const int jlast = j;
    j = imax(x);
    if (!(hmod(x[jlast]) != hmod(x[j]) && iter < 5)) break;
    iter++;
  }
  double alt = 1.0;
  for (int i = 0; i < n; i++) { x[i] = cz{alt * (1.0 + (double)i / (double)(n - 1)), 0.0}; alt = -alt; }
  if ((rc = apply(1, x))) return rc;
  const double temp = 2.0 * (asum(x) / (3.0 * n));
  *est = temp > e ? temp : e;
  return MFB_OK;
}

// host vector -> device, x := M x on the device, device -> host
template <typename T>
static int dev_apply(std::vector<T>& x, T* dx, T* tmp, const std::function<int(T*)>& op,
                     hipStream_t s) {
  const int n = (int)x.size();
  MFB_HIP(hipMemcpyAsync(dx, x.data(), sizeof(T) * n, hipMemcpyHostToDevice, s));
  int rc = op(dx);
  if (rc) return rc;
  MFB_HIP(hipMemcpyAsync(x.data(), dx, sizeof(T) * n, hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  return MFB_OK;
}

// ?gecon (norm '1'): rcond of the matrix whose FP64 factors are f, anorm its 1-norm
template <typename T>
static int gecon(const LU<T>& f, double anorm, double* rcond, hipStream_t s) {
  const int n = f.n;
  *rcond = 0.0;
  if (n == 0) { *rcond = 1.0; return MFB_OK; }
  if (anorm == 0.0) return MFB_OK;
  T *dx = nullptr, *tmp = nullptr;
  int rc = ws_get(ctx(), WS_X, sizeof(T) * n, &dx);
  if (!rc) rc = ws_get(ctx(), WS_Y, sizeof(T) * n, &tmp);
  if (rc) return rc;
  double ainvnm = 0.0;
  rc = lacn2(n, [&](int kase, std::vector<T>& x) {
    return dev_apply<T>(x, dx, tmp, [&](T* d) { return lu_apply(f, d, tmp, kase == 1 ? 0 : 1, s); }, s);
  }, &ainvnm);
  if (rc) return rc;
  if (ainvnm != 0.0) *rcond = (1.0 / ainvnm) / anorm;
  return MFB_OK;
}

// ?gerfs (trans 'N'): refine the solutions X (device, n x nrhs, ld n) of Ao X = B with the FP64
// factors f; berr, ferr (host, nrhs) as LAPACK defines them; *steps = largest number of
// corrections applied to one right-hand side
template <typename T>
static int gerfs(const T* Ao, const LU<T>& f, const T* B, T* X, int nrhs, double* berr,
                 double* ferr, int* steps, hipStream_t s) {
  const int n = f.n;
  const double eps = 1.1102230246251565e-16;          // dlamch('E')
  const double safmin = 2.2250738585072014e-308;      // dlamch('S')
  const double nz = n + 1.0, safe1 = nz * safmin, safe2 = safe1 / eps;
  T *dr = nullptr, *tmp = nullptr, *dv = nullptr;
  double* dw = nullptr;
  int rc = ws_get(ctx(), WS_R, sizeof(T) * n, &dr);
  if (!rc) rc = ws_get(ctx(), WS_Y, sizeof(T) * n, &tmp);
  if (!rc) rc = ws_get(ctx(), WS_RX, sizeof(T) * n, &dv);
  if (!rc) rc = ws_get(ctx(), WS_RS, sizeof(double) * n, &dw);
  if (rc) return rc;
  std::vector<T> hr(n), hx(n);
  std::vector<double> w(n);
  *steps = 0;
  for (int j = 0; j < nrhs; j++) {
    T* x = X + (long long)n * j;
    const T* b = B + (long long)n * j;
    int count = 1;
    double lstres = 3.0, be = 0.0;
    while (true) {
      hipLaunchKernelGGL(k_resid_abs<T>, dim3((n + 255) / 256), dim3(256), 0, s, Ao, (long long)n,
                         n, (const T*)x, b, dr, dw);
      MFB_HIP(hipGetLastError());
      MFB_HIP(hipMemcpyAsync(hr.data(), dr, sizeof(T) * n, hipMemcpyDeviceToHost, s));
      MFB_HIP(hipMemcpyAsync(w.data(), dw, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      MFB_HIP(hipStreamSynchronize(s));
      be = 0.0;
      for (int i = 0; i < n; i++)
        be = std::max(be, w[i] > safe2 ? cabs1(hr[i]) / w[i]
                                       : (cabs1(hr[i]) + safe1) / (w[i] + safe1));
      if (be > eps && 2.0 * be <= lstres && count <= 5) {
        // x += A^{-1} r
        if ((rc = lu_apply(f, dr, tmp, 2, s))) return rc;
        hipLaunchKernelGGL((k_axpy_conv<T, T>), dim3((n + 255) / 256), dim3(256), 0, s,
                           (const T*)dr, x, (long long)n);
        lstres = be;
        count++;
        continue;
      }
      break;
    }
    berr[j] = be;
    *steps = std::max(*steps, count - 1);
    // error bound: ||inv(A) diag(W)||_inf estimated, W = |r| + nz eps (|A| |x| + |b|)
    for (int i = 0; i < n; i++)
      w[i] = w[i] > safe2 ? cabs1(hr[i]) + nz * eps * w[i] : cabs1(hr[i]) + nz * eps * w[i] + safe1;
    double fe = 0.0;
    rc = lacn2(n, [&](int kase, std::vector<T>& v) {
      if (kase == 1) {  // diag(W) inv(A^H)
        const int r1 = dev_apply<T>(v, dv, tmp, [&](T* d) { return lu_apply(f, d, tmp, 3, s); }, s);
        for (int i = 0; i < n; i++) v[i] = tscale_h(v[i], w[i]);
        return r1;
      }
      for (int i = 0; i < n; i++) v[i] = tscale_h(v[i], w[i]);   // inv(A) diag(W)
      return dev_apply<T>(v, dv, tmp, [&](T* d) { return lu_apply(f, d, tmp, 2, s); }, s);
    }, &fe);
    if (rc) return rc;
    MFB_HIP(hipMemcpy(hx.data(), x, sizeof(T) * n, hipMemcpyDeviceToHost));
    double xm = 0.0;
    for (int i = 0; i < n; i++) xm = std::max(xm, cabs1(hx[i]));
    ferr[j] = xm != 0.0 ? fe / xm : fe;
  }
  return MFB_OK;
}

template <typename T>
static int solve_lse(Kept<T>& k, int n, T* A, int* ipiv, int aodim, T* Ao, char* equed, double* r,
                     double* cvec, int n_rhs, T* b, int factorize, int scaling, int condition,
                     int refine, int precision_mode, mfb_solve_report* rep) {
  typedef typename LowOf<T>::T Tl;
  MFB_ARG(n > 0 && n_rhs > 0 && A && b, "invalid solve_lse arguments");
  MFB_ARG(precision_mode == MFB_PREC_FP64 || is_mixed(precision_mode), "unknown precision_mode");
  MFB_ARG(!scaling || (equed && r && cvec), "scaling needs equed, r and c");
  // the condition and refine stages work on the unfactored matrix Ao (the caller's, Aodim = n)
  // and on FP64 factors: with either of them the factorisation is FP64 whatever the mode
  const bool cr = condition || refine;
  MFB_ARG(!cr || (aodim == n && Ao), "invalid value of Aodim (condition / refine need Aodim = n_dof)");
  if (rep) std::memset(rep, 0, sizeof *rep);
  hipStream_t s = nullptr;
  const size_t nn = (size_t)n * n;
  auto t0 = std::chrono::steady_clock::now();
  if (factorize) {
    kept_release(k);
    MFB_HIP(hipMalloc(&k.dA, sizeof(T) * nn));
    MFB_HIP(hipMemcpy(k.dA, A, sizeof(T) * nn, hipMemcpyHostToDevice));
    k.key = A;
    k.n = n;
    k.mode = precision_mode;
    int rc = MFB_OK;
    if (scaling) rc = equilibrate(k.dA, n, r, cvec, equed, s);
    if (!rc && cr) {
      // Ao = A (after the scaling), on the device for the stages and in the caller's array
      hipError_t he = hipMalloc(&k.dAo, sizeof(T) * nn);
      if (he == hipSuccess) he = hipMemcpy(k.dAo, k.dA, sizeof(T) * nn, hipMemcpyDeviceToDevice);
      if (he == hipSuccess) he = hipMemcpy(Ao, k.dAo, sizeof(T) * nn, hipMemcpyDeviceToHost);
      if (he != hipSuccess) {
        set_error("solve_lse: %s", hipGetErrorString(he));
        rc = MFB_ERR_HIP;
      }
    }
    if (!rc) rc = kept_factor(k, cr ? MFB_PREC_FP64 : precision_mode, s);
    if (rc) {
      if (rep) rep->info = k.f.info;
      kept_release(k);
      return rc;
    }
    if (condition) {
      // anorm = ||Ao||_1, then ?gecon on the FP64 factors
      double* cs = nullptr;
      rc = ws_get(ctx(), WS_CS, sizeof(double) * n, &cs);
      if (rc) return rc;
      hipLaunchKernelGGL(k_colabs<T>, dim3(n), dim3(256), 0, s, (const T*)k.dAo, (long long)n, n,
                         cs);
      std::vector<double> hcs(n);
      MFB_HIP(hipMemcpy(hcs.data(), cs, sizeof(double) * n, hipMemcpyDeviceToHost));
      double anorm = 0.0;
      for (double v : hcs) anorm = std::max(anorm, v);
      double rcond = 0.0;
      rc = gecon(k.f, anorm, &rcond, s);
      if (rc) return rc;
      if (rep) rep->rcond = rcond;
    }
  } else {
    if (!(k.key == A && k.n == n && k.mode == precision_mode && k.dA)) {
      set_error("factorize = 0 needs the factor of an earlier factorize = 1 call on the same A "
                "(same array, n and precision mode)");
      return MFB_ERR_ARG;
    }
    if (scaling && !(*equed == 'N' || *equed == 'R' || *equed == 'C' || *equed == 'B')) {
      set_error("invalid value of equed");
      return MFB_ERR_ARG;
    }
    if (refine && k.eff != MFB_PREC_FP64) {
      set_error("refine on a re-solve needs the FP64 factor of a factorize = 1 call made with "
                "condition or refine");
      return MFB_ERR_ARG;
    }
    if (refine && !k.dAo) {   // the caller's Ao (the matrix as factorised)
      MFB_HIP(hipMalloc(&k.dAo, sizeof(T) * nn));
      MFB_HIP(hipMemcpy(k.dAo, Ao, sizeof(T) * nn, hipMemcpyHostToDevice));
    }
  }
  if (rep) rep->factor_ms = factorize ? ms_since(t0) : 0.0;
  auto t1 = std::chrono::steady_clock::now();
  T* dB = nullptr;
  MFB_HIP(hipMalloc(&dB, sizeof(T) * (size_t)n * n_rhs));
  struct FreeB {
    T* p;
    ~FreeB() { (void)hipFree(p); }
  } fb{dB};
  MFB_HIP(hipMemcpy(dB, b, sizeof(T) * (size_t)n * n_rhs, hipMemcpyHostToDevice));
  int rc = MFB_OK;
  if (scaling && (*equed == 'R' || *equed == 'B')) rc = scale_rows(dB, n, n_rhs, r, s);
  if (rc) return rc;
  int it = 0;
  double berr = 0;
  if (is_mixed(k.eff)) {
    bool ok = false;
    rc = mixed_refine(k.mx, dB, n, n_rhs, &ok, &it, &berr, s);
    if (rc) return rc;
    if (!ok) {
      // refinement stalled: the kept matrix gets the FP64 factorisation (from now on as well)
      for (void* p : k.owned)
        if (p) (void)hipFree(p);
      k.owned.clear();
      k.mx = MixedLU<T, Tl>();
      rc = kept_factor(k, MFB_PREC_FP64, s);
      if (rc) return rc;
      if (rep) rep->fell_back = 1;
    }
  }
  if (k.eff == MFB_PREC_FP64) {
    T* tmp = nullptr;
    T* dBc = nullptr;   // the right-hand sides before the solve (refine stage)
    if (refine) MFB_HIP(hipMalloc(&dBc, sizeof(T) * (size_t)n * n_rhs));
    struct FreeC {
      T* p;
      ~FreeC() { if (p) (void)hipFree(p); }
    } fc{dBc};
    if (refine)
      MFB_HIP(hipMemcpy(dBc, dB, sizeof(T) * (size_t)n * n_rhs, hipMemcpyDeviceToDevice));
    rc = ws_get(ctx(), WS_TMP, sizeof(T) * (size_t)n * n_rhs, &tmp);
    if (!rc) rc = getrs(k.f, dB, n, n_rhs, tmp, s);
    if (rc) return rc;
    if (refine) {
      std::vector<double> be(n_rhs), fe(n_rhs);
      int steps = 0;
      rc = gerfs((const T*)k.dAo, k.f, (const T*)dBc, dB, n_rhs, be.data(), fe.data(), &steps, s);
      if (rc) return rc;
      berr = *std::max_element(be.begin(), be.end());
      it = steps;
      if (rep) rep->ferr = *std::max_element(fe.begin(), fe.end());
    }
  }
  if (scaling && (*equed == 'C' || *equed == 'B')) rc = scale_rows(dB, n, n_rhs, cvec, s);
  if (rc) return rc;
  MFB_HIP(hipMemcpy(b, dB, sizeof(T) * (size_t)n * n_rhs, hipMemcpyDeviceToHost));
  if (factorize) {
    // A out: LAPACK-format factors (FP64 mode: ?getrf's; mixed: the FP32 factors of the scaled
    // matrix written as FP64 factors of A, U times C^{-1}), ipiv: 1-based pivots
    const bool mixed = is_mixed(k.eff);
    if (mixed) {
      T* out = nullptr;
      MFB_HIP(hipMalloc(&out, sizeof(T) * nn));
      hipLaunchKernelGGL((k_factor_out<Tl, T>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                         (const Tl*)k.mx.f.A, n, (const double*)k.mx.cs, out);
      hipError_t he = hipMemcpy(A, out, sizeof(T) * nn, hipMemcpyDeviceToHost);
      (void)hipFree(out);
      if (he != hipSuccess) {
        set_error("solve_lse: %s", hipGetErrorString(he));
        return MFB_ERR_HIP;
      }
    } else {
      MFB_HIP(hipMemcpy(A, k.dA, sizeof(T) * nn, hipMemcpyDeviceToHost));
    }
    if (ipiv) {
      const int* dp = mixed ? k.mx.f.ipiv : k.f.ipiv;
      MFB_HIP(hipMemcpy(ipiv, dp, sizeof(int) * n, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; i++) ipiv[i] += 1;
    }
  }
  if (rep) {
    rep->precision_mode = k.eff;
    rep->refinement_steps = it;
    rep->backward_error = berr;
    rep->solve_ms = ms_since(t1);
    rep->info = 0;
  }
  return MFB_OK;
}

// Release everything the solver keeps between calls: kept factors, rocBLAS handles, the second
// stream and its events, workspaces and the persistent kernels' posting buffers.
static void shutdown_solver() {
  kept_release(kept_r);
  kept_release(kept_c);
  Ctx& c = ctx();
  if (c.device >= 0) (void)hipSetDevice(c.device);
  if (c.s2) (void)hipStreamSynchronize(c.s2);
  (void)hipDeviceSynchronize();
  if (c.blas) rocblas_destroy_handle(c.blas);
  if (c.blas2) rocblas_destroy_handle(c.blas2);
  if (c.s2) (void)hipStreamDestroy(c.s2);
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_join) (void)hipEventDestroy(c.ev_join);
  void* bufs[] = {c.ps.val, c.ps.idx, c.ps.counter, c.ps.info, c.moves, c.nmoves, c.pan.rec,
                  c.pan.err, c.pan.crow, c.pan.rowj, c.sol.rec, c.sol.err};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < WS_COUNT; k++)
    if (c.ws[k]) (void)hipFree(c.ws[k]);
  c = Ctx();
}

}  // namespace mfb

extern "C" int mfb_solve_lse_r(int n_dof, double* A, int* ipiv, int aodim, double* Ao, char* equed,
                               double* r, double* c, int n_rhs, double* b, int factorize,
                               int scaling, int condition, int refine, int precision_mode,
                               mfb_solve_report* report) {
  if (equed && !scaling) *equed = 'N';
  return solve_lse(kept_r, n_dof, A, ipiv, aodim, Ao, equed, r, c, n_rhs, b, factorize, scaling,
                   condition, refine, precision_mode, report);
}

extern "C" int mfb_solve_lse_c(int n_dof, double* A, int* ipiv, int aodim, double* Ao, char* equed,
                               double* r, double* c, int n_rhs, double* b, int factorize,
                               int scaling, int condition, int refine, int precision_mode,
                               mfb_solve_report* report) {
  if (equed && !scaling) *equed = 'N';
  return solve_lse(kept_c, n_dof, (cz*)A, ipiv, aodim, (cz*)Ao, equed, r, c, n_rhs, (cz*)b,
                   factorize, scaling, condition, refine, precision_mode, report);
}

extern "C" int mfb_shutdown(void) {
  shutdown_solver();
  return MFB_OK;
}

extern "C" int mfb_solver_set_schedule(int lookahead, int la_chunk, int trsv_parts) {
  MFB_ARG(lookahead >= -1 && lookahead <= 1, "invalid value of lookahead (-1, 0 or 1)");
  MFB_ARG(la_chunk >= -1, "invalid value of la_chunk (-1, 0 or a column count)");
  MFB_ARG(trsv_parts == -1 || trsv_parts == 1 || trsv_parts == 2 || trsv_parts == 4,
          "invalid value of trsv_parts (-1, 1, 2 or 4)");
  Schedule& g = schedule();
  g.lookahead = lookahead;
  g.la_chunk = la_chunk > 0 ? std::max(NB, la_chunk / NB * NB) : la_chunk;
  g.trsv_parts = trsv_parts;
  return MFB_OK;
}

extern "C" int mfb_solver_get_schedule(int* lookahead, int* la_chunk, int* trsv_parts) {
  const Schedule& g = schedule();
  if (lookahead) *lookahead = g.lookahead;
  if (la_chunk) *la_chunk = g.la_chunk;
  if (trsv_parts) *trsv_parts = g.trsv_parts;
  return MFB_OK;
}

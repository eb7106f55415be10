// Device core of the layered LDPC decoders (shared by the per-wave and per-workgroup
// decoder kernels of decode.hip and the mixed-code queue kernel of multi.hip).
//
// Semantics restated from src/LDPCDec/LDPCdecode.cpp:79-201 and
// decode_functions.cpp:11-137 (per check row r of layer l, edge j on column c_j
// with relabeled shift s_j, variable v_j = c_j*Z + (s_j + r) mod Z):
//   Q_j = sat(L[v_j] - R_j);  A_j = max(wrap(wabs(Q_j) - beta), 0);  S = xor of signs
//   M_j = min(cap, min_{k!=j} A_k)          (min-sum, exact whatever the order)
//   M_j = 2 atanh(prod-tree_{k!=j} tanh(A_k/2), cap clamps)   (sum-product)
//   R_j = (S xor sign Q_j) ? wrap(-M_j) : M_j;  L[v_j] = sat(Q_j + R_j)
// then after every full iteration the syndrome (sign bits) decides early exit.
//
// MI355X mapping: one wave (64 lanes) owns `cw_per_wave` codewords; a codeword's
// posterior and C2V messages live in that wave's LDS slice, its Z check rows are
// spread over `lanes_per_cw` lanes (rows_per_lane each).  A codeword never spans two
// waves, so layers need no workgroup barrier: LDS operations of one wave complete in
// order, and a wavefront-scope fence keeps the compiler from reordering them.
// Early exit is a 64-bit ballot per iteration; lanes of finished codewords idle.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <limits>
#include <type_traits>

#include "fecmath.h"
#include "internal.h"
#include "prodtree_gen.h"

namespace fecsim {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// element policies
// ---------------------------------------------------------------------------

template <int BITS>
struct IntPolicy {
  using S = typename std::conditional<BITS == 8, int8_t,
                                      typename std::conditional<BITS == 16, int16_t, int32_t>::type>::type;
  using C = typename std::conditional<BITS == 32, long long, int>::type;   // compute type
  using K = typename std::conditional<BITS == 32, long long, int>::type;   // min key type
  static constexpr bool kFloat = false;
  static constexpr bool kProd = false;
  static constexpr C LO = -(C(1) << (BITS - 1));
  static constexpr C HI = (C(1) << (BITS - 1)) - 1;
  __device__ static __forceinline__ C sat(C x) { return x < LO ? LO : (x > HI ? HI : x); }
  __device__ static __forceinline__ C wrap(C x) { return (C)(S)x; }
  __device__ static __forceinline__ C wabs(C x) { return x == LO ? LO : (x < 0 ? -x : x); }
};

template <typename F, bool PROD>
struct FloatPolicy {
  using S = F;
  static constexpr bool kFloat = true;
  static constexpr bool kProd = PROD;
};

__device__ __forceinline__ bool sign_of(float x) { return (__float_as_uint(x) >> 31) != 0; }
__device__ __forceinline__ bool sign_of(double x) { return __double_as_longlong(x) < 0; }

template <class P>
__device__ __forceinline__ bool is_neg(typename P::S x) {
  if constexpr (P::kFloat) return sign_of(x);
  else return x < 0;
}

// ---------------------------------------------------------------------------
// codeword transfer between global memory and LDS
// ---------------------------------------------------------------------------

// Rows r0, r0 + step, ... of every column, eight columns per batch so eight loads
// are in flight before the first store waits (the last batch re-reads the last
// column instead of branching).  `gs` is the column stride in global memory; LDS
// columns are packed (stride Z).
template <class S>
__device__ __forceinline__ void cols_to_lds(S* __restrict__ post, const S* __restrict__ g, long long base,
                                            long long gs, int mcol, int Z, int r0, int step) {
  for (int r = r0; r < Z; r += step)
    for (int c0 = 0; c0 < mcol; c0 += 8) {
      S t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = g[base + min(c0 + u, mcol - 1) * gs + r];   // unconditional: one wait
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (c0 + u < mcol) post[(c0 + u) * Z + r] = t[u];
    }
}

template <class S>
__device__ __forceinline__ void cols_from_lds(S* __restrict__ g, const S* __restrict__ post, long long base,
                                              long long gs, int mcol, int Z, int r0, int step) {
  for (int r = r0; r < Z; r += step)
    for (int c0 = 0; c0 < mcol; c0 += 8) {
      S t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = post[min(c0 + u, mcol - 1) * Z + r];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (c0 + u < mcol) g[base + (c0 + u) * gs + r] = t[u];
    }
}

// ---------------------------------------------------------------------------
// edge tables
// ---------------------------------------------------------------------------

// Each block first copies the code's edge words and layer weights into the front of
// its LDS ([edges | layer weights]); the row loops then read them with broadcast LDS
// loads instead of global loads on their critical path.
__device__ __forceinline__ void tables_to_lds(uint32_t* __restrict__ tab, const uint32_t* __restrict__ edges, int ne,
                                              const int32_t* __restrict__ lw, int mrow) {
  for (int i = threadIdx.x; i < ne; i += blockDim.x) tab[i] = edges[i];
  for (int i = threadIdx.x; i < mrow; i += blockDim.x) tab[ne + i] = (uint32_t)lw[i];
}

// LDS bytes of that copy for one code, 16-byte aligned
inline int table_bytes(int nedge, int mrow) { return ((nedge + mrow) * 4 + 15) / 16 * 16; }

// variable index of edge word `ed` for row r: (col*z) + (shift + r) mod z
__device__ __forceinline__ int edge_var(uint32_t ed, int r, int Z) {
  int v = (int)(ed & 0xffffu) + r;
  v = v >= Z ? v - Z : v;
  return v + (int)(ed >> 16);
}

// ---------------------------------------------------------------------------
// one check row, integer elements (min-sum)
// ---------------------------------------------------------------------------

template <class P, int MAXW>
__device__ __forceinline__ void row_int(typename P::S* __restrict__ post, typename P::S* __restrict__ c2v,
                                        const uint32_t* __restrict__ edges, int e0, int w, int r, int Z,
                                        typename P::C beta, typename P::C cap) {
  using C = typename P::C;
  using K = typename P::K;
  C q[MAXW];
  int vi[MAXW];
  K m1 = std::numeric_limits<K>::max(), m2 = std::numeric_limits<K>::max();
  int sg = 0;
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      const int v = edge_var(edges[e0 + j], r, Z);
      vi[j] = v;
      const C qq = P::sat((C)post[v] - (C)c2v[(e0 + j) * Z + r]);
      q[j] = qq;
      C a = P::wrap(P::wabs(qq) - beta);
      a = a > 0 ? a : 0;
      sg ^= (qq < 0);
      const K key = ((K)a << 6) | j;
      const K hi = key > m1 ? key : m1;
      m2 = hi < m2 ? hi : m2;
      m1 = key < m1 ? key : m1;
    }
  }
  const int im = (int)(m1 & 63);
  C M1 = (C)(m1 >> 6), M2 = (C)(m2 >> 6);
  M1 = cap < M1 ? cap : M1;
  M2 = cap < M2 ? cap : M2;
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      const C m = (j == im) ? M2 : M1;
      const bool neg = sg ^ (q[j] < 0);
      const C rn = neg ? P::wrap(-m) : m;
      post[vi[j]] = (typename P::S)P::sat(q[j] + rn);
      c2v[(e0 + j) * Z + r] = (typename P::S)rn;
    }
  }
}

// ---------------------------------------------------------------------------
// one check row, int8 min-sum with beta >= 1 and the row weight W known at compile
// time: straight-line code, all of the row's LDS reads issued before the first use.
// Same results as row_int: with beta >= 1, wrap8(wabs8(Q) - beta) = |Q| - beta also
// for Q = -128, so A_j = max(|Q_j| - beta, 0) with the true |Q_j| <= 128, and
//   mv1 = min(cap, max(m1, beta) - beta),  mv2 = min(max(m2, beta), cap + beta) - beta,
//   M_j = (j is the argmin ? mv2 : mv1) = max(mv1, m1 + mv2 - |Q_j|)
// (m1 <= m2 the two smallest |Q|): at the argmin the term is mv2 >= mv1; at any other
// edge |Q_j| >= m2 makes it <= m1 + mv2 - m2, which is <= mv1 (<= m1 - beta when the
// cap does not bind, <= cap = mv1 = mv2 when it does, < 0 = mv1 when m1 < beta, and
// <= 0 = mv1 = mv2 when m2 < beta).  Ties included.
// ---------------------------------------------------------------------------

// Edge words come from the code's table in global memory: the row's words are uniform
// across the wave, so they load through the scalar cache into SGPRs and the per-edge
// wrap test (s + r >= Z) is one compare of an SGPR against a lane constant.
typedef __attribute__((address_space(4))) const uint32_t ConstU32;
// clamp as one v_med3_i32 (left to itself the compiler rewrites the int8 saturations
// as 16-bit shift / saturating-subtract / shift sequences)
__device__ __forceinline__ int med3_i32(int a, int b, int c) {
  int d;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
__device__ __forceinline__ ConstU32* uniform_table(const uint32_t* p) {
  const unsigned long long v = (unsigned long long)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return (ConstU32*)(((unsigned long long)hi << 32) | lo);
}

template <int W>
__device__ __forceinline__ void row_i8_w(int8_t* __restrict__ post, int8_t* __restrict__ c2v,
                                         const uint32_t* __restrict__ gedges, int e0, int r, int Z, int beta,
                                         int cap) {
  int8_t* vp[W];
  int lv[W], rv[W];
  int8_t* cr = c2v + e0 * Z + r;
  int8_t* const pr = post + r;         // variable (s + r) of column c: pr + c*Z + s, or pr - Z + ... on wrap
  int8_t* const prz = post + r - Z;
  const int zr = Z - r;
  ConstU32* const ge = uniform_table(gedges + e0);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t ed = ge[j];
    const int sh = (int)(ed & 0xffffu), cz = (int)(ed >> 16);
    vp[j] = (sh >= zr ? prz : pr) + (cz + sh);
  }
#pragma unroll
  for (int j = 0; j < W; ++j) {
    lv[j] = *vp[j];
    rv[j] = cr[j * Z];
  }
  int q[W];
  uint32_t u[W];
  int S = 0;
  uint32_t m1 = 255u, m2 = 255u;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const int qq = med3_i32(lv[j] - rv[j], -128, 127);   // sat8(L - R)
    q[j] = qq;
    const uint32_t a = (uint32_t)(qq < 0 ? -qq : qq);
    u[j] = a;
    S ^= qq;
    const uint32_t hi = m1 > a ? m1 : a;
    m2 = hi < m2 ? hi : m2;   // median of (m1, a, m2)
    m1 = a < m1 ? a : m1;
  }
  const int b = beta;
  const int m1p = (int)m1 > b ? (int)m1 : b, m2p = (int)m2 > b ? (int)m2 : b;
  const int mv1 = m1p - b < cap ? m1p - b : cap;
  const int mv2 = (m2p < cap + b ? m2p : cap + b) - b;
  const int sumb = (int)m1 + mv2;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const int t = sumb - (int)u[j];
    const int m = t > mv1 ? t : mv1;
    const int neg = (S ^ q[j]) >> 31;   // 0 or -1
    const int rn = (m ^ neg) - neg;
    const int g = q[j] + rn;
    *vp[j] = (int8_t)med3_i32(g, -128, 127);
    cr[j * Z] = (int8_t)rn;
  }
}

// the same check row's parity (sign bits of its int8 variables)
template <int W>
__device__ __forceinline__ bool row_parity_i8_w(const int8_t* __restrict__ post, const uint32_t* __restrict__ gedges,
                                                int e0, int r, int Z) {
  const int8_t* const pr = post + r;
  const int8_t* const prz = post + r - Z;
  const int zr = Z - r;
  ConstU32* const ge = uniform_table(gedges + e0);
  int x = 0;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t ed = ge[j];
    const int sh = (int)(ed & 0xffffu), cz = (int)(ed >> 16);
    x ^= *((sh >= zr ? prz : pr) + (cz + sh));
  }
  return x < 0;
}

// dispatch on the run-time row weight w (uniform across the wave); false when w is
// outside 2..WMAX
template <int W, int WMAX>
__device__ __forceinline__ bool row_i8_sw(int w, int8_t* post, int8_t* c2v, const uint32_t* edges, int e0, int r,
                                          int Z, int beta, int cap) {
  if constexpr (W > WMAX) {
    return false;
  } else {
    if (w == W) {
      row_i8_w<W>(post, c2v, edges, e0, r, Z, beta, cap);
      return true;
    }
    return row_i8_sw<W + 1, WMAX>(w, post, c2v, edges, e0, r, Z, beta, cap);
  }
}
template <int W, int WMAX>
__device__ __forceinline__ bool parity_i8_sw(int w, const int8_t* post, const uint32_t* edges, int e0, int r, int Z,
                                             bool* fail) {
  if constexpr (W > WMAX) {
    return false;
  } else {
    if (w == W) {
      *fail |= row_parity_i8_w<W>(post, edges, e0, r, Z);
      return true;
    }
    return parity_i8_sw<W + 1, WMAX>(w, post, edges, e0, r, Z, fail);
  }
}

// parity of row r of a layer (sign bits of its variables)
template <class P, int MAXW>
__device__ __forceinline__ bool row_parity(const typename P::S* __restrict__ post, const uint32_t* __restrict__ edges,
                                           int e0, int w, int r, int Z) {
  bool par = false;
#pragma unroll
  for (int j = 0; j < MAXW; ++j)
    if (j < w) par ^= is_neg<P>(post[edge_var(edges[e0 + j], r, Z)]);
  return par;
}

// ---------------------------------------------------------------------------
// one check row, float elements (min-sum or sum-product)
// ---------------------------------------------------------------------------

// The sum-product factors tanh(A/2) and messages 2 atanh(p): float32 from the hardware
// approximations (within 4 units of the correctly rounded values), double from the
// series that the oracle restates bit for bit.
__device__ __forceinline__ float sp_tanh_half(float a) { return fecmath::tanh_half_fast(a); }
__device__ __forceinline__ double sp_tanh_half(double a) { return fecmath::tanh_own(a * 0.5); }
__device__ __forceinline__ float sp_two_atanh(float p) { return fecmath::two_atanh_fast(p); }
__device__ __forceinline__ double sp_two_atanh(double p) { return 2.0 * fecmath::atanh_own(p); }

template <class P, int MAXW>
__device__ __forceinline__ void row_float(typename P::S* __restrict__ post, typename P::S* __restrict__ c2v,
                                          const uint32_t* __restrict__ edges, int e0, int w, int r, int Z,
                                          typename P::S beta, typename P::S cap) {
  using F = typename P::S;
  F q[MAXW], A[MAXW], M[MAXW];
  int vi[MAXW];
  int sg = 0;
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      const int v = edge_var(edges[e0 + j], r, Z);
      vi[j] = v;
      const F qq = post[v] - c2v[(e0 + j) * Z + r];
      q[j] = qq;
      const F a = fabs(qq) - beta;
      A[j] = a > F(0) ? a : F(0);   // _mm_max_ps(a, 0)
      sg ^= (int)sign_of(qq);
    }
  }
  if constexpr (P::kProd) {
    F t[MAXW];
#pragma unroll
    for (int j = 0; j < MAXW; ++j) t[j] = (j < w) ? sp_tanh_half(A[j]) : F(0);
    F o[MAXW];
#pragma unroll
    for (int j = 0; j < MAXW; ++j) o[j] = F(0);
    prod_tree_dyn<F>(w, t, cap, o);
#pragma unroll
    for (int j = 0; j < MAXW; ++j) M[j] = (j < w) ? sp_two_atanh(o[j]) : F(0);
  } else {
    F m1 = F(INFINITY), m2 = F(INFINITY);
    int im = -1;
#pragma unroll
    for (int j = 0; j < MAXW; ++j) {
      if (j < w) {
        const F a = A[j];
        const F hi = a > m1 ? a : m1;
        m2 = hi < m2 ? hi : m2;
        im = a < m1 ? j : im;
        m1 = a < m1 ? a : m1;
      }
    }
    const F M1 = cap < m1 ? cap : m1, M2 = cap < m2 ? cap : m2;
#pragma unroll
    for (int j = 0; j < MAXW; ++j) M[j] = (j == im) ? M2 : M1;
  }
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      const bool neg = sg ^ (int)sign_of(q[j]);
      const F rn = neg ? -M[j] : M[j];
      post[vi[j]] = q[j] + rn;
      c2v[(e0 + j) * Z + r] = rn;
    }
  }
}

// The same row with the row weight W known at compile time: straight-line code, the
// row's edge words through the scalar cache (as row_i8_w), tanh/atanh evaluated for
// the W edges only.  The arithmetic and its order are row_float's, so the results are
// identical bit for bit.
template <class P, int W>
__device__ __forceinline__ void row_float_w(typename P::S* __restrict__ post, typename P::S* __restrict__ c2v,
                                           const uint32_t* __restrict__ gedges, int e0, int r, int Z,
                                           typename P::S beta, typename P::S cap) {
  using F = typename P::S;
  F* vp[W];
  F lv[W], rv[W];
  F* cr = c2v + e0 * Z + r;
  F* const pr = post + r;
  F* const prz = post + r - Z;
  const int zr = Z - r;
  ConstU32* const ge = uniform_table(gedges + e0);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t ed = ge[j];
    const int sh = (int)(ed & 0xffffu), cz = (int)(ed >> 16);
    vp[j] = (sh >= zr ? prz : pr) + (cz + sh);
  }
#pragma unroll
  for (int j = 0; j < W; ++j) {
    lv[j] = *vp[j];
    rv[j] = cr[j * Z];
  }
  F q[W], A[W], M[W];
  int sg = 0;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const F qq = lv[j] - rv[j];
    q[j] = qq;
    const F a = fabs(qq) - beta;
    A[j] = a > F(0) ? a : F(0);   // _mm_max_ps(a, 0)
    sg ^= (int)sign_of(qq);
  }
  if constexpr (P::kProd) {
    F t[W], o[W];
#pragma unroll
    for (int j = 0; j < W; ++j) t[j] = sp_tanh_half(A[j]);
    prod_tree_dyn<F>(W, t, cap, o);
#pragma unroll
    for (int j = 0; j < W; ++j) M[j] = sp_two_atanh(o[j]);
  } else {
    F m1 = F(INFINITY), m2 = F(INFINITY);
    int im = -1;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const F a = A[j];
      const F hi = a > m1 ? a : m1;
      m2 = hi < m2 ? hi : m2;
      im = a < m1 ? j : im;
      m1 = a < m1 ? a : m1;
    }
    const F M1 = cap < m1 ? cap : m1, M2 = cap < m2 ? cap : m2;
#pragma unroll
    for (int j = 0; j < W; ++j) M[j] = (j == im) ? M2 : M1;
  }
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const bool neg = sg ^ (int)sign_of(q[j]);
    const F rn = neg ? -M[j] : M[j];
    *vp[j] = q[j] + rn;
    cr[j * Z] = rn;
  }
}

// parity of a float row of compile-time weight (sign bits, -0.0 counts as negative)
template <class P, int W>
__device__ __forceinline__ bool row_parity_float_w(const typename P::S* __restrict__ post,
                                                   const uint32_t* __restrict__ gedges, int e0, int r, int Z) {
  using F = typename P::S;
  const F* const pr = post + r;
  const F* const prz = post + r - Z;
  const int zr = Z - r;
  ConstU32* const ge = uniform_table(gedges + e0);
  bool par = false;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t ed = ge[j];
    const int sh = (int)(ed & 0xffffu), cz = (int)(ed >> 16);
    par ^= sign_of(*((sh >= zr ? prz : pr) + (cz + sh)));
  }
  return par;
}

template <class P, int W, int WMAX>
__device__ __forceinline__ bool row_float_sw(int w, typename P::S* post, typename P::S* c2v, const uint32_t* edges,
                                             int e0, int r, int Z, typename P::S beta, typename P::S cap) {
  if constexpr (W > WMAX) {
    return false;
  } else {
    if (w == W) {
      row_float_w<P, W>(post, c2v, edges, e0, r, Z, beta, cap);
      return true;
    }
    return row_float_sw<P, W + 1, WMAX>(w, post, c2v, edges, e0, r, Z, beta, cap);
  }
}
template <class P, int W, int WMAX>
__device__ __forceinline__ bool parity_float_sw(int w, const typename P::S* post, const uint32_t* edges, int e0, int r,
                                                int Z, bool* fail) {
  if constexpr (W > WMAX) {
    return false;
  } else {
    if (w == W) {
      *fail |= row_parity_float_w<P, W>(post, edges, e0, r, Z);
      return true;
    }
    return parity_float_sw<P, W + 1, WMAX>(w, post, edges, e0, r, Z, fail);
  }
}

// float rows are specialized per weight for kernels whose weight bound is <= 8 (the
// codes of the float workloads: 16e, 11n and 11ad rate 1/2, weights 5..8)
template <class P, int MAXW>
constexpr bool kRowFloatFast = P::kFloat && MAXW <= 8;

// ---------------------------------------------------------------------------
// the per-wave decode loop: every lane returns its codeword's iteration count
// ---------------------------------------------------------------------------

// the straight-line int8 rows are compiled for kernels whose row weight bound is <= 24
// (one instance per weight 2..MAXW)
template <class P, int MAXW>
constexpr bool kRowI8Fast = std::is_same<P, IntPolicy<8>>::value && MAXW <= 24;

// `tab`: the code's [edges | layer weights] copy in LDS
template <class P, int MAXW>
__device__ __forceinline__ uint32_t decode_loop(const DecCore& a, const uint32_t* __restrict__ tab,
                                                typename P::S* post, typename P::S* c2v, int slot, int lr, bool valid) {
  using S = typename P::S;
  const int L = a.lanes_per_cw, Z = a.z;
  const uint32_t* edges = tab;
  const int32_t* layer_w = reinterpret_cast<const int32_t*>(tab + a.nedge);
  if (valid)
    for (int i = lr; i < a.e; i += L) c2v[i] = S(0);   // LDPCdecode.cpp:115-119
  wave_sync();
  const S fbeta = P::kFloat ? (S)a.fbeta : S(0), fcap = P::kFloat ? (S)a.fcap : S(0);
  uint32_t iter = 0;
  bool active = valid;
  const unsigned long long gmask = (L >= 64) ? ~0ull : ((1ull << L) - 1ull);
  while (__ballot(active) != 0ull) {
    bool fail = false;
    if (active) {
      int e0 = 0;
      for (int l = 0; l < a.mrow; ++l) {
        const int w = layer_w[l];
        S fb = fbeta;
        long long ib = a.ibeta;
        if (a.beta_tab) {   // per-(iteration, layer) offsets
          if constexpr (P::kFloat) fb = (S)reinterpret_cast<const double*>(a.beta_tab)[iter * a.mrow + l];
          else ib = reinterpret_cast<const long long*>(a.beta_tab)[iter * a.mrow + l];
        }
        // int8 with beta >= 1 and cap >= 0: straight-line rows of compile-time weight
        bool fast = false;
        if constexpr (kRowI8Fast<P, MAXW>) fast = ib >= 1 && a.icap >= 0 && w >= 2 && w <= MAXW;
        for (int k = 0; k < a.rows_per_lane; ++k) {
          const int r = lr + k * L;
          if (r < Z) {
            if constexpr (P::kFloat) {
              if constexpr (kRowFloatFast<P, MAXW>)
                if (row_float_sw<P, 2, MAXW>(w, post, c2v, a.edges, e0, r, Z, fb, fcap)) continue;
              row_float<P, MAXW>(post, c2v, edges, e0, w, r, Z, fb, fcap);
            } else {
              if constexpr (kRowI8Fast<P, MAXW>)
                if (fast && row_i8_sw<2, MAXW>(w, post, c2v, a.edges, e0, r, Z, (int)ib, (int)a.icap)) continue;
              row_int<P, MAXW>(post, c2v, edges, e0, w, r, Z, (typename P::C)ib, (typename P::C)a.icap);
            }
          }
        }
        e0 += w;
        wave_sync();
      }
      e0 = 0;
      for (int l = 0; l < a.mrow; ++l) {                  // LDPCdecode.cpp:172-195
        const int w = layer_w[l];
        for (int k = 0; k < a.rows_per_lane; ++k) {
          const int r = lr + k * L;
          if (r < Z) {
            if constexpr (kRowI8Fast<P, MAXW>)
              if (parity_i8_sw<2, MAXW>(w, post, a.edges, e0, r, Z, &fail)) continue;
            if constexpr (kRowFloatFast<P, MAXW>)
              if (parity_float_sw<P, 2, MAXW>(w, post, a.edges, e0, r, Z, &fail)) continue;
            fail |= row_parity<P, MAXW>(post, edges, e0, w, r, Z);
          }
        }
        e0 += w;
      }
    }
    const unsigned long long fb = __ballot(active && fail);
    if (active) {
      ++iter;
      const bool cwfail = ((fb >> (slot * L)) & gmask) != 0ull;
      if (!cwfail || iter == a.max_iter) active = false;   // LDPCdecode.cpp:192-197
    }
  }
  wave_sync();
  return iter;
}

}  // namespace fecsim

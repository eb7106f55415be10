// Layered LDPC decoders for gfx950: offset min-sum on int8/int16/int32 with the
// reference's saturating arithmetic, and float/double min-sum and tanh/atanh
// sum-product.
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
  using S = typename std::conditional<BITS == 8, int8_t, typename std::conditional<BITS == 16, int16_t, int32_t>::type>::type;
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
      const uint32_t ed = edges[e0 + j];
      int v = (int)(ed & 0xffffu) + r;
      v = v >= Z ? v - Z : v;
      v += (int)(ed >> 16);
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
// one check row, float elements (min-sum or sum-product)
// ---------------------------------------------------------------------------

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
      const uint32_t ed = edges[e0 + j];
      int v = (int)(ed & 0xffffu) + r;
      v = v >= Z ? v - Z : v;
      v += (int)(ed >> 16);
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
    for (int j = 0; j < MAXW; ++j) t[j] = (j < w) ? fecmath::tanh_own(A[j] * F(0.5)) : F(0);
    F o[MAXW];
#pragma unroll
    for (int j = 0; j < MAXW; ++j) o[j] = F(0);
    prod_tree_dyn<F>(w, t, cap, o);
#pragma unroll
    for (int j = 0; j < MAXW; ++j) M[j] = F(2) * fecmath::atanh_own(o[j]);
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

template <class P>
__device__ __forceinline__ bool is_neg(typename P::S x) {
  if constexpr (P::kFloat) return sign_of(x);
  else return x < 0;
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------

template <class P, int MAXW>
__global__ void __launch_bounds__(256) k_decode(const KernelArgs a) {
  using S = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = a.lanes_per_cw, CPW = a.cw_per_wave, Z = a.z;
  const int slot = lane / L;
  const int lr = lane - slot * L;
  const long long cw = ((long long)blockIdx.x * a.waves + wave) * CPW + slot;
  const bool valid = slot < CPW && cw < (long long)a.num_cw;
  const int sl = slot < CPW ? slot : 0;
  char* wbase = smem + (size_t)wave * a.wave_bytes;
  S* post = reinterpret_cast<S*>(wbase) + (size_t)sl * a.post_stride;
  S* c2v = reinterpret_cast<S*>(wbase) + (size_t)CPW * a.post_stride + (size_t)sl * a.c2v_stride;
  const S* gin = reinterpret_cast<const S*>(a.metrics_in);
  S* gout = reinterpret_cast<S*>(a.metrics_out);

  if (valid) {
    const long long cbase = a.base + cw * a.cw_stride;
    for (int c = 0; c < a.mcol; ++c)
      for (int r = lr; r < Z; r += L) post[c * Z + r] = gin[cbase + c * a.col_stride + r];
    for (int i = lr; i < a.e; i += L) c2v[i] = S(0);
  }
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
        const int w = a.layer_w[l];
        for (int k = 0; k < a.rows_per_lane; ++k) {
          const int r = lr + k * L;
          if (r < Z) {
            if constexpr (P::kFloat) row_float<P, MAXW>(post, c2v, a.edges, e0, w, r, Z, fbeta, fcap);
            else row_int<P, MAXW>(post, c2v, a.edges, e0, w, r, Z, (typename P::C)a.ibeta, (typename P::C)a.icap);
          }
        }
        e0 += w;
        wave_sync();
      }
      e0 = 0;
      for (int l = 0; l < a.mrow; ++l) {
        const int w = a.layer_w[l];
        for (int k = 0; k < a.rows_per_lane; ++k) {
          const int r = lr + k * L;
          if (r < Z) {
            bool par = false;
            for (int j = 0; j < w; ++j) {
              const uint32_t ed = a.edges[e0 + j];
              int v = (int)(ed & 0xffffu) + r;
              v = v >= Z ? v - Z : v;
              v += (int)(ed >> 16);
              par ^= is_neg<P>(post[v]);
            }
            fail |= par;
          }
        }
        e0 += w;
      }
    }
    const unsigned long long fb = __ballot(active && fail);
    if (active) {
      ++iter;
      const bool cwfail = ((fb >> (slot * L)) & gmask) != 0ull;
      if (!cwfail || iter == a.max_iter) active = false;
    }
  }
  wave_sync();

  if (valid) {
    const long long cbase = a.base + cw * a.cw_stride;
    for (int c = 0; c < a.mcol; ++c)
      for (int r = lr; r < Z; r += L) gout[cbase + c * a.col_stride + r] = post[c * Z + r];
    if (a.c2v) {
      S* gc = reinterpret_cast<S*>(a.c2v) + cw * (long long)a.e;
      for (int i = lr; i < a.e; i += L) gc[i] = c2v[i];
    }
    if (lr == 0) {
      if (a.iters) a.iters[cw] = iter;
    }
  }
  if (a.tot_iter) {
    // one atomic per wave: lane 0 sums the slots' iteration counts
    unsigned long long s = 0;
    for (int q = 0; q < CPW; ++q) {
      const unsigned it = (unsigned)__shfl((int)iter, q * L, 64);
      const long long cq = ((long long)blockIdx.x * a.waves + wave) * CPW + q;
      if (cq < (long long)a.num_cw) s += it;
    }
    if (lane == 0 && s) atomicAdd(a.tot_iter, s);
  }
}

// ---------------------------------------------------------------------------
// shape choice and dispatch
// ---------------------------------------------------------------------------

static int maxw_bucket(int w) { return w <= 8 ? 8 : w <= 16 ? 16 : w <= 24 ? 24 : w <= 32 ? 32 : 0; }

int choose_shape(const HdecHost& h, int elem_bytes, DecShape* s, std::string* why) {
  const int Z = h.z;
  s->maxw = maxw_bucket(h.max_row_weight);
  if (!s->maxw) { *why = "row weight above 32"; return FECSIM_E_UNSUPPORTED; }
  const int N = h.mcol * Z, E = h.total_hc_ones * Z;
  if (N > 65535) { *why = "mcol*z above 65535"; return FECSIM_E_UNSUPPORTED; }
  // rows per lane: best lane utilization Z*cpw/(64*K); ties -> fewer rows per lane
  double best = -1.0;
  for (int K = 1; K <= 16; ++K) {
    const int L = (Z + K - 1) / K;
    if (L > 64) continue;
    const int cpw = 64 / L;
    const double util = (double)Z * cpw / (64.0 * K);
    if (util > best + 1e-9) { best = util; s->rows_per_lane = K; s->lanes_per_cw = L; s->cw_per_wave = cpw; }
  }
  if (best < 0) { *why = "z above 1024"; return FECSIM_E_UNSUPPORTED; }
  const int pad = 4 / (elem_bytes < 4 ? elem_bytes : 4);
  s->post_stride = (N + pad - 1) / pad * pad;
  s->c2v_stride = (E + pad - 1) / pad * pad;
  const int lds_cap = 64 * 1024;
  for (;;) {
    s->wave_bytes = s->cw_per_wave * (s->post_stride + s->c2v_stride) * elem_bytes;
    if (s->wave_bytes <= lds_cap || s->cw_per_wave == 1) break;
    --s->cw_per_wave;
  }
  if (s->wave_bytes > lds_cap) { *why = "codeword does not fit one wave's LDS slice"; return FECSIM_E_UNSUPPORTED; }
  s->waves = lds_cap / s->wave_bytes;
  if (s->waves > 4) s->waves = 4;
  return FECSIM_OK;
}

template <class P>
static hipError_t launch_p(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  const long long per_block = (long long)a.waves * a.cw_per_wave;
  const unsigned blocks = (unsigned)((a.num_cw + per_block - 1) / per_block);
  if (blocks == 0) return hipSuccess;
  const dim3 grid(blocks), block(64 * a.waves);
  const size_t lds = (size_t)a.waves * a.wave_bytes;
  switch (d.shape.maxw) {
  case 8: hipLaunchKernelGGL((k_decode<P, 8>), grid, block, lds, st, a); break;
  case 16: hipLaunchKernelGGL((k_decode<P, 16>), grid, block, lds, st, a); break;
  case 24: hipLaunchKernelGGL((k_decode<P, 24>), grid, block, lds, st, a); break;
  default: hipLaunchKernelGGL((k_decode<P, 32>), grid, block, lds, st, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_decode(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  switch (d.elem) {
  case FECSIM_I8: return launch_p<IntPolicy<8>>(d, a, st);
  case FECSIM_I16: return launch_p<IntPolicy<16>>(d, a, st);
  case FECSIM_I32: return launch_p<IntPolicy<32>>(d, a, st);
  case FECSIM_F32:
    return d.mode == FECSIM_PROD ? launch_p<FloatPolicy<float, true>>(d, a, st)
                                 : launch_p<FloatPolicy<float, false>>(d, a, st);
  default:
    return d.mode == FECSIM_PROD ? launch_p<FloatPolicy<double, true>>(d, a, st)
                                 : launch_p<FloatPolicy<double, false>>(d, a, st);
  }
}

}  // namespace fecsim

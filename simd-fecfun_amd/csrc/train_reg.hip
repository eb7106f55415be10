// Differentiable layered offset / normalized min-sum decoder (the forward and adjoint of
// train.hip) with the C2V messages in registers, for compiled code shapes.
//
// No counterpart in the reference; the forward row is the reference's float min-sum row
// (src/LDPCDec/decode_functions.cpp:11-103) with per-(iteration, layer) offset and scale,
// as in train.hip.  train.hip keeps a codeword's C2V messages (and their adjoints) in a
// global array, so every layer of every iteration waits for them on each wave's serial
// chain; here they stay in VGPRs:
//
// * One wave per codeword (block of one wave).  Lane l owns rows l and l + 64 (RPL = 1 or
//   2 rows per lane, Z <= 64 RPL) of every layer and keeps their C2V values, one per edge
//   of the shape (`ne`), in registers: rpl * ne of them (172 for 802.11n 1/2 z81).  The
//   posteriors (and in the adjoint their adjoints) sit in LDS.
// * The layer loop is unrolled at compile time over the shape's row weights, so every
//   register index is static.
// * The adjoint uses the same registers for the C2V adjoints (negated) during its reverse sweeps,
//   and parks them in the global gc array while it runs a segment's forward again from
//   its checkpoint (the recomputation needs the C2V messages themselves).
// * Instead of a row's W V2C values the forward keeps, for the iterations the adjoint
//   sweeps, the row's reverse state: its two smallest magnitudes m1, m2, their edges, and
//   per edge the V2C sign, whether the offset left a positive magnitude and (int8
//   emulation) whether the V2C value and the posterior saturated -- 3 words a row (4 with
//   saturation) in place of W.  The reverse sweep reads these words, one layer ahead of
//   the layer it processes, instead of re-deriving the minima from the V2C values.
// * Per (iteration, layer) the parameter gradients of the codeword's rows are summed
//   across the wave (fixed order: deterministic) and written per codeword as in
//   train.hip, whose k_train_sum then sums over the batch.
//
// Every row runs the float operations of train.hip's fwd_row / rev_row in the same
// order (the reverse state holds exactly the minima, indices and predicates rev_row
// derives from the kept V2C values), so forward posteriors and gradients equal
// train.hip's kernels bit for bit.  The checkpoint and gradient buffers keep train.hip's
// layouts (edge e, row r at e * Z + r); the reverse state of iteration i of a segment,
// layer l, word k, row r sits at ((i * mrow + l) * words + k) * Z + r of the codeword's
// slot of the kept-state buffer (K * mrow * words * Z values a codeword).
#include <hip/hip_runtime.h>

#include "decode_core.h"
#include "internal.h"
#include "train_core.h"
#include "train_shapes.h"

namespace fecsim {

namespace {

using namespace train_core;

// first edge of layer L of shape SH (compile time)
template <class SH>
constexpr int layer_start(int L) {
  int e = 0;
  for (int k = 0; k < L; ++k) e += SH::w[k];
  return e;
}

// words of a row's reverse state: m1, m2, packed bits (+ saturation bits in int8 emulation)
template <bool QT>
constexpr int kStateWords = QT ? 4 : 3;

// packed bits: [0, W) the C2V sign of each edge (edge j at bit W - 1 - j: the row's sign
// parity and the V2C sign), [12] the V2C sign of the m1 edge, [13] of the m2 edge,
// [16, 20) index of m1, [20, 24) index of m2 (15: none), [24, 32) V2C within int8 (int8
// emulation, edge j at bit 24 + j); word 3: posterior saturated (edge j at bit j).  Whether
// the two minima's offsets were active (|q| - beta > 0) is m1 > 0 and m2 > 0.
template <int W>
__device__ __forceinline__ uint32_t pack_state(uint32_t qm, uint32_t sgb, int im, int im2, uint32_t inm) {
  const uint32_t fl = (qm ^ (uint32_t)((int32_t)sgb >> 31)) & ((1u << W) - 1u);   // qm: V2C signs, sgb: parity
  const uint32_t q1 = (qm >> (W - 1 - im)) & 1u, q2 = (qm >> (W - 1 - im2)) & 1u;   // im = -1: shift by W, 0
  return fl | (q1 << 12) | (q2 << 13) | ((uint32_t)(im & 15) << 16) | ((uint32_t)(im2 & 15) << 20) | (inm << 24);
}

constexpr uint32_t kSign = 0x80000000u;

// LDS edge table of the register kernels, built at kernel start from the decoder's edge
// words
template <class SH>
constexpr uint32_t reg_table_bytes() {
  return (2u * SH::ne * 4u + 15u) / 16u * 16u;
}

typedef __attribute__((address_space(3))) float lds_float;

// LDS address (byte offset in the LDS space) of the block's dynamic LDS
__device__ __forceinline__ uint32_t lds_address(char* smem) {
  return (uint32_t)(size_t)(__attribute__((address_space(3))) char*)smem;
}
__device__ __forceinline__ float lds_ld(uint32_t ad) { return *(const lds_float*)(size_t)ad; }
__device__ __forceinline__ void lds_st(uint32_t ad, float v) { *(lds_float*)(size_t)ad = v; }

// word per edge: 4 * shift | (LDS address of the column's first value) << 16, for the
// posteriors and for their adjoints (LDS addresses below 64 KiB)
template <class SH>
__device__ __forceinline__ void reg_tables(uint32_t* t, const uint32_t* __restrict__ edges, uint32_t p_ad,
                                           uint32_t g_ad) {
  for (int i = threadIdx.x; i < SH::ne; i += blockDim.x) {
    const uint32_t ed = edges[i];
    t[i] = (ed & 0xffffu) * 4u | ((p_ad + (ed >> 16) * 4u) << 16);
    t[SH::ne + i] = (ed & 0xffffu) * 4u | ((g_ad + (ed >> 16) * 4u) << 16);
  }
}

// LDS address of row r's variable on an edge of table word w: 4 ((r + shift) mod Z) as
// min(4 (r + shift), 4 (r + shift - Z)) in unsigned arithmetic (the second wraps above
// the first when r + shift < Z) plus the column's address; r4 = 4 r, rz4 = 4 (r - Z)
__device__ __forceinline__ uint32_t row_addr(uint32_t w, uint32_t r4, uint32_t rz4) {
  const uint32_t t = r4 + (w & 0xffffu), u = rz4 + (w & 0xffffu);
  return (t < u ? t : u) + (w >> 16);
}

// one check row of the forward with its W C2V values in registers c[0..W-1] and its edges'
// table words in tw: C_old -> C_new, P updated in place.  The values equal train.hip's
// fwd_row bit for bit: the running second minimum is a median of three (m1 <= m2: it picks
// the input fwd_row's compare-and-swap keeps), the capped and scaled magnitude is formed
// once per row for m1 and m2 (the same product of the same factors), and the C2V sign is
// an exclusive or of sign bits (IEEE negation).  KEEP also writes the row's reverse state
// to St[k * Z + r].
template <int W, bool QT, bool KEEP>
__device__ __forceinline__ void fwd_row_reg(const uint32_t (&tw)[W], float* c, float* St, int r, uint32_t r4,
                                            uint32_t rz4, int Z, float al, float be, float cap) {
  static_assert(W <= 8, "reverse state packs 8 edges a row");
  uint32_t ad[W];
  float q[W], pv[W];
#pragma unroll
  for (int j = 0; j < W; ++j) ad[j] = row_addr(tw[j], r4, rz4);
#pragma unroll
  for (int j = 0; j < W; ++j) pv[j] = lds_ld(ad[j]);
  float m1 = INFINITY, m2 = INFINITY;
  int im = -1, im2 = -1;
  uint32_t sgb = 0, qm = 0, inm = 0, ym = 0;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const float xq = pv[j] - c[j];
    q[j] = QT ? sat8f(xq) : xq;
    const float A = mag_a<float, QT>(q[j], be);
    sgb ^= __float_as_uint(q[j]);
    const bool lt1 = A < m1;
    if constexpr (KEEP) {
      im2 = lt1 ? im : (A < m2 ? j : im2);
      qm = __builtin_amdgcn_alignbit(qm, __float_as_uint(q[j]), 31);   // (qm << 1) | sign of q
      if constexpr (QT) inm |= (uint32_t)(q[j] == xq) << j;   // xq within [-128, 127]
    }
    im = lt1 ? j : im;
    m2 = __builtin_amdgcn_fmed3f(m1, m2, A);
    m1 = lt1 ? A : m1;
  }
  const uint32_t S = sgb & kSign;
  uint32_t s1 = __float_as_uint(al * (cap < m1 ? cap : m1)) ^ S;
  uint32_t s2 = __float_as_uint(al * (cap < m2 ? cap : m2)) ^ S;
  asm volatile("" : "+v"(s1), "+v"(s2));   // formed once per row, not per edge
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const float rn = __uint_as_float(((j == im) ? s2 : s1) ^ (__float_as_uint(q[j]) & kSign));
    const float y = q[j] + rn;
    if constexpr (QT) {
      lds_st(ad[j], sat8f(y));
      if constexpr (KEEP) ym |= (uint32_t)(y < -128.f || y > 127.f) << j;
    } else {
      lds_st(ad[j], y);
    }
    c[j] = rn;
  }
  if constexpr (KEEP) {
    St[r] = m1;
    St[Z + r] = m2;
    reinterpret_cast<uint32_t*>(St)[2 * Z + r] = pack_state<W>(qm, sgb, im, im2, inm);
    if constexpr (QT) reinterpret_cast<uint32_t*>(St)[3 * Z + r] = ym;
  }
}

// one check row of the reverse sweep with its W C2V adjoints in registers, held negated
// (ng[j] = -gC_j, so gv + gC_j is gv - ng[j]), its edges' table words (adjoint columns)
// in tw and its reverse state in st: train.hip's rev_row operation for operation, its
// conditionals as selects and its multiplications by +-1 as sign changes; sets the row's
// parameter-gradient terms *ga, *gb
template <int W, bool QT>
__device__ __forceinline__ void rev_row_reg(const uint32_t (&tw)[W], float* ng, const uint32_t (&st)[kStateWords<QT>],
                                            uint32_t r4, uint32_t rz4, float al, float cap, float* ga, float* gb) {
  uint32_t ad[W];
  float gq[W];
#pragma unroll
  for (int j = 0; j < W; ++j) ad[j] = row_addr(tw[j], r4, rz4);
#pragma unroll
  for (int j = 0; j < W; ++j) gq[j] = lds_ld(ad[j]);
  const float m1 = __uint_as_float(st[0]), m2 = __uint_as_float(st[1]);
  const uint32_t b = st[2];
  const int im = (int)((b >> 16) & 15u), im2 = (int)((b >> 20) & 15u);
  const bool cp1 = cap < m1, cp2 = cap < m2;
  const float c1 = cp1 ? cap : m1, c2 = cp2 ? cap : m2;
  float S = 0.f, gmi = 0.f;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const bool isim = j == im;
    if constexpr (QT) gq[j] = ((st[3] >> j) & 1u) ? 0.f : gq[j];
    const float gm = __uint_as_float(__float_as_uint(gq[j] - ng[j]) ^ ((b << (32 - W + j)) & kSign));
    const float Sn = S + gm;
    S = isim ? S : Sn;
    gmi = isim ? gm : gmi;
  }
  *ga = c1 * S + c2 * gmi;
  const float g_im = cp1 ? 0.f : al * S, g_im2 = cp2 ? 0.f : al * gmi;
  const bool act1 = m1 > 0.f, act2 = m2 > 0.f;
  // the terms the two minima's edges take back (-0 adds nothing)
  const float T1 = act1 ? __uint_as_float(__float_as_uint(g_im) ^ ((b << 19) & kSign)) : -0.f;
  const float T2 = act2 ? __uint_as_float(__float_as_uint(g_im2) ^ ((b << 18) & kSign)) : -0.f;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    float g = gq[j] + ((j == im) ? T1 : ((j == im2) ? T2 : -0.f));   // + -0 leaves gq as it is
    if constexpr (QT) g = ((b >> (24 + j)) & 1u) ? g : 0.f;
    lds_st(ad[j], g);
    ng[j] = g;
  }
  *gb = -((act1 ? g_im : 0.f) + (act2 ? g_im2 : 0.f));
}

// the block's threads (NWV waves, one codeword) at a layer boundary
template <int NWV>
__device__ __forceinline__ void block_sync() {
  if constexpr (NWV == 1) wave_sync();
  else __syncthreads();
}

// one forward iteration t, layers L.. of shape SH (compile-time recursion: static register
// indices) for the thread's row r; KEEP writes the iteration's reverse state at St
template <class SH, int L, int NWV, bool QT, bool KEEP>
__device__ __forceinline__ void fwd_layers(const TrainArgs& a, const uint32_t* tab, float (&c)[SH::ne], float* St, int t,
                                           int row, float cap) {
  if constexpr (L < SH::mrow) {
    constexpr int W = SH::w[L], E0 = layer_start<SH>(L);
    const float al = reinterpret_cast<const float*>(a.alpha)[t * SH::mrow + L];
    const float be = offset_of<float, QT>(reinterpret_cast<const float*>(a.beta)[t * SH::mrow + L]);
    float* Sl = KEEP ? St + (size_t)L * kStateWords<QT> * a.z : nullptr;
    uint32_t tw[W];
#pragma unroll
    for (int j = 0; j < W; ++j) tw[j] = tab[E0 + j];
    int r = row;
    uint32_t r4 = 4u * row, rz4 = 4u * (row - a.z);
    asm volatile("" : "+v"(r), "+v"(r4), "+v"(rz4));   // edge offsets formed here, not kept across iterations
    if (r < a.z) fwd_row_reg<W, QT, KEEP>(tw, &c[E0], Sl, r, r4, rz4, a.z, al, be, cap);
    block_sync<NWV>();
    fwd_layers<SH, L + 1, NWV, QT, KEEP>(a, tab, c, St, t, row, cap);
  }
}

// the reverse state words of layer base S for row r (0 past Z)
template <bool QT>
__device__ __forceinline__ void load_state(uint32_t (&w)[kStateWords<QT>], const float* S, int row, int Z) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(S);
  int r = row;
  asm volatile("" : "+v"(r));
#pragma unroll
  for (int x = 0; x < kStateWords<QT>; ++x) w[x] = r < Z ? s[x * Z + r] : 0u;
}

// one reverse iteration t, layers L down to 0, for the thread's row, with layer L's
// reverse state in cur; each layer first issues the loads of the next layer's state (at
// layer 0: layer mrow - 1 of iteration t - 1 at Snext, if not null), which it leaves in
// cur.  Writes the codeword's parameter gradients: per (iteration, layer) the sum over
// lanes l of (row l's sum [+ row l + 64's sum]), each row's terms summed from 0 in edge
// order -- train.hip's order.  With two waves each sums one of the two, the other handing
// it its rows' terms through xch: two slots, alternating over the layers in the order
// they are swept, so a wave's next write never meets the other wave's read of the last.
template <class SH, int L, int NWV, bool QT>
__device__ __forceinline__ void rev_layers(const TrainArgs& a, const uint32_t* tab, float (&g)[SH::ne],
                                           const float* St, const float* Snext, uint32_t (&cur)[kStateWords<QT>],
                                           float* part, float* xch, int t, int row, float cap) {
  if constexpr (L >= 0) {
    constexpr int W = SH::w[L], E0 = layer_start<SH>(L), NW = kStateWords<QT>;
    uint32_t nxt[NW];
    if constexpr (L > 0) load_state<QT>(nxt, St + (size_t)(L - 1) * NW * a.z, row, a.z);
    else if (Snext) load_state<QT>(nxt, Snext + (size_t)(SH::mrow - 1) * NW * a.z, row, a.z);
    const float al = reinterpret_cast<const float*>(a.alpha)[t * SH::mrow + L];
    uint32_t tw[W];
#pragma unroll
    for (int j = 0; j < W; ++j) tw[j] = tab[SH::ne + E0 + j];
    float ga = 0.f, gb = 0.f;
    int r = row;
    uint32_t r4 = 4u * row, rz4 = 4u * (row - a.z);
    asm volatile("" : "+v"(r), "+v"(r4), "+v"(rz4));
    if (r < a.z) rev_row_reg<W, QT>(tw, &g[E0], cur, r4, rz4, al, cap, &ga, &gb);
    if constexpr (NWV == 2) {
      // the first wave sums the scale's terms, the second the offset's: each hands the
      // other its rows' terms of the other sum
      float* x = xch + ((t * SH::mrow + L) & 1) * 128;
      if (row < 64) x[64 + row] = gb;
      else x[row - 64] = ga;
      block_sync<NWV>();
      if (row < 64) {
        if (row + 64 < a.z) ga = ga + x[row];
        ga = wave_sum(ga);
        if (row == 0) part[(t * SH::mrow + L) * 2 + 0] = ga;
      } else {
        gb = row < a.z ? x[row] + gb : x[row];
        gb = wave_sum(gb);
        if (row == 64) part[(t * SH::mrow + L) * 2 + 1] = gb;
      }
    } else {
      block_sync<NWV>();
      ga = wave_sum(ga);
      gb = wave_sum(gb);
      if (row == 0) {
        part[(t * SH::mrow + L) * 2 + 0] = ga;
        part[(t * SH::mrow + L) * 2 + 1] = gb;
      }
    }
    if (L > 0 || Snext) {
#pragma unroll
      for (int x = 0; x < NW; ++x) cur[x] = nxt[x];
    }
    rev_layers<SH, L - 1, NWV, QT>(a, tab, g, St, Snext, cur, part, xch, t, row, cap);
  }
}

// the thread's row of a global [E] array laid out e * Z + r <-> registers
template <class SH>
__device__ __forceinline__ void regs_to_global(float* dst, const float (&c)[SH::ne], int row, int Z) {
  int r = row;
  asm volatile("" : "+v"(r));
  if (r < Z) {
#pragma unroll
    for (int e = 0; e < SH::ne; ++e) dst[e * Z + r] = c[e];
  }
}
template <class SH>
__device__ __forceinline__ void global_to_regs(float (&c)[SH::ne], const float* src, int row, int Z) {
  int r = row;
  asm volatile("" : "+v"(r));
#pragma unroll
  for (int e = 0; e < SH::ne; ++e) c[e] = r < Z ? src[e * Z + r] : 0.f;
}

// bytes of the gradient hand-over between the two waves of a codeword
template <int NWV>
constexpr uint32_t xch_bytes() {
  return NWV == 2 ? 2u * 128u * 4u : 0u;
}

template <class SH, int NWV, bool QT>
__global__ void __launch_bounds__(64 * NWV) k_train_fwd_reg(const TrainArgs a) {
  HIP_DYNAMIC_SHARED(char, smem);
  const int row = threadIdx.x;
  constexpr uint32_t p_off = reg_table_bytes<SH>();
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem);
  reg_tables<SH>(tab, a.edges, lds_address(smem) + p_off, lds_address(smem) + p_off);
  const uint64_t cw = blockIdx.x;
  float* P = reinterpret_cast<float*>(smem + p_off);
  const int N = a.n, E = a.e, Z = a.z, T = a.iters;
  const float* llr = reinterpret_cast<const float*>(a.llr) + cw * (uint64_t)N;
  const int K = a.seg, NS = (T + K - 1) / K;
  float* ckP = reinterpret_cast<float*>(a.ckP) + cw * (uint64_t)NS * N;
  float* ckC = reinterpret_cast<float*>(a.ckC) + cw * (uint64_t)NS * E;
  const float cap = (float)a.cap;
  for (int i = row; i < N; i += 64 * NWV) P[i] = llr[i];
  float c[SH::ne];
#pragma unroll
  for (int e = 0; e < SH::ne; ++e) c[e] = 0.f;
  block_sync<NWV>();
  const int tlast = seg_start(NS - 1, T, K, NS);   // the adjoint starts with this segment: its state is kept now
  const int f = seg_start(1, T, K, NS);           // first segment's length (T when NS = 1)
  constexpr int SW = SH::mrow * kStateWords<QT>;  // reverse-state words per row and iteration
  float* Sseg = reinterpret_cast<float*>(a.qs) + cw * (uint64_t)K * SW * Z;
  for (int t = 0; t < T; ++t) {
    const int s = t == 0 ? 0 : (t >= f && (t - f) % K == 0 ? (t - f) / K + 1 : -1);
    if (s >= 0) {   // checkpoint at the start of segment s
      float* cp = ckP + (uint64_t)s * N;
      for (int i = row; i < N; i += 64 * NWV) cp[i] = P[i];
      regs_to_global<SH>(ckC + (uint64_t)s * E, c, row, Z);
    }
    if (t == tlast) break;
    fwd_layers<SH, 0, NWV, QT, false>(a, tab, c, nullptr, t, row, cap);
  }
  for (int t = tlast; t < T; ++t)
    fwd_layers<SH, 0, NWV, QT, true>(a, tab, c, Sseg + (uint64_t)(t - tlast) * SW * Z, t, row, cap);
  float* out = reinterpret_cast<float*>(a.post) + cw * (uint64_t)N;
  for (int i = row; i < N; i += 64 * NWV) out[i] = P[i];
}

template <class SH, int NWV, bool QT>
__global__ void __launch_bounds__(64 * NWV) k_train_bwd_reg(const TrainArgs a) {
  HIP_DYNAMIC_SHARED(char, smem);
  const int row = threadIdx.x;
  constexpr uint32_t p_off = reg_table_bytes<SH>();
  const uint32_t g_off = p_off + (uint32_t)a.n_pad * 4u;
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem);
  reg_tables<SH>(tab, a.edges, lds_address(smem) + p_off, lds_address(smem) + g_off);
  const uint64_t cw = blockIdx.x;
  float* P = reinterpret_cast<float*>(smem + p_off);
  float* gP = reinterpret_cast<float*>(smem + g_off);
  float* xch = reinterpret_cast<float*>(smem + g_off + (uint32_t)a.n_pad * 4u);
  const int N = a.n, E = a.e, Z = a.z, T = a.iters;
  const int K = a.seg, NS = (T + K - 1) / K;
  const float* ckP = reinterpret_cast<const float*>(a.ckP) + cw * (uint64_t)NS * N;
  const float* ckC = reinterpret_cast<const float*>(a.ckC) + cw * (uint64_t)NS * E;
  constexpr int SW = SH::mrow * kStateWords<QT>;
  float* Sseg = reinterpret_cast<float*>(a.qs) + cw * (uint64_t)K * SW * Z;
  float* gcp = reinterpret_cast<float*>(a.gc) + cw * (uint64_t)E;   // the negated C2V adjoints while a segment is recomputed
  float* part = reinterpret_cast<float*>(a.partial) + cw * (uint64_t)T * SH::mrow * 2;
  const float cap = (float)a.cap;
  const float* gpost = reinterpret_cast<const float*>(a.grad_post) + cw * (uint64_t)N;
  for (int i = row; i < N; i += 64 * NWV) gP[i] = gpost[i];
  // the C2V adjoints, negated (-0 for an adjoint of +0); the C2V messages while a segment
  // is recomputed
  float g[SH::ne];
#pragma unroll
  for (int e = 0; e < SH::ne; ++e) g[e] = -0.f;
  block_sync<NWV>();
  for (int sg = NS - 1; sg >= 0; --sg) {
    const int t0 = seg_start(sg, T, K, NS), t1 = sg == NS - 1 ? T : seg_start(sg + 1, T, K, NS);
    if (sg < NS - 1) {
      // the segment's forward again from its checkpoint, keeping every iteration's reverse
      // state (the last segment's was kept by the forward)
      regs_to_global<SH>(gcp, g, row, Z);
      const float* cp = ckP + (uint64_t)sg * N;
      for (int i = row; i < N; i += 64 * NWV) P[i] = cp[i];
      global_to_regs<SH>(g, ckC + (uint64_t)sg * E, row, Z);
      block_sync<NWV>();
      for (int t = t0; t < t1; ++t)
        fwd_layers<SH, 0, NWV, QT, true>(a, tab, g, Sseg + (uint64_t)(t - t0) * SW * Z, t, row, cap);
      global_to_regs<SH>(g, gcp, row, Z);
    }
    uint32_t cur[kStateWords<QT>];
    load_state<QT>(cur, Sseg + ((uint64_t)(t1 - 1 - t0) * SW + (SH::mrow - 1) * kStateWords<QT>) * Z, row, Z);
    for (int t = t1 - 1; t >= t0; --t) {
      const float* St = Sseg + (uint64_t)(t - t0) * SW * Z;
      rev_layers<SH, SH::mrow - 1, NWV, QT>(a, tab, g, St, t > t0 ? St - (uint64_t)SW * Z : nullptr, cur, part,
                                            xch, t, row, cap);
    }
  }
  block_sync<NWV>();
  if (a.grad_llr) {
    float* gl = reinterpret_cast<float*>(a.grad_llr) + cw * (uint64_t)N;
    for (int i = row; i < N; i += 64 * NWV) gl[i] = gP[i];
  }
}

template <class SH, int NWV>
hipError_t launch_reg(const TrainArgs& a, int bwd, hipStream_t st) {
  const size_t lds = reg_table_bytes<SH>() + (size_t)(bwd ? 2 : 1) * a.n_pad * sizeof(float) + (bwd ? xch_bytes<NWV>() : 0);
  const dim3 grid((unsigned)a.batch), block(64 * NWV);
  if (a.quant) {
    if (bwd) hipLaunchKernelGGL((k_train_bwd_reg<SH, NWV, true>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((k_train_fwd_reg<SH, NWV, true>), grid, block, lds, st, a);
  } else {
    if (bwd) hipLaunchKernelGGL((k_train_bwd_reg<SH, NWV, false>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((k_train_fwd_reg<SH, NWV, false>), grid, block, lds, st, a);
  }
  return hipGetLastError();
}

// one wave per codeword up to Z = 64, two from there to 128 (one row per thread)
template <class SH>
hipError_t launch_shape(const TrainArgs& a, int bwd, hipStream_t st) {
  if (a.z <= 64) return launch_reg<SH, 1>(a, bwd, st);
  if (a.z <= 128) return launch_reg<SH, 2>(a, bwd, st);
  return hipErrorInvalidValue;
}

}  // namespace

int train_reg_shape(const HdecHost& h, int is_double) {
  if (is_double || h.z > 128 || h.z < 1 || h.max_row_weight > 8) return -1;
  // the edge table's 16-bit column offsets reach past the posteriors' adjoints
  if (4096 + 8 * (size_t)h.mcol * h.z >= 65536) return -1;
  return train_shape_index(h.mrow, h.row_weights.data());
}

#define FECSIM_TRAIN_LAUNCH(SH) return launch_shape<SH>(a, bwd, st);
#include "train_shapes_dispatch.h"

int train_reg_state_words(int quant) { return quant ? kStateWords<true> : kStateWords<false>; }

hipError_t launch_train_reg(int shape, const TrainArgs& a, int bwd, hipStream_t st) {
  if (a.batch == 0) return hipSuccess;
  return train_reg_dispatch(shape, a, bwd, st);
}

}  // namespace fecsim

// Row arithmetic of the code-specialized int8 decoders (code_i8.hip: one codeword per
// workgroup; multi_code.hip: a mixed-code queue, several codewords per workgroup).  The
// code -- lifting size, layer weights, columns and shifts -- is a compile-time table
// (code_tables.h), so every LDS address is a base register plus an immediate offset and a
// row's C2V messages stay in registers.  Same results as the generic kernel
// (decode_core.h row_int, after src/LDPCDec/decode_functions.cpp:11-103), bit for bit.
//
// LDS layout of one codeword: posteriors as bytes biased by 128 (b = v + 128, 0..255).
// Column c keeps two copies, at Z + 3Zc and Z + 3Zc + Z, with a Z-byte gap after each
// column (and one before the first): (3 mcol + 1) Z bytes.  Row r of edge (c, s) reads its
// variable (s + r) mod Z at Z + 3Zc + s + r without a wrap-around test, and writes the new
// value there and Z below and above: two of the three are the variable's copies, the third
// lands in a gap.
//
// Per edge: q = sat8(L - R) is med3(b - R, 0, 255) on the biased value; |q| = |b - 128| is
// one v_sad_u8; min1/min2 take a v_med3_u32 and a v_min_u32; the sign product is the XOR of
// the biased bytes (bit 7 set = non-negative).  With beta >= 1 the exclusive minimum follows
// from min1/min2 of |q| without an argmin (see code_layer).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fecsim {
namespace {   // internal linkage: the compile-time plans fold into each kernel

typedef __attribute__((address_space(3))) uint8_t LdsU8;
__device__ __forceinline__ LdsU8* lds8(uint32_t byte_addr) { return (LdsU8*)(size_t)byte_addr; }

__device__ __forceinline__ uint32_t med3u(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t d;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
__device__ __forceinline__ int med3i(int a, int b, int c) {
  int d;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
__device__ __forceinline__ int clamp_byte(int x) { return x < 0 ? 0 : (x > 255 ? 255 : x); }

// byte e & 3 of the packed C2V word of edge e, sign-extended
__device__ __forceinline__ int c2v_byte(uint32_t word, int b) { return __builtin_amdgcn_sbfe(word, 8 * b, 8); }

// v_perm selector that replaces byte b of the second operand by byte 0 of the first
constexpr uint32_t insert_sel(int b) {
  return b == 0 ? 0x03020104u : b == 1 ? 0x03020400u : b == 2 ? 0x03040100u : 0x04020100u;
}

template <class CT>
constexpr int edge_off(int e) {   // LDS byte offset (from the row's base) of edge e's read
  return CT::z + 3 * CT::z * CT::col[e] + CT::sh[e];
}

// LDS byte at offset K from the thread's row base; one base register per 32 KiB keeps
// every offset in the instruction's immediate field
template <int NB>
__device__ __forceinline__ LdsU8* lds_at(const uint32_t (&vb)[NB], int K) {
  return lds8(vb[K >> 15] + (uint32_t)(K & 32767));
}

// Where the workgroup barriers go within an iteration.  Every layer writes all the
// columns it reads, so layer l needs a barrier before it only when it reads a column
// written since the last barrier (consecutive orthogonal layers share none: base graph
// 1's extension rows pair up this way).  Before such a barrier a thread can already
// issue the reads of the columns not written since the last one.
// The reads of layer l + 1 that need no barrier after layer l (all of them when there is
// none, else those of columns not written since the last one) are issued earlier still,
// at the start of layer l, so that they overlap its arithmetic; only where the two
// layers' weights leave the registers for it.
template <class CT>
struct CodePlan {
  bool bar[CT::mrow];   // barrier before layer l
  bool early[CT::ne];   // edge e read before that barrier
  bool pre[CT::ne];     // edge e (of layer l >= 1) read at the start of layer l - 1
};

constexpr int kPreReadMaxW = 24;   // largest w[l-1] + w[l] with layer l's reads issued during layer l - 1
constexpr int kKeepAbsMaxW = 12;   // largest row weight whose |q| values stay in registers

template <class CT>
constexpr CodePlan<CT> code_plan() {
  static_assert(CT::mcol <= 128, "column sets are two 64-bit words");
  CodePlan<CT> p{};
  uint64_t dirty[2] = {0, 0};   // columns written since the last barrier
  int e0 = 0;
  for (int l = 0; l < CT::mrow; ++l) {
    uint64_t cols[2] = {0, 0};
    for (int j = 0; j < CT::w[l]; ++j) {
      const int c = CT::col[e0 + j];
      cols[c >> 6] |= 1ull << (c & 63);
    }
    p.bar[l] = ((cols[0] & dirty[0]) | (cols[1] & dirty[1])) != 0;
    for (int j = 0; j < CT::w[l]; ++j) {
      const int c = CT::col[e0 + j];
      p.early[e0 + j] = p.bar[l] && ((dirty[c >> 6] >> (c & 63)) & 1) == 0;
    }
    for (int j = 0; j < CT::w[l]; ++j)
      p.pre[e0 + j] = l > 0 && CT::w[l - 1] + CT::w[l] <= kPreReadMaxW && (!p.bar[l] || p.early[e0 + j]);
    if (p.bar[l]) dirty[0] = dirty[1] = 0;
    dirty[0] |= cols[0];
    dirty[1] |= cols[1];
    e0 += CT::w[l];
  }
  return p;
}

template <class CT, int L>
constexpr int wnext() {
  return L + 1 < CT::mrow ? CT::w[L + 1] : 1;
}

template <class CT>
constexpr CodePlan<CT> kCodePlan = code_plan<CT>();

// One layer of row r (edges E0 .. E0 + w - 1, compiled-in weight, columns and shifts),
// with the workgroup barrier in front of it when the plan has one.  A thread with
// act == false (no codeword in its slot) reaches the barrier and does nothing else.
template <class CT, int L, int E0, int NB, int NW>
__device__ __forceinline__ void code_layer(const uint32_t (&vb)[NB], uint32_t (&c2v)[NW], uint32_t be, int cap,
                                           const int (&pin)[CT::w[L]], int (&pout)[wnext<CT, L>()], bool act) {
  constexpr int Z = CT::z, W = CT::w[L];
  constexpr const CodePlan<CT>& P = kCodePlan<CT>;
  static_assert(W >= 2, "row weight below 2");
  int pb[W];
#pragma unroll
  for (int j = 0; j < W; ++j)
    if (P.pre[E0 + j]) pb[j] = pin[j];   // read during the previous layer
  if constexpr (P.bar[L]) {
    if (act) {
#pragma unroll
      for (int j = 0; j < W; ++j)
        if (P.early[E0 + j] && !P.pre[E0 + j]) pb[j] = *lds_at(vb, edge_off<CT>(E0 + j));
    }
    __syncthreads();   // reached by every thread of the workgroup, with or without a row
    if (!act) return;
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (!P.early[E0 + j]) pb[j] = *lds_at(vb, edge_off<CT>(E0 + j));
  } else {
    if (!act) return;
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (!P.pre[E0 + j]) pb[j] = *lds_at(vb, edge_off<CT>(E0 + j));
  }
  if constexpr (L + 1 < CT::mrow) {   // the next layer's reads that may go now
#pragma unroll
    for (int j = 0; j < CT::w[L + 1]; ++j)
      if (P.pre[E0 + W + j]) pout[j] = *lds_at(vb, edge_off<CT>(E0 + W + j));
  }
  // V2C: q_j = sat8(L_j - R_j) (biased), |q_j|, min1/min2 of |q|, XOR of the bytes
  int qb[W];
  // |q_j| is kept for the output pass in light layers and recomputed (one v_sad_u8) in
  // heavy ones, whose registers set the kernel's peak
  constexpr bool keep_u = W <= kKeepAbsMaxW;
  uint32_t u[keep_u ? W : 1];
  uint32_t m1 = 255u, m2 = 255u, X = (W & 1) ? 0u : 0x80u;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const int e = E0 + j;
    const int q = clamp_byte(pb[j] - c2v_byte(c2v[e >> 2], e & 3));
    qb[j] = q;
    const uint32_t uu = __builtin_amdgcn_sad_u8((uint32_t)q, 128u, 0u);
    if constexpr (keep_u) u[j] = uu;
    X ^= (uint32_t)q;
    m2 = med3u(m1, uu, m2);
    m1 = m1 < uu ? m1 : uu;
  }
  // A_j = max(|q_j| - beta, 0) (|q| <= 128, beta in 1..127: no int8 wrap).  With
  // m' = max(m, beta) and mv = min(cap, m' - beta) the capped exclusive minimum is
  //   M_j = (j is the argmin ? mv2 : mv1) = med3(mv1, m1' + m2' - beta - |q_j|, mv2):
  // at the argmin |q_j| = m1 <= m1', so the middle term is >= m2' - beta >= mv2; at any
  // other edge |q_j| >= m2, so it is <= m1' - beta when m2 >= beta, and when m2 < beta
  // mv1 = mv2 = 0.  Ties included.
  const uint32_t m1p = m1 > be ? m1 : be, m2p = m2 > be ? m2 : be;
  const int mv1 = (int)(m1p - be) < cap ? (int)(m1p - be) : cap;
  const int mv2 = (int)(m2p - be) < cap ? (int)(m2p - be) : cap;
  int sumb = (int)(m1p + m2p - be);
  asm volatile("" : "+v"(sumb));   // one subtraction per edge, not re-associated
  // C2V_j is negative when bit 7 of X ^ qb_j is set: X holds the XOR of the biased bytes
  // (bit 7 = non-negative) flipped once more for an even row weight, so this is the
  // parity of the other edges' negative signs.
  int rn[W];
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t uj = keep_u ? u[keep_u ? j : 0] : __builtin_amdgcn_sad_u8((uint32_t)qb[j], 128u, 0u);
    const int m = med3i(mv1, sumb - (int)uj, mv2);
    const int neg = __builtin_amdgcn_sbfe(X ^ (uint32_t)qb[j], 7, 1);   // 0 or -1
    const int rv = (m ^ neg) - neg;
    rn[j] = rv;
    const uint8_t pn = (uint8_t)clamp_byte(qb[j] + rv);   // sat8(q + C2V), biased
    const int K = edge_off<CT>(E0 + j);
    *lds_at(vb, K - Z) = pn;
    *lds_at(vb, K) = pn;
    *lds_at(vb, K + Z) = pn;
    // the new C2V byte into its word at once (few live values): a word whose four edges
    // all belong to this layer is built when its last byte is ready, others byte by byte
    const int e = E0 + j;
    const bool whole = ((e & ~3) >= E0) && ((e | 3) < E0 + W);
    if (whole && (e & 3) == 3) {
      const uint32_t lo = __builtin_amdgcn_perm((uint32_t)rn[j - 2], (uint32_t)rn[j - 3], 0x0c0c0400u);
      const uint32_t hi = __builtin_amdgcn_perm((uint32_t)rv, (uint32_t)rn[j - 1], 0x04000c0cu);
      c2v[e >> 2] = lo | hi;
    } else if (!whole) {
      c2v[e >> 2] = __builtin_amdgcn_perm((uint32_t)rv, c2v[e >> 2], insert_sel(e & 3));
    }
    // opaque: otherwise the compiler sees through the packing to the separate int32
    // values (the next read of the byte is their sign-extended low byte) and keeps all
    // of them live, one register per edge instead of one per four
    if ((e & 3) == 3 || j == W - 1) asm volatile("" : "+v"(c2v[e >> 2]));
  }
}

// layers L .. mrow-1 of one iteration, then the barrier before the syndrome
template <class CT, int L, int E0, int NB, int NW>
__device__ __forceinline__ void code_layers(const uint32_t (&vb)[NB], uint32_t (&c2v)[NW], const long long* bt,
                                            uint32_t beta, int cap, const int (&pin)[CT::w[L]], bool act) {
  const uint32_t be = (act && bt) ? (uint32_t)bt[L] : beta;
  int pout[wnext<CT, L>()];
  code_layer<CT, L, E0>(vb, c2v, be, cap, pin, pout, act);
  __builtin_amdgcn_sched_barrier(0);   // layer by layer: bounded live ranges
  if constexpr (L + 1 < CT::mrow) code_layers<CT, L + 1, E0 + CT::w[L]>(vb, c2v, bt, beta, cap, pout, act);
  else __syncthreads();
}

// OR over layers L .. mrow-1 of the row's parity in bit 7 (set: check not satisfied):
// the XOR of the biased bytes has bit 7 = parity of the non-negative values, flipped for
// an odd row weight into the parity of the negative ones
template <class CT, int L, int E0, int NB>
__device__ __forceinline__ uint32_t code_syndrome(const uint32_t (&vb)[NB]) {
  constexpr int W = CT::w[L];
  uint32_t par = (W & 1) ? 0x80u : 0u;
#pragma unroll
  for (int j = 0; j < W; ++j) par ^= *lds_at(vb, edge_off<CT>(E0 + j));
  if constexpr (L + 1 < CT::mrow) return par | code_syndrome<CT, L + 1, E0 + W>(vb);
  else return par;
}


}  // namespace
}  // namespace fecsim

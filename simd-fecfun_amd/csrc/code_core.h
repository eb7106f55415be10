// Row arithmetic of the code-specialized int8 decoders (code_i8.hip: one codeword per
// workgroup; multi.hip's k_multi_code: a mixed-code queue, several codewords per
// workgroup).  The
// code -- lifting size, layer weights, columns and shifts -- is a compile-time table
// (code_tables.h), so every LDS address is a base register plus an immediate offset and a
// row's C2V messages stay in registers.  Same results as the generic kernel
// (decode_core.h row_int, after src/LDPCDec/decode_functions.cpp:11-103), bit for bit.
//
// LDS layout of one codeword: posteriors as bytes (int8, or biased by 128 in the one-edge
// form of the row arithmetic, see code_layer).  Column c keeps two copies, at
// Z + 3Zc and Z + 3Zc + Z, with a Z-byte gap after each column (and one before the first):
// (3 mcol + 1) Z bytes.  Row r of edge (c, s) reads its variable (s + r) mod Z at
// Z + 3Zc + s + r without a wrap-around test, and writes the new value there and Z below
// and above: two of the three are the variable's copies, the third lands in a gap.
//
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

// v_perm selector that replaces byte b of the second operand by byte 0 of the first
constexpr uint32_t insert_sel(int b) {
  return b == 0 ? 0x03020104u : b == 1 ? 0x03020400u : b == 2 ? 0x03040100u : 0x04020100u;
}

typedef short __attribute__((ext_vector_type(2))) s16x2;
typedef unsigned short __attribute__((ext_vector_type(2))) u16x2;
__device__ __forceinline__ uint32_t as_u(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t as_u(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s16x2 as_s(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ u16x2 as_us(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t swap16(uint32_t v) { return __builtin_amdgcn_alignbit(v, v, 16); }

// byte e & 3 of the packed C2V word of edge e, sign-extended
__device__ __forceinline__ int c2v_byte(uint32_t word, int b) { return __builtin_amdgcn_sbfe(word, 8 * b, 8); }

// v_perm selector that keeps bytes 0..3 of the second operand but byte b, which takes
// byte s (4..7: bytes of the first operand)
constexpr uint32_t put_sel(int b, uint32_t s) { return (0x03020100u & ~(0xffu << (8 * b))) | (s << (8 * b)); }
// ... bytes b and b + 1 (b < 3), which take bytes s0 and s1
constexpr uint32_t put_sel2(int b, uint32_t s0, uint32_t s1) {
  return (0x03020100u & ~(0xffffu << (8 * b))) | (s0 << (8 * b)) | (s1 << (8 * (b + 1)));
}

// The C2V messages of edges e and e + 1 (two) or of edge e alone, as int16 halves v << 8
// (edge e low, e + 1 high; low bytes 0).  e is a compile-time index after unrolling.
template <int NW>
__device__ __forceinline__ uint32_t c2v_pair(const uint32_t (&c2v)[NW], int e, bool two) {
  const int w = e >> 2, b = e & 3;
  if (!two) return __builtin_amdgcn_perm(0u, c2v[w], 0x0c0c000cu | ((uint32_t)b << 8));
  if (b < 3) return __builtin_amdgcn_perm(0u, c2v[w], 0x000c000cu | ((uint32_t)b << 8) | ((uint32_t)(b + 1) << 24));
  return __builtin_amdgcn_perm(c2v[w + 1], c2v[w], 0x040c030cu);   // byte 3 of word w, byte 0 of word w + 1
}

// edges e, e + 1 (two) or edge e alone take the new messages rn (int16 halves v << 8);
// the touched words pass through an empty asm, so the compiler keeps them packed (four
// edges per register) instead of following each byte back to its int32 value
template <int NW>
__device__ __forceinline__ void c2v_put_pair(uint32_t (&c2v)[NW], int e, bool two, uint32_t rn) {
  const int w = e >> 2, b = e & 3;
  if (!two) {
    c2v[w] = __builtin_amdgcn_perm(rn, c2v[w], put_sel(b, 5u));
  } else if (b < 3) {
    c2v[w] = __builtin_amdgcn_perm(rn, c2v[w], put_sel2(b, 5u, 7u));
  } else {
    c2v[w] = __builtin_amdgcn_perm(rn, c2v[w], put_sel(3, 5u));
    c2v[w + 1] = __builtin_amdgcn_perm(rn, c2v[w + 1], put_sel(0, 7u));
    asm volatile("" : "+v"(c2v[w + 1]));
  }
  asm volatile("" : "+v"(c2v[w]));
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
constexpr int kKeepAbsMaxW = 12;   // biased form: largest row weight whose |q| values stay in registers

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
// One layer of row r (edges E0 .. E0 + w - 1, compiled-in weight, columns and shifts),
// with the workgroup barrier in front of it when the plan has one.  A thread with
// act == false (no codeword in its slot) reaches the barrier and does nothing else.
// Two forms of the row arithmetic, with the same results:
//  * PK = false: one edge per 32-bit lane operation; posteriors in LDS as bytes biased by
//    128 (b = v + 128).  q = sat8(L - R) is med3(b - R, 0, 255) on the biased value; |q| =
//    |b - 128| one v_sad_u8; min1/min2 a v_med3_u32 and a v_min_u32; the sign product the
//    XOR of the biased bytes (bit 7 set = non-negative).
//  * PK = true: two edges per 32-bit lane operation, as int16 halves; posteriors in LDS as
//    plain int8 bytes.  Fewer VALU instructions per edge (BG1: 4184 against 5159 per
//    iteration) but a longer dependency chain per layer (the halves' min1/min2 are merged
//    once per row), which pays in layers of many edges and costs in short ones.
template <class CT, int L, int E0, bool PK, int NB, int NW>
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
      if (P.pre[E0 + W + j]) {
        pout[j] = *lds_at(vb, edge_off<CT>(E0 + W + j));
        // opaque (two-edge form): kept as the loaded 32-bit value across the next barrier,
        // not narrowed to a byte and widened again (one AND per edge) on the other side
        if constexpr (PK) asm volatile("" : "+v"(pout[j]));
      }
  }
  if constexpr (PK) {
    // Two edges per 32-bit lane operation: edges j, j + 1 of the row (j even) are the low
    // and high int16 halves of one register, each value v held as v << 8 (low byte 0); an
    // odd row weight leaves the last edge alone in its low half.
    //  * Q = sat8(L - R): the saturating v_pk_sub of L << 8 and R << 8 is 256 (L - R) when
    //    that fits int16 and 32767 / -32768 otherwise, so its high byte is sat8(L - R); an
    //    AND clears the low byte 0xff of 32767.
    //  * |Q| = max(Q, -Q) read as unsigned 16-bit: |-128 << 8| = 0x8000 = 128 << 8 exactly.
    //    A lone edge's high half gets |Q| = 0xffff, above every real magnitude, and Q = 0.
    //  * min1/min2 of |Q| per half, then merged across the halves once per row; the sign
    //    parity is bit 15 xor bit 31 of the XOR of the Q registers.
    //  * With beta >= 1, A_j = max(|Q_j| - beta, 0), and the capped exclusive minimum is,
    //    per row, mv1 = min(cap, max(m1, beta) - beta), mv2 = min(max(m2, beta), cap + beta)
    //    - beta, and per edge M_j = (j is the argmin ? mv2 : mv1) = max(mv1, m1 + mv2 -
    //    |Q_j|): at the argmin |Q_j| = m1 and the term is mv2 (>= mv1); at any other edge
    //    |Q_j| >= m2, so the term is <= m1 + mv2 - m2, which is <= m1 - beta <= mv1 when m2 >=
    //    beta and the cap does not bind, <= mv2 = mv1 = cap when it does, below 0 = mv1 when
    //    m1 < beta, and <= 0 = mv1 = mv2 when m2 < beta.  Ties included.  The term lies in
    //    [m1 + mv2 - 128, mv2] (scaled by 256), so it is exact as a wrapped int16.
    //  * L' = sat8(Q + C2V): the saturating v_pk_add, whose high byte is the new posterior.
    constexpr int NP = (W + 1) / 2;
    uint32_t Qp[NP], Ap[NP];
    uint32_t S = 0u;
    u16x2 m1 = u16x2{0xffff, 0xffff}, m2 = u16x2{0xffff, 0xffff};
  #pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int j = 2 * k;
      const bool two = j + 1 < W;
      const uint32_t Lp = two ? __builtin_amdgcn_perm((uint32_t)pb[j + 1], (uint32_t)pb[j], 0x040c000cu)
                              : __builtin_amdgcn_perm(0u, (uint32_t)pb[j], 0x0c0c000cu);
      const uint32_t Rp = c2v_pair(c2v, E0 + j, two);
      const uint32_t qq = as_u(__builtin_elementwise_sub_sat(as_s(Lp), as_s(Rp))) & 0xff00ff00u;
      Qp[k] = qq;
      uint32_t a = as_u(__builtin_elementwise_max(as_s(qq), -as_s(qq)));
      if (!two) a |= 0xffff0000u;
      Ap[k] = a;
      if (k & 1) S = __builtin_amdgcn_bitop3_b32(S, Qp[k - 1], qq, 0x96);   // XOR of three
      else if (k + 1 == NP) S ^= qq;
      m2 = __builtin_elementwise_min(m2, __builtin_elementwise_max(m1, as_us(a)));
      m1 = __builtin_elementwise_min(m1, as_us(a));
    }
    // the row's min1 / min2 in both halves
    const u16x2 m1x = as_us(swap16(as_u(m1))), m2x = as_us(swap16(as_u(m2)));
    const u16x2 m1r = __builtin_elementwise_min(m1, m1x);
    const u16x2 m2r = __builtin_elementwise_min(__builtin_elementwise_max(m1, m1x), __builtin_elementwise_min(m2, m2x));
    S ^= swap16(S);   // bits 15 and 31: parity of the row's negative Q
    const uint32_t ce = (uint32_t)cap + be < 128u ? (uint32_t)cap + be : 128u;
    const u16x2 b2 = as_us(be * 0x01000100u), cb2 = as_us(ce * 0x01000100u);
    const s16x2 cap2 = as_s((uint32_t)cap * 0x01000100u);
    const s16x2 mv1 = __builtin_elementwise_min(cap2, as_s(as_u(__builtin_elementwise_max(m1r, b2) - b2)));
    const u16x2 mv2 = __builtin_elementwise_min(__builtin_elementwise_max(m2r, b2), cb2) - b2;
    uint32_t sumb_u = as_u(m1r + mv2);
    asm volatile("" : "+v"(sumb_u));   // one subtraction per edge pair, not re-associated into two
    const s16x2 sumb = as_s(sumb_u);
    // the layer's results: the new C2V pair and the new posteriors
    auto result = [&](int k, uint32_t* rn_out, uint32_t* gs_out) {
      const uint32_t qq = Qp[k];
      const uint32_t a = Ap[k];
      const s16x2 m = __builtin_elementwise_max(mv1, sumb - as_s(a));   // argmin ? min2 : min1, capped
      const uint32_t sm = as_u(as_s(S ^ qq) >> s16x2{15, 15});          // C2V sign: other edges' parity
      *rn_out = as_u(as_us(as_u(m) ^ sm) - as_us(sm));
      *gs_out = as_u(__builtin_elementwise_add_sat(as_s(qq), as_s(*rn_out))) >> 8;   // posteriors in bytes 0, 2
    };
    // the stores of pair k: C2V bytes (a word whose four edges are two pairs of this layer is
    // built from both at once, other pairs are inserted into their words) and posteriors
    auto store = [&](int k, uint32_t rnk, uint32_t rnprev, uint32_t g) {
      const int j = 2 * k, e = E0 + j;
      const bool two = j + 1 < W;
      const bool whole_lo = (e & 3) == 0 && two && e + 3 < E0 + W;
      const bool whole_hi = (e & 3) == 2 && two && e - 2 >= E0;
      if (whole_hi) {
        c2v[e >> 2] = __builtin_amdgcn_perm(rnk, rnprev, 0x07050301u);
        asm volatile("" : "+v"(c2v[e >> 2]));
      } else if (!whole_lo) {
        c2v_put_pair(c2v, e, two, rnk);
      }
      // opaque: the high edge's byte is then stored from bits 16..23, not shifted once more
      asm volatile("" : "+v"(g));
      const int K0 = edge_off<CT>(e);
      *lds_at(vb, K0 - Z) = (uint8_t)g;
      *lds_at(vb, K0) = (uint8_t)g;
      *lds_at(vb, K0 + Z) = (uint8_t)g;
      if (two) {
        const int K1 = edge_off<CT>(e + 1);
        *lds_at(vb, K1 - Z) = (uint8_t)(g >> 16);
        *lds_at(vb, K1) = (uint8_t)(g >> 16);
        *lds_at(vb, K1 + Z) = (uint8_t)(g >> 16);
      }
    };
    {   // all results first: the pairs' chains are independent, so the scheduler
               // interleaves them (no back-to-back dependent packed ops, which wait)
      uint32_t rn[NP], gs[NP];
  #pragma unroll
      for (int k = 0; k < NP; ++k) result(k, &rn[k], &gs[k]);
  #pragma unroll
      for (int k = 0; k < NP; ++k) store(k, rn[k], k ? rn[k - 1] : 0u, gs[k]);
    }
  } else {
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
}

// layers L .. mrow-1 of one iteration, then the barrier before the syndrome
template <class CT, int L, int E0, bool PK, int NB, int NW>
__device__ __forceinline__ void code_layers(const uint32_t (&vb)[NB], uint32_t (&c2v)[NW], const long long* bt,
                                            uint32_t beta, int cap, const int (&pin)[CT::w[L]], bool act) {
  const uint32_t be = (act && bt) ? (uint32_t)bt[L] : beta;
  int pout[wnext<CT, L>()];
  code_layer<CT, L, E0, PK>(vb, c2v, be, cap, pin, pout, act);
  __builtin_amdgcn_sched_barrier(0);   // layer by layer: bounded live ranges
  if constexpr (L + 1 < CT::mrow) code_layers<CT, L + 1, E0 + CT::w[L], PK>(vb, c2v, bt, beta, cap, pout, act);
  else __syncthreads();
}

// OR over layers L .. mrow-1 of the row's parity in bit 7 (set: check not satisfied):
// bit 7 of the XOR of the row's posterior bytes is the parity of its negative values (PK),
// or of its non-negative ones (biased bytes), flipped for an odd row weight
template <class CT, int L, int E0, bool PK, int NB>
__device__ __forceinline__ uint32_t code_syndrome(const uint32_t (&vb)[NB]) {
  constexpr int W = CT::w[L];
  uint32_t par = (!PK && (W & 1)) ? 0x80u : 0u;
  if constexpr (PK) {
#pragma unroll
    for (int j = 0; j + 1 < W; j += 2)
      par = __builtin_amdgcn_bitop3_b32(par, (uint32_t)*lds_at(vb, edge_off<CT>(E0 + j)),
                                        (uint32_t)*lds_at(vb, edge_off<CT>(E0 + j + 1)), 0x96);   // XOR of three
    if constexpr (W & 1) par ^= *lds_at(vb, edge_off<CT>(E0 + W - 1));
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) par ^= *lds_at(vb, edge_off<CT>(E0 + j));
  }
  if constexpr (L + 1 < CT::mrow) return par | code_syndrome<CT, L + 1, E0 + W, PK>(vb);
  else return par;
}


}  // namespace
}  // namespace fecsim

// Code-specialized int8 offset min-sum decoder, one codeword per workgroup, for codes
// with a large lifting size (5G NR base graph 1 at Z = 384).  The whole code -- lifting
// size, layer weights, columns and shifts -- is compiled in (code_tables.h, generated by
// fecsim/gen_code_tables.py), so every LDS address is a register plus an immediate offset
// and the C2V messages can live in registers.  Same results as the generic kernel
// (decode_core.h row_int, after src/LDPCDec/decode_functions.cpp:11-103), bit for bit:
//
// * A workgroup of Z threads owns one codeword.  Thread i owns check row
//   r = 4 (i mod Z/4) + i div (Z/4) of every layer, so the 64 lanes of a wave touch 64
//   distinct LDS dwords.
// * Posteriors live in LDS as bytes biased by 128 (b = v + 128, 0..255).  Column c keeps
//   two copies, at Z + 3Zc and Z + 3Zc + Z, with a Z-byte gap after each column (and one
//   before the first).  Row r of edge (c, s) reads its variable (s + r) mod Z at
//   Z + 3Zc + s + r without a wrap-around test, and writes the new value there and Z
//   below and above: two of the three are the variable's copies, the third lands in a gap.
// * A row's C2V messages stay in VGPRs as int8 bytes, four edges per register.
// * Per edge: q = sat8(L - R) is med3(b - R, 0, 255) on the biased value; |q| = |b - 128|
//   is one v_sad_u8; min1/min2 take a v_med3_u32 and a v_min_u32; the sign product is the
//   XOR of the biased bytes (bit 7 set = non-negative).  With beta >= 1 the exclusive
//   minimum follows from min1/min2 of |q| without an argmin (see the layer body).
// * A workgroup barrier separates two layers that share a column (a compile-time plan;
//   base graph 1 needs 32 per iteration for its 46 layers), and reads of columns not
//   written since the last barrier are issued before it.  The syndrome after each
//   iteration is the XOR of each row's biased bytes, and a workgroup OR ends the loop
//   (src/LDPCDec/LDPCdecode.cpp:172-197).
#include <mutex>

#include "decode_core.h"
#include "code_tables.h"

namespace fecsim {

struct CodeArgs {
  uint32_t num_cw, max_iter;
  int beta, cap;                // int8 beta (1..127) and cap (0..127)
  const long long* beta_tab;    // [max_iter][mrow] per-(iteration, layer) offsets, each 1..127, or null
  const int8_t* llr_in;         // dense [num_cw][mcol*z], 16-byte aligned
  int8_t* post_out;             // same layout (may equal llr_in)
  uint32_t* iters;
  int8_t* c2v;                  // [num_cw][ne*z] (edge-major, row r of edge e at e*z + r) or null
  unsigned long long* tot_iter;
};

namespace {

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
// with the workgroup barrier in front of it when the plan has one.
template <class CT, int L, int E0, int NB, int NW>
__device__ __forceinline__ void code_layer(const uint32_t (&vb)[NB], uint32_t (&c2v)[NW], uint32_t be, int cap,
                                           const int (&pin)[CT::w[L]], int (&pout)[wnext<CT, L>()]) {
  constexpr int Z = CT::z, W = CT::w[L];
  constexpr const CodePlan<CT>& P = kCodePlan<CT>;
  static_assert(W >= 2, "row weight below 2");
  int pb[W];
#pragma unroll
  for (int j = 0; j < W; ++j)
    if (P.pre[E0 + j]) pb[j] = pin[j];   // read during the previous layer
  if constexpr (P.bar[L]) {
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (P.early[E0 + j] && !P.pre[E0 + j]) pb[j] = *lds_at(vb, edge_off<CT>(E0 + j));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (!P.early[E0 + j]) pb[j] = *lds_at(vb, edge_off<CT>(E0 + j));
  } else {
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
                                            uint32_t beta, int cap, const int (&pin)[CT::w[L]]) {
  const uint32_t be = bt ? (uint32_t)bt[L] : beta;
  int pout[wnext<CT, L>()];
  code_layer<CT, L, E0>(vb, c2v, be, cap, pin, pout);
  __builtin_amdgcn_sched_barrier(0);   // layer by layer: bounded live ranges
  if constexpr (L + 1 < CT::mrow) code_layers<CT, L + 1, E0 + CT::w[L]>(vb, c2v, bt, beta, cap, pout);
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

template <class CT>
__global__ void __launch_bounds__(CT::z) __attribute__((amdgpu_waves_per_eu(4))) k_code_i8(const CodeArgs a) {
  constexpr int Z = CT::z, Q = Z / 4, NE = CT::ne, NW = (NE + 3) / 4, MC = CT::mcol;
  constexpr int LDS = (3 * MC + 1) * Z;
  constexpr int NB = (LDS + 32767) / 32768;   // base registers, one per 32 KiB: offsets stay < 32768
  constexpr int CH = MC * Z / 16, CPC = Z / 16;
  static_assert(Z % 16 == 0 && Z <= 1024, "Z: multiple of 16, at most 1024");
  HIP_DYNAMIC_SHARED(uint8_t, smem);
  const int i = threadIdx.x;
  const int r = 4 * (i % Q) + i / Q;
  const uint32_t lbase = (uint32_t)(size_t)(LdsU8*)smem;
  uint32_t vb[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    vb[k] = lbase + (uint32_t)r + 32768u * (uint32_t)k;
    asm volatile("" : "+v"(vb[k]));   // kept as separate bases, not re-folded into one
  }
  unsigned long long tot = 0;

  for (long long cw = blockIdx.x; cw < (long long)a.num_cw; cw += gridDim.x) {   // uniform per block
    // ---- codeword in: both copies of each column, biased ----
    const uint4* gi = reinterpret_cast<const uint4*>(a.llr_in + cw * (long long)(MC * Z));
    for (int k = i; k < CH; k += Z) {
      uint4 v = gi[k];
      v.x ^= 0x80808080u;
      v.y ^= 0x80808080u;
      v.z ^= 0x80808080u;
      v.w ^= 0x80808080u;
      const int c = k / CPC, o = (k - c * CPC) * 16;
      *reinterpret_cast<uint4*>(smem + Z + 3 * Z * c + o) = v;
      *reinterpret_cast<uint4*>(smem + 2 * Z + 3 * Z * c + o) = v;
    }
    __syncthreads();

    uint32_t c2v[NW];   // C2V of edges 4k..4k+3 (row r), int8 bytes; LDPCdecode.cpp:115-119 zeroes them
#pragma unroll
    for (int k = 0; k < NW; ++k) c2v[k] = 0u;
    uint32_t iter = 0;
    for (;;) {
      const long long* bt = a.beta_tab ? a.beta_tab + (size_t)iter * CT::mrow : nullptr;
      int pin0[CT::w[0]] = {};   // layer 0 reads nothing ahead
      code_layers<CT, 0, 0>(vb, c2v, bt, (uint32_t)a.beta, a.cap, pin0);
      const uint32_t F = code_syndrome<CT, 0, 0>(vb);   // syndrome after the full iteration
      ++iter;
      const int any = __syncthreads_or((int)(F & 0x80u));
      if (!any || iter == a.max_iter) break;
    }

    // ---- codeword out ----
    uint4* go = reinterpret_cast<uint4*>(a.post_out + cw * (long long)(MC * Z));
    for (int k = i; k < CH; k += Z) {
      const int c = k / CPC, o = (k - c * CPC) * 16;
      uint4 v = *reinterpret_cast<const uint4*>(smem + Z + 3 * Z * c + o);
      v.x ^= 0x80808080u;
      v.y ^= 0x80808080u;
      v.z ^= 0x80808080u;
      v.w ^= 0x80808080u;
      go[k] = v;
    }
    if (a.c2v) {
      int8_t* gc = a.c2v + cw * (long long)(NE * Z) + r;
#pragma unroll
      for (int e = 0; e < NE; ++e) gc[(size_t)e * Z] = (int8_t)c2v_byte(c2v[e >> 2], e & 3);
    }
    if (i == 0) {
      if (a.iters) a.iters[cw] = iter;
      tot += iter;
    }
    __syncthreads();   // LDS reads of this codeword done before the next one is loaded
  }
  if (a.tot_iter && i == 0 && tot) atomicAdd(a.tot_iter, tot);
}

template <class CT>
hipError_t launch_code(const CodeArgs& a, hipStream_t st) {
  if (a.num_cw == 0) return hipSuccess;
  constexpr size_t lds = (size_t)(3 * CT::mcol + 1) * CT::z;
  // the LDS attribute is set once per device (mutex: host threads may launch
  // concurrently on distinct handles)
  static std::mutex mu;
  static bool ready_by_dev[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!ready_by_dev[dev]) {
      if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_code_i8<CT>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
      }
      ready_by_dev[dev] = true;
    }
  }
  // one workgroup per codeword: the dispatcher hands a finished workgroup's CU the next
  // codeword, so codewords that stop early free their CU at once
  const uint32_t blocks = a.num_cw < 0x7fffffffu ? a.num_cw : 0x7fffffffu;
  hipLaunchKernelGGL((k_code_i8<CT>), dim3(blocks), dim3(CT::z), lds, st, a);
  return hipGetLastError();
}

}  // namespace

int code_i8_index(const DecoderDev& d) {
  const HdecHost& h = d.h;
  if (d.elem != FECSIM_I8 || d.mode != FECSIM_MIN) return -1;
  const float beta = d.params.beta_offset, mrm = d.params.max_row_metric;
  if (mrm < 1e-6f) return -1;   // cap 1 - max_row_metric: generic kernel
  const int b8 = (int)(int8_t)(int)beta, c8 = (int)(int8_t)(int)mrm;
  if (b8 < 1 || c8 < 0) return -1;
  return code_tab_index(h.z, h.mcol, h.mrow, h.row_weights.data(), h.row_active_cols.data(), h.row_offset.data());
}

bool code_i8_offsets_ok(const std::vector<long long>& tab) {
  for (long long b : tab)
    if (b < 1 || b > 127) return false;
  return true;
}

hipError_t launch_code_i8(const DecoderDev& d, const KernelArgs& ka, hipStream_t st) {
  CodeArgs a{};
  a.num_cw = ka.num_cw;
  a.max_iter = d.params.max_iter;
  a.beta = (int)(int8_t)(int)d.params.beta_offset;
  a.cap = (int)(int8_t)(int)d.params.max_row_metric;
  a.beta_tab = reinterpret_cast<const long long*>(d.d_beta_tab);
  a.llr_in = reinterpret_cast<const int8_t*>(ka.metrics_in);
  a.post_out = reinterpret_cast<int8_t*>(ka.metrics_out);
  a.iters = ka.iters;
  a.c2v = reinterpret_cast<int8_t*>(ka.c2v);
  a.tot_iter = ka.tot_iter;
  return code_dispatch(d.code_tab, a, st);
}

#define FECSIM_CODE_LAUNCH(CT) return launch_code<CT>(a, st);
#include "code_tables_dispatch.h"

}  // namespace fecsim

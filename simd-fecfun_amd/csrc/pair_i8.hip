// Row-pair int8 offset min-sum decoder for even Z, specialized per code shape
// (layer count and row weights known at compile time; shifts and columns are read
// from the decoder's table at run time).  Same results as the generic kernel
// (decode_core.h), bit for bit; it only reorganizes the work for gfx950:
//
// * Rows r and r + Z/2 of a layer are handled by one lane as the two 16-bit halves of
//   a VGPR, so every VALU op is a packed v_pk_* op on two rows.  For Z = 42 a
//   codeword takes 21 lanes and a wave decodes 3 codewords.
// * Posteriors live in LDS as int16 pairs D[c][t] = (L[c][t], L[c][(t + Z/2) mod Z]);
//   each variable is stored twice (low half of pair t, high half of pair t - Z/2), so
//   the lane reads its two rows' variables with one ds_read_b32 and no rotation fix-up,
//   and writes the pair and its half-swapped twin.
// * The slots of a wave share its LDS columns: column c of a wave is two rows of 64
//   dwords (wave bases are multiples of 512 bytes); pair t of slot s sits in row
//   (t >= Z/2) at dword (t mod Z/2) + s*Z/2, so a column holds all slots' 2*Z/2 pairs
//   and the twin of pair t (pair t + Z/2 mod Z) is at byte address (address of t) xor
//   256.  A slot's lanes read Z/2 pairs with distinct t mod Z/2, hence distinct banks.
// * The per-lane LDS byte address of every edge's pair stays in a VGPR for the whole
//   decode; the C2V messages too, as int8 bytes, two edges per VGPR.
// Values are int8 in the high byte of each int16 half.  Q, |Q| and C2V messages are
// v << 8; a posterior is (v << 8) | l with l = 0, or l = 0xff where a saturating add
// stopped at 32767 (v = 127):
// * Q = sat8(L - R): the int16 saturating v_pk_sub (clamp bit) of (L << 8 | l) and
//   R << 8 is 256 (L - R) + l when L - R lies in [-128, 127] (it fits int16) and
//   saturates to 32767 / -32768 otherwise, so its high byte is sat8(L - R); one AND
//   with 0xff00ff00 leaves Q << 8.
// * L = sat8(Q + R): the saturating v_pk_add of Q << 8 and R << 8, whose high byte is
//   sat8(Q + R) and whose low byte is 0 or, at 32767, 0xff.
// * |Q| is max(Q, -Q) read as unsigned 16-bit, so |-128 << 8| = 0x8000 = 128 << 8
//   exactly; R = (S xor sign Q) ? -M : M.
#include <mutex>

#include "decode_core.h"
#include "pair_shapes.h"

namespace fecsim {

typedef short __attribute__((ext_vector_type(2))) short2v;
typedef unsigned short __attribute__((ext_vector_type(2))) ushort2v;

__device__ __forceinline__ uint32_t as_u(short2v v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ short2v as_s(uint32_t v) { return __builtin_bit_cast(short2v, v); }
__device__ __forceinline__ ushort2v as_us(uint32_t v) { return __builtin_bit_cast(ushort2v, v); }
__device__ __forceinline__ uint32_t as_u(ushort2v v) { return __builtin_bit_cast(uint32_t, v); }

// LDS word at an absolute LDS byte address held in a VGPR (no base added per access)
typedef __attribute__((address_space(3))) uint32_t LdsU32;
__device__ __forceinline__ LdsU32* lds32(uint32_t byte_addr) { return (LdsU32*)(size_t)byte_addr; }
typedef __attribute__((address_space(3))) uint16_t LdsU16;
__device__ __forceinline__ LdsU16* lds16(uint32_t byte_addr) { return (LdsU16*)(size_t)byte_addr; }

// C2V messages of edges 2k and 2k+1 share a VGPR as int8 bytes: edge 2k in bytes 1 and 3
// (rows lr, lr + Z/2), edge 2k+1 in bytes 0 and 2.  c2v_get returns edge e's pair as
// int16 halves v << 8 (low bytes 0).  e is a compile-time index in the unrolled loops.
template <int N>
__device__ __forceinline__ uint32_t c2v_get(const uint32_t (&p)[N], int e) {
  return (e & 1) ? __builtin_amdgcn_perm(0u, p[e >> 1], 0x020C000Cu) : (p[e >> 1] & 0xff00ff00u);
}
// store edge e's new pair r (int16 halves v << 8), keeping its partner's bytes
template <int N>
__device__ __forceinline__ void c2v_put(uint32_t (&p)[N], int e, uint32_t r) {
  p[e >> 1] = (e & 1) ? __builtin_amdgcn_perm(r, p[e >> 1], 0x03070105u) : __builtin_amdgcn_perm(r, p[e >> 1], 0x07020500u);
}
// both edges of a word at once: r_even, r_odd as int16 halves v << 8
__device__ __forceinline__ uint32_t c2v_pack(uint32_t r_even, uint32_t r_odd) {
  return __builtin_amdgcn_perm(r_even, r_odd, 0x07030501u);
}
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ short2v smax(short2v a, short2v b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ short2v smin(short2v a, short2v b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ ushort2v umin(ushort2v a, ushort2v b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ ushort2v umax(ushort2v a, ushort2v b) { return __builtin_elementwise_max(a, b); }

struct PairArgs {
  const uint32_t* adrtab;  // [cw_per_wave][half][ne]: per slot, lane and edge, the LDS byte
                           // offset (in the wave's columns) of the pair the lane reads
  int z, half, mcol;       // Z, Z/2
  int cw_per_wave, waves;
  int col_dwords;          // dwords per wave column (128; passed at run time so the refill and
                           // write-out loops are not unrolled into extra live registers)
  int wave_dwords;         // dwords per wave (mcol * col_dwords)
  uint32_t num_cw, max_iter;
  long long chunk;         // codewords per wave (persistent grid)
  int max_blocks;          // resident-grid size bound (0: one block per cw_per_wave*waves codewords)
  int beta, cap;           // int8 beta (1..127) and cap (0..127)
  const int8_t* beta_tab;  // [max_iter][mrow] per-(iteration, layer) int8 offsets (0..127) or null
  const int8_t* llr_in;    // dense [num_cw][mcol*z]
  int8_t* post_out;
  uint32_t* iters;
  int8_t* c2v;             // [num_cw][E] or null
  unsigned long long* tot_iter;
  int ct;                  // the decoder's edge table is the shape's compiled-in one (k_pair_i8c)
  unsigned* ctr;           // k_pair_i8c work queue: counter g of codeword range g at ctr[64 g]
  int qgroups;             // number of ranges (and counters), at most 64
  int lead[4];             // k_pair_i8c: lane of each slot's row 0, and each slot's lanes
  unsigned long long smask[4];
};

// waves per SIMD the register allocation aims at: 4 when the edge addresses, the packed
// C2V bytes and a layer's working set fit 128 VGPRs, else no constraint
template <class SH>
constexpr int pair_waves_per_eu() {
  return (SH::ne + 1) / 2 + SH::ne + 3 * SH::maxw + 20 <= 128 ? 4 : 1;
}

// ZC: the lifting size compiled in (SH::zc, with SH::mc columns), or 0 for a run-time Z.
// With ZC the codeword's loads and stores take immediate offsets from one base address
// and its column loops unroll (a refill or write-out costs 1-2 VALU per column instead of
// the 64-bit address arithmetic of each byte).
// BT: per-(iteration, layer) offsets from a.beta_tab, copied behind the waves' columns in
// LDS; each lane reads its slot's offset once per layer.
template <class SH, int ZC, bool BT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(pair_waves_per_eu<SH>())))
k_pair_i8(const PairArgs a) {
  constexpr int NE = SH::ne;
  constexpr bool FZ = ZC != 0;
  HIP_DYNAMIC_SHARED(uint32_t, smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = FZ ? ZC / 2 : a.half, Z = FZ ? ZC : a.z, CPW = a.cw_per_wave;
  const bool real = lane < CPW * L;          // lanes past the last slot mirror lane 0
  const int slot = real ? lane / L : 0;
  const int lr = real ? lane - slot * L : 0;
  uint32_t* wbase = smem + (size_t)wave * a.wave_dwords;
  const int pos = lr + slot * L;             // this lane's pair position in a column row

  // per-edge LDS byte address of the pair this lane reads (t = (s + lr) mod Z), absolute
  // in the block's LDS; its twin is at address ^ 256
  const uint32_t wb_off = (uint32_t)(size_t)(LdsU32*)wbase;   // multiple of 512
  uint32_t adr[NE];
  const uint32_t* at = a.adrtab + ((size_t)slot * L + lr) * NE;
#pragma unroll
  for (int e = 0; e < NE; ++e) adr[e] = at[e] + wb_off;

  // persistent work: this wave decodes codewords [w0, w1); each slot takes the next one
  // as soon as its current codeword stops
  const long long gw = (long long)blockIdx.x * a.waves + wave;
  const long long w0 = gw * a.chunk;
  long long w1 = w0 + a.chunk;
  if (w1 > (long long)a.num_cw) w1 = a.num_cw;
  long long next = w0;
  long long cw = -1;                 // codeword of this lane's slot, -1 when idle
  bool active = false;
  uint32_t iter = 0;
  // C2V of edges 2k and 2k+1, int8 bytes [2k row lr, 2k+1 row lr, 2k row lr+L, 2k+1 row lr+L]
  uint32_t c2vp[(NE + 1) / 2];
#pragma unroll
  for (int k = 0; k < (NE + 1) / 2; ++k) c2vp[k] = 0u;
  const ushort2v beta2 = ushort2v{(unsigned short)(a.beta << 8), (unsigned short)(a.beta << 8)};
  const short2v cap2 = short2v{(short)(a.cap << 8), (short)(a.cap << 8)};
  // cap + beta, at most 128 (|Q| <= 128, so a larger cap never binds): min(cap, m' - beta)
  // = min(m', capb) - beta
  const int capb_i = (a.cap + a.beta < 128 ? a.cap + a.beta : 128) << 8;
  const ushort2v capb2 = ushort2v{(unsigned short)capb_i, (unsigned short)capb_i};
  const unsigned long long gmask = (L >= 64) ? ~0ull : ((1ull << L) - 1ull);
  unsigned long long my_iters = 0;   // summed by lane 0 of each slot
  // offset table: [max_iter][mrow] bytes behind the waves' columns
  const uint8_t* btab = reinterpret_cast<const uint8_t*>(smem + (size_t)a.waves * a.wave_dwords);
  if constexpr (BT) {
    uint8_t* bt = const_cast<uint8_t*>(btab);
    const int nb = (int)a.max_iter * SH::mrow;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bt[i] = (uint8_t)a.beta_tab[i];
    __syncthreads();
  }

  for (;;) {
    // ---- refill: idle slots take the next codewords of the wave's range ----
    const unsigned long long want = __ballot(real && lr == 0 && !active);
    if (want != 0ull && next < w1) {
      const int rank = __popcll(want & ((1ull << (slot * L)) - 1ull));
      const long long mine = next + rank;
      const bool take = !active && mine < w1;
      next += __popcll(want);
      if (take) {
        cw = mine;
        active = true;
        iter = 0;
        if (real) {
          if constexpr (FZ) {
            // columns in groups of four: eight byte loads at immediate offsets, then per
            // column two byte permutes, the pair (lo, hi) and its twin (hi, lo) as int16
            // halves v << 8 (groups bound the registers live beside the decode state)
            const uint8_t* g = reinterpret_cast<const uint8_t*>(a.llr_in) + cw * (long long)(SH::mc * ZC) + lr;
            uint32_t* wp = wbase + pos;
            static_assert(SH::mc % 4 == 0, "column groups of four");
#pragma unroll
            for (int c0 = 0; c0 < SH::mc; c0 += 4) {
              uint32_t lo[4], hi[4];
#pragma unroll
              for (int k = 0; k < 4; ++k) {
                lo[k] = g[(c0 + k) * ZC];
                hi[k] = g[(c0 + k) * ZC + ZC / 2];
              }
#pragma unroll
              for (int k = 0; k < 4; ++k) {
                wp[(c0 + k) * 128] = __builtin_amdgcn_perm(hi[k], lo[k], 0x040c000cu);
                wp[(c0 + k) * 128 + 64] = __builtin_amdgcn_perm(lo[k], hi[k], 0x040c000cu);
              }
              __builtin_amdgcn_sched_barrier(0);
            }
          } else {
            const int8_t* g = a.llr_in + cw * (long long)(a.mcol * Z);
            for (int c = 0; c < a.mcol; ++c) {
              const int lo = g[c * Z + lr], hi = g[c * Z + lr + L];
              const uint32_t p = ((uint32_t)(lo & 0xff) << 8) | ((uint32_t)(hi & 0xff) << 24);
              wbase[c * a.col_dwords + pos] = p;
              wbase[c * a.col_dwords + pos + 64] = (p >> 16) | (p << 16);
            }
          }
        }
      }
      const uint32_t keep = take ? 0u : ~0u;   // fresh codeword: C2V messages start at zero
#pragma unroll
      for (int k = 0; k < (NE + 1) / 2; ++k) c2vp[k] &= keep;
    }
    if (__ballot(active) == 0ull) break;
    wave_sync();

    // ---- one layered iteration; lanes of idle slots compute on their own (dead) slot ----
#pragma unroll
    for (int e = 0; e < NE; ++e) asm volatile("" : "+v"(adr[e]));   // unpack at each use, not hoisted
    int e0 = 0;
    // this slot's offsets in iteration `iter` (idle slots keep counting iterations: clamped
    // so that their reads stay inside the table)
    const uint8_t* bt_it = btab + (iter < a.max_iter ? iter : a.max_iter - 1u) * SH::mrow;
#pragma unroll
    for (int l = 0; l < SH::mrow; ++l) {
      constexpr int MAXW = SH::maxw;
      const int w = SH::w[l];
      ushort2v beta2l = beta2, capb2l = capb2;
      // |Q| mask: with an offset of 0 the reference's abs(-128) = -128 gives A = max(-128, 0)
      // = 0 (decode_functions.cpp:56-61), and |Q| = 0x8000 is the only magnitude with bit 15
      uint32_t amask = 0xffffffffu;
      if constexpr (BT) {   // the layer's offset b: beta2 = b << 8, capb2 = min(cap + b, 128) << 8
        const uint32_t b = bt_it[l];
        const uint32_t cb = b + (uint32_t)a.cap;
        beta2l = as_us(b * 0x01000100u);
        capb2l = as_us((cb < 128u ? cb : 128u) * 0x01000100u);
        amask = b ? 0xffffffffu : 0x7fff7fffu;
      }
      // Min1/min2 are tracked on |Q_j| (unsigned, |-128| = 128).  With beta >= 1,
      // A_j = max(|Q_j| - beta, 0) and the capped exclusive minimum is, per row,
      //   mv1 = min(cap, max(m1, beta) - beta),  mv2 = min(max(m2, beta), cap + beta) - beta,
      //   M_j = (j is the argmin ? mv2 : mv1) = max(mv1, m1 + mv2 - |Q_j|):
      // at the argmin |Q_j| = m1 and the term is mv2 (>= mv1); at any other edge
      // |Q_j| >= m2, so the term is <= m1 + mv2 - m2, which is <= m1 - beta <= mv1 when
      // m2 >= beta and the cap does not bind, <= mv2 = mv1 = cap when it does, below
      // 0 = mv1 when m1 < beta, and <= 0 = mv1 = mv2 when m2 < beta.  Ties included.
      // The term lies in [m1 + mv2 - 128, mv2] (scaled by 256), so it is exact as a
      // wrapped int16 for every input.
      short2v q[MAXW];
      ushort2v A[MAXW];
      uint32_t S = 0;   // XOR of the layer's Q (sign bits)
      ushort2v m1 = ushort2v{0xffff, 0xffff}, m2 = ushort2v{0xffff, 0xffff};
#pragma unroll
      for (int j = 0; j < MAXW; ++j) {
        if (j < w) {
          const uint32_t x = *lds32(adr[e0 + j]);
          const short2v qs = __builtin_elementwise_sub_sat(as_s(x), as_s(c2v_get(c2vp, e0 + j)));
          const short2v qq = as_s(as_u(qs) & 0xff00ff00u);   // Q << 8, Q = sat8(L - R)
          q[j] = qq;
          const ushort2v av = BT ? as_us(as_u(smax(qq, -qq)) & amask) : as_us(as_u(smax(qq, -qq)));
          A[j] = av;
          if (j & 1) S = xor3(S, as_u(q[j - 1]), as_u(qq));
          else if (j + 1 == w) S ^= as_u(qq);
          m2 = umin(m2, umax(m1, av));
          m1 = umin(m1, av);
        }
      }
      const short2v mv1 = smin(cap2, as_s(as_u(umax(m1, beta2l) - beta2l)));
      const short2v mv2 = as_s(as_u(umin(umax(m2, beta2l), capb2l) - beta2l));
      uint32_t sumb_u = as_u(m1 + as_us(as_u(mv2)));
      asm volatile("" : "+v"(sumb_u));   // one subtraction per edge, not re-associated into two
      const short2v sumb = as_s(sumb_u);
      // all of the layer's results first, then the stores: the per-edge chains are
      // independent, so the scheduler can interleave them (no back-to-back dependent
      // packed ops waiting on each other)
      uint32_t gu[MAXW];
      uint32_t pend = 0;   // new C2V of an even edge whose odd partner follows in this layer
#pragma unroll
      for (int j = 0; j < MAXW; ++j) {
        if (j < w) {
          const short2v m = smax(mv1, sumb - as_s(as_u(A[j])));   // argmin ? min2 : min1, capped
          const uint32_t sm = as_u(as_s(S ^ as_u(q[j])) >> short2v{15, 15});
          const uint32_t rn = as_u(as_us(as_u(m) ^ sm) - as_us(sm));
          // L = sat8(Q + R) in the high bytes
          const uint32_t g = as_u(__builtin_elementwise_add_sat(q[j], as_s(rn)));
          const int e = e0 + j;   // compile-time after unrolling
          if ((e & 1) == 0 && j + 1 < w) pend = rn;
          else if ((e & 1) == 1 && j > 0) c2vp[e >> 1] = c2v_pack(pend, rn);   // both bytes new
          else c2v_put(c2vp, e, rn);
          gu[j] = g;
        }
      }
#pragma unroll
      for (int j = 0; j < MAXW; ++j) {
        if (j < w) {
          const uint32_t ad = adr[e0 + j];
          *lds32(ad) = gu[j];
          *lds32(ad ^ 256u) = __builtin_amdgcn_alignbit(gu[j], gu[j], 16);   // twin: halves swapped
        }
      }
      e0 += w;
      wave_sync();
      __builtin_amdgcn_sched_barrier(0);   // schedule layer by layer: bounded live ranges
    }
    // ---- syndrome after the full iteration, LDPCdecode.cpp:172-195 ----
#pragma unroll
    for (int e = 0; e < NE; ++e) asm volatile("" : "+v"(adr[e]));   // no reuse of the layers' unpacked addresses
    // F = OR over rows of the XOR of the row's posteriors (sign bits 15 and 31 matter)
    uint32_t F = 0;
    e0 = 0;
#pragma unroll
    for (int l = 0; l < SH::mrow; ++l) {
      const int w = SH::w[l];
      uint32_t xs[SH::maxw];
#pragma unroll
      for (int j = 0; j < SH::maxw; ++j)
        if (j < w) xs[j] = *lds32(adr[e0 + j]);
      uint32_t par = xs[0];
#pragma unroll
      for (int j = 1; j + 1 < SH::maxw; j += 2)
        if (j + 1 < w) par = xor3(par, xs[j], xs[j + 1]);
      if (w & 1) F |= par;
      else F = __builtin_amdgcn_bitop3_b32(F, par, xs[w - 1], 0xF6);   // F | (par ^ x)
      e0 += w;
      __builtin_amdgcn_sched_barrier(0);
    }
    const bool fail = (F & 0x80008000u) != 0u;
    const unsigned long long fb = __ballot(real && active && fail);
    ++iter;
    const bool cwfail = ((fb >> (slot * L)) & gmask) != 0ull;
    const bool finish = active && (!cwfail || iter == a.max_iter);   // LDPCdecode.cpp:192-197
    if (finish && real) {
      if constexpr (FZ) {
        // columns in groups of eight: the LDS reads first, then the byte stores at
        // immediate offsets
        const uint32_t* wp = wbase + pos;
        uint8_t* g = reinterpret_cast<uint8_t*>(a.post_out) + cw * (long long)(SH::mc * ZC) + lr;
        static_assert(SH::mc % 8 == 0, "column groups of eight");
#pragma unroll
        for (int c0 = 0; c0 < SH::mc; c0 += 8) {
          uint32_t v[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) v[k] = wp[(c0 + k) * 128];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const uint32_t x = v[k] >> 8;
            g[(c0 + k) * ZC] = (uint8_t)x;
            g[(c0 + k) * ZC + ZC / 2] = (uint8_t)(x >> 16);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        int8_t* g = a.post_out + cw * (long long)(a.mcol * Z);
        for (int c = 0; c < a.mcol; ++c) {
          const uint32_t p = wbase[c * a.col_dwords + pos];
          g[c * Z + lr] = (int8_t)(p >> 8);
          g[c * Z + lr + L] = (int8_t)(p >> 24);
        }
      }
      if (a.c2v) {
        // opaque row offset: keeps the NE store offsets from being hoisted out of the
        // decode loop as loop-invariant live registers
        int zo = Z;
        asm volatile("" : "+v"(zo));
        int8_t* gc = a.c2v + cw * (long long)(NE * Z) + lr;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          const uint32_t r = c2v_get(c2vp, e);
          gc[e * zo] = (int8_t)(r >> 8);
          gc[e * zo + L] = (int8_t)(r >> 24);
        }
      }
      if (lr == 0) {
        if (a.iters) a.iters[cw] = iter;
        my_iters += iter;
      }
    }
    active = active && !finish;
  }
  if (a.tot_iter) {
    unsigned long long s = 0;
    for (int q = 0; q < CPW; ++q) s += (unsigned long long)__shfl((long long)my_iters, q * L, 64);
    if (lane == 0 && s) atomicAdd(a.tot_iter, s);
  }
}

// ---------------------------------------------------------------------------
// Instance with the code's edge table compiled in (fixed-Z shapes whose decoder table is
// the P = 64 int8 one).  The same decode as k_pair_i8, bit for bit, with fewer registers
// and VALU per edge:
// * Edge addresses.  For an edge of shift s the lane reads pair t = (s + lr) mod Z, at
//   position (c + lr) mod Z/2 of its column with c = s mod Z/2, in row
//   [(s + lr) mod Z >= Z/2].  Per shift class c the lane keeps that address in both rows,
//   ga[k] (row [c + lr >= Z/2]) and gb[k] = ga[k] ^ 256: an edge with s < Z/2 reads ga
//   and writes its twin at gb, one with s >= Z/2 the other way round, and the column is
//   an immediate offset.  2 registers per class (17 for 802.11ad 1/2) instead of one per
//   edge (52), and no address arithmetic in the loop.
// * C2V messages stay unpacked, one register per edge (both rows' messages as int16
//   halves v << 8), subtracted and replaced as they are: no byte unpack or repack.
// * The signed message is R = M * k, a 16-bit multiply by k = (Sa ^ sign Q_j) | 1 = +-1
//   per half, Sa the sign mask of the row's XOR: (m << 8) * 0xffff = -(m << 8) mod 2^16.
// * mv2 = min(cap, sat(m2 - beta)), which equals min(max(m2, beta), cap + beta) - beta for
//   |Q| <= 128.
// Per edge: its column is written by no later layer of the iteration, so its posterior
// after the edge's own layer is the one the syndrome reads (taken from the register, not
// read back); per layer: it has such an edge.
template <class SH>
struct PairFinal {
  struct Tab {
    bool e[SH::ne];
    bool l[SH::mrow];
  };
  static constexpr Tab make() {
    Tab t{};
    int e0 = 0;
    for (int l = 0; l < SH::mrow; ++l) {
      t.l[l] = false;
      for (int j = 0; j < SH::w[l]; ++j) {
        bool fin = true;
        int x = e0 + SH::w[l];
        for (int l2 = l + 1; l2 < SH::mrow; ++l2) {
          for (int k = 0; k < SH::w[l2]; ++k)
            if (SH::ecol[x + k] == SH::ecol[e0 + j]) fin = false;
          x += SH::w[l2];
        }
        t.e[e0 + j] = fin;
        t.l[l] = t.l[l] || fin;
      }
      e0 += SH::w[l];
    }
    return t;
  }
  static constexpr Tab tab = make();
};

// 4 waves per SIMD when the C2V registers, the class addresses and a layer's working set
// fit 128 VGPRs (802.11ad 1/2), else no constraint (the other shapes take 129-147)
template <class SH>
constexpr int pair_c_waves_per_eu() {
  return SH::ne + 2 * SH::ncls + 3 * SH::maxw + 14 < 128 ? 4 : 1;
}

template <class SH, bool BT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(pair_c_waves_per_eu<SH>())))
k_pair_i8c(const PairArgs a) {
  constexpr int NE = SH::ne, ZC = SH::zc, L = ZC / 2, NC = SH::ncls;
  static_assert(SH::ct && ZC % 2 == 0 && L <= 32, "compiled edge table, even Z <= 64");
  HIP_DYNAMIC_SHARED(uint32_t, smem);
  const int lane = threadIdx.x;              // one wave per block
  // per-lane layout from pair_layout (a.adrtab): the class addresses (absolute LDS byte
  // addresses, the wave's columns start at 0), the lane's pair position in a column row,
  // its row lr and slot (a.lead / a.smask give each slot's row-0 lane and lanes).  Lanes,
  // slots and pair positions are placed to keep LDS bank conflicts low; the spare lane
  // mirrors a lane of its own half-wave (same addresses, same values: a broadcast)
  const uint32_t* lt = a.adrtab + (size_t)lane * (NC + 2);
  uint32_t ga[NC], gb[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    ga[k] = lt[k];
    gb[k] = ga[k] ^ 256u;
  }
  const int pos = (int)lt[NC];
  const uint32_t info = lt[NC + 1];
  const int lr = (int)(info & 0xffu);
  const int sl = (int)((info >> 8) & 0xffu);   // slot
  const bool real = ((info >> 16) & 1u) != 0u;
  uint32_t* wbase = smem;

  // work queue: the grid is the resident blocks, and the blocks of group g = blockIdx %
  // qgroups take batches of QB codewords of the group's share of the range from counter g;
  // the next batch is fetched one ahead.  Every block works until its group's range is
  // done, so no block idles while another still holds a long static chunk; one counter
  // per 64 blocks keeps the same-address atomics far below their ~100 ns serial rate.
  constexpr int QB = 8;
  const int qg = (int)(blockIdx.x % (unsigned)a.qgroups);
  const long long gs = (long long)a.num_cw * qg / a.qgroups, ge = (long long)a.num_cw * (qg + 1) / a.qgroups;
  unsigned vq = 0;
  if (lane == 0) vq = atomicAdd(a.ctr + qg * 64, (unsigned)QB);
  long long bcur = gs + (long long)__builtin_amdgcn_readfirstlane(vq);
  long long bend = bcur + QB < ge ? bcur + QB : ge;
  bool exhausted = bcur >= ge;
  if (!exhausted && lane == 0) vq = atomicAdd(a.ctr + qg * 64, (unsigned)QB);   // next batch, in flight
  uint32_t cw = 0;                           // codeword of this lane's slot (when active)
  bool active = false;
  uint32_t iter = 0;
  typedef uint32_t __attribute__((ext_vector_type(2))) u32x2;
  u32x2 rp[(NE + 1) / 2];                    // C2V of edges 2k, 2k+1: rows lr, lr + L as int16 halves v << 8
#pragma unroll
  for (int k = 0; k < (NE + 1) / 2; ++k) rp[k] = u32x2{0u, 0u};
  const ushort2v beta2 = ushort2v{(unsigned short)(a.beta << 8), (unsigned short)(a.beta << 8)};
  const ushort2v cap2 = ushort2v{(unsigned short)(a.cap << 8), (unsigned short)(a.cap << 8)};
  unsigned long long my_iters = 0;
  const uint8_t* btab = reinterpret_cast<const uint8_t*>(smem + (size_t)SH::mc * 128);
  if constexpr (BT) {
    uint8_t* bt = const_cast<uint8_t*>(btab);
    const int nb = (int)a.max_iter * SH::mrow;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bt[i] = (uint8_t)a.beta_tab[i];
    __syncthreads();
  }

  for (;;) {
    // ---- refill: idle slots take the next codewords of the block's range ----
    int lead = a.lead[0];
#pragma unroll
    for (int k = 1; k < 64 / L; ++k) lead = sl == k ? a.lead[k] : lead;
    const unsigned long long want = __ballot(real && lane == lead && !active);
    if (want != 0ull && !exhausted) {
      const int nwant = __popcll(want);
      const int rank = __popcll(want & ((1ull << lead) - 1ull));
      const long long avail = bend - bcur;
      long long mine = bcur + rank;
      bool valid = rank < avail;
      if (nwant > avail) {   // the batch runs out: continue with the prefetched one
        const long long bn = gs + (long long)__builtin_amdgcn_readfirstlane(vq);
        const long long be = bn + QB < ge ? bn + QB : ge;
        if (!valid) {
          mine = bn + (rank - avail);
          valid = mine < be;
        }
        bcur = bn + (nwant - avail);
        bend = be;
        exhausted = bn >= ge;
        if (!exhausted && lane == 0) vq = atomicAdd(a.ctr + qg * 64, (unsigned)QB);
      } else {
        bcur += nwant;
      }
      const bool take = !active && valid;
      if (take) {
        cw = (uint32_t)mine;
        active = true;
        iter = 0;
        if (real) {
          int lro = lr;
          asm volatile("" : "+v"(lro));   // per-lane pointers formed here, not kept across the loop
          const uint8_t* g = reinterpret_cast<const uint8_t*>(a.llr_in) + (long long)cw * (SH::mc * ZC) + lro;
          uint32_t* wp = wbase + pos;
          // columns in two groups of mc / 2: the byte loads of a group in flight together (two
          // memory round trips per refill; all 2 * mc at once would spill beside the C2V and
          // address registers), then per column the pair (lo, hi) and its twin (hi, lo) as
          // int16 halves v << 8
          static_assert(SH::mc % 2 == 0, "two column groups");
          constexpr int CG = SH::mc / 2;
#pragma unroll
          for (int c0 = 0; c0 < SH::mc; c0 += CG) {
            uint32_t lo[CG], hi[CG];
#pragma unroll
            for (int k = 0; k < CG; ++k) {
              lo[k] = g[(c0 + k) * ZC];
              hi[k] = g[(c0 + k) * ZC + L];
            }
#pragma unroll
            for (int k = 0; k < CG; ++k) {
              wp[(c0 + k) * 128] = __builtin_amdgcn_perm(hi[k], lo[k], 0x040c000cu);
              wp[(c0 + k) * 128 + 64] = __builtin_amdgcn_perm(lo[k], hi[k], 0x040c000cu);
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      if (take) {   // fresh codeword: C2V messages start at zero (64-bit moves on its lanes)
#pragma unroll
        for (int k = 0; k < (NE + 1) / 2; ++k) asm volatile("v_mov_b64 %0, 0" : "=v"(rp[k]));
      }
    }
    if (__ballot(active) == 0ull) break;
    wave_sync();

    // ---- one layered iteration; lanes of idle slots compute on their own (dead) slot ----
#pragma unroll
    for (int k = 0; k < NC; ++k) asm volatile("" : "+v"(ga[k]), "+v"(gb[k]));   // edge addresses formed at each use
    int e0 = 0;
    uint32_t psyn[SH::mrow];   // per layer, XOR of the posteriors of its final edges
    const uint8_t* bt_it = btab + (iter < a.max_iter ? iter : a.max_iter - 1u) * SH::mrow;
#pragma unroll
    for (int l = 0; l < SH::mrow; ++l) {
      constexpr int MAXW = SH::maxw;
      const int w = SH::w[l];
      ushort2v beta2l = beta2;
      uint32_t amask = 0xffffffffu;
      if constexpr (BT) {   // the layer's offset b; an offset of 0 keeps abs(-128) = -128 (A = 0)
        const uint32_t b = bt_it[l];
        beta2l = as_us(b * 0x01000100u);
        amask = b ? 0xffffffffu : 0x7fff7fffu;
      }
      short2v q[MAXW];
      ushort2v A[MAXW];
      uint32_t S = 0;
      ushort2v m1 = ushort2v{0xffff, 0xffff}, m2 = ushort2v{0xffff, 0xffff};
#pragma unroll
      for (int j = 0; j < MAXW; ++j) {
        if (j < w) {
          const int e = e0 + j;
          const uint32_t ad = (SH::eflip[e] ? gb[SH::ecls[e]] : ga[SH::ecls[e]]) + SH::ecol[e] * 512;
          const uint32_t x = *lds32(ad);
          const short2v qs = __builtin_elementwise_sub_sat(as_s(x), as_s(rp[e >> 1][e & 1]));
          const short2v qq = as_s(as_u(qs) & 0xff00ff00u);   // Q << 8, Q = sat8(L - R)
          q[j] = qq;
          const ushort2v av = BT ? as_us(as_u(smax(qq, -qq)) & amask) : as_us(as_u(smax(qq, -qq)));
          A[j] = av;
          if (j & 1) S = xor3(S, as_u(q[j - 1]), as_u(qq));
          else if (j + 1 == w) S ^= as_u(qq);
          m2 = umin(m2, umax(m1, av));
          m1 = umin(m1, av);
        }
      }
      // capped exclusive minimum as in k_pair_i8: M_j = max(mv1, m1 + mv2 - |Q_j|)
      const short2v mv1 = as_s(as_u(umin(cap2, __builtin_elementwise_sub_sat(m1, beta2l))));
      const ushort2v mv2 = umin(cap2, __builtin_elementwise_sub_sat(m2, beta2l));
      uint32_t sumb_u = as_u(m1 + mv2);
      asm volatile("" : "+v"(sumb_u));   // one subtraction per edge, not re-associated into two
      const short2v sumb = as_s(sumb_u);
      const uint32_t sa = as_u(as_s(S) >> short2v{15, 15});   // 0 or 0xffff per half
      uint32_t gu[MAXW];
#pragma unroll
      for (int j = 0; j < MAXW; ++j) {
        if (j < w) {
          const short2v m = smax(mv1, sumb - as_s(as_u(A[j])));   // argmin ? min2 : min1, capped
          const uint32_t qa = as_u(q[j] >> short2v{15, 15});
          const uint32_t k = __builtin_amdgcn_bitop3_b32(sa, qa, 0x00010001u, 0xBE);   // (sa ^ qa) | 1
          const uint32_t rn = as_u(as_us(as_u(m)) * as_us(k));
          gu[j] = as_u(__builtin_elementwise_add_sat(q[j], as_s(rn)));   // L = sat8(Q + R)
          rp[(e0 + j) >> 1][(e0 + j) & 1] = rn;
        }
      }
      psyn[l] = 0u;
#pragma unroll
      for (int j = 0; j < MAXW; ++j)
        if (j < w && PairFinal<SH>::tab.e[e0 + j]) psyn[l] ^= gu[j];
#pragma unroll
      for (int j = 0; j < MAXW; ++j) {
        if (j < w) {
          const int e = e0 + j;
          const uint32_t c = SH::ecol[e] * 512;
          const uint32_t ad = (SH::eflip[e] ? gb[SH::ecls[e]] : ga[SH::ecls[e]]) + c;
          const uint32_t tw = (SH::eflip[e] ? ga[SH::ecls[e]] : gb[SH::ecls[e]]) + c;
          *lds32(ad) = gu[j];
          *lds32(tw) = __builtin_amdgcn_alignbit(gu[j], gu[j], 16);   // twin: halves swapped
        }
      }
      e0 += w;
      wave_sync();
      __builtin_amdgcn_sched_barrier(0);   // schedule layer by layer: bounded live ranges
    }
    // ---- syndrome after the full iteration, LDPCdecode.cpp:172-195 ----
#pragma unroll
    for (int k = 0; k < NC; ++k) asm volatile("" : "+v"(ga[k]), "+v"(gb[k]));
    uint32_t F = 0;
    e0 = 0;
#pragma unroll
    for (int l = 0; l < SH::mrow; ++l) {
      const int w = SH::w[l];
      uint32_t xs[SH::maxw];
      int n = 0;
#pragma unroll
      for (int j = 0; j < SH::maxw; ++j)
        if (j < w && !PairFinal<SH>::tab.e[e0 + j]) {
          const int e = e0 + j;
          xs[n++] = *lds32((SH::eflip[e] ? gb[SH::ecls[e]] : ga[SH::ecls[e]]) + SH::ecol[e] * 512);
        }
      if (PairFinal<SH>::tab.l[l]) xs[n++] = psyn[l];
      uint32_t par = xs[0];
#pragma unroll
      for (int j = 1; j + 1 < SH::maxw; j += 2)
        if (j + 1 < n) par = xor3(par, xs[j], xs[j + 1]);
      if (n & 1) F |= par;
      else F = __builtin_amdgcn_bitop3_b32(F, par, xs[n - 1], 0xF6);   // F | (par ^ x)
      e0 += w;
      __builtin_amdgcn_sched_barrier(0);
    }
    const bool fail = (F & 0x80008000u) != 0u;
    const unsigned long long fb = __ballot(real && active && fail);
    ++iter;
    unsigned long long smask = a.smask[0];
#pragma unroll
    for (int k = 1; k < 64 / L; ++k) smask = sl == k ? a.smask[k] : smask;
    const bool cwfail = (fb & smask) != 0ull;
    const bool finish = active && (!cwfail || iter == a.max_iter);   // LDPCdecode.cpp:192-197
    if (finish && real) {
      const uint32_t* wp = wbase + pos;
      int lro = lr;
      asm volatile("" : "+v"(lro));
      uint8_t* g = reinterpret_cast<uint8_t*>(a.post_out) + (long long)cw * (SH::mc * ZC) + lro;
      static_assert(SH::mc % 8 == 0, "column groups of eight");
#pragma unroll
      for (int c0 = 0; c0 < SH::mc; c0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = wp[(c0 + k) * 128];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const uint32_t x = v[k] >> 8;
          g[(c0 + k) * ZC] = (uint8_t)x;
          g[(c0 + k) * ZC + L] = (uint8_t)(x >> 16);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if (a.c2v) {
        int zo = ZC;
        asm volatile("" : "+v"(zo));
        int8_t* gc = a.c2v + (long long)cw * (NE * ZC) + lro;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
          gc[e * zo] = (int8_t)(rp[e >> 1][e & 1] >> 8);
          gc[e * zo + L] = (int8_t)(rp[e >> 1][e & 1] >> 24);
        }
      }
      if (lr == 0) {
        if (a.iters) a.iters[cw] = iter;
        my_iters += iter;
      }
    }
    active = active && !finish;
  }
  if (a.tot_iter) {   // my_iters is nonzero on the slots' row-0 lanes only
    long long s = (long long)my_iters;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0 && s) atomicAdd(a.tot_iter, (unsigned long long)s);
  }
}

// ---------------------------------------------------------------------------

bool pair_i8_table_ok(const DecoderDev& d, size_t n) {
  return d.pair_shape >= 0 && n <= 4096 && pair_shape_fixed_z(d.pair_shape, d.h.z, d.h.mcol);
}

int pair_i8_shape(const DecoderDev& d) {
  const HdecHost& h = d.h;
  if (d.elem != FECSIM_I8 || d.mode != FECSIM_MIN) return -1;
  if (h.z % 2 != 0 || h.z / 2 > 32 || h.z / 2 < 2) return -1;
  // one wave's columns (two rows of 64 dwords each) must fit the 64 KiB block LDS
  if (h.mcol * 128 * 4 > 64 * 1024) return -1;
  const float beta = d.params.beta_offset, mrm = d.params.max_row_metric;
  const int b8 = (int)(int8_t)(int)beta;
  if (mrm < 1e-6f) return -1;
  const int c8 = (int)(int8_t)(int)mrm;
  if (b8 < 1 || c8 < 0) return -1;
  return pair_shape_index(h.mrow, h.row_weights.data());
}

// clears k_pair_i8c's 64 queue counters (one per 256 bytes)
__global__ void k_pair_queue_reset(unsigned* c) {
  c[threadIdx.x * 64] = 0u;
}

// CT: the instance with the code's edge table compiled in (k_pair_i8c)
template <class SH, int ZC, bool BT, bool CT>
static hipError_t launch_shape_z(const PairArgs& in, hipStream_t st) {
  PairArgs a = in;
  if (a.num_cw == 0) return hipSuccess;
  void (*kfn)(const PairArgs) = nullptr;
  if constexpr (CT) {
    if (a.waves != 1) return hipErrorInvalidValue;
    kfn = k_pair_i8c<SH, BT>;
  } else {
    kfn = k_pair_i8<SH, ZC, BT>;
  }
  const size_t lds = (size_t)a.waves * a.wave_dwords * 4 + (BT ? ((size_t)a.max_iter * SH::mrow + 3) / 4 * 4 : 0);
  // grid of 8x the resident blocks, each wave working through a contiguous chunk of
  // codewords with per-slot refill: chunks of ~30 codewords per wave keep the spread of
  // iteration counts from leaving a long tail (measured on 11ad 1/2: 2x -> 8x is +5.6%,
  // 16x loses again)
  // resident blocks per device, measured once per device (mutex: host threads may
  // launch concurrently on distinct handles)
  static std::mutex mu;
  static int resident_by_dev[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  int resident;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!resident_by_dev[dev]) {
      // the twin addressing (byte address xor 256) needs the waves' columns at multiples
      // of 512 bytes, i.e. the dynamic LDS starting at 0: no static LDS in the kernel
      hipFuncAttributes fa;
      if (hipFuncGetAttributes(&fa, (const void*)kfn) != hipSuccess || fa.sharedSizeBytes != 0)
        return hipErrorInvalidDeviceFunction;
      int cus = 0, per_cu = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 64 * a.waves, lds) != hipSuccess)
        return hipErrorInvalidValue;
      resident_by_dev[dev] = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    }
    resident = resident_by_dev[dev];
  }
  const long long slots_per_block = (long long)a.waves * a.cw_per_wave;
  long long blocks = (a.num_cw + slots_per_block - 1) / slots_per_block;
  if constexpr (CT) {
    // persistent grid of the resident blocks; the counters (one of the decoder's sets, used
    // in rotation) are cleared by a kernel, not a memset, so that the pair can be captured
    // in a graph
    if (!a.ctr) return hipErrorInvalidValue;
    if (blocks > resident) blocks = resident;
    a.qgroups = blocks < 64 ? (int)blocks : 64;
    hipLaunchKernelGGL(k_pair_queue_reset, dim3(1), dim3(64), 0, st, a.ctr);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
  } else {
    const long long cap_blocks = a.max_blocks > 0 ? a.max_blocks : 8LL * resident;
    if (blocks > cap_blocks) blocks = cap_blocks;
  }
  const long long waves = blocks * a.waves;
  a.chunk = (a.num_cw + waves - 1) / waves;
  hipLaunchKernelGGL(kfn, dim3((unsigned)blocks), dim3(64 * a.waves), lds, st, a);
  return hipGetLastError();
}

template <class SH>
static hipError_t launch_shape(const PairArgs& a, hipStream_t st) {
  if constexpr (SH::zc != 0) {
    if (a.z == SH::zc && a.mcol == SH::mc && a.col_dwords == 128) {
      if constexpr (SH::ct) {
        if (a.ct)
          return a.beta_tab ? launch_shape_z<SH, SH::zc, true, true>(a, st)
                            : launch_shape_z<SH, SH::zc, false, true>(a, st);
      }
      return a.beta_tab ? launch_shape_z<SH, SH::zc, true, false>(a, st)
                        : launch_shape_z<SH, SH::zc, false, false>(a, st);
    }
  }
  if (a.beta_tab) return hipErrorInvalidValue;   // tables run on the fixed-Z instances only
  return launch_shape_z<SH, 0, false, false>(a, st);
}

// Host-side layout of the pair kernel for decoder `d` (tables in d.h).
void pair_layout(const HdecHost& h, PairLayout* pl) {
  pl->half = h.z / 2;
  pl->cw_per_wave = 64 / pl->half;
  pl->col_dwords = 128;
  pl->wave_dwords = h.mcol * 128;   // wave bases stay multiples of 512 bytes
  // one wave per block: a wave that finishes its chunk frees its SIMD slot at once
  // instead of holding it until the block's other waves finish (11ad 1/2: +1.3% over
  // four waves per block)
  pl->waves = 1;
  const int L = pl->half, Z = h.z, NE = h.total_hc_ones;
  pl->adrtab.assign((size_t)pl->cw_per_wave * L * NE, 0u);
  for (int slot = 0; slot < pl->cw_per_wave; ++slot)
    for (int lr = 0; lr < L; ++lr)
      for (int e = 0; e < NE; ++e) {
        const int c = h.row_active_cols[e], s = h.row_offset[e];
        const int t = (s + lr) % Z;
        const int d = c * 128 + (t >= L ? 64 : 0) + (t % L) + slot * L;
        pl->adrtab[((size_t)slot * L + lr) * NE + e] = (uint32_t)(d * 4);
      }
  const int shape = pair_shape_index(h.mrow, h.row_weights.data());
  pl->ct = shape >= 0 && pair_shape_fixed_z(shape, h.z, h.mcol) &&
           pair_shape_table(shape, NE, h.row_active_cols.data(), h.row_offset.data());
  pl->ct_off = pl->adrtab.size();
  if (!pl->ct) return;
  // k_pair_i8c's per-lane table, behind the run-time table: per hardware lane, the
  // addresses of the shift classes (distinct s mod Z/2 in order of first use, as the shape
  // header lists them) in row [c + lr >= Z/2], the pair position and lr | slot << 8 |
  // real << 16; per slot its row-0 lane and the 64-bit mask of its lanes (kernel arguments)
  std::vector<int> cls;
  for (int e = 0; e < NE; ++e) {
    const int c = h.row_offset[e] % L;
    bool seen = false;
    for (int x : cls) seen = seen || x == c;
    if (!seen) cls.push_back(c);
  }
  const int nc = (int)cls.size(), cpw = pl->cw_per_wave;
  // lane -> (slot, lr) item and (slot, position) -> dword: the shape's bank-conflict
  // layout when it has one, else slots in lane order and positions in slot order
  int32_t lane_item[64], dword[64];
  if (!pair_shape_lds_layout(shape, lane_item, dword)) {
    for (int x = 0; x < 64; ++x) lane_item[x] = x < cpw * L ? x : -1;
    for (int x = 0; x < cpw * L; ++x) dword[x] = x;
  }
  // a spare lane mirrors the first real lane of its own half-wave
  int item_of[64];
  bool real[64];
  for (int x = 0; x < 64; ++x) {
    real[x] = lane_item[x] >= 0;
    item_of[x] = lane_item[x];
    if (!real[x])
      for (int y = x & 32; y < (x & 32) + 32; ++y)
        if (lane_item[y] >= 0) {
          item_of[x] = lane_item[y];
          break;
        }
  }
  int lead_of_slot[64] = {};
  unsigned long long mask_of_slot[64] = {};
  for (int x = 0; x < 64; ++x)
    if (real[x]) {
      const int sl = item_of[x] / L;
      if (item_of[x] % L == 0) lead_of_slot[sl] = x;
      mask_of_slot[sl] |= 1ull << x;
    }
  const size_t stride = (size_t)nc + 2;
  pl->adrtab.resize(pl->ct_off + 64 * stride);
  uint32_t* t = pl->adrtab.data() + pl->ct_off;
  for (int x = 0; x < 64; ++x) {
    const int sl = item_of[x] / L, lr = item_of[x] % L;
    for (int k = 0; k < nc; ++k) {
      const int t0 = cls[k] + lr, o = t0 >= L;
      t[x * stride + k] = (uint32_t)(o * 256 + 4 * dword[sl * L + t0 - o * L]);
    }
    t[x * stride + nc] = (uint32_t)dword[sl * L + lr];
    t[x * stride + nc + 1] = (uint32_t)lr | ((uint32_t)sl << 8) | ((real[x] ? 1u : 0u) << 16);
  }
  for (int sl = 0; sl < 4; ++sl) {
    pl->lead[sl] = sl < cpw ? lead_of_slot[sl] : 0;
    pl->smask[sl] = sl < cpw ? mask_of_slot[sl] : 0ull;
  }
}

hipError_t launch_pair_i8(const DecoderDev& d, const KernelArgs& ka, hipStream_t st, bool run_time_table) {
  const HdecHost& h = d.h;
  const PairLayout& pl = d.pair;
  PairArgs a{};
  a.ct = pl.ct && !run_time_table ? 1 : 0;
  if (a.ct) {   // the next counter set: up to kPairQueueSets decodes of one decoder may overlap
    const unsigned k = __atomic_fetch_add(&d.pair_queue_next, 1u, __ATOMIC_RELAXED) % kPairQueueSets;
    a.ctr = d.d_pair_queue + (size_t)k * (kPairQueueSetBytes / sizeof(unsigned));
  }
  for (int sl = 0; sl < 4; ++sl) {
    a.lead[sl] = pl.lead[sl];
    a.smask[sl] = pl.smask[sl];
  }
  a.adrtab = d.d_pair_adr + (a.ct ? pl.ct_off : 0);
  a.z = h.z;
  a.half = pl.half;
  a.mcol = h.mcol;
  a.cw_per_wave = pl.cw_per_wave;
  a.col_dwords = pl.col_dwords;
  a.wave_dwords = pl.wave_dwords;
  a.waves = pl.waves;
  a.num_cw = ka.num_cw;
  a.max_iter = d.params.max_iter;
  a.beta = (int)(int8_t)(int)d.params.beta_offset;
  a.cap = (int)(int8_t)(int)d.params.max_row_metric;
  a.beta_tab = d.d_beta_tab8;
  a.llr_in = reinterpret_cast<const int8_t*>(ka.metrics_in);
  a.post_out = reinterpret_cast<int8_t*>(ka.metrics_out);
  a.iters = ka.iters;
  a.c2v = reinterpret_cast<int8_t*>(ka.c2v);
  a.tot_iter = ka.tot_iter;
  return pair_dispatch(d.pair_shape, a, st);
}

#define FECSIM_PAIR_LAUNCH(SH) return launch_shape<SH>(a, st);
#include "pair_shapes_dispatch.h"

}  // namespace fecsim

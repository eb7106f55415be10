// Float32 layered decoder for compiled code shapes, with the C2V messages in VGPRs.
//
// The generic float kernels keep a codeword's posteriors and C2V messages in LDS, so a
// CU holds few codewords (16e 1/2 at Z = 96: 38 KB each, 4 per CU) and many lanes sit
// idle (Z rows on a multiple of 64 lanes).  Here a workgroup of NW waves (1, 2, 4) holds
// CPB = floor(64 NW / Z) codewords; thread t owns row t mod Z of every layer of codeword
// slot t / Z and keeps that row's C2V values (one per edge of its layers, `ne` in all)
// in registers through every iteration, so LDS holds only the posteriors (N floats per
// codeword).  The layer loop is unrolled over the shape's row weights (compile time),
// which makes every C2V index static.
//
// Per layer: each active thread runs its row (the arithmetic of row_float_w, operation
// for operation, so posteriors, C2V and iteration counts equal the generic kernels' bit
// for bit), then a block barrier.  Per iteration: the rows' parities are OR-ed per slot
// through LDS flags (two sets, alternating by iteration, so one barrier suffices), and a
// slot whose codeword passes the syndrome check or reaches max_iter writes it out and
// loads its next codeword (slot s of block b decodes b*CPB + s, then strides by
// gridDim.x*CPB), so finished codewords do not leave lanes idle while the others of
// the block continue.  The block leaves when no slot has a codeword left.
//
// Reference semantics: LDPCdecode.cpp:115-197 (layered loop, syndrome exit) and
// decode_functions.cpp:17-38 (the float row), as the generic float kernels.
#include <algorithm>
#include <mutex>
#include <vector>

#include "decode_core.h"
#include "pair_shapes.h"

namespace fecsim {

namespace {

// one check row with its W C2V values in c[0..W-1] (row_float_w with the C2V array
// replaced by registers)
template <class P, int W>
__device__ __forceinline__ void row_reg(typename P::S* __restrict__ post, typename P::S* c, ConstU32* ge, int r,
                                       int Z, typename P::S beta, typename P::S cap) {
  using F = typename P::S;
  F* vp[W];
  F* const pr = post + r;
  F* const prz = post + r - Z;
  const int zr = Z - r;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t ed = ge[j];
    const int sh = (int)(ed & 0xffffu), cz = (int)(ed >> 16);
    vp[j] = (sh >= zr ? prz : pr) + (cz + sh);
  }
  F q[W], A[W], M[W];
  int sg = 0;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const F qq = *vp[j] - c[j];
    q[j] = qq;
    const F a = fabs(qq) - beta;
    A[j] = a > F(0) ? a : F(0);
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
    c[j] = rn;
  }
}

// Codeword transfer for thread r: row r of every column (global column stride gs).  The
// strides pass through an empty asm first, so their address arithmetic stays inside the
// (rarely taken) branch that moves a codeword instead of being hoisted out of the
// decode loop into VGPRs that would stay live beside the C2V registers.
template <class F>
__device__ __forceinline__ void load_row(F* __restrict__ post, const F* __restrict__ g, long long gs, int mcol,
                                         int Z, int r) {
  asm volatile("" : "+s"(gs));
  const F* src = g + r;
  F* dst = post + r;
  for (int c0 = 0; c0 < mcol; c0 += 8) {
    F t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = src[(long long)min(c0 + u, mcol - 1) * gs];   // unconditional: one wait
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c0 + u < mcol) dst[(c0 + u) * Z] = t[u];
  }
}

template <class F>
__device__ __forceinline__ void store_row(F* __restrict__ g, const F* __restrict__ post, long long gs, int mcol,
                                          int Z, int r) {
  asm volatile("" : "+s"(gs));
  F* dst = g + r;
  const F* src = post + r;
  for (int c0 = 0; c0 < mcol; c0 += 8) {
    F t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = src[min(c0 + u, mcol - 1) * Z];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c0 + u < mcol) dst[(long long)(c0 + u) * gs] = t[u];
  }
}

// layers L.. of one iteration; E0: the first edge of layer L.  `act`: the thread's slot
// has a codeword; every thread reaches every barrier.
template <class P, class SH, int L, int E0>
__device__ __forceinline__ void layers(typename P::S* post, typename P::S* c, const uint32_t* edges, int r, int Z,
                                      bool act, typename P::S fbeta, const double* btab, int bidx,
                                      typename P::S cap) {
  using F = typename P::S;
  if constexpr (L < SH::mrow) {
    constexpr int W = SH::w[L];
    if (act) {
      const F fb = btab ? (F)btab[bidx + L] : fbeta;   // per-(iteration, layer) offsets
      row_reg<P, W>(post, c + E0, uniform_table(edges + E0), r, Z, fb, cap);
    }
    __syncthreads();
    layers<P, SH, L + 1, E0 + W>(post, c, edges, r, Z, act, fbeta, btab, bidx, cap);
  }
}

template <class P, class SH, int L, int E0>
__device__ __forceinline__ bool parity(const typename P::S* post, const uint32_t* edges, int r, int Z) {
  if constexpr (L == SH::mrow) {
    return false;
  } else {
    const bool p = row_parity_float_w<P, SH::w[L]>(post, edges, E0, r, Z);
    return p | parity<P, SH, L + 1, E0 + SH::w[L]>(post, edges, r, Z);
  }
}

// shapes of up to 76 edges (802.16e 1/2) bounded to 128 VGPRs, 4 waves per SIMD: unbounded
// they take 133-136 and run 3 (a few registers spill; the 16e 1/2 sweep decodes 7% faster)
template <class SH>
constexpr int float_reg_min_waves() {
  return SH::ne <= 76 ? 4 : 1;
}

// core.cw_per_wave carries CPB (codewords per block) and core.post_stride the padded N
template <class P, class SH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(float_reg_min_waves<SH>())))
k_float_reg(const KernelArgs a) {
  using F = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const DecCore& d = a.core;
  const int Z = d.z, tid = threadIdx.x, cpb = d.cw_per_wave;
  const int slot = tid / Z;
  const int r = tid - slot * Z;
  const bool has_slot = slot < cpb;
  F* post = reinterpret_cast<F*>(smem) + (size_t)(has_slot ? slot : 0) * d.post_stride;
  int* flags = reinterpret_cast<int*>(reinterpret_cast<F*>(smem) + (size_t)cpb * d.post_stride);   // [2][cpb]
  const F* gin = reinterpret_cast<const F*>(a.metrics_in);
  F* gout = reinterpret_cast<F*>(a.metrics_out);
  const F fbeta = (F)d.fbeta, fcap = (F)d.fcap;
  const double* btab = reinterpret_cast<const double*>(d.beta_tab);
  const long long stride = (long long)gridDim.x * cpb;
  long long cw = (long long)blockIdx.x * cpb + slot;
  bool act = has_slot && cw < (long long)a.num_cw;
  F c[SH::ne];
#pragma unroll
  for (int i = 0; i < SH::ne; ++i) c[i] = F(0);   // LDPCdecode.cpp:115-119
  if (tid < 2 * cpb) flags[tid] = 0;
  if (act) load_row(post, gin + (a.base + cw * a.cw_stride), a.col_stride, d.mcol, Z, r);
  uint32_t iter = 0;
  unsigned long long tot = 0;
  int par = 0;
  bool any = __syncthreads_or(act);
  while (any) {
    // the edge table's address made opaque per iteration: the per-edge LDS addresses are
    // then recomputed from scalar-loaded edge words in each iteration, not hoisted out of
    // the loop into one VGPR per edge beside the C2V registers
    const uint32_t* edges = d.edges;
    asm volatile("" : "+s"(edges));
    layers<P, SH, 0, 0>(post, c, edges, r, Z, act, fbeta, btab, (int)iter * SH::mrow, fcap);
    // LDPCdecode.cpp:172-195: the syndrome of the slot's codeword
    if (act && parity<P, SH, 0, 0>(post, edges, r, Z)) flags[par * cpb + slot] = 1;
    __syncthreads();
    const bool cwfail = act && flags[par * cpb + slot] != 0;
    if (has_slot && r == 0) flags[(par ^ 1) * cpb + slot] = 0;   // read again two iterations on
    par ^= 1;
    if (act) {
      ++iter;
      if (!cwfail || iter == d.max_iter) {   // LDPCdecode.cpp:192-197
        // thread r writes and then reads row r of every column only: no barrier needed
        // between this codeword's write-out and the next one's load
        store_row(gout + (a.base + cw * a.cw_stride), post, a.col_stride, d.mcol, Z, r);
        if (a.c2v) {
          int zs = Z;
          asm volatile("" : "+s"(zs));
          F* gc = reinterpret_cast<F*>(a.c2v) + cw * (long long)d.e + r;
#pragma unroll
          for (int i = 0; i < SH::ne; ++i) {
            *gc = c[i];
            gc += zs;
          }
        }
        if (r == 0) {
          if (a.iters) a.iters[cw] = iter;
          tot += iter;
        }
        cw += stride;
        act = cw < (long long)a.num_cw;
        iter = 0;
        if (act) {
          load_row(post, gin + (a.base + cw * a.cw_stride), a.col_stride, d.mcol, Z, r);
#pragma unroll
          for (int i = 0; i < SH::ne; ++i) c[i] = F(0);
        }
      }
    }
    any = __syncthreads_or(act);
  }
  if (a.tot_iter && r == 0 && has_slot && tot) atomicAdd(a.tot_iter, tot);
}

#define FECSIM_FREG_LAUNCH(P, SH) return launch_reg<P, SH>(a, st);

struct RegPlan {
  int nw = 0, cpb = 0, post_stride = 0;
  size_t lds = 0;
};

// Waves per block: 1, 2 or 4, whichever decodes the most codewords per wave (ties: fewer
// waves).  The kernel is VALU-bound with every wave issuing its rows' instructions
// whatever number of its lanes own a row, so throughput follows codewords per resident
// wave (16e 1/2, decode only: 17.5 M cw/s at one codeword per wave, 12.2 M at 3 per
// 4-wave block, 8.4 M at one per 2-wave block).  3-wave blocks pack more rows into
// their lanes but a CU then holds 3 of them, not 4: measured 16.3 vs 20.3 M cw/s at
// Z = 44 (3 waves x 4 codewords vs 4 x 5) and 36.6 vs 38.7 M at Z = 24.
RegPlan plan_for(int Z, int N) {
  RegPlan p;
  double best = 0.0;
  for (int nw = 1; nw <= 4; nw *= 2) {
    const int cpb = 64 * nw / Z;
    if (cpb < 1) continue;
    const double u = (double)cpb / nw;
    if (u > best + 1e-9) {
      best = u;
      p.nw = nw;
      p.cpb = cpb;
    }
  }
  p.post_stride = (N + 3) / 4 * 4;
  p.lds = (size_t)p.cpb * p.post_stride * 4 + (size_t)2 * p.cpb * 4;
  return p;
}

template <class P, class SH>
hipError_t launch_reg(const KernelArgs& in, hipStream_t st) {
  const int Z = in.core.z, N = in.core.mcol * Z;
  const RegPlan p = plan_for(Z, N);
  if (p.cpb < 1 || p.lds > 65536) return hipErrorInvalidValue;
  if (in.num_cw == 0) return hipSuccess;
  // resident blocks per device and plan (LDS and VGPRs bound them), measured once per
  // device: host threads may launch concurrently on distinct handles
  static std::mutex mu;
  static std::vector<std::vector<std::pair<int, int>>> resident;   // [device] -> (Z, blocks)
  int dev = 0, blocks = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  {
    std::lock_guard<std::mutex> lk(mu);
    if ((int)resident.size() <= dev) resident.resize(dev + 1);
    for (const auto& zb : resident[dev])
      if (zb.first == Z) blocks = zb.second;
    if (!blocks) {
      int per_cu = 0, cus = 0;
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_float_reg<P, SH>, 64 * p.nw, p.lds);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (e != hipSuccess) return e;
      blocks = std::max(1, per_cu) * std::max(1, cus);
      resident[dev].push_back({Z, blocks});
    }
  }
  KernelArgs a = in;
  a.core.cw_per_wave = p.cpb;
  a.core.post_stride = p.post_stride;
  a.core.waves = p.nw;
  const long long need = ((long long)a.num_cw + p.cpb - 1) / p.cpb;
  const unsigned grid = (unsigned)std::min<long long>(need, blocks);
  hipLaunchKernelGGL((k_float_reg<P, SH>), dim3(grid), dim3(64 * p.nw), p.lds, st, a);
  return hipGetLastError();
}

#include "float_shapes_dispatch.h"

}  // namespace

int float_reg_shape(const DecoderDev& d) {
  const HdecHost& h = d.h;
  if (d.elem != FECSIM_F32) return -1;
  if (h.z < 2 || h.z > 256) return -1;
  const int s = pair_shape_index(h.mrow, h.row_weights.data());
  if (s < 0 || !float_shape_ok(s)) return -1;
  if (plan_for(h.z, h.mcol * h.z).lds > 65536) return -1;
  return s;
}

hipError_t launch_float_reg(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  if (d.freg_shape < 0) return hipErrorInvalidValue;
  return d.mode == FECSIM_PROD ? float_dispatch<FloatPolicy<float, true>>(d.freg_shape, a, st)
                               : float_dispatch<FloatPolicy<float, false>>(d.freg_shape, a, st);
}

}  // namespace fecsim

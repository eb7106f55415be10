// Mixed-code batched decoder: one persistent launch decodes a queue of codewords of
// different codes (different Z, base graphs, iteration limits).  No counterpart in the
// reference (each of its processes decodes one code); per codeword the result is
// exactly that code's single-code decoder (LDPCdecode.cpp:79-201 semantics via
// decode_loop).
//
//   k_multi_hist     histogram of the queue's code ids
//   k_multi_scan     prefix sums: codeword start and work-item start per code; resets
//                    the work counter and the scatter cursors (graph-safe: no memset)
//   k_multi_scatter  groups the queue by code (perm[])
//   k_multi_decode   persistent: each wave takes work items (cw_per_wave codewords of
//                    one code) from an atomic counter until the queue is drained
//   k_multi_code     persistent, for queues whose codes are all compiled in (int8 offset
//                    min-sum): each workgroup decodes one code at a time with that code's
//                    compiled layers, several codewords at once, refilling each slot from
//                    the code's segment of the grouped queue
#include <hip/hip_runtime.h>

#include <type_traits>

#include "code_core.h"
#include "code_tables.h"
#include "decode_core.h"
#include "internal.h"

namespace fecsim {

namespace {

// Block-private counts in LDS, then one global atomic per code and block (the queue has
// few codes, so per-entry global atomics would all contend on a handful of words).
__global__ void __launch_bounds__(256) k_multi_hist(MultiArgs a) {
  __shared__ uint32_t cnt[64];
  if (threadIdx.x < 64) cnt[threadIdx.x] = 0u;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.num_cw; i += gridDim.x * blockDim.x) {
    const int c = a.code[i];
    if (c >= 0 && c < a.ncode) atomicAdd(&cnt[c], 1u);   // unknown code ids are skipped
  }
  __syncthreads();
  if ((int)threadIdx.x < a.ncode && cnt[threadIdx.x]) atomicAdd(&a.cnt[threadIdx.x], cnt[threadIdx.x]);
}

__global__ void k_multi_scan(MultiArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t cs = 0, is = 0;
  for (int c = 0; c < a.ncode; ++c) {
    a.code_start[c] = cs;
    a.item_start[c] = is;
    const uint32_t n = a.cnt[c];
    const uint32_t cpw = (uint32_t)a.cores[c].cw_per_wave;
    cs += n;
    is += (n + cpw - 1) / cpw;
    a.cursor[c] = 0;
    a.take[c] = 0;
    a.cnt[c] = 0;            // ready for the next call
  }
  a.code_start[a.ncode] = cs;
  a.item_start[a.ncode] = is;
  *a.counter = 0;
}

// Per tile of 256 entries: ranks within the tile from LDS atomics, one global atomic per
// code reserves the tile's range of that code's segment.  The order inside a segment is
// arbitrary; every codeword decodes the same wherever it sits.
__global__ void __launch_bounds__(256) k_multi_scatter(MultiArgs a) {
  __shared__ uint32_t cnt[64], base[64];
  for (uint32_t t0 = blockIdx.x * blockDim.x; t0 < a.num_cw; t0 += gridDim.x * blockDim.x) {   // uniform per block
    if (threadIdx.x < 64) cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t i = t0 + threadIdx.x;
    const int c = i < a.num_cw ? a.code[i] : -1;
    const bool ok = c >= 0 && c < a.ncode;
    const uint32_t rank = ok ? atomicAdd(&cnt[c], 1u) : 0u;
    __syncthreads();
    if ((int)threadIdx.x < a.ncode && cnt[threadIdx.x])
      base[threadIdx.x] = atomicAdd(&a.cursor[threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (ok) a.perm[a.code_start[c] + base[c] + rank] = i;
    __syncthreads();
  }
}

// int8 rows: at most 168 VGPRs, 3 waves per SIMD (the weight-24 instance takes 197 and
// 2 waves unbounded; bounded, its heaviest rows spill a few registers and the mixed
// queue of config 3 decodes 10% faster)
template <class P>
constexpr int multi_min_waves() {
  return std::is_same<P, IntPolicy<8>>::value ? 3 : 1;
}

template <class P, int MAXW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(multi_min_waves<P>())))
k_multi_decode(const MultiArgs a) {
  using S = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t* tabs = reinterpret_cast<uint32_t*>(smem);
  for (int i = threadIdx.x; i < a.tab_words; i += blockDim.x) tabs[i] = a.tabs[i];
  __syncthreads();
  char* wbase = smem + a.tab_bytes + (size_t)wave * a.wave_bytes;
  const S* gin = reinterpret_cast<const S*>(a.in);
  S* gout = reinterpret_cast<S*>(a.out);
  const uint32_t nitems = a.item_start[a.ncode];
  unsigned long long tot = 0;
  for (;;) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(a.counter, 1u);
    item = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)item, 0, 64));
    if (item >= nitems) break;
    int c = 0;
    while (item >= a.item_start[c + 1]) ++c;
    const DecCore d = a.cores[c];
    const int L = d.lanes_per_cw, CPW = d.cw_per_wave;
    const int slot = lane / L, lr = lane - slot * L;
    const int sl = slot < CPW ? slot : 0;
    const uint32_t first = a.code_start[c] + (item - a.item_start[c]) * (uint32_t)CPW;
    const uint32_t end = a.code_start[c + 1];
    const bool valid = slot < CPW && first + (uint32_t)slot < end;
    const uint32_t cw = valid ? a.perm[first + slot] : 0u;
    S* post = reinterpret_cast<S*>(wbase) + (size_t)sl * d.post_stride;
    S* c2v = reinterpret_cast<S*>(wbase) + (size_t)CPW * d.post_stride + (size_t)sl * d.c2v_stride;
    const unsigned long long base = valid ? a.off[cw] : 0ull;
    if (valid)
      cols_to_lds(post, gin, (long long)base, d.z, d.mcol, d.z, lr, L);
    const uint32_t iter = decode_loop<P, MAXW>(d, tabs + d.tab_off, post, c2v, slot, lr, valid);
    if (valid) {
      cols_from_lds(gout, post, (long long)base, d.z, d.mcol, d.z, lr, L);
      if (lr == 0 && a.iters) a.iters[cw] = iter;
    }
    if (a.tot_iter) {
      for (int q = 0; q < CPW; ++q) {
        const unsigned it = (unsigned)__shfl((int)iter, q * L, 64);
        if (first + (uint32_t)q < end) tot += it;
      }
    }
    wave_sync();
  }
  if (a.tot_iter && lane == 0 && tot) atomicAdd(a.tot_iter, tot);
}

// ---------------------------------------------------------------------------
// k_multi_code: the queue's codes compiled in (code_tables.h MultiTab*)
// ---------------------------------------------------------------------------
//
// A workgroup of 256 threads decodes one code at a time.  For code CT it holds
// CPB = floor(256 / Z) codeword slots; thread t owns check row t mod Z of every layer of
// slot t / Z, keeps that row's C2V bytes in registers and runs code_core.h's layers
// (LDS addresses are a base register plus compiled-in offsets; the same arithmetic as
// k_code_i8, so the same results as the generic kernels).  Every slot stops on its own
// syndrome and max_iter: it writes its codeword out and loads the next one of its code at
// once, so a codeword that stops early does not leave lanes idle while the others of the
// workgroup continue.  A slot's next codeword index is taken from the code's counter
// (take[c]) one codeword ahead and parked in LDS (two entries per slot, alternating), so
// the slot's threads read it without another barrier.  When a code's segment is drained
// and its slots are empty the workgroup moves on to the next code with codewords left;
// it leaves when there is none.

constexpr int kMcThreads = 256;
constexpr uint32_t kNoCw = 0xffffffffu;

template <class CT>
struct McLayout {
  static constexpr int cpb = kMcThreads / CT::z;   // codeword slots per workgroup
  static constexpr int slot_bytes = ((3 * CT::mcol + 1) * CT::z + 15) / 16 * 16;
  static constexpr int flags_off = cpb * slot_bytes;   // int [2][cpb]: syndrome flags by iteration parity
  static constexpr int next_off = flags_off + 8 * cpb;   // uint32 [2][cpb]: next codeword per slot
  static constexpr int lds = next_off + 8 * cpb;
  static_assert(cpb >= 1, "Z above 256");
  static_assert(slot_bytes < 32768, "LDS offsets beyond the instruction's immediate field");
};

template <int I>
constexpr int mc_lds_max() {
  if constexpr (I == kMultiTabs) {
    return 0;
  } else {
    constexpr int a = McLayout<typename MultiTabAt<I>::T>::lds, b = mc_lds_max<I + 1>();
    return a > b ? a : b;
  }
}

// row r of every column of a codeword into both LDS copies (int8: the two-edge form of the
// row arithmetic), and back
template <class CT>
__device__ __forceinline__ void mc_load(uint32_t vb, const int8_t* __restrict__ g) {
  constexpr int Z = CT::z, MC = CT::mcol;
  uint32_t v[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) v[c] = (uint8_t)g[c * Z];
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    *lds8(vb + (uint32_t)(Z + 3 * Z * c)) = (uint8_t)v[c];
    *lds8(vb + (uint32_t)(2 * Z + 3 * Z * c)) = (uint8_t)v[c];
  }
}

template <class CT>
__device__ __forceinline__ void mc_store(int8_t* __restrict__ g, uint32_t vb) {
  constexpr int Z = CT::z, MC = CT::mcol;
  uint32_t v[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) v[c] = *lds8(vb + (uint32_t)(Z + 3 * Z * c));
#pragma unroll
  for (int c = 0; c < MC; ++c) g[c * Z] = (int8_t)(uint8_t)v[c];
}

// decode code c's segment of the grouped queue until it is drained; every thread of the
// workgroup calls it (c is uniform) and leaves it at the same barrier
template <class CT>
__device__ __forceinline__ void mc_run(const MultiArgs& a, int c, uint8_t* smem, unsigned long long& tot) {
  using LY = McLayout<CT>;
  constexpr int Z = CT::z, CPB = LY::cpb, NW = (CT::ne + 3) / 4;
  const int tid = threadIdx.x;
  const int slot = tid / Z, r = tid - slot * Z;
  const bool has_slot = slot < CPB;
  const int sl = has_slot ? slot : 0;
  int* flags = reinterpret_cast<int*>(smem + LY::flags_off);
  uint32_t* nxt = reinterpret_cast<uint32_t*>(smem + LY::next_off);
  const uint32_t cs = a.code_start[c], count = a.code_start[c + 1] - cs;
  const DecCore& d = a.cores[c];
  const uint32_t max_iter = d.max_iter, beta = (uint32_t)d.ibeta;
  const int cap = (int)d.icap;
  const long long* btab = reinterpret_cast<const long long*>(d.beta_tab);
  uint32_t* take = a.take + c;
  const int8_t* gin = reinterpret_cast<const int8_t*>(a.in);
  int8_t* gout = reinterpret_cast<int8_t*>(a.out);
  // each slot's first two codewords (queue positions past the segment: none)
  if (has_slot && r == 0) {
    const uint32_t i0 = atomicAdd(take, 2u);
    nxt[sl] = i0 < count ? a.perm[cs + i0] : kNoCw;
    nxt[CPB + sl] = i0 + 1 < count ? a.perm[cs + i0 + 1] : kNoCw;
    flags[sl] = 0;
    flags[CPB + sl] = 0;
  }
  __syncthreads();
  uint32_t cw = has_slot ? nxt[sl] : kNoCw;
  uint32_t gen = 1;   // nxt[gen & 1] holds the slot's next codeword
  bool act = cw != kNoCw;
  uint32_t vb[1] = {(uint32_t)(size_t)(LdsU8*)smem + (uint32_t)(sl * LY::slot_bytes + r)};
  unsigned long long base = act ? a.off[cw] : 0ull;
  if (act) mc_load<CT>(vb[0], gin + base + r);
  uint32_t c2v[NW];   // LDPCdecode.cpp:115-119
#pragma unroll
  for (int k = 0; k < NW; ++k) c2v[k] = 0u;
  uint32_t iter = 0;
  int par = 0;
  bool any = __syncthreads_or(act);
  while (any) {
    const long long* bt = btab ? btab + (size_t)iter * CT::mrow : nullptr;
    int pin0[CT::w[0]] = {};
    code_layers<CT, 0, 0, true>(vb, c2v, bt, beta, cap, pin0, act);   // ends at a barrier (two edges per op)
    // LDPCdecode.cpp:172-195: the syndrome of the slot's codeword
    if (act && (code_syndrome<CT, 0, 0, true>(vb) & 0x80u)) flags[par * CPB + sl] = 1;
    __syncthreads();
    const bool cwfail = act && flags[par * CPB + sl] != 0;
    if (has_slot && r == 0) flags[(par ^ 1) * CPB + sl] = 0;   // read again two iterations on
    par ^= 1;
    if (act) {
      ++iter;
      if (!cwfail || iter == max_iter) {   // LDPCdecode.cpp:192-197
        // thread r reads and then writes variables (c, r) only: no barrier between this
        // codeword's write-out and the next one's load
        mc_store<CT>(gout + base + r, vb[0]);
        if (r == 0) {
          if (a.iters) a.iters[cw] = iter;
          tot += iter;
        }
        cw = nxt[(gen & 1) * CPB + sl];
        if (r == 0 && cw != kNoCw) {
          const uint32_t i = atomicAdd(take, 1u);
          nxt[((gen + 1) & 1) * CPB + sl] = i < count ? a.perm[cs + i] : kNoCw;
        }
        ++gen;
        act = cw != kNoCw;
        iter = 0;
        if (act) {
          base = a.off[cw];
          mc_load<CT>(vb[0], gin + base + r);
#pragma unroll
          for (int k = 0; k < NW; ++k) c2v[k] = 0u;
        }
      }
    }
    any = __syncthreads_or(act);
  }
}

template <int I>
__device__ __forceinline__ void mc_dispatch(int t, const MultiArgs& a, int c, uint8_t* smem, unsigned long long& tot) {
  if constexpr (I < kMultiTabs) {
    if (t == I) mc_run<typename MultiTabAt<I>::T>(a, c, smem, tot);
    else mc_dispatch<I + 1>(t, a, c, smem, tot);
  }
}

__global__ void __launch_bounds__(kMcThreads) __attribute__((amdgpu_waves_per_eu(4))) k_multi_code(const MultiArgs a) {
  HIP_DYNAMIC_SHARED(char, lds);
  uint8_t* smem = reinterpret_cast<uint8_t*>(lds);
  __shared__ int pick;
  unsigned long long tot = 0;
  int c = (int)(blockIdx.x % (unsigned)a.ncode);
  for (;;) {
    // the next code with codewords not yet taken, from this workgroup's current one on
    if (threadIdx.x == 0) {
      int found = -1;
      for (int k = 0; k < a.ncode && found < 0; ++k) {
        const int cc = c + k < a.ncode ? c + k : c + k - a.ncode;
        const uint32_t n = a.code_start[cc + 1] - a.code_start[cc];
        if (__hip_atomic_load(a.take + cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n) found = cc;
      }
      pick = found;
    }
    __syncthreads();
    c = __builtin_amdgcn_readfirstlane(pick);
    if (c < 0) break;
    // mc_run starts with a barrier, so `pick` is not rewritten before every thread read it
    mc_dispatch<0>(__builtin_amdgcn_readfirstlane(a.ctab[c]), a, c, smem, tot);
  }
  if (a.tot_iter && tot) atomicAdd(a.tot_iter, tot);
}

template <class P>
const void* multi_kernel(int maxw) {
  switch (maxw) {
  case 8: return (const void*)k_multi_decode<P, 8>;
  case 16: return (const void*)k_multi_decode<P, 16>;
  case 24: return (const void*)k_multi_decode<P, 24>;
  default: return (const void*)k_multi_decode<P, 32>;
  }
}

const void* multi_kernel_for(int elem, int mode, int maxw) {
  switch (elem) {
  case FECSIM_I8: return multi_kernel<IntPolicy<8>>(maxw);
  case FECSIM_I16: return multi_kernel<IntPolicy<16>>(maxw);
  case FECSIM_I32: return multi_kernel<IntPolicy<32>>(maxw);
  case FECSIM_F32:
    return mode == FECSIM_PROD ? multi_kernel<FloatPolicy<float, true>>(maxw) : multi_kernel<FloatPolicy<float, false>>(maxw);
  default:
    return mode == FECSIM_PROD ? multi_kernel<FloatPolicy<double, true>>(maxw)
                               : multi_kernel<FloatPolicy<double, false>>(maxw);
  }
}

}  // namespace

// the compiled code of the queue kernel k_multi_code that runs decoder d's codewords, or -1
// (int8 offset min-sum with 1 <= beta, 0 <= cap, any offset table with entries 1..127)
int multi_code_index(const DecoderDev& d) {
  const HdecHost& h = d.h;
  if (d.elem != FECSIM_I8 || d.mode != FECSIM_MIN) return -1;
  const float beta = d.params.beta_offset, mrm = d.params.max_row_metric;
  if (mrm < 1e-6f) return -1;   // cap 1 - max_row_metric: generic kernel
  const int b8 = (int)(int8_t)(int)beta, c8 = (int)(int8_t)(int)mrm;
  if (b8 < 1 || c8 < 0) return -1;
  if (d.d_beta_tab && !d.beta_tab_code_ok) return -1;
  return multi_tab_index(h.z, h.mcol, h.mrow, h.row_weights.data(), h.row_active_cols.data(), h.row_offset.data());
}

size_t multi_code_lds() { return (size_t)mc_lds_max<0>(); }

int multi_code_resident_blocks() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_multi_code, kMcThreads,
                                                   multi_code_lds()) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * cus;
}

int multi_resident_blocks(int elem, int mode, int maxw, int waves, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, multi_kernel_for(elem, mode, maxw), 64 * waves, lds) !=
          hipSuccess || per_cu < 1)
    per_cu = 1;
  return per_cu * cus;
}

template <class P>
static void launch_mp(int maxw, dim3 g, dim3 b, size_t lds, hipStream_t st, const MultiArgs& a) {
  switch (maxw) {
  case 8: hipLaunchKernelGGL((k_multi_decode<P, 8>), g, b, lds, st, a); break;
  case 16: hipLaunchKernelGGL((k_multi_decode<P, 16>), g, b, lds, st, a); break;
  case 24: hipLaunchKernelGGL((k_multi_decode<P, 24>), g, b, lds, st, a); break;
  default: hipLaunchKernelGGL((k_multi_decode<P, 32>), g, b, lds, st, a); break;
  }
}

hipError_t launch_multi(int elem, int mode, int maxw, const MultiArgs& a, uint32_t grid, hipStream_t st,
                        bool code_kernel) {
  if (a.num_cw == 0) return hipSuccess;
  const uint32_t g1 = (a.num_cw + 255) / 256 < 1024 ? (a.num_cw + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_multi_hist, dim3(g1), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_multi_scan, dim3(1), dim3(64), 0, st, a);
  hipLaunchKernelGGL(k_multi_scatter, dim3(g1), dim3(256), 0, st, a);
  if (code_kernel) {
    hipLaunchKernelGGL(k_multi_code, dim3(grid), dim3(kMcThreads), multi_code_lds(), st, a);
    return hipGetLastError();
  }
  const size_t lds = (size_t)a.tab_bytes + (size_t)a.waves * a.wave_bytes;
  const dim3 g(grid), b(64 * a.waves);
  using F1 = FloatPolicy<float, true>;
  using F0 = FloatPolicy<float, false>;
  using D1 = FloatPolicy<double, true>;
  using D0 = FloatPolicy<double, false>;
  switch (elem) {
  case FECSIM_I8: launch_mp<IntPolicy<8>>(maxw, g, b, lds, st, a); break;
  case FECSIM_I16: launch_mp<IntPolicy<16>>(maxw, g, b, lds, st, a); break;
  case FECSIM_I32: launch_mp<IntPolicy<32>>(maxw, g, b, lds, st, a); break;
  case FECSIM_F32:
    if (mode == FECSIM_PROD) launch_mp<F1>(maxw, g, b, lds, st, a);
    else launch_mp<F0>(maxw, g, b, lds, st, a);
    break;
  default:
    if (mode == FECSIM_PROD) launch_mp<D1>(maxw, g, b, lds, st, a);
    else launch_mp<D0>(maxw, g, b, lds, st, a);
    break;
  }
  return hipGetLastError();
}

}  // namespace fecsim

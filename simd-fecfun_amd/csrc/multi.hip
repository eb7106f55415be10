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
#include <hip/hip_runtime.h>

#include <type_traits>

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

hipError_t launch_multi(int elem, int mode, int maxw, const MultiArgs& a, uint32_t grid, hipStream_t st) {
  if (a.num_cw == 0) return hipSuccess;
  const uint32_t g1 = (a.num_cw + 255) / 256 < 1024 ? (a.num_cw + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_multi_hist, dim3(g1), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_multi_scan, dim3(1), dim3(64), 0, st, a);
  hipLaunchKernelGGL(k_multi_scatter, dim3(g1), dim3(256), 0, st, a);
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

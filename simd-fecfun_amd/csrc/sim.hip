// Simulation kernels around the decoder: counter-based data bits, the QC encoder,
// BPSK + AWGN -> quantized channel LLRs, and error counting.
//
//   k_encode  src/Top/gen_rand_databits.cpp:11-40 + src/LDPCEnc/LDPCencode.cpp:40-100
//   k_llr     src/RNGfunc/prng_functions.cpp:148-186 (normdist2) + src/Top/gen_metrics.cpp:18-121
//   k_count   src/Top/LDPC_count_errors.cpp:40-136
//
// Bits of a codeword are [mcol][col_words] 32-bit words, row r of a column at word
// r >> 5, bit r & 31 (with col_words = 16 this is byte-for-byte the reference's
// 512-bit cw_bits word per column, bit r at byte r >> 3, bit r & 7).
//
// Random numbers are counter based, so every codeword can be generated by any
// thread and a batch split over launches (or GPUs) reproduces the same stream:
//   data word:   splitmix64((seed ^ DATA_KEY) + (q + 1) * GAMMA), q = (cw*mcol + c)*8 + k
//                (64-bit chunk k of data column c)
//   noise pair:  splitmix64(seed + (p + 1) * GAMMA), p = (cw*ceil(mcol/2) + col/2)*z + row,
//                turned into two normal samples by the reference's Box-Muller
//                transform (sin for even columns, cos for odd columns).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "decode_core.h"
#include "fecmath.h"
#include "internal.h"

namespace fecsim {

namespace {

constexpr uint64_t kGamma = 0x9e3779b97f4a7c15ull;
constexpr uint64_t kDataKey = 0xd1b54a32d192ed03ull;
constexpr uint64_t kExtraKey = 0x5851f42d4c957f2dull;

__device__ __forceinline__ uint32_t zmask_word(int z, int w) {
  const int valid = z - 32 * w;
  return valid >= 32 ? 0xffffffffu : (valid <= 0 ? 0u : (0xffffffffu >> (32 - valid)));
}

// 32 bits of a Z-bit column starting at row a (0 <= a < z), wrapping modulo z.
// col[] holds the column with bits >= z cleared and one zero word after it.
__device__ __forceinline__ uint32_t cyc32(const uint32_t* col, int z, int a) {
  if (z >= 32) {
    const int i = a >> 5;
    const uint64_t pair = ((uint64_t)col[i + 1] << 32) | col[i];
    uint32_t x = (uint32_t)(pair >> (a & 31));
    const int rem = z - a;
    if (rem < 32) x |= col[0] << rem;
    return x;
  }
  const uint32_t c0 = col[0];
  uint32_t x = 0;
  int p = a;
  for (int t = 0; t < 32; ++t) {
    x |= ((c0 >> p) & 1u) << t;
    if (++p == z) p = 0;
  }
  return x;
}

// ---------------------------------------------------------------------------
// data bits + encoder: one wave holds `cpw` codewords, `zw` lanes per codeword
// (lane w owns word w of every column); the columns live in the wave's LDS slice
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_encode(EncArgs a) {
  extern __shared__ uint32_t lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int zw = a.zw, cstride = zw + 1;
  const int slot = lane / zw, w = lane - slot * zw;
  const bool lane_ok = slot < a.cpw;
  uint32_t* col0 = lds + (size_t)wv * a.cpw * a.slot_words + (size_t)(lane_ok ? slot : 0) * a.slot_words;
  const uint32_t zm = zmask_word(a.z, w);
  const unsigned long long cw_base = a.cw0 + (a.cw0_dev ? *a.cw0_dev : 0ull);
  const uint64_t dkey = a.seed ^ kDataKey;
  const uint64_t step = (uint64_t)gridDim.x * a.waves * a.cpw;
  for (uint64_t first = ((uint64_t)blockIdx.x * a.waves + wv) * a.cpw; first < a.num_cw; first += step) {
    const uint64_t cw = first + slot;
    const bool valid = lane_ok && cw < a.num_cw;
    uint32_t* g = a.bits + (valid ? cw : 0) * (uint64_t)a.mcol * a.col_words;
    // data columns
    for (int c = 0; c < a.ndata; ++c) {
      uint32_t v;
      if (a.gen) {
        const uint64_t q = ((uint64_t)(cw_base + cw) * (uint64_t)a.mcol + (uint64_t)c) * 8u + (uint64_t)(w >> 1);
        const uint64_t r = fecmath::splitmix64(dkey + (q + 1) * kGamma);
        v = (w & 1) ? (uint32_t)(r >> 32) : (uint32_t)r;
        v &= zm & ~a.fill[c * zw + w];
      } else {
        v = valid ? g[(size_t)c * a.col_words + w] & zm : 0u;
      }
      if (lane_ok) {
        col0[c * cstride + w] = v;
        if (w == 0) col0[c * cstride + zw] = 0u;
      }
      if (a.gen && valid)
        for (int x = w; x < a.col_words; x += zw) g[(size_t)c * a.col_words + x] = x < zw ? v : 0u;
    }
    wave_sync();
    // parity columns in schedule order: XOR of rotated columns
    int idx = 0;
    for (int i = 0; i < a.mrow; ++i) {
      uint32_t acc = 0;
      const int cnt = a.cnt[i];
      for (int k = 0; k < cnt; ++k) {
        const uint32_t s = a.src[idx + k];
        const int c = (int)(s >> 16), sh = (int)(s & 0xffffu);
        int at = 32 * w + sh;
        if (at >= a.z) at -= a.z;
        acc ^= cyc32(col0 + c * cstride, a.z, at);
      }
      idx += cnt;
      acc &= zm;
      const int pc = a.ndata + a.colnum[i];
      if (lane_ok) {
        col0[pc * cstride + w] = acc;
        if (w == 0) col0[pc * cstride + zw] = 0u;
      }
      wave_sync();
    }
    if (valid) {
      for (int pc = a.ndata; pc < a.mcol; ++pc) {
        const uint32_t v = col0[pc * cstride + w];
        for (int x = w; x < a.col_words; x += zw) g[(size_t)pc * a.col_words + x] = x < zw ? v : 0u;
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// noise + quantized LLRs: one thread per (codeword, column pair, row)
// ---------------------------------------------------------------------------
// hp = (p + 1) * GAMMA for pair counter p
__device__ __forceinline__ void noise_pair(uint64_t seed, uint64_t hp, float* s_out, float* c_out) {
  // u = float((r & 0x7fffffff) ^ 0x7fffffff) * 2^-31, an extra draw scaled by 2^-62
  // (plus 2^-63) when the top 28 bits are zero, g = sqrt(-2 ln u), angle = int32 * 2^-31
  // (prng_functions.cpp:148-186, one pair per counter)
  const uint64_t r = fecmath::splitmix64(seed + hp);
  const int32_t lo = (int32_t)(uint32_t)r, hi = (int32_t)(r >> 32);
  const int32_t v = (lo & 0x7fffffff) ^ 0x7fffffff;
  float u = (float)v * 0x1.0p-31f;
  if ((v & 0x7ffffff8) == 0) {
    const uint64_t r2 = fecmath::splitmix64((seed ^ kExtraKey) + hp);
    const int32_t v2 = ((int32_t)(uint32_t)r2 & 0x7fffffff) ^ 0x7fffffff;
    u = u + __builtin_fmaf((float)v2, 0x1.0p-62f, 0x1.0p-63f);
  }
  const float g = __builtin_sqrtf(fecmath::fastm2log(u));
  float s, c;
  fecmath::sincos_norm((float)hi * 0x1.0p-31f, &s, &c);
  *s_out = s * g;
  *c_out = c * g;
}

// cvtps2dq: nearest-even, out of range or NaN -> INT32_MIN (gen_metrics.cpp:98 roundi)
__device__ __forceinline__ int32_t roundi_x86(float x) {
  if (!(fabsf(x) < 2147483648.0f)) return INT32_MIN;
  return (int32_t)__builtin_rintf(x);
}

template <typename T>
__device__ __forceinline__ T to_elem(float l) {
  if constexpr (std::is_floating_point<T>::value) return (T)l;
  else return (T)roundi_x86(l);   // truncation to the element width, as compress() does
}

__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long* red) {
  // wave sum with DPP-free shuffles, then across the block's waves
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  return t;
}

template <typename T>
__global__ void __launch_bounds__(256) k_llr(LlrArgs a) {
  __shared__ unsigned long long red[4];
  const uint32_t per_cw = (uint32_t)a.hm * (uint32_t)a.z;
  const unsigned long long cw_base = a.cw0 + (a.cw0_dev ? *a.cw0_dev : 0ull);
  T* out = reinterpret_cast<T*>(a.out);
  unsigned long long err = 0;
  const uint32_t total = a.num_cw * per_cw;   // host splits launches below 2^30
  // The pair counter is linear in the flat index t = (cw*hm + k)*z + r: p = p0 + t.  The
  // loop advances (cw, k, r) and (p + 1) * GAMMA by the grid stride instead of dividing
  // and multiplying for every item.
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t cw = fdiv(t, a.div_cw);
  uint32_t k = fdiv(t - cw * per_cw, a.div_z);
  uint32_t r = t - cw * per_cw - k * (uint32_t)a.z;
  const uint32_t dcw = fdiv(stride, a.div_cw);
  const uint32_t dk = fdiv(stride - dcw * per_cw, a.div_z);
  const uint32_t dr = stride - dcw * per_cw - dk * (uint32_t)a.z;
  const uint64_t p0 = (uint64_t)(cw_base + a.cw_off) * per_cw;
  uint64_t hp = (p0 + 1 + t) * kGamma;
  const uint64_t dhp = (uint64_t)stride * kGamma;
  // element offsets of column 2k of codeword cw_off + cw in the bit words and in the
  // output, advanced with (cw, k, r) like hp (no 64-bit multiplies in the loop)
  const uint64_t cwbits = (uint64_t)a.mcol * a.col_words;
  uint64_t bidx = (a.cw_off + cw) * cwbits + (uint64_t)(2 * k) * a.col_words;
  long long oidx = (long long)(a.cw_off + cw) * a.cw_stride + (long long)(2 * k) * a.col_stride + a.base;
  const uint64_t d_b = (uint64_t)dcw * cwbits + (uint64_t)(2 * dk) * a.col_words;
  const long long d_o = (long long)dcw * a.cw_stride + (long long)(2 * dk) * a.col_stride;
  const uint64_t rk_b = 2ull * a.col_words;                                  // r wrapped: k + 1
  const long long rk_o = 2LL * a.col_stride;
  const uint64_t kc_b = cwbits - (uint64_t)(2 * a.hm) * a.col_words;        // k wrapped: cw + 1
  const long long kc_o = a.cw_stride - 2LL * a.hm * a.col_stride;
  for (; t < total; t += stride) {
    float ns, nc;
    noise_pair(a.seed, hp, &ns, &nc);
    const uint64_t gcw = a.cw_off + cw;
    float sigma = a.sigma, factor = a.factor;
    if (a.sf) {
      sigma = a.sf[2 * gcw];
      factor = a.sf[2 * gcw + 1];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = 2 * (int)k + h;
      if (col >= a.mcol) break;
      const uint32_t word = a.bits[bidx + (uint64_t)h * a.col_words + (r >> 5)];
      const uint32_t bit = (word >> (r & 31)) & 1u;
      const float orig = bit ? -1.0f : 1.0f;
      const float rx = __builtin_fmaf(h ? nc : ns, sigma, orig);
      err += (__float_as_uint(orig) ^ __float_as_uint(rx)) >> 31;
      float l = rx * factor;
      l = a.max_abs < l ? a.max_abs : l;
      l = -a.max_abs > l ? -a.max_abs : l;
      if (a.has_mask) {
        const int mw = col * a.zw + (r >> 5);
        if ((a.punct[mw] >> (r & 31)) & 1u) l = 0.0f;
        if ((a.fill[mw] >> (r & 31)) & 1u) l = a.max_abs;
      }
      out[oidx + h * a.col_stride + r] = to_elem<T>(l);
    }
    hp += dhp;
    r += dr;
    k += dk;
    cw += dcw;
    bidx += d_b;
    oidx += d_o;
    if (r >= (uint32_t)a.z) {
      r -= a.z;
      ++k;
      bidx += rk_b;
      oidx += rk_o;
    }
    if (k >= (uint32_t)a.hm) {
      k -= a.hm;
      ++cw;
      bidx += kc_b;
      oidx += kc_o;
    }
  }
  const unsigned long long s = block_sum(err, red);
  if (threadIdx.x == 0 && s && a.uncoded) atomicAdd(a.uncoded, s);
}

// ---------------------------------------------------------------------------
// error counting: a block takes a tile of `cpb` codewords and walks their data
// columns one 32-row word per thread: the word's hard decisions (value < 0,
// LDPC_count_errors.cpp:84) against the data bits, one XOR and a popcount for 32
// bits; errors are rare, so they are tallied per codeword with LDS atomics
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_count(CountArgs a) {
  __shared__ uint32_t s_cnt[kCountMaxTile];
  __shared__ unsigned long long s_mask[kCountMaxTile];
  __shared__ unsigned long long red[4];
  const T* post = reinterpret_cast<const T*>(a.post);
  const uint32_t zw = ((uint32_t)a.z + 31u) >> 5;            // words holding rows of a column
  const uint32_t dw = (uint32_t)a.ndata * zw;                // words per codeword
  const uint32_t items = (uint32_t)a.cpb * dw;
  for (int i = threadIdx.x; i < a.cpb; i += blockDim.x) { s_cnt[i] = 0; s_mask[i] = 0; }
  __syncthreads();
  unsigned long long bit_err = 0, cw_err = 0;
  const uint64_t ntiles = ((uint64_t)a.num_cw + a.cpb - 1) / a.cpb;
  // item f of a tile = (codeword lc, data column c, word w), advanced by blockDim per step
  const uint32_t dlc = fdiv(blockDim.x, a.div_dz);
  const uint32_t dc = fdiv(blockDim.x - dlc * dw, a.div_z);
  const uint32_t dwd = blockDim.x - dlc * dw - dc * zw;
  const uint32_t lc0 = fdiv(threadIdx.x, a.div_dz);
  const uint32_t c0 = fdiv(threadIdx.x - lc0 * dw, a.div_z);
  const uint32_t w0 = threadIdx.x - lc0 * dw - c0 * zw;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t cw_first = tile * a.cpb;
    uint32_t lc = lc0, c = c0, w = w0;
    for (uint32_t f = threadIdx.x; f < items; f += blockDim.x) {
      const uint64_t cw = cw_first + lc;
      if (cw < a.num_cw) {
        const T* col = post + ((long long)cw * a.cw_stride + (long long)c * a.col_stride + a.base + 32 * (long long)w);
        const int nb = a.z - 32 * (int)w;   // rows of this word (32 but in a column's last word)
        uint32_t hard = 0;
#pragma unroll
        for (int b = 0; b < 32; ++b)
          if (b < nb) hard |= (col[b] < T(0) ? 1u : 0u) << b;
        const uint32_t err =
            (hard ^ a.bits[(cw * a.mcol + c) * a.col_words + w]) & zmask_word(a.z, (int)w);
        if (err) {
          atomicAdd(&s_cnt[lc], (uint32_t)__popc(err));
          unsigned long long m = 1ull;   // rows r with an error mark bit r % num_par
          if (a.num_par > 1) {
            m = 0ull;
            for (uint32_t x = err; x; x &= x - 1u)
              m |= 1ull << ((32u * w + (uint32_t)__ffs(x) - 1u) % (uint32_t)a.num_par);
          }
          atomicOr(&s_mask[lc], m);
        }
      }
      w += dwd;
      c += dc;
      lc += dlc;
      if (w >= zw) {
        w -= zw;
        ++c;
      }
      if (c >= (uint32_t)a.ndata) {
        c -= a.ndata;
        ++lc;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.cpb; i += blockDim.x) {
      const uint64_t cw = cw_first + i;
      const uint32_t n = s_cnt[i];
      if (cw < a.num_cw) {
        if (a.cw_bit_err) a.cw_bit_err[cw] = n;
        bit_err += n;
        cw_err += (unsigned long long)__popcll(s_mask[i]);
      }
      s_cnt[i] = 0;
      s_mask[i] = 0;
    }
    __syncthreads();
  }
  const unsigned long long sb = block_sum(bit_err, red);
  const unsigned long long sc = block_sum(cw_err, red);
  if (threadIdx.x == 0) {
    if (sb) atomicAdd(&a.counts[0], sb);
    if (sc) atomicAdd(&a.counts[1], sc);
    if (blockIdx.x == 0 && a.cw_counter) atomicAdd(a.cw_counter, (unsigned long long)a.num_cw);
  }
}

// hard decisions (value < 0 -> 1, LDPC_count_errors.cpp:84; -0.0 counts as 0 there) of
// every column, packed like the encoder's codeword bits: one thread per output word
template <typename T>
__global__ void __launch_bounds__(256) k_hard(CountArgs a, uint32_t* out) {
  const T* post = reinterpret_cast<const T*>(a.post);
  const uint64_t per_cw = (uint64_t)a.mcol * (uint64_t)a.col_words;
  const uint64_t words = a.num_cw * per_cw;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t cw = i / per_cw;
    const uint32_t rest = (uint32_t)(i - cw * per_cw);
    const uint32_t c = rest / (uint32_t)a.col_words, wd = rest - c * (uint32_t)a.col_words;
    const T* col = post + (long long)cw * a.cw_stride + (long long)c * a.col_stride + a.base;
    uint32_t v = 0;
    const int r0 = (int)wd * 32;
    for (int b = 0; b < 32; ++b)
      if (r0 + b < a.z) v |= (col[r0 + b] < T(0) ? 1u : 0u) << b;
    out[i] = v;
  }
}

// histogram of per-codeword iteration counts: block-private LDS counts, then one global
// atomic per bin and block
constexpr int kHistMaxBins = 1024;
__global__ void __launch_bounds__(256) k_iter_hist(const uint32_t* iters, uint32_t n, uint32_t nbins,
                                                   unsigned long long* hist) {
  __shared__ uint32_t h[kHistMaxBins];
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t b = iters[i] < nbins - 1 ? iters[i] : nbins - 1;
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// counters and the codeword offset of a simulation loop (fill and advance kernels
// rather than memset nodes, so the loop can be captured in a graph)
__global__ void k_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_advance(unsigned long long* p, unsigned long long by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += by;
}

int resident_blocks(const void* fn, int block, size_t lds) {
  int dev = 0, per_cu = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  return per_cu * cus;
}

}  // namespace

int encode_shape(int z, int mcol, EncArgs* a) {
  a->zw = (z + 31) / 32;
  if (a->zw > 64) return FECSIM_E_UNSUPPORTED;
  a->cpw = 64 / a->zw;
  a->slot_words = mcol * (a->zw + 1);
  const int wave_bytes = a->cpw * a->slot_words * 4;
  if (wave_bytes > 65536) {
    a->cpw = 65536 / (a->slot_words * 4);
    if (a->cpw < 1) return FECSIM_E_UNSUPPORTED;
  }
  a->waves = 65536 / (a->cpw * a->slot_words * 4);
  if (a->waves > 4) a->waves = 4;
  return FECSIM_OK;
}

hipError_t launch_encode(const EncArgs& a, hipStream_t st) {
  if (a.num_cw == 0) return hipSuccess;
  const size_t lds = (size_t)a.waves * a.cpw * a.slot_words * 4;
  const uint64_t per_block = (uint64_t)a.waves * a.cpw;
  uint64_t grid = (a.num_cw + per_block - 1) / per_block;
  const uint64_t cap = (uint64_t)resident_blocks((const void*)k_encode, a.waves * 64, lds) * 4;
  if (grid > cap) grid = cap;
  hipLaunchKernelGGL(k_encode, dim3((uint32_t)grid), dim3(a.waves * 64), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_llr(int elem, const LlrArgs& a0, hipStream_t st) {
  const uint64_t per_cw = (uint64_t)a0.hm * a0.z;
  const uint32_t chunk = (uint32_t)((1ull << 30) / per_cw);
  LlrArgs a = a0;
  a.div_cw = make_fastdiv((uint32_t)per_cw);
  a.div_z = make_fastdiv((uint32_t)a0.z);
  for (uint64_t done = 0; done < a0.num_cw; done += chunk) {
    const uint32_t n = (uint32_t)((a0.num_cw - done) < chunk ? (a0.num_cw - done) : chunk);
    a.cw_off = a0.cw_off + done;
    a.num_cw = n;
    const uint64_t threads = (uint64_t)n * per_cw;
    uint64_t grid = (threads + 255) / 256;
    const void* fn;
    switch (elem) {
    case FECSIM_I8: fn = (const void*)k_llr<int8_t>; break;
    case FECSIM_I16: fn = (const void*)k_llr<int16_t>; break;
    case FECSIM_I32: fn = (const void*)k_llr<int32_t>; break;
    case FECSIM_F32: fn = (const void*)k_llr<float>; break;
    default: fn = (const void*)k_llr<double>; break;
    }
    // at most 8 rounds of resident blocks; but at least 16 pairs a thread beyond one round:
    // every block adds its uncoded-error count to one counter as it ends, and in a short
    // launch the blocks end together, so their atomics queue (~4 ns each)
    const uint64_t res = (uint64_t)resident_blocks(fn, 256, 0);
    const uint64_t lean = (threads + 256 * 16 - 1) / (256 * 16);
    grid = std::min(grid, std::min(res * 8, std::max(res, lean)));
    switch (elem) {
    case FECSIM_I8: hipLaunchKernelGGL(k_llr<int8_t>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
    case FECSIM_I16: hipLaunchKernelGGL(k_llr<int16_t>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
    case FECSIM_I32: hipLaunchKernelGGL(k_llr<int32_t>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
    case FECSIM_F32: hipLaunchKernelGGL(k_llr<float>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL(k_llr<double>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_count(int elem, const CountArgs& a0, hipStream_t st) {
  if (a0.num_cw == 0) return hipSuccess;
  CountArgs a = a0;
  const uint32_t zw = ((uint32_t)a.z + 31u) >> 5;
  const uint32_t dw = (uint32_t)a.ndata * zw;   // 32-row words per codeword
  a.cpb = (int)(1024 / dw);
  if (a.cpb < 1) a.cpb = 1;
  if (a.cpb > kCountMaxTile) a.cpb = kCountMaxTile;
  a.div_dz = make_fastdiv(dw);
  a.div_z = make_fastdiv(zw);
  const uint64_t tiles = ((uint64_t)a.num_cw + a.cpb - 1) / a.cpb;
  // at most 4 rounds of resident blocks; but at least 8 tiles a block beyond one round (each
  // block's two counter atomics at its end queue when the blocks of a short launch end
  // together)
  const uint64_t res = (uint64_t)resident_blocks((const void*)k_count<int8_t>, 256, 0);
  const uint64_t grid = std::min(tiles, std::min(res * 4, std::max(res, (tiles + 7) / 8)));
  switch (elem) {
  case FECSIM_I8: hipLaunchKernelGGL(k_count<int8_t>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
  case FECSIM_I16: hipLaunchKernelGGL(k_count<int16_t>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
  case FECSIM_I32: hipLaunchKernelGGL(k_count<int32_t>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
  case FECSIM_F32: hipLaunchKernelGGL(k_count<float>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
  default: hipLaunchKernelGGL(k_count<double>, dim3((uint32_t)grid), dim3(256), 0, st, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_hard(int elem, const CountArgs& a, uint32_t* out, hipStream_t st) {
  if (a.num_cw == 0) return hipSuccess;
  const uint64_t words = a.num_cw * (uint64_t)a.mcol * a.col_words;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((words + 255) / 256, 65536);
  switch (elem) {
  case FECSIM_I8: hipLaunchKernelGGL(k_hard<int8_t>, dim3(grid), dim3(256), 0, st, a, out); break;
  case FECSIM_I16: hipLaunchKernelGGL(k_hard<int16_t>, dim3(grid), dim3(256), 0, st, a, out); break;
  case FECSIM_I32: hipLaunchKernelGGL(k_hard<int32_t>, dim3(grid), dim3(256), 0, st, a, out); break;
  case FECSIM_F32: hipLaunchKernelGGL(k_hard<float>, dim3(grid), dim3(256), 0, st, a, out); break;
  default: hipLaunchKernelGGL(k_hard<double>, dim3(grid), dim3(256), 0, st, a, out); break;
  }
  return hipGetLastError();
}

hipError_t launch_iter_hist(const uint32_t* iters, uint32_t n, uint32_t nbins, unsigned long long* hist,
                            hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (nbins < 1 || nbins > (uint32_t)kHistMaxBins) return hipErrorInvalidValue;
  const uint32_t grid = std::min<uint32_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_iter_hist, dim3(grid), dim3(256), 0, st, iters, n, nbins, hist);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_advance(unsigned long long* p, unsigned long long by, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, p, by);
  return hipGetLastError();
}

}  // namespace fecsim

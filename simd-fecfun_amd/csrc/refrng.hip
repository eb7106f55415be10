// The reference's own sequential random streams, for replaying one of its runs bit for
// bit (the product's simulation loop uses the counter-based streams of sim.hip instead).
//
// src/RNGfunc/prng_functions.cpp keeps s[8] four-lane 64-bit words seeded from the
// 32 rand_state values (init_randstate :22-26): sfc64 in s[0..3] draws the noise
// (rand_sfc64_N :33-41, normdist2 :148-186), xoshiro256++ in s[4..7] the data bits
// (rand_xoshiro_N :44-61, src/Top/gen_rand_databits.cpp:11-40).  The four lanes of a
// stream advance together, and normdist2 takes an extra draw of all four when any of
// the eight 31-bit values is tiny, so the streams are sequential: one wave replays them,
// lane k holding lane k of the state.  Not a throughput path.
//
// k_llr_noise applies gen_metrics (src/Top/gen_metrics.cpp:18-121) to a noise buffer in
// normdist2's order: per codeword, column and group of eight rows, eight values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "fecmath.h"
#include "internal.h"

struct fecsim_refrng {
  unsigned long long* d_state = nullptr;   // [8][4]: s[i] lane k at 4 i + k
  int device = -1;
};

namespace fecsim {

namespace {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Sfc {
  uint64_t a, b, c, d;
  __device__ __forceinline__ uint64_t next() {   // prng_functions.cpp:33-41
    const uint64_t t = a + b + d;
    d = d + 1;
    a = b ^ (b >> 11);
    b = c + (c << 3);
    c = rotl64(c, 24) + t;
    return t;
  }
};

struct Xoshiro {
  uint64_t s0, s1, s2, s3;
  __device__ __forceinline__ uint64_t next() {   // prng_functions.cpp:44-61
    const uint64_t res = rotl64(s0 + s3, 23) + s0;
    const uint64_t t1 = s1 << 17;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t1;
    s3 = rotl64(s3, 45);
    return res;
  }
};

__device__ __forceinline__ int32_t u31(int32_t r) { return (r & 0x7fffffff) ^ 0x7fffffff; }

// normdist2(result, simd_count): iteration `it` writes result[16 it + i] = sin * g and
// result[16 it + 8 + i] = cos * g for the eight values i of one Vec8f pair; lane k
// owns values 2k (low half of its 64-bit draw) and 2k + 1 (high half)
__global__ void __launch_bounds__(64) k_ref_normdist2(unsigned long long* st, float* out, uint32_t iters) {
  const int k = threadIdx.x;
  const bool on = k < 4;
  Sfc g{};
  if (on) g = Sfc{st[k], st[4 + k], st[8 + k], st[12 + k]};
  for (uint32_t it = 0; it < iters; ++it) {
    const uint64_t r = on ? g.next() : 0ull;
    const int32_t v0 = u31((int32_t)(uint32_t)r), v1 = u31((int32_t)(r >> 32));
    float u0 = (float)v0 * 0x1.0p-31f, u1 = (float)v1 * 0x1.0p-31f;
    const bool tiny = on && (((v0 & 0x7ffffff8) == 0) || ((v1 & 0x7ffffff8) == 0));
    if (__ballot(tiny) != 0ull) {   // an extra draw for all lanes (:161-173)
      const uint64_t r2 = on ? g.next() : 0ull;
      u0 = u0 + __builtin_fmaf((float)u31((int32_t)(uint32_t)r2), 0x1.0p-62f, 0x1.0p-63f);
      u1 = u1 + __builtin_fmaf((float)u31((int32_t)(r2 >> 32)), 0x1.0p-62f, 0x1.0p-63f);
    }
    const uint64_t rc = on ? g.next() : 0ull;
    if (on) {
      const float uu[2] = {u0, u1};
      const int32_t ang[2] = {(int32_t)(uint32_t)rc, (int32_t)(rc >> 32)};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float gg = __builtin_sqrtf(fecmath::fastm2log(uu[h]));
        float s, c;
        fecmath::sincos_norm((float)ang[h] * 0x1.0p-31f, &s, &c);
        out[16ull * it + 2 * k + h] = s * gg;
        out[16ull * it + 8 + 2 * k + h] = c * gg;
      }
    }
  }
  if (on) {
    st[k] = g.a;
    st[4 + k] = g.b;
    st[8 + k] = g.c;
    st[12 + k] = g.d;
  }
}

// gen_rand_databits: per codeword and data column, two draws of the four lanes form the
// 512-bit column word (lane k: bits 64k.. of the first draw, 64(4+k).. of the second),
// masked to Z rows and to the non-filler rows; parity columns are left to the encoder
__global__ void __launch_bounds__(64) k_ref_databits(unsigned long long* st, uint32_t* bits, uint32_t num_cw, int z,
                                                     int mrow, int mcol, int col_words, const uint32_t* fill, int zw) {
  const int k = threadIdx.x;
  const bool on = k < 4;
  Xoshiro x{};
  if (on) x = Xoshiro{st[16 + k], st[20 + k], st[24 + k], st[28 + k]};
  for (uint32_t cw = 0; cw < num_cw; ++cw) {
    for (int dc = 0; dc < mcol - mrow; ++dc) {
      const uint64_t a = on ? x.next() : 0ull;
      const uint64_t b = on ? x.next() : 0ull;
      if (!on) continue;
      uint32_t* col = bits + ((uint64_t)cw * mcol + dc) * col_words;
      const uint64_t w2[2] = {a, b};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int q = 4 * h + k;   // 64-bit word of the column
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          const int wd = 2 * q + half;   // 32-bit word
          if (wd >= col_words) continue;
          uint32_t v = (uint32_t)(w2[h] >> (32 * half));
          const int r0 = 32 * wd;
          const uint32_t zm = r0 >= z ? 0u : (z - r0 >= 32 ? ~0u : ((1u << (z - r0)) - 1u));
          const uint32_t fm = wd < zw ? fill[dc * zw + wd] : 0u;
          col[wd] = v & zm & ~fm;
        }
      }
    }
  }
  if (on) {
    st[16 + k] = x.s0;
    st[20 + k] = x.s1;
    st[24 + k] = x.s2;
    st[28 + k] = x.s3;
  }
}

// cvtps2dq: nearest-even, out of range or NaN -> INT32_MIN (gen_metrics.cpp:98 roundi)
__device__ __forceinline__ int32_t roundi_x86(float x) {
  if (!(fabsf(x) < 2147483648.0f)) return INT32_MIN;
  return (int32_t)__builtin_rintf(x);
}

// gen_metrics from a given noise buffer: one thread per noise value (codeword, column,
// row r < 8 ceil(Z/8)); rows past Z are cut off (:62-65) and not stored
template <typename T>
__global__ void __launch_bounds__(256) k_llr_noise(LlrArgs a, const float* wgn) {
  __shared__ unsigned long long red[4];
  const uint32_t nb8 = (uint32_t)((a.z + 7) / 8) * 8u;
  const uint64_t total = (uint64_t)a.num_cw * a.mcol * nb8;
  T* out = reinterpret_cast<T*>(a.out);
  unsigned long long err = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t cc = i / nb8;
    const uint32_t r = (uint32_t)(i - cc * nb8);
    const uint64_t cw = cc / (uint64_t)a.mcol;
    const int col = (int)(cc - cw * a.mcol);
    if (r >= (uint32_t)a.z) continue;   // cutoff: y = 0, no error, not stored
    const uint32_t word = a.bits[(cw * a.mcol + col) * a.col_words + (r >> 5)];
    const float orig = ((word >> (r & 31)) & 1u) ? -1.0f : 1.0f;
    const float rx = __builtin_fmaf(wgn[i], a.sigma, orig);
    err += (__float_as_uint(orig) ^ __float_as_uint(rx)) >> 31;
    float l = rx * a.factor;
    l = a.max_abs < l ? a.max_abs : l;
    l = -a.max_abs > l ? -a.max_abs : l;
    const int mw = col * a.zw + (int)(r >> 5);
    if ((a.punct[mw] >> (r & 31)) & 1u) l = 0.0f;
    if ((a.fill[mw] >> (r & 31)) & 1u) l = a.max_abs;
    T v;
    if constexpr (std::is_floating_point<T>::value) v = (T)l;
    else v = (T)roundi_x86(l);   // truncation to the element width, as compress() does
    out[(long long)cw * a.cw_stride + (long long)col * a.col_stride + a.base + r] = v;
  }
  for (int o = 32; o > 0; o >>= 1) err += __shfl_xor(err, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = err;
  __syncthreads();
  if (threadIdx.x == 0 && a.uncoded) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    if (t) atomicAdd(a.uncoded, t);
  }
}

template <typename T>
hipError_t launch_noise_t(const LlrArgs& a, const float* wgn, hipStream_t st) {
  const uint64_t total = (uint64_t)a.num_cw * a.mcol * (uint64_t)((a.z + 7) / 8) * 8u;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL((k_llr_noise<T>), dim3(grid), dim3(256), 0, st, a, wgn);
  return hipGetLastError();
}

}  // namespace

}  // namespace fecsim

using namespace fecsim;

extern "C" {

int fecsim_refrng_create(const uint64_t* rand_state, fecsim_refrng** out) {
  if (!rand_state || !out) return fail(FECSIM_E_INVAL, "null argument");
  fecsim_refrng* g = new (std::nothrow) fecsim_refrng();
  if (!g) return fail(FECSIM_E_NOMEM, "host allocation");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipMalloc((void**)&g->d_state, 32 * sizeof(unsigned long long));
  if (e == hipSuccess)
    e = hipMemcpy(g->d_state, rand_state, 32 * sizeof(unsigned long long), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(g->d_state);
    delete g;
    return hip_fail(e, "reference stream state");
  }
  g->device = dev;
  *out = g;
  return FECSIM_OK;
}

void fecsim_refrng_free(fecsim_refrng* g) {
  if (!g) return;
  (void)hipFree(g->d_state);
  delete g;
}

int fecsim_refrng_normdist2(fecsim_refrng* g, float* wgn, uint32_t simd_count, void* stream) {
  if (!g || (!wgn && simd_count)) return fail(FECSIM_E_INVAL, "null argument");
  if (simd_count & 1u) return fail(FECSIM_E_INVAL, "simd_count must be even (normdist2 draws Vec8f pairs)");
  if (!simd_count) return FECSIM_OK;
  hipLaunchKernelGGL(k_ref_normdist2, dim3(1), dim3(64), 0, (hipStream_t)stream, g->d_state, wgn, simd_count / 2);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "normdist2 launch");
}

int fecsim_refrng_databits(fecsim_refrng* g, fecsim_henc* h, uint32_t num_cw, int col_words, uint32_t* bits,
                           void* stream) {
  if (!g || !h || (!bits && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (h->h.z > 512) return fail(FECSIM_E_UNSUPPORTED, "the reference's data words hold at most 512 rows");
  if (!num_cw) return FECSIM_OK;
  const int rc = ensure_device(h);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ref_databits, dim3(1), dim3(64), 0, (hipStream_t)stream, g->d_state, bits, num_cw, h->h.z,
                     h->h.mrow, h->h.mcol, col_words, (const uint32_t*)h->d_fill, h->h.zw);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "databits launch");
}

int fecsim_gen_llr_from_noise(fecsim_henc* h, const fecsim_chan_params* p, const float* wgn, uint32_t num_cw,
                              int col_words, const uint32_t* bits, int elem, int layout_p, void* llr,
                              unsigned long long* uncoded_err, void* stream) {
  if (!h || !p || ((!wgn || !bits || !llr) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (elem < FECSIM_I8 || elem > FECSIM_F64) return fail(FECSIM_E_INVAL, "unknown element kind");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  if (!num_cw) return FECSIM_OK;
  const int rc = ensure_device(h);
  if (rc) return rc;
  const HencHost& hh = h->h;
  LlrArgs a{};
  a.bits = bits;
  a.punct = h->d_punct;
  a.fill = h->d_fill;
  a.z = hh.z;
  a.zw = hh.zw;
  a.mcol = hh.mcol;
  a.col_words = col_words;
  a.num_cw = num_cw;
  chan_sigma(p->snr_db, p->llr_scale, &a.sigma, &a.factor);
  a.max_abs = p->max_abs_llr;
  a.out = llr;
  layout_strides(hh.z, hh.mcol, layout_p, &a.cw_stride, &a.col_stride, &a.base);
  a.uncoded = uncoded_err;
  hipError_t e;
  switch (elem) {
  case FECSIM_I8: e = launch_noise_t<int8_t>(a, wgn, (hipStream_t)stream); break;
  case FECSIM_I16: e = launch_noise_t<int16_t>(a, wgn, (hipStream_t)stream); break;
  case FECSIM_I32: e = launch_noise_t<int32_t>(a, wgn, (hipStream_t)stream); break;
  case FECSIM_F32: e = launch_noise_t<float>(a, wgn, (hipStream_t)stream); break;
  default: e = launch_noise_t<double>(a, wgn, (hipStream_t)stream); break;
  }
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "llr launch");
}

}  // extern "C"

// Generic layered decoder kernels: any .dat table, any element kind, dense or the
// reference's padded memory layout (see decode_core.h for the semantics and the
// wave/LDS mapping): codewords packed into waves (k_decode) or one codeword per
// workgroup for large codewords (k_decode_wg), and the shape choice between them.
#include <algorithm>
#include <cstring>

#include "decode_core.h"

namespace fecsim {

// Codewords are packed into waves (lanes_per_cw lanes each); the grid walks the
// batch with a stride so any grid size covers it.
template <class P, int MAXW>
__global__ void __launch_bounds__(256) k_decode(const KernelArgs a) {
  using S = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const DecCore& d = a.core;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = d.lanes_per_cw, CPW = d.cw_per_wave, Z = d.z;
  const int slot = lane / L;
  const int lr = lane - slot * L;
  const int sl = slot < CPW ? slot : 0;
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem);
  tables_to_lds(tab, d.edges, d.nedge, d.layer_w, d.mrow);
  __syncthreads();
  char* wbase = smem + d.tab_bytes + (size_t)wave * d.wave_bytes;
  S* post = reinterpret_cast<S*>(wbase) + (size_t)sl * d.post_stride;
  S* c2v = reinterpret_cast<S*>(wbase) + (size_t)CPW * d.post_stride + (size_t)sl * d.c2v_stride;
  const S* gin = reinterpret_cast<const S*>(a.metrics_in);
  S* gout = reinterpret_cast<S*>(a.metrics_out);
  unsigned long long tot = 0;

  for (long long first = ((long long)blockIdx.x * d.waves + wave) * CPW; first < (long long)a.num_cw;
       first += (long long)gridDim.x * d.waves * CPW) {   // uniform per wave
    const long long cw = first + slot;
    const bool valid = slot < CPW && cw < (long long)a.num_cw;
    if (valid) cols_to_lds(post, gin, a.base + cw * a.cw_stride, a.col_stride, d.mcol, Z, lr, L);
    const uint32_t iter = decode_loop<P, MAXW>(d, tab, post, c2v, slot, lr, valid);

    if (valid) {
      cols_from_lds(gout, post, a.base + cw * a.cw_stride, a.col_stride, d.mcol, Z, lr, L);
      if (a.c2v) {
        S* gc = reinterpret_cast<S*>(a.c2v) + cw * (long long)d.e;
        for (int i = lr; i < d.e; i += L) gc[i] = c2v[i];
      }
      if (lr == 0 && a.iters) a.iters[cw] = iter;
    }
    if (a.tot_iter) {
      // lane 0 sums the slots' iteration counts
      for (int q = 0; q < CPW; ++q) {
        const unsigned it = (unsigned)__shfl((int)iter, q * L, 64);
        if (first + q < (long long)a.num_cw) tot += it;
      }
    }
    wave_sync();
  }
  if (a.tot_iter && lane == 0 && tot) atomicAdd(a.tot_iter, tot);   // one atomic per wave
}

// One codeword per workgroup, for codes whose messages do not fit a wave's LDS slice
// (5G NR BG1 at Z=384): the posteriors sit in LDS shared by the block, whose threads
// split the rows of every layer (thread t owns rows t, t + blockDim, ...), with a block
// barrier between layers and a block-wide OR of the rows' parities for the syndrome
// exit.  C2V sits in LDS after the posteriors when the block's LDS holds both, else
// (GC2V) in a global scratch slot per block, with the grid capped at the resident
// blocks and walking the batch with a stride.
template <class P, int MAXW, bool GC2V>
__global__ void __launch_bounds__(512) k_decode_wg(const KernelArgs a) {
  using S = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const DecCore& d = a.core;
  const int Z = d.z, tid = threadIdx.x, nt = blockDim.x, ne = d.nedge;
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem);
  tables_to_lds(tab, d.edges, ne, d.layer_w, d.mrow);
  __syncthreads();
  const uint32_t* edges = tab;
  const int32_t* layer_w = reinterpret_cast<const int32_t*>(tab + ne);
  S* post = reinterpret_cast<S*>(smem + d.tab_bytes);
  S* c2v;
  if constexpr (GC2V) c2v = reinterpret_cast<S*>(d.c2v_scratch) + (size_t)blockIdx.x * d.c2v_stride;
  else c2v = post + d.post_stride;
  const S* gin = reinterpret_cast<const S*>(a.metrics_in);
  S* gout = reinterpret_cast<S*>(a.metrics_out);
  const S fbeta = P::kFloat ? (S)d.fbeta : S(0), fcap = P::kFloat ? (S)d.fcap : S(0);
  unsigned long long tot = 0;
  for (long long cw = blockIdx.x; cw < (long long)a.num_cw; cw += gridDim.x) {   // uniform per block
    const long long cbase = a.base + cw * a.cw_stride;
    cols_to_lds(post, gin, cbase, a.col_stride, d.mcol, Z, tid, nt);
    for (int j = 0; j < ne; ++j)   // LDPCdecode.cpp:115-119
      for (int r = tid; r < Z; r += nt) c2v[j * Z + r] = S(0);
    __syncthreads();
    uint32_t iter = 0;
    for (;;) {
      int e0 = 0;
      for (int l = 0; l < d.mrow; ++l) {
        const int w = layer_w[l];
        S fb = fbeta;
        long long ib = d.ibeta;
        if (d.beta_tab) {   // per-(iteration, layer) offsets
          if constexpr (P::kFloat) fb = (S)reinterpret_cast<const double*>(d.beta_tab)[iter * d.mrow + l];
          else ib = reinterpret_cast<const long long*>(d.beta_tab)[iter * d.mrow + l];
        }
        for (int r = tid; r < Z; r += nt) {
          if constexpr (P::kFloat) {
            if constexpr (kRowFloatFast<P, MAXW> && !GC2V)
              if (row_float_sw<P, 2, MAXW>(w, post, c2v, d.edges, e0, r, Z, fb, fcap)) continue;
            row_float<P, MAXW>(post, c2v, edges, e0, w, r, Z, fb, fcap);
          } else {
            row_int<P, MAXW>(post, c2v, edges, e0, w, r, Z, (typename P::C)ib, (typename P::C)d.icap);
          }
        }
        e0 += w;
        __syncthreads();
      }
      int fail = 0;                                      // LDPCdecode.cpp:172-195
      e0 = 0;
      for (int l = 0; l < d.mrow; ++l) {
        const int w = layer_w[l];
        for (int r = tid; r < Z; r += nt) {
          if constexpr (kRowFloatFast<P, MAXW>) {
            bool f = false;
            if (parity_float_sw<P, 2, MAXW>(w, post, d.edges, e0, r, Z, &f)) {
              fail |= f;
              continue;
            }
          }
          fail |= row_parity<P, MAXW>(post, edges, e0, w, r, Z);
        }
        e0 += w;
      }
      ++iter;
      const int any = __syncthreads_or(fail);
      if (!any || iter == d.max_iter) break;             // LDPCdecode.cpp:192-197
    }
    cols_from_lds(gout, post, cbase, a.col_stride, d.mcol, Z, tid, nt);
    if (a.c2v) {
      S* gc = reinterpret_cast<S*>(a.c2v) + cw * (long long)d.e;
      for (int j = 0; j < ne; ++j)
        for (int r = tid; r < Z; r += nt) gc[j * Z + r] = c2v[j * Z + r];
    }
    if (tid == 0) {
      if (a.iters) a.iters[cw] = iter;
      tot += iter;
    }
    __syncthreads();
  }
  if (a.tot_iter && tid == 0 && tot) atomicAdd(a.tot_iter, tot);
}

// ---------------------------------------------------------------------------
// shape choice and dispatch
// ---------------------------------------------------------------------------

static int maxw_bucket(int w) { return w <= 8 ? 8 : w <= 16 ? 16 : w <= 24 ? 24 : w <= 32 ? 32 : 0; }

// Shape with K rows per lane: codewords per wave reduced until a wave's slice
// (posteriors and C2V) fits the 64 KB a block may use beside the edge tables.
static bool fit_shape(const HdecHost& h, int elem_bytes, int K, DecShape* s) {
  const int Z = h.z, N = h.mcol * Z, E = h.total_hc_ones * Z;
  s->tab_bytes = table_bytes(h.total_hc_ones, h.mrow);
  const int room = 65536 - s->tab_bytes;
  const int L = (Z + K - 1) / K;
  if (L > 64) return false;
  const int pad = 16 / elem_bytes;  // 16-byte aligned slots
  s->rows_per_lane = K;
  s->lanes_per_cw = L;
  s->post_stride = (N + pad - 1) / pad * pad;
  s->c2v_stride = (E + pad - 1) / pad * pad;
  s->c2v_global = false;
  const int per_slot = (s->post_stride + s->c2v_stride) * elem_bytes;
  s->cw_per_wave = 64 / L;
  while (s->cw_per_wave > 1 && s->cw_per_wave * per_slot > room) --s->cw_per_wave;
  s->wave_bytes = s->cw_per_wave * per_slot;
  if (s->wave_bytes > room) return false;
  // waves per block: most waves resident in a CU's 160 KB (at most 32); ties -> larger blocks
  int best = 0;
  for (int w = 1; w <= std::min(4, room / s->wave_bytes); ++w) {
    const int blocks = 160 * 1024 / (s->tab_bytes + w * s->wave_bytes);
    const int resident = std::min(32, blocks * w);
    if (resident >= best) {
      best = resident;
      s->waves = w;
    }
  }
  return true;
}

// Lanes doing row work per CU: waves a CU holds within its 160 KB of LDS (at most 32)
// times the active lanes of a wave.
static double lanes_per_cu(const DecShape& s) {
  const int block = s.tab_bytes + s.waves * s.wave_bytes;
  const int blocks = std::max(1, 160 * 1024 / block);
  const int waves = std::min(32, blocks * s.waves);
  return (double)waves * s.cw_per_wave * s.lanes_per_cw;
}

// Best per-wave shape: rows per lane K giving the most row work in flight per CU once
// the LDS limits are applied; ties -> fewer rows per lane.  Returns its score (0: none).
static double best_wave_shape(const HdecHost& h, int elem_bytes, int maxw, DecShape* s) {
  double best = 0.0;
  for (int K = 1; K <= 16; ++K) {
    DecShape t;
    t.maxw = maxw;
    if (!fit_shape(h, elem_bytes, K, &t)) continue;
    const double sc = lanes_per_cu(t);
    if (sc > best * (1.0 + 1e-9)) {
      best = sc;
      *s = t;
    }
  }
  return best;
}

// One codeword per workgroup: C2V in LDS when the block's LDS holds it beside the
// posteriors, else in global memory.  Returns the rows in flight per CU (0: does not fit).
static double wg_shape(const HdecHost& h, int elem_bytes, int maxw, DecShape* s) {
  int lds_max = 65536;
  {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      lds_max = v;
  }
  const int Z = h.z, N = h.mcol * Z, E = h.total_hc_ones * Z;
  const int pad = 16 / elem_bytes;
  DecShape t;
  t.maxw = maxw;
  t.post_stride = (N + pad - 1) / pad * pad;
  t.c2v_stride = (E + pad - 1) / pad * pad;
  t.tab_bytes = table_bytes(h.total_hc_ones, h.mrow);
  t.wg = true;
  t.wg_threads = std::min(512, (Z + 63) / 64 * 64);   // 512 threads leave 256 VGPRs per lane
  t.waves = t.wg_threads / 64;
  t.rows_per_lane = (Z + t.wg_threads - 1) / t.wg_threads;
  t.lanes_per_cw = t.wg_threads;
  t.cw_per_wave = 1;
  const long long posts = (long long)t.post_stride * elem_bytes;
  const long long both = (long long)(t.post_stride + t.c2v_stride) * elem_bytes;
  lds_max -= t.tab_bytes;
  if (posts > lds_max) return 0.0;
  t.c2v_global = both > lds_max;
  t.wave_bytes = (int)(t.c2v_global ? posts : both);
  *s = t;
  const long long blocks = std::max(1LL, std::min(160LL * 1024 / (t.tab_bytes + t.wave_bytes), 32LL / t.waves));
  return (double)blocks * std::min(Z, t.wg_threads);
}

int choose_shape(const HdecHost& h, int elem_bytes, DecShape* s, std::string* why) {
  s->maxw = maxw_bucket(h.max_row_weight);
  if (!s->maxw) { *why = "row weight above 32"; return FECSIM_E_UNSUPPORTED; }
  if (h.mcol * h.z > 65535) { *why = "mcol*z above 65535"; return FECSIM_E_UNSUPPORTED; }
  if (h.z > 1024) { *why = "z above 1024"; return FECSIM_E_UNSUPPORTED; }
  DecShape wave, wg;
  const double sw = best_wave_shape(h, elem_bytes, s->maxw, &wave);
  const double sg = wg_shape(h, elem_bytes, s->maxw, &wg);
  // a workgroup per codeword pays a block barrier per layer: take it when no wave slice
  // holds the codeword, or when it keeps clearly more rows in flight (codewords so large
  // that a CU's LDS holds only a few, each on a partly idle wave)
  if (sw > 0.0 && (sg == 0.0 || wg.c2v_global || sg < 1.25 * sw)) {
    *s = wave;
    return FECSIM_OK;
  }
  if (sg == 0.0) {
    *why = "codeword posteriors do not fit a workgroup's LDS";
    return FECSIM_E_UNSUPPORTED;
  }
  *s = wg;
  return FECSIM_OK;
}

int shape_with_rows(const HdecHost& h, int elem_bytes, int K, DecShape* s) {
  if (K == 0) {   // the per-wave choice of choose_shape (workgroup shapes excluded)
    s->maxw = maxw_bucket(h.max_row_weight);
    if (!s->maxw || h.mcol * h.z > 65535) return 1;
    return best_wave_shape(h, elem_bytes, s->maxw, s) > 0.0 ? 0 : 1;
  }
  s->maxw = maxw_bucket(h.max_row_weight);
  const int Z = h.z, N = h.mcol * Z, E = h.total_hc_ones * Z;
  const int L = (Z + K - 1) / K;
  if (!s->maxw || K < 1 || L > 64 || N > 65535) return 1;
  s->rows_per_lane = K;
  s->lanes_per_cw = L;
  s->cw_per_wave = 64 / L;
  const int pad = 16 / elem_bytes;
  s->post_stride = (N + pad - 1) / pad * pad;
  s->c2v_stride = (E + pad - 1) / pad * pad;
  s->tab_bytes = table_bytes(h.total_hc_ones, h.mrow);
  const int room = 65536 - s->tab_bytes;
  for (;;) {
    s->wave_bytes = s->cw_per_wave * (s->post_stride + s->c2v_stride) * elem_bytes;
    if (s->wave_bytes <= room || s->cw_per_wave == 1) break;
    --s->cw_per_wave;
  }
  if (s->wave_bytes > room) return 1;
  s->waves = room / s->wave_bytes;
  if (s->waves > 4) s->waves = 4;
  s->c2v_global = false;
  return 0;
}

static long long wrap_to(long long x, int bits) {
  const unsigned long long m = (1ull << bits) - 1;
  unsigned long long u = (unsigned long long)x & m;
  if (u >> (bits - 1)) u |= ~m;
  return (long long)u;
}

void fill_core(const DecoderDev& dv, DecCore* c) {
  const HdecHost& h = dv.h;
  c->edges = dv.d_edges;
  c->layer_w = dv.d_layer_w;
  c->mrow = h.mrow;
  c->mcol = h.mcol;
  c->z = h.z;
  c->n = h.mcol * h.z;
  c->e = h.total_hc_ones * h.z;
  c->lanes_per_cw = dv.shape.lanes_per_cw;
  c->cw_per_wave = dv.shape.cw_per_wave;
  c->waves = dv.shape.waves;
  c->rows_per_lane = dv.shape.rows_per_lane;
  c->post_stride = dv.shape.post_stride;
  c->c2v_stride = dv.shape.c2v_stride;
  c->wave_bytes = dv.shape.wave_bytes;
  c->nedge = h.total_hc_ones;
  c->tab_off = 0;
  c->tab_bytes = dv.shape.tab_bytes;
  c->max_iter = dv.params.max_iter;
  c->c2v_scratch = dv.d_c2v_scratch;
  c->beta_tab = dv.d_beta_tab;
  const float mrm = dv.params.max_row_metric, beta = dv.params.beta_offset;
  c->ibeta = c->icap = 0;
  c->fbeta = c->fcap = 0.0;
  if (dv.elem <= FECSIM_I32) {
    // decode_functions.cpp:28-32,60: float -> int truncation, then element broadcast
    const int bits = 8 * dv.elem_bytes;
    c->ibeta = wrap_to((long long)(int)beta, bits);
    if (mrm < 1e-6f) {
      const long long t = wrap_to((long long)(int)mrm, bits);
      c->icap = wrap_to((long long)(int)(1.0 - (double)t), bits);
    } else {
      c->icap = wrap_to((long long)(int)mrm, bits);
    }
  } else if (dv.elem == FECSIM_F32) {
    c->fbeta = (double)beta;
    c->fcap = (double)(mrm < 1e-6f ? (float)(1.0 - (double)mrm) : mrm);
  } else {
    c->fbeta = (double)beta;
    c->fcap = mrm < 1e-6f ? 1.0 - (double)mrm : (double)mrm;
  }
}

void offsets_table(const DecoderDev& dv, const float* beta, size_t n, std::vector<long long>* out) {
  out->resize(n);
  for (size_t i = 0; i < n; ++i) {
    if (dv.elem <= FECSIM_I32) {
      (*out)[i] = wrap_to((long long)(int)beta[i], 8 * dv.elem_bytes);   // as beta_offset
    } else {
      const double v = (double)beta[i];
      memcpy(&(*out)[i], &v, sizeof(v));
    }
  }
}

template <class P>
static const void* wave_kernel_of(int maxw) {
  switch (maxw) {
  case 8: return (const void*)k_decode<P, 8>;
  case 16: return (const void*)k_decode<P, 16>;
  case 24: return (const void*)k_decode<P, 24>;
  default: return (const void*)k_decode<P, 32>;
  }
}

template <class P, bool G>
static const void* wg_kernel_of(int maxw) {
  switch (maxw) {
  case 8: return (const void*)k_decode_wg<P, 8, G>;
  case 16: return (const void*)k_decode_wg<P, 16, G>;
  case 24: return (const void*)k_decode_wg<P, 24, G>;
  default: return (const void*)k_decode_wg<P, 32, G>;
  }
}

template <class P>
static const void* kernel_for(const DecoderDev& d) {
  if (!d.shape.wg) return wave_kernel_of<P>(d.shape.maxw);
  return d.shape.c2v_global ? wg_kernel_of<P, true>(d.shape.maxw) : wg_kernel_of<P, false>(d.shape.maxw);
}

static const void* generic_kernel(const DecoderDev& d) {
  switch (d.elem) {
  case FECSIM_I8: return kernel_for<IntPolicy<8>>(d);
  case FECSIM_I16: return kernel_for<IntPolicy<16>>(d);
  case FECSIM_I32: return kernel_for<IntPolicy<32>>(d);
  case FECSIM_F32:
    return d.mode == FECSIM_PROD ? kernel_for<FloatPolicy<float, true>>(d) : kernel_for<FloatPolicy<float, false>>(d);
  default:
    return d.mode == FECSIM_PROD ? kernel_for<FloatPolicy<double, true>>(d) : kernel_for<FloatPolicy<double, false>>(d);
  }
}

static int block_threads(const DecShape& s) { return s.wg ? s.wg_threads : 64 * s.waves; }
static size_t block_lds(const DecShape& s) {
  return (size_t)s.tab_bytes + (s.wg ? (size_t)s.wave_bytes : (size_t)s.waves * s.wave_bytes);
}

int generic_resident_blocks(const DecoderDev& d) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (prepare_generic(d) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, generic_kernel(d), block_threads(d.shape),
                                                   block_lds(d.shape)) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * cus;
}

hipError_t prepare_generic(const DecoderDev& d) {
  const size_t lds = block_lds(d.shape);
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(generic_kernel(d), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <class P>
static hipError_t launch_p(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  const long long per_block = d.shape.wg ? 1LL : (long long)a.core.waves * a.core.cw_per_wave;
  long long blocks = (a.num_cw + per_block - 1) / per_block;
  if (blocks == 0) return hipSuccess;
  if (d.max_grid && blocks > (long long)d.max_grid) blocks = d.max_grid;
  const dim3 grid((unsigned)blocks), block(block_threads(d.shape));
  const size_t lds = block_lds(d.shape);
  if (!d.shape.wg) {
    switch (d.shape.maxw) {
    case 8: hipLaunchKernelGGL((k_decode<P, 8>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((k_decode<P, 16>), grid, block, lds, st, a); break;
    case 24: hipLaunchKernelGGL((k_decode<P, 24>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_decode<P, 32>), grid, block, lds, st, a); break;
    }
  } else if (d.shape.c2v_global) {
    if (!d.d_c2v_scratch || !d.max_grid) return hipErrorInvalidValue;   // one scratch slot per block
    switch (d.shape.maxw) {
    case 8: hipLaunchKernelGGL((k_decode_wg<P, 8, true>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((k_decode_wg<P, 16, true>), grid, block, lds, st, a); break;
    case 24: hipLaunchKernelGGL((k_decode_wg<P, 24, true>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_decode_wg<P, 32, true>), grid, block, lds, st, a); break;
    }
  } else {
    switch (d.shape.maxw) {
    case 8: hipLaunchKernelGGL((k_decode_wg<P, 8, false>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((k_decode_wg<P, 16, false>), grid, block, lds, st, a); break;
    case 24: hipLaunchKernelGGL((k_decode_wg<P, 24, false>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_decode_wg<P, 32, false>), grid, block, lds, st, a); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_decode(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  switch (d.elem) {
  case FECSIM_I8: return launch_p<IntPolicy<8>>(d, a, st);
  case FECSIM_I16: return launch_p<IntPolicy<16>>(d, a, st);
  case FECSIM_I32: return launch_p<IntPolicy<32>>(d, a, st);
  case FECSIM_F32:
    return d.mode == FECSIM_PROD ? launch_p<FloatPolicy<float, true>>(d, a, st)
                                 : launch_p<FloatPolicy<float, false>>(d, a, st);
  default:
    return d.mode == FECSIM_PROD ? launch_p<FloatPolicy<double, true>>(d, a, st)
                                 : launch_p<FloatPolicy<double, false>>(d, a, st);
  }
}

namespace {
__global__ void k_sp_math(int fn, const float* __restrict__ x, float* __restrict__ y, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fn == 0 ? sp_tanh_half(x[i]) : sp_two_atanh(x[i]);
}
}  // namespace

hipError_t launch_sp_math(int fn, const float* x, float* y, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sp_math, dim3((n + 255) / 256), dim3(256), 0, st, fn, x, y, n);
  return hipGetLastError();
}

}  // namespace fecsim

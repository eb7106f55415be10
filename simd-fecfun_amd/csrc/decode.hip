// Generic layered decoder kernel: any .dat table, any element kind, dense or the
// reference's padded memory layout (see decode_core.h for the semantics and the
// wave/LDS mapping).
#include <algorithm>
#include <cstring>

#include "decode_core.h"

namespace fecsim {

// GC2V: the C2V messages live in a global scratch slot per (block, wave, codeword
// slot) instead of the wave's LDS slice; the grid then walks the codewords with a
// stride so the scratch is sized by the grid, not by the batch.
template <class P, int MAXW, bool GC2V>
__global__ void __launch_bounds__(256) k_decode(const KernelArgs a) {
  using S = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const DecCore& d = a.core;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = d.lanes_per_cw, CPW = d.cw_per_wave, Z = d.z;
  const int slot = lane / L;
  const int lr = lane - slot * L;
  const int sl = slot < CPW ? slot : 0;
  char* wbase = smem + (size_t)wave * d.wave_bytes;
  S* post = reinterpret_cast<S*>(wbase) + (size_t)sl * d.post_stride;
  S* c2v;
  if constexpr (GC2V)
    c2v = reinterpret_cast<S*>(d.c2v_scratch) +
          (((size_t)blockIdx.x * d.waves + wave) * CPW + sl) * (size_t)d.c2v_stride;
  else
    c2v = reinterpret_cast<S*>(wbase) + (size_t)CPW * d.post_stride + (size_t)sl * d.c2v_stride;
  const S* gin = reinterpret_cast<const S*>(a.metrics_in);
  S* gout = reinterpret_cast<S*>(a.metrics_out);
  unsigned long long tot = 0;

  for (long long first = ((long long)blockIdx.x * d.waves + wave) * CPW; first < (long long)a.num_cw;
       first += (long long)gridDim.x * d.waves * CPW) {   // uniform per wave
    const long long cw = first + slot;
    const bool valid = slot < CPW && cw < (long long)a.num_cw;
    if (valid) {
      const long long cbase = a.base + cw * a.cw_stride;
      for (int c = 0; c < d.mcol; ++c)
        for (int r = lr; r < Z; r += L) post[c * Z + r] = gin[cbase + c * a.col_stride + r];
    }
    const uint32_t iter = decode_loop<P, MAXW>(d, post, c2v, slot, lr, valid);

    if (valid) {
      const long long cbase = a.base + cw * a.cw_stride;
      for (int c = 0; c < d.mcol; ++c)
        for (int r = lr; r < Z; r += L) gout[cbase + c * a.col_stride + r] = post[c * Z + r];
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

// ---------------------------------------------------------------------------
// shape choice and dispatch
// ---------------------------------------------------------------------------

static int maxw_bucket(int w) { return w <= 8 ? 8 : w <= 16 ? 16 : w <= 24 ? 24 : w <= 32 ? 32 : 0; }

// Shape with K rows per lane: codewords per wave reduced until a wave's slice fits the
// 64 KB a block may use; C2V in LDS, or in global memory when `global_c2v`.
static bool fit_shape(const HdecHost& h, int elem_bytes, int extra_slot_bytes, int K, bool global_c2v, DecShape* s) {
  const int Z = h.z, N = h.mcol * Z, E = h.total_hc_ones * Z;
  const int L = (Z + K - 1) / K;
  if (L > 64) return false;
  const int pad = 16 / elem_bytes;  // 16-byte aligned slots
  s->rows_per_lane = K;
  s->lanes_per_cw = L;
  s->post_stride = (N + pad - 1) / pad * pad;
  s->c2v_stride = (E + pad - 1) / pad * pad;
  s->c2v_global = global_c2v;
  const int per_slot = (s->post_stride + (global_c2v ? 0 : s->c2v_stride)) * elem_bytes + extra_slot_bytes;
  s->cw_per_wave = 64 / L;
  while (s->cw_per_wave > 1 && s->cw_per_wave * per_slot > 65536) --s->cw_per_wave;
  s->wave_bytes = s->cw_per_wave * per_slot;
  if (s->wave_bytes > 65536) return false;
  s->waves = std::min(4, 65536 / s->wave_bytes);
  return true;
}

// Lanes doing row work per CU: waves a CU holds within its 160 KB of LDS (at most 32)
// times the active lanes of a wave.
static double lanes_per_cu(const DecShape& s) {
  const int block = s.waves * s.wave_bytes;
  const int blocks = std::max(1, 160 * 1024 / block);
  const int waves = std::min(32, blocks * s.waves);
  return (double)waves * s.cw_per_wave * s.lanes_per_cw;
}

int choose_shape(const HdecHost& h, int elem_bytes, int extra_slot_bytes, DecShape* s, std::string* why) {
  s->maxw = maxw_bucket(h.max_row_weight);
  if (!s->maxw) { *why = "row weight above 32"; return FECSIM_E_UNSUPPORTED; }
  if (h.mcol * h.z > 65535) { *why = "mcol*z above 65535"; return FECSIM_E_UNSUPPORTED; }
  if (h.z > 1024) { *why = "z above 1024"; return FECSIM_E_UNSUPPORTED; }
  // rows per lane K: most row work in flight per CU once the LDS limits are applied;
  // ties -> fewer rows per lane.  C2V stays in LDS whenever some K allows it.
  for (int global = 0; global < 2; ++global) {
    double best = -1.0;
    DecShape pick;
    for (int K = 1; K <= 16; ++K) {
      DecShape t;
      t.maxw = s->maxw;
      if (!fit_shape(h, elem_bytes, extra_slot_bytes, K, global != 0, &t)) continue;
      const double sc = lanes_per_cu(t);
      if (sc > best * (1.0 + 1e-9)) {
        best = sc;
        pick = t;
      }
    }
    if (best > 0) {
      *s = pick;
      return FECSIM_OK;
    }
  }
  *why = "codeword posteriors do not fit one wave's LDS slice";
  return FECSIM_E_UNSUPPORTED;
}

int shape_with_rows(const HdecHost& h, int elem_bytes, int K, DecShape* s) {
  if (K == 0) {
    std::string why;
    const int rc = choose_shape(h, elem_bytes, 0, s, &why);
    return rc || s->c2v_global;
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
  for (;;) {
    s->wave_bytes = s->cw_per_wave * (s->post_stride + s->c2v_stride) * elem_bytes;
    if (s->wave_bytes <= 65536 || s->cw_per_wave == 1) break;
    --s->cw_per_wave;
  }
  if (s->wave_bytes > 65536) return 1;
  s->waves = 65536 / s->wave_bytes;
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

template <class P, bool G>
static const void* kernel_of(int maxw) {
  switch (maxw) {
  case 8: return (const void*)k_decode<P, 8, G>;
  case 16: return (const void*)k_decode<P, 16, G>;
  case 24: return (const void*)k_decode<P, 24, G>;
  default: return (const void*)k_decode<P, 32, G>;
  }
}

template <class P>
static const void* kernel_for(const DecoderDev& d) {
  return d.shape.c2v_global ? kernel_of<P, true>(d.shape.maxw) : kernel_of<P, false>(d.shape.maxw);
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

int generic_resident_blocks(const DecoderDev& d) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const size_t lds = (size_t)d.shape.waves * d.shape.wave_bytes;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, generic_kernel(d), 64 * d.shape.waves, lds) !=
          hipSuccess || per_cu < 1)
    per_cu = 1;
  return per_cu * cus;
}

template <class P>
static hipError_t launch_p(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  const long long per_block = (long long)a.core.waves * a.core.cw_per_wave;
  long long blocks = (a.num_cw + per_block - 1) / per_block;
  if (blocks == 0) return hipSuccess;
  if (d.max_grid && blocks > (long long)d.max_grid) blocks = d.max_grid;
  const dim3 grid((unsigned)blocks), block(64 * a.core.waves);
  const size_t lds = (size_t)a.core.waves * a.core.wave_bytes;
  if (d.shape.c2v_global) {
    if (!d.d_c2v_scratch) return hipErrorInvalidValue;
    FECSIM_MAXW_DISPATCH(d.shape.maxw, k_decode, P, true, grid, block, lds, st, a);
  } else {
    FECSIM_MAXW_DISPATCH(d.shape.maxw, k_decode, P, false, grid, block, lds, st, a);
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

}  // namespace fecsim

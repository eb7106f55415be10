// Generic layered decoder kernel: any .dat table, any element kind, dense or the
// reference's padded memory layout (see decode_core.h for the semantics and the
// wave/LDS mapping).
#include "decode_core.h"

namespace fecsim {

template <class P, int MAXW>
__global__ void __launch_bounds__(256) k_decode(const KernelArgs a) {
  using S = typename P::S;
  HIP_DYNAMIC_SHARED(char, smem);
  const DecCore& d = a.core;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = d.lanes_per_cw, CPW = d.cw_per_wave, Z = d.z;
  const int slot = lane / L;
  const int lr = lane - slot * L;
  const long long cw = ((long long)blockIdx.x * d.waves + wave) * CPW + slot;
  const bool valid = slot < CPW && cw < (long long)a.num_cw;
  const int sl = slot < CPW ? slot : 0;
  char* wbase = smem + (size_t)wave * d.wave_bytes;
  S* post = reinterpret_cast<S*>(wbase) + (size_t)sl * d.post_stride;
  S* c2v = reinterpret_cast<S*>(wbase) + (size_t)CPW * d.post_stride + (size_t)sl * d.c2v_stride;
  const S* gin = reinterpret_cast<const S*>(a.metrics_in);
  S* gout = reinterpret_cast<S*>(a.metrics_out);

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
    // one atomic per wave: lane 0 sums the slots' iteration counts
    unsigned long long s = 0;
    for (int q = 0; q < CPW; ++q) {
      const unsigned it = (unsigned)__shfl((int)iter, q * L, 64);
      const long long cq = ((long long)blockIdx.x * d.waves + wave) * CPW + q;
      if (cq < (long long)a.num_cw) s += it;
    }
    if (lane == 0 && s) atomicAdd(a.tot_iter, s);
  }
}

// ---------------------------------------------------------------------------
// shape choice and dispatch
// ---------------------------------------------------------------------------

static int maxw_bucket(int w) { return w <= 8 ? 8 : w <= 16 ? 16 : w <= 24 ? 24 : w <= 32 ? 32 : 0; }

int choose_shape(const HdecHost& h, int elem_bytes, int extra_slot_bytes, DecShape* s, std::string* why) {
  const int Z = h.z;
  s->maxw = maxw_bucket(h.max_row_weight);
  if (!s->maxw) { *why = "row weight above 32"; return FECSIM_E_UNSUPPORTED; }
  const int N = h.mcol * Z, E = h.total_hc_ones * Z;
  if (N > 65535) { *why = "mcol*z above 65535"; return FECSIM_E_UNSUPPORTED; }
  // rows per lane: best lane utilization Z*cpw/(64*K); ties -> fewer rows per lane
  double best = -1.0;
  for (int K = 1; K <= 16; ++K) {
    const int L = (Z + K - 1) / K;
    if (L > 64) continue;
    const int cpw = 64 / L;
    const double util = (double)Z * cpw / (64.0 * K);
    if (util > best + 1e-9) { best = util; s->rows_per_lane = K; s->lanes_per_cw = L; s->cw_per_wave = cpw; }
  }
  if (best < 0) { *why = "z above 1024"; return FECSIM_E_UNSUPPORTED; }
  const int pad = 16 / elem_bytes;  // 16-byte aligned slots
  s->post_stride = (N + pad - 1) / pad * pad;
  s->c2v_stride = (E + pad - 1) / pad * pad;
  const int lds_cap = 64 * 1024;
  for (;;) {
    s->wave_bytes = s->cw_per_wave * ((s->post_stride + s->c2v_stride) * elem_bytes + extra_slot_bytes);
    if (s->wave_bytes <= lds_cap || s->cw_per_wave == 1) break;
    --s->cw_per_wave;
  }
  if (s->wave_bytes > lds_cap) { *why = "codeword does not fit one wave's LDS slice"; return FECSIM_E_UNSUPPORTED; }
  s->waves = lds_cap / s->wave_bytes;
  if (s->waves > 4) s->waves = 4;
  return FECSIM_OK;
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

template <class P>
static hipError_t launch_p(const DecoderDev& d, const KernelArgs& a, hipStream_t st) {
  const long long per_block = (long long)a.core.waves * a.core.cw_per_wave;
  const unsigned blocks = (unsigned)((a.num_cw + per_block - 1) / per_block);
  if (blocks == 0) return hipSuccess;
  const dim3 grid(blocks), block(64 * a.core.waves);
  const size_t lds = (size_t)a.core.waves * a.core.wave_bytes;
  FECSIM_MAXW_DISPATCH(d.shape.maxw, k_decode, P, grid, block, lds, st, a);
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

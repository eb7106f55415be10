// C ABI of the differentiable decoder (declared in include/fecsim.h).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace fecsim;

struct fecsim_trainer {
  HdecHost h;
  fecsim_train_params p{};
  int elem_bytes = 4;
  int n_pad = 0, tab_bytes = 0;
  int waves[2] = {4, 4}, global_p[2] = {0, 0};   // per kernel: [forward, backward]
  int reg_shape = -1;   // compiled shape of the register-C2V kernels, -1: train.hip's kernels
  uint32_t* d_edges = nullptr;
  int32_t* d_layer_w = nullptr;
  int seg = 1;   // checkpoint every seg iterations
  void* ckP = nullptr;
  void* ckC = nullptr;
  void* wc = nullptr;
  void* qs = nullptr;
  void* gc = nullptr;
  void* partial = nullptr;
  void* gscratch = nullptr;
  // the last forward's batch and its alpha, beta (copied on the forward's stream, so the
  // adjoint does not depend on the caller keeping its buffers)
  uint32_t last_batch = 0;
  int have_fwd = 0;
  void* alpha = nullptr;
  void* beta = nullptr;
};

namespace {

constexpr int kTrainMaxW = 32;

void free_all(fecsim_trainer* t) {
  (void)hipFree(t->d_edges);
  (void)hipFree(t->d_layer_w);
  (void)hipFree(t->ckP);
  (void)hipFree(t->ckC);
  (void)hipFree(t->wc);
  (void)hipFree(t->qs);
  (void)hipFree(t->gc);
  (void)hipFree(t->partial);
  (void)hipFree(t->gscratch);
  (void)hipFree(t->alpha);
  (void)hipFree(t->beta);
}

TrainArgs base_args(const fecsim_trainer* t, uint32_t batch, int bwd) {
  TrainArgs a{};
  const HdecHost& h = t->h;
  a.edges = t->d_edges;
  a.layer_w = t->d_layer_w;
  a.mrow = h.mrow;
  a.z = h.z;
  a.n = h.mcol * h.z;
  a.e = h.total_hc_ones * h.z;
  a.n_pad = t->n_pad;
  a.nedge = t->h.total_hc_ones;
  a.tab_bytes = t->tab_bytes;
  a.iters = (int)t->p.iters;
  a.maxw = h.max_row_weight;
  a.waves = t->waves[bwd];
  a.is_double = t->p.elem == FECSIM_F64;
  a.quant = t->p.elem == FECSIM_I8;
  a.reg_shape = t->reg_shape;
  a.global_p = t->global_p[bwd];
  a.cap = (double)t->p.cap;
  a.batch = batch;
  a.seg = t->seg;
  a.ckP = t->ckP;
  a.ckC = t->ckC;
  a.wc = t->wc;
  a.qs = t->qs;
  a.gc = t->gc;
  a.partial = t->partial;
  a.gscratch = t->gscratch;
  return a;
}

}  // namespace

extern "C" {

int fecsim_trainer_create(const fecsim_hdec* hd, const fecsim_train_params* p, fecsim_trainer** out) {
  if (!hd || !p || !out) return fail(FECSIM_E_INVAL, "null argument");
  if (p->elem != FECSIM_F32 && p->elem != FECSIM_F64 && p->elem != FECSIM_I8)
    return fail(FECSIM_E_INVAL, "elem must be FECSIM_F32, FECSIM_F64 or FECSIM_I8");
  if (p->iters == 0 || p->iters > 1000 || p->max_batch == 0) return fail(FECSIM_E_INVAL, "need 1..1000 iterations and max_batch >= 1");
  if (p->flags & ~FECSIM_TRAIN_GLOBAL_C2V) return fail(FECSIM_E_INVAL, "unknown trainer flags");
  const HdecHost& h = hd->h;
  if (h.max_row_weight > kTrainMaxW) return fail(FECSIM_E_UNSUPPORTED, "row weight above 32");
  fecsim_trainer* t = new (std::nothrow) fecsim_trainer();
  if (!t) return fail(FECSIM_E_NOMEM, "host allocation");
  t->h = h;
  t->p = *p;
  t->elem_bytes = p->elem == FECSIM_F64 ? 8 : 4;
  // C2V messages in registers (one wave per codeword, posteriors in LDS) when the code's
  // shape is compiled for it: no global C2V traffic on the layers' serial chain
  t->reg_shape = (p->flags & FECSIM_TRAIN_GLOBAL_C2V) ? -1 : train_reg_shape(h, p->elem == FECSIM_F64);
  const size_t N = (size_t)h.mcol * h.z, E = (size_t)h.total_hc_ones * h.z, T = p->iters, B = p->max_batch;
  t->n_pad = (int)((N + 3) / 4 * 4);
  t->tab_bytes = ((h.total_hc_ones + h.mrow) * 4 + 15) / 16 * 16;   // edge words + layer weights in LDS
  const size_t wave_bytes = 2 * (size_t)t->n_pad * t->elem_bytes;
  const size_t room = 65536 - (size_t)t->tab_bytes;
  // Posteriors and their adjoints in LDS (waves per block: the most waves a CU's 160 KB
  // holds, ties to larger blocks) or in a global scratch slot per codeword; each kernel
  // takes the form with more waves resident per CU.  For 802.11n z81 at T = 10 both take
  // global posteriors: the adjoint (104 VGPRs) then holds 16 waves per CU against 10 with
  // LDS and runs 11% faster (5.15 -> 4.57 ms per 4096 codewords); the forward is
  // latency-bound either way (2.02 -> 2.06 ms).
  int lds_waves = 0;
  if (wave_bytes <= room) {
    int best = 0;
    for (int w = 1; w <= 4 && (size_t)w * wave_bytes <= room; ++w) {
      const int resident = w * (int)(160 * 1024 / (t->tab_bytes + w * wave_bytes));
      if (resident >= best) {
        best = resident;
        lds_waves = w;
      }
    }
  }
  std::vector<uint32_t> edges(h.total_hc_ones);
  for (int e = 0; e < h.total_hc_ones; ++e)
    edges[e] = ((uint32_t)(h.row_active_cols[e] * h.z) << 16) | (uint32_t)h.row_offset[e];
  const size_t eb = t->elem_bytes;
  hipError_t he = hipMalloc((void**)&t->d_edges, sizeof(uint32_t) * edges.size());
  if (he == hipSuccess) he = hipMalloc((void**)&t->d_layer_w, sizeof(int32_t) * h.mrow);
  if (he == hipSuccess) he = hipMemcpy(t->d_edges, edges.data(), sizeof(uint32_t) * edges.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(t->d_layer_w, h.row_weights.data(), sizeof(int32_t) * h.mrow, hipMemcpyHostToDevice);
  // checkpoints (posteriors and C2V, N + E values) at the starts of ceil(T / K) segments, the
  // last one K long; the adjoint recomputes one segment of at most K iterations at a time,
  // keeping S values per iteration (the V2C messages, S = E, or the register kernels'
  // reverse state).  K is the segment length of least memory ceil(T / K) (N + E) + K S
  // (the larger K on a tie), of order sqrt(T (N + E) / S)
  const size_t S = t->reg_shape >= 0 ? (size_t)h.mrow * train_reg_state_words(p->elem == FECSIM_I8) * h.z : E;
  size_t K = 1, best = SIZE_MAX;
  for (size_t k = 1; k <= T; ++k) {
    const size_t m = (T + k - 1) / k * (N + E) + k * S;
    if (m <= best) {
      best = m;
      K = k;
    }
  }
  t->seg = (int)K;
  const size_t NS = (T + K - 1) / K;
  if (he == hipSuccess) he = hipMalloc(&t->ckP, B * NS * N * eb);
  if (he == hipSuccess) he = hipMalloc(&t->ckC, B * NS * E * eb);
  if (he == hipSuccess && t->reg_shape < 0) he = hipMalloc(&t->wc, B * E * eb);   // working C2V of train.hip
  if (he == hipSuccess) he = hipMalloc(&t->qs, B * K * S * eb);
  if (he == hipSuccess) he = hipMalloc(&t->gc, B * E * eb);
  if (he == hipSuccess) he = hipMalloc(&t->partial, B * T * h.mrow * 2 * eb);
  for (int bwd = 0; bwd < 2 && he == hipSuccess; ++bwd) {
    t->global_p[bwd] = 1;
    t->waves[bwd] = 4;
    if (lds_waves) {
      TrainArgs g = base_args(t, 1, bwd);
      t->global_p[bwd] = 0;
      t->waves[bwd] = lds_waves;
      TrainArgs l = base_args(t, 1, bwd);
      const int wg = train_resident_waves(g, bwd), wl = train_resident_waves(l, bwd);
      if (wg > wl) {
        t->global_p[bwd] = 1;
        t->waves[bwd] = 4;
      }
    }
  }
  if (he == hipSuccess && (t->global_p[0] || t->global_p[1])) he = hipMalloc(&t->gscratch, B * 2 * N * eb);
  if (he == hipSuccess) he = hipMalloc(&t->alpha, T * h.mrow * eb);
  if (he == hipSuccess) he = hipMalloc(&t->beta, T * h.mrow * eb);
  if (he != hipSuccess) {
    free_all(t);
    delete t;
    return hip_fail(he, "trainer buffers");
  }
  *out = t;
  return FECSIM_OK;
}

int fecsim_train_fwd(fecsim_trainer* t, uint32_t batch, const void* llr, const void* alpha, const void* beta,
                     void* post, void* stream) {
  if (!t || ((!llr || !alpha || !beta || !post) && batch)) return fail(FECSIM_E_INVAL, "null argument");
  if (batch > t->p.max_batch) return fail(FECSIM_E_INVAL, "batch above max_batch");
  TrainArgs a = base_args(t, batch, 0);
  const size_t pbytes = (size_t)t->p.iters * t->h.mrow * t->elem_bytes;
  hipError_t e = hipMemcpyAsync(t->alpha, alpha, pbytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e == hipSuccess) e = hipMemcpyAsync(t->beta, beta, pbytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "train forward parameter copy");
  a.llr = llr;
  a.alpha = t->alpha;
  a.beta = t->beta;
  a.post = post;
  e = launch_train(a, 0, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "train forward launch");
  t->last_batch = batch;
  t->have_fwd = 1;
  return FECSIM_OK;
}

int fecsim_train_bwd(fecsim_trainer* t, uint32_t batch, const void* grad_post, void* grad_alpha, void* grad_beta,
                     void* grad_llr, void* stream) {
  if (!t || !grad_alpha || !grad_beta || (!grad_post && batch)) return fail(FECSIM_E_INVAL, "null argument");
  if (batch != t->last_batch || !t->have_fwd) return fail(FECSIM_E_INVAL, "backward needs a forward of the same batch since the last backward");
  TrainArgs a = base_args(t, batch, 1);
  a.alpha = t->alpha;
  a.beta = t->beta;
  a.grad_post = grad_post;
  a.grad_alpha = grad_alpha;
  a.grad_beta = grad_beta;
  a.grad_llr = grad_llr;
  const hipError_t e = launch_train(a, 1, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "train backward launch");
  t->have_fwd = 0;   // the recomputation overwrote the kept V2C values of the last segment
  return FECSIM_OK;
}

void fecsim_trainer_free(fecsim_trainer* t) {
  if (!t) return;
  free_all(t);
  delete t;
}

}  // extern "C"

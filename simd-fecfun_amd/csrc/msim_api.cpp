// C ABI of the mixed-code simulation loop (declared in include/fecsim.h): every loop
// generates new codewords of several codes (a fixed random code per queue entry and a
// fixed random Eb/N0 per entry), their LLRs, decodes the interleaved queue with one
// persistent launch and counts errors per code, on the device and optionally as one
// replayed HIP graph.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace fecsim;

struct fecsim_msim {
  std::vector<fecsim_henc*> hencs;
  fecsim_multi* m = nullptr;
  fecsim_msim_params p{};
  int ncode = 0, elem = 0, elem_bytes = 1;
  std::vector<uint32_t> count, first;       // queue entries per code, first entry of a code in code order
  std::vector<unsigned long long> llr_base; // element offset of a code's LLR region
  int32_t* d_code = nullptr;                // [num_cw] queue: code id
  unsigned long long* d_off = nullptr;      // [num_cw] queue: LLR offset
  float* d_sf = nullptr;                    // [num_cw] (sigma, factor), in code order
  std::vector<uint32_t*> d_bits;            // per code [count][mcol][zw]
  void* d_llr = nullptr;                    // packed LLRs / posteriors
  unsigned long long llr_bytes = 0;         // its size
  unsigned long long* d_ctr = nullptr;      // [ncode][4] cw, uncoded, bit err, cw err; [ncode*4] iterations; cw0
  hipStream_t cap_stream = nullptr;
  // per-code launches (encoder and channel, counter) fork over side streams (default
  // priority), code c on side[c % side.size()], and join back before the decoder / the advance
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

constexpr int kMStat = 4;
constexpr int kMaxSideStreams = 8;

uint64_t mix64(uint64_t z) {   // SplitMix64 output function (Steele, Lea, Flood 2014)
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double unit(uint64_t seed, uint64_t i) { return (double)(mix64(seed + (i + 1) * 0x9e3779b97f4a7c15ull) >> 11) * 0x1.0p-53; }

void free_msim(fecsim_msim* s) {
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
  for (hipStream_t x : s->side) (void)hipStreamDestroy(x);
  for (hipEvent_t x : s->joins) (void)hipEventDestroy(x);
  if (s->fork) (void)hipEventDestroy(s->fork);
  (void)hipFree(s->d_code);
  (void)hipFree(s->d_off);
  (void)hipFree(s->d_sf);
  for (uint32_t* b : s->d_bits) (void)hipFree(b);
  (void)hipFree(s->d_llr);
  (void)hipFree(s->d_ctr);
}

// the side streams start after the work already on st
int fork_sides(fecsim_msim* s, hipStream_t st) {
  hipError_t e = hipEventRecord(s->fork, st);
  for (size_t k = 0; k < s->side.size() && e == hipSuccess; ++k) e = hipStreamWaitEvent(s->side[k], s->fork, 0);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "fork to the side streams");
}

// st continues after everything on the side streams
int join_sides(fecsim_msim* s, hipStream_t st) {
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < s->side.size() && e == hipSuccess; ++k) {
    e = hipEventRecord(s->joins[k], s->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, s->joins[k], 0);
  }
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "join from the side streams");
}

hipStream_t side_of(const fecsim_msim* s, int c) { return s->side[(size_t)c % s->side.size()]; }

// encoder and channel of every code for the loop at the device's codeword counter
int enqueue_channel(fecsim_msim* s, hipStream_t st0) {
  unsigned long long* cw0_dev = s->d_ctr + (size_t)s->ncode * kMStat + 1;
  hipError_t e;
  int rc = fork_sides(s, st0);
  if (rc) return rc;
  for (int c = 0; c < s->ncode; ++c) {
    if (!s->count[c]) continue;
    const hipStream_t st = side_of(s, c);
    fecsim_henc* h = s->hencs[c];
    unsigned long long* ctr = s->d_ctr + (size_t)c * kMStat;
    // codeword indices of code c this loop: *cw0_dev + first[c] .. + count[c]
    EncArgs ea;
    int rc = enc_args(h, 1, s->p.seed, s->first[c], cw0_dev, s->count[c], h->h.zw, s->d_bits[c], &ea);
    if (rc) return rc;
    if ((e = launch_encode(ea, st)) != hipSuccess) return hip_fail(e, "encode launch");
    LlrArgs la;
    llr_args(h, 0.f, 0.f, s->p.max_abs_llr, s->p.seed, s->first[c], cw0_dev, s->count[c], h->h.zw, s->d_bits[c], 0,
             static_cast<char*>(s->d_llr) + s->llr_base[c] * s->elem_bytes, ctr + 1, &la);
    la.sf = s->d_sf + 2 * (size_t)s->first[c];
    if ((e = launch_llr(s->elem, la, st)) != hipSuccess) return hip_fail(e, "llr launch");
  }
  return join_sides(s, st0);
}

int enqueue(fecsim_msim* s, hipStream_t st) {
  hipError_t e;
  int rc = enqueue_channel(s, st);
  if (rc) return rc;
  rc = fecsim_multi_decode(s->m, s->p.num_cw, s->d_code, s->d_off, s->d_llr, s->d_llr, nullptr,
                               s->d_ctr + (size_t)s->ncode * kMStat, st);
  if (rc) return rc;
  unsigned long long* cw0_dev = s->d_ctr + (size_t)s->ncode * kMStat + 1;
  if ((rc = fork_sides(s, st))) return rc;
  for (int c = 0; c < s->ncode; ++c) {
    if (!s->count[c]) continue;
    fecsim_henc* h = s->hencs[c];
    unsigned long long* ctr = s->d_ctr + (size_t)c * kMStat;
    CountArgs ca;
    count_args(h, s->count[c], h->h.zw, s->d_bits[c], 0, static_cast<char*>(s->d_llr) + s->llr_base[c] * s->elem_bytes,
               ctr + 2, nullptr, ctr + 0, &ca);
    if ((e = launch_count(s->elem, ca, side_of(s, c))) != hipSuccess) return hip_fail(e, "count launch");
  }
  if ((rc = join_sides(s, st))) return rc;
  if ((e = launch_advance(cw0_dev, s->p.num_cw, st)) != hipSuccess) return hip_fail(e, "advance launch");
  return FECSIM_OK;
}

}  // namespace

extern "C" {

int fecsim_msim_create(fecsim_henc* const* hencs, fecsim_multi* m, const fecsim_msim_params* p, fecsim_msim** out) {
  if (!hencs || !m || !p || !out) return fail(FECSIM_E_INVAL, "null argument");
  const int n = (int)m->decs.size();
  if (p->num_cw == 0 || p->num_cw > m->max_cw) return fail(FECSIM_E_INVAL, "num_cw must be 1..max_cw of the queue decoder");
  fecsim_msim* s = new (std::nothrow) fecsim_msim();
  if (!s) return fail(FECSIM_E_NOMEM, "host allocation");
  s->m = m;
  s->p = *p;
  s->ncode = n;
  s->elem = m->elem;
  s->elem_bytes = m->decs[0]->d.elem_bytes;
  s->count.assign(n, 0);
  s->first.assign(n, 0);
  s->llr_base.assign(n, 0);
  for (int c = 0; c < n; ++c) {
    fecsim_henc* h = hencs[c];
    const HdecHost& hd = m->decs[c]->d.h;
    if (!h || h->h.z != hd.z || h->h.mcol != hd.mcol || h->h.mrow != hd.mrow) {
      delete s;
      return fail(FECSIM_E_INVAL, "encoder and decoder tables describe different codes");
    }
    int rc = ensure_device(h);
    if (rc) {
      delete s;
      return rc;
    }
    s->hencs.push_back(h);
  }
  // fixed queue: code of entry i uniform over the codes; entries are numbered per code
  std::vector<int32_t> code(p->num_cw);
  std::vector<uint32_t> rank(p->num_cw);
  for (uint32_t i = 0; i < p->num_cw; ++i) {
    code[i] = std::min(n - 1, (int)(unit(p->seed ^ 0x243f6a8885a308d3ull, i) * n));
    rank[i] = s->count[code[i]]++;
  }
  uint32_t acc = 0;
  unsigned long long base = 0;
  for (int c = 0; c < n; ++c) {
    s->first[c] = acc;
    s->llr_base[c] = base;
    acc += s->count[c];
    base += (unsigned long long)s->count[c] * hencs[c]->h.mcol * hencs[c]->h.z;
  }
  std::vector<unsigned long long> off(p->num_cw);
  std::vector<float> sf(2 * (size_t)p->num_cw);
  for (uint32_t i = 0; i < p->num_cw; ++i) {
    const int c = code[i];
    const HencHost& hh = hencs[c]->h;
    off[i] = s->llr_base[c] + (unsigned long long)rank[i] * hh.mcol * hh.z;
    const double rate = (double)(hh.mcol - hh.mrow) / hh.mcol;
    const float ebn0 = (float)(p->ebn0_lo + (p->ebn0_hi - p->ebn0_lo) * unit(p->seed ^ 0x13198a2e03707344ull, i));
    const float snr = ebn0 + (float)(10.0 * std::log10(rate));
    const size_t j = s->first[c] + rank[i];   // code order
    chan_sigma(snr, p->llr_scale, &sf[2 * j], &sf[2 * j + 1]);
  }
  hipError_t e = hipMalloc((void**)&s->d_code, sizeof(int32_t) * p->num_cw);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_off, sizeof(unsigned long long) * p->num_cw);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_sf, sizeof(float) * sf.size());
  s->llr_bytes = base * (unsigned long long)s->elem_bytes;
  if (e == hipSuccess) e = hipMalloc(&s->d_llr, (size_t)s->llr_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_ctr, sizeof(unsigned long long) * ((size_t)n * kMStat + 2));
  for (int c = 0; c < n && e == hipSuccess; ++c) {
    uint32_t* b = nullptr;
    e = hipMalloc((void**)&b, sizeof(uint32_t) * std::max<size_t>(1, (size_t)s->count[c] * hencs[c]->h.mcol * hencs[c]->h.zw));
    s->d_bits.push_back(b);
  }
  if (e == hipSuccess) e = hipMemcpy(s->d_code, code.data(), sizeof(int32_t) * p->num_cw, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->d_off, off.data(), sizeof(unsigned long long) * p->num_cw, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->d_sf, sf.data(), sizeof(float) * sf.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(s->d_ctr, 0, sizeof(unsigned long long) * ((size_t)n * kMStat + 2));
  const int nside = std::min(n, kMaxSideStreams);
  for (int k = 0; k < nside && e == hipSuccess; ++k) {
    hipStream_t x = nullptr;
    hipEvent_t j = nullptr;
    e = hipStreamCreateWithFlags(&x, hipStreamNonBlocking);
    if (e == hipSuccess) s->side.push_back(x);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if (e == hipSuccess) s->joins.push_back(j);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    free_msim(s);
    delete s;
    return hip_fail(e, "mixed simulation buffers");
  }
  *out = s;
  return FECSIM_OK;
}

int fecsim_msim_run(fecsim_msim* s, uint32_t loops, void* stream) {
  if (!s) return fail(FECSIM_E_INVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  if (!s->p.use_graph) {
    for (uint32_t i = 0; i < loops; ++i) {
      const int rc = enqueue(s, st);
      if (rc) return rc;
    }
    return FECSIM_OK;
  }
  if (!s->exec) {
    hipError_t e = hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "capture stream");
    e = hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "begin capture");
    const int rc = enqueue(s, s->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s->cap_stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "end capture");
    s->graph = g;
    e = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      s->exec = nullptr;
      return hip_fail(e, "graph instantiate");
    }
  }
  for (uint32_t i = 0; i < loops; ++i) {
    const hipError_t e = hipGraphLaunch(s->exec, st);
    if (e != hipSuccess) return hip_fail(e, "graph launch");
  }
  return FECSIM_OK;
}

int fecsim_msim_stats(fecsim_msim* s, fecsim_sim_stats* per_code, unsigned long long* iters) {
  if (!s || !per_code) return fail(FECSIM_E_INVAL, "null argument");
  std::vector<unsigned long long> c((size_t)s->ncode * kMStat + 1);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(c.data(), s->d_ctr, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "stats copy");
  for (int k = 0; k < s->ncode; ++k) {
    per_code[k].cw = c[(size_t)k * kMStat + 0];
    per_code[k].uncoded_bit_err = c[(size_t)k * kMStat + 1];
    per_code[k].iters = 0;
    per_code[k].bit_err = c[(size_t)k * kMStat + 2];
    per_code[k].cw_err = c[(size_t)k * kMStat + 3];
  }
  if (iters) *iters = c[(size_t)s->ncode * kMStat];
  return FECSIM_OK;
}

int fecsim_msim_queue(const fecsim_msim* s, const int32_t** code, const unsigned long long** off, void** llr,
                      const float** sigma_factor) {
  if (!s) return fail(FECSIM_E_INVAL, "null argument");
  if (code) *code = s->d_code;
  if (off) *off = s->d_off;
  if (llr) *llr = s->d_llr;
  if (sigma_factor) *sigma_factor = s->d_sf;
  return FECSIM_OK;
}

int fecsim_msim_llr_bytes(const fecsim_msim* s, unsigned long long* bytes) {
  if (!s || !bytes) return fail(FECSIM_E_INVAL, "null argument");
  *bytes = s->llr_bytes;
  return FECSIM_OK;
}

int fecsim_msim_fill(fecsim_msim* s, void* stream) {
  if (!s) return fail(FECSIM_E_INVAL, "null argument");
  return enqueue_channel(s, (hipStream_t)stream);
}

int fecsim_msim_reset(fecsim_msim* s, void* stream) {
  if (!s) return fail(FECSIM_E_INVAL, "null argument");
  const hipError_t e = launch_fill_u64(s->d_ctr, (uint32_t)(s->ncode * kMStat + 1), 0ull, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "reset launch");
}

void fecsim_msim_free(fecsim_msim* s) {
  if (!s) return;
  free_msim(s);
  delete s;
}

}  // extern "C"

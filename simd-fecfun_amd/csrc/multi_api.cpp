// C ABI of the mixed-code queue decoder and its tile-shape autotuner (declared in
// include/fecsim.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace fecsim;

namespace {

// Workspace of fecsim_multi_decode_static: the queue permutation (FECSIM_MULTI_STATIC_MAX_CW
// entries) and the per-code counters (histogram, scatter cursors, codeword and work-item
// starts for up to 64 codes, the work counter, codewords taken per code), 4 MiB + 1292
// bytes.  Every GPU the library runs on holds its own copy until the process ends; all
// static calls on one GPU share it.
constexpr uint32_t kStaticMaxCw = FECSIM_MULTI_STATIC_MAX_CW;
constexpr size_t kStaticWords = (size_t)kStaticMaxCw + 64 + 64 + 65 + 65 + 1 + 64;
static __device__ __align__(256) uint32_t scratch[kStaticWords];

// Device address of `scratch` per device, looked up once with that device current (in host
// code the symbol's name is not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

int static_workspace(uint32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(FECSIM_E_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FECSIM_OK;
}

void free_multi(fecsim_multi* m) {
  (void)hipFree(m->d_cores);
  (void)hipFree(m->d_ws);
  (void)hipFree(m->d_tabs);
  (void)hipFree(m->d_ctab);
}

// recompute the per-code cores, the LDS slice and the persistent grid for m->shapes
int update(fecsim_multi* m) {
  const int n = (int)m->decs.size();
  std::vector<DecCore> cores(n);
  int wb = 0;
  for (int c = 0; c < n; ++c) {
    DecoderDev tmp = m->decs[c]->d;
    tmp.shape = m->shapes[c];
    fill_core(tmp, &cores[c]);
    cores[c].tab_off = m->tab_off[c];
    cores[c].tab_bytes = m->tab_bytes;
    wb = std::max(wb, m->shapes[c].wave_bytes);
  }
  m->wave_bytes = wb;
  m->waves = std::max(1, std::min(m->waves_pref, std::max(1, 65536 - m->tab_bytes) / wb));
  for (int c = 0; c < n; ++c) cores[c].waves = m->waves;
  const hipError_t e = hipMemcpy(m->d_cores, cores.data(), sizeof(DecCore) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "multi-code tables to device");
  const int blocks =
      multi_resident_blocks(m->elem, m->mode, m->maxw, m->waves, (size_t)m->tab_bytes + (size_t)m->waves * wb);
  if (blocks < 1) return fail(FECSIM_E_HIP, "occupancy query failed");
  m->grid = (uint32_t)blocks;
  return FECSIM_OK;
}

// ws: [ws_cw] queue permutation followed by the per-code counters
MultiArgs args(const fecsim_multi* m, uint32_t* ws, uint32_t ws_cw, uint32_t num_cw, const int32_t* code,
               const unsigned long long* off, const void* in, void* out, uint32_t* iters, unsigned long long* tot) {
  const int n = (int)m->decs.size();
  MultiArgs a{};
  a.cores = m->d_cores;
  a.ncode = n;
  a.num_cw = num_cw;
  a.code = code;
  a.off = off;
  a.in = in;
  a.out = out;
  a.iters = iters;
  a.tot_iter = tot;
  a.perm = ws;
  a.cnt = a.perm + ws_cw;
  a.cursor = a.cnt + n;
  a.code_start = a.cursor + n;
  a.item_start = a.code_start + n + 1;
  a.counter = a.item_start + n + 1;
  a.take = a.counter + 1;
  a.ctab = m->d_ctab;
  a.wave_bytes = m->wave_bytes;
  a.waves = m->waves;
  a.tabs = m->d_tabs;
  a.tab_words = m->tab_words;
  a.tab_bytes = m->tab_bytes;
  return a;
}

int run(fecsim_multi* m, uint32_t* ws, uint32_t ws_cw, uint32_t num_cw, const int32_t* code,
        const unsigned long long* off, const void* in, void* out, uint32_t* iters, unsigned long long* tot,
        hipStream_t st) {
  const MultiArgs a = args(m, ws, ws_cw, num_cw, code, off, in, out, iters, tot);
  if (m->use_code) {   // persistent grid; a workgroup without codewords leaves at once
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(m->code_grid, num_cw));
    const hipError_t e = launch_multi(m->elem, m->mode, m->maxw, a, grid, st, true);
    return e == hipSuccess ? FECSIM_OK : hip_fail(e, "multi-code decode launch");
  }
  // persistent grid, no larger than the work could use
  uint32_t min_cpw = 64;
  for (const DecShape& s : m->shapes) min_cpw = std::min<uint32_t>(min_cpw, (uint32_t)s.cw_per_wave);
  const uint64_t need = ((uint64_t)num_cw + (uint64_t)min_cpw * m->waves - 1) / ((uint64_t)min_cpw * m->waves);
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(m->grid, need));
  const hipError_t e = launch_multi(m->elem, m->mode, m->maxw, a, grid, st, false);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "multi-code decode launch");
}

// seconds per decode of the queue, best of `rounds`
double time_queue(fecsim_multi* m, uint32_t num_cw, const int32_t* code, const unsigned long long* off,
                  const void* in, void* out, int rounds, hipStream_t st) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1.0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1.0;
  }
  double best = std::numeric_limits<double>::infinity();
  for (int r = 0; r < rounds + 1; ++r) {   // the first run warms up
    (void)hipEventRecord(e0, st);
    if (run(m, m->d_ws, m->max_cw, num_cw, code, off, in, out, nullptr, nullptr, st)) {
      best = -1.0;
      break;
    }
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) {
      best = -1.0;
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0) best = std::min(best, (double)ms * 1e-3);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

extern "C" {

int fecsim_multi_create(fecsim_decoder* const* decs, int ncode, uint32_t max_cw, fecsim_multi** out) {
  if (!decs || !out || ncode < 1 || ncode > 64 || max_cw == 0) return fail(FECSIM_E_INVAL, "need 1..64 decoders and max_cw >= 1");
  fecsim_multi* m = new (std::nothrow) fecsim_multi();
  if (!m) return fail(FECSIM_E_NOMEM, "host allocation");
  m->max_cw = max_cw;
  for (int c = 0; c < ncode; ++c) {
    fecsim_decoder* d = decs[c];
    if (!d) {
      delete m;
      return fail(FECSIM_E_INVAL, "null decoder");
    }
    if (c == 0) {
      m->elem = d->d.elem;
      m->mode = d->d.mode;
    } else if (d->d.elem != m->elem || d->d.mode != m->mode) {
      delete m;
      return fail(FECSIM_E_INVAL, "all decoders of a queue need the same element kind and mode");
    }
    DecShape s;
    if (shape_with_rows(d->d.h, d->d.elem_bytes, 0, &s)) {
      delete m;
      return fail(FECSIM_E_UNSUPPORTED, "a code's messages do not fit one wave's LDS slice");
    }
    m->decs.push_back(d);
    m->shapes.push_back(s);
    m->maxw = std::max(m->maxw, s.maxw);
  }
  // every code's [edge words | layer weights], copied to the front of each block's LDS
  std::vector<uint32_t> tabs;
  for (fecsim_decoder* d : m->decs) {
    const HdecHost& h = d->d.h;
    m->tab_off.push_back((int)tabs.size());
    for (int e = 0; e < h.total_hc_ones; ++e)
      tabs.push_back(((uint32_t)(h.row_active_cols[e] * h.z) << 16) | (uint32_t)h.row_offset[e]);
    for (int l = 0; l < h.mrow; ++l) tabs.push_back((uint32_t)h.row_weights[l]);
  }
  m->tab_words = (int)tabs.size();
  m->tab_bytes = (m->tab_words * 4 + 15) / 16 * 16;
  if (m->tab_bytes > 32768) {
    delete m;
    return fail(FECSIM_E_UNSUPPORTED, "the codes' edge tables exceed 32 KB of LDS");
  }
  // the queue kernel with the codes compiled in, when every decoder's code is
  std::vector<int32_t> ctab(ncode);
  m->code_ok = true;
  for (int c = 0; c < ncode; ++c) {
    ctab[c] = multi_code_index(m->decs[c]->d);
    m->code_ok = m->code_ok && ctab[c] >= 0;
  }
  const size_t ws_words = (size_t)max_cw + 5 * (size_t)ncode + 3;
  hipError_t e = hipMalloc((void**)&m->d_cores, sizeof(DecCore) * ncode);
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_ws, sizeof(uint32_t) * ws_words);
  if (e == hipSuccess) e = hipMemset(m->d_ws, 0, sizeof(uint32_t) * ws_words);
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_tabs, sizeof(uint32_t) * tabs.size());
  if (e == hipSuccess)
    e = hipMemcpy(m->d_tabs, tabs.data(), sizeof(uint32_t) * tabs.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_ctab, sizeof(int32_t) * ncode);
  if (e == hipSuccess) e = hipMemcpy(m->d_ctab, ctab.data(), sizeof(int32_t) * ncode, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_multi(m);
    delete m;
    return hip_fail(e, "multi-code buffers");
  }
  int rc = update(m);
  if (!rc && m->code_ok) {
    const int blocks = multi_code_resident_blocks();
    if (blocks < 1) rc = fail(FECSIM_E_HIP, "occupancy query failed");
    m->code_grid = (uint32_t)blocks;
    m->use_code = true;
  }
  if (rc) {
    free_multi(m);
    delete m;
    return rc;
  }
  for (fecsim_decoder* d : m->decs) d->holders.fetch_add(1);
  *out = m;
  return FECSIM_OK;
}

int fecsim_multi_decode(fecsim_multi* m, uint32_t num_cw, const int32_t* code, const unsigned long long* off,
                        const void* llr_in, void* post_out, uint32_t* iters, unsigned long long* tot_iter,
                        void* stream) {
  if (!m || ((!code || !off || !llr_in || !post_out) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (num_cw > m->max_cw) return fail(FECSIM_E_INVAL, "queue longer than max_cw");
  if (!num_cw) return FECSIM_OK;
  return run(m, m->d_ws, m->max_cw, num_cw, code, off, llr_in, post_out, iters, tot_iter, (hipStream_t)stream);
}

int fecsim_multi_decode_static(fecsim_multi* m, uint32_t num_cw, const int32_t* code, const unsigned long long* off,
                               const void* llr_in, void* post_out, uint32_t* iters, unsigned long long* tot_iter,
                               void* stream) {
  if (!m || ((!code || !off || !llr_in || !post_out) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (num_cw > kStaticMaxCw)
    return fail(FECSIM_E_INVAL, "queue longer than the static workspace (FECSIM_MULTI_STATIC_MAX_CW entries)");
  if (!num_cw) return FECSIM_OK;
  uint32_t* ws = nullptr;
  int rc = static_workspace(&ws);
  if (rc) return rc;
  const hipStream_t st = (hipStream_t)stream;
  // the histogram must start at zero: a fill kernel (no memset, so a captured graph
  // replays it correctly); the cursors and the work counter are reset by k_multi_scan
  const hipError_t e = launch_fill_u64(reinterpret_cast<unsigned long long*>(ws + kStaticMaxCw), 32, 0ull, st);
  if (e != hipSuccess) return hip_fail(e, "static workspace clear");
  return run(m, ws, kStaticMaxCw, num_cw, code, off, llr_in, post_out, iters, tot_iter, st);
}

int fecsim_multi_autotune(fecsim_multi* m, uint32_t num_cw, const int32_t* code, const unsigned long long* off,
                          const void* llr_in, void* post_out, int rounds, void* stream) {
  if (!m || !code || !off || !llr_in || !post_out || num_cw == 0) return fail(FECSIM_E_INVAL, "null argument or empty queue");
  if (num_cw > m->max_cw) return fail(FECSIM_E_INVAL, "queue longer than max_cw");
  hipStream_t st = (hipStream_t)stream;
  if (rounds < 1) rounds = 1;
  if (m->kernel_pref == FECSIM_KERNEL_CODE) return FECSIM_OK;   // nothing to tune
  int rc = update(m);
  if (rc) return rc;
  m->use_code = false;   // the generic kernel's shapes first
  double best = time_queue(m, num_cw, code, off, llr_in, post_out, rounds, st);
  if (best < 0) return fail(FECSIM_E_HIP, "autotune timing failed");
  // coordinate descent: rows per lane (hence codewords per wave) of every code, then waves per block
  for (size_t c = 0; c < m->decs.size(); ++c) {
    const HdecHost& h = m->decs[c]->d.h;
    const int k0 = (h.z + 63) / 64;
    DecShape keep = m->shapes[c];
    for (int K = k0; K <= k0 + 4; ++K) {
      DecShape s;
      if (shape_with_rows(h, m->decs[c]->d.elem_bytes, K, &s)) continue;
      if (s.rows_per_lane == keep.rows_per_lane && s.cw_per_wave == keep.cw_per_wave) continue;
      m->shapes[c] = s;
      m->maxw = std::max(m->maxw, s.maxw);
      if ((rc = update(m))) return rc;
      const double t = time_queue(m, num_cw, code, off, llr_in, post_out, rounds, st);
      if (t < 0) return fail(FECSIM_E_HIP, "autotune timing failed");
      if (t < best) {
        best = t;
        keep = s;
      }
    }
    m->shapes[c] = keep;
  }
  int keep_w = m->waves_pref;
  for (int w : {1, 2, 4}) {
    if (w == keep_w) continue;
    m->waves_pref = w;
    if ((rc = update(m))) return rc;
    const double t = time_queue(m, num_cw, code, off, llr_in, post_out, rounds, st);
    if (t < 0) return fail(FECSIM_E_HIP, "autotune timing failed");
    if (t < best) {
      best = t;
      keep_w = w;
    }
  }
  m->waves_pref = keep_w;
  if ((rc = update(m))) return rc;
  // then the kernel: the compiled-code one when it is faster than the tuned generic one
  if (m->code_ok && m->kernel_pref == FECSIM_KERNEL_AUTO) {
    m->use_code = true;
    const double t = time_queue(m, num_cw, code, off, llr_in, post_out, rounds, st);
    if (t < 0) return fail(FECSIM_E_HIP, "autotune timing failed");
    m->use_code = t < best;
  }
  return FECSIM_OK;
}

int fecsim_multi_select_kernel(fecsim_multi* m, int kernel) {
  if (!m) return fail(FECSIM_E_INVAL, "null argument");
  if (kernel != FECSIM_KERNEL_AUTO && kernel != FECSIM_KERNEL_GENERIC && kernel != FECSIM_KERNEL_CODE)
    return fail(FECSIM_E_INVAL, "kernel must be FECSIM_KERNEL_AUTO, _GENERIC or _CODE");
  if (kernel == FECSIM_KERNEL_CODE && !m->code_ok)
    return fail(FECSIM_E_UNSUPPORTED, "a code of the queue is not compiled into the queue kernel");
  m->kernel_pref = kernel;
  m->use_code = m->code_ok && kernel != FECSIM_KERNEL_GENERIC;
  return FECSIM_OK;
}

int fecsim_multi_kernel(const fecsim_multi* m) {
  if (!m) return fail(FECSIM_E_INVAL, "null argument");
  return m->use_code ? FECSIM_KERNEL_CODE : FECSIM_KERNEL_GENERIC;
}

int fecsim_multi_shape(const fecsim_multi* m, int c, int* rows_per_lane, int* cw_per_wave, int* waves_per_block) {
  if (!m || c < 0 || c >= (int)m->shapes.size()) return fail(FECSIM_E_INVAL, "bad code index");
  if (rows_per_lane) *rows_per_lane = m->shapes[c].rows_per_lane;
  if (cw_per_wave) *cw_per_wave = m->shapes[c].cw_per_wave;
  if (waves_per_block) *waves_per_block = m->waves;
  return FECSIM_OK;
}

void fecsim_multi_free(fecsim_multi* m) {
  if (!m) return;
  for (fecsim_decoder* d : m->decs) d->holders.fetch_sub(1);
  free_multi(m);
  delete m;
}

}  // extern "C"

// C ABI of the fecsim library (declared in include/fecsim.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace fecsim;


namespace fecsim {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(FECSIM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

const int kElemBytes[5] = {1, 2, 4, 4, 8};

}  // namespace fecsim

namespace {

// Parse the int32 words of LDPCdecode.raw.dat (decode_matrix_allocate.cpp:22-50)
int parse_hdec(const int32_t* w, size_t n, fecsim_hdec* out) {
  if (n < 7) return fail(FECSIM_E_FORMAT, "decoder table: header truncated");
  HdecHost& h = out->h;
  h.mrow = w[0];
  h.mcol = w[1];
  h.z = w[2];
  h.total_hc_ones = w[3];
  h.max_row_weight = w[4];
  const int32_t lf = w[5], lb = w[6];
  if (h.mrow <= 0 || h.mcol <= h.mrow || h.z <= 0 || h.total_hc_ones <= 0 || lf < 0 || lb < 0 ||
      h.mrow > 4096 || h.mcol > 4096 || h.z > 65535)
    return fail(FECSIM_E_FORMAT, "decoder table: bad header");
  const size_t need = 7 + (size_t)h.mrow + 2 * (size_t)h.total_hc_ones + (size_t)h.mrow + 2 * (size_t)lf +
                      (size_t)h.mrow + 1 + 2 * (size_t)lb;
  if (n < need) return fail(FECSIM_E_FORMAT, "decoder table: truncated (Error in reading decode Matrix data)");
  const int32_t* p = w + 7;
  auto take = [&](std::vector<int32_t>& v, size_t cnt) {
    v.assign(p, p + cnt);
    p += cnt;
  };
  take(h.row_weights, h.mrow);
  take(h.row_active_cols, h.total_hc_ones);
  take(h.row_offset, h.total_hc_ones);
  take(h.f2b_copy_cnt, h.mrow);
  take(h.f2b_cols, lf);
  take(h.f2b_offsets, lf);
  take(h.b2f_copy_cnt, (size_t)h.mrow + 1);
  take(h.b2f_cols, lb);
  take(h.b2f_offsets, lb);
  out->len_f2b = lf;
  out->len_b2f = lb;
  long long s = 0;
  int mx = 0;
  for (int l = 0; l < h.mrow; ++l) {
    if (h.row_weights[l] < 2 || h.row_weights[l] > 53)
      return fail(FECSIM_E_FORMAT, "decoder table: row weight outside 2..53 (min_find_wrap.pl:6)");
    s += h.row_weights[l];
    mx = h.row_weights[l] > mx ? h.row_weights[l] : mx;
  }
  if (s != h.total_hc_ones) return fail(FECSIM_E_FORMAT, "decoder table: row weights do not sum to total_Hc_ones");
  if (mx > h.max_row_weight) return fail(FECSIM_E_FORMAT, "decoder table: max_row_weight too small");
  int e = 0;
  for (int l = 0; l < h.mrow; ++l) {
    std::vector<char> seen(h.mcol, 0);
    for (int j = 0; j < h.row_weights[l]; ++j, ++e) {
      const int c = h.row_active_cols[e], o = h.row_offset[e];
      if (c < 0 || c >= h.mcol || seen[c]) return fail(FECSIM_E_FORMAT, "decoder table: bad column in a layer");
      if (o < 0 || o >= h.z) return fail(FECSIM_E_FORMAT, "decoder table: shift outside 0..z-1");
      seen[c] = 1;
    }
  }
  return FECSIM_OK;
}

// "LDPCtop_Vec64c_min" / "Vec16f_prod" -> element kind and mode
int parse_kind(const char* kind, int* elem, int* mode) {
  if (!kind) return FECSIM_E_INVAL;
  const char* v = strstr(kind, "Vec");
  if (!v) return FECSIM_E_INVAL;
  v += 3;
  while (*v >= '0' && *v <= '9') ++v;
  switch (*v) {
  case 'c': *elem = FECSIM_I8; break;
  case 's': *elem = FECSIM_I16; break;
  case 'w': *elem = FECSIM_I32; break;
  case 'f': *elem = FECSIM_F32; break;
  case 'd': *elem = FECSIM_F64; break;
  default: return FECSIM_E_INVAL;
  }
  ++v;
  if (strcmp(v, "_min") == 0) *mode = FECSIM_MIN;
  else if (strcmp(v, "_prod") == 0) *mode = FECSIM_PROD;
  else return FECSIM_E_INVAL;
  if (*mode == FECSIM_PROD && (*elem != FECSIM_F32 && *elem != FECSIM_F64)) return FECSIM_E_UNSUPPORTED;
  return FECSIM_OK;
}

}  // namespace

extern "C" {

int fecsim_abi_version(void) { return FECSIM_ABI_VERSION; }

const char* fecsim_strerror(int s) {
  switch (s) {
  case FECSIM_OK: return "ok";
  case FECSIM_E_INVAL: return "invalid argument";
  case FECSIM_E_IO: return "i/o error";
  case FECSIM_E_FORMAT: return "malformed table file";
  case FECSIM_E_NOMEM: return "out of memory";
  case FECSIM_E_HIP: return "HIP runtime error";
  case FECSIM_E_UNSUPPORTED: return "unsupported";
  default: return "unknown error";
  }
}

const char* fecsim_last_error(void) { return fecsim::g_err.c_str(); }

int fecsim_hdec_from_words(const int32_t* words, size_t nwords, fecsim_hdec** out) {
  if (!words || !out) return fail(FECSIM_E_INVAL, "null argument");
  fecsim_hdec* h = new (std::nothrow) fecsim_hdec();
  if (!h) return fail(FECSIM_E_NOMEM, "host allocation");
  int rc = parse_hdec(words, nwords, h);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return FECSIM_OK;
}

int fecsim_hdec_load(const char* path, fecsim_hdec** out) {
  if (!path || !out) return fail(FECSIM_E_INVAL, "null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(FECSIM_E_IO, std::string("cannot open ") + path);
  std::vector<int32_t> w;
  int32_t buf[4096];
  size_t got;
  while ((got = fread(buf, sizeof(int32_t), 4096, f)) > 0) w.insert(w.end(), buf, buf + got);
  fclose(f);
  return fecsim_hdec_from_words(w.data(), w.size(), out);
}

int fecsim_hdec_info(const fecsim_hdec* h, fecsim_hdec_info_t* info) {
  if (!h || !info) return fail(FECSIM_E_INVAL, "null argument");
  info->mrow = h->h.mrow;
  info->mcol = h->h.mcol;
  info->z = h->h.z;
  info->total_hc_ones = h->h.total_hc_ones;
  info->max_row_weight = h->h.max_row_weight;
  info->len_f2b = h->len_f2b;
  info->len_b2f = h->len_b2f;
  return FECSIM_OK;
}

void fecsim_hdec_free(fecsim_hdec* h) { delete h; }

int fecsim_decoder_create(const fecsim_hdec* hd, const char* kind, const fecsim_dec_params* prm,
                          fecsim_decoder** out) {
  if (!hd || !prm || !out) return fail(FECSIM_E_INVAL, "null argument");
  int elem, mode;
  int rc = parse_kind(kind, &elem, &mode);
  if (rc == FECSIM_E_UNSUPPORTED) return fail(rc, "sum-product needs a float or double element kind");
  if (rc) return fail(rc, std::string("bad decoder kind: ") + (kind ? kind : "(null)"));
  if (prm->max_iter == 0) return fail(FECSIM_E_INVAL, "max_iter must be at least 1");
  fecsim_decoder* dec = new (std::nothrow) fecsim_decoder();
  if (!dec) return fail(FECSIM_E_NOMEM, "host allocation");
  DecoderDev& d = dec->d;
  d.h = hd->h;
  d.elem = elem;
  d.mode = mode;
  d.elem_bytes = kElemBytes[elem];
  d.params = *prm;
  std::string why;
  rc = choose_shape(d.h, d.elem_bytes, &d.shape, &why);
  if (rc) {
    delete dec;
    return fail(rc, "decoder shape: " + why);
  }
  const HdecHost& h = d.h;
  std::vector<uint32_t> edges(h.total_hc_ones);
  for (int e = 0; e < h.total_hc_ones; ++e)
    edges[e] = ((uint32_t)(h.row_active_cols[e] * h.z) << 16) | (uint32_t)h.row_offset[e];
  hipError_t he = hipMalloc(&d.d_edges, sizeof(uint32_t) * edges.size());
  if (he == hipSuccess) he = hipMalloc(&d.d_layer_w, sizeof(int32_t) * h.mrow);
  if (he == hipSuccess)
    he = hipMemcpy(d.d_edges, edges.data(), sizeof(uint32_t) * edges.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess)
    he = hipMemcpy(d.d_layer_w, h.row_weights.data(), sizeof(int32_t) * h.mrow, hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    fecsim_decoder_free(dec);
    return hip_fail(he, "decoder tables to device");
  }
  he = prepare_generic(d);
  if (he != hipSuccess) {
    fecsim_decoder_free(dec);
    return hip_fail(he, "decoder LDS limit");
  }
  if (d.shape.c2v_global) {
    // one C2V slot per resident workgroup; the kernel strides over the batch
    const int blocks = generic_resident_blocks(d);
    if (blocks < 1) {
      fecsim_decoder_free(dec);
      return fail(FECSIM_E_HIP, "occupancy query failed");
    }
    d.max_grid = (uint32_t)blocks;
    const size_t bytes = (size_t)blocks * d.shape.c2v_stride * d.elem_bytes;
    he = hipMalloc(&d.d_c2v_scratch, bytes);
    if (he != hipSuccess) {
      fecsim_decoder_free(dec);
      return hip_fail(he, "C2V scratch");
    }
  }
  d.code_tab = code_i8_index(d);
  d.freg_shape = float_reg_shape(d);
  d.pair_shape = pair_i8_shape(d);
  if (d.pair_shape >= 0) {
    pair_layout(h, &d.pair);
    he = hipMalloc(&d.d_pair_adr, sizeof(uint32_t) * d.pair.adrtab.size());
    if (he == hipSuccess)
      he = hipMemcpy(d.d_pair_adr, d.pair.adrtab.data(), sizeof(uint32_t) * d.pair.adrtab.size(),
                     hipMemcpyHostToDevice);
    if (he == hipSuccess && d.pair.ct) he = hipMalloc(&d.d_pair_queue, kPairQueueSets * kPairQueueSetBytes);
    if (he != hipSuccess) {
      fecsim_decoder_free(dec);
      return hip_fail(he, "pair kernel tables to device");
    }
  }
  *out = dec;
  return FECSIM_OK;
}

int fecsim_decoder_select_kernel(fecsim_decoder* d, int kernel) {
  if (!d) return fail(FECSIM_E_INVAL, "null argument");
  if (d->holders.load() > 0)
    return fail(FECSIM_E_INVAL, "decoder is held by a queue decoder or simulation loop; select its kernel first");
  if (kernel < FECSIM_KERNEL_AUTO || kernel > FECSIM_KERNEL_PAIR_RT) return fail(FECSIM_E_INVAL, "unknown kernel");
  if ((kernel == FECSIM_KERNEL_PAIR || kernel == FECSIM_KERNEL_PAIR_RT) && d->d.pair_shape < 0)
    return fail(FECSIM_E_UNSUPPORTED, "no row-pair kernel for this code / element / parameters");
  if (kernel == FECSIM_KERNEL_CODE && d->d.code_tab < 0)
    return fail(FECSIM_E_UNSUPPORTED, "no code-specialized kernel for this code / element / parameters");
  if (kernel == FECSIM_KERNEL_FLOAT && d->d.freg_shape < 0)
    return fail(FECSIM_E_UNSUPPORTED, "no register-C2V float kernel for this code / element");
  d->d.kernel = kernel;
  return FECSIM_OK;
}

// the specialized kernel a dense decode would run (FECSIM_KERNEL_GENERIC if none);
// `aligned`: the buffers are 16-byte aligned (the code kernel moves 16-byte chunks)
static int dense_kernel(const DecoderDev& d, bool aligned) {
  if (d.kernel == FECSIM_KERNEL_GENERIC) return FECSIM_KERNEL_GENERIC;
  if (d.pair_shape >= 0 && (!d.d_beta_tab || d.d_beta_tab8))
    return d.kernel == FECSIM_KERNEL_PAIR_RT ? FECSIM_KERNEL_PAIR_RT : FECSIM_KERNEL_PAIR;
  if (d.code_tab >= 0 && aligned && (!d.d_beta_tab || d.beta_tab_code_ok)) return FECSIM_KERNEL_CODE;
  if (d.freg_shape >= 0) return FECSIM_KERNEL_FLOAT;
  return FECSIM_KERNEL_GENERIC;
}

int fecsim_decoder_kernel(const fecsim_decoder* d) {
  if (!d) return fail(FECSIM_E_INVAL, "null argument");
  return dense_kernel(d->d, true);
}

int fecsim_decoder_set_offsets(fecsim_decoder* d, const float* beta) {
  if (!d) return fail(FECSIM_E_INVAL, "null argument");
  if (d->holders.load() > 0)
    return fail(FECSIM_E_INVAL, "decoder is held by a queue decoder or simulation loop; set its offsets first");
  DecoderDev& dv = d->d;
  if (!beta) {
    if (dv.d_beta_tab) (void)hipFree(dv.d_beta_tab);
    if (dv.d_beta_tab8) (void)hipFree(dv.d_beta_tab8);
    dv.d_beta_tab = nullptr;
    dv.d_beta_tab8 = nullptr;
    return FECSIM_OK;
  }
  const size_t n = (size_t)dv.params.max_iter * dv.h.mrow;
  std::vector<long long> tab;
  offsets_table(dv, beta, n, &tab);
  dv.beta_tab_code_ok = dv.elem == FECSIM_I8 && code_i8_offsets_ok(tab);
  void* p = dv.d_beta_tab;
  hipError_t he = p ? hipSuccess : hipMalloc(&p, n * sizeof(long long));
  if (he == hipSuccess) he = hipMemcpy(p, tab.data(), n * sizeof(long long), hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    if (p && !dv.d_beta_tab) (void)hipFree(p);
    return hip_fail(he, "offset table to device");
  }
  dv.d_beta_tab = p;
  // the row-pair kernel's int8 copy (its entries are the int8 offsets themselves, 0..127)
  bool pair_ok = dv.elem == FECSIM_I8 && pair_i8_table_ok(dv, n);
  for (long long b : tab) pair_ok = pair_ok && b >= 0 && b <= 127;
  if (pair_ok) {
    std::vector<int8_t> t8(tab.begin(), tab.end());
    int8_t* q = dv.d_beta_tab8;
    he = q ? hipSuccess : hipMalloc((void**)&q, n);
    if (he == hipSuccess) he = hipMemcpy(q, t8.data(), n, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
      if (q && !dv.d_beta_tab8) (void)hipFree(q);
      return hip_fail(he, "offset table to device");
    }
    dv.d_beta_tab8 = q;
  } else if (dv.d_beta_tab8) {
    (void)hipFree(dv.d_beta_tab8);
    dv.d_beta_tab8 = nullptr;
  }
  return FECSIM_OK;
}

int fecsim_decoder_info(const fecsim_decoder* d, int* elem, int* mode, int* elem_bytes) {
  if (!d) return fail(FECSIM_E_INVAL, "null argument");
  if (elem) *elem = d->d.elem;
  if (mode) *mode = d->d.mode;
  if (elem_bytes) *elem_bytes = d->d.elem_bytes;
  return FECSIM_OK;
}

void fecsim_decoder_free(fecsim_decoder* d) {
  if (!d) return;
  if (d->d.d_edges) (void)hipFree(d->d.d_edges);
  if (d->d.d_layer_w) (void)hipFree(d->d.d_layer_w);
  if (d->d.d_pair_adr) (void)hipFree(d->d.d_pair_adr);
  if (d->d.d_pair_queue) (void)hipFree(d->d.d_pair_queue);
  if (d->d.d_c2v_scratch) (void)hipFree(d->d.d_c2v_scratch);
  if (d->d.d_beta_tab) (void)hipFree(d->d.d_beta_tab);
  if (d->d.d_beta_tab8) (void)hipFree(d->d.d_beta_tab8);
  delete d;
}

int fecsim_decode(fecsim_decoder* dec, uint32_t num_cw, int layout_p, const void* llr_in, void* post_out,
                  uint32_t* iters, void* c2v, unsigned long long* tot_iter, void* stream) {
  if (!dec || ((!llr_in || !post_out) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  if (num_cw == 0) return FECSIM_OK;
  const DecoderDev& d = dec->d;
  const HdecHost& h = d.h;
  KernelArgs a{};
  fill_core(d, &a.core);
  a.num_cw = num_cw;
  a.metrics_in = llr_in;
  a.metrics_out = post_out;
  layout_strides(h.z, h.mcol, layout_p, &a.cw_stride, &a.col_stride, &a.base);
  a.iters = iters;
  a.c2v = c2v;
  a.tot_iter = tot_iter;
  const bool aligned = ((reinterpret_cast<uintptr_t>(llr_in) | reinterpret_cast<uintptr_t>(post_out)) & 15u) == 0;
  const int kern = layout_p == 0 ? dense_kernel(d, aligned) : FECSIM_KERNEL_GENERIC;
  hipError_t he = kern == FECSIM_KERNEL_PAIR || kern == FECSIM_KERNEL_PAIR_RT
                      ? launch_pair_i8(d, a, (hipStream_t)stream, kern == FECSIM_KERNEL_PAIR_RT)
                  : kern == FECSIM_KERNEL_CODE  ? launch_code_i8(d, a, (hipStream_t)stream)
                  : kern == FECSIM_KERNEL_FLOAT ? launch_float_reg(d, a, (hipStream_t)stream)
                                                : launch_decode(d, a, (hipStream_t)stream);
  if (he != hipSuccess) return hip_fail(he, "decode launch");
  return FECSIM_OK;
}

int fecsim_decode_set(fecsim_decoder* dec, uint32_t num_cw, int layout_p, void* metrics, uint32_t* iters,
                      void* c2v, unsigned long long* tot_iter, void* stream) {
  return fecsim_decode(dec, num_cw, layout_p, metrics, metrics, iters, c2v, tot_iter, stream);
}

int fecsim_decode_set_host(fecsim_decoder* dec, uint32_t num_cw, int layout_p, void* metrics_host,
                           uint32_t* iters_host, unsigned long long* tot_iter) {
  if (!dec || (!metrics_host && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  const HdecHost& h = dec->d.h;
  size_t elems;
  if (layout_p == 0) {
    elems = (size_t)num_cw * h.mcol * h.z;
  } else {
    const size_t spz = (h.z + layout_p - 1) / layout_p;
    elems = (size_t)num_cw * h.mcol * (spz + 2) * layout_p;
  }
  const size_t bytes = elems * dec->d.elem_bytes;
  void* dm = nullptr;
  uint32_t* di = nullptr;
  unsigned long long* dt = nullptr;
  hipError_t he = hipMalloc(&dm, bytes ? bytes : 1);
  if (he == hipSuccess) he = hipMalloc((void**)&di, sizeof(uint32_t) * (num_cw ? num_cw : 1));
  if (he == hipSuccess) he = hipMalloc((void**)&dt, sizeof(unsigned long long));
  if (he == hipSuccess) he = hipMemcpy(dm, metrics_host, bytes, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemset(dt, 0, sizeof(unsigned long long));
  int rc = FECSIM_OK;
  if (he == hipSuccess) {
    rc = fecsim_decode_set(dec, num_cw, layout_p, dm, di, nullptr, dt, nullptr);
    if (rc == FECSIM_OK) {
      he = hipMemcpy(metrics_host, dm, bytes, hipMemcpyDeviceToHost);
      unsigned long long t = 0;
      if (he == hipSuccess) he = hipMemcpy(&t, dt, sizeof(t), hipMemcpyDeviceToHost);
      if (he == hipSuccess && iters_host && num_cw)
        he = hipMemcpy(iters_host, di, sizeof(uint32_t) * num_cw, hipMemcpyDeviceToHost);
      if (tot_iter) *tot_iter = t;
    }
  }
  if (dm) (void)hipFree(dm);
  if (di) (void)hipFree(di);
  if (dt) (void)hipFree(dt);
  if (rc) return rc;
  if (he != hipSuccess) return hip_fail(he, "decode_set_host");
  return FECSIM_OK;
}

int fecsim_sp_math(int fn, const float* x, float* y, uint32_t n, void* stream) {
  if (fn != 0 && fn != 1) return fail(FECSIM_E_INVAL, "fn must be 0 (tanh(x/2)) or 1 (2 atanh(x))");
  if ((!x || !y) && n) return fail(FECSIM_E_INVAL, "null argument");
  const hipError_t he = launch_sp_math(fn, x, y, n, (hipStream_t)stream);
  return he == hipSuccess ? FECSIM_OK : hip_fail(he, "sp_math launch");
}

}  // extern "C"

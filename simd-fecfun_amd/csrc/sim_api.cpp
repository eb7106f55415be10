// C ABI of the encoder, channel, error-count and simulation-loop entry points
// (declared in include/fecsim.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace fecsim;

struct fecsim_sim {
  fecsim_henc* h = nullptr;
  fecsim_decoder* d = nullptr;
  bool holds_d = false;                     // counted in d->holders
  uint32_t num_cw = 0;
  std::vector<float> snr, sigma, factor;
  float llr_scale = 0, max_abs = 0;
  uint64_t seed = 0;
  int use_graph = 0;
  uint32_t* bits = nullptr;                 // [num_cw][mcol][zw]
  void* llr = nullptr;                      // [num_cw][mcol*z] decoder elements
  unsigned long long* ctr = nullptr;        // [num_snr][5] + codeword offset
  uint32_t* iters = nullptr;                // [num_cw] per-codeword iterations of the last decode
  unsigned long long* hist = nullptr;       // [num_snr][nbins] iteration histograms
  uint32_t nbins = 0;                       // max_iter + 1, at most 1024
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

constexpr int kEncMaxZ = 512;  // one 512-bit ENC_CLASS word per column (enc_defs.h)
constexpr int kStat = 5;       // cw, uncoded, iters, bit errors, codeword errors

// Parse the int32 words of LDPCencode.raw.dat (encode_matrix_allocate.cpp:51-84) and
// decode the schedule's rotation-area addresses col*1024 + (shift&7)*128 + (shift>>3)
// (gen_encode.m:38,57).
int parse_henc(const int32_t* w, size_t n, HencHost* out) {
  if (n < 6) return fail(FECSIM_E_FORMAT, "encoder table: header truncated (Error in fread of initial encode integers)");
  HencHost& h = *out;
  h.z = w[0];
  h.punc_cnt = w[1];
  h.fill_cnt = w[2];
  h.num_par = w[3];
  h.mrow = w[4];
  h.mcol = w[5];
  if (h.z <= 0 || h.z > kEncMaxZ || h.mrow <= 0 || h.mcol <= h.mrow || h.mcol > 4096 || h.num_par < 1 ||
      h.num_par > 64 || h.punc_cnt < 0 || h.fill_cnt < 0 || (h.punc_cnt & 1) || (h.fill_cnt & 1))
    return fail(FECSIM_E_FORMAT, "encoder table: bad header (z must be 1..512, num_par 1..64, even list lengths)");
  size_t pos = 6;
  if (n < pos + 2 * (size_t)h.mrow) return fail(FECSIM_E_FORMAT, "encoder table: truncated");
  h.par_cnts.assign(w + pos, w + pos + h.mrow);
  pos += h.mrow;
  h.par_colnum.assign(w + pos, w + pos + h.mrow);
  pos += h.mrow;
  long long nsrc = 0;
  for (int i = 0; i < h.mrow; ++i) {
    if (h.par_cnts[i] < 0 || h.par_cnts[i] > 65536) return fail(FECSIM_E_FORMAT, "encoder table: bad source count");
    nsrc += h.par_cnts[i];
  }
  if (n < pos + (size_t)nsrc + (size_t)h.punc_cnt + (size_t)h.fill_cnt)
    return fail(FECSIM_E_FORMAT, "encoder table: truncated");
  h.par_offset.assign(w + pos, w + pos + nsrc);
  pos += nsrc;
  h.punc_list.assign(w + pos, w + pos + h.punc_cnt);
  pos += h.punc_cnt;
  h.fill_list.assign(w + pos, w + pos + h.fill_cnt);
  const int ndata = h.mcol - h.mrow;
  std::vector<char> done(h.mrow, 0);
  h.src.resize(nsrc);
  size_t idx = 0;
  for (int i = 0; i < h.mrow; ++i) {
    const int pc = h.par_colnum[i];
    if (pc < 0 || pc >= h.mrow || done[pc]) return fail(FECSIM_E_FORMAT, "encoder table: parity columns are not a permutation");
    for (int k = 0; k < h.par_cnts[i]; ++k, ++idx) {
      const int32_t a = h.par_offset[idx];
      if (a < 0) return fail(FECSIM_E_FORMAT, "encoder table: negative rotation address");
      const int col = a / 1024, rem = a % 1024, mod8 = rem / 128, byte = rem % 128;
      const int shift = byte * 8 + mod8;
      if (col >= h.mcol || shift >= h.z) return fail(FECSIM_E_FORMAT, "encoder table: source outside the code");
      if (col >= ndata && !done[col - ndata])
        return fail(FECSIM_E_FORMAT, "encoder table: parity column used before it is computed");
      h.src[idx] = ((uint32_t)col << 16) | (uint32_t)shift;
    }
    done[pc] = 1;
  }
  h.zw = (h.z + 31) / 32;
  // enc_fill_mask (encode_matrix_allocate.cpp:13-36): inclusive ranges toggle bits
  auto masks = [&](const std::vector<int32_t>& lst, std::vector<uint32_t>& m, int* bits) -> int {
    m.assign((size_t)h.mcol * h.zw, 0u);
    long long sum = 0;
    for (size_t i = 0; i + 1 < lst.size(); i += 2) {
      const long long s = lst[i], e = lst[i + 1];
      if (s < 0 || e < s || e >= (long long)h.mcol * h.z) return 1;
      sum += e - s + 1;
      for (long long x = s; x <= e; ++x) {
        const int col = (int)(x / h.z), r = (int)(x % h.z);
        m[(size_t)col * h.zw + (r >> 5)] ^= 1u << (r & 31);
      }
    }
    *bits = (int)sum;
    return 0;
  };
  if (masks(h.punc_list, h.punct_words, &h.punc_bits) || masks(h.fill_list, h.fill_words, &h.fill_bits))
    return fail(FECSIM_E_FORMAT, "encoder table: puncture/filler range outside the codeword");
  return FECSIM_OK;
}

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& v) {
  hipError_t e = hipMalloc((void**)dst, sizeof(T) * (v.empty() ? 1 : v.size()));
  if (e == hipSuccess && !v.empty()) e = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return e;
}

void free_tables(fecsim_henc* h) {
  (void)hipFree(h->d_cnt);
  (void)hipFree(h->d_colnum);
  (void)hipFree(h->d_src);
  (void)hipFree(h->d_punct);
  (void)hipFree(h->d_fill);
  h->d_cnt = h->d_colnum = nullptr;
  h->d_src = h->d_punct = h->d_fill = nullptr;
  h->device = -1;
}

}  // namespace

namespace fecsim {

int ensure_device(fecsim_henc* h) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (h->device == dev) return FECSIM_OK;
  if (h->device >= 0) return fail(FECSIM_E_UNSUPPORTED, "encoder tables live on another device");
  e = upload(&h->d_cnt, h->h.par_cnts);
  if (e == hipSuccess) e = upload(&h->d_colnum, h->h.par_colnum);
  if (e == hipSuccess) e = upload(&h->d_src, h->h.src);
  if (e == hipSuccess) e = upload(&h->d_punct, h->h.punct_words);
  if (e == hipSuccess) e = upload(&h->d_fill, h->h.fill_words);
  if (e != hipSuccess) {
    free_tables(h);
    return hip_fail(e, "encoder tables to device");
  }
  h->device = dev;
  return FECSIM_OK;
}

int enc_args(fecsim_henc* h, int gen, uint64_t seed, uint64_t cw0, const unsigned long long* cw0_dev,
             uint32_t num_cw, int col_words, uint32_t* bits, EncArgs* a) {
  const HencHost& hh = h->h;
  *a = EncArgs{};
  if (encode_shape(hh.z, hh.mcol, a)) return fail(FECSIM_E_UNSUPPORTED, "codeword does not fit the encoder's LDS slice");
  a->cnt = h->d_cnt;
  a->colnum = h->d_colnum;
  a->src = h->d_src;
  a->fill = h->d_fill;
  a->z = hh.z;
  a->mrow = hh.mrow;
  a->mcol = hh.mcol;
  a->ndata = hh.mcol - hh.mrow;
  a->col_words = col_words;
  a->gen = gen;
  a->seed = seed;
  a->cw0 = cw0;
  a->cw0_dev = cw0_dev;
  a->num_cw = num_cw;
  a->bits = bits;
  return FECSIM_OK;
}

void chan_sigma(float snr, float llr_scale, float* sigma, float* factor) {
  // gen_metrics.cpp:28-30, evaluated in the same float/double mix
  const float three_db = 10 * log10f(2.0);
  const float s = powf(10.0, ((snr + three_db) / -20.0));
  *sigma = s;
  *factor = 2 * llr_scale / (s * s);
}

int llr_args(fecsim_henc* h, float sigma, float factor, float max_abs, uint64_t seed, uint64_t cw0,
             const unsigned long long* cw0_dev, uint32_t num_cw, int col_words, const uint32_t* bits, int layout_p,
             void* llr, unsigned long long* uncoded, LlrArgs* a) {
  const HencHost& hh = h->h;
  *a = LlrArgs{};
  a->bits = bits;
  a->punct = h->d_punct;
  a->fill = h->d_fill;
  a->z = hh.z;
  a->zw = hh.zw;
  a->mcol = hh.mcol;
  a->hm = (hh.mcol + 1) / 2;
  a->col_words = col_words;
  a->seed = seed;
  a->cw0 = cw0;
  a->cw0_dev = cw0_dev;
  a->cw_off = 0;
  a->num_cw = num_cw;
  a->sigma = sigma;
  a->factor = factor;
  a->max_abs = max_abs;
  a->out = llr;
  layout_strides(hh.z, hh.mcol, layout_p, &a->cw_stride, &a->col_stride, &a->base);
  a->uncoded = uncoded;
  a->has_mask = hh.punc_bits > 0 || hh.fill_bits > 0;
  return FECSIM_OK;
}

void count_args(const fecsim_henc* h, uint32_t num_cw, int col_words, const uint32_t* bits, int layout_p,
                const void* post, unsigned long long* counts, uint32_t* cw_bit_err, unsigned long long* cw_counter,
                CountArgs* a) {
  const HencHost& hh = h->h;
  *a = CountArgs{};
  a->bits = bits;
  a->col_words = col_words;
  a->z = hh.z;
  a->mcol = hh.mcol;
  a->ndata = hh.mcol - hh.mrow;
  a->num_par = hh.num_par;
  a->num_cw = num_cw;
  a->post = post;
  layout_strides(hh.z, hh.mcol, layout_p, &a->cw_stride, &a->col_stride, &a->base);
  a->counts = counts;
  a->cw_bit_err = cw_bit_err;
  a->cw_counter = cw_counter;
}

}  // namespace fecsim

namespace {

bool elem_ok(int elem) { return elem >= FECSIM_I8 && elem <= FECSIM_F64; }

// one inner loop of the simulation on `st`
int sim_enqueue(fecsim_sim* s, hipStream_t st) {
  fecsim_henc* h = s->h;
  const HencHost& hh = h->h;
  const int nsnr = (int)s->snr.size();
  unsigned long long* cw0_dev = s->ctr + (size_t)nsnr * kStat;
  EncArgs ea;
  int rc = enc_args(h, 1, s->seed, 0, cw0_dev, s->num_cw, hh.zw, s->bits, &ea);
  if (rc) return rc;
  hipError_t e = launch_encode(ea, st);
  if (e != hipSuccess) return hip_fail(e, "encode launch");
  const int elem = s->d->d.elem;
  for (int i = 0; i < nsnr; ++i) {
    unsigned long long* c = s->ctr + (size_t)i * kStat;
    LlrArgs la;
    llr_args(h, s->sigma[i], s->factor[i], s->max_abs, s->seed, 0, cw0_dev, s->num_cw, hh.zw, s->bits, 0, s->llr,
             c + 1, &la);
    e = launch_llr(elem, la, st);
    if (e != hipSuccess) return hip_fail(e, "llr launch");
    rc = fecsim_decode_set(s->d, s->num_cw, 0, s->llr, s->iters, nullptr, c + 2, st);
    if (rc) return rc;
    e = launch_iter_hist(s->iters, s->num_cw, s->nbins, s->hist + (size_t)i * s->nbins, st);
    if (e != hipSuccess) return hip_fail(e, "histogram launch");
    CountArgs ca;
    count_args(h, s->num_cw, hh.zw, s->bits, 0, s->llr, c + 3, nullptr, c + 0, &ca);
    e = launch_count(elem, ca, st);
    if (e != hipSuccess) return hip_fail(e, "count launch");
  }
  e = launch_advance(cw0_dev, s->num_cw, st);
  if (e != hipSuccess) return hip_fail(e, "advance launch");
  return FECSIM_OK;
}

}  // namespace

extern "C" {

int fecsim_henc_from_words(const int32_t* words, size_t nwords, fecsim_henc** out) {
  if (!words || !out) return fail(FECSIM_E_INVAL, "null argument");
  fecsim_henc* h = new (std::nothrow) fecsim_henc();
  if (!h) return fail(FECSIM_E_NOMEM, "host allocation");
  const int rc = parse_henc(words, nwords, &h->h);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return FECSIM_OK;
}

int fecsim_henc_load(const char* path, fecsim_henc** out) {
  if (!path || !out) return fail(FECSIM_E_INVAL, "null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(FECSIM_E_IO, std::string("Cannot open file ") + path + " for reading");
  std::vector<int32_t> w;
  int32_t buf[4096];
  size_t got;
  while ((got = fread(buf, sizeof(int32_t), 4096, f)) > 0) w.insert(w.end(), buf, buf + got);
  fclose(f);
  return fecsim_henc_from_words(w.data(), w.size(), out);
}

int fecsim_henc_info(const fecsim_henc* h, fecsim_henc_info_t* info) {
  if (!h || !info) return fail(FECSIM_E_INVAL, "null argument");
  const HencHost& hh = h->h;
  info->z = hh.z;
  info->mrow = hh.mrow;
  info->mcol = hh.mcol;
  info->num_par = hh.num_par;
  info->punc_cnt = hh.punc_cnt;
  info->fill_cnt = hh.fill_cnt;
  info->punc_bits = hh.punc_bits;
  info->fill_bits = hh.fill_bits;
  info->col_words = hh.zw;
  info->num_src = (int32_t)hh.src.size();
  return FECSIM_OK;
}

int fecsim_henc_masks(const fecsim_henc* h, uint32_t* punct, uint32_t* fill) {
  if (!h) return fail(FECSIM_E_INVAL, "null argument");
  const size_t n = h->h.punct_words.size();
  if (punct) memcpy(punct, h->h.punct_words.data(), n * sizeof(uint32_t));
  if (fill) memcpy(fill, h->h.fill_words.data(), n * sizeof(uint32_t));
  return FECSIM_OK;
}

void fecsim_henc_free(fecsim_henc* h) {
  if (!h) return;
  if (h->device >= 0) free_tables(h);
  delete h;
}

int fecsim_gen_codewords(fecsim_henc* h, uint64_t seed, uint64_t cw0, uint32_t num_cw, int col_words, uint32_t* bits,
                         void* stream) {
  if (!h || (!bits && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (!num_cw) return FECSIM_OK;
  int rc = ensure_device(h);
  if (rc) return rc;
  EncArgs a;
  rc = enc_args(h, 1, seed, cw0, nullptr, num_cw, col_words, bits, &a);
  if (rc) return rc;
  const hipError_t e = launch_encode(a, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "encode launch");
}

int fecsim_encode(fecsim_henc* h, uint32_t num_cw, int col_words, uint32_t* bits, void* stream) {
  if (!h || (!bits && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (!num_cw) return FECSIM_OK;
  int rc = ensure_device(h);
  if (rc) return rc;
  EncArgs a;
  rc = enc_args(h, 0, 0, 0, nullptr, num_cw, col_words, bits, &a);
  if (rc) return rc;
  const hipError_t e = launch_encode(a, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "encode launch");
}

int fecsim_chan_sigma(const fecsim_chan_params* p, float* sigma, float* factor) {
  if (!p || !sigma || !factor) return fail(FECSIM_E_INVAL, "null argument");
  chan_sigma(p->snr_db, p->llr_scale, sigma, factor);
  return FECSIM_OK;
}

int fecsim_gen_llr(fecsim_henc* h, const fecsim_chan_params* p, uint64_t seed, uint64_t cw0, uint32_t num_cw,
                   int col_words, const uint32_t* bits, int elem, int layout_p, void* llr,
                   unsigned long long* uncoded_err, void* stream) {
  if (!h || !p || ((!bits || !llr) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (!elem_ok(elem)) return fail(FECSIM_E_INVAL, "unknown element kind");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  if (!num_cw) return FECSIM_OK;
  int rc = ensure_device(h);
  if (rc) return rc;
  float sigma, factor;
  chan_sigma(p->snr_db, p->llr_scale, &sigma, &factor);
  LlrArgs a;
  llr_args(h, sigma, factor, p->max_abs_llr, seed, cw0, nullptr, num_cw, col_words, bits, layout_p, llr, uncoded_err,
           &a);
  const hipError_t e = launch_llr(elem, a, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "llr launch");
}

int fecsim_gen_llr_var(fecsim_henc* h, const float* sigma_factor, float max_abs_llr, uint64_t seed, uint64_t cw0,
                       uint32_t num_cw, int col_words, const uint32_t* bits, int elem, int layout_p, void* llr,
                       unsigned long long* uncoded_err, void* stream) {
  if (!h || ((!bits || !llr || !sigma_factor) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (!elem_ok(elem)) return fail(FECSIM_E_INVAL, "unknown element kind");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  if (!num_cw) return FECSIM_OK;
  int rc = ensure_device(h);
  if (rc) return rc;
  LlrArgs a;
  llr_args(h, 0.0f, 0.0f, max_abs_llr, seed, cw0, nullptr, num_cw, col_words, bits, layout_p, llr, uncoded_err, &a);
  a.sf = sigma_factor;
  const hipError_t e = launch_llr(elem, a, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "llr launch");
}

int fecsim_count_errors(fecsim_henc* h, uint32_t num_cw, int col_words, const uint32_t* bits, int elem, int layout_p,
                        const void* post, unsigned long long* counts, uint32_t* cw_bit_err, void* stream) {
  if (!h || !counts || ((!bits || !post) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (!elem_ok(elem)) return fail(FECSIM_E_INVAL, "unknown element kind");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  if (!num_cw) return FECSIM_OK;
  CountArgs a;
  count_args(h, num_cw, col_words, bits, layout_p, post, counts, cw_bit_err, nullptr, &a);
  const hipError_t e = launch_count(elem, a, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "count launch");
}

void fecsim_sim_free(fecsim_sim* s) {
  if (!s) return;
  if (s->holds_d) s->d->holders.fetch_sub(1);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
  (void)hipFree(s->bits);
  (void)hipFree(s->llr);
  (void)hipFree(s->ctr);
  (void)hipFree(s->iters);
  (void)hipFree(s->hist);
  delete s;
}

int fecsim_sim_create(fecsim_henc* h, fecsim_decoder* d, const fecsim_sim_params* p, fecsim_sim** out) {
  if (!h || !d || !p || !out || (p->num_snr && !p->snr_db)) return fail(FECSIM_E_INVAL, "null argument");
  if (p->num_cw == 0 || p->num_snr == 0 || p->num_snr > 1024) return fail(FECSIM_E_INVAL, "need num_cw >= 1 and 1..1024 SNR points");
  const HencHost& he = h->h;
  const HdecHost& hd = d->d.h;
  if (he.z != hd.z || he.mcol != hd.mcol || he.mrow != hd.mrow)
    return fail(FECSIM_E_INVAL, "encoder and decoder tables describe different codes");
  int rc = ensure_device(h);
  if (rc) return rc;
  fecsim_sim* s = new (std::nothrow) fecsim_sim();
  if (!s) return fail(FECSIM_E_NOMEM, "host allocation");
  s->h = h;
  s->d = d;
  s->num_cw = p->num_cw;
  s->snr.assign(p->snr_db, p->snr_db + p->num_snr);
  s->llr_scale = p->llr_scale;
  s->max_abs = p->max_abs_llr;
  s->seed = p->seed;
  s->use_graph = p->use_graph;
  for (float x : s->snr) {
    float sg, f;
    chan_sigma(x, p->llr_scale, &sg, &f);
    s->sigma.push_back(sg);
    s->factor.push_back(f);
  }
  const size_t nbits = (size_t)p->num_cw * he.mcol * he.zw;
  const size_t nllr = (size_t)p->num_cw * hd.mcol * hd.z * d->d.elem_bytes;
  const size_t nctr = (size_t)p->num_snr * kStat + 1;
  const uint32_t mi = d->d.params.max_iter;
  s->nbins = mi < 1023u ? mi + 1u : 1024u;
  const size_t nhist = (size_t)p->num_snr * s->nbins;
  hipError_t e = hipMalloc((void**)&s->bits, nbits * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&s->llr, nllr);
  if (e == hipSuccess) e = hipMalloc((void**)&s->ctr, nctr * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(s->ctr, 0, nctr * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void**)&s->iters, (size_t)p->num_cw * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&s->hist, nhist * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(s->hist, 0, nhist * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    fecsim_sim_free(s);
    return hip_fail(e, "simulation buffers");
  }
  d->holders.fetch_add(1);
  s->holds_d = true;
  *out = s;
  return FECSIM_OK;
}

int fecsim_sim_run(fecsim_sim* s, uint32_t loops, void* stream) {
  if (!s) return fail(FECSIM_E_INVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  if (!s->use_graph) {
    for (uint32_t i = 0; i < loops; ++i) {
      const int rc = sim_enqueue(s, st);
      if (rc) return rc;
    }
    return FECSIM_OK;
  }
  if (!s->exec) {
    hipError_t e = hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "capture stream");
    e = hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "begin capture");
    const int rc = sim_enqueue(s, s->cap_stream);
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

int fecsim_sim_stats_get(fecsim_sim* s, fecsim_sim_stats* out) {
  if (!s || !out) return fail(FECSIM_E_INVAL, "null argument");
  hipError_t e = hipDeviceSynchronize();
  std::vector<unsigned long long> c(s->snr.size() * kStat);
  if (e == hipSuccess) e = hipMemcpy(c.data(), s->ctr, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "stats copy");
  for (size_t i = 0; i < s->snr.size(); ++i) {
    out[i].cw = c[i * kStat + 0];
    out[i].uncoded_bit_err = c[i * kStat + 1];
    out[i].iters = c[i * kStat + 2];
    out[i].bit_err = c[i * kStat + 3];
    out[i].cw_err = c[i * kStat + 4];
  }
  return FECSIM_OK;
}

int fecsim_sim_reset(fecsim_sim* s, void* stream) {
  if (!s) return fail(FECSIM_E_INVAL, "null argument");
  hipError_t e = launch_fill_u64(s->ctr, (uint32_t)(s->snr.size() * kStat), 0ull, (hipStream_t)stream);
  if (e == hipSuccess) e = launch_fill_u64(s->hist, (uint32_t)(s->snr.size() * s->nbins), 0ull, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "reset launch");
}

int fecsim_sim_iter_hist(fecsim_sim* s, unsigned long long* hist_out, uint32_t nbins) {
  if (!s || !hist_out) return fail(FECSIM_E_INVAL, "null argument");
  if (nbins != s->nbins) return fail(FECSIM_E_INVAL, "nbins must be the decoder's max_iter + 1 (at most 1024)");
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(hist_out, s->hist, s->snr.size() * s->nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "histogram copy");
}

int fecsim_hard_decisions(fecsim_henc* h, uint32_t num_cw, int col_words, int elem, int layout_p, const void* post,
                          uint32_t* bits_out, void* stream) {
  if (!h || ((!post || !bits_out) && num_cw)) return fail(FECSIM_E_INVAL, "null argument");
  if (col_words < h->h.zw || col_words > 4096) return fail(FECSIM_E_INVAL, "col_words below ceil(z/32)");
  if (!elem_ok(elem)) return fail(FECSIM_E_INVAL, "unknown element kind");
  if (layout_p < 0 || layout_p > 4096) return fail(FECSIM_E_INVAL, "layout_p outside 0..4096");
  if (!num_cw) return FECSIM_OK;
  CountArgs a;
  count_args(h, num_cw, col_words, nullptr, layout_p, post, nullptr, nullptr, nullptr, &a);
  const hipError_t e = launch_hard(elem, a, bits_out, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "hard-decision launch");
}

int fecsim_iter_histogram(const uint32_t* iters, uint32_t num_cw, uint32_t nbins, unsigned long long* hist,
                          void* stream) {
  if ((!iters && num_cw) || !hist) return fail(FECSIM_E_INVAL, "null argument");
  if (nbins < 1 || nbins > 1024) return fail(FECSIM_E_INVAL, "nbins outside 1..1024");
  const hipError_t e = launch_iter_hist(iters, num_cw, nbins, hist, (hipStream_t)stream);
  return e == hipSuccess ? FECSIM_OK : hip_fail(e, "histogram launch");
}

}  // extern "C"

// Drop-in replacement of the reference's simulation worker LDPCtop_Vec<P><t>_<mode>
// (src/Top/LDPCtop.cpp:41-139) running on the GPU through libfecsim.
//
// Protocol kept from the reference, so build/run_LDPCsim.pl drives it unchanged:
//   stdin   the parameter blob of run_LDPCsim.pl:79-82 (pack 'L6 f7 Q64 Z128 Z128'),
//           read field by field as read_sim_params.cpp:12-27 does
//   argv[0] the executable name selects the decoder (LDPCtop_Vec64c_min, ...); a name
//           without "Vec" takes the kind from argv[1]
//   stdout  "CW Count: ...", optional "rand seed: ..." lines, "Matrix Params: ..." and
//           one "loop %d: SNR=%f  inp_bit_err=... tot_dec_ticks=..." line per SNR point
//           and outer loop (LDPCtop.cpp:127-129); tot_dec_ticks are TSC ticks around
//           the decode call (LDPCtop.cpp:107-110), the decode being synchronised
//
// Data bits and noise come from the library's counter-based generator keyed by a seed
// folded from the 64 rand_state words (SplitMix64 output function over them in order)
// and the codeword index (outer * loop_count + inner) * num_codeword + i.
#include <hip/hip_runtime.h>
#include <x86intrin.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/fecsim.h"

namespace {

struct SimParams {   // sim_struct (sim_struct_LDPC.h), in blob order
  uint32_t outer_count, loop_count, num_codeword, max_iter;
  int32_t puncture_8023ca_flag, use_given_rand_state;
  float max_row_metric, llr_scaling_factor, max_abs_llr, beta_offset, snr_start, snr_end, snr_step;
  uint64_t rand_state[64];
  char enc_raw_file[128], dec_raw_file[128];
};

uint64_t mix64(uint64_t z) {   // SplitMix64 output function (Steele, Lea, Flood 2014)
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

void die(const char* what, int rc) {
  fprintf(stderr, "%s: %s (%d) %s\n", what, fecsim_strerror(rc), rc, fecsim_last_error());
  exit(1);
}

#define CK(call)                  \
  do {                            \
    const int rc_ = (call);       \
    if (rc_) die(#call, rc_);     \
  } while (0)

#define HK(call)                                                              \
  do {                                                                        \
    const hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));              \
      exit(1);                                                                \
    }                                                                         \
  } while (0)

bool read_params(SimParams* s) {
  // read_sim_params.cpp:12-27: one fread per field, 4+2+7+64+128+128 items
  size_t n = 0;
  n += fread(&s->outer_count, 4, 1, stdin);
  n += fread(&s->loop_count, 4, 1, stdin);
  n += fread(&s->num_codeword, 4, 1, stdin);
  n += fread(&s->max_iter, 4, 1, stdin);
  n += fread(&s->puncture_8023ca_flag, 4, 1, stdin);
  n += fread(&s->use_given_rand_state, 4, 1, stdin);
  n += fread(&s->max_row_metric, 4, 1, stdin);
  n += fread(&s->llr_scaling_factor, 4, 1, stdin);
  n += fread(&s->max_abs_llr, 4, 1, stdin);
  n += fread(&s->beta_offset, 4, 1, stdin);
  n += fread(&s->snr_start, 4, 1, stdin);
  n += fread(&s->snr_end, 4, 1, stdin);
  n += fread(&s->snr_step, 4, 1, stdin);
  n += fread(s->rand_state, 8, 64, stdin);
  n += fread(s->enc_raw_file, 1, 128, stdin);
  n += fread(s->dec_raw_file, 1, 128, stdin);
  s->enc_raw_file[127] = s->dec_raw_file[127] = 0;
  return n == 4 + 2 + 7 + 64 + 128 + 128;
}

}  // namespace

int main(int argc, char** argv) {
  std::string kind = argv[0];
  const size_t slash = kind.find_last_of('/');
  if (slash != std::string::npos) kind = kind.substr(slash + 1);
  if (kind.find("Vec") == std::string::npos) {
    if (argc < 2) {
      fprintf(stderr, "usage: LDPCtop_Vec<P><c|s|w|f|d>_<min|prod> < params, or %s <kind> < params\n", argv[0]);
      return 1;
    }
    kind = argv[1];
  }

  SimParams sp;
  if (!read_params(&sp)) {
    printf("Error in reading sim parameters from stdin!\n");
    return 1;
  }
  printf("CW Count: %d codewords in inner loop * num_codeword\n", (int)(sp.loop_count * sp.num_codeword));
  if (sp.use_given_rand_state != 1) {
    std::random_device rd;
    for (int i = 0; i < 64; ++i) {
      sp.rand_state[i] = ((uint64_t)rd() << 32) ^ rd();
      printf("rand seed: %016lx\n", (unsigned long)sp.rand_state[i]);
    }
  }
  // the reference reads the flag but never applies it (gen_metrics.cpp:47-49,88); a
  // punctured 802.3ca code comes from an encoder file with a puncture list instead
  if (sp.puncture_8023ca_flag)
    fprintf(stderr, "warning: puncture_8023ca_flag is ignored, as in the reference; write the encoder file "
                    "with the 802.3ca puncture list to puncture\n");
  uint64_t seed = 0;
  for (int i = 0; i < 64; ++i) seed = mix64(seed ^ sp.rand_state[i]);

  fecsim_henc* he = nullptr;
  fecsim_hdec* hd = nullptr;
  CK(fecsim_henc_load(sp.enc_raw_file, &he));
  CK(fecsim_hdec_load(sp.dec_raw_file, &hd));
  fecsim_henc_info_t ei;
  CK(fecsim_henc_info(he, &ei));
  printf("Matrix Params: Z=%d  Hcol=%d  Hrow=%d num_par_cw=%d fill_cnt=%d\n", ei.z, ei.mcol, ei.mrow, ei.num_par,
         ei.fill_bits);
  fflush(stdout);
  fecsim_dec_params dp = {sp.max_iter, sp.max_row_metric, sp.beta_offset};
  fecsim_decoder* dec = nullptr;
  CK(fecsim_decoder_create(hd, kind.c_str(), &dp, &dec));
  int elem = 0, mode = 0, eb = 1;
  CK(fecsim_decoder_info(dec, &elem, &mode, &eb));

  const uint32_t ncw = sp.num_codeword;
  const size_t n_el = (size_t)ei.mcol * ei.z;
  uint32_t* bits = nullptr;
  void* llr = nullptr;
  unsigned long long* ctr = nullptr;   // uncoded, iterations, bit errors, codeword errors
  HK(hipMalloc((void**)&bits, sizeof(uint32_t) * (size_t)ncw * ei.mcol * ei.col_words + 4));
  HK(hipMalloc(&llr, n_el * ncw * eb + 16));
  HK(hipMalloc((void**)&ctr, 4 * sizeof(unsigned long long)));
  hipStream_t st;
  HK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));

  const unsigned snr_iters = (unsigned)((sp.snr_end - sp.snr_start + 1e-6) / sp.snr_step) + 1;   // LDPCtop.cpp:77
  std::vector<unsigned> inp(snr_iters), iters(snr_iters), bit_err(snr_iters), cw_err(snr_iters);
  std::vector<uint64_t> ticks(snr_iters);
  for (uint32_t outloop = 0; outloop < sp.outer_count; ++outloop) {
    for (uint32_t inloop = 0; inloop < sp.loop_count; ++inloop) {
      const uint64_t cw0 = ((uint64_t)outloop * sp.loop_count + inloop) * ncw;
      CK(fecsim_gen_codewords(he, seed, cw0, ncw, ei.col_words, bits, st));
      for (unsigned k = 0; k < snr_iters; ++k) {
        const float snr = k * sp.snr_step + sp.snr_start;
        HK(hipMemsetAsync(ctr, 0, 4 * sizeof(unsigned long long), st));
        fecsim_chan_params cp = {snr, sp.llr_scaling_factor, sp.max_abs_llr};
        CK(fecsim_gen_llr(he, &cp, seed, cw0, ncw, ei.col_words, bits, elem, 0, llr, ctr + 0, st));
        HK(hipStreamSynchronize(st));
        const uint64_t t0 = __rdtsc();
        CK(fecsim_decode_set(dec, ncw, 0, llr, nullptr, nullptr, ctr + 1, st));
        HK(hipStreamSynchronize(st));
        const uint64_t t1 = __rdtsc();
        CK(fecsim_count_errors(he, ncw, ei.col_words, bits, elem, 0, llr, ctr + 2, nullptr, st));
        unsigned long long c[4];
        HK(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, st));
        HK(hipStreamSynchronize(st));
        inp[k] += (unsigned)c[0];
        iters[k] += (unsigned)c[1];
        bit_err[k] += (unsigned)c[2];
        cw_err[k] += (unsigned)c[3];
        ticks[k] += t1 - t0;
      }
    }
    for (unsigned k = 0; k < snr_iters; ++k) {   // LDPCtop.cpp:124-136
      const float snr = k * sp.snr_step + sp.snr_start;
      printf("loop %d: SNR=%f  inp_bit_err=%d tot_iter=%d tot_bit_err=%d tot_cw_err=%d tot_dec_ticks=%ld\n", outloop,
             snr, inp[k], iters[k], bit_err[k], cw_err[k], (long)ticks[k]);
      fflush(stdout);
      inp[k] = iters[k] = bit_err[k] = cw_err[k] = 0;
      ticks[k] = 0;
    }
  }
  (void)hipStreamDestroy(st);
  (void)hipFree(bits);
  (void)hipFree(llr);
  (void)hipFree(ctr);
  fecsim_decoder_free(dec);
  fecsim_hdec_free(hd);
  fecsim_henc_free(he);
  return 0;
}

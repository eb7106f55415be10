// Internal structures shared by the C ABI (api.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fecsim.h"

namespace fecsim {

// Host copy of LDPCdecode.raw.dat (dec_struct_LDPC.h:3-24)
struct HdecHost {
  int32_t mrow = 0, mcol = 0, z = 0, total_hc_ones = 0, max_row_weight = 0;
  std::vector<int32_t> row_weights, row_active_cols, row_offset;
  std::vector<int32_t> f2b_copy_cnt, f2b_cols, f2b_offsets, b2f_copy_cnt, b2f_cols, b2f_offsets;
};

// Host copy of LDPCencode.raw.dat (enc_struct_LDPC.h:12-28), schedule decoded
struct HencHost {
  int32_t z = 0, punc_cnt = 0, fill_cnt = 0, num_par = 1, mrow = 0, mcol = 0;
  std::vector<int32_t> par_cnts, par_colnum, par_offset, punc_list, fill_list;
  std::vector<int32_t> src_col, src_shift;       // per schedule entry
  std::vector<uint32_t> punct_words, fill_words; // [mcol][wpc] 32-bit masks, bit r of column at word r>>5
  int wpc = 0;                                    // 32-bit words per column = ceil(z/32)
  int fill_bits = 0;                              // number of filler bits (fill_cnt of "Matrix Params")
};

// Launch shape of the layered decoder: each wave owns `cw_per_wave` codeword
// slots; a slot spreads the Z check rows of a layer over `lanes_per_cw` lanes,
// `rows_per_lane` rows each (lane lr owns rows lr, lr + lanes_per_cw, ...).
struct DecShape {
  int rows_per_lane = 1, lanes_per_cw = 64, cw_per_wave = 1, waves = 4;
  int maxw = 8;
  int post_stride = 0, c2v_stride = 0;  // elements per slot, padded
  int wave_bytes = 0;                    // LDS bytes per wave
};

// What the decode loop needs (device pointers + parameters)
struct DecCore {
  const uint32_t* edges;    // per edge: (col*z) << 16 | shift
  const int32_t* layer_w;   // row weight per layer
  int mrow, mcol, z, n, e;  // layers, columns, lifting size, mcol*z, total_hc_ones*z
  int lanes_per_cw, cw_per_wave, waves, rows_per_lane;
  int post_stride, c2v_stride, wave_bytes;
  uint32_t max_iter;
  long long ibeta, icap;    // integer-element beta and cap (already truncated/wrapped)
  double fbeta, fcap;       // float-element beta and cap
};

struct KernelArgs {
  DecCore core;
  uint32_t num_cw;
  const void* metrics_in;   // channel LLRs
  void* metrics_out;        // posteriors (may equal metrics_in)
  long long cw_stride, col_stride, base;  // layout of the metrics in elements
  uint32_t* iters;
  void* c2v;
  unsigned long long* tot_iter;
};

// Fused simulation step: data bits -> encoder -> BPSK + AWGN -> quantized LLRs ->
// decoder -> error counts, all inside one wave's LDS slice per codeword.
struct SimArgs {
  DecCore core;
  // encoder
  const int32_t* par_cnts;     // [mrow]
  const int32_t* par_colnum;   // [mrow]
  const int32_t* src_col;      // [nsrc]
  const int32_t* src_shift;    // [nsrc]
  const uint32_t* punct;       // [mcol][wpc]
  const uint32_t* fill;        // [mcol][wpc]
  int wpc, dwords;             // words per column; words per doubled column (2*wpc + 1)
  int enc_bytes;               // LDS bytes per slot for the doubled columns
  int num_par;                 // interleaved codewords per super codeword (error slots)
  int ndata;                   // data columns
  // channel
  uint64_t seed;
  unsigned long long cw0;      // global index of the first codeword
  uint32_t num_cw;
  const float* sigma;          // [nsnr]
  const float* factor;         // [nsnr]
  float max_abs;
  int nsnr;
  unsigned long long* counters;   // [nsnr][FECSIM_NCOUNT]
};

// Layout of the row-pair int8 kernel (pair_i8.hip)
struct PairLayout {
  int half = 0, cw_per_wave = 0, col_dwords = 0, slot_dwords = 0, waves = 0;
  std::vector<uint32_t> adrtab;   // [half][total_hc_ones]
};

struct DecoderDev {
  HdecHost h;
  int elem = FECSIM_I8, mode = FECSIM_MIN, elem_bytes = 1;
  fecsim_dec_params params{};
  DecShape shape;
  uint32_t* d_edges = nullptr;
  int32_t* d_layer_w = nullptr;
  int pair_shape = -1;            // compiled code shape for the pair kernel, -1 if none
  PairLayout pair;
  uint32_t* d_pair_adr = nullptr;
  int kernel = FECSIM_KERNEL_AUTO;
};

int pair_i8_shape(const DecoderDev& d);
void pair_layout(const HdecHost& h, PairLayout* pl);
hipError_t launch_pair_i8(const DecoderDev& d, const KernelArgs& a, hipStream_t stream);
int choose_shape(const HdecHost& h, int elem_bytes, int extra_slot_bytes, DecShape* s, std::string* why);
void fill_core(const DecoderDev& d, DecCore* c);
hipError_t launch_decode(const DecoderDev& d, const KernelArgs& a, hipStream_t stream);
hipError_t launch_sim(const DecoderDev& d, const SimArgs& a, hipStream_t stream);

}  // namespace fecsim

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

// Launch shape of the layered decoder: each wave owns `cw_per_wave` codeword
// slots; a slot spreads the Z check rows of a layer over `lanes_per_cw` lanes,
// `rows_per_lane` rows each (lane lr owns rows lr, lr + lanes_per_cw, ...).
struct DecShape {
  int rows_per_lane = 1, lanes_per_cw = 64, cw_per_wave = 1, waves = 4;
  int maxw = 8;
  int post_stride = 0, c2v_stride = 0;  // elements per slot, padded
  int wave_bytes = 0;                    // LDS bytes per wave
};

struct KernelArgs {
  const uint32_t* edges;    // per edge: (col*z) << 16 | shift
  const int32_t* layer_w;   // row weight per layer
  int mrow, z, n, e;        // layers, lifting size, mcol*z, total_hc_ones*z
  int mcol;
  int lanes_per_cw, cw_per_wave, waves, rows_per_lane;
  int post_stride, c2v_stride, wave_bytes;
  uint32_t num_cw, max_iter;
  long long ibeta, icap;    // integer-element beta and cap (already truncated/wrapped)
  double fbeta, fcap;       // float-element beta and cap
  const void* metrics_in;   // channel LLRs
  void* metrics_out;        // posteriors (may equal metrics_in)
  long long cw_stride, col_stride, base;  // layout of `metrics` in elements
  uint32_t* iters;
  void* c2v;
  unsigned long long* tot_iter;
};

struct DecoderDev {
  HdecHost h;
  int elem = FECSIM_I8, mode = FECSIM_MIN, elem_bytes = 1;
  fecsim_dec_params params{};
  DecShape shape;
  uint32_t* d_edges = nullptr;
  int32_t* d_layer_w = nullptr;
};

int choose_shape(const HdecHost& h, int elem_bytes, DecShape* s, std::string* why);
hipError_t launch_decode(const DecoderDev& d, const KernelArgs& a, hipStream_t stream);

}  // namespace fecsim

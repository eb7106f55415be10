// Internal structures shared by the C ABI (api.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
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
  std::vector<uint32_t> src;                      // per schedule entry: (column << 16) | shift
  std::vector<uint32_t> punct_words, fill_words;  // [mcol][zw] masks, row r of a column at word r>>5
  int zw = 0;                                     // 32-bit words per column = ceil(z/32)
  int punc_bits = 0, fill_bits = 0;               // bits set in the masks
};

// Launch shape of the layered decoder: each wave owns `cw_per_wave` codeword
// slots; a slot spreads the Z check rows of a layer over `lanes_per_cw` lanes,
// `rows_per_lane` rows each (lane lr owns rows lr, lr + lanes_per_cw, ...).
struct DecShape {
  int rows_per_lane = 1, lanes_per_cw = 64, cw_per_wave = 1, waves = 4;
  int maxw = 8;
  int post_stride = 0, c2v_stride = 0;  // elements per slot, padded
  int wave_bytes = 0;                    // LDS bytes per wave (per block when wg)
  int tab_bytes = 0;                     // LDS bytes of the edge/layer tables in front of the slices
  bool c2v_global = false;               // C2V in a global scratch (codewords too large for LDS)
  bool wg = false;                       // one codeword per workgroup of wg_threads threads
  int wg_threads = 0;
};

// What the decode loop needs (device pointers + parameters)
struct DecCore {
  const uint32_t* edges;    // per edge: (col*z) << 16 | shift
  const int32_t* layer_w;   // row weight per layer
  int mrow, mcol, z, n, e;  // layers, columns, lifting size, mcol*z, total_hc_ones*z
  int nedge;                // total_hc_ones
  int tab_off;              // word offset of this code's [edges | layer_w] in the block's LDS tables
  int tab_bytes;            // LDS bytes in front of the codeword slices (all codes' tables)
  int lanes_per_cw, cw_per_wave, waves, rows_per_lane;
  int post_stride, c2v_stride, wave_bytes;
  uint32_t max_iter;
  long long ibeta, icap;    // integer-element beta and cap (already truncated/wrapped)
  double fbeta, fcap;       // float-element beta and cap
  void* c2v_scratch;        // [grid slots][c2v_stride] when the shape keeps C2V in global memory
  const void* beta_tab;     // [max_iter][mrow] offsets (long long / double) or null
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

// n / d for n < 2^31 by a 64-bit multiply: q = (n * m) >> s with m = ceil(2^s / d),
// s = 31 + ceil(log2 d); exact because m*d - 2^s < d <= 2^(s-31) (Granlund-Montgomery)
struct FastDiv {
  uint32_t m = 1, s = 0;
};
inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  FastDiv f;
  f.s = 31 + l;
  f.m = (uint32_t)(((1ull << f.s) + d - 1) / d);
  return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
  return (uint32_t)(((uint64_t)n * f.m) >> f.s);
}

// Data bits + encoder (sim.hip, k_encode)
struct EncArgs {
  const int32_t* cnt;        // [mrow] sources per parity column
  const int32_t* colnum;     // [mrow] parity column index (among the parity columns)
  const uint32_t* src;       // [nsrc] (source column << 16) | shift
  const uint32_t* fill;      // [mcol][zw] filler mask words
  int z, zw, mrow, mcol, ndata;
  int cpw, slot_words, waves;  // codewords per wave, LDS words per codeword, waves per block
  int col_words;             // words per column in the caller's bit buffer (>= zw)
  int gen;                   // 1: generate the data columns, 0: read them from bits
  uint64_t seed;
  unsigned long long cw0;    // index of the first codeword (+ *cw0_dev when given)
  const unsigned long long* cw0_dev;
  uint64_t num_cw;
  uint32_t* bits;            // [num_cw][mcol][col_words]
};

// Noise + quantized LLRs (sim.hip, k_llr)
struct LlrArgs {
  const uint32_t* bits;      // [num_cw][mcol][col_words]
  const uint32_t* punct;     // [mcol][zw]
  const uint32_t* fill;      // [mcol][zw]
  int z, zw, mcol, hm, col_words;   // hm = ceil(mcol / 2) column pairs
  uint64_t seed;
  unsigned long long cw0;
  const unsigned long long* cw0_dev;
  uint64_t cw_off;           // first codeword of this launch within the batch
  uint32_t num_cw;
  float sigma, factor, max_abs;
  void* out;                 // LLRs of the decoder's element type
  long long cw_stride, col_stride, base;  // output layout in elements
  unsigned long long* uncoded;            // device counter (added to), may be null
  int has_mask;              // any puncture or filler bit
  FastDiv div_cw, div_z;     // by hm*z and by z (set by the launcher)
  const float* sf;           // per-codeword (sigma, factor) pairs [num_cw][2] replacing sigma/factor, or null
};

// Error counting (sim.hip, k_count)
constexpr int kCountMaxTile = 64;
struct CountArgs {
  const uint32_t* bits;
  int col_words, z, mcol, ndata, num_par;
  int cpb;                   // codewords per block tile (set by the launcher)
  uint64_t num_cw;
  const void* post;
  long long cw_stride, col_stride, base;
  unsigned long long* counts;      // device [2]: bit errors, codeword errors (added to)
  uint32_t* cw_bit_err;            // device [num_cw] or null
  unsigned long long* cw_counter;  // device counter of codewords (added to), or null
  FastDiv div_dz, div_z;           // by ndata*zw and by zw, zw = ceil(z / 32) (set by the launcher)
};

// Differentiable min-sum forward / adjoint (train.hip); F = float or double
struct TrainArgs {
  const uint32_t* edges;     // decoder edge table ((col*z) << 16 | shift)
  const int32_t* layer_w;
  int mrow, z, n, e, n_pad;  // n_pad: LDS elements per posterior array
  int nedge, tab_bytes;      // total_hc_ones; LDS bytes of the [edges | layer_w] copy in front
  int iters;                 // T
  int maxw;                  // largest row weight
  int waves;                 // codewords (waves) per block
  int is_double, global_p;   // element type; posteriors in global scratch instead of LDS
  int quant;                 // float32 emulating the int8 decoder (saturation, integer offsets)
  int reg_shape;             // compiled shape of the register-C2V kernels (train_reg.hip), -1 if none
  double cap;
  uint64_t batch;
  const void* llr;           // [batch][n]
  const void* alpha;         // [T][mrow]
  const void* beta;          // [T][mrow]
  void* post;                // [batch][n]
  int seg;                   // K: checkpoint every K iterations (the K of least adjoint memory)
  void* ckP;                 // [batch][ceil(T/K)][n] posteriors at the start of iterations 0, K, 2K, ...
  void* ckC;                 // [batch][ceil(T/K)][e] C2V at those iterations
  void* wc;                  // [batch][e] working C2V
  void* qs;                  // [batch][K][S] kept state of the segment being reversed: the V2C
                             // messages (S = e), or the register kernels' reverse state
                             // (S = mrow * train_reg_state_words * z)
  void* gc;                  // [batch][e] adjoint of the C2V messages
  void* partial;             // [batch][T][mrow][2] per-codeword parameter gradients
  void* gscratch;            // [batch][2][n] posteriors + adjoints when global_p
  const void* grad_post;     // [batch][n]
  void* grad_alpha;          // [T][mrow]
  void* grad_beta;           // [T][mrow]
  void* grad_llr;            // [batch][n] or null
};

// Mixed-code queue decoder (multi.hip)
struct MultiArgs {
  const DecCore* cores;      // [ncode] per-code tables and shapes (C2V in LDS)
  int ncode;
  uint32_t num_cw;
  const int32_t* code;       // [num_cw] code id of every queue entry
  const unsigned long long* off;  // [num_cw] element offset of its LLRs / posteriors (dense)
  const void* in;
  void* out;
  uint32_t* iters;           // [num_cw] or null
  unsigned long long* tot_iter;
  uint32_t* perm;            // [max_cw] queue indices grouped by code
  uint32_t* cnt;             // [ncode] histogram (zero between calls)
  uint32_t* cursor;          // [ncode]
  uint32_t* code_start;      // [ncode + 1]
  uint32_t* item_start;      // [ncode + 1]
  uint32_t* counter;         // next work item
  uint32_t* take;            // [ncode] codewords of each code's segment taken (k_multi_code)
  int wave_bytes, waves;     // LDS per wave (max over codes), waves per block
  const uint32_t* tabs;      // every code's [edges | layer_w], copied to the front of LDS
  int tab_words, tab_bytes;
  const int32_t* ctab;       // [ncode] compiled code of each decoder (k_multi_code)
};

// k_pair_i8c's counter sets per decoder: that many of its decodes may overlap on the device
constexpr int kPairQueueSets = 8;
constexpr size_t kPairQueueSetBytes = 64 * 256;

// Layout of the row-pair int8 kernel (pair_i8.hip)
struct PairLayout {
  int half = 0, cw_per_wave = 0, col_dwords = 0, wave_dwords = 0, waves = 0;
  bool ct = false;                // the edge table is compiled into the shape's k_pair_i8c
  size_t ct_off = 0;              // k_pair_i8c's per-lane table in adrtab (when ct)
  int lead[4] = {};               // k_pair_i8c: lane of each slot's row 0 and each slot's lanes
  unsigned long long smask[4] = {};
  std::vector<uint32_t> adrtab;   // [half][total_hc_ones], then k_pair_i8c's table
};

struct DecoderDev {
  HdecHost h;
  int elem = FECSIM_I8, mode = FECSIM_MIN, elem_bytes = 1;
  fecsim_dec_params params{};
  DecShape shape;
  uint32_t* d_edges = nullptr;
  int32_t* d_layer_w = nullptr;
  int pair_shape = -1;            // compiled code shape for the pair kernel, -1 if none
  int freg_shape = -1;            // compiled code shape for the register-C2V float kernel, -1 if none
  PairLayout pair;
  uint32_t* d_pair_adr = nullptr;
  unsigned* d_pair_queue = nullptr;  // k_pair_i8c's work-queue counters: kPairQueueSets sets of 64
                                     // (one per 256 bytes), one set per launch in rotation
  mutable unsigned pair_queue_next = 0;
  int kernel = FECSIM_KERNEL_AUTO;
  void* d_c2v_scratch = nullptr;  // generic kernel, c2v_global shapes only
  uint32_t max_grid = 0;          // blocks the generic kernel is launched with at most (0: no cap)
  void* d_beta_tab = nullptr;     // per-(iteration, layer) offsets: long long (int kinds) or double
  int code_tab = -1;              // code compiled into the code-specialized int8 kernel, -1 if none
  bool beta_tab_code_ok = false;  // the offset table suits that kernel (every entry 1..127)
  int8_t* d_beta_tab8 = nullptr;  // the same table as int8 for the row-pair kernel, when it suits it
                                  // (entries 0..127, fixed-Z shape, at most 4096 entries)
};

int pair_i8_shape(const DecoderDev& d);
// register-C2V float32 kernel (csrc/float_reg.hip): the compiled shape that runs decoder
// d (float32, Z <= 256, shape compiled in), -1 if none; and its launch
int float_reg_shape(const DecoderDev& d);
hipError_t launch_float_reg(const DecoderDev& d, const KernelArgs& a, hipStream_t stream);
// the row-pair kernel can run decoder d with an offset table of n entries
bool pair_i8_table_ok(const DecoderDev& d, size_t n);
void pair_layout(const HdecHost& h, PairLayout* pl);
// run_time_table: the instance that reads the edge table at run time even where the
// shape has it compiled in
hipError_t launch_pair_i8(const DecoderDev& d, const KernelArgs& a, hipStream_t stream, bool run_time_table);
// code-specialized int8 kernel (one codeword per workgroup, code compiled in): the code's
// index for decoder d, or -1; whether an offset table suits it; the launch
int code_i8_index(const DecoderDev& d);
bool code_i8_offsets_ok(const std::vector<long long>& tab);
hipError_t launch_code_i8(const DecoderDev& d, const KernelArgs& a, hipStream_t stream);
int choose_shape(const HdecHost& h, int elem_bytes, DecShape* s, std::string* why);
void fill_core(const DecoderDev& d, DecCore* c);
hipError_t launch_decode(const DecoderDev& d, const KernelArgs& a, hipStream_t stream);
// blocks of the generic kernel resident on the whole device for decoder d's shape
int generic_resident_blocks(const DecoderDev& d);
// raises the kernel's dynamic LDS limit when a workgroup shape needs more than 64 KB
hipError_t prepare_generic(const DecoderDev& d);
// device form of a per-(iteration, layer) offset table (long long for integer kinds,
// the bits of a double for float kinds)
void offsets_table(const DecoderDev& d, const float* beta, size_t n, std::vector<long long>* out);
// mixed-code decoder: blocks resident for its kernel, and the launch (sort + decode)
int multi_resident_blocks(int elem, int mode, int maxw, int waves, size_t lds);
// code_kernel: k_multi_code (every decoder's multi_code_index >= 0, a.ctab holds them)
hipError_t launch_multi(int elem, int mode, int maxw, const MultiArgs& a, uint32_t grid, hipStream_t stream,
                        bool code_kernel);
// the queue kernel with the codes compiled in: decoder d's compiled code or -1, its LDS
// per workgroup and resident workgroups on the current device
int multi_code_index(const DecoderDev& d);
size_t multi_code_lds();
int multi_code_resident_blocks();
// a DecShape with `rows_per_lane` rows per lane (0: the best lane utilisation), C2V in LDS;
// returns non-zero when it does not fit
int shape_with_rows(const HdecHost& h, int elem_bytes, int rows_per_lane, DecShape* s);
// forward (bwd = 0) or adjoint + parameter-gradient sum (bwd = 1) of the trainable decoder
hipError_t launch_train(const TrainArgs& a, int bwd, hipStream_t stream);
// waves of that kernel (with a's global_p and waves) resident per CU (0: query failed)
int train_resident_waves(const TrainArgs& a, int bwd);
// register-C2V kernels of the differentiable decoder (train_reg.hip): the compiled shape
// for code h (float32 elements, Z <= 128), or -1; and their launch (forward or adjoint)
int train_reg_shape(const HdecHost& h, int is_double);
hipError_t launch_train_reg(int shape, const TrainArgs& a, int bwd, hipStream_t stream);
// words a row and layer of the register kernels' kept reverse state (quant: int8 emulation)
int train_reg_state_words(int quant);
int encode_shape(int z, int mcol, EncArgs* a);
hipError_t launch_encode(const EncArgs& a, hipStream_t stream);
hipError_t launch_llr(int elem, const LlrArgs& a, hipStream_t stream);
hipError_t launch_count(int elem, const CountArgs& a, hipStream_t stream);
hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t stream);
// the float32 sum-product decoder's tanh(x/2) (fn 0) or 2 atanh(x) (fn 1), elementwise
hipError_t launch_sp_math(int fn, const float* x, float* y, uint32_t n, hipStream_t stream);
hipError_t launch_hard(int elem, const CountArgs& a, uint32_t* out, hipStream_t stream);
hipError_t launch_iter_hist(const uint32_t* iters, uint32_t n, uint32_t nbins, unsigned long long* hist,
                            hipStream_t stream);
hipError_t launch_advance(unsigned long long* p, unsigned long long by, hipStream_t stream);

// error reporting shared by the C ABI translation units (api.cpp)
extern thread_local std::string g_err;
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
extern const int kElemBytes[5];

// element strides of a metrics buffer: dense (layout_p = 0) or the reference's
// padded P-lane layout [cw][mcol][(ceil(z/P)+2)*P] with data at +P
inline void layout_strides(int z, int mcol, int layout_p, long long* cw_stride, long long* col_stride,
                           long long* base) {
  if (layout_p == 0) {
    *col_stride = z;
    *base = 0;
  } else {
    const long long spz = (z + layout_p - 1) / layout_p;
    *col_stride = (spz + 2) * layout_p;
    *base = layout_p;
  }
  *cw_stride = *col_stride * mcol;
}

}  // namespace fecsim

struct fecsim_decoder {
  fecsim::DecoderDev d;
  // queue decoders and simulation loops holding this decoder: they keep its tables,
  // offset table and kernel choice (copied into device cores or a captured graph), so
  // those cannot change while holders > 0
  std::atomic<int> holders{0};
};

struct fecsim_hdec {
  fecsim::HdecHost h;
  int32_t len_f2b = 0, len_b2f = 0;
};

struct fecsim_henc {
  fecsim::HencHost h;
  int device = -1;  // device holding the tables below (-1: not uploaded yet)
  int32_t* d_cnt = nullptr;
  int32_t* d_colnum = nullptr;
  uint32_t* d_src = nullptr;
  uint32_t* d_punct = nullptr;
  uint32_t* d_fill = nullptr;
};

struct fecsim_multi {
  std::vector<fecsim_decoder*> decs;
  int elem = 0, mode = 0, maxw = 8;
  std::vector<fecsim::DecShape> shapes;  // tile shape per code
  int waves_pref = 4;                     // waves per block asked for
  int waves = 4, wave_bytes = 0;          // in use
  uint32_t grid = 0, max_cw = 0;
  fecsim::DecCore* d_cores = nullptr;
  uint32_t* d_ws = nullptr;               // perm | cnt | cursor | code_start | item_start | counter | take
  uint32_t* d_tabs = nullptr;             // all codes' [edges | layer_w]
  std::vector<int> tab_off;               // word offset of each code's tables
  int tab_words = 0, tab_bytes = 0;
  int32_t* d_ctab = nullptr;              // [ncode] compiled code per decoder (k_multi_code)
  bool code_ok = false;                   // every decoder's code is compiled into k_multi_code
  int kernel_pref = FECSIM_KERNEL_AUTO;   // FECSIM_KERNEL_AUTO, _GENERIC or _CODE
  bool use_code = false;                  // k_multi_code runs the queue
  uint32_t code_grid = 0;                 // its persistent grid
};

namespace fecsim {
// encoder / channel / count launch helpers shared by the simulation loops (sim_api.cpp)
int ensure_device(fecsim_henc* h);
int enc_args(fecsim_henc* h, int gen, uint64_t seed, uint64_t cw0, const unsigned long long* cw0_dev,
             uint32_t num_cw, int col_words, uint32_t* bits, EncArgs* a);
void chan_sigma(float snr, float llr_scale, float* sigma, float* factor);
int llr_args(fecsim_henc* h, float sigma, float factor, float max_abs, uint64_t seed, uint64_t cw0,
             const unsigned long long* cw0_dev, uint32_t num_cw, int col_words, const uint32_t* bits, int layout_p,
             void* llr, unsigned long long* uncoded, LlrArgs* a);
void count_args(const fecsim_henc* h, uint32_t num_cw, int col_words, const uint32_t* bits, int layout_p,
                const void* post, unsigned long long* counts, uint32_t* cw_bit_err, unsigned long long* cw_counter,
                CountArgs* a);
}  // namespace fecsim

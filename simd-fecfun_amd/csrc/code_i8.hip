// Code-specialized int8 offset min-sum decoder, one codeword per workgroup, for codes
// with a large lifting size (5G NR base graph 1 at Z = 384, 802.3ca at Z = 256).  The
// whole code is compiled in (code_tables.h, generated by fecsim/gen_code_tables.py); the
// row arithmetic, the LDS layout of a codeword and the barrier plan are code_core.h's.
// Same results as the generic kernel, bit for bit:
//
// * A workgroup of Z threads owns one codeword.  Thread i owns check row
//   r = 4 (i mod Z/4) + i div (Z/4) of every layer, so the 64 lanes of a wave touch 64
//   distinct LDS dwords.
// * A row's C2V messages stay in VGPRs as int8 bytes, four edges per register.
// * A workgroup barrier separates two layers that share a column (a compile-time plan;
//   base graph 1 needs 32 per iteration for its 46 layers), and reads of columns not
//   written since the last barrier are issued before it.  The syndrome after each
//   iteration is the XOR of each row's posterior bytes, and a workgroup OR ends the loop
//   (src/LDPCDec/LDPCdecode.cpp:172-197).
#include <mutex>

#include "code_core.h"
#include "code_tables.h"
#include "decode_core.h"

namespace fecsim {

struct CodeArgs {
  uint32_t num_cw, max_iter;
  int beta, cap;                // int8 beta (1..127) and cap (0..127)
  const long long* beta_tab;    // [max_iter][mrow] per-(iteration, layer) offsets, each 1..127, or null
  const int8_t* llr_in;         // dense [num_cw][mcol*z], 16-byte aligned
  int8_t* post_out;             // same layout (may equal llr_in)
  uint32_t* iters;
  int8_t* c2v;                  // [num_cw][ne*z] (edge-major, row r of edge e at e*z + r) or null
  unsigned long long* tot_iter;
};

namespace {

// the row arithmetic's form per code (code_core.h code_layer): two edges per operation for
// codes whose layers are heavy on average (802.3ca, 22.9 edges: 7.70 -> 7.02 ms per 32768
// codewords), one for the others (base graph 1, 6.9 edges over 46 mostly short layers:
// 17.8 against 18.9 ms per 16384)
template <class CT>
constexpr bool code_i8_packed() {
  return CT::ne >= 12 * CT::mrow;
}

template <class CT>
__global__ void __launch_bounds__(CT::z) __attribute__((amdgpu_waves_per_eu(4))) k_code_i8(const CodeArgs a) {
  constexpr bool PK = code_i8_packed<CT>();
  constexpr uint32_t BIAS = PK ? 0u : 0x80808080u;   // posteriors in LDS: int8 or biased by 128
  constexpr int Z = CT::z, Q = Z / 4, NE = CT::ne, NW = (NE + 3) / 4, MC = CT::mcol;
  constexpr int LDS = (3 * MC + 1) * Z;
  constexpr int NB = (LDS + 32767) / 32768;   // base registers, one per 32 KiB: offsets stay < 32768
  constexpr int CH = MC * Z / 16, CPC = Z / 16;
  static_assert(Z % 16 == 0 && Z <= 1024, "Z: multiple of 16, at most 1024");
  HIP_DYNAMIC_SHARED(uint8_t, smem);
  const int i = threadIdx.x;
  const int r = 4 * (i % Q) + i / Q;
  const uint32_t lbase = (uint32_t)(size_t)(LdsU8*)smem;
  uint32_t vb[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    vb[k] = lbase + (uint32_t)r + 32768u * (uint32_t)k;
    asm volatile("" : "+v"(vb[k]));   // kept as separate bases, not re-folded into one
  }
  unsigned long long tot = 0;

  for (long long cw = blockIdx.x; cw < (long long)a.num_cw; cw += gridDim.x) {   // uniform per block
    // ---- codeword in: both copies of each column ----
    const uint4* gi = reinterpret_cast<const uint4*>(a.llr_in + cw * (long long)(MC * Z));
    for (int k = i; k < CH; k += Z) {
      uint4 v = gi[k];
      v.x ^= BIAS;
      v.y ^= BIAS;
      v.z ^= BIAS;
      v.w ^= BIAS;
      const int c = k / CPC, o = (k - c * CPC) * 16;
      *reinterpret_cast<uint4*>(smem + Z + 3 * Z * c + o) = v;
      *reinterpret_cast<uint4*>(smem + 2 * Z + 3 * Z * c + o) = v;
    }
    __syncthreads();

    uint32_t c2v[NW];   // C2V of edges 4k..4k+3 (row r), int8 bytes; LDPCdecode.cpp:115-119 zeroes them
#pragma unroll
    for (int k = 0; k < NW; ++k) c2v[k] = 0u;
    uint32_t iter = 0;
    for (;;) {
      const long long* bt = a.beta_tab ? a.beta_tab + (size_t)iter * CT::mrow : nullptr;
      int pin0[CT::w[0]] = {};   // layer 0 reads nothing ahead
      code_layers<CT, 0, 0, PK>(vb, c2v, bt, (uint32_t)a.beta, a.cap, pin0, true);
      const uint32_t F = code_syndrome<CT, 0, 0, PK>(vb);   // syndrome after the full iteration
      ++iter;
      const int any = __syncthreads_or((int)(F & 0x80u));
      if (!any || iter == a.max_iter) break;
    }

    // ---- codeword out ----
    uint4* go = reinterpret_cast<uint4*>(a.post_out + cw * (long long)(MC * Z));
    for (int k = i; k < CH; k += Z) {
      const int c = k / CPC, o = (k - c * CPC) * 16;
      uint4 v = *reinterpret_cast<const uint4*>(smem + Z + 3 * Z * c + o);
      v.x ^= BIAS;
      v.y ^= BIAS;
      v.z ^= BIAS;
      v.w ^= BIAS;
      go[k] = v;
    }
    if (a.c2v) {
      int8_t* gc = a.c2v + cw * (long long)(NE * Z) + r;
#pragma unroll
      for (int e = 0; e < NE; ++e) gc[(size_t)e * Z] = (int8_t)c2v_byte(c2v[e >> 2], e & 3);
    }
    if (i == 0) {
      if (a.iters) a.iters[cw] = iter;
      tot += iter;
    }
    __syncthreads();   // LDS reads of this codeword done before the next one is loaded
  }
  if (a.tot_iter && i == 0 && tot) atomicAdd(a.tot_iter, tot);
}

template <class CT>
hipError_t launch_code(const CodeArgs& a, hipStream_t st) {
  if (a.num_cw == 0) return hipSuccess;
  constexpr size_t lds = (size_t)(3 * CT::mcol + 1) * CT::z;
  // the LDS attribute is set once per device (mutex: host threads may launch
  // concurrently on distinct handles)
  static std::mutex mu;
  static bool ready_by_dev[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!ready_by_dev[dev]) {
      if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_code_i8<CT>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
      }
      ready_by_dev[dev] = true;
    }
  }
  // one workgroup per codeword: the dispatcher hands a finished workgroup's CU the next
  // codeword, so codewords that stop early free their CU at once
  const uint32_t blocks = a.num_cw < 0x7fffffffu ? a.num_cw : 0x7fffffffu;
  hipLaunchKernelGGL((k_code_i8<CT>), dim3(blocks), dim3(CT::z), lds, st, a);
  return hipGetLastError();
}

}  // namespace

int code_i8_index(const DecoderDev& d) {
  const HdecHost& h = d.h;
  if (d.elem != FECSIM_I8 || d.mode != FECSIM_MIN) return -1;
  const float beta = d.params.beta_offset, mrm = d.params.max_row_metric;
  if (mrm < 1e-6f) return -1;   // cap 1 - max_row_metric: generic kernel
  const int b8 = (int)(int8_t)(int)beta, c8 = (int)(int8_t)(int)mrm;
  if (b8 < 1 || c8 < 0) return -1;
  return code_tab_index(h.z, h.mcol, h.mrow, h.row_weights.data(), h.row_active_cols.data(), h.row_offset.data());
}

bool code_i8_offsets_ok(const std::vector<long long>& tab) {
  for (long long b : tab)
    if (b < 1 || b > 127) return false;
  return true;
}

hipError_t launch_code_i8(const DecoderDev& d, const KernelArgs& ka, hipStream_t st) {
  CodeArgs a{};
  a.num_cw = ka.num_cw;
  a.max_iter = d.params.max_iter;
  a.beta = (int)(int8_t)(int)d.params.beta_offset;
  a.cap = (int)(int8_t)(int)d.params.max_row_metric;
  a.beta_tab = reinterpret_cast<const long long*>(d.d_beta_tab);
  a.llr_in = reinterpret_cast<const int8_t*>(ka.metrics_in);
  a.post_out = reinterpret_cast<int8_t*>(ka.metrics_out);
  a.iters = ka.iters;
  a.c2v = reinterpret_cast<int8_t*>(ka.c2v);
  a.tot_iter = ka.tot_iter;
  return code_dispatch(d.code_tab, a, st);
}

#define FECSIM_CODE_LAUNCH(CT) return launch_code<CT>(a, st);
#include "code_tables_dispatch.h"

}  // namespace fecsim

// Device helpers shared by the differentiable decoder's kernels (train.hip, train_reg.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace fecsim {
namespace train_core {

// IEEE sign bit as 0/1 (-0.0 counts as negative, like the reference's sign_combine)
__device__ __forceinline__ int sbit(float x) { return (int)(__float_as_uint(x) >> 31); }
__device__ __forceinline__ int sbit(double x) { return (int)((unsigned long long)__double_as_longlong(x) >> 63); }

// int8 decoder emulation (QT): saturation of V2C and posteriors to int8, and the
// magnitude of an offset-0 row: abs(-128) = -128 gives A = max(-128, 0) = 0
// (decode_functions.cpp:48-61,88-90).  On integer-valued inputs with integer offsets
// every value stays an integer, exact in float.
template <typename F>
__device__ __forceinline__ F sat8f(F x) { return x < F(-128) ? F(-128) : (x > F(127) ? F(127) : x); }
template <typename F, bool QT>
__device__ __forceinline__ F mag_a(F q, F be) {
  const F x = fabs(q) - be;
  const F A = x > F(0) ? x : F(0);
  if constexpr (QT) return (be == F(0) && q == F(-128)) ? F(0) : A;
  return A;
}

// sum over the 64 lanes of a wave (all lanes active), wave-uniform.  float: four DPP steps
// within each row of 16 lanes (partner lane ^ 1, ^ 2, the mirrored quad, the mirrored
// half-row) leave row sums R0..R3 in their rows; two row broadcasts then form
// (R3 + R2) + (R1 + R0) in lane 63: a fixed order, so every kernel that sums with it rounds
// alike
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_plus(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v) {
  v = dpp_plus<0xB1, 0xF>(v);    // quad_perm [1, 0, 3, 2]
  v = dpp_plus<0x4E, 0xF>(v);    // quad_perm [2, 3, 0, 1]
  v = dpp_plus<0x141, 0xF>(v);   // row_half_mirror
  v = dpp_plus<0x140, 0xF>(v);   // row_mirror
  v = dpp_plus<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3: R1 + R0, R3 + R2
  v = dpp_plus<0x143, 0x8>(v);   // row_bcast:31 into row 3: (R3 + R2) + (R1 + R0)
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// QT offsets: rounded to integers (the int8 decoder's table takes integers)
template <typename F, bool QT>
__device__ __forceinline__ F offset_of(F b) {
  if constexpr (QT) return rint(b);
  return b;
}

// first iteration of segment s: [0, f), [f, f + K), ... with f = T - (NS - 1) K
__device__ __forceinline__ int seg_start(int s, int T, int K, int NS) {
  return s == 0 ? 0 : T - (NS - s) * K;
}

}  // namespace train_core
}  // namespace fecsim

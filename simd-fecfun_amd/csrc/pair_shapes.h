// Generated by fecsim/gen_pair_shapes.py. Do not edit.
#pragma once
#include <stdint.h>

namespace fecsim {

struct PairShape0 {  // 11ad 1/2
  static constexpr int mrow = 8, ne = 52, maxw = 8;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[8] = {5, 6, 6, 6, 7, 7, 7, 8};
};
struct PairShape1 {  // 11ad 13/16
  static constexpr int mrow = 3, ne = 45, maxw = 16;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[3] = {14, 15, 16};
};
struct PairShape2 {  // 11ad 3/4
  static constexpr int mrow = 4, ne = 56, maxw = 15;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[4] = {13, 14, 14, 15};
};
struct PairShape3 {  // 11ad 5/8
  static constexpr int mrow = 6, ne = 50, maxw = 10;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[6] = {10, 10, 7, 8, 8, 7};
};
struct PairShape4 {  // 11n 1/2 z54
  static constexpr int mrow = 12, ne = 86, maxw = 8;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {7, 8, 7, 7, 7, 7, 7, 8, 7, 7, 7, 7};
};
struct PairShape5 {  // 11n 2/3 z54
  static constexpr int mrow = 8, ne = 88, maxw = 11;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[8] = {11, 11, 11, 11, 11, 11, 11, 11};
};
struct PairShape6 {  // 11n 3/4 z54
  static constexpr int mrow = 6, ne = 88, maxw = 15;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[6] = {15, 14, 15, 15, 15, 14};
};
struct PairShape7 {  // 11n 1/2 z27
  static constexpr int mrow = 12, ne = 88, maxw = 8;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {7, 8, 7, 7, 7, 8, 7, 7, 8, 7, 8, 7};
};
struct PairShape8 {  // 16e 1/2
  static constexpr int mrow = 12, ne = 76, maxw = 7;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {6, 7, 7, 6, 6, 7, 6, 6, 7, 6, 6, 6};
};
struct PairShape9 {  // 16e 2/3A
  static constexpr int mrow = 8, ne = 80, maxw = 10;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[8] = {10, 10, 10, 10, 10, 10, 10, 10};
};
struct PairShape10 {  // ITU G.hn 1/2 z14
  static constexpr int mrow = 12, ne = 77, maxw = 7;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {5, 7, 7, 6, 7, 6, 6, 6, 7, 6, 7, 7};
};

inline int pair_shape_index(int mrow, const int32_t* w) {
  if (mrow == 8 && w[0] == 5 && w[1] == 6 && w[2] == 6 && w[3] == 6 && w[4] == 7 && w[5] == 7 && w[6] == 7 && w[7] == 8) return 0;
  if (mrow == 3 && w[0] == 14 && w[1] == 15 && w[2] == 16) return 1;
  if (mrow == 4 && w[0] == 13 && w[1] == 14 && w[2] == 14 && w[3] == 15) return 2;
  if (mrow == 6 && w[0] == 10 && w[1] == 10 && w[2] == 7 && w[3] == 8 && w[4] == 8 && w[5] == 7) return 3;
  if (mrow == 12 && w[0] == 7 && w[1] == 8 && w[2] == 7 && w[3] == 7 && w[4] == 7 && w[5] == 7 && w[6] == 7 && w[7] == 8 && w[8] == 7 && w[9] == 7 && w[10] == 7 && w[11] == 7) return 4;
  if (mrow == 8 && w[0] == 11 && w[1] == 11 && w[2] == 11 && w[3] == 11 && w[4] == 11 && w[5] == 11 && w[6] == 11 && w[7] == 11) return 5;
  if (mrow == 6 && w[0] == 15 && w[1] == 14 && w[2] == 15 && w[3] == 15 && w[4] == 15 && w[5] == 14) return 6;
  if (mrow == 12 && w[0] == 7 && w[1] == 8 && w[2] == 7 && w[3] == 7 && w[4] == 7 && w[5] == 8 && w[6] == 7 && w[7] == 7 && w[8] == 8 && w[9] == 7 && w[10] == 8 && w[11] == 7) return 7;
  if (mrow == 12 && w[0] == 6 && w[1] == 7 && w[2] == 7 && w[3] == 6 && w[4] == 6 && w[5] == 7 && w[6] == 6 && w[7] == 6 && w[8] == 7 && w[9] == 6 && w[10] == 6 && w[11] == 6) return 8;
  if (mrow == 8 && w[0] == 10 && w[1] == 10 && w[2] == 10 && w[3] == 10 && w[4] == 10 && w[5] == 10 && w[6] == 10 && w[7] == 10) return 9;
  if (mrow == 12 && w[0] == 5 && w[1] == 7 && w[2] == 7 && w[3] == 6 && w[4] == 7 && w[5] == 6 && w[6] == 6 && w[7] == 6 && w[8] == 7 && w[9] == 6 && w[10] == 7 && w[11] == 7) return 10;
  return -1;
}

// the shape has a kernel with lifting size z and mcol columns compiled in
inline bool pair_shape_fixed_z(int shape, int z, int mcol) {
  switch (shape) {
  case 0: return z == 42 && mcol == 16;
  case 1: return z == 42 && mcol == 16;
  case 2: return z == 42 && mcol == 16;
  case 3: return z == 42 && mcol == 16;
  default: return false;
  }
}

// the shape has a register-C2V float32 decoder
inline bool float_shape_ok(int shape) {
  switch (shape) {
  case 0:
  case 4:
  case 7:
  case 8:
    return true;
  default: return false;
  }
}

struct PairArgs;
hipError_t pair_dispatch(int shape, const PairArgs& a, hipStream_t st);

}  // namespace fecsim

// Generated by fecsim/gen_pair_shapes.py. Do not edit.
#pragma once
#include <stdint.h>

namespace fecsim {

// code shapes (layer count and row weights) of the register-C2V differentiable decoder
// (csrc/train_reg.hip); shifts and columns stay run-time data
struct TrainShape0 {  // 11n 1/2 z81
  static constexpr int mrow = 12, ne = 86, maxw = 8;
  static constexpr int w[12] = {7, 7, 7, 7, 7, 7, 8, 7, 7, 7, 7, 8};
};
struct TrainShape1 {  // 11n 1/2 z54
  static constexpr int mrow = 12, ne = 86, maxw = 8;
  static constexpr int w[12] = {7, 8, 7, 7, 7, 7, 7, 8, 7, 7, 7, 7};
};
struct TrainShape2 {  // 11n 1/2 z27
  static constexpr int mrow = 12, ne = 88, maxw = 8;
  static constexpr int w[12] = {7, 8, 7, 7, 7, 8, 7, 7, 8, 7, 8, 7};
};
struct TrainShape3 {  // 11ad 1/2
  static constexpr int mrow = 8, ne = 52, maxw = 8;
  static constexpr int w[8] = {5, 6, 6, 6, 7, 7, 7, 8};
};
struct TrainShape4 {  // 16e 1/2
  static constexpr int mrow = 12, ne = 76, maxw = 7;
  static constexpr int w[12] = {6, 7, 7, 6, 6, 7, 6, 6, 7, 6, 6, 6};
};

inline int train_shape_index(int mrow, const int32_t* w) {
  if (mrow == 12 && w[0] == 7 && w[1] == 7 && w[2] == 7 && w[3] == 7 && w[4] == 7 && w[5] == 7 && w[6] == 8 && w[7] == 7 && w[8] == 7 && w[9] == 7 && w[10] == 7 && w[11] == 8) return 0;
  if (mrow == 12 && w[0] == 7 && w[1] == 8 && w[2] == 7 && w[3] == 7 && w[4] == 7 && w[5] == 7 && w[6] == 7 && w[7] == 8 && w[8] == 7 && w[9] == 7 && w[10] == 7 && w[11] == 7) return 1;
  if (mrow == 12 && w[0] == 7 && w[1] == 8 && w[2] == 7 && w[3] == 7 && w[4] == 7 && w[5] == 8 && w[6] == 7 && w[7] == 7 && w[8] == 8 && w[9] == 7 && w[10] == 8 && w[11] == 7) return 2;
  if (mrow == 8 && w[0] == 5 && w[1] == 6 && w[2] == 6 && w[3] == 6 && w[4] == 7 && w[5] == 7 && w[6] == 7 && w[7] == 8) return 3;
  if (mrow == 12 && w[0] == 6 && w[1] == 7 && w[2] == 7 && w[3] == 6 && w[4] == 6 && w[5] == 7 && w[6] == 6 && w[7] == 6 && w[8] == 7 && w[9] == 6 && w[10] == 6 && w[11] == 6) return 4;
  return -1;
}

}  // namespace fecsim

// Generated by fecsim/gen_pair_shapes.py. Do not edit.
#pragma once
#include <stdint.h>

namespace fecsim {

struct PairShape0 {  // 11ad 1/2
  static constexpr int mrow = 8, ne = 52, maxw = 8;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[8] = {5, 6, 6, 6, 7, 7, 7, 8};
  // edge table compiled in (P = 64 int8 layout): column, shift, shift class (shift mod Z/2) and shift >= Z/2 per edge
  static constexpr bool ct = true;
  static constexpr int ncls = 17;
  static constexpr int ecol[52] = {0, 2, 4, 6, 8, 0, 2, 4, 7, 8, 9, 1, 3, 5, 7, 9, 10, 1, 3, 5, 6, 10, 11, 0, 2, 4, 6, 8, 11, 12, 0, 2, 5, 7, 9, 11, 13, 1, 3, 5, 7, 10, 13, 14, 1, 3, 4, 6, 8, 12, 14, 15};
  static constexpr int esft[52] = {0, 40, 15, 7, 20, 39, 40, 32, 35, 7, 6, 37, 32, 8, 35, 11, 0, 37, 28, 22, 30, 25, 16, 39, 3, 2, 1, 32, 7, 32, 29, 0, 22, 4, 28, 27, 23, 32, 24, 22, 21, 13, 1, 14, 32, 2, 41, 24, 14, 23, 32, 34};
  static constexpr int ecls[52] = {0, 1, 2, 3, 4, 5, 1, 6, 7, 3, 8, 9, 6, 10, 7, 6, 0, 9, 3, 11, 12, 13, 9, 5, 14, 15, 11, 6, 3, 6, 10, 0, 11, 13, 3, 8, 15, 6, 14, 11, 0, 16, 11, 7, 6, 15, 4, 14, 7, 15, 6, 16};
  static constexpr int eflip[52] = {0, 1, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 0, 0, 1, 1, 1, 1, 1, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 1, 0, 1, 1, 0, 1, 1, 1};
  static constexpr int cls[17] = {0, 19, 15, 7, 20, 18, 11, 14, 6, 16, 8, 1, 9, 4, 3, 2, 13};
};
struct PairShape1 {  // 11ad 13/16
  static constexpr int mrow = 3, ne = 45, maxw = 16;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[3] = {14, 15, 16};
  // edge table compiled in (P = 64 int8 layout): column, shift, shift class (shift mod Z/2) and shift >= Z/2 per edge
  static constexpr bool ct = true;
  static constexpr int ncls = 18;
  static constexpr int ecol[45] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
  static constexpr int esft[45] = {2, 3, 15, 23, 6, 37, 32, 19, 0, 1, 35, 0, 39, 38, 24, 18, 5, 10, 40, 8, 35, 7, 19, 38, 41, 16, 39, 29, 0, 1, 40, 22, 10, 7, 21, 32, 33, 22, 20, 32, 36, 31, 31, 40, 0};
  static constexpr int ecls[45] = {0, 1, 2, 0, 3, 4, 5, 6, 7, 8, 9, 7, 10, 11, 1, 10, 12, 13, 6, 14, 9, 15, 6, 11, 16, 4, 10, 14, 7, 8, 6, 8, 13, 15, 7, 5, 17, 8, 16, 5, 2, 13, 13, 6, 7};
  static constexpr int eflip[45] = {0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1, 0, 1, 1, 1, 0, 0, 0, 1, 0, 1, 0, 0, 1, 1, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 0};
  static constexpr int cls[18] = {2, 3, 15, 6, 16, 11, 19, 0, 1, 14, 18, 17, 5, 10, 8, 7, 20, 12};
};
struct PairShape2 {  // 11ad 3/4
  static constexpr int mrow = 4, ne = 56, maxw = 15;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[4] = {13, 14, 14, 15};
  // edge table compiled in (P = 64 int8 layout): column, shift, shift class (shift mod Z/2) and shift >= Z/2 per edge
  static constexpr bool ct = true;
  static constexpr int ncls = 18;
  static constexpr int ecol[56] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15};
  static constexpr int esft[56] = {36, 20, 0, 23, 41, 0, 40, 7, 29, 19, 18, 4, 29, 6, 7, 19, 27, 10, 41, 36, 23, 4, 5, 39, 4, 1, 0, 37, 31, 18, 23, 11, 21, 6, 20, 32, 9, 12, 29, 0, 13, 1, 40, 22, 10, 7, 21, 32, 33, 22, 32, 36, 31, 31, 40, 0};
  static constexpr int ecls[56] = {0, 1, 2, 3, 1, 2, 4, 5, 6, 4, 7, 8, 6, 9, 5, 4, 9, 10, 1, 0, 3, 8, 11, 7, 8, 12, 2, 13, 10, 7, 3, 14, 2, 9, 1, 14, 15, 16, 6, 2, 17, 12, 4, 12, 10, 5, 2, 14, 16, 12, 14, 0, 10, 10, 4, 2};
  static constexpr int eflip[56] = {1, 0, 0, 1, 1, 0, 1, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 1, 0, 0, 1, 0, 0, 0, 1, 1, 0, 1, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0};
  static constexpr int cls[18] = {15, 20, 0, 2, 19, 7, 8, 18, 4, 6, 10, 5, 1, 16, 11, 9, 12, 13};
};
struct PairShape3 {  // 11ad 5/8
  static constexpr int mrow = 6, ne = 50, maxw = 10;
  static constexpr int zc = 42, mc = 16;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[6] = {10, 10, 7, 8, 8, 7};
  // edge table compiled in (P = 64 int8 layout): column, shift, shift class (shift mod Z/2) and shift >= Z/2 per edge
  static constexpr bool ct = true;
  static constexpr int ncls = 17;
  static constexpr int ecol[50] = {0, 1, 2, 3, 4, 5, 6, 7, 9, 10, 0, 1, 3, 5, 6, 7, 8, 9, 10, 11, 0, 2, 4, 6, 8, 11, 12, 0, 2, 5, 7, 9, 11, 12, 13, 1, 3, 5, 7, 9, 10, 13, 14, 1, 3, 4, 6, 8, 14, 15};
  static constexpr int esft[50] = {21, 37, 35, 32, 21, 8, 0, 35, 11, 0, 40, 37, 28, 22, 30, 24, 12, 35, 25, 16, 40, 4, 3, 2, 33, 8, 33, 29, 0, 22, 4, 28, 27, 24, 23, 32, 24, 22, 21, 10, 13, 1, 14, 32, 2, 41, 24, 14, 32, 34};
  static constexpr int ecls[50] = {0, 1, 2, 3, 0, 4, 0, 2, 3, 0, 5, 1, 6, 7, 8, 9, 10, 2, 11, 1, 5, 11, 9, 12, 10, 4, 10, 4, 0, 7, 11, 6, 13, 9, 12, 3, 9, 7, 0, 14, 15, 7, 2, 3, 12, 16, 9, 2, 3, 15};
  static constexpr int eflip[50] = {1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 1, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 1, 0, 1, 1, 0, 1, 1};
  static constexpr int cls[17] = {0, 16, 14, 11, 8, 19, 7, 1, 9, 3, 12, 4, 2, 6, 10, 13, 20};
};
struct PairShape4 {  // 11n 1/2 z54
  static constexpr int mrow = 12, ne = 86, maxw = 8;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {7, 8, 7, 7, 7, 7, 7, 8, 7, 7, 7, 7};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
};
struct PairShape5 {  // 11n 2/3 z54
  static constexpr int mrow = 8, ne = 88, maxw = 11;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[8] = {11, 11, 11, 11, 11, 11, 11, 11};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
};
struct PairShape6 {  // 11n 3/4 z54
  static constexpr int mrow = 6, ne = 88, maxw = 15;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[6] = {15, 14, 15, 15, 15, 14};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
};
struct PairShape7 {  // 11n 1/2 z27
  static constexpr int mrow = 12, ne = 88, maxw = 8;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {7, 8, 7, 7, 7, 8, 7, 7, 8, 7, 8, 7};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
};
struct PairShape8 {  // 16e 1/2
  static constexpr int mrow = 12, ne = 76, maxw = 7;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {6, 7, 7, 6, 6, 7, 6, 6, 7, 6, 6, 6};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
};
struct PairShape9 {  // 16e 2/3A
  static constexpr int mrow = 8, ne = 80, maxw = 10;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[8] = {10, 10, 10, 10, 10, 10, 10, 10};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
};
struct PairShape10 {  // ITU G.hn 1/2 z14
  static constexpr int mrow = 12, ne = 77, maxw = 7;
  static constexpr int zc = 0, mc = 0;  // Z and columns of the fixed-Z kernel (0: none)
  static constexpr int w[12] = {5, 7, 7, 6, 7, 6, 6, 6, 7, 6, 7, 7};
  static constexpr bool ct = false;
  static constexpr int ncls = 0;
  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};
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

// the shape's compiled-in edge table equals (cols, shifts) of a decoder
inline bool pair_shape_table(int shape, int ne, const int32_t* cols, const int32_t* shifts) {
  switch (shape) {
  case 0: {
    static const int32_t c[52] = {0, 2, 4, 6, 8, 0, 2, 4, 7, 8, 9, 1, 3, 5, 7, 9, 10, 1, 3, 5, 6, 10, 11, 0, 2, 4, 6, 8, 11, 12, 0, 2, 5, 7, 9, 11, 13, 1, 3, 5, 7, 10, 13, 14, 1, 3, 4, 6, 8, 12, 14, 15};
    static const int32_t s[52] = {0, 40, 15, 7, 20, 39, 40, 32, 35, 7, 6, 37, 32, 8, 35, 11, 0, 37, 28, 22, 30, 25, 16, 39, 3, 2, 1, 32, 7, 32, 29, 0, 22, 4, 28, 27, 23, 32, 24, 22, 21, 13, 1, 14, 32, 2, 41, 24, 14, 23, 32, 34};
    if (ne != 52) return false;
    for (int e = 0; e < ne; ++e)
      if (cols[e] != c[e] || shifts[e] != s[e]) return false;
    return true;
  }
  case 1: {
    static const int32_t c[45] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    static const int32_t s[45] = {2, 3, 15, 23, 6, 37, 32, 19, 0, 1, 35, 0, 39, 38, 24, 18, 5, 10, 40, 8, 35, 7, 19, 38, 41, 16, 39, 29, 0, 1, 40, 22, 10, 7, 21, 32, 33, 22, 20, 32, 36, 31, 31, 40, 0};
    if (ne != 45) return false;
    for (int e = 0; e < ne; ++e)
      if (cols[e] != c[e] || shifts[e] != s[e]) return false;
    return true;
  }
  case 2: {
    static const int32_t c[56] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15};
    static const int32_t s[56] = {36, 20, 0, 23, 41, 0, 40, 7, 29, 19, 18, 4, 29, 6, 7, 19, 27, 10, 41, 36, 23, 4, 5, 39, 4, 1, 0, 37, 31, 18, 23, 11, 21, 6, 20, 32, 9, 12, 29, 0, 13, 1, 40, 22, 10, 7, 21, 32, 33, 22, 32, 36, 31, 31, 40, 0};
    if (ne != 56) return false;
    for (int e = 0; e < ne; ++e)
      if (cols[e] != c[e] || shifts[e] != s[e]) return false;
    return true;
  }
  case 3: {
    static const int32_t c[50] = {0, 1, 2, 3, 4, 5, 6, 7, 9, 10, 0, 1, 3, 5, 6, 7, 8, 9, 10, 11, 0, 2, 4, 6, 8, 11, 12, 0, 2, 5, 7, 9, 11, 12, 13, 1, 3, 5, 7, 9, 10, 13, 14, 1, 3, 4, 6, 8, 14, 15};
    static const int32_t s[50] = {21, 37, 35, 32, 21, 8, 0, 35, 11, 0, 40, 37, 28, 22, 30, 24, 12, 35, 25, 16, 40, 4, 3, 2, 33, 8, 33, 29, 0, 22, 4, 28, 27, 24, 23, 32, 24, 22, 21, 10, 13, 1, 14, 32, 2, 41, 24, 14, 32, 34};
    if (ne != 50) return false;
    for (int e = 0; e < ne; ++e)
      if (cols[e] != c[e] || shifts[e] != s[e]) return false;
    return true;
  }
  default: return false;
  }
}

// k_pair_i8c's LDS layout of the shape (lds_layout.py): per hardware lane the
// (slot, lr) item slot * Z/2 + lr (-1: spare lane), per (slot, position) item its
// dword in a column row; false: slots in lane order, positions in slot order
inline bool pair_shape_lds_layout(int shape, int32_t* lane_item, int32_t* dword) {
  switch (shape) {
  case 0: {  // 11ad 1/2: 2.346 LDS cycles per access (3.808 in lane order)
    static const int32_t li[64] = {0, 1, 8, 9, 10, 11, 12, 18, 19, 20, 26, 27, 28, 29, 30, 31, 37, 38, 39, 40, 41, 42, 43, 49, 50, 51, 52, 53, 59, 60, 61, 62, 2, 3, 4, 5, 6, 7, 13, 14, 15, 16, 17, 21, 22, 23, 24, 25, 32, 33, 34, 35, 36, 44, 45, 46, 47, 48, 54, 55, 56, 57, 58, -1};
    static const int32_t dw[63] = {8, 39, 10, 5, 2, 54, 43, 38, 52, 3, 28, 9, 6, 23, 57, 47, 50, 24, 53, 19, 4, 62, 46, 26, 40, 7, 30, 31, 34, 13, 1, 16, 0, 51, 60, 41, 48, 55, 37, 15, 27, 56, 11, 49, 20, 35, 61, 59, 17, 32, 25, 29, 18, 44, 21, 63, 36, 45, 12, 42, 14, 58, 22};
    for (int x = 0; x < 64; ++x) lane_item[x] = li[x];
    for (int x = 0; x < 63; ++x) dword[x] = dw[x];
    return true;
  }
  case 1: {  // 11ad 13/16: 2.311 LDS cycles per access (3.733 in lane order)
    static const int32_t li[64] = {2, 3, 4, 5, 6, 8, 9, 10, 11, 21, 22, 23, 34, 35, 36, 37, 38, 39, 40, 41, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 61, 0, 1, 7, 12, 13, 14, 15, 16, 17, 18, 19, 20, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 42, 43, 44, 56, 57, 58, 59, 60, 62, -1};
    static const int32_t dw[63] = {2, 28, 41, 3, 59, 50, 7, 48, 36, 10, 31, 46, 62, 1, 54, 0, 23, 15, 17, 24, 6, 38, 5, 40, 58, 13, 43, 25, 21, 51, 44, 20, 37, 8, 26, 45, 11, 29, 57, 53, 19, 12, 42, 63, 14, 30, 33, 22, 32, 55, 47, 49, 56, 52, 34, 60, 9, 35, 27, 18, 39, 16, 4};
    for (int x = 0; x < 64; ++x) lane_item[x] = li[x];
    for (int x = 0; x < 63; ++x) dword[x] = dw[x];
    return true;
  }
  case 2: {  // 11ad 3/4: 2.339 LDS cycles per access (3.679 in lane order)
    static const int32_t li[64] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 24, 25, 27, 29, 31, 33, 35, 37, 39, 41, 21, 23, 26, 28, 30, 32, 34, 36, 38, 40, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 62, -1};
    static const int32_t dw[63] = {31, 39, 22, 40, 50, 6, 45, 56, 15, 52, 62, 4, 25, 27, 60, 49, 23, 14, 34, 41, 12, 61, 1, 35, 3, 42, 10, 37, 5, 0, 32, 26, 58, 11, 43, 53, 16, 48, 51, 19, 33, 29, 24, 28, 13, 36, 55, 8, 38, 9, 46, 7, 44, 2, 17, 20, 18, 54, 59, 63, 30, 47, 57};
    for (int x = 0; x < 64; ++x) lane_item[x] = li[x];
    for (int x = 0; x < 63; ++x) dword[x] = dw[x];
    return true;
  }
  case 3: {  // 11ad 5/8: 2.480 LDS cycles per access (3.740 in lane order)
    static const int32_t li[64] = {0, 5, 6, 7, 12, 13, 14, 18, 19, 20, 21, 25, 26, 27, 32, 33, 34, 39, 40, 41, 44, 45, 46, 47, 51, 52, 53, 54, 58, 59, 60, 61, 1, 2, 3, 4, 8, 9, 10, 11, 15, 16, 17, 22, 23, 24, 28, 29, 30, 31, 35, 36, 37, 38, 42, 43, 48, 49, 50, 55, 56, 57, 62, -1};
    static const int32_t dw[63] = {42, 34, 25, 14, 21, 7, 18, 0, 36, 38, 62, 60, 41, 63, 48, 3, 33, 61, 40, 55, 47, 28, 39, 31, 32, 17, 51, 26, 52, 23, 24, 10, 2, 1, 29, 8, 44, 15, 37, 22, 6, 11, 50, 58, 4, 12, 30, 59, 54, 19, 43, 20, 45, 46, 5, 49, 13, 16, 35, 57, 56, 53, 9};
    for (int x = 0; x < 64; ++x) lane_item[x] = li[x];
    for (int x = 0; x < 63; ++x) dword[x] = dw[x];
    return true;
  }
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

// Local orderings of simplex sub-entities, in constant memory.
//
// These tables must equal the reference's (tables.cpp:77-193) because they
// define the numbering of every derived and generated entity.  canon(h,l,w,j)
// is row j of the (h,l,w) table: the w-dimensional sub-entities (local
// indices in the h-simplex) of the j-th l-dimensional sub-entity.
#pragma once

#include "ohx_common.hpp"

namespace ohx {

static __constant__ u32 k_down_deg[4][4] = {{1, 0, 0, 0}, {2, 1, 0, 0}, {3, 3, 1, 0}, {4, 6, 4, 1}};
static __constant__ u32 k_f_e_v[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static __constant__ u32 k_r_e_v[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static __constant__ u32 k_r_f_v[4][3] = {{0, 1, 2}, {1, 0, 3}, {2, 1, 3}, {0, 2, 3}};
static __constant__ u32 k_r_f_e[4][3] = {{0, 1, 2}, {0, 3, 4}, {1, 4, 5}, {2, 5, 3}};
static __constant__ u32 k_iota[6] = {0, 1, 2, 3, 4, 5};
// the_opposite_orders
static __constant__ u32 k_o_e_v[2] = {1, 0};
static __constant__ u32 k_o_f_v[3] = {1, 2, 0};
static __constant__ u32 k_o_f_e[3] = {2, 0, 1};
static __constant__ u32 k_o_r_v[4] = {2, 3, 1, 0};
static __constant__ u32 k_o_r_e[6] = {5, 3, 4, 1, 2, 0};
static __constant__ u32 k_o_r_f[4] = {3, 2, 0, 1};

__device__ inline u32 const* canon(u32 h, u32 l, u32 w, u32 j) {
  if (w == l) return &k_iota[(h == l) ? 0 : j];
  if (h == l) return k_iota;
  switch (h * 100 + l * 10 + w) {
    case 210: return k_f_e_v[j];
    case 310: return k_r_e_v[j];
    case 320: return k_r_f_v[j];
    default: return k_r_f_e[j];  // 321
  }
}

__device__ inline u32 const* opp(u32 h, u32 l) {
  switch (h * 10 + l) {
    case 10: return k_o_e_v;
    case 20: return k_o_f_v;
    case 21: return k_o_f_e;
    case 30: return k_o_r_v;
    case 31: return k_o_r_e;
    default: return k_o_r_f;  // 32
  }
}

// host copy of the degrees
static constexpr u32 h_down_deg[4][4] = {{1, 0, 0, 0}, {2, 1, 0, 0}, {3, 3, 1, 0}, {4, 6, 4, 1}};

}  // namespace ohx

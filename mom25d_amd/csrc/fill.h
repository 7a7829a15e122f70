// Data layouts shared by the fill kernels (fill.hip) and the plan object
// that prepares their work (plan.hip).
#pragma once

#include "common.h"

namespace mom {

constexpr int kBlock = 32;     // test triangles per block
constexpr int kChunk = 64;     // source triangles per work item (one per lane)
constexpr double kFourPi = 12.566370614359172954;

struct Bucket {        // interval lookup entry (16 bytes; see spline_interval)
  double split;        // the node inside the bucket's range (+inf: none)
  int32_t i0;          // interval at the start of the range
  int32_t pad_;
};

struct PairTab {       // spline of one (observer level, source level) pair at one frequency
  int64_t x_off;       // offset of the rho nodes
  int64_t c_off;       // offset (in doubles) of the coefficients, 16 per interval
  int64_t b_off;       // offset (in entries) of the bucket table
  int32_t n;           // number of nodes
  int32_t nbucket;
  double inv_dx;       // 1 / bucket width
  double dz;           // vertical distance of the two levels
  double cA_re, cA_im; // coefficient of 1/(4 pi R) in gA
  double cV_re, cV_im; // coefficient of 1/(4 pi R) in gV
  double cW_re, cW_im; // coefficient of 1/(4 pi R) in gW (line voltage)
  int32_t zero;        // levels in different stack sections: no coupling
  int32_t pad_;
};

struct FillArgs {
  // mesh plan
  const double* V;     // (Nt,3,2) vertices
  const double* qp7;   // (Nt,7,2)
  const double* qw7;   // (Nt,7)
  const double* qp3;   // (Nt,3,2)
  const double* qw3;   // (Nt,3)
  const double* cen;   // (Nt,2)
  const double* area;  // (Nt)
  const double* cA;    // (Nt,3)
  const double* cV;    // (Nt,3)
  const int32_t* tb;   // (Nt,3) basis index or -1
  const int32_t* pad_pos;   // (Nt) row in IVpad or -1
  const int32_t* blk_tris;  // (nblk, 32), -1 padded
  const int32_t* items;     // (n_items, 6): block, pair index, src offset, src count,
                            // column offset, column count
  const int32_t* src;       // source triangle lists
  int32_t nt;
  // frequency data
  const PairTab* pairs;
  const double* tab_x;
  const double* tab_c;
  const Bucket* tab_b;
  double jwmu_re, jwmu_im;
  double ijwe_re, ijwe_im;
  double* Z;      // complex, row-major
  int64_t ldz;    // leading dimension in complex elements
  double* IVpad;  // complex (npad, Nt) or null
  // scatter through LDS: the distinct rows (test basis functions) of every
  // block and columns (source basis functions) of every work item
  const int32_t* blk_row_off;  // (nblk + 1) offsets into blk_rows
  const int32_t* blk_rows;     // basis index of each block row
  const int32_t* blk_rloc;     // (nblk, 32, 3) row of each block triangle's edge, or -1
  const int32_t* item_cols;    // basis index of each item column (offsets in items[4])
  const int32_t* src_cloc;     // (len(src), 3) column of each source triangle's edge, or -1
  int64_t acc_bytes;           // LDS accumulator bytes of the launch (max over its items)
  // per work item, bounds [lo, hi] on its source-test distances R (padded
  // for rounding): the spline intervals between them are staged in LDS when
  // there are at most stage_max (<= kStageMax) of them
  const double* item_r;
  int32_t stage_max;
  int32_t pad2_;
  // near tier: the frequency-independent closed-form parts S0 = I0 / 4 pi,
  // S1 = ((r - o) I0 + I1) / 4 pi of every (test point, source triangle),
  // computed once per mesh by near_static_kernel; work item it keeps them at
  // near_static + item_soff[it] as [test triangle][test point][S0, S1x, S1y]
  // [source lane] with a lane stride of the item's source count
  const double* near_static;
  const int64_t* item_soff;
  // ownership of Z entries within one tier's launch: entry (row, col) of a
  // work item is touched by no other item of the launch when both its block
  // row (row_own, per blk_rows entry) and its item column (col_own, per
  // item_cols entry) are 1; the flush then adds it with a plain load and
  // store instead of atomics.  Null: every entry through atomics.
  const uint8_t* row_own;
  const uint8_t* col_own;
};

}  // namespace mom

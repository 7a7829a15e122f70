// Plan object of the impedance fill for one mesh.
//
// Built from the arrays of the reference's frequency-independent set-up,
// AssemblyGeometry (Engine/mom25d/solver.py:307-426): the quadrature points
// and weights of every triangle (7-point test / near rule, 3-point mid rule,
// centroid far rule), the RWG coefficients cA, cV and basis map, the layer
// heights, the Morton-ordered test blocks of up to 32 triangles and, per
// block, the source groups (source layer, tier, source triangles) in which
// it classified every (block, source) pair.  The plan turns them into the
// work of the GPU fill -- items of <= 64 source triangles, the distinct rows
// and columns each block and item adds into, distance bounds for the spline
// staging -- uploads it once and computes the near tier's closed-form
// parts once.  Per frequency it then adds the RWG block of Z and the pad
// potentials IVpad that the reference's loop of _fast.chunk_vals + scatter
// produces (Engine/mom25d/solver.py:472-483, _fast.py:145-211), from either
// caller-provided device tables or splines given pair by pair in the
// reference's own PPoly layout (_fast.spline_data, _fast.py:35-45).

#include "fill.h"

#include "../../include/mom25d.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

namespace mom {
namespace {

constexpr int64_t kAccMaxBytes = 136 * 1024;  // LDS accumulator of one fill workgroup
constexpr int kTiers = 3;

struct Dev {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Plan {
  // host arrays
  int32_t nt = 0, nb = 0, nl = 0, n_blocks = 0, n_pad = 0;
  std::vector<double> V, qp7, qw7, qp3, qw3, cen, area, cA, cV, zl;
  std::vector<int32_t> tb, pad_pos, blk_tris, blk_ntest, blk_layer, src;
  std::vector<int32_t> items[kTiers];
  std::vector<double> item_r[kTiers];
  int64_t acc_bytes[kTiers] = {0, 0, 0};
  int64_t tier_pairs[kTiers] = {0, 0, 0};
  std::vector<int32_t> blk_row_off, blk_rows, blk_rloc, item_cols, src_cloc;
  std::vector<uint8_t> row_own, col_own;  // entries with one writer per launch (FillArgs)
  std::vector<int64_t> item_soff;
  int64_t near_static_len = 0;
  std::vector<char> pair_used;  // (nl * nl) pairs some work item reads

  // the plan's own spline tables (mom_plan_set_pair)
  std::vector<PairTab> rec;
  std::vector<std::vector<double>> own_x, own_c;
  std::vector<std::vector<Bucket>> own_b;
  std::vector<char> own_set;
  bool own_dirty = true, own_uploaded = false;
  std::vector<double> cat_x, cat_c;  // the own tables concatenated (own_tables)
  std::vector<Bucket> cat_b;
  std::vector<PairTab> cat_rec;

  // device
  int device = -1;
  bool on_device = false;
  Dev d_V, d_qp7, d_qw7, d_qp3, d_qw3, d_cen, d_area, d_cA, d_cV, d_tb, d_pad_pos, d_blk_tris, d_src;
  Dev d_items[kTiers], d_item_r[kTiers];
  Dev d_blk_row_off, d_blk_rows, d_blk_rloc, d_item_cols, d_src_cloc, d_item_soff, d_near_static;
  Dev d_row_own, d_col_own;
  Dev d_pairs, d_tab_x, d_tab_c, d_tab_b;  // own tables
  Dev d_Z, d_IV;                           // staging of mom_plan_fill_host
  hipStream_t host_stream = nullptr;

  std::vector<Dev*> all() {
    std::vector<Dev*> v = {&d_V, &d_qp7, &d_qw7, &d_qp3, &d_qw3, &d_cen, &d_area, &d_cA, &d_cV, &d_tb,
                           &d_pad_pos, &d_blk_tris, &d_src, &d_blk_row_off, &d_blk_rows, &d_blk_rloc,
                           &d_item_cols, &d_src_cloc, &d_item_soff, &d_near_static, &d_pairs, &d_tab_x,
                           &d_tab_c, &d_tab_b, &d_Z, &d_IV, &d_row_own, &d_col_own};
    for (int t = 0; t < kTiers; ++t) {
      v.push_back(&d_items[t]);
      v.push_back(&d_item_r[t]);
    }
    return v;
  }
};

// ------------------------------------------------------------- host set-up
// Stable argsort of the Z-order code of 16-bit quantised points
// (Engine/mom25d/solver.py:294-304): q = trunc((p - min) / span * 65535).
std::vector<int32_t> morton_order(const std::vector<int32_t>& tri, const std::vector<double>& cen) {
  const size_t m = tri.size();
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int32_t t : tri)
    for (int d = 0; d < 2; ++d) {
      lo[d] = std::min(lo[d], cen[2 * (size_t)t + d]);
      hi[d] = std::max(hi[d], cen[2 * (size_t)t + d]);
    }
  double span = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), 1e-30);
  auto spread = [](uint64_t v) {
    const uint64_t masks[4] = {0x00FF00FF, 0x0F0F0F0F, 0x33333333, 0x55555555};
    const int shifts[4] = {8, 4, 2, 1};
    for (int k = 0; k < 4; ++k) v = (v | (v << shifts[k])) & masks[k];
    return v;
  };
  std::vector<uint64_t> code(m);
  for (size_t i = 0; i < m; ++i) {
    uint64_t q[2];
    for (int d = 0; d < 2; ++d) q[d] = (uint64_t)((cen[2 * (size_t)tri[i] + d] - lo[d]) / span * 65535.0);
    code[i] = spread(q[0]) | (spread(q[1]) << 1);
  }
  std::vector<int32_t> order(m);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
  return order;
}

// Sorted distinct basis indices of the edges of triangles tri[0..n) and the
// position of each edge among them (-1: no basis).
void distinct_edges(const int32_t* tb, const int32_t* tri, int n, std::vector<int32_t>& u, int32_t* loc) {
  u.clear();
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 3; ++i) {
      int32_t e = tb[3 * (size_t)tri[k] + i];
      if (e >= 0) u.push_back(e);
    }
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 3; ++i) {
      int32_t e = tb[3 * (size_t)tri[k] + i];
      loc[3 * k + i] = e < 0 ? -1 : (int32_t)(std::lower_bound(u.begin(), u.end(), e) - u.begin());
    }
}

// Centre and radius of the bounding box of n points.
void circle(const double* pts, size_t n, double& cx, double& cy, double& r) {
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i)
    for (int d = 0; d < 2; ++d) {
      lo[d] = std::min(lo[d], pts[2 * i + d]);
      hi[d] = std::max(hi[d], pts[2 * i + d]);
    }
  cx = 0.5 * (lo[0] + hi[0]);
  cy = 0.5 * (lo[1] + hi[1]);
  r = 0.5 * std::hypot(hi[0] - lo[0], hi[1] - lo[1]);
}

int build(Plan& P, const MomAssemblyGeometry& g) {
  const int32_t nt = g.nt, nb = g.nb, nl = g.n_levels, nblk = g.n_blocks, ngrp = g.n_groups;
  if (nt <= 0 || nb < 0 || nl <= 0 || nblk <= 0 || ngrp < 0) return fail("mom_plan_create: empty or negative sizes");
  if (!g.V || !g.qpts || !g.qw || !g.mpts || !g.mw || !g.fpts || !g.fw || !g.cA || !g.cV || !g.tri_basis || !g.zl ||
      !g.blk_off || !g.blk_tris || !g.blk_layer || (ngrp && (!g.grp_block || !g.grp_layer || !g.grp_tier ||
                                                             !g.grp_off || !g.grp_src)))
    return fail("mom_plan_create: null array");
  P.nt = nt, P.nb = nb, P.nl = nl, P.n_blocks = nblk;
  P.V.assign(g.V, g.V + 6 * (size_t)nt);
  P.qp7.assign(g.qpts, g.qpts + 14 * (size_t)nt);
  P.qw7.assign(g.qw, g.qw + 7 * (size_t)nt);
  P.qp3.assign(g.mpts, g.mpts + 6 * (size_t)nt);
  P.qw3.assign(g.mw, g.mw + 3 * (size_t)nt);
  P.cen.assign(g.fpts, g.fpts + 2 * (size_t)nt);
  P.area.assign(g.fw, g.fw + (size_t)nt);
  P.cA.assign(g.cA, g.cA + 3 * (size_t)nt);
  P.cV.assign(g.cV, g.cV + 3 * (size_t)nt);
  P.zl.assign(g.zl, g.zl + nl);
  P.tb.resize(3 * (size_t)nt);
  for (size_t i = 0; i < 3 * (size_t)nt; ++i) {
    int64_t e = g.tri_basis[i];
    if (e < -1 || e >= nb) return fail("mom_plan_create: tri_basis entry outside [-1, nb)");
    P.tb[i] = (int32_t)e;
  }
  P.pad_pos.assign(nt, -1);
  int32_t npad = 0;
  if (g.pad_pos) {
    for (int32_t t = 0; t < nt; ++t) {
      int64_t q = g.pad_pos[t];
      if (q < -1 || q >= nt) return fail("mom_plan_create: pad_pos entry outside [-1, nt)");
      P.pad_pos[t] = (int32_t)q;
      npad = std::max<int32_t>(npad, (int32_t)q + 1);
    }
  }
  P.n_pad = npad;

  // test blocks: every triangle in exactly one block of <= 32 on one layer
  std::vector<int32_t> tri_layer(nt, -1);
  P.blk_tris.assign((size_t)nblk * kBlock, -1);
  P.blk_ntest.resize(nblk);
  P.blk_layer.resize(nblk);
  if (g.blk_off[0] != 0) return fail("mom_plan_create: blk_off must start at 0");
  for (int32_t b = 0; b < nblk; ++b) {
    int64_t a = g.blk_off[b], e = g.blk_off[b + 1];
    if (e <= a || e - a > kBlock) return fail("mom_plan_create: a test block must hold 1 to 32 triangles");
    int32_t lt = g.blk_layer[b];
    if (lt < 0 || lt >= nl) return fail("mom_plan_create: block layer outside [0, n_levels)");
    P.blk_layer[b] = lt;
    P.blk_ntest[b] = (int32_t)(e - a);
    for (int64_t k = a; k < e; ++k) {
      int64_t t = g.blk_tris[k];
      if (t < 0 || t >= nt) return fail("mom_plan_create: block triangle outside [0, nt)");
      if (tri_layer[t] >= 0) return fail("mom_plan_create: a triangle is in two test blocks");
      tri_layer[t] = lt;
      P.blk_tris[(size_t)b * kBlock + (k - a)] = (int32_t)t;
    }
  }
  for (int32_t t = 0; t < nt; ++t)
    if (tri_layer[t] < 0) return fail("mom_plan_create: a triangle is in no test block");

  // Morton rank of every triangle within its layer: sources are chunked in
  // this order, so that a chunk is a compact patch with few columns
  std::vector<int64_t> mrank(nt, 0);
  for (int32_t l = 0; l < nl; ++l) {
    std::vector<int32_t> on;
    for (int32_t t = 0; t < nt; ++t)
      if (tri_layer[t] == l) on.push_back(t);
    if (on.empty()) continue;
    std::vector<int32_t> ord = morton_order(on, P.cen);
    for (size_t i = 0; i < ord.size(); ++i) mrank[on[ord[i]]] = (int64_t)i;
  }

  // source groups -> source lists and chunks of <= 64 per tier
  struct Chunk {
    int32_t blk, pp, so, cnt;
  };
  std::vector<Chunk> per_tier[kTiers];
  P.pair_used.assign((size_t)nl * nl, 0);
  if (ngrp && g.grp_off[0] != 0) return fail("mom_plan_create: grp_off must start at 0");
  std::vector<char> seen((size_t)nt);
  for (int32_t q = 0; q < ngrp; ++q) {
    int32_t b = g.grp_block[q], ls = g.grp_layer[q], tier = g.grp_tier[q];
    if (b < 0 || b >= nblk || ls < 0 || ls >= nl || tier < 0 || tier >= kTiers)
      return fail("mom_plan_create: group block, layer or tier out of range");
    int64_t a = g.grp_off[q], e = g.grp_off[q + 1];
    if (e < a) return fail("mom_plan_create: grp_off must not decrease");
    if (e == a) continue;
    std::vector<int32_t> sel;
    sel.reserve(e - a);
    for (int64_t k = a; k < e; ++k) {
      int64_t s = g.grp_src[k];
      if (s < 0 || s >= nt) return fail("mom_plan_create: source triangle outside [0, nt)");
      if (tri_layer[s] != ls) return fail("mom_plan_create: a group's source triangle is not on its layer");
      sel.push_back((int32_t)s);
    }
    std::stable_sort(sel.begin(), sel.end(), [&](int32_t x, int32_t y) { return mrank[x] < mrank[y]; });
    if (P.src.size() + sel.size() > (size_t)std::numeric_limits<int32_t>::max())
      return fail("mom_plan_create: more than 2^31 source entries");
    const int32_t so0 = (int32_t)P.src.size();
    P.src.insert(P.src.end(), sel.begin(), sel.end());
    const int32_t pp = P.blk_layer[b] * nl + ls;
    P.pair_used[pp] = 1;
    P.tier_pairs[tier] += (int64_t)sel.size() * P.blk_ntest[b];
    for (int32_t c0 = 0; c0 < (int32_t)sel.size(); c0 += kChunk)
      per_tier[tier].push_back({b, pp, so0 + c0, std::min<int32_t>(kChunk, (int32_t)sel.size() - c0)});
  }

  // distinct rows of every block, and the row of each block triangle's edge
  P.blk_row_off.assign(nblk + 1, 0);
  P.blk_rloc.assign((size_t)nblk * kBlock * 3, -1);
  std::vector<int32_t> nrow(nblk), u;
  for (int32_t b = 0; b < nblk; ++b) {
    distinct_edges(P.tb.data(), &P.blk_tris[(size_t)b * kBlock], P.blk_ntest[b], u, &P.blk_rloc[(size_t)b * kBlock * 3]);
    P.blk_rows.insert(P.blk_rows.end(), u.begin(), u.end());
    nrow[b] = (int32_t)u.size();
    P.blk_row_off[b + 1] = P.blk_row_off[b] + nrow[b];
  }
  // distinct columns of every work item; an item whose accumulator would
  // not fit the LDS budget is halved (first half processed first)
  P.src_cloc.assign(3 * P.src.size(), -1);
  int64_t col_off = 0;
  std::vector<int32_t> loc;
  for (int tier = 0; tier < kTiers; ++tier) {
    std::deque<Chunk> todo(per_tier[tier].begin(), per_tier[tier].end());
    int64_t acc = 0;
    while (!todo.empty()) {
      Chunk c = todo.front();
      todo.pop_front();
      loc.assign(3 * (size_t)c.cnt, -1);
      distinct_edges(P.tb.data(), &P.src[c.so], c.cnt, u, loc.data());
      const int64_t nbytes = 16 * (int64_t)nrow[c.blk] * (int64_t)u.size();
      if (nbytes > kAccMaxBytes && c.cnt > 1) {
        const int32_t h = c.cnt / 2;
        todo.push_front({c.blk, c.pp, c.so + h, c.cnt - h});
        todo.push_front({c.blk, c.pp, c.so, h});
        continue;
      }
      if (nbytes > kAccMaxBytes) return fail("mom_plan_create: one source triangle's accumulator exceeds the LDS");
      std::copy(loc.begin(), loc.end(), P.src_cloc.begin() + 3 * (size_t)c.so);
      if (col_off + (int64_t)u.size() > std::numeric_limits<int32_t>::max())
        return fail("mom_plan_create: more than 2^31 item columns");
      const int32_t row[6] = {c.blk, c.pp, c.so, c.cnt, (int32_t)col_off, (int32_t)u.size()};
      P.items[tier].insert(P.items[tier].end(), row, row + 6);
      P.item_cols.insert(P.item_cols.end(), u.begin(), u.end());
      col_off += (int64_t)u.size();
      acc = std::max(acc, nbytes);
    }
    P.acc_bytes[tier] = acc;
  }

  // ownership: a block row whose basis function no other block holds, and
  // an item column whose basis function no other item of the same block and
  // tier holds (items of other blocks reach the entry only through their own
  // rows), so that the item is the only writer of such an entry in its
  // tier's launch
  {
    std::vector<int32_t> cnt((size_t)P.nb, 0);
    for (int32_t m : P.blk_rows) ++cnt[m];
    P.row_own.resize(P.blk_rows.size());
    for (size_t e = 0; e < P.blk_rows.size(); ++e) P.row_own[e] = cnt[P.blk_rows[e]] == 1;
    P.col_own.assign(P.item_cols.size(), 0);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int tier = 0; tier < kTiers; ++tier) {
      const int32_t* it = P.items[tier].data();
      const size_t n_items = P.items[tier].size() / 6;
      std::vector<std::vector<int32_t>> by_blk((size_t)nblk);
      for (size_t k = 0; k < n_items; ++k) by_blk[it[6 * k]].push_back((int32_t)k);
      for (const auto& ks : by_blk) {
        for (int32_t k : ks)
          for (int32_t e = it[6 * k + 4]; e < it[6 * k + 4] + it[6 * k + 5]; ++e) ++cnt[P.item_cols[e]];
        for (int32_t k : ks)
          for (int32_t e = it[6 * k + 4]; e < it[6 * k + 4] + it[6 * k + 5]; ++e)
            P.col_own[e] = cnt[P.item_cols[e]] == 1;
        for (int32_t k : ks)
          for (int32_t e = it[6 * k + 4]; e < it[6 * k + 4] + it[6 * k + 5]; ++e) cnt[P.item_cols[e]] = 0;
      }
    }
  }

  // per work item, bounds [lo, hi] on the distance between any test point of
  // its block and any source point of its tier, padded for rounding
  std::vector<double> bc(2 * (size_t)nblk), br(nblk);
  for (int32_t b = 0; b < nblk; ++b) {
    std::vector<double> pts;
    for (int k = 0; k < P.blk_ntest[b]; ++k) {
      const double* q = &P.qp7[14 * (size_t)P.blk_tris[(size_t)b * kBlock + k]];
      pts.insert(pts.end(), q, q + 14);
    }
    circle(pts.data(), pts.size() / 2, bc[2 * b], bc[2 * b + 1], br[b]);
  }
  for (int tier = 0; tier < kTiers; ++tier) {
    const std::vector<double>& sp = tier == 0 ? P.qp7 : (tier == 1 ? P.qp3 : P.cen);
    const int K = tier == 0 ? 7 : (tier == 1 ? 3 : 1);
    const size_t n_items = P.items[tier].size() / 6;
    P.item_r[tier].resize(2 * n_items);
    std::vector<double> pts;
    for (size_t it = 0; it < n_items; ++it) {
      const int32_t* row = &P.items[tier][6 * it];
      pts.clear();
      for (int k = 0; k < row[3]; ++k) {
        const double* q = &sp[2 * K * (size_t)P.src[row[2] + k]];
        pts.insert(pts.end(), q, q + 2 * K);
      }
      double sx, sy, sr;
      circle(pts.data(), pts.size() / 2, sx, sy, sr);
      const double D = std::hypot(sx - bc[2 * row[0]], sy - bc[2 * row[0] + 1]);
      const double r = br[row[0]] + sr;
      const double lo = std::max(0.0, (D - r) * (1 - 1e-9) - 1e-12);
      P.item_r[tier][2 * it] = D - r > 0 ? lo : 0.0;
      P.item_r[tier][2 * it + 1] = (D + r) * (1 + 1e-9) + 1e-12;
    }
  }

  // near items: offsets of their statics (test triangles x 7 points x 3 values x sources)
  const size_t n_near = P.items[0].size() / 6;
  P.item_soff.resize(n_near);
  int64_t off = 0;
  for (size_t it = 0; it < n_near; ++it) {
    P.item_soff[it] = off;
    off += 21 * (int64_t)P.blk_ntest[P.items[0][6 * it]] * P.items[0][6 * it + 3];
  }
  P.near_static_len = off;

  P.rec.assign((size_t)nl * nl, PairTab{});
  P.own_x.assign((size_t)nl * nl, {});
  P.own_c.assign((size_t)nl * nl, {});
  P.own_b.assign((size_t)nl * nl, {});
  P.own_set.assign((size_t)nl * nl, 0);
  return 0;
}

// ------------------------------------------------------------- device side
template <typename T>
int upload(Dev& d, const std::vector<T>& h, hipStream_t s) {
  const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
  if (d.bytes < bytes) {
    if (d.p) MOM_TRY(hipFree(d.p));
    d.p = nullptr;
    d.bytes = 0;
    MOM_TRY(hipMalloc(&d.p, bytes));
    d.bytes = bytes;
  }
  if (!h.empty()) MOM_TRY(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

int reserve(Dev& d, size_t bytes) {
  if (d.bytes >= bytes) return 0;
  if (d.p) MOM_TRY(hipFree(d.p));
  d.p = nullptr;
  d.bytes = 0;
  MOM_TRY(hipMalloc(&d.p, bytes));
  d.bytes = bytes;
  return 0;
}

FillArgs base_args(Plan& P) {
  FillArgs a;
  std::memset(&a, 0, sizeof(a));
  a.V = (const double*)P.d_V.p;
  a.qp7 = (const double*)P.d_qp7.p;
  a.qw7 = (const double*)P.d_qw7.p;
  a.qp3 = (const double*)P.d_qp3.p;
  a.qw3 = (const double*)P.d_qw3.p;
  a.cen = (const double*)P.d_cen.p;
  a.area = (const double*)P.d_area.p;
  a.cA = (const double*)P.d_cA.p;
  a.cV = (const double*)P.d_cV.p;
  a.tb = (const int32_t*)P.d_tb.p;
  a.pad_pos = (const int32_t*)P.d_pad_pos.p;
  a.blk_tris = (const int32_t*)P.d_blk_tris.p;
  a.src = (const int32_t*)P.d_src.p;
  a.nt = P.nt;
  a.blk_row_off = (const int32_t*)P.d_blk_row_off.p;
  a.blk_rows = (const int32_t*)P.d_blk_rows.p;
  a.blk_rloc = (const int32_t*)P.d_blk_rloc.p;
  a.item_cols = (const int32_t*)P.d_item_cols.p;
  a.src_cloc = (const int32_t*)P.d_src_cloc.p;
  a.near_static = (const double*)P.d_near_static.p;
  a.item_soff = (const int64_t*)P.d_item_soff.p;
  a.row_own = (const uint8_t*)P.d_row_own.p;
  a.col_own = (const uint8_t*)P.d_col_own.p;
  return a;
}

int to_device(Plan& P, hipStream_t s) {
  MOM_TRY(hipGetDevice(&P.device));
  if (int rc = upload(P.d_V, P.V, s)) return rc;
  if (int rc = upload(P.d_qp7, P.qp7, s)) return rc;
  if (int rc = upload(P.d_qw7, P.qw7, s)) return rc;
  if (int rc = upload(P.d_qp3, P.qp3, s)) return rc;
  if (int rc = upload(P.d_qw3, P.qw3, s)) return rc;
  if (int rc = upload(P.d_cen, P.cen, s)) return rc;
  if (int rc = upload(P.d_area, P.area, s)) return rc;
  if (int rc = upload(P.d_cA, P.cA, s)) return rc;
  if (int rc = upload(P.d_cV, P.cV, s)) return rc;
  if (int rc = upload(P.d_tb, P.tb, s)) return rc;
  if (int rc = upload(P.d_pad_pos, P.pad_pos, s)) return rc;
  if (int rc = upload(P.d_blk_tris, P.blk_tris, s)) return rc;
  if (int rc = upload(P.d_src, P.src, s)) return rc;
  for (int t = 0; t < kTiers; ++t) {
    if (int rc = upload(P.d_items[t], P.items[t], s)) return rc;
    if (int rc = upload(P.d_item_r[t], P.item_r[t], s)) return rc;
  }
  if (int rc = upload(P.d_blk_row_off, P.blk_row_off, s)) return rc;
  if (int rc = upload(P.d_blk_rows, P.blk_rows, s)) return rc;
  if (int rc = upload(P.d_blk_rloc, P.blk_rloc, s)) return rc;
  if (int rc = upload(P.d_item_cols, P.item_cols, s)) return rc;
  if (int rc = upload(P.d_src_cloc, P.src_cloc, s)) return rc;
  if (int rc = upload(P.d_item_soff, P.item_soff, s)) return rc;
  if (int rc = upload(P.d_row_own, P.row_own, s)) return rc;
  if (int rc = upload(P.d_col_own, P.col_own, s)) return rc;
  if (int rc = reserve(P.d_near_static, std::max<int64_t>(P.near_static_len, 1) * sizeof(double))) return rc;
  const int n_near = (int)(P.items[0].size() / 6);
  Dev dz;
  std::vector<double> pair_dz;  // alive until the stream is synchronised
  int rc = 0;
  if (n_near) {
    pair_dz.resize((size_t)P.nl * P.nl);
    for (int i = 0; i < P.nl; ++i)
      for (int j = 0; j < P.nl; ++j) pair_dz[(size_t)i * P.nl + j] = std::fabs(P.zl[i] - P.zl[j]);
    rc = upload(dz, pair_dz, s);
    if (!rc) {
      FillArgs a = base_args(P);
      a.items = (const int32_t*)P.d_items[0].p;
      rc = mom_near_statics(n_near, &a, (const double*)dz.p, (double*)P.d_near_static.p, s);
    }
  }
  const hipError_t e = hipStreamSynchronize(s);
  if (dz.p) (void)hipFree(dz.p);
  if (rc) return rc;
  if (e != hipSuccess) return fail(std::string("mom_plan_create: ") + hipGetErrorString(e));
  P.on_device = true;
  return 0;
}

// Interval lookup of the fill kernel for the nodes x (the same table the
// host side of this library builds for its own sweeps): entry b serves the
// distances with (int)(R / dxb) == b in double arithmetic, holding the
// spline interval at the start of that range and the one node inside it.
int bucket_table(const double* x, int n, double dxb, std::vector<Bucket>& out) {
  const double inv = 1.0 / dxb;
  const int64_t nb = (int64_t)(x[n - 1] * inv) + 2;
  if (nb > (int64_t)1 << 26) return fail("mom_plan_set_pair: node spacing too fine for the bucket table");
  std::vector<double> lo(nb);
  for (int64_t b = 0; b < nb; ++b) {
    const double bd = (double)b;
    double v = bd * dxb;
    for (int it = 0; it < 64; ++it) {  // the smallest double with floor(v * inv) >= b
      double dn = std::nextafter(v, -INFINITY);
      if (std::floor(dn * inv) >= bd && dn >= 0) v = dn;
      else break;
    }
    for (int it = 0; it < 64; ++it) {
      if (std::floor(v * inv) < bd) v = std::nextafter(v, INFINITY);
      else break;
    }
    lo[b] = v;
  }
  out.assign(nb, Bucket{});
  for (int64_t b = 0; b < nb; ++b) {
    const double hi = b + 1 < nb ? lo[b + 1] : INFINITY;
    const int64_t raw = (int64_t)(std::upper_bound(x, x + n, lo[b]) - x) - 1;
    const int64_t j = raw + 1;
    const int64_t inside = (int64_t)(std::lower_bound(x, x + n, hi) - x) - j;
    if (inside > 1) return fail("mom_plan_set_pair: a bucket holds more than one node");
    out[b].i0 = (int32_t)std::min<int64_t>(std::max<int64_t>(raw, 0), n - 2);
    const bool moves = inside >= 1 && raw >= 0 && j <= n - 2;
    out[b].split = moves ? x[std::min<int64_t>(j, n - 1)] : INFINITY;
    out[b].pad_ = 0;
  }
  return 0;
}

// The plan's own tables concatenated in the layout of the fill kernel.
int own_tables(Plan& P) {
  const size_t np = (size_t)P.nl * P.nl;
  for (size_t q = 0; q < np; ++q)
    if (P.pair_used[q] && !P.own_set[q]) return fail("mom_plan_fill: a level pair the fill reads has no spline set");
    else if (P.pair_used[q] && P.rec[q].zero)  // the reference forms no groups for uncoupled pairs
      return fail("mom_plan_fill: a level pair the fill reads is marked as not coupled");
  if (!P.own_dirty) return 0;
  P.cat_x.clear(), P.cat_c.clear(), P.cat_b.clear();
  P.cat_rec = P.rec;
  for (size_t q = 0; q < np; ++q) {
    if (!P.own_set[q] || P.cat_rec[q].zero) continue;
    P.cat_rec[q].x_off = (int64_t)P.cat_x.size();
    P.cat_rec[q].c_off = (int64_t)P.cat_c.size();
    P.cat_rec[q].b_off = (int64_t)P.cat_b.size();
    P.cat_x.insert(P.cat_x.end(), P.own_x[q].begin(), P.own_x[q].end());
    P.cat_c.insert(P.cat_c.end(), P.own_c[q].begin(), P.own_c[q].end());
    P.cat_b.insert(P.cat_b.end(), P.own_b[q].begin(), P.own_b[q].end());
  }
  P.own_dirty = false;
  P.own_uploaded = false;
  return 0;
}

int upload_own_tables(Plan& P, hipStream_t s) {
  if (int rc = own_tables(P)) return rc;
  if (P.own_uploaded) return 0;
  if (int rc = upload(P.d_pairs, P.cat_rec, s)) return rc;
  if (int rc = upload(P.d_tab_x, P.cat_x, s)) return rc;
  if (int rc = upload(P.d_tab_c, P.cat_c, s)) return rc;
  if (int rc = upload(P.d_tab_b, P.cat_b, s)) return rc;
  MOM_TRY(hipStreamSynchronize(s));  // the copies read host memory that the next set_pair may change
  P.own_uploaded = true;
  return 0;
}

int check_plan(void* plan, const char* what) {
  if (!plan) return fail(std::string(what) + ": null plan");
  Plan& P = *static_cast<Plan*>(plan);
  if (!P.on_device) return fail(std::string(what) + ": the plan was created without a device copy");
  int dev = -1;
  MOM_TRY(hipGetDevice(&dev));
  if (dev != P.device) return fail(std::string(what) + ": the current device is not the plan's");
  return 0;
}

}  // namespace
}  // namespace mom

using namespace mom;

extern "C" int mom_plan_struct_sizes(int32_t* out) {
  if (!out) return fail("mom_plan_struct_sizes: null output");
  out[0] = (int32_t)sizeof(MomAssemblyGeometry);
  out[1] = (int32_t)sizeof(MomFillCall);
  out[2] = (int32_t)sizeof(MomPlanInfo);
  return 0;
}

extern "C" int mom_plan_create(const MomAssemblyGeometry* g, int on_device, void* stream, void** plan) {
  if (!g || !plan) return fail("mom_plan_create: null argument");
  *plan = nullptr;
  Plan* P = new (std::nothrow) Plan();
  if (!P) return fail("mom_plan_create: out of host memory");
  int rc = 0;
  try {
    rc = build(*P, *g);
  } catch (const std::bad_alloc&) {
    rc = fail("mom_plan_create: out of host memory");
  }
  if (!rc && on_device) rc = to_device(*P, (hipStream_t)stream);
  if (rc) {
    mom_plan_destroy(P);
    return rc;
  }
  *plan = P;
  return 0;
}

extern "C" int mom_plan_destroy(void* plan) {
  if (!plan) return 0;
  Plan* P = static_cast<Plan*>(plan);
  int rc = 0;
  for (Dev* d : P->all())
    if (d->p && hipFree(d->p) != hipSuccess) rc = fail("mom_plan_destroy: hipFree failed");
  if (P->host_stream && hipStreamDestroy(P->host_stream) != hipSuccess) rc = fail("mom_plan_destroy: stream");
  delete P;
  return rc;
}

extern "C" int mom_plan_info(const void* plan, MomPlanInfo* info) {
  if (!plan || !info) return fail("mom_plan_info: null argument");
  const Plan& P = *static_cast<const Plan*>(plan);
  std::memset(info, 0, sizeof(*info));
  for (int t = 0; t < kTiers; ++t) {
    info->n_items[t] = (int32_t)(P.items[t].size() / 6);
    info->acc_bytes[t] = P.acc_bytes[t];
    info->tier_pairs[t] = P.tier_pairs[t];
  }
  info->n_pad = P.n_pad;
  info->near_static_len = P.near_static_len;
  info->n_src = (int64_t)P.src.size();
  return 0;
}

extern "C" int mom_plan_host_array(const void* plan, const char* name, const void** data, int64_t* count) {
  if (!plan || !name || !data || !count) return fail("mom_plan_host_array: null argument");
  const Plan& P = *static_cast<const Plan*>(plan);
  const std::string n(name);
  auto out = [&](const auto& v) {
    *data = v.data();
    *count = (int64_t)v.size();
    return 0;
  };
  if (n == "src") return out(P.src);
  if (n == "items0") return out(P.items[0]);
  if (n == "items1") return out(P.items[1]);
  if (n == "items2") return out(P.items[2]);
  if (n == "item_r0") return out(P.item_r[0]);
  if (n == "item_r1") return out(P.item_r[1]);
  if (n == "item_r2") return out(P.item_r[2]);
  if (n == "blk_tris") return out(P.blk_tris);
  if (n == "blk_row_off") return out(P.blk_row_off);
  if (n == "blk_rows") return out(P.blk_rows);
  if (n == "blk_rloc") return out(P.blk_rloc);
  if (n == "item_cols") return out(P.item_cols);
  if (n == "src_cloc") return out(P.src_cloc);
  if (n == "item_soff") return out(P.item_soff);
  if (n == "row_own") return out(P.row_own);
  if (n == "col_own") return out(P.col_own);
  if (n == "tab_x" || n == "tab_c" || n == "tab_b" || n == "pairs") {  // the own tables (mom_plan_set_pair)
    Plan& Q = const_cast<Plan&>(P);
    if (int rc = own_tables(Q)) return rc;
    if (n == "tab_x") return out(Q.cat_x);
    if (n == "tab_c") return out(Q.cat_c);
    if (n == "tab_b") return out(Q.cat_b);
    return out(Q.cat_rec);
  }
  return fail("mom_plan_host_array: unknown array " + n);
}

extern "C" int mom_plan_set_pair(void* plan, int lt, int ls, int n, const double* x, const double* c, double dz,
                                 double cA_re, double cA_im, double cV_re, double cV_im, int zero) {
  if (!plan) return fail("mom_plan_set_pair: null plan");
  Plan& P = *static_cast<Plan*>(plan);
  if (lt < 0 || lt >= P.nl || ls < 0 || ls >= P.nl) return fail("mom_plan_set_pair: level outside [0, n_levels)");
  const size_t q = (size_t)lt * P.nl + ls;
  PairTab r;
  std::memset(&r, 0, sizeof(r));
  r.dz = dz;
  r.cA_re = cA_re, r.cA_im = cA_im, r.cV_re = cV_re, r.cV_im = cV_im;
  r.zero = zero ? 1 : 0;
  if (!zero) {
    if (n < 2 || !x || !c) return fail("mom_plan_set_pair: a spline needs n >= 2 nodes and its coefficients");
    double dmin = INFINITY;
    for (int i = 0; i + 1 < n; ++i) {
      if (!(x[i + 1] > x[i])) return fail("mom_plan_set_pair: nodes must increase");
      dmin = std::min(dmin, x[i + 1] - x[i]);
    }
    if (!(x[0] >= 0)) return fail("mom_plan_set_pair: nodes must be distances (>= 0)");
    const double dxb = 0.5 * dmin;
    std::vector<Bucket> b;
    if (int rc = bucket_table(x, n, dxb, b)) return rc;
    // c: PPoly coefficients (4, n-1, 2) complex -> per interval [power][kernel][re, im]
    std::vector<double> cc(16 * (size_t)(n - 1));
    for (int p = 0; p < 4; ++p)
      for (int i = 0; i + 1 < n; ++i)
        for (int m = 0; m < 2; ++m)
          for (int k = 0; k < 2; ++k)
            cc[16 * (size_t)i + 4 * p + 2 * m + k] = c[(((size_t)p * (n - 1) + i) * 2 + m) * 2 + k];
    P.own_x[q].assign(x, x + n);
    P.own_c[q] = std::move(cc);
    P.own_b[q] = std::move(b);
    r.n = n;
    r.nbucket = (int32_t)P.own_b[q].size();
    r.inv_dx = 1.0 / dxb;
  }
  P.rec[q] = r;
  P.own_set[q] = 1;
  P.own_dirty = true;
  return 0;
}

extern "C" int mom_plan_fill(void* plan, const MomFillCall* c, void* stream) {
  if (int rc = check_plan(plan, "mom_plan_fill")) return rc;
  if (!c || !c->Z) return fail("mom_plan_fill: null call block or Z");
  Plan& P = *static_cast<Plan*>(plan);
  hipStream_t s = (hipStream_t)stream;
  if (c->ldz < P.nb) return fail("mom_plan_fill: ldz below the number of RWG functions");
  if (P.n_pad && !c->IVpad) return fail("mom_plan_fill: the mesh has pad triangles: IVpad is required");
  if (c->stage_max < 0 || c->stage_max > 64) return fail("mom_plan_fill: stage_max outside [0, 64]");
  if ((uintptr_t)c->Z % 16) return fail("mom_plan_fill: Z must be 16-byte aligned (complex128)");
  FillArgs a = base_args(P);
  if (c->pairs) {
    if (!c->tab_x || !c->tab_c || !c->tab_b) return fail("mom_plan_fill: incomplete device tables");
    a.pairs = (const PairTab*)c->pairs;
    a.tab_x = c->tab_x;
    a.tab_c = c->tab_c;
    a.tab_b = (const Bucket*)c->tab_b;
  } else {
    if (int rc = upload_own_tables(P, s)) return rc;
    a.pairs = (const PairTab*)P.d_pairs.p;
    a.tab_x = (const double*)P.d_tab_x.p;
    a.tab_c = (const double*)P.d_tab_c.p;
    a.tab_b = (const Bucket*)P.d_tab_b.p;
  }
  a.jwmu_re = c->jwmu_re, a.jwmu_im = c->jwmu_im;
  a.ijwe_re = c->ijwe_re, a.ijwe_im = c->ijwe_im;
  a.Z = (double*)c->Z;
  a.ldz = c->ldz;
  a.IVpad = (double*)c->IVpad;
  a.stage_max = c->stage_max ? c->stage_max : 64;
  const int mask = c->tier_mask ? c->tier_mask : 7;
  for (int t = 0; t < kTiers; ++t) {
    const int n_items = (int)(P.items[t].size() / 6);
    if (!(mask & (1 << t)) || !n_items) continue;
    a.items = (const int32_t*)P.d_items[t].p;
    a.item_r = (const double*)P.d_item_r[t].p;
    a.acc_bytes = P.acc_bytes[t];
    if (int rc = mom_fill_tier(t, n_items, &a, s)) return rc;
  }
  return 0;
}

extern "C" int mom_plan_fill_host(void* plan, const MomFillCall* c) {
  if (int rc = check_plan(plan, "mom_plan_fill_host")) return rc;
  if (!c || !c->Z) return fail("mom_plan_fill_host: null call block or Z");
  Plan& P = *static_cast<Plan*>(plan);
  if (c->ldz < P.nb) return fail("mom_plan_fill_host: ldz below the number of RWG functions");
  if (P.n_pad && !c->IVpad) return fail("mom_plan_fill_host: the mesh has pad triangles: IVpad is required");
  if (!P.host_stream) MOM_TRY(hipStreamCreateWithFlags(&P.host_stream, hipStreamNonBlocking));
  hipStream_t s = P.host_stream;
  const size_t zb = (size_t)P.nb * P.nb * 16, ivb = (size_t)P.n_pad * P.nt * 16;
  if (int rc = reserve(P.d_Z, std::max<size_t>(zb, 16))) return rc;
  if (int rc = reserve(P.d_IV, std::max<size_t>(ivb, 16))) return rc;
  MOM_TRY(hipMemsetAsync(P.d_Z.p, 0, zb, s));
  if (ivb) MOM_TRY(hipMemsetAsync(P.d_IV.p, 0, ivb, s));
  MomFillCall d = *c;
  d.Z = P.d_Z.p;
  d.ldz = P.nb;
  d.IVpad = ivb ? P.d_IV.p : nullptr;
  if (int rc = mom_plan_fill(plan, &d, s)) return rc;
  if (zb)
    MOM_TRY(hipMemcpy2DAsync(c->Z, (size_t)c->ldz * 16, P.d_Z.p, (size_t)P.nb * 16, (size_t)P.nb * 16, P.nb,
                             hipMemcpyDeviceToHost, s));
  if (ivb) MOM_TRY(hipMemcpyAsync(c->IVpad, P.d_IV.p, ivb, hipMemcpyDeviceToHost, s));
  MOM_TRY(hipStreamSynchronize(s));
  return 0;
}

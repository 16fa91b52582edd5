// batch_layout.hpp - host side of the batched engine's data layout: per-LP offsets, the
// concatenated and padded matrix arrays, and every LP's work plans in every workgroup geometry.
// Host code only (the device reads LpDesc / LpPlan as laid out here); kept apart from the HIP
// calls so that it runs, and is tested, without a device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "plan_host.hpp"

namespace sankhya_hip {
namespace batch {

constexpr int kGeos = 3;
constexpr int kGeoThreads[kGeos] = {256, 512, 1024};

// One LP's plan in one geometry: ranges in that geometry's concatenated plan arrays.
struct LpPlan {
  int32_t item0, nitems, lr0, nlr, grp0, ngrp, seg0, pad;
};

// One LP: sizes, offsets into the concatenated arrays, and its plans per geometry.
struct LpDesc {
  int32_t m, n, row0, col0;
  int64_t nnz0;    // first nonzero of the LP in the concatenated col/val arrays (even)
  int32_t arp0;    // first entry of its rows' pointers (m + 1 entries)
  int32_t trp0;    // first entry of its columns' pointers (n + 1 entries)
  LpPlan a[kGeos], t[kGeos];
};

struct PlanSet {  // one matrix side in one geometry, all LPs
  std::vector<int4> items, lr_info, groups;
  int64_t segs = 0;
};

struct BatchLayout {
  std::vector<LpDesc> desc;
  int64_t sum_m = 0, sum_n = 0, sum_nnz = 0, padded_nnz = 0;
  std::vector<int32_t> a_col, t_col;  // padded_nnz + 2 entries each
  std::vector<double> a_val, t_val;
  PlanSet plans[2][kGeos];  // [0]: rows of Ahat, [1]: rows of Ahat^T
};

// Checks the batch's arrays (layout as in include/pdhg_batch.h) and builds the layout. Returns
// an empty string, or what is wrong.
inline std::string build_batch_layout(int32_t nlp, const int32_t* m, const int32_t* n,
                                      const int64_t* nnz, const int32_t* a_rowptr,
                                      const int32_t* a_col, const double* a_val,
                                      const int32_t* at_rowptr, const int32_t* at_col,
                                      const double* at_val, BatchLayout* L) {
  if (nlp <= 0 || !m || !n || !nnz || !a_rowptr || !at_rowptr) return "bad batch arguments";
  L->desc.assign(static_cast<size_t>(nlp), LpDesc{});
  int64_t sm = 0, sn = 0, nz = 0, rawnz = 0, arp = 0, trp = 0;
  for (int32_t k = 0; k < nlp; ++k) {
    if (m[k] < 0 || n[k] < 0 || nnz[k] < 0 || nnz[k] > 0x7fffffffLL)
      return "bad dimensions of LP " + std::to_string(k);
    LpDesc& d = L->desc[k];
    d.m = m[k];
    d.n = n[k];
    d.row0 = static_cast<int32_t>(sm);
    d.col0 = static_cast<int32_t>(sn);
    d.nnz0 = nz;
    d.arp0 = static_cast<int32_t>(arp);
    d.trp0 = static_cast<int32_t>(trp);
    sm += m[k];
    sn += n[k];
    nz += (nnz[k] + 1) & ~int64_t{1};
    arp += m[k] + 1;
    trp += n[k] + 1;
    rawnz += nnz[k];
    if (sm > 0x7fffffffLL || sn > 0x7fffffffLL || arp > 0x7fffffffLL || trp > 0x7fffffffLL)
      return "batch too large for 32-bit offsets";
  }
  if (rawnz > 0 && (!a_col || !a_val || !at_col || !at_val)) return "null matrix pointer";
  L->sum_m = sm;
  L->sum_n = sn;
  L->sum_nnz = rawnz;
  L->padded_nnz = nz;
  {
    int64_t off = 0, ap = 0, tp = 0;
    for (int32_t k = 0; k < nlp; ++k) {
      const int32_t* rp = a_rowptr + ap;
      const int32_t* tpp = at_rowptr + tp;
      if (rp[0] != 0 || rp[m[k]] != nnz[k] || tpp[0] != 0 || tpp[n[k]] != nnz[k])
        return "row pointers of LP " + std::to_string(k) + " do not match nnz";
      for (int32_t i = 0; i < m[k]; ++i)
        if (rp[i + 1] < rp[i]) return "row pointer not monotone in LP " + std::to_string(k);
      for (int32_t j = 0; j < n[k]; ++j)
        if (tpp[j + 1] < tpp[j]) return "transpose pointer not monotone in LP " + std::to_string(k);
      for (int64_t q = 0; q < nnz[k]; ++q)
        if (a_col[off + q] < 0 || a_col[off + q] >= n[k] || at_col[off + q] < 0 ||
            at_col[off + q] >= m[k])
          return "index out of range in LP " + std::to_string(k);
      off += nnz[k];
      ap += m[k] + 1;
      tp += n[k] + 1;
    }
  }
  // Each LP's nonzeros start at an even offset so pair loads stay aligned; two zero entries
  // close the arrays.
  const size_t total = static_cast<size_t>(nz) + 2;
  L->a_col.assign(total, 0);
  L->t_col.assign(total, 0);
  L->a_val.assign(total, 0.0);
  L->t_val.assign(total, 0.0);
  {
    int64_t src = 0;
    for (int32_t k = 0; k < nlp; ++k) {
      const int64_t dst = L->desc[k].nnz0;
      std::copy(a_col + src, a_col + src + nnz[k], L->a_col.begin() + dst);
      std::copy(at_col + src, at_col + src + nnz[k], L->t_col.begin() + dst);
      std::copy(a_val + src, a_val + src + nnz[k], L->a_val.begin() + dst);
      std::copy(at_val + src, at_val + src + nnz[k], L->t_val.begin() + dst);
      src += nnz[k];
    }
  }
  // Plans: every LP in every geometry, no column bands (one workgroup walks an LP's plan).
  for (int g = 0; g < kGeos; ++g) {
    const int nt = kGeoThreads[g];
    PlanGeo pg;
    pg.tile = 8 * nt;
    pg.seg = 8 * nt;
    pg.max_rows = 2 * nt;
    pg.tail = nt;
    pg.bands = 1;
    pg.threads = nt;
    for (int side = 0; side < 2; ++side) {
      PlanSet& ps = L->plans[side][g];
      int64_t src = 0, rp_off = 0;
      for (int32_t k = 0; k < nlp; ++k) {
        LpDesc& d = L->desc[k];
        const int32_t rows = side == 0 ? m[k] : n[k];
        const int32_t cols = side == 0 ? n[k] : m[k];
        const int32_t* rp = (side == 0 ? a_rowptr : at_rowptr) + rp_off;
        const int32_t* ci = (side == 0 ? a_col : at_col) + src;
        std::vector<int4> logical;
        const HostPlan hp = build_plan(rows, cols, rp, ci, &logical, pg);
        if (!plan_ok(hp, logical, rows, rp, pg))
          return "internal: batch plan does not cover LP " + std::to_string(k);
        LpPlan& p = side == 0 ? d.a[g] : d.t[g];
        p.item0 = static_cast<int32_t>(ps.items.size());
        p.nitems = static_cast<int32_t>(hp.items.size());
        p.lr0 = static_cast<int32_t>(ps.lr_info.size());
        p.nlr = static_cast<int32_t>(hp.lr_nseg.size());
        p.grp0 = static_cast<int32_t>(ps.groups.size());
        p.ngrp = static_cast<int32_t>(hp.groups.size());
        p.seg0 = static_cast<int32_t>(ps.segs);
        ps.items.insert(ps.items.end(), hp.items.begin(), hp.items.end());
        for (size_t q = 0; q < hp.lr_nseg.size(); ++q)
          ps.lr_info.push_back(make_int4(hp.lr_row[q], hp.lr_base[q], hp.lr_nseg[q], 0));
        ps.groups.insert(ps.groups.end(), hp.groups.begin(), hp.groups.end());
        ps.segs += hp.nseg;
        src += nnz[k];
        rp_off += rows + 1;
      }
    }
  }
  return "";
}

}  // namespace batch
}  // namespace sankhya_hip

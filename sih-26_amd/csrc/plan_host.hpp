// plan_host.hpp - host-side work plans over a CSR matrix (see pdhg_kernels.hpp for how the
// kernels walk them): stream items of whole rows, long-row segments, XCD column bands, tail
// groups, and the coverage check every plan passes before a kernel sees it.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "pdhg_kernels.hpp"

namespace sankhya_hip {

struct HostPlan {
  std::vector<int4> items;
  std::vector<int32_t> lr_nseg, lr_base, lr_row;
  std::vector<int4> groups;  // {first, last) long rows, first segment, segment count}: rows
                             // whose segment sums fit one LDS round, at most kTail of them
  int32_t nseg = 0;
};

// Work items over a CSR matrix: whole rows packed below kTile nonzeros and up to kMaxRows rows
// per item; rows of kTile nonzeros or more split into segments of fewer than kSeg nonzeros that
// never cross a column band (ncols split into kXcds equal bands). Staying one short of kTile and
// kSeg keeps an item's span from its pair-aligned first nonzero within kTile (kSeg), so the
// kernels cover it in exactly kTile / (2 threads) two-nonzero steps: an odd start would
// otherwise need one step more, which costs every lane its registers whether used or not.
//
// XCD-aware dispatch: workgroup b runs on XCD b % kXcds, and each XCD has its own L2. The
// segments of band k are placed at item positions congruent to k, so every XCD's L2 only has to
// hold its band of the gathered vector rather than all of it. The `logical` order (rows
// ascending, segments of a row in column order) is kept for the coverage check.
// Item geometry on the host side: nonzeros per stream item, per long-row segment, rows per
// stream item, tail workgroup size (rows per tail group), and the number of column bands
// (1 = no XCD banding, for plans that one workgroup walks alone).
struct PlanGeo {
  int tile = kTile, seg = kSeg, max_rows = kMaxRows, tail = kTail, bands = 8, threads = kThreads;
};

// Column bands of a single-LP plan: 8 (one per XCD) times the smallest power of two that keeps
// one band of the gathered vector of (x, x') or (y, y') pairs within kBandBytes, half of an
// XCD's 4 MB L2 (the other half is left to the streamed matrix). On the Zipf LPs: 200k x 500k
// in FP64 (8 MB of x pairs) runs k_dual fastest with 8 bands; 1M x 2M in FP64 (32 MB) with 16,
// 16% faster than with 8; 1M x 2M in FP32 (16 MB) with 8.
constexpr int64_t kBandBytes = int64_t{2} << 20;
inline PlanGeo single_geo(int64_t gathered_entries, int pair_bytes) {
  PlanGeo g;
  int per_xcd = 1;
  while (per_xcd < 64 && gathered_entries * pair_bytes > int64_t{8} * per_xcd * kBandBytes) per_xcd *= 2;
  g.bands = 8 * per_xcd;
  return g;
}

// A stream item's rows are summed by L lanes each, L the largest power of two (at most 64) with
// rows * L <= threads; one lane then walks ceil(len / L) products of a row serially. An item is
// closed before a row that would take that past kSerialSum (twice what a single full-tile row
// costs at 256 threads), so a long row never lands among hundreds of tiny ones with one lane to
// sum it: such an item, 300 empty rows and a 1,451-nonzero row, made a whole Zipf-LP launch 4x
// slower. Plans whose rows come in length order (generated Zipf rows, uniform rows) never reach
// the cap and are unchanged by it.
constexpr int kSerialSum = 64;
inline int lanes_per_row(int rows, int threads) {
  int L = 64;
  while (L > 1 && rows * L > threads) L >>= 1;
  return L;
}

inline HostPlan build_plan(int32_t rows, int32_t ncols, const int32_t* rowptr, const int32_t* col,
                    std::vector<int4>* logical, const PlanGeo& g = PlanGeo{}) {
  // Column bands: g.bands of them, band b on XCD b % xcds; an XCD works through its bands in
  // band order (all of band x's segments, then band x + xcds's, ...), so its L2 holds one band
  // of the gathered vector at a time.
  const int nbands = std::max(1, g.bands);
  const int kXcds = std::min(nbands, 8);
  HostPlan hp;
  std::vector<int4> stream;
  std::vector<std::vector<int4>> band(kXcds);  // per XCD, its bands' segments in band order
  std::vector<std::vector<int4>> by_band(nbands);
  const int64_t width = std::max<int64_t>(1, (static_cast<int64_t>(ncols) + nbands - 1) / nbands);
  int32_t r = 0;
  while (r < rows) {
    const int32_t len = rowptr[r + 1] - rowptr[r];
    if (len >= g.tile) {
      const int32_t lr = static_cast<int32_t>(hp.lr_nseg.size());
      int32_t nseg = 0;
      int32_t k = rowptr[r];
      while (k < rowptr[r + 1]) {
        const int b = static_cast<int>(std::min<int64_t>(col[k] / width, nbands - 1));
        const int64_t band_end_col = (b + 1) * width;
        int32_t kb = k;  // first nonzero of the next band
        while (kb < rowptr[r + 1] && col[kb] < band_end_col) ++kb;
        for (int32_t k0 = k; k0 < kb; k0 += g.seg - 1) {
          const int32_t k1 = std::min(k0 + g.seg - 1, kb);
          const int4 item = make_int4(hp.nseg + nseg, -lr - 1, k0, k1);
          by_band[b].push_back(item);
          logical->push_back(item);
          ++nseg;
        }
        k = kb;
      }
      hp.lr_nseg.push_back(nseg);
      hp.lr_base.push_back(hp.nseg);
      hp.lr_row.push_back(r);
      hp.nseg += nseg;
      ++r;
      continue;
    }
    const int32_t r0 = r;
    int64_t acc = 0;
    int32_t longest = 0;
    while (r < rows && r - r0 < g.max_rows) {
      const int32_t l = rowptr[r + 1] - rowptr[r];
      if (acc + l >= g.tile) break;
      const int32_t lg = std::max(longest, l);
      const int L = lanes_per_row(r - r0 + 1, g.threads);
      if (r > r0 && (lg + L - 1) / L > kSerialSum) break;
      longest = lg;
      acc += l;
      ++r;
    }
    const int4 item = make_int4(r0, r, rowptr[r0], rowptr[r]);
    stream.push_back(item);
    logical->push_back(item);
  }
  for (int b = 0; b < nbands; ++b)
    band[b % kXcds].insert(band[b % kXcds].end(), by_band[b].begin(), by_band[b].end());
  // Dispatch order: stream items spread evenly through the segments (a stream item is placed
  // whenever the share of stream items placed so far is at most the share of segments), so the
  // latency-bound short rows run beside the bandwidth-bound long-row segments instead of after
  // them. A segment goes to position p on XCD p % kXcds: that band's next segment, else a
  // stream item, else the fullest band's.
  std::vector<size_t> next(kXcds, 0);
  size_t ns = 0;
  const size_t total = logical->size();
  const size_t nsegs = total - stream.size();
  size_t segs_placed = 0;
  for (size_t p = 0; p < total; ++p) {
    const int x = static_cast<int>(p % kXcds);
    if (ns < stream.size() &&
        (nsegs == 0 || static_cast<double>(ns) * static_cast<double>(nsegs) <=
                           static_cast<double>(segs_placed) * static_cast<double>(stream.size()))) {
      hp.items.push_back(stream[ns++]);
      continue;
    }
    ++segs_placed;
    if (next[x] < band[x].size()) {
      hp.items.push_back(band[x][next[x]++]);
    } else if (ns < stream.size()) {
      hp.items.push_back(stream[ns++]);
    } else {
      int best = 0;
      for (int b = 1; b < kXcds; ++b)
        if (band[b].size() - next[b] > band[best].size() - next[best]) best = b;
      hp.items.push_back(band[best][next[best]++]);
    }
  }
  const int32_t nlr = static_cast<int32_t>(hp.lr_nseg.size());
  for (int32_t lr = 0; lr < nlr;) {
    int32_t end = lr + 1;
    while (end < nlr && end - lr < g.tail &&
           hp.lr_base[end] + hp.lr_nseg[end] - hp.lr_base[lr] <= kChunk)
      ++end;
    hp.groups.push_back(make_int4(lr, end, hp.lr_base[lr],
                                  hp.lr_base[end - 1] + hp.lr_nseg[end - 1] - hp.lr_base[lr]));
    lr = end;
  }
  return hp;
}


// Every row covered exactly once (in the logical order), stream items below the LDS tile and
// within the row cap, long segments below kSeg and numbered consecutively per row, and the
// dispatch order a permutation of the logical one: checked on the host before any kernel sees
// the plan.
inline bool plan_ok(const HostPlan& hp, const std::vector<int4>& logical, int32_t rows,
             const int32_t* rowptr, const PlanGeo& g = PlanGeo{}) {
  int32_t next_row = 0, next_k = -1;
  for (const int4& it : logical) {
    if (it.y >= 0) {
      if (next_k >= 0) return false;
      if (it.x != next_row || it.y <= it.x || it.y - it.x > g.max_rows) return false;
      if (it.z != rowptr[it.x] || it.w != rowptr[it.y] || it.w - it.z >= g.tile) return false;
      next_row = it.y;
    } else {
      const int32_t lr = -it.y - 1;
      if (lr < 0 || lr >= static_cast<int32_t>(hp.lr_nseg.size())) return false;
      const int32_t row = hp.lr_row[lr];
      if (row != next_row) return false;
      if (it.x < hp.lr_base[lr] || it.x >= hp.lr_base[lr] + hp.lr_nseg[lr]) return false;
      if (it.z != (next_k < 0 ? rowptr[row] : next_k) || it.w <= it.z || it.w - it.z >= g.seg) return false;
      next_k = it.w;
      if (it.w == rowptr[row + 1]) {
        next_row = row + 1;
        next_k = -1;
      }
    }
  }
  if (hp.items.size() != logical.size()) return false;
  // A single long row must fit one LDS round of the tail kernels (kChunk segment sums).
  for (const int32_t ns : hp.lr_nseg)
    if (ns > kChunk) return false;
  // Tail groups: consecutive, covering every long row, at most one row per tail lane and at
  // most kChunk segment sums each.
  int32_t next_lr = 0;
  for (const int4& q : hp.groups) {
    if (q.x != next_lr || q.y <= q.x || q.y - q.x > g.tail || q.w > kChunk) return false;
    if (q.z != hp.lr_base[q.x] || q.w != hp.lr_base[q.y - 1] + hp.lr_nseg[q.y - 1] - q.z)
      return false;
    next_lr = q.y;
  }
  if (next_lr != static_cast<int32_t>(hp.lr_nseg.size())) return false;
  return next_row == rows && next_k < 0;
}


}  // namespace sankhya_hip

// BVH construction on the device with the reference's split rule (RC/src/spatial/bvh.rs:126
// build_with): at depth d the elements of a node are ordered along axis d mod 3 by (total order key
// of the coordinate, element index); the split k minimises diam²(box of the first k) +
// diam²(box of the rest) over k in [⌈n/4⌉, n − ⌈n/4⌉], the first minimum winning; nodes of at most
// leaf_size elements are leaves.  The result is bitwise the host build's (box unions are exact,
// the diameters are evaluated with the same operations, no contraction).
//
// Device form, level by level: each axis is sorted once by (key, index) with a radix sort; a level
// orders every splitting node's elements along its axis by a stable radix sort of that global
// order on the node index (a stable partition), then one workgroup per node scans the suffix boxes,
// the prefix boxes and the split costs and reduces to the first minimum.  The host keeps the node
// list (a few kilobytes per level) and numbers the nodes in pre-order at the end.
#pragma once
#include "runtime.hpp"
#include "trees.hpp"

namespace ddgrs {

struct BvhLevel;

struct GpuBvhBuilder {
  size_t cap = 0;
  DevBuf<uint64_t> key_in, key_out;
  DevBuf<uint32_t> iota, axis_order, order, seg_of, part_key_in, part_key_out, part_val_in, part_val_out;
  DevBuf<uint32_t> act_start, act_end, act_prefix, act_split;
  DevBuf<double> suf;  // 6 per element: suffix box lo, hi
  DevBuf<double> box_lo, box_hi;  // element boxes when the caller has none on the device
  // cluster aggregates of a built tree on the device: its topology (order, start, end, right,
  // depth, leaves) and the per-node sums and records
  DevBuf<uint32_t> cl_topo;
  DevBuf<double> cl_buf;
  DevBuf<uint8_t> temp;
  // chunked split of large nodes: per chunk its node and first element, per node its first chunk
  DevBuf<uint32_t> ch_seg, ch_c0, seg_chunk0;
  DevBuf<double> ch_box;    // per chunk 3 boxes of 6 doubles: total, earlier chunks, later chunks
  DevBuf<uint64_t> ch_best; // per chunk (cost, split): 2 words
  std::vector<uint32_t> c_seg, c_c0, c_first;
  // device-driven levels: node records (first element, end, element prefix, split; n + 1 each),
  // per-level counts, and the largest node of each level of the previous build (sizes the launches)
  static constexpr uint32_t kMaxBvhLevels = 96;
  DevBuf<uint32_t> rec;
  // the cut: per axis the cut nodes' elements in axis order, per cut node its splits and their count
  DevBuf<uint32_t> cut_lists, sub_split, sub_count;
  DevBuf<struct BvhLevel> lv_dev;
  std::vector<uint32_t> prev_levels;
  // Builds over n elements with centres pts and boxes [lo, hi] (3 FP64 each, device pointers).  The
  // levels are enqueued without reading back the splits (one read-back when the previous build's
  // depth is reached, the splits once at the end); DDGRS_BVH_LEVELS_ON_HOST=1 reads each level's
  // splits back and keeps the node list on the host (the same tree).
  Topology build(ddgrs_ctx* ctx, size_t n, const double* pts, const double* lo, const double* hi, size_t leaf_size);
  // Levels above a cut depth as below, the subtrees of the cut's nodes one workgroup each (from the
  // second build on, the cut placed from the previous build's node sizes; with allow_cut false every
  // level runs, and a cut node too large for one workgroup makes the level loop go on from the cut).
  Topology build_levels_on_device(ddgrs_ctx* ctx, size_t n, const double* lo, const double* hi, size_t leaf_size,
                                  bool allow_cut = true);
};

}  // namespace ddgrs

// Continuous collision detection on the device (crates/ddg-repulsive-surfaces/src/ccd.rs), mesh
// against mesh:
//   * the broad phase per time window: swept boxes of vertices, faces and edges, and two static box
//     trees (faces, edges) built on the device as the reference builds them (Morton keys of the box
//     centres sorted with their index, blocks of `leaf` boxes merged pairwise: BoxTree :268);
//   * the narrow phase: one thread per query (vertex against the face tree, edge against the edge
//     tree) walks the tree in the reference's order (query :359: explicit stack, right child first)
//     and advances each candidate pair conservatively (advance :529) with the horizon of the
//     query's best contact so far, in FP64 without contraction, so every pair's result and count is
//     the reference's;
//   * the combination: the earliest contact by (τ, primitive), integer sums of the counts.
// The static queries of the step's exempt band (intersecting_pairs :1229, IntersectionBand::new
// :1292) run on the device too.  Static obstacles (CcdObstacle) are not on this path.
#pragma once
#include <cmath>
#include <vector>

#include "runtime.hpp"

namespace ddgrs {

struct CcdBox {
  double lo[3], hi[3];
};

// CcdOptions (ccd.rs:86); reference_length NaN is None (the mean edge length of the start).
struct CcdOpts {
  double gap = 0.05, reference_length = NAN, close_factor = 0.5;
  uint32_t max_iterations = 100, leaf_size = 8;
};
// CcdOptions::validate (ccd.rs:126)
void check_ccd_options(const CcdOpts& o);

// CcdReport (ccd.rs:214): kind 0 vertex a against triangle b, 1 edges a < b of the sorted edge list.
struct CcdResult {
  double tau_max = 0.0, gap_abs = 0.0, contact_tau = 0.0;
  bool has_contact = false;
  int kind = 0;
  uint32_t a = 0, b = 0;
  uint64_t iterations = 0, pairs_tested = 0, close_pairs = 0, touching_pairs = 0, exempt_pairs = 0;
  uint32_t windows = 0;
};

constexpr uint32_t kCcdMaxLevels = 60;
struct CcdLevels {
  uint32_t off[kCcdMaxLevels], cnt[kCcdMaxLevels];
  uint32_t nlev;
};

// One query's result in a window: its earliest contact (flags bit 0: found, bit 1: edge-edge) and
// its counts.
struct CcdQuery {
  double tau;
  uint32_t flags, a, b;
  uint32_t iterations, pairs, close, touching, exempt;
};
struct CcdSummary {
  double tau;
  uint32_t flags, a, b;
  unsigned long long iterations, pairs, close, touching, exempt;
};

// A static box tree on the device (ccd.rs:268 BoxTree).
struct CcdTreeDev {
  uint32_t n = 0, leaf = 8;
  CcdLevels lv{};
  DevBuf<unsigned long long> key_in, key_out;
  DevBuf<uint32_t> idx_in, order;
  DevBuf<CcdBox> sboxes, nodes;
  DevBuf<double> part, bounds;
  DevBuf<uint8_t> temp;
  void build(ddgrs_ctx* ctx, const CcdBox* boxes, uint32_t n, uint32_t leaf);
};

struct CcdDev {
  uint32_t nv = 0, nf = 0, ne = 0;
  std::vector<uint32_t> edges_h;  // 2 per edge, sorted and unique (Surface::edges)
  DevBuf<uint32_t> edges;
  DevBuf<double> speed;
  DevBuf<CcdBox> vbox, fbox, ebox, fq, eq;  // fq, eq: inflated by half the gap
  CcdTreeDev ftree, etree;
  DevBuf<CcdQuery> out;
  DevBuf<CcdSummary> sum;
  DevBuf<unsigned long long> scal;  // [0] max speed (bits), [1] intersecting pairs found
  DevBuf<uint32_t> pairs;
  DevBuf<uint8_t> core, band;
  size_t pair_cap = 0;

  // The sorted unique edges of the faces (once per mesh).
  void set_mesh(ddgrs_ctx* ctx, const std::vector<uint32_t>& faces_h, uint32_t nv);
  // first_contact_with_band (ccd.rs:781) of start + t·(vscale·vel) on [0, tau_end]: start, vel and
  // band (nv flags or nullptr) on the device; start_h the same start positions on the host (the
  // mean edge length is summed there, in edge order, as the reference sums it).  Synchronizes
  // once per window.
  CcdResult first_contact(ddgrs_ctx* ctx, const uint32_t* faces, const double* start, const double* start_h,
                          const double* vel, double vscale, double tau_end, const CcdOpts& o, const uint8_t* band);
  // intersecting_pairs (ccd.rs:1229) at the device positions: the pairs stay in `pairs`
  // (unsorted); returns their number.  Synchronizes.
  size_t find_intersecting_pairs(ddgrs_ctx* ctx, const uint32_t* faces, const double* pos);
  // IntersectionBand::new (ccd.rs:1292): the flags into `band` (device, nv); returns the number
  // of intersecting pairs (0: the band is empty).
  size_t band_flags(ddgrs_ctx* ctx, const uint32_t* faces, const double* pos);
};

// The mean edge length over the edge list (RS/src/geometry.rs:249), summed in edge order.
double mean_edge_length_host(const std::vector<uint32_t>& edges, const double* pos);

}  // namespace ddgrs

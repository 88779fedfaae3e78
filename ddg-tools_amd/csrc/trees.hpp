// Host-side tree structures of one descent step: the face / vertex BVH topology, the Barnes–Hut
// interaction lists and the block cluster tree structure.  The arithmetic on them (node data,
// pair kernels, far and near fields, products) runs on the device.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace ddgrs {

// A BVH over n weighted points with boxes (RC/src/spatial/bvh.rs:126 build_with, min split fraction
// 1/4): nodes in pre-order, left child of node i is i + 1, `order` the element permutation.
struct Topology {
  std::vector<uint32_t> order, start, end, right, parent, depth;
  std::vector<uint32_t> leaves, leaf_ordinal;  // leaf_ordinal[node] = k or UINT32_MAX
  uint32_t height = 0;
  size_t n_nodes() const { return start.size(); }
  bool is_leaf(size_t a) const { return right[a] == 0; }
};

Topology build_topology(size_t n, const double* pts, const double* lo, const double* hi, size_t leaf_size);

// A node of a BVH under construction (creation order: children are appended after their parent).
struct TreeSegment {
  uint32_t start, end, depth;
  int64_t left = -1, right = -1;
};
// Pre-order ids (left child = i + 1), parents, depths and leaves from the segments.
Topology finish_topology(const std::vector<TreeSegment>& segs, std::vector<uint32_t> order);

// Build-time cluster data (RS/src/spatial/mod.rs:80 Cluster), summed bottom-up in the reference's
// order: area, centroid, box of the element boxes, radius to the farthest box corner.
// A tree and its clusters already on the device (the face tree of a step): the block tree's build
// copies them on the device instead of uploading the host copies.
struct DeviceTree {
  const uint32_t *order, *start, *end, *right, *parent, *leaves;
  const double *mass, *centroid, *lo, *hi, *radius, *projector;
};

struct Clusters {
  std::vector<double> mass, centroid, lo, hi, radius;  // n, 3n, 3n, 3n, n
  std::vector<double> projector;  // 6n: P̄ = Σ a N Nᵀ / Σ a, upper triangle xx xy xz yy yz zz (face trees)
};
// `normals` (3 per element) adds the projectors (RS/src/spatial/mod.rs:275), nullptr skips them.
Clusters build_clusters(const Topology& t, const double* pts, const double* lo, const double* hi, const double* mass,
                        const double* normals = nullptr);

// Barnes–Hut lists (RS/src/energy/barnes_hut.rs:743 build_lists): per leaf ordinal k the lumped
// nodes and exact leaves in traversal order, and their transposes.
struct BhLists {
  std::vector<uint32_t> lumped_off, lumped, exact_off, exact;
  std::vector<uint32_t> lumped_t_off, lumped_t;  // per node: entries (indices into lumped), increasing
  std::vector<uint32_t> exact_t_off, exact_t;    // per leaf ordinal: source leaf ordinals, entry order
  uint64_t pairs_exact = 0;
};
BhLists build_bh_lists(const Topology& t, const Clusters& c, double theta);

// Block cluster tree structure (RS/src/spatial/block_tree.rs:129 build_impl) for a separation test
// on the clusters: far blocks and near blocks grouped by row node, near storage offsets (canonical
// a ≤ b blocks first, then mirrors), chains root → leaf.
struct BlockStructure {
  std::vector<uint32_t> far_ptr, far_col;          // per row node
  std::vector<uint32_t> near_ptr, near_col;        // per row node
  std::vector<uint64_t> near_off;                  // per near block
  std::vector<uint32_t> near_row;                  // per near block
  std::vector<uint32_t> chain_ptr, chain;          // per leaf (pre-order leaves)
  std::vector<uint8_t> far_needed;                 // per node: appears as a far column
  uint64_t near_total = 0, near_entries = 0, admissible_pairs = 0;
  uint32_t depth = 0;
};
BlockStructure build_block_structure(const Topology& t, const Clusters& c, double chi, size_t leaf_pairs);

}  // namespace ddgrs

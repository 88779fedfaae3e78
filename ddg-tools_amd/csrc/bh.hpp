// The device Barnes–Hut evaluator (see bh.hip).
#pragma once
#include <algorithm>

#include "kernels.hpp"
#include "tree_gpu.hpp"
#include "trees.hpp"

namespace ddgrs {

struct BhArgs;

// Steps (of 8 list entries) per work item of the Barnes–Hut pair kernels.  The BVH keeps a few long
// thin leaves along mesh seams whose exact lists hold tens of thousands of faces (about 100 times
// the mean); cut into items, they spread over many waves.
constexpr uint32_t kBhItemSteps = 256;

struct BhEval {
  double theta = 0.5;
  size_t leaf_size = 8;
  Topology topo;
  Clusters clusters;  // build-time cluster data (boxes, radii: frozen)
  BhLists lists;      // host lists (DDGRS_BH_LISTS_ON_HOST=1 only; the lists live on the device)
  int nf = 0, nn = 0, nl = 0;
  uint64_t n_lumped = 0, n_exact = 0, pairs_exact = 0;
  // structure on the device
  DevBuf<uint32_t> d_order, d_start, d_end, d_right, d_leaves, d_lumped_off, d_lumped, d_exact_off, d_exact,
      d_exact_t_off, d_exact_t, d_lumped_t_off, d_lumped_t, d_leaf_of_pos, d_chain_ptr, d_chain, d_levels;
  std::vector<std::vector<uint32_t>> levels;
  std::vector<uint32_t> level_off;
  // data of the current positions
  DevBuf<float4> pxhi, pxlo, pnrm, chi, clo, node_bu;
  DevBuf<double> mass, moment, cen;
  DevBuf<double4> sums;
  DevBuf<float> part;
  // device list build
  DevBuf<uint32_t> d_leaf_ordinal, d_parent, d_task_root, d_task_depth, d_task_ptr, d_cut_base, l_count, e_count,
      l_task_off, e_task_off;
  DevBuf<uint32_t> l_idx, l_key_s, e_tgt, e_src, e_tgt_s;
  // flattened exact lists: per leaf the BVH positions of the faces of its exact nodes (own side)
  // and of the leaves that open it (partner side), with per-entry sizes and offsets
  DevBuf<uint32_t> d_ex_off, d_ex_pos, d_pt_off, d_pt_pos, ex_size, ex_fo, pt_size, pt_fo;
  DevBuf<uint8_t> d_bhnode;  // per node a 64-byte record for the list walks
  // Work items of the pair kernels: a leaf's walk (lumped, exact and partner entries, 8 per step)
  // cut into pieces of at most kBhItemSteps steps; per item its leaf and first step, per leaf its
  // first item.  The per-face partials and energies are written per item and summed in item order.
  int n_items = 0;
  DevBuf<uint32_t> item_leaf, item_g0, item_ptr, item_cnt;
  DevBuf<double> item_energy;
  DevBuf<unsigned long long> d_pairs;
  DevBuf<uint8_t> temp;

  // Tree and lists at the surface's current positions (face geometry must be current).
  void build(ddgrs_surface* s, double theta, size_t leaf_size);
  // Lists on a face tree of the surface's current positions.
  void build_from(ddgrs_surface* s, const Topology& t, const Clusters& c, double theta, size_t leaf_size);
  // The same on the tree already in topo and clusters (no copy).
  void build_from_members(ddgrs_surface* s, double theta, size_t leaf_size);
  // Face packs in BVH order and node aggregates of the surface's current positions.
  void refresh(ddgrs_surface* s);
  // Ẽ into total[0] (device).
  void energy(ddgrs_ctx* ctx, const Exps& e, double* total);
  // Ẽ into total[0] and the per-face partials (7·nf FP64, face order) on the device.
  void gradient(ddgrs_surface* s, const Exps& e, double* total, double* partials);
  // The lists on the host (copied from the device).
  BhLists download_lists(ddgrs_ctx* ctx) const;
  // The work items of the current lists.
  void build_items(ddgrs_ctx* ctx);
  void trace_lists(ddgrs_ctx* ctx) const;

 private:
  void upload_structure(ddgrs_ctx* ctx);
  // Lists, transposes and counts on the device from topo and clusters (bh_lists_gpu.hip).
  void build_lists_device(ddgrs_ctx* ctx);
  // The exact lists flattened to face positions (bh_lists_gpu.hip).
  void flatten_lists(ddgrs_ctx* ctx);
  BhArgs args() const;
};

}  // namespace ddgrs

namespace ddgrs {
// The surface's current positions, face areas and vertex areas on the host (one synchronisation).
struct HostGeometry {
  std::vector<double> pos, area;  // 3 nv, nf
  std::vector<double> vmass;      // vertex areas a_v = ⅓ Σ_{S∋v} a_S
};
HostGeometry download_geometry(ddgrs_surface* s);
// Face BVH (topology built on the device) and its build-time clusters with projectors (computed on
// the device, read back).
void build_face_tree(ddgrs_surface* s, size_t leaf_size, GpuBvhBuilder& b, Topology* topo, Clusters* clusters);
// The tree and clusters of b's last build_face_tree, on the device (valid until b builds again).
DeviceTree device_tree(const GpuBvhBuilder& b, const Topology& t);
// Face BVH at the surface's current positions.
void build_face_tree(ddgrs_surface* s, size_t leaf_size, Topology* topo, Clusters* clusters);
// Vertex BVH (points f_v) at the surface's current positions, built on the device.
Topology build_vertex_tree(ddgrs_surface* s, size_t leaf_size, GpuBvhBuilder& b);
// Vertex areas a_v = ⅓ Σ_{S∋v} a_S (RS/src/geometry.rs:194) and positions of the surface (host).
void vertex_areas_host(ddgrs_surface* s, std::vector<double>* pos, std::vector<double>* mass);
}  // namespace ddgrs

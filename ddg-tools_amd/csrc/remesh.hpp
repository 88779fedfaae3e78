// One remeshing pass on the device (RS/src/remesh/mod.rs:235 `remesh`): the mesh is kept in HBM
// as face slots and per-vertex face lists; the data-parallel parts (edge tables, length and
// Delaunay classification, the smoothing's circumcenters, vertex normals, displacements and
// guarded moves, the compaction) run as wide kernels, and the edits whose order defines the
// result (splits in edge order, collapses shortest first, the flips' FIFO queue) run as one
// workgroup that visits only the edges the wide kernels flagged, in the reference's order.
#pragma once
#include <vector>

#include "runtime.hpp"

struct ddgrs_remeshed {
  std::vector<double> positions;     // 3·nv
  std::vector<uint32_t> faces;       // 3·nf
  std::vector<uint32_t> parents;     // nf
  std::vector<uint32_t> map_offsets, map_sources, pinned;
  std::vector<double> map_weights;
};

namespace ddgrs {

struct RemeshInput {
  int64_t nv = 0, nf = 0;
  const uint32_t* faces = nullptr;
  const double* positions = nullptr;
  std::vector<uint32_t> pinned;
  const uint8_t* band = nullptr;  // nv flags or null
};

// One pass; returns the result numbered as the reference's EditMesh::finish numbers it.
void remesh_pass(ddgrs_ctx* ctx, const RemeshInput& in, double l0, const ddgrs_remesh_options& o,
                 ddgrs_remeshed* out, ddgrs_remesh_report* rep);
// One operation (0 split, 1 collapse, 2 flip) and finish; returns 0 or 1 + the refusal kind.
int remesh_single_op(ddgrs_ctx* ctx, const RemeshInput& in, int op, uint32_t a, uint32_t b, bool delaunay,
                     bool foldover, bool* reordered, ddgrs_remeshed* out);

}  // namespace ddgrs

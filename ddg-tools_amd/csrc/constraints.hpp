// The step's nonlinear constraint rows (total area, enclosed volume) and their Schur complement,
// on the host in FP64: k ≤ 2 rows, so S is at most 2 × 2 and the per-row work is O(F); the k solves
// Z = A′⁻¹ C_Nᵀ run on the device with the step's operator and preconditioner.
//   crates/ddg-repulsive-surfaces/src/constraints/mod.rs   area_of :462, volume_of :474,
//       area_gradient :606, volume_gradient :627, tolerance :435, project_positional :657
//   crates/ddg-repulsive-surfaces/src/solve/schur.rs        from_solutions :142, factor :185,
//       solve_n :251, combine :281, direction_from :318, correction :341, project :495
#pragma once
#include <cstdint>
#include <vector>

namespace ddgrs {

enum { kRowNone = 0, kRowArea = 1, kRowVolume = 2 };

struct HostMesh {
  const uint32_t* faces;  // 3 per face
  int64_t nf, nv;
};

double row_value(const HostMesh& m, const double* pos, int kind);
// ∇ of the row at pos (3 nv, interleaved)
std::vector<double> row_gradient(const HostMesh& m, const double* pos, int kind);
// 1e-8 |target|, or 1e-8 for a zero target (NONLINEAR_TOLERANCE)
double row_tolerance(double target);
// the area-weighted barycenter (vertex areas a_S / 3 per face, faces in order)
void area_barycenter(const HostMesh& m, const double* pos, double b[3]);

// S = −C_N Z, N = −½ (S + Sᵀ) equilibrated to unit diagonal and factored by Cholesky with the
// largest remaining pivot first (ties by row order); rows whose pivot is at most 1e-10 are dropped.
struct SchurRows {
  std::vector<std::vector<double>> c, z;  // C_N rows and Z columns (position blocks, 3 nv each)
  std::vector<double> s, d, l;
  std::vector<size_t> kept;  // pivot order
  double min_pivot = 1.0;

  void build(std::vector<std::vector<double>> c_rows, std::vector<std::vector<double>> z_cols);
  std::vector<double> apply_c(const double* v) const;              // C_N v
  std::vector<double> solve_n(const std::vector<double>& v) const;  // N_kept⁻¹ v
  std::vector<double> combine(const std::vector<double>& c) const;  // Σ_r c_r Z_r
  std::vector<double> direction(const std::vector<double>& y) const;    // y − Z N⁻¹ C_N y
  std::vector<double> correction(const std::vector<double>& phi) const;  // Z N⁻¹ Φ_N
};

}  // namespace ddgrs

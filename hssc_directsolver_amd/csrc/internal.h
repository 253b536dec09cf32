// Internal declarations shared by the host builder, the GPU orchestration and the kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hssc.h"

namespace hssc {

// Records a failure message for hssc_last_error() and returns the status.
int fail(int status, const std::string& msg);

// Per-level data.  In a uniform tree every box of a level has the same n, k and M, so the
// GPU works on strided batches of equal-size blocks.  Matrices (J, T, F^-1, E) are stored
// once per "key": one key per box for NTI kernels, one per level for TI kernels.
struct Level {
  int lev = 0;
  int64_t n = 0, k = 0, M = 0;  // |I_src|, |I_sk|, proxy size
  int64_t nbox = 0, nkey = 0;   // boxes on the level, distinct (J,T) sets
  int64_t q = 0;                // columns of the stored T (n-k for dense T)

  // Host copies (index sets are built on the host).
  std::vector<int32_t> J;      // nkey x n   local permutation, skeleton first
  std::vector<int32_t> Jinv;   // nkey x n   inverse permutation
  std::vector<double> Px, Py;  // nkey x n   points of I_src in J order
  std::vector<double> Pw;      // nkey x n   per-point kernel weight b(x) = c(x)
  std::vector<double> PNx, PNy, PNw;  // nkey x n   the same points in natural I_src order
  std::vector<double> Xpx, Xpy, Xpw;  // nkey x M proxy points and weights

  // Device copies (owned by the tree).
  int32_t* dJ = nullptr;
  int32_t* dJinv = nullptr;
  double* dPx = nullptr;
  double* dPy = nullptr;
  double* dPw = nullptr;
  double* dPNx = nullptr;
  double* dPNy = nullptr;
  double* dPNw = nullptr;
  double* dXpx = nullptr;
  double* dXpy = nullptr;
  double* dXpw = nullptr;
  double* dT = nullptr;    // nkey x (k x (n-k)) column-major, FP64
  float* dT32 = nullptr;   // FP32 copy for the FP32 factorization
};

}  // namespace hssc

struct hssc_tree {
  hssc_params p{};
  int depth = 0;
  int64_t N = 0;
  std::vector<double> X;                     // N x 2 column-major
  std::vector<double> cent;                  // (nboxes+1) x 2, row b = centre of box b
  std::vector<std::vector<int64_t>> Isrc;    // box -> global I_src (0-based)
  std::vector<std::vector<int64_t>> Isk;     // box -> global I_sk
  std::vector<hssc::Level> levels;           // index = level
  std::vector<int32_t> leafG;                // nleaf x n_leaf: global index of J-ordered leaf point
  int32_t* dLeafG = nullptr;
  bool has_T = false;
  bool on_device = false;
  int device = 0;
  ~hssc_tree();
};

namespace hssc {

// Host builders (tree.cpp).
int build_index_sets(const hssc_params& p, hssc_tree* t);
void source_points(int Ng, int np, double* X /* N x 2 col-major */);

// Kernel weight of a point (Common/Kernel_Eval.m:111-159, params.proxy == 0).
double point_weight(const hssc_params& p, double x, double y);

}  // namespace hssc

// Internal declarations shared by the host builder, the GPU orchestration and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hssc.h"

namespace hssc {

// Records a failure message for hssc_last_error() and returns the status.
int fail(int status, const std::string& msg);

// Allocator whose value-initialisation is a no-op: resize() leaves new elements
// uninitialised, so large host arrays are first touched (and paged in) by the threads
// that fill them instead of being zeroed serially.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};
template <class T>
using hvec = std::vector<T, NoInitAlloc<T>>;

// Per-level data.  In a uniform tree every box of a level has the same n, k and M, so the
// GPU works on strided batches of equal-size blocks.  Matrices (J, T, F^-1, E) are stored
// once per "key": one key per box for NTI kernels, one per level for TI kernels.
struct Level {
  int lev = 0;
  int64_t n = 0, k = 0, M = 0;  // |I_src|, |I_sk|, proxy size
  int64_t nbox = 0, nkey = 0;   // boxes on the level, distinct (J,T) sets
  int64_t q = 0;                // columns of the stored T (n-k for dense T)
  bool id_skel = false;         // skeletons chosen by an ID (kh > 0, two bottom levels)
  std::vector<int32_t> kbox;    // nkey: each box's own rank (ID) or 'bdry' count before
                                // padding to the level's k

  // Host copies (index sets are built on the host).
  hvec<int32_t> J;      // nkey x n   local permutation, skeleton first
  hvec<int32_t> Jinv;   // nkey x n   inverse permutation
  hvec<double> Px, Py;  // nkey x n   points of I_src in J order
  hvec<double> Pw;      // nkey x n   per-point kernel weight b(x) = c(x)
  hvec<double> PNx, PNy, PNw;  // nkey x n   the same points in natural I_src order
  hvec<double> Xpx, Xpy, Xpw;  // nkey x M proxy points and weights

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
  // T: nkey x (k x (n-k)) column-major, FP64 (double, or interleaved complex for the
  // Helmholtz kernel), plus the FP32 copy (float / complex float) of the FP32 factorization
  void* dT = nullptr;
  void* dT32 = nullptr;
  // recorded once T of this level is complete (the build runs on its own streams and
  // returns before the GPU has finished; readers of T wait on this)
  hipEvent_t ready = nullptr;
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
  // interpolation operators still being computed: the build's streams and workspaces,
  // released by hssc::tree_sync once the device work is done
  std::vector<hipStream_t> bstreams;
  std::vector<void*> bpending;
  bool cx = false;  // complex (Helmholtz) kernel: every matrix and vector is complex
  int device = 0;
  // Lifetime: the caller's handle is one reference, and every inverse and operator made from
  // the tree (they read its device data) holds another; hssc_tree_destroy drops the caller's,
  // and whoever drops the last one frees the tree.
  std::atomic<int> refs{1};
  // counters of the skeleton IDs (ID-selected levels): blocks, blocks whose sketched ID failed
  // the acceptance test and were redone by the plain ID, sketch rows, rows of the blocks
  // factored by the plain ID
  int64_t id_boxes = 0, id_fallbacks = 0, id_sketch_rows = 0, id_full_rows = 0;
  ~hssc_tree();
};

namespace hssc {
// Takes / drops a reference on a tree (dropping the last one frees it).
void tree_ref(hssc_tree* t);
void tree_unref(hssc_tree* t);
}  // namespace hssc

namespace hssc {

// One level's batch of skeleton IDs: for every key box the matrix K(Q, S) of its M proxy
// points Q against its n source points S (I_src order), with the per-point weights.
struct IdJob {
  int lev;
  int64_t nbox, n, M;
  const double *Sx, *Sy, *Sw;  // nbox x n
  const double *Qx, *Qy, *Qw;  // nbox x M
  int64_t pin = 0;  // > 0: the first pin columns are a preset skeleton that the job extends
                    // with the residual's skeleton points (params.skel_extra_pts)
};
// Fills order (nbox x n, the ID's column order, skeleton first) and rank (nbox adaptive
// ranks); returns an HSSC_* status.
using IdSelect = std::function<int(const hssc_params&, const IdJob&, int32_t* order,
                                   int32_t* rank)>;

// HSS2D/predict_rank.m:5-23: the reference's closed-form rank guess of an n-column block.
double predict_rank_table(int n, double acc);

// Number of features per box fed to the rank network.
constexpr int kRankFeatures = 8;
// Features of the boxes of one ID job (nbox x kRankFeatures, column-major): level, log2 n,
// log2 M, kh times the box width, mean / min / max of the points' kernel weights, -log10 acc.
void id_job_features(const hssc_params& p, const IdJob& jb, float* X);

// Called with each level once its index sets, proxies and points are final (levels in build
// order, the leaves first); a nonzero HSSC_* status stops the build and is returned.
using LevelDone = std::function<int(int lev)>;
// Host builders (tree.cpp).  sel runs the IDs of the levels whose skeletons come from one.
int build_index_sets(const hssc_params& p, hssc_tree* t, const IdSelect& sel = nullptr,
                     const LevelDone& done = nullptr);
void source_points(int Ng, int np, double* X /* N x 2 col-major */);

// Waits for the tree's asynchronous build (interpolation operators) and releases its
// streams and workspaces.  Idempotent.
void tree_sync(hssc_tree* t);

// Kernel weight of a point (Common/Kernel_Eval.m:111-159, params.proxy == 0): b(x) of a
// row point or c(x) of a column point (the same for a symmetric kernel).
double point_weight(const hssc_params& p, double x, double y, bool row = true);

// Checks the options of a params struct that every entry point relies on; returns an
// HSSC_* status.
int check_params(const hssc_params& p);

}  // namespace hssc

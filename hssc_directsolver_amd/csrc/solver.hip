// GPU orchestration of the HSS2D compress / invert / apply path and the C ABI (hssc.h).
//
// Level-by-level batched launches: every box of a level (or the single per-level operator
// of a translation-invariant tree) is processed by one launch of each kernel.  Matrices of
// a level are stored as strided batches in J order (skeleton first), so the skeleton /
// residual blocks of F^-1 used by E = inv(R F^-1 R^T) and by the solve sweeps are plain
// sub-blocks.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <type_traits>
#include <new>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "gridop.h"
#include "internal.h"
#include "kernels.h"
#include "pool.h"
#include "stats.h"

namespace hssc {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

namespace {

struct HipError {
  int status;
};

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      g_last_error = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
      throw HipError{e_ == hipErrorOutOfMemory ? HSSC_ERR_OUT_OF_MEMORY : HSSC_ERR_HIP};  \
    }                                                                                     \
  } while (0)

template <typename T>
T* dalloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  HIPCHK(pool::alloc(&p, count * sizeof(T)));
  return static_cast<T*>(p);
}

template <class T, class A>
T* dupload(const std::vector<T, A>& h) {
  T* d = dalloc<T>(h.size());
  if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

void dfree(void* p) { pool::release(p); }

// Runs f and maps exceptions to status codes.
template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return e.status;
  } catch (const std::bad_alloc&) {
    return fail(HSSC_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

int check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(HSSC_ERR_HIP, "no HIP device is available");
  return HSSC_OK;
}

// y = alpha * op(A_key) x_box + gamma * z_box for every box of a level.
//   A: per-key matrix (rows x cols, leading dimension lda, key stride sA; sA = 0 for TI)
//   x, y, z: per-box blocks with leading dimensions ldx, ldy, ldz and box stride ld*m.
// TI levels share A, so all boxes form one GEMM with nbox*m columns.
template <typename T>
void level_matvec(hipStream_t st, bool trans, int64_t rows, int64_t cols, T alpha, const T* A,
                  int64_t lda, int64_t sA, const T* x, int64_t ldx, T gamma, const T* z,
                  int64_t ldz, T* y, int64_t ldy, int64_t nbox, int64_t m, bool shared,
                  T* work) {
  const int64_t out = trans ? cols : rows, inner = trans ? rows : cols;
  if (out <= 0) return;
  if (!shared && m == 1) {
    k::gemv<T>(st, trans, rows, cols, alpha, A, lda, sA, x, ldx, gamma, z, ldz, y, ldy, nbox,
               work);
    return;
  }
  if (gamma != T(0) || inner <= 0) {
    if (z != y || gamma != T(1))
      k::add2d<T>(st, out, nbox * m, gamma, z, ldz, 0, T(0), z, ldz, 0, y, ldy, 0, 1);
  }
  if (inner <= 0) return;
  const T beta = gamma != T(0) ? T(1) : T(0);
  if (shared) {
    k::gemm<T>(st, trans, false, out, nbox * m, inner, alpha, A, lda, 0, x, ldx, 0, beta, y, ldy,
               0, 1);
  } else {
    k::gemm<T>(st, trans, false, out, m, inner, alpha, A, lda, sA, x, ldx, ldx * m, beta, y, ldy,
               ldy * m, nbox);
  }
}

// y = A x + gamma z for the n x n blocks A of a level (stride n^2): from the upper triangle
// when sym (exactly symmetric blocks, one right-hand side, one block per box), else as
// level_matvec.
template <typename T>
void sym_matvec(hipStream_t st, bool sym, int64_t n, const T* A, const T* x, int64_t ldx,
                T gamma, const T* z, int64_t ldz, T* y, int64_t ldy, int64_t nbox, int64_t m,
                bool shared, T* work) {
  if (sym)
    k::symv_upper<T>(st, n, T(1), A, n, n * n, x, ldx, gamma, z, ldz, y, ldy, nbox, work);
  else
    level_matvec<T>(st, false, n, n, T(1), A, n, n * n, x, ldx, gamma, z, ldz, y, ldy, nbox, m,
                    shared, work);
}

int64_t gemv_work_elems(int64_t out, int64_t inner, int64_t batch) {
  return batch * ((inner + k::kGemvChunk - 1) / k::kGemvChunk + 1) * out;
}

}  // namespace
}  // namespace hssc

using namespace hssc;

// ---------------------------------------------------------------------------------------
// Object definitions
// ---------------------------------------------------------------------------------------

hssc_tree::~hssc_tree() {
  tree_sync(this);
  for (auto& L : levels) {
    dfree(L.dJ), dfree(L.dJinv), dfree(L.dPx), dfree(L.dPy), dfree(L.dPw);
    dfree(L.dPNx), dfree(L.dPNy), dfree(L.dPNw);
    dfree(L.dXpx), dfree(L.dXpy), dfree(L.dXpw), dfree(L.dT), dfree(L.dT32);
    if (L.ready) (void)hipEventDestroy(L.ready);
  }
  dfree(dLeafG);
}

namespace hssc {
int check_params(const hssc_params& p) {
  if (p.flag_pot != HSSC_POT_SL_L_2D && p.flag_pot != HSSC_POT_SL_H_2D)
    return fail(HSSC_ERR_INVALID_ARG, "unknown kernel");
  if (p.flag_pot == HSSC_POT_SL_H_2D && !(p.kh > 0))
    return fail(HSSC_ERR_INVALID_ARG, "SL_H_2D needs kh > 0");
  if (!(p.kernel_scale > 0) || !std::isfinite(p.kernel_scale) || !(p.length_scale > 0) ||
      !std::isfinite(p.length_scale))
    return fail(HSSC_ERR_INVALID_ARG, "kernel_scale and length_scale must be positive");
  if (p.flag_pot == HSSC_POT_SL_H_2D && (p.kernel_scale != 1.0 || p.length_scale != 1.0))
    return fail(HSSC_ERR_UNSUPPORTED, "kernel_scale / length_scale apply to SL_L_2D only");
  if (p.contrast != HSSC_CONTRAST_PLATEAU && p.contrast != HSSC_CONTRAST_GAUSSIAN)
    return fail(HSSC_ERR_INVALID_ARG, "unknown contrast");
  if (p.contrast == HSSC_CONTRAST_GAUSSIAN && (!(p.contrast_amp >= 0) || !(p.contrast_width > 0)))
    return fail(HSSC_ERR_INVALID_ARG, "Gaussian contrast needs amp >= 0 and width > 0");
  if (p.tree_precision != HSSC_PREC_FP64 && p.tree_precision != HSSC_PREC_FP32)
    return fail(HSSC_ERR_INVALID_ARG, "unknown tree precision");
  if (p.proxy_rule != HSSC_PROXY_BDRY && p.proxy_rule != HSSC_PROXY_RAND)
    return fail(HSSC_ERR_INVALID_ARG, "unknown proxy rule");
  if (p.proxy_rule == HSSC_PROXY_RAND && !(p.lambda > 0))
    return fail(HSSC_ERR_INVALID_ARG, "the 'rand' proxy rule needs lambda > 0");
  return HSSC_OK;
}

void tree_ref(hssc_tree* t) {
  if (t) t->refs.fetch_add(1);
}
void tree_unref(hssc_tree* t) {
  if (t && t->refs.fetch_sub(1) == 1) delete t;
}

void tree_sync(hssc_tree* t) {
  if (!t || t->bstreams.empty()) return;
  for (hipStream_t s : t->bstreams) (void)hipStreamSynchronize(s);
  stats::flush();
  for (hipStream_t s : t->bstreams) (void)hipStreamDestroy(s);
  for (void* p : t->bpending) dfree(p);
  t->bstreams.clear();
  t->bpending.clear();
}
}  // namespace hssc

namespace hssc {
// FP64 type of the same field as T: double for real T, complex double for complex T.
template <typename T>
using hi_t = typename std::conditional<is_complex_v<T>, zdouble, double>::type;
// FP32 type of the same field as T.
template <typename T>
using lo_t = typename std::conditional<is_complex_v<T>, zfloat, float>::type;

// Per-RHS-count solve workspace (levels x {Q, PU, PD}) plus scratch.  It either owns its
// buffers (ensure: pool allocations kept for the next call with the same m) or lays them out
// in memory the caller provides (borrow: nothing is allocated or freed).
template <typename T>
struct SolveWork {
  int64_t m = 0;
  bool borrowed = false;
  std::vector<T*> Q, PU, PD;
  T* S1 = nullptr;
  T* S2 = nullptr;
  T* work = nullptr;
  void release() {
    if (!borrowed) {
      for (T* p : Q) dfree(p);
      for (T* p : PU) dfree(p);
      for (T* p : PD) dfree(p);
      dfree(S1), dfree(S2), dfree(work);
    }
    Q.clear(), PU.clear(), PD.clear();
    S1 = S2 = work = nullptr;
    m = 0;
    borrowed = false;
  }
  // Visits every buffer of the workspace for mm right-hand sides with its element count.
  template <typename F>
  static void layout(const hssc_tree* t, int64_t mm, F&& f) {
    const int D = t->depth;
    int64_t smax = 0, wmax = 0;
    for (int l = 0; l <= D; ++l) {
      const Level& L = t->levels[l];
      f(0, l, L.nbox * L.n * mm);
      f(1, l, L.nbox * L.k * mm);
      f(2, l, L.nbox * L.k * mm);
      smax = std::max(smax, L.nbox * L.n * mm);
      wmax = std::max(wmax, gemv_work_elems(L.n, L.n, L.nbox));
      wmax = std::max(wmax, k::symv_work_elems(L.n, L.nbox));
    }
    f(3, 0, smax);
    f(4, 0, smax);
    f(5, 0, wmax);
  }
  void assign(int which, int l, T* p) {
    if (which == 0) Q[l] = p;
    else if (which == 1) PU[l] = p;
    else if (which == 2) PD[l] = p;
    else if (which == 3) S1 = p;
    else if (which == 4) S2 = p;
    else work = p;
  }
  void ensure(const hssc_tree* t, int64_t mm) {
    if (m == mm) return;
    release();
    const int D = t->depth;
    Q.assign(D + 1, nullptr), PU.assign(D + 1, nullptr), PD.assign(D + 1, nullptr);
    layout(t, mm, [&](int which, int l, int64_t cnt) { assign(which, l, dalloc<T>(cnt)); });
    m = mm;
  }
  // Bytes borrow() needs for mm right-hand sides (every buffer 256-byte aligned).
  static int64_t borrow_bytes(const hssc_tree* t, int64_t mm) {
    int64_t total = 0;
    layout(t, mm, [&](int, int, int64_t cnt) {
      total += (std::max<int64_t>(cnt, 1) * int64_t(sizeof(T)) + 255) / 256 * 256;
    });
    return total;
  }
  // Lays the workspace out in [base, base + borrow_bytes) (base 256-byte aligned).
  void borrow(const hssc_tree* t, int64_t mm, void* base) {
    release();
    const int D = t->depth;
    Q.assign(D + 1, nullptr), PU.assign(D + 1, nullptr), PD.assign(D + 1, nullptr);
    char* cur = static_cast<char*>(base);
    layout(t, mm, [&](int which, int l, int64_t cnt) {
      assign(which, l, reinterpret_cast<T*>(cur));
      cur += (std::max<int64_t>(cnt, 1) * int64_t(sizeof(T)) + 255) / 256 * 256;
    });
    borrowed = true;
    m = mm;
  }
};

// One workspace per scalar type.
struct SolveWorks {
  SolveWork<double> d;
  SolveWork<float> f;
  SolveWork<zdouble> z;
  SolveWork<zfloat> c;
  template <typename T>
  SolveWork<T>& get();
  void release() { d.release(), f.release(), z.release(), c.release(); }
};
template <>
SolveWork<double>& SolveWorks::get<double>() {
  return d;
}
template <>
SolveWork<float>& SolveWorks::get<float>() {
  return f;
}
template <>
SolveWork<zdouble>& SolveWorks::get<zdouble>() {
  return z;
}
template <>
SolveWork<zfloat>& SolveWorks::get<zfloat>() {
  return c;
}

// The GPU form of Common/Kernel_Eval.m for the tree's kernel.
k::KernelSpec kernel_spec(const hssc_params& p) {
  k::KernelSpec ks;
  ks.pot = p.flag_pot;
  const double h2 = p.h * p.h;
  if (p.flag_pot == HSSC_POT_SL_H_2D) {
    const double k2 = p.kh * p.kh;
    ks.scale = h2 * k2 * 0.25;
    ks.kh = p.kh;
    // w = 1 + k^2 (i/4) corr0
    ks.wre = 1.0 - 0.25 * k2 * p.dr_weights_im;
    ks.wim = 0.25 * k2 * p.dr_weights_re;
  } else {
    const double L = p.length_scale;
    ks.scale = 0.5 * h2 * p.kernel_scale * L * L;
    ks.lshift = 2.0 * std::log(L);
  }
  return ks;
}

// Device buffers released together when the owner goes out of scope.
struct DevBufs {
  std::vector<void*> v;
  template <typename T>
  T* get(size_t count) {
    T* d = dalloc<T>(count);
    v.push_back(d);
    return d;
  }
  template <typename T>
  T* upload(const T* h, size_t count) {
    T* d = get<T>(count);
    if (count) HIPCHK(hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
  ~DevBufs() {
    for (void* p : v) dfree(p);
  }
};

// A non-blocking stream of one call, drained and released when the owner goes out of scope.
// Work that must not wait for the tree's level streams runs on one: commands on the null stream
// are ordered after the work of every other stream of the device.
struct OwnStream {
  hipStream_t s = nullptr;
  OwnStream() { HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~OwnStream() {
    (void)hipStreamSynchronize(s);
    stats::flush();
    (void)hipStreamDestroy(s);
  }
  OwnStream(const OwnStream&) = delete;
  OwnStream& operator=(const OwnStream&) = delete;
};

void id_job_features(const hssc_params& p, const IdJob& jb, float* X) {
  const int64_t nb = jb.nbox, n = jb.n;
  for (int64_t b = 0; b < nb; ++b) {
    const double* sx = jb.Sx + b * n;
    const double* sy = jb.Sy + b * n;
    const double* sw = jb.Sw + b * n;
    double x0 = sx[0], x1 = sx[0], y0 = sy[0], y1 = sy[0], wm = 0, wlo = sw[0], whi = sw[0];
    for (int64_t i = 0; i < n; ++i) {
      x0 = std::min(x0, sx[i]), x1 = std::max(x1, sx[i]);
      y0 = std::min(y0, sy[i]), y1 = std::max(y1, sy[i]);
      wm += sw[i], wlo = std::min(wlo, sw[i]), whi = std::max(whi, sw[i]);
    }
    const double width = std::max(x1 - x0, y1 - y0) + p.h;
    const float f[kRankFeatures] = {float(jb.lev), float(std::log2(double(n))),
                                    float(std::log2(double(jb.M))), float(p.kh * width),
                                    float(wm / double(n)), float(wlo), float(whi),
                                    float(-std::log10(p.acc))};
    for (int q = 0; q < kRankFeatures; ++q) X[b + q * nb] = f[q];
  }
}

// How the skeleton IDs of a tree are computed: mode 0 = the plain ID of every block; 1 / 2 =
// a randomized ID whose sketch size comes from the closed-form rank guess / from the rank
// network, judged on the full block, with the plain ID as the fallback.
struct IdPolicy {
  int mode = 0;
  hssc_mlp* mlp = nullptr;
  int oversample = 8;
  hssc_tree* tree = nullptr;  // receives the counters
};

// Skeleton selection of one ID level (LOCAL_Non_Preset_Skeleton_Info,
// HSS2D/HSS2D_build_tree_fsym.m:686-699): A_b = K(X_proxy, X_src) of every key box is
// generated on the GPU and factored by the batched column-pivoted QR ID with the adaptive
// rank of Common/ID.m:40-65 (|R_ii| > acc |R_11|, then the randomized check).  With a rank
// policy the ID first runs on Y_b = G_b A_b (Common/LRID_rand.m:10-40): G_b is the first s_b
// rows of a Gaussian shared by the level's boxes, s_b = kSketchMargin x the predicted rank of
// box b (at most n) plus oversampling, rounded up to 8 rows (the box's rank bucket); one
// batched MFMA product for the level at the largest s_b, then each box's rows are packed to its
// own leading dimension.  A box whose sketched ID fails the check of ID.m:55-59 on its full
// block (or whose sketch saturated), or whose bucket has no fewer rows than the block, gets the
// plain ID.  Returns the pivot order of every box (skeleton first) and its rank.
constexpr double kSketchMargin = 1.05;  // covers the rank network's typical error (a few percent)

static int id_select_chunk(const hssc_params& p, const IdJob& jb, int32_t* order, int32_t* rank,
                           const IdPolicy& pol);

// The level's boxes go through in chunks whose device buffers (the blocks, their copies, the
// sketches and the ID work) stay under kIdChunkBytes.
constexpr int64_t kIdChunkBytes = int64_t(8) << 30;

// Extra skeleton points of a preset level (params.skel_extra_pts; LOCAL_ID_rand_update,
// HSS2D/HSS2D_build_tree_fsym.m:1069-1124).  A_b = K(X_proxy, X_src) of every key box, columns
// in the preset J order, is factored by the column-pivoted QR with its first jb.pin (skeleton)
// columns taken in order: after them the trailing block is the interpolation residual
// K(Xp, X_rs) - K(Xp, X_sk) T (T the least-squares interpolation operator) in an orthonormal
// basis, so the later pivots and |R_jj| are the residual's own column-pivoted QR.  The
// reference counts the ranks of a Gaussian sketch of the residual with q0 = 8 layers
// ceil(log2 k0) samples against the absolute tolerance acc (QR_findrank.m:45-92); a sketch
// scales the residual's singular values by about sqrt(q0), so the residual's own pivots count
// while |R_jj| > acc / sqrt(q0).  order = the pivot order, rank = pin + the added points.
template <typename S>
static int extra_points_chunk(const hssc_params& p, const IdJob& jb, int32_t* order,
                              int32_t* rank) {
  return guarded([&]() {
    const int64_t nb = jb.nbox, n = jb.n, M = jb.M, pin = jb.pin;
    const int64_t kq = std::min(M, n);
    const int64_t tsz = std::max<int64_t>(1, kq * (n - kq)), wsz = 6 * n + M;
    DevBufs B;
    const double* Sx = B.upload(jb.Sx, nb * n);
    const double* Sy = B.upload(jb.Sy, nb * n);
    const double* Sw = B.upload(jb.Sw, nb * n);
    const double* Qx = B.upload(jb.Qx, nb * M);
    const double* Qy = B.upload(jb.Qy, nb * M);
    const double* Qw = B.upload(jb.Qw, nb * M);
    std::vector<int64_t> off(nb), poff(nb), toff(nb), woff(nb);
    std::vector<int32_t> mr(nb, int32_t(M)), nc(nb, int32_t(n)), kfix(nb, int32_t(kq)),
        kpin(nb, int32_t(pin));
    for (int64_t b = 0; b < nb; ++b)
      off[b] = b * M * n, poff[b] = b * n, toff[b] = b * tsz, woff[b] = b * wsz;
    S* A = B.get<S>(nb * M * n);
    S* Ac = B.get<S>(nb * M * n);
    S* Tt = B.get<S>(nb * tsz);
    S* W = B.get<S>(nb * wsz);
    int32_t* dperm = B.get<int32_t>(nb * n);
    int32_t* dk = B.get<int32_t>(nb);
    const hipStream_t st = 0;
    k::gen_block<S>(st, M, n, nb, Qx, Qy, Qw, M, Sx, Sy, Sw, n, A, M, M * n, kernel_spec(p));
    HIPCHK(hipMemcpyAsync(Ac, A, size_t(nb * M * n) * sizeof(S), hipMemcpyDeviceToDevice, st));
    k::cpqr_id<S>(st, int(nb), B.upload(off.data(), nb), B.upload(mr.data(), nb),
                  B.upload(nc.data(), nb), A, Ac, -1.0, B.upload(kfix.data(), nb), dperm,
                  B.upload(poff.data(), nb), dk, Tt, B.upload(toff.data(), nb), W,
                  B.upload(woff.data(), nb), 0x5EEDull ^ uint64_t(jb.lev), 0, 0,
                  B.upload(kpin.data(), nb));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    std::vector<S> w(size_t(nb * wsz));
    HIPCHK(hipMemcpy(w.data(), W, w.size() * sizeof(S), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(order, dperm, size_t(nb * n) * sizeof(int32_t), hipMemcpyDeviceToHost));
    const double q0 = 8.0 * p.layers * std::ceil(std::log2(double(std::max<int64_t>(pin, 2))));
    const double thr = p.acc / std::sqrt(q0);
    for (int64_t b = 0; b < nb; ++b) {
      int64_t k1 = 0;
      for (int64_t j = pin; j < kq; ++j) k1 += re_s(w[b * wsz + n + j]) > thr;
      rank[b] = int32_t(pin + k1);
    }
    return HSSC_OK;
  });
}

static int extra_points_gpu(const hssc_params& p, const IdJob& jb, int32_t* order,
                            int32_t* rank) {
  if (jb.n < 1 || jb.M < 1 || jb.n > INT32_MAX || jb.M > INT32_MAX || jb.pin > jb.n)
    return fail(HSSC_ERR_INVALID_ARG, "empty or oversized extra-points block");
  const bool cx = p.flag_pot == HSSC_POT_SL_H_2D;
  const int64_t es = cx ? int64_t(sizeof(zdouble)) : int64_t(sizeof(double));
  const int64_t per_box = es * (2 * jb.M * jb.n + jb.n * jb.n / 4 + 7 * jb.n + jb.M);
  const int64_t chunk = std::max<int64_t>(1, kIdChunkBytes / per_box);
  for (int64_t b0 = 0; b0 < jb.nbox; b0 += chunk) {
    IdJob sub = jb;
    sub.nbox = std::min(chunk, jb.nbox - b0);
    sub.Sx += b0 * jb.n, sub.Sy += b0 * jb.n, sub.Sw += b0 * jb.n;
    sub.Qx += b0 * jb.M, sub.Qy += b0 * jb.M, sub.Qw += b0 * jb.M;
    const int s = cx ? extra_points_chunk<zdouble>(p, sub, order + b0 * jb.n, rank + b0)
                     : extra_points_chunk<double>(p, sub, order + b0 * jb.n, rank + b0);
    if (s) return s;
  }
  return HSSC_OK;
}

int id_select_gpu(const hssc_params& p, const IdJob& jb, int32_t* order, int32_t* rank,
                  const IdPolicy& pol) {
  if (int s = check_device()) return s;
  if (jb.pin > 0) return extra_points_gpu(p, jb, order, rank);
  if (p.flag_pot != HSSC_POT_SL_H_2D)
    return fail(HSSC_ERR_UNSUPPORTED, "ID-selected skeletons are defined for kh > 0 only");
  if (jb.n < 1 || jb.M < 1 || jb.n > INT32_MAX || jb.M > INT32_MAX)
    return fail(HSSC_ERR_INVALID_ARG, "empty or oversized ID block");
  const int64_t n = jb.n, M = jb.M;
  const int64_t per_box = int64_t(sizeof(zdouble)) * (5 * M * n + n * n / 4 + 7 * n + M);
  const int64_t chunk = std::max<int64_t>(1, kIdChunkBytes / per_box);
  for (int64_t b0 = 0; b0 < jb.nbox; b0 += chunk) {
    IdJob sub = jb;
    sub.nbox = std::min(chunk, jb.nbox - b0);
    sub.Sx += b0 * n, sub.Sy += b0 * n, sub.Sw += b0 * n;
    sub.Qx += b0 * M, sub.Qy += b0 * M, sub.Qw += b0 * M;
    if (int s = id_select_chunk(p, sub, order + b0 * n, rank + b0, pol)) return s;
  }
  return HSSC_OK;
}

static int id_select_chunk(const hssc_params& p, const IdJob& jb, int32_t* order, int32_t* rank,
                           const IdPolicy& pol) {
  return guarded([&]() {
    const int64_t nb = jb.nbox, n = jb.n, M = jb.M;
    const uint64_t seed = 0x5EEDull ^ uint64_t(jb.lev);
    // sketch rows of every box (0: plain ID)
    std::vector<int32_t> sb(nb, 0);
    int64_t s = 0;  // the largest
    if (pol.mode != 0) {
      std::vector<double> pred(nb, 1.0);
      if (pol.mode == 1) {
        std::fill(pred.begin(), pred.end(), predict_rank_table(int(n), p.acc));
      } else {
        std::vector<float> X(nb * kRankFeatures), y(nb);
        id_job_features(p, jb, X.data());
        if (int st = hssc_mlp_predict(pol.mlp, X.data(), int32_t(nb), y.data()))
          throw HipError{st};
        for (int64_t b = 0; b < nb; ++b) pred[b] = std::exp2(double(y[b]));
      }
      for (int64_t b = 0; b < nb; ++b) {
        // the rank is at most n: a sketch of n + oversample rows keeps the whole column space
        const double r = std::min(std::max(1.0, kSketchMargin * pred[b]), double(n)) +
                         pol.oversample;
        const int64_t sr = (int64_t(std::ceil(r)) + 7) / 8 * 8;
        sb[b] = sr < M ? int32_t(sr) : 0;  // no fewer rows than the block: the plain ID
        s = std::max<int64_t>(s, sb[b]);
      }
    }
    const int64_t tsz = n * n / 4 + n, wsz = 6 * n + M;  // T is k x (n-k) <= n^2/4
    DevBufs B;
    const double* Sx = B.upload(jb.Sx, nb * n);
    const double* Sy = B.upload(jb.Sy, nb * n);
    const double* Sw = B.upload(jb.Sw, nb * n);
    const double* Qx = B.upload(jb.Qx, nb * M);
    const double* Qy = B.upload(jb.Qy, nb * M);
    const double* Qw = B.upload(jb.Qw, nb * M);
    std::vector<int64_t> off(nb), poff(nb), toff(nb), woff(nb);
    std::vector<int32_t> mr(nb, int32_t(M)), nc(nb, int32_t(n)), kfix(nb, 0);
    for (int64_t b = 0; b < nb; ++b)
      off[b] = b * M * n, poff[b] = b * n, toff[b] = b * tsz, woff[b] = b * wsz;
    const int64_t* dpoff = B.upload(poff.data(), nb);
    const int64_t* dtoff = B.upload(toff.data(), nb);
    const int64_t* dwoff = B.upload(woff.data(), nb);
    const int32_t* dmr = B.upload(mr.data(), nb);
    const int32_t* dnc = B.upload(nc.data(), nb);
    const int32_t* dkfix = B.upload(kfix.data(), nb);
    zdouble* A = B.get<zdouble>(nb * M * n);
    zdouble* Tt = B.get<zdouble>(nb * tsz);
    zdouble* W = B.get<zdouble>(nb * wsz);
    int32_t* dperm = B.get<int32_t>(nb * n);
    const hipStream_t st = 0;
    k::gen_block<zdouble>(st, M, n, nb, Qx, Qy, Qw, M, Sx, Sy, Sw, n, A, M, M * n, kernel_spec(p));
    std::vector<int> ok(nb, 0);
    std::vector<int64_t> sj;  // the boxes that get a sketch
    for (int64_t b = 0; b < nb; ++b)
      if (sb[b] > 0) sj.push_back(b);
    const int64_t ns = int64_t(sj.size());
    int64_t sketch_rows = 0;
    if (ns > 0) {
      // Ytmp_b = G A_b (G: s x M Gaussian shared by the level), then box b keeps its first
      // s_b rows in Y at leading dimension s_b
      std::vector<zdouble> G(s * M);
      std::mt19937_64 rng(seed ^ 0x9E3779B97F4A7C15ull);
      std::normal_distribution<double> nd;
      for (auto& g : G) g = zdouble(nd(rng));
      const zdouble* dG = B.upload(G.data(), s * M);
      zdouble* Ytmp = B.get<zdouble>(nb * s * n);
      k::gemm<zdouble>(st, false, false, s, n, M, zdouble(1), dG, s, 0, A, M, M * n, zdouble(0),
                       Ytmp, s, s * n, nb);
      std::vector<int64_t> soff(ns), yoff(ns), aoff(ns), jpoff(ns), jtoff(ns), jwoff(ns);
      std::vector<int32_t> sr(ns);
      for (int64_t j = 0; j < ns; ++j) {
        const int64_t b = sj[j];
        soff[j] = b * s * n, yoff[j] = sketch_rows * n, sr[j] = sb[b];
        aoff[j] = off[b], jpoff[j] = poff[b], jtoff[j] = toff[b], jwoff[j] = woff[b];
        sketch_rows += sb[b];
      }
      zdouble* Y = B.get<zdouble>(sketch_rows * n);
      zdouble* Yc = B.get<zdouble>(sketch_rows * n);
      const int32_t* dsr = B.upload(sr.data(), ns);
      const int64_t* dyoff = B.upload(yoff.data(), ns);
      k::pack_rows<zdouble>(st, int(ns), B.upload(soff.data(), ns), s, dyoff, dsr, n, Ytmp, Y);
      HIPCHK(hipMemcpyAsync(Yc, Y, size_t(sketch_rows * n) * sizeof(zdouble),
                            hipMemcpyDeviceToDevice, st));
      const int64_t* djp = B.upload(jpoff.data(), ns);
      const int64_t* djt = B.upload(jtoff.data(), ns);
      int32_t* dks = B.get<int32_t>(ns);
      int32_t* dok = B.get<int32_t>(ns);
      k::cpqr_id<zdouble>(st, int(ns), dyoff, dsr, dnc, Y, Yc, p.acc, dkfix, dperm, djp, dks, Tt,
                          djt, W, B.upload(jwoff.data(), ns), seed, s, n);
      k::id_verify<zdouble>(st, int(ns), B.upload(aoff.data(), ns), dmr, dnc, dsr, A, dperm, djp,
                            dks, Tt, djt, p.acc, seed ^ 0xABCDull, dok);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
      std::vector<int32_t> oks(ns), ks(ns);
      HIPCHK(hipMemcpy(oks.data(), dok, size_t(ns) * sizeof(int32_t), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(ks.data(), dks, size_t(ns) * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int64_t j = 0; j < ns; ++j) ok[sj[j]] = oks[j], rank[sj[j]] = ks[j];
    }
    // the plain ID (on the full blocks) for every box without an accepted sketch
    std::vector<int64_t> fall;
    for (int64_t b = 0; b < nb; ++b)
      if (!ok[b]) fall.push_back(b);
    if (!fall.empty()) {
      const int64_t nf = int64_t(fall.size());
      std::vector<int64_t> foff(nf), fpoff(nf), ftoff(nf), fwoff(nf);
      for (int64_t j = 0; j < nf; ++j)
        foff[j] = off[fall[j]], fpoff[j] = poff[fall[j]], ftoff[j] = toff[fall[j]],
        fwoff[j] = woff[fall[j]];
      zdouble* Ac = B.get<zdouble>(nb * M * n);
      HIPCHK(hipMemcpyAsync(Ac, A, size_t(nb * M * n) * sizeof(zdouble), hipMemcpyDeviceToDevice, st));
      int32_t* dkf = B.get<int32_t>(nf);
      k::cpqr_id<zdouble>(st, int(nf), B.upload(foff.data(), nf), dmr, dnc, A, Ac, p.acc, dkfix,
                          dperm, B.upload(fpoff.data(), nf), dkf, Tt, B.upload(ftoff.data(), nf), W,
                          B.upload(fwoff.data(), nf), seed, M, n);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
      std::vector<int32_t> kf(nf);
      HIPCHK(hipMemcpy(kf.data(), dkf, size_t(nf) * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int64_t j = 0; j < nf; ++j) rank[fall[j]] = kf[j];
    }
    HIPCHK(hipMemcpy(order, dperm, size_t(nb * n) * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t b = 0; b < nb; ++b)
      if (rank[b] < 0 || rank[b] > n) return fail(HSSC_ERR_HIP, "skeleton ID returned a bad rank");
    if (pol.tree) {
      int64_t failed = 0;
      for (int64_t b : sj) failed += ok[b] ? 0 : 1;
      pol.tree->id_boxes += nb;
      pol.tree->id_fallbacks += failed;
      pol.tree->id_sketch_rows += sketch_rows;
      pol.tree->id_full_rows += int64_t(fall.size()) * M;
    }
    return HSSC_OK;
  });
}
// Refinement steps replayed from HIP graphs, for repeated solves with one inverse and one
// residual operator: "first" = {Y = Ainv B; R = B - A Y; partial sums of |R|^2 and |B|^2} and
// "step" = {D = Ainv R; Y += D; R = B - A Y; partial sums of |R|^2}, on buffers (and a solve
// workspace) of their own; the caller's B is copied in and Y out.
struct RefineGraphs {
  uint64_t op = 0;  // the residual operator's uid
  int64_t m = 0;
  int calls = 0;    // solves with this operator and m
  hipStream_t cs = nullptr;
  hipGraphExec_t first = nullptr, step = nullptr;
  void *B = nullptr, *Y = nullptr, *R = nullptr, *D = nullptr;
  double* part = nullptr;  // 512 partial sums: |R|^2, then |B|^2
  SolveWorks w;
  void release() {
    if (cs) (void)hipStreamSynchronize(cs);
    if (first) (void)hipGraphExecDestroy(first);
    if (step) (void)hipGraphExecDestroy(step);
    if (cs) (void)hipStreamDestroy(cs);
    dfree(B), dfree(Y), dfree(R), dfree(D), dfree(part);
    w.release();
    cs = nullptr, first = step = nullptr;
    B = Y = R = D = nullptr, part = nullptr;
    op = 0, m = 0, calls = 0;
  }
  ~RefineGraphs() { release(); }
};
}  // namespace hssc

struct hssc_inverse {
  hssc_tree* tree = nullptr;
  int prec = HSSC_PREC_FP64;
  std::vector<void*> FI, E;  // per level, nkey x n x n and nkey x k x k (J order)
  mutable hssc::SolveWorks w;
  mutable hssc::RefineGraphs graphs;
  explicit hssc_inverse(hssc_tree* t) : tree(t) { hssc::tree_ref(t); }
  ~hssc_inverse() {
    graphs.release();
    for (void* p : FI) dfree(p);
    for (void* p : E) dfree(p);
    w.release();
    hssc::tree_unref(tree);
  }
};

struct hssc_operator {
  hssc_tree* tree = nullptr;
  // level depth: Bself (nkey x n x n, J order); others: B12 (kc x kc); FP64 real or complex
  std::vector<void*> B;
  mutable hssc::SolveWorks w;
  explicit hssc_operator(hssc_tree* t) : tree(t) { hssc::tree_ref(t); }
  ~hssc_operator() {
    for (void* p : B) dfree(p);
    w.release();
    hssc::tree_unref(tree);
  }
};

namespace hssc {
namespace {

// ---------------------------------------------------------------------------------------
// Tree upload and interpolation operators
// ---------------------------------------------------------------------------------------

// Uploads one level's index sets: J, J^-1, the points in J and in natural order, the proxies.
void upload_level(Level& L) {
  L.dJ = dupload(L.J), L.dJinv = dupload(L.Jinv);
  L.dPx = dupload(L.Px), L.dPy = dupload(L.Py), L.dPw = dupload(L.Pw);
  L.dPNx = dupload(L.PNx), L.dPNy = dupload(L.PNy), L.dPNw = dupload(L.PNw);
  L.dXpx = dupload(L.Xpx), L.dXpy = dupload(L.Xpy), L.dXpw = dupload(L.Xpw);
}

// T = K(Xp, X_sk) \ K(Xp, X_rs) for every key of every level
// (HSS2D/interpolation_operator.m:57-66 and the same operator to tolerance for :80-127),
// by blocked Householder QR of the augmented M x n block [K11 K21] on its first k
// columns, then back substitution with R11.  The root's T is never used by invert or the
// sweeps (the root has no E), so it is not formed.  S = double or complex double.
//
// The build follows the host's index sets level by level (the leaves first): add(l) uploads
// level l, allocates its workspace and hands its whole least-squares sequence (hundreds of
// launches on the upper levels) to a launcher thread of its own, which queues it on a stream
// of its own.  The GPU factors the lower levels while the host still builds the upper ones,
// and the launches of the levels do not queue behind each other on the host: the top level,
// the longest chain, starts as soon as its index sets exist.  finish() joins the launchers,
// uploads the leaf map and hands the streams and workspaces to the tree, whose readers drain
// them (tree_sync); a build destroyed unfinished joins, drains and frees its own.
class InterpBuild {
 public:
  virtual ~InterpBuild() = default;
  virtual void add(int l) = 0;
  virtual void finish() = 0;
};

template <typename S>
class InterpBuildT final : public InterpBuild {
 public:
  explicit InterpBuildT(hssc_tree* t) : t_(t), ws_(t->depth + 1) {
    HIPCHK(hipGetDevice(&t->device));
    // Four streams: levels 1 and 2, whose chains of dependent panels are the longest, get one
    // each, and the levels below alternate over the other two.  Every stream in flight is a
    // hardware queue whose kernels compete for the CUs with the top levels' small,
    // latency-bound ones; the lower levels, which finish well before the top one, absorb the
    // serialisation (512² Helmholtz: complex64 tree 0.223 -> 0.215 s, complex128 0.290 ->
    // 0.280 s, against a stream per level; one shared stream for the lower levels is slower).
    streams_.assign(4, nullptr);
  }
  ~InterpBuildT() override {
    join();
    if (handed_) return;
    for (auto& s : streams_)
      if (s) (void)hipStreamSynchronize(s);
    stats::flush();
    for (auto& s : streams_)
      if (s) (void)hipStreamDestroy(s);
    for (auto& w : ws_)
      for (void* p : w.all()) dfree(p);
  }
  void add(int l) override;
  void finish() override {
    join();
    if (err_) throw HipError{fail(err_, err_msg_)};
    t_->dLeafG = dupload(t_->leafG);
    t_->on_device = true;
    for (auto& s : streams_)
      if (s) t_->bstreams.push_back(s);
    for (auto& w : ws_)
      for (void* p : w.all())
        if (p) t_->bpending.push_back(p);
    handed_ = true;
    t_->has_T = true;
  }

 private:
  struct Ws {
    S *A = nullptr, *V = nullptr, *W1 = nullptr, *W2 = nullptr, *Tf = nullptr, *part = nullptr,
      *Ts = nullptr, *GV = nullptr;
    wide_t<S>* cq = nullptr;  // CholeskyQR workspace (FP64 arithmetic for every storage type)
    std::vector<void*> all() const { return {A, V, W1, W2, Tf, part, Ts, GV, cq}; }
  };
  void join() {
    for (auto& th : launchers_)
      if (th.joinable()) th.join();
    launchers_.clear();
  }
  hssc_tree* t_;
  std::vector<Ws> ws_;
  std::vector<hipStream_t> streams_;
  std::vector<std::thread> launchers_;
  std::mutex err_mu_;
  int err_ = HSSC_OK;  // first launcher failure (status and message), guarded by err_mu_
  std::string err_msg_;
  bool handed_ = false;
};

template <typename S>
void InterpBuildT<S>::add(int l) {
  hssc_tree* t = t_;
  Level& L = t->levels[l];
  upload_level(L);
  if (l == 0) return;  // the root's T is not formed
  const k::KernelSpec ks = kernel_spec(t->p);
  const int64_t nb = k::kQrNb;
  const int64_t kchunk = 512;
  // K chunk of the thin products (at most 128 x 128 outputs): more, shorter chunks
  const int64_t kchunk_thin = 128;
  // outer panel width (a multiple of the 32-column sub-panel; HSSC_LS_OUTER_NB overrides it)
  static const int64_t kOuterNb = [] {
    const char* v = std::getenv("HSSC_LS_OUTER_NB");
    const int64_t w = v ? std::atoll(v) : 128;
    return std::min<int64_t>(512, std::max<int64_t>(32, w / 32 * 32));
  }();
  // regularisation weight of FP32 trees relative to the RMS column norm of K(Xp, X_sk)
  // (HSSC_TIKHONOV_ETA overrides it)
  static const double kTikhonovEta = [] {
    const char* v = std::getenv("HSSC_TIKHONOV_ETA");
    return v ? std::atof(v) : 1e-5;
  }();
  constexpr int64_t kMergeCols = 128;  // trailing blocks up to this width take rank-32 updates
  // FP32 trees solve the Tikhonov-regularised least squares
  //   min ||K(Xp, X_sk) T - K(Xp, X_rs)||^2 + lambda^2 ||T||^2
  // as the plain one of the block with k more rows [lambda I 0] below it: directions in
  // which K(Xp, X_sk) is smaller than FP32 rounding can resolve (its singular values reach
  // acc ||K||) are damped instead of amplifying rounding noise into T.
  const bool tikhonov = sizeof(real_t<S>) == 4;
  const int64_t Mq = L.M, n = L.n, kk = L.k, r = n - kk, E = L.nkey;
  const int64_t M = tikhonov ? Mq + kk : Mq;  // rows of the least squares
  S* dT = dalloc<S>(E * kk * std::max<int64_t>(r, 1));
  // an FP32 tree keeps only the FP32 operator (the copy the FP32 factorization reads)
  if constexpr (sizeof(real_t<S>) == 4) L.dT32 = dT;
  else L.dT = dT;
  L.q = r;
  if (r == 0) return;
  if (Mq < kk) throw HipError{fail(HSSC_ERR_UNSUPPORTED, "proxy smaller than skeleton (M < k)")};
  Ws& w = ws_[l];
  const int64_t sA = M * n;
  // Outer panels of up to 128 columns made of 32-column CholeskyQR3 sub-panels, so the
  // trailing matrix (the HBM-bound part) is streamed once per outer panel.
  const int64_t nbo = std::min<int64_t>(kOuterNb, (kk + nb - 1) / nb * nb);
  w.A = dalloc<S>(E * sA);
  w.V = dalloc<S>(E * M * nbo);
  w.W1 = dalloc<S>(E * nbo * n);
  w.W2 = dalloc<S>(E * nbo * n);
  w.Tf = dalloc<S>(E * nbo * nbo);
  const int64_t part_cap = E * k::ls_part_elems(M, n, nbo, kchunk, kchunk_thin);
  w.part = dalloc<S>(part_cap);
  w.cq = dalloc<wide_t<S>>(E * k::cqr_ws_elems(M));
  w.Ts = dalloc<S>(E * kOuterNb * nb);
  w.GV = dalloc<S>(E * nbo * nbo);
  S* A = w.A;
  std::vector<std::function<void(hipStream_t)>> sl;
  // K(Xp, P) with P in J order: columns [0,k) = skeleton, [k,n) = residual
  sl.push_back([=, &L](hipStream_t st) {
    k::gen_block<S>(st, Mq, n, E, L.dXpx, L.dXpy, L.dXpw, Mq, L.dPx, L.dPy, L.dPw, n, A, M, sA,
                    ks);
    if (tikhonov) {
      k::zero2d<S>(st, kk, n, A + Mq, M, sA, E);
      k::tikhonov_rows<S>(st, Mq, kk, A, M, sA, E, kTikhonovEta);
    }
    k::zero2d<S>(st, M, nbo, w.V, M, M * nbo, E);  // V's rows the panels never write
  });
  for (int64_t j0 = 0; j0 < kk; j0 += nbo) {
    sl.push_back([=](hipStream_t st) {
      const int jbo = int(std::min(nbo, kk - j0));
      const int64_t nc = n - j0 - jbo;
      // V keeps leading dimension M for every outer panel: the rows above each sub-panel's
      // diagonal block and the regularisation rows below its reach are then never written
      // and stay zero from the level's first step
      const int64_t Mp = M;
      const int64_t sV = M * nbo;
      // Rows that can be nonzero below row j of the columns >= j once the columns before
      // jend are factored: every data row, and of the Tikhonov rows only those of the
      // factored columns (the reflector of column j reaches the Tikhonov rows of the
      // columns <= j only, so Tikhonov row i stays lambda e_i until column i is reached).
      // The panels and their products skip the untouched rows.
      auto rows_to = [=](int64_t jend) { return tikhonov ? Mq + std::min(jend, kk) : M; };
      const int64_t Mpe = rows_to(j0 + jbo) - j0;
      // a narrow trailing block (the last panels of the small levels) takes each
      // sub-panel's reflectors directly, without the outer panel's block reflector
      const bool merge = nc <= kMergeCols;
      for (int s0 = 0; s0 < jbo; s0 += int(nb)) {
        const int64_t j1 = j0 + s0;
        const int jb = int(std::min<int64_t>(nb, jbo - s0));
        const int64_t Mps = rows_to(j1 + jb) - j1;
        S* Vs = w.V + s0 + s0 * Mp;
        S* Tss = w.Ts + (s0 / nb) * nb * nb;
        k::cqr_panel<S>(st, j1 + Mps, j1, jb, A, M, sA, Vs, Mp, sV, Tss, kOuterNb * nb, w.cq,
                        E);
        // this sub-panel's reflectors on the rest of the outer panel (and the trailing
        // block when merged)
        const int64_t nin = jbo - s0 - jb + (merge ? std::max<int64_t>(nc, 0) : 0);
        if (nin <= 0) continue;
        S* Cin = A + j1 + (j1 + jb) * M;
        // W2 = Ts^H V^H C_in
        k::gemm_tn_splitk_th<S>(st, jb, nin, Mps, Vs, Mp, sV, Cin, M, sA, Tss, nb, kOuterNb * nb,
                                w.W1, w.W2, jb * nin, E, w.part, kchunk_thin, part_cap);
        k::gemm<S>(st, false, false, Mps, nin, jb, S(-1), Vs, Mp, sV, w.W2, jb, jb * nin,
                   S(1), Cin, M, sA, E);
      }
      if (nc <= 0 || merge) return;
      // T of the outer panel from the sub-panel factors and V^H V
      k::gemm_tn_splitk<S>(st, jbo, jbo, Mpe, w.V, Mp, sV, w.V, Mp, sV, w.GV, jbo, jbo * jbo,
                           E, w.part, kchunk_thin, part_cap);
      k::t_assemble<S>(st, jbo, w.Ts, kOuterNb * nb, w.GV, jbo, jbo * jbo, w.Tf, jbo,
                       nbo * nbo, w.W2, nbo * n, E);
      const int64_t ldt = jbo;
      S* C = A + j0 + (j0 + jbo) * M;
      // C <- Q_panel^H C:  W1 = V^H C ; W2 = Tf^H W1 ; C -= V W2
      k::gemm_tn_splitk<S>(st, jbo, nc, Mpe, w.V, Mp, sV, C, M, sA, w.W1, jbo, jbo * nc, E,
                           w.part, kchunk, part_cap);
      k::gemm_h<S>(st, jbo, nc, jbo, S(1), w.Tf, ldt, nbo * nbo, w.W1, jbo, jbo * nc, S(0),
                   w.W2, jbo, jbo * nc, E);
      k::gemm<S>(st, false, false, Mpe, nc, jbo, S(-1), w.V, Mp, sV, w.W2, jbo, jbo * nc,
                 S(1), C, M, sA, E);
    });
  }
  // Back substitution R11 X = C1 (C1 = A[0:k, k:n]) from the bottom block up.  The inverses
  // of R11's 32 x 32 diagonal blocks are formed first, all in one launch, into W1 (free by
  // then), so the chain of dependent launches is two products per block; a level too small
  // for that layout inverts each block in its own step.
  S* C1 = A + kk * M;
  const int64_t last = ((kk - 1) / nb) * nb;
  const int64_t sInv = ((kk + nb - 1) / nb) * nb * nb;  // per matrix
  const bool inv_all = sInv <= nbo * n;
  if (inv_all)
    sl.push_back([=](hipStream_t st) { k::trinv_upper_all<S>(st, kk, A, M, sA, w.W1, nbo * n, E); });
  for (int64_t p = last; p >= 0; p -= nb) {
    sl.push_back([=](hipStream_t st) {
      const int bsz = int(std::min(nb, kk - p));
      const S* Rinv = inv_all ? w.W1 + (p / nb) * nb * nb : w.Tf;
      const int64_t sRinv = inv_all ? nbo * n : nb * nb;
      if (!inv_all) k::trinv_upper<S>(st, p, bsz, A, M, sA, w.Tf, nb * nb, E);
      // X_b = inv(R_bb) C1_b  (in place; one workgroup reads its columns before storing)
      k::gemm<S>(st, false, false, bsz, r, bsz, S(1), Rinv, bsz, sRinv, C1 + p, M, sA, S(0),
                 C1 + p, M, sA, E);
      if (p > 0)
        k::gemm<S>(st, false, false, p, r, bsz, S(-1), A + p * M, M, sA, C1 + p, M, sA, S(1),
                   C1, M, sA, E);
    });
  }
  sl.push_back([=](hipStream_t st) { k::copy2d<S, S>(st, kk, r, C1, M, sA, dT, kk, kk * r, E); });
  // the level's uploads (pageable copies on the null stream) have landed before its stream
  // reads them
  HIPCHK(hipStreamSynchronize(nullptr));
  hipStream_t& sref = streams_[l <= 2 ? 2 + (l - 1) : l % 2];
  if (!sref) HIPCHK(hipStreamCreateWithFlags(&sref, hipStreamNonBlocking));
  if (!L.ready) HIPCHK(hipEventCreateWithFlags(&L.ready, hipEventDisableTiming));
  const hipStream_t st = sref;
  const hipEvent_t ev = L.ready;
  const int dev = t->device;
  launchers_.emplace_back([this, st, ev, dev, steps = std::move(sl)]() {
    try {
      HIPCHK(hipSetDevice(dev));  // the current device is per host thread
      for (auto& f : steps) f(st);
      HIPCHK(hipEventRecord(ev, st));
      HIPCHK(hipGetLastError());
    } catch (const HipError& e) {
      std::lock_guard<std::mutex> g(err_mu_);
      if (!err_) err_ = e.status, err_msg_ = g_last_error;
    } catch (...) {
      std::lock_guard<std::mutex> g(err_mu_);
      if (!err_) err_ = HSSC_ERR_HIP, err_msg_ = "least-squares launcher failed";
    }
  });
}

// The least squares in the tree's precision: FP64, or FP32 / complex64 on MFMA (the skeleton
// IDs stay FP64 either way).
std::unique_ptr<InterpBuild> interp_build(hssc_tree* t) {
  const bool lo = t->p.tree_precision == HSSC_PREC_FP32;
  if (t->cx) {
    if (lo) return std::unique_ptr<InterpBuild>(new InterpBuildT<zfloat>(t));
    return std::unique_ptr<InterpBuild>(new InterpBuildT<zdouble>(t));
  }
  if (lo) return std::unique_ptr<InterpBuild>(new InterpBuildT<float>(t));
  return std::unique_ptr<InterpBuild>(new InterpBuildT<double>(t));
}

// FP64 interpolation operators of an FP32 tree (widened once, for the FP64 readers: the FP64
// factorization, the compressed operator's sweeps, hssc_tree_get_T).  Waits for the build.
void ensure_T64(hssc_tree* t) {
  tree_sync(t);
  for (int l = 1; l <= t->depth; ++l) {
    Level& L = t->levels[l];
    if (L.dT || !L.dT32) continue;
    const int64_t cnt = L.nkey * L.k * std::max<int64_t>(L.n - L.k, 1);
    if (t->cx) {
      zdouble* d = dalloc<zdouble>(cnt);
      k::axpby<zfloat, zdouble>(0, cnt, zdouble(1), static_cast<const zfloat*>(L.dT32), zdouble(0), d);
      L.dT = d;
    } else {
      double* d = dalloc<double>(cnt);
      k::axpby<float, double>(0, cnt, 1.0, static_cast<const float*>(L.dT32), 0.0, d);
      L.dT = d;
    }
  }
  HIPCHK(hipStreamSynchronize(0));
}

// ---------------------------------------------------------------------------------------
// Inversion (HSS2D/HSS2D_invert_fsym.m:73-195, dense F^-1 and E)
// ---------------------------------------------------------------------------------------

// T of a level in the precision of T (FP64 or the FP32 copy).
template <typename T>
const T* level_T(const Level& L) {
  return static_cast<const T*>(sizeof(real_t<T>) == 8 ? L.dT : L.dT32);
}

template <typename T>
void invert_impl(hssc_tree* t, hssc_inverse* inv, hipStream_t st) {
  const int D = t->depth;
  const k::KernelSpec ks = kernel_spec(t->p);
  const bool TI = t->p.transinv != 0;
  const bool tree_lo = t->p.tree_precision == HSSC_PREC_FP32;
  // The build's level streams are drained before the inversion starts.  An inversion running
  // beside them gave complex128 factors that differed between identical runs (every T
  // identical to a drained build's; E and F^-1 wrong from a random level up, in about one run
  // in three after another path's build had run in the process), also when the host waited
  // for each level's T before queueing its readers; inverting the same tree again once the
  // build had drained gave the exact factors every time.
  tree_sync(t);
  if (tree_lo && sizeof(real_t<T>) == 8) ensure_T64(t);  // FP64 factors of an FP32 tree
  inv->FI.assign(D + 1, nullptr);
  inv->E.assign(D + 1, nullptr);
  int64_t wmax = 0, gmax = 0;
  for (int l = 0; l <= D; ++l) {
    const Level& L = t->levels[l];
    if (l > 0) wmax = std::max(wmax, L.nkey * L.k * (L.n - L.k));
    gmax = std::max({gmax, k::invert_ws_elems<T>(L.n, L.nkey), k::invert_ws_elems<T>(L.k, L.nkey)});
  }
  T* Wk = dalloc<T>(wmax);
  T* Wg = gmax > 0 ? dalloc<T>(gmax) : nullptr;  // blocked Gauss-Jordan workspace
  int* dflag = dalloc<int>(1);  // set by the sweeps on a singular diagonal block
  HIPCHK(hipMemsetAsync(dflag, 0, sizeof(int), st));
  // second stream for the look-ahead of the blocked Gauss-Jordan sweeps
  hipStream_t aux = nullptr;
  HIPCHK(hipGetLastError());
  if (hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) != hipSuccess) {
    aux = nullptr;  // the sweeps run without look-ahead
    (void)hipGetLastError();
  }
  for (int l = D; l >= 0; --l) {
    const Level& L = t->levels[l];
    const int64_t n = L.n, kk = L.k, r = n - kk, E = L.nkey;
    T* FI = dalloc<T>(E * n * n);
    inv->FI[l] = FI;
    if (l == D) {
      k::gen_block<T>(st, n, n, E, L.dPx, L.dPy, L.dPw, n, L.dPx, L.dPy, L.dPw, n, FI, n, n * n,
                      ks);
    } else {
      const Level& C = t->levels[l + 1];
      const int64_t kc = C.k;
      const int64_t sEbox = TI ? 0 : kc * kc, sEpar = TI ? 0 : 2 * kc * kc;
      k::gen_parent<T>(st, n, kc, E, L.dJ, TI ? 0 : n, L.dPx, L.dPy, L.dPw, TI ? 0 : n,
                       static_cast<const T*>(inv->E[l + 1]), sEbox, sEpar, FI, n * n, ks);
    }
    k::invert_batched<T>(st, n, FI, n, n * n, E, aux, Wg, true, dflag);  // F is symmetric
    if (l > 0) {
      // EI = Phi_ss + T Phi_rs + (Phi_sr + T Phi_rr) T^T ;  E = inv(EI)
      T* EI = dalloc<T>(E * kk * kk);
      inv->E[l] = EI;
      k::copy2d<T, T>(st, kk, kk, FI, n, n * n, EI, kk, kk * kk, E);
      if (r > 0) {
        if (L.ready) HIPCHK(hipStreamWaitEvent(st, L.ready, 0));
        if (sizeof(real_t<T>) == 4 && !tree_lo) {  // an FP32 tree already holds the FP32 T
          Level& Lm = t->levels[l];
          if (!Lm.dT32) Lm.dT32 = dalloc<T>(Lm.nkey * Lm.k * r);
          k::axpby<hi_t<T>, T>(st, Lm.nkey * Lm.k * r, T(1), static_cast<const hi_t<T>*>(Lm.dT),
                               T(0), static_cast<T*>(Lm.dT32));
        }
        const T* Tm = level_T<T>(L);
        k::copy2d<T, T>(st, kk, r, FI + kk * n, n, n * n, Wk, kk, kk * r, E);
        k::gemm<T>(st, false, false, kk, r, r, T(1), Tm, kk, kk * r, FI + kk + kk * n, n, n * n,
                   T(1), Wk, kk, kk * r, E);
        // EI is symmetric: the two kk x kk products form its upper 64 x 64 tiles only and the
        // lower ones are their transposes (Phi_ss, copied whole, is symmetric itself)
        const int64_t tk = (kk + 63) / 64, big = INT64_MAX;
        k::gemm_skip<T>(st, false, false, kk, kk, r, T(1), Tm, kk, kk * r, FI + kk, n, n * n, T(1),
                        EI, kk, kk * kk, E, tk, tk, big, 0, big, 0, 1);
        k::gemm_skip<T>(st, false, true, kk, kk, r, T(1), Wk, kk, kk * r, Tm, kk, kk * r, T(1), EI,
                        kk, kk * kk, E, tk, tk, big, 0, big, 0, 1);
        k::sym_lower_fill<T>(st, kk, EI, kk, kk * kk, E);
      }
      k::invert_batched<T>(st, kk, EI, kk, kk * kk, E, aux, Wg, true, dflag);  // EI = R F^-1 R^T
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  if (aux) {
    (void)hipStreamSynchronize(aux);
    stats::flush();
    (void)hipStreamDestroy(aux);
  }
  HIPCHK(hipGetLastError());
  int singular = 0;
  HIPCHK(hipMemcpy(&singular, dflag, sizeof(int), hipMemcpyDeviceToHost));
  dfree(Wk);
  dfree(Wg);
  dfree(dflag);
  tree_sync(t);  // every level's T has been waited for: release the build's resources
  if (singular)
    throw HipError{fail(HSSC_ERR_SINGULAR,
                        "singular or non-finite pivot in the inversion of F or R F^-1 R^T")};
}

// ---------------------------------------------------------------------------------------
// Tree sweeps
// ---------------------------------------------------------------------------------------

// Forms the J-ordered input of every box of level l (< depth) from the children's
// upward vectors PU[l+1] (child blocks kc x m, two per parent) into Q (n x m per box).
template <typename T>
void form_parent_q(const hssc_tree* t, int l, const T* PUc, T* Q, T* S, int64_t m,
                   hipStream_t st) {
  const Level& L = t->levels[l];
  const int64_t n = L.n, kc = t->levels[l + 1].k;
  const bool TI = t->p.transinv != 0;
  // natural order: rows [0,kc) = child 1, [kc,n) = child 2
  k::copy2d<T, T>(st, kc, m, PUc, kc, 2 * kc * m, S, n, n * m, L.nbox);
  k::copy2d<T, T>(st, kc, m, PUc + kc * m, kc, 2 * kc * m, S + kc, n, n * m, L.nbox);
  k::gather<T, T>(st, n, m, L.dJ, TI ? 0 : n, S, n, n * m, Q, n, n * m, L.nbox);
}

// Splits J-ordered box vectors PHI (n x m) into the children's downward vectors PD[l+1].
template <typename T>
void split_to_children(const hssc_tree* t, int l, const T* PHI, T* PDc, T* S, int64_t m,
                       hipStream_t st, bool accumulate_into_S = false) {
  const Level& L = t->levels[l];
  const int64_t n = L.n, kc = t->levels[l + 1].k;
  const bool TI = t->p.transinv != 0;
  k::scatter<T, T>(st, n, m, L.dJ, TI ? 0 : n, PHI, n, n * m, S, n, n * m, L.nbox,
                   accumulate_into_S);
  k::copy2d<T, T>(st, kc, m, S, n, n * m, PDc, kc, 2 * kc * m, L.nbox);
  k::copy2d<T, T>(st, kc, m, S + kc, n, n * m, PDc + kc * m, kc, 2 * kc * m, L.nbox);
}

// HSS2D/HSS2D_applyinv_TI.m:63-225 and HSS2D_applyinv_fsym.m (Algorithm 6 of the paper).
// X, Y: N x m in the FP64 type of the field (double or complex double).
template <typename T>
void applyinv_impl(const hssc_inverse* inv, SolveWork<T>& w, const hi_t<T>* X, hi_t<T>* Y,
                   int64_t m, hipStream_t st) {
  using H = hi_t<T>;
  const hssc_tree* t = inv->tree;
  const int D = t->depth;
  const bool TI = t->p.transinv != 0;
  const int64_t N = t->N;
  w.ensure(t, m);
  // F^-1 and E of a complex symmetric operator are exactly symmetric (the symmetric sweep
  // of invert_batched), so their products with one vector read the upper triangle only
  const bool symv = is_complex_v<T> && m == 1 && !TI;
  // upward pass
  for (int l = D; l >= 1; --l) {
    const Level& L = t->levels[l];
    const int64_t n = L.n, kk = L.k, r = n - kk, nb = L.nbox;
    const T* FI = static_cast<const T*>(inv->FI[l]);
    const T* E = static_cast<const T*>(inv->E[l]);
    const T* Tm = level_T<T>(L);
    T* Q = w.Q[l];
    if (l == D) k::gather<H, T>(st, n, m, t->dLeafG, n, X, N, 0, Q, n, n * m, nb);
    else form_parent_q<T>(t, l, w.PU[l + 1], Q, w.S1, m, st);
    // psi = FI q
    T* PSI = w.S2;
    sym_matvec<T>(st, symv, n, FI, Q, n, T(0), nullptr, n, PSI, n, nb, m, TI, w.work);
    // u = psi_s + T psi_r
    T* U = w.S1;
    if (r > 0)
      level_matvec<T>(st, false, kk, r, T(1), Tm, kk, kk * r, PSI + kk, n, T(1), PSI, n, U, kk,
                      nb, m, TI, w.work);
    else
      k::copy2d<T, T>(st, kk, nb * m, PSI, n, 0, U, kk, 0, 1);
    // phi_up = E u
    sym_matvec<T>(st, symv, kk, E, U, kk, T(0), nullptr, kk, w.PU[l], kk, nb, m, TI, w.work);
  }
  // root: phi = FI q
  {
    const Level& L = t->levels[0];
    const int64_t n = L.n;
    T* Q = w.Q[0];
    if (D == 0) k::gather<H, T>(st, n, m, t->dLeafG, n, X, N, 0, Q, n, n * m, 1);
    else form_parent_q<T>(t, 0, w.PU[1], Q, w.S1, m, st);
    T* PHI = w.S2;
    sym_matvec<T>(st, symv, n, static_cast<const T*>(inv->FI[0]), Q, n, T(0), nullptr, n, PHI,
                  n, 1, m, TI, w.work);
    if (D == 0) k::scatter<T, H>(st, n, m, t->dLeafG, n, PHI, n, n * m, Y, N, 0, 1, false);
    else split_to_children<T>(t, 0, PHI, w.PD[1], w.S1, m, st);
  }
  // downward pass
  for (int l = 1; l <= D; ++l) {
    const Level& L = t->levels[l];
    const int64_t n = L.n, kk = L.k, r = n - kk, nb = L.nbox;
    const T* FI = static_cast<const T*>(inv->FI[l]);
    const T* E = static_cast<const T*>(inv->E[l]);
    const T* Tm = level_T<T>(L);
    T* Q = w.Q[l];
    // wv = E phi_dn - phi_up
    T* WV = w.S1;
    sym_matvec<T>(st, symv, kk, E, w.PD[l], kk, T(-1), w.PU[l], kk, WV, kk, nb, m, TI, w.work);
    // v = q + [wv ; T^T wv]
    T* V = w.S2;
    k::add2d<T>(st, kk, nb * m, T(1), Q, n, 0, T(1), WV, kk, 0, V, n, 0, 1);
    if (r > 0)
      level_matvec<T>(st, true, kk, r, T(1), Tm, kk, kk * r, WV, kk, T(1), Q + kk, n, V + kk, n,
                      nb, m, TI, w.work);
    // phi = FI v  (into Q, which is no longer needed)
    sym_matvec<T>(st, symv, n, FI, V, n, T(0), nullptr, n, Q, n, nb, m, TI, w.work);
    if (l == D) k::scatter<T, H>(st, n, m, t->dLeafG, n, Q, n, n * m, Y, N, 0, nb, false);
    else split_to_children<T>(t, l, Q, w.PD[l + 1], w.S1, m, st);
  }
}

// HSS2D/HSS2D_apply_TI.m:74-189 and HSS2D_apply_fsym.m: telescoping forward operator in
// FP64 (S = double or complex double).
template <typename S>
void apply_impl(const hssc_operator* op, const S* X, S* Y, int64_t m, hipStream_t st) {
  const hssc_tree* t = op->tree;
  SolveWork<S>& w = op->w.get<S>();
  const int D = t->depth;
  const bool TI = t->p.transinv != 0;
  const int64_t N = t->N;
  w.ensure(t, m);
  auto Bl = [&](int l) { return static_cast<const S*>(op->B[l]); };
  // upward: phi_up = q_s + T q_r
  for (int l = D; l >= 1; --l) {
    const Level& L = t->levels[l];
    const int64_t n = L.n, kk = L.k, r = n - kk, nb = L.nbox;
    const S* Tm = static_cast<const S*>(L.dT);
    S* Qd = w.Q[l];
    if (l == D) k::gather<S, S>(st, n, m, t->dLeafG, n, X, N, 0, Qd, n, n * m, nb);
    else form_parent_q<S>(t, l, w.PU[l + 1], Qd, w.S1, m, st);
    if (r > 0)
      level_matvec<S>(st, false, kk, r, S(1), Tm, kk, kk * r, Qd + kk, n, S(1), Qd, n, w.PU[l],
                      kk, nb, m, TI, w.work);
    else
      k::copy2d<S, S>(st, kk, nb * m, Qd, n, 0, w.PU[l], kk, 0, 1);
  }
  if (D == 0) {
    const Level& L = t->levels[0];
    k::gather<S, S>(st, L.n, m, t->dLeafG, L.n, X, N, 0, w.Q[0], L.n, L.n * m, 1);
    level_matvec<S>(st, false, L.n, L.n, S(1), Bl(0), L.n, L.n * L.n, w.Q[0], L.n, S(0), nullptr,
                    L.n, w.S1, L.n, 1, m, TI, w.work);
    k::scatter<S, S>(st, L.n, m, t->dLeafG, L.n, w.S1, L.n, L.n * m, Y, N, 0, 1, false);
    return;
  }
  // downward
  for (int l = 0; l < D; ++l) {
    const Level& L = t->levels[l];
    const Level& C = t->levels[l + 1];
    const int64_t n = L.n, kk = L.k, r = n - kk, nb = L.nbox, kc = C.k;
    const S* B12 = Bl(l);
    const S* Tm = static_cast<const S*>(L.dT);
    const int64_t sB = TI ? 0 : kc * kc;
    S* Sv = w.S1;  // natural-order eta (n x m per box)
    if (l > 0) {
      // eta_J = [phi_dn ; T^T phi_dn], then natural order
      S* EJ = w.S2;
      k::copy2d<S, S>(st, kk, nb * m, w.PD[l], kk, 0, EJ, n, 0, 1);
      if (r > 0)
        level_matvec<S>(st, true, kk, r, S(1), Tm, kk, kk * r, w.PD[l], kk, S(0), nullptr, n,
                        EJ + kk, n, nb, m, TI, w.work);
      k::scatter<S, S>(st, n, m, L.dJ, TI ? 0 : n, EJ, n, n * m, Sv, n, n * m, nb, false);
    }
    // children: PD[2e] = B12 PU[2e+1] + eta[0:kc];  PD[2e+1] = B12^T PU[2e] + eta[kc:n]
    S* PUc = w.PU[l + 1];
    S* PDc = w.PD[l + 1];
    for (int s = 0; s < 2; ++s) {
      const S* xin = PUc + (1 - s) * kc * m;
      S* yout = PDc + s * kc * m;
      const S* z = Sv + s * kc;
      const S gamma = l > 0 ? S(1) : S(0);
      // per parent box: x block stride 2*kc*m, y block stride 2*kc*m, z stride n*m
      if (!TI && m == 1) {
        k::gemv<S>(st, s == 1, kc, kc, S(1), B12, kc, sB, xin, 2 * kc, gamma, z, n, yout, 2 * kc,
                   nb, w.work);
      } else {
        if (l > 0)
          k::add2d<S>(st, kc, m, S(1), z, n, n * m, S(0), z, n, n * m, yout, kc, 2 * kc * m, nb);
        k::gemm<S>(st, s == 1, false, kc, m, kc, S(1), B12, kc, sB, xin, kc, 2 * kc * m, gamma,
                   yout, kc, 2 * kc * m, nb);
      }
    }
  }
  // leaves: phi_J = Bself q_J + [phi_dn ; T^T phi_dn]
  {
    const Level& L = t->levels[D];
    const int64_t n = L.n, kk = L.k, r = n - kk, nb = L.nbox;
    const S* Tm = static_cast<const S*>(L.dT);
    S* EJ = w.S2;
    k::copy2d<S, S>(st, kk, nb * m, w.PD[D], kk, 0, EJ, n, 0, 1);
    if (r > 0)
      level_matvec<S>(st, true, kk, r, S(1), Tm, kk, kk * r, w.PD[D], kk, S(0), nullptr, n,
                      EJ + kk, n, nb, m, TI, w.work);
    level_matvec<S>(st, false, n, n, S(1), Bl(D), n, n * n, w.Q[D], n, S(1), EJ, n, w.S1, n, nb,
                    m, TI, w.work);
    k::scatter<S, S>(st, n, m, t->dLeafG, n, w.S1, n, n * m, Y, N, 0, nb, false);
  }
}

void apply_dispatch(const hssc_operator* op, const void* X, void* Y, int64_t m, hipStream_t st) {
  ensure_T64(op->tree);  // the sweeps read the FP64 T
  if (op->tree->cx)
    apply_impl<zdouble>(op, static_cast<const zdouble*>(X), static_cast<zdouble*>(Y), m, st);
  else
    apply_impl<double>(op, static_cast<const double*>(X), static_cast<double*>(Y), m, st);
}

template <typename S>
double device_norm(const S* x, int64_t len, hipStream_t st, double* part_d) {
  const int nblk = 256;
  k::sumsq<S>(st, len, x, part_d, nblk);
  double part[256];
  HIPCHK(hipMemcpyAsync(part, part_d, sizeof(part), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double s = 0;
  for (int i = 0; i < nblk; ++i) s += part[i];
  return std::sqrt(s);
}

// applyinv with the solve workspace ws (the inverse's own, or a graph's).
void applyinv_with(const hssc_inverse* inv, SolveWorks& ws, const void* X, void* Y, int64_t m,
                   hipStream_t st) {
  tree_sync(inv->tree);
  const bool fp64 = inv->prec == HSSC_PREC_FP64;
  if (inv->tree->cx) {
    const auto* Xz = static_cast<const zdouble*>(X);
    auto* Yz = static_cast<zdouble*>(Y);
    if (fp64) applyinv_impl<zdouble>(inv, ws.get<zdouble>(), Xz, Yz, m, st);
    else applyinv_impl<zfloat>(inv, ws.get<zfloat>(), Xz, Yz, m, st);
  } else {
    const auto* Xd = static_cast<const double*>(X);
    auto* Yd = static_cast<double*>(Y);
    if (fp64) applyinv_impl<double>(inv, ws.get<double>(), Xd, Yd, m, st);
    else applyinv_impl<float>(inv, ws.get<float>(), Xd, Yd, m, st);
  }
}

void applyinv_dispatch(const hssc_inverse* inv, const void* X, void* Y, int64_t m,
                       hipStream_t st) {
  applyinv_with(inv, inv->w, X, Y, m, st);
}

// The workspace of applyinv_with for m right-hand sides, allocated ahead (graph capture
// allocates nothing).
void ensure_works(const hssc_inverse* inv, SolveWorks& ws, int64_t m) {
  const bool fp64 = inv->prec == HSSC_PREC_FP64;
  if (inv->tree->cx) fp64 ? ws.get<zdouble>().ensure(inv->tree, m) : ws.get<zfloat>().ensure(inv->tree, m);
  else fp64 ? ws.get<double>().ensure(inv->tree, m) : ws.get<float>().ensure(inv->tree, m);
}

// Solves replayed from graphs: on (HSSC_SOLVE_GRAPHS unset or 1) unless kernel statistics
// are being collected (their events do not belong in a graph).
bool solve_graphs_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("HSSC_SOLVE_GRAPHS");
    return !v || std::atoi(v) != 0;
  }();
  return on && !stats::enabled();
}

// applyinv with the solve workspace laid out in caller-provided device memory of cap bytes.
template <typename T>
int applyinv_borrowed(const hssc_inverse* inv, const void* X, void* Y, int64_t m, hipStream_t st,
                      void* base, int64_t cap) {
  const int64_t need = SolveWork<T>::borrow_bytes(inv->tree, m);
  if (need > cap)
    return fail(HSSC_ERR_INVALID_ARG,
                "the solve workspace of this tree and m needs " + std::to_string(need) +
                    " bytes; the fixed buffer holds " + std::to_string(cap));
  SolveWork<T> w;
  w.borrow(inv->tree, m, base);
  applyinv_impl<T>(inv, w, static_cast<const hi_t<T>*>(X), static_cast<hi_t<T>*>(Y), m, st);
  w.release();  // borrowed: nothing is freed
  return HSSC_OK;
}

// Common/Iterative_Refinement.m:20-35: Y = Ainv(B); while ||B - A(Y)|| > tol ||B||: Y +=
// Ainv(B - A(Y)).  resid(Y, R) forms R = B - A(Y) for the residual operator A.
template <typename S, typename Resid>
int refine_impl(const hssc_inverse* inv, const S* B, S* Y, int64_t m, double tol, int32_t maxit,
                int32_t* iters, double* relres, hipStream_t st, Resid&& resid) {
  const int64_t len = inv->tree->N * m;
  S* R = dalloc<S>(len);
  S* D = dalloc<S>(len);
  double* part = dalloc<double>(256);
  applyinv_dispatch(inv, B, Y, m, st);
  const double nb = device_norm<S>(B, len, st, part);
  int it = 1;
  double rel = 0;
  for (;;) {
    resid(Y, R);
    rel = device_norm<S>(R, len, st, part) / (nb > 0 ? nb : 1.0);
    if (rel <= tol || it >= maxit) break;
    applyinv_dispatch(inv, R, D, m, st);
    k::axpby<S, S>(st, len, S(1), D, S(1), Y);
    ++it;
  }
  HIPCHK(hipStreamSynchronize(st));
  dfree(R), dfree(D), dfree(part);
  if (iters) *iters = it;
  if (relres) *relres = rel;
  return rel <= tol ? HSSC_OK : fail(HSSC_ERR_NOT_CONVERGED, "refinement did not reach tol");
}

// The refinement with A = the FP64 compressed operator.
template <typename S>
int solve_impl(const hssc_inverse* inv, const hssc_operator* op, const S* B, S* Y, int64_t m,
               double tol, int32_t maxit, int32_t* iters, double* relres, hipStream_t st) {
  const int64_t len = inv->tree->N * m;
  return refine_impl<S>(inv, B, Y, m, tol, maxit, iters, relres, st, [&](const S* y, S* r) {
    apply_dispatch(op, y, r, m, st);
    k::axpby<S, S>(st, len, S(1), B, S(-1), r);  // R = B - A Y
  });
}

// The refinement of refine_impl replayed from the inverse's graphs (captured on first use
// for this residual operator and m).  resid(y, b, r, s) forms r = b - A y on stream s.
template <typename S, typename Resid>
int refine_graphed(const hssc_inverse* inv, RefineGraphs& G, const S* B, S* Y, int64_t m,
                   double tol, int32_t maxit, int32_t* iters, double* relres, hipStream_t st,
                   Resid&& resid) {
  const int64_t len = inv->tree->N * m;
  const size_t bytes = size_t(len) * sizeof(S);
  if (!G.first) {
    tree_sync(inv->tree);
    HIPCHK(hipStreamCreateWithFlags(&G.cs, hipStreamNonBlocking));
    G.B = dalloc<S>(len), G.Y = dalloc<S>(len), G.R = dalloc<S>(len), G.D = dalloc<S>(len);
    G.part = dalloc<double>(512);
    ensure_works(inv, G.w, m);
    S* gB = static_cast<S*>(G.B);
    S* gY = static_cast<S*>(G.Y);
    S* gR = static_cast<S*>(G.R);
    S* gD = static_cast<S*>(G.D);
    auto capture = [&](hipGraphExec_t* out, auto&& body) {
      HIPCHK(hipStreamBeginCapture(G.cs, hipStreamCaptureModeThreadLocal));
      hipGraph_t g = nullptr;
      try {
        body();
      } catch (...) {
        (void)hipStreamEndCapture(G.cs, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
      }
      HIPCHK(hipStreamEndCapture(G.cs, &g));
      const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(e);
    };
    capture(&G.first, [&]() {
      applyinv_with(inv, G.w, gB, gY, m, G.cs);
      resid(gY, gB, gR, G.cs);
      k::sumsq<S>(G.cs, len, gR, G.part, 256);
      k::sumsq<S>(G.cs, len, gB, G.part + 256, 256);
    });
    capture(&G.step, [&]() {
      applyinv_with(inv, G.w, gR, gD, m, G.cs);
      k::axpby<S, S>(G.cs, len, S(1), gD, S(1), gY);
      resid(gY, gB, gR, G.cs);
      k::sumsq<S>(G.cs, len, gR, G.part, 256);
    });
  }
  // after the caller's work on st
  hipEvent_t ev = nullptr;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIPCHK(hipEventRecord(ev, st));
  HIPCHK(hipStreamWaitEvent(G.cs, ev, 0));
  (void)hipEventDestroy(ev);
  HIPCHK(hipMemcpyAsync(G.B, B, bytes, hipMemcpyDeviceToDevice, G.cs));
  HIPCHK(hipGraphLaunch(G.first, G.cs));
  double part[512];
  HIPCHK(hipMemcpyAsync(part, G.part, sizeof(part), hipMemcpyDeviceToHost, G.cs));
  HIPCHK(hipStreamSynchronize(G.cs));
  double sr = 0, sb = 0;
  for (int i = 0; i < 256; ++i) sr += part[i], sb += part[256 + i];
  const double nb = std::sqrt(sb) > 0 ? std::sqrt(sb) : 1.0;
  double rel = std::sqrt(sr) / nb;
  int it = 1;
  while (!(rel <= tol) && it < maxit) {
    HIPCHK(hipGraphLaunch(G.step, G.cs));
    HIPCHK(hipMemcpyAsync(part, G.part, 256 * sizeof(double), hipMemcpyDeviceToHost, G.cs));
    HIPCHK(hipStreamSynchronize(G.cs));
    sr = 0;
    for (int i = 0; i < 256; ++i) sr += part[i];
    rel = std::sqrt(sr) / nb;
    ++it;
  }
  HIPCHK(hipMemcpyAsync(Y, G.Y, bytes, hipMemcpyDeviceToDevice, G.cs));
  HIPCHK(hipStreamSynchronize(G.cs));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  return rel <= tol ? HSSC_OK : fail(HSSC_ERR_NOT_CONVERGED, "refinement did not reach tol");
}

// The refinement with A = the exact grid operator (FFT convolution).  The first solve with a
// given operator and m runs its launches directly; later ones replay the inverse's graphs.
template <typename S>
int solve_exact_impl(const hssc_inverse* inv, const hssc_grid_op* A, const S* B, S* Y,
                     int64_t m, double tol, int32_t maxit, int32_t* iters, double* relres,
                     hipStream_t st) {
  RefineGraphs& G = inv->graphs;
  const uint64_t uid = grid_op_uid(A);
  if (G.op != uid || G.m != m) {
    G.release();
    G.op = uid, G.m = m;
  }
  if (++G.calls >= 2 && solve_graphs_enabled())
    return refine_graphed<S>(inv, G, B, Y, m, tol, maxit, iters, relres, st,
                             [&](const S* y, const S* b, S* r, hipStream_t s) {
                               grid_op_apply(A, y, b, r, m, s);
                             });
  return refine_impl<S>(inv, B, Y, m, tol, maxit, iters, relres, st,
                        [&](const S* y, S* r) { grid_op_apply(A, y, B, r, m, st); });
}

}  // namespace
}  // namespace hssc

// ---------------------------------------------------------------------------------------
// Fixed device workspace of hssc_hss2d_applyinv_device_noalloc
// ---------------------------------------------------------------------------------------
// Every GPU the library runs on holds its own copy of this array from the module load until
// the process ends.  In host code its name only stands in for the symbol: the device address
// comes from hipGetSymbolAddress, looked up once per device with that device current.
static constexpr int64_t kNoallocScratchBytes = int64_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace hssc {
namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device ordinal

// Device address of the current device's scratch array (looked up on its first use there).
int scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (int64_t(g_scratch_ptr.size()) <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(HSSC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HSSC_OK;
}
}  // namespace
}  // namespace hssc

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

// Dense block inverses of hssc_dense_invert_device, on streams of the call.
template <typename T>
static void dense_invert(bool sym, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch,
                         int* singular) {
  OwnStream os, aux;
  const int64_t ws = k::invert_ws_elems<T>(n, batch);
  T* W = ws > 0 ? dalloc<T>(ws) : nullptr;
  int* dflag = dalloc<int>(1);
  HIPCHK(hipMemsetAsync(dflag, 0, sizeof(int), os.s));
  k::invert_batched<T>(os.s, n, A, lda, sA, batch, aux.s, W, sym, dflag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(os.s));
  HIPCHK(hipStreamSynchronize(aux.s));
  HIPCHK(hipMemcpy(singular, dflag, sizeof(int), hipMemcpyDeviceToHost));
  dfree(W);
  dfree(dflag);
}

extern "C" {

const char* hssc_last_error(void) { return hssc::g_last_error.c_str(); }

const char* hssc_version(void) { return "hssc-mi355x 0.1 (gfx950)"; }

int hssc_release_cached_memory(void) {
  if (int s = check_device()) return s;
  pool::trim();
  return HSSC_OK;
}

int hssc_cached_memory(int64_t* bytes) {
  if (!bytes) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *bytes = int64_t(pool::cached_bytes());
  return HSSC_OK;
}

int hssc_params_init(hssc_params* p, int32_t pot, double kh, int32_t sym, int32_t Ng,
                     int32_t np, double acc, int32_t lay, int32_t n_cut, int32_t TI) {
  if (!p) return fail(HSSC_ERR_INVALID_ARG, "null params");
  if (pot != HSSC_POT_SL_L_2D && pot != HSSC_POT_SL_H_2D)
    return fail(HSSC_ERR_INVALID_ARG, "unknown kernel");
  if (Ng < 1 || np < 2 || (np & (np - 1)) != 0)
    return fail(HSSC_ERR_INVALID_ARG, "need Ng >= 1 and np a power of two >= 2");
  if (!(acc > 0 && acc < 1)) return fail(HSSC_ERR_INVALID_ARG, "acc must be in (0,1)");
  if (lay < 1) return fail(HSSC_ERR_INVALID_ARG, "layers must be >= 1");
  std::memset(p, 0, sizeof(*p));
  p->flag_pot = pot;
  p->kh = kh;
  p->sym = sym;
  p->Ng = Ng;
  p->np = np;
  p->acc = acc;
  p->layers = lay;
  p->n_cut = n_cut;
  p->transinv = TI ? 1 : 0;
  // Common/HSS_tree_parameters.m:113,118: depth = round(2(log2 np - 1) + 3) - 1
  const int nlevels = int(std::lround(2 * (std::log(double(np)) / std::log(2.0) - 1) + 3));
  p->depth = nlevels - 1;
  p->h = 2.0 / (double(Ng) * np);
  p->max_particles = Ng * Ng;
  p->interform = HSSC_INTER_LOWRANK;
  p->k_cut = 2500;
  p->factor_precision = HSSC_PREC_FP64;
  p->tree_precision = HSSC_PREC_FP64;
  p->contrast = HSSC_CONTRAST_PLATEAU;
  p->contrast_amp = 0.5;
  p->contrast_width = 0.15;
  p->kernel_scale = 1.0;
  p->length_scale = 1.0;
  p->proxy_rule = HSSC_PROXY_BDRY;
  p->lambda = 2.0;
  if (kh > 0) {
    // :167-175, params.order = 4: dr_weights = corr0(h, kh)
    if (int s = hssc_dr_corr0(p->h, kh, &p->dr_weights_re, &p->dr_weights_im)) return s;
  }
  return HSSC_OK;
}

int hssc_params_set_dr_weights(hssc_params* p, double re, double im) {
  if (!p) return fail(HSSC_ERR_INVALID_ARG, "null params");
  p->dr_weights_re = re;
  p->dr_weights_im = im;
  return HSSC_OK;
}

int64_t hssc_params_npoints(const hssc_params* p) {
  if (!p) return -1;
  return int64_t(p->np) * p->np * p->Ng * p->Ng;
}

int hssc_source_points(const hssc_params* p, double* X) {
  if (!p || !X) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  hssc::source_points(p->Ng, p->np, X);
  return HSSC_OK;
}

int hssc_kernel_eval(const hssc_params* p, const double* X1, int64_t m, const double* X2,
                     int64_t n, double* A, int64_t lda) {
  if (!p || !X1 || !X2 || !A || m < 0 || n < 0 || lda < m)
    return fail(HSSC_ERR_INVALID_ARG, "invalid kernel_eval arguments");
  if (int s = check_params(*p)) return s;
  if (int s = check_device()) return s;
  const bool cx = p->flag_pot == HSSC_POT_SL_H_2D;
  const size_t es = cx ? sizeof(zdouble) : sizeof(double);
  return guarded([&]() {
    std::vector<double> w1(m), w2(n), x1(m), y1(m), x2(n), y2(n);
    for (int64_t i = 0; i < m; ++i)
      x1[i] = X1[i], y1[i] = X1[m + i], w1[i] = point_weight(*p, x1[i], y1[i], true);
    for (int64_t j = 0; j < n; ++j)
      x2[j] = X2[j], y2[j] = X2[n + j], w2[j] = point_weight(*p, x2[j], y2[j], false);
    double *dx1 = dupload(x1), *dy1 = dupload(y1), *dw1 = dupload(w1);
    double *dx2 = dupload(x2), *dy2 = dupload(y2), *dw2 = dupload(w2);
    void* dA = dalloc<zdouble>(std::max<int64_t>(1, m * n));
    const k::KernelSpec ks = kernel_spec(*p);
    if (cx)
      k::gen_block<zdouble>(0, m, n, 1, dx1, dy1, dw1, 0, dx2, dy2, dw2, 0,
                            static_cast<zdouble*>(dA), std::max<int64_t>(m, 1), 0, ks);
    else
      k::gen_block<double>(0, m, n, 1, dx1, dy1, dw1, 0, dx2, dy2, dw2, 0,
                           static_cast<double*>(dA), std::max<int64_t>(m, 1), 0, ks);
    HIPCHK(hipDeviceSynchronize());
    if (m > 0 && n > 0)
      HIPCHK(hipMemcpy2D(A, lda * es, dA, m * es, m * es, n, hipMemcpyDeviceToHost));
    dfree(dx1), dfree(dy1), dfree(dw1), dfree(dx2), dfree(dy2), dfree(dw2), dfree(dA);
    return HSSC_OK;
  });
}

static int build_index_sets_policy(
    const hssc_params* p, hssc_tree** tree, IdPolicy pol,
    const std::function<int(hssc_tree*, int)>& level_done = nullptr) {
  if (!p || !tree) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *tree = nullptr;
  if (int s = check_params(*p)) return s;
  if (pol.mode < 0 || pol.mode > 2 || (pol.mode == 2 && !pol.mlp) || pol.oversample < 0)
    return fail(HSSC_ERR_INVALID_ARG, "invalid rank prediction arguments");
  return guarded([&]() {
    auto* t = new hssc_tree();
    pol.tree = t;
    LevelDone done = nullptr;
    if (level_done) done = [&](int lev) { return level_done(t, lev); };
    const int s = build_index_sets(
        *p, t,
        [&](const hssc_params& pp, const IdJob& jb, int32_t* o, int32_t* r) {
          return id_select_gpu(pp, jb, o, r, pol);
        },
        done);
    if (s != HSSC_OK) {
      delete t;
      return s;
    }
    *tree = t;
    return HSSC_OK;
  });
}

static int build_tree_policy(const hssc_params* p, hssc_tree** tree, const IdPolicy& pol) {
  if (!p || !tree) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *tree = nullptr;
  if (int s = check_device()) return s;
  pool::reclaim();  // the device is expected idle here: released blocks become reusable
  // Each level's interpolation operators are queued as soon as its index sets are final, so
  // the GPU works on the levels below while the host builds the ones above.
  std::unique_ptr<InterpBuild> ib;
  int s;
  {
    pool::NoSync ns;  // allocations meanwhile do not wait for the queued levels
    s = build_index_sets_policy(p, tree, pol, [&](hssc_tree* t, int lev) {
      return guarded([&]() {
        if (!ib) ib = interp_build(t);
        ib->add(lev);
        return HSSC_OK;
      });
    });
    if (s == HSSC_OK) {
      hssc_tree* t = *tree;
      s = guarded([&]() {
        if (!ib) ib = interp_build(t);
        ib->finish();
        return HSSC_OK;
      });
      if (s != HSSC_OK) {
        ib.reset();
        delete t;
        *tree = nullptr;
      }
    }
  }
  // a failed build drains its queued levels before their workspaces return to the pool
  ib.reset();
  return s;
}

int hssc_hss2d_build_index_sets(const hssc_params* p, hssc_tree** tree) {
  return build_index_sets_policy(p, tree, IdPolicy{});
}

int hssc_hss2d_build_tree(const hssc_params* p, hssc_tree** tree) {
  return build_tree_policy(p, tree, IdPolicy{});
}

int hssc_hss2d_build_tree_ranked(const hssc_params* p, int32_t rank_mode, hssc_mlp* mlp,
                                 int32_t oversample, hssc_tree** tree) {
  IdPolicy pol;
  pol.mode = rank_mode, pol.mlp = mlp, pol.oversample = oversample;
  return build_tree_policy(p, tree, pol);
}

int hssc_tree_id_stats(const hssc_tree* t, int64_t* boxes, int64_t* fallbacks,
                       int64_t* sketch_rows, int64_t* full_rows) {
  if (!t) return fail(HSSC_ERR_INVALID_ARG, "null tree");
  if (boxes) *boxes = t->id_boxes;
  if (fallbacks) *fallbacks = t->id_fallbacks;
  if (sketch_rows) *sketch_rows = t->id_sketch_rows;
  if (full_rows) *full_rows = t->id_full_rows;
  return HSSC_OK;
}

int hssc_tree_box_features(const hssc_tree* t, int32_t lev, float* X) {
  if (!t || !X || lev < 0 || lev > t->depth) return fail(HSSC_ERR_INVALID_ARG, "bad level");
  const Level& L = t->levels[lev];
  IdJob jb{lev, L.nkey, L.n, L.M, L.PNx.data(), L.PNy.data(), L.PNw.data(),
           L.Xpx.data(), L.Xpy.data(), L.Xpw.data()};
  id_job_features(t->p, jb, X);
  return HSSC_OK;
}

int hssc_tree_depth(const hssc_tree* t, int32_t* depth) {
  if (!t || !depth) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *depth = t->depth;
  return HSSC_OK;
}

int hssc_tree_level_info(const hssc_tree* t, int32_t lev, int64_t* n, int64_t* k, int64_t* M,
                         int64_t* nbox, int64_t* nkey, int64_t* q) {
  if (!t || lev < 0 || lev > t->depth) return fail(HSSC_ERR_INVALID_ARG, "bad level");
  const Level& L = t->levels[lev];
  if (n) *n = L.n;
  if (k) *k = L.k;
  if (M) *M = L.M;
  if (nbox) *nbox = L.nbox;
  if (nkey) *nkey = L.nkey;
  if (q) *q = L.q;
  return HSSC_OK;
}

static int box_level(const hssc_tree* t, int64_t box, int* lev) {
  if (!t || box < 1 || box >= (int64_t(1) << (t->depth + 1)))
    return fail(HSSC_ERR_INVALID_ARG, "box number out of range");
  int l = 0;
  while ((int64_t(1) << (l + 1)) <= box) ++l;
  *lev = l;
  return HSSC_OK;
}

static int64_t box_key(const hssc_tree* t, int64_t box, int lev) {
  return t->p.transinv ? 0 : box - (int64_t(1) << lev);
}

int hssc_tree_get_isrc(const hssc_tree* t, int64_t box, int64_t* out) {
  int lev;
  if (int s = box_level(t, box, &lev)) return s;
  if (!out) return fail(HSSC_ERR_INVALID_ARG, "null output");
  std::copy(t->Isrc[box].begin(), t->Isrc[box].end(), out);
  return HSSC_OK;
}

int hssc_tree_box_ranks(const hssc_tree* t, int32_t lev, int32_t* ranks, int32_t* id_level) {
  if (!t || lev < 0 || lev > t->depth) return fail(HSSC_ERR_INVALID_ARG, "level out of range");
  const Level& L = t->levels[lev];
  if (ranks) std::copy(L.kbox.begin(), L.kbox.end(), ranks);
  if (id_level) *id_level = L.id_skel ? 1 : 0;
  return HSSC_OK;
}

int hssc_tree_get_isk(const hssc_tree* t, int64_t box, int64_t* out) {
  int lev;
  if (int s = box_level(t, box, &lev)) return s;
  if (!out) return fail(HSSC_ERR_INVALID_ARG, "null output");
  std::copy(t->Isk[box].begin(), t->Isk[box].end(), out);
  return HSSC_OK;
}

int hssc_tree_get_J(const hssc_tree* t, int64_t box, int64_t* out) {
  int lev;
  if (int s = box_level(t, box, &lev)) return s;
  if (!out) return fail(HSSC_ERR_INVALID_ARG, "null output");
  const Level& L = t->levels[lev];
  const int64_t key = box_key(t, box, lev);
  for (int64_t a = 0; a < L.n; ++a) out[a] = L.J[key * L.n + a];
  return HSSC_OK;
}

int hssc_tree_get_proxy(const hssc_tree* t, int64_t box, double* out) {
  int lev;
  if (int s = box_level(t, box, &lev)) return s;
  if (!out) return fail(HSSC_ERR_INVALID_ARG, "null output");
  const Level& L = t->levels[lev];
  const int64_t key = box_key(t, box, lev);
  for (int64_t r = 0; r < L.M; ++r) {
    out[r] = L.Xpx[key * L.M + r];
    out[L.M + r] = L.Xpy[key * L.M + r];
  }
  return HSSC_OK;
}

int hssc_tree_get_T(const hssc_tree* t, int64_t box, double* out) {
  int lev;
  if (int s = box_level(t, box, &lev)) return s;
  if (!out) return fail(HSSC_ERR_INVALID_ARG, "null output");
  if (!t->has_T) return fail(HSSC_ERR_INVALID_ARG, "tree has no interpolation operators");
  ensure_T64(const_cast<hssc_tree*>(t));
  const Level& L = t->levels[lev];
  if (lev == 0) return fail(HSSC_ERR_INVALID_ARG, "the root has no interpolation operator");
  const int64_t key = box_key(t, box, lev), cnt = L.k * (L.n - L.k);
  if (cnt == 0) return HSSC_OK;
  const size_t es = t->cx ? sizeof(zdouble) : sizeof(double);
  return guarded([&]() {
    HIPCHK(hipMemcpy(out, static_cast<const char*>(L.dT) + key * cnt * es, cnt * es,
                     hipMemcpyDeviceToHost));
    return HSSC_OK;
  });
}

void hssc_tree_destroy(hssc_tree* t) {
  hssc::tree_unref(t);  // drops the caller's reference; the last inverse or operator frees it
}

int hssc_levhss(const hssc_tree* t, int32_t n_cut, int32_t* lhss) {
  // HSS2D/levHSS.m:33-65: first level whose skeletons are all <= n_cut (TI: the level k).
  if (!t || !lhss) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  int lev = 0;
  int64_t n_min = t->levels[0].k;
  if (!t->p.transinv && t->depth >= 1) n_min = std::max(n_min, t->levels[1].k);
  while (n_min > n_cut && lev < t->depth - 1) {
    ++lev;
    n_min = t->levels[lev].k;  // uniform tree: every box of a level has the same k
  }
  if (lev > 0) *lhss = (lev % 2 == 1) ? lev - 1 : lev;
  else *lhss = -1;
  return HSSC_OK;
}

int hssc_dense_invert_device(int32_t complex_entries, int32_t precision, int32_t sym, int64_t n,
                             void* A, int64_t lda, int64_t sA, int64_t batch) {
  if (!A || n < 1 || lda < n || batch < 1 || (batch > 1 && sA < lda * n))
    return fail(HSSC_ERR_INVALID_ARG, "invalid dense inverse arguments");
  if (precision != HSSC_PREC_FP64 && precision != HSSC_PREC_FP32)
    return fail(HSSC_ERR_INVALID_ARG, "unknown precision");
  if (int s = check_device()) return s;
  int singular = 0;
  int st = guarded([&]() {
    const bool s = sym != 0;
    if (complex_entries && precision == HSSC_PREC_FP64)
      dense_invert(s, n, static_cast<zdouble*>(A), lda, sA, batch, &singular);
    else if (complex_entries)
      dense_invert(s, n, static_cast<zfloat*>(A), lda, sA, batch, &singular);
    else if (precision == HSSC_PREC_FP64)
      dense_invert(false, n, static_cast<double*>(A), lda, sA, batch, &singular);
    else
      dense_invert(false, n, static_cast<float*>(A), lda, sA, batch, &singular);
    return HSSC_OK;
  });
  if (st != HSSC_OK) return st;
  if (singular) return fail(HSSC_ERR_SINGULAR, "singular or non-finite pivot in a dense inverse");
  return HSSC_OK;
}

int hssc_hss2d_invert(hssc_tree* t, int32_t precision, hssc_inverse** out) {
  if (!t || !out) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!t->has_T) return fail(HSSC_ERR_INVALID_ARG, "tree has no interpolation operators");
  if (precision != HSSC_PREC_FP64 && precision != HSSC_PREC_FP32)
    return fail(HSSC_ERR_INVALID_ARG, "unknown precision");
  return guarded([&]() {
    auto* inv = new hssc_inverse(t);
    inv->prec = precision;
    try {
      OwnStream os;
      if (t->cx) {
        if (precision == HSSC_PREC_FP64) invert_impl<zdouble>(t, inv, os.s);
        else invert_impl<zfloat>(t, inv, os.s);
      } else {
        if (precision == HSSC_PREC_FP64) invert_impl<double>(t, inv, os.s);
        else invert_impl<float>(t, inv, os.s);
      }
    } catch (...) {
      delete inv;
      throw;
    }
    *out = inv;
    return HSSC_OK;
  });
}

static int inverse_block(const hssc_inverse* inv, int64_t box, double* out, bool wantE) {
  const hssc_tree* t = inv ? inv->tree : nullptr;
  int lev;
  if (int s = box_level(t, box, &lev)) return s;
  if (!out) return fail(HSSC_ERR_INVALID_ARG, "null output");
  const Level& L = t->levels[lev];
  if (wantE && lev == 0) return fail(HSSC_ERR_INVALID_ARG, "the root has no E");
  const int64_t key = box_key(t, box, lev);
  const int64_t n = wantE ? L.k : L.n;
  // complex blocks come back as interleaved complex double (2 doubles per entry)
  const int64_t nr = t->cx ? 2 : 1;
  return guarded([&]() {
    std::vector<double> J(n * n * nr);
    const void* src = wantE ? inv->E[lev] : inv->FI[lev];
    if (inv->prec == HSSC_PREC_FP64) {
      HIPCHK(hipMemcpy(J.data(), static_cast<const double*>(src) + key * n * n * nr,
                       n * n * nr * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> f(n * n * nr);
      HIPCHK(hipMemcpy(f.data(), static_cast<const float*>(src) + key * n * n * nr,
                       n * n * nr * 4, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n * n * nr; ++i) J[i] = f[i];
    }
    if (wantE) {
      std::copy(J.begin(), J.end(), out);
    } else {
      // stored in J order: FI(J[a], J[c]) = stored(a, c)
      const int32_t* Jp = L.J.data() + key * L.n;
      for (int64_t c = 0; c < n; ++c)
        for (int64_t a = 0; a < n; ++a)
          for (int64_t z = 0; z < nr; ++z)
            out[(Jp[a] + int64_t(Jp[c]) * n) * nr + z] = J[(a + c * n) * nr + z];
    }
    return HSSC_OK;
  });
}

int hssc_inverse_get_FI(const hssc_inverse* inv, int64_t box, double* out) {
  return inverse_block(inv, box, out, false);
}

int hssc_inverse_get_E(const hssc_inverse* inv, int64_t box, double* out) {
  return inverse_block(inv, box, out, true);
}

void hssc_inverse_destroy(hssc_inverse* inv) { delete inv; }

int hssc_hss2d_compress(hssc_tree* t, hssc_operator** out) {
  // HSS2D/HSS2D_compress_fsym.m:73-138: Bself = K(I_src, I_src) at leaves (J order here),
  // B12 = K(I_sk(c1), I_sk(c2)) at every other box (skeleton order).
  if (!t || !out) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!t->has_T) return fail(HSSC_ERR_INVALID_ARG, "tree has no interpolation operators");
  return guarded([&]() {
    auto* op = new hssc_operator(t);
    const int D = t->depth;
    op->B.assign(D + 1, nullptr);
    const k::KernelSpec ks = kernel_spec(t->p);
    hipStream_t st = nullptr;
    auto build = [&](auto tag) {
      using S = decltype(tag);
      for (int l = 0; l <= D; ++l) {
        const Level& L = t->levels[l];
        if (l == D) {
          S* B = dalloc<S>(L.nkey * L.n * L.n);
          op->B[l] = B;
          k::gen_block<S>(st, L.n, L.n, L.nkey, L.dPx, L.dPy, L.dPw, L.n, L.dPx, L.dPy, L.dPw, L.n,
                          B, L.n, L.n * L.n, ks);
        } else {
          const int64_t kc = t->levels[l + 1].k;
          S* B = dalloc<S>(L.nkey * kc * kc);
          op->B[l] = B;
          k::gen_block<S>(st, kc, kc, L.nkey, L.dPNx, L.dPNy, L.dPNw, L.n, L.dPNx + kc,
                          L.dPNy + kc, L.dPNw + kc, L.n, B, kc, kc * kc, ks);
        }
      }
    };
    try {
      // a stream of its own: the tree's interpolation operators may still be in the making on
      // the tree's streams, and the blocks made here do not need them
      OwnStream os;
      st = os.s;
      if (t->cx) build(zdouble());
      else build(double());
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipGetLastError());
    } catch (...) {
      delete op;
      throw;
    }
    *out = op;
    return HSSC_OK;
  });
}

void hssc_operator_destroy(hssc_operator* op) { delete op; }

int hssc_hss2d_applyinv_device(const hssc_inverse* inv, const void* dX, void* dY, int64_t m,
                               void* stream) {
  if (!inv || !dX || !dY || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid applyinv arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return guarded([&]() {
    applyinv_dispatch(inv, dX, dY, m, st);
    HIPCHK(hipGetLastError());
    return HSSC_OK;
  });
}

int hssc_hss2d_applyinv_device_noalloc(const hssc_inverse* inv, const void* dX, void* dY,
                                       int64_t m, void* stream) {
  if (!inv || !dX || !dY || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid applyinv arguments");
  const hssc_tree* t = inv->tree;
  // the inversion has waited for the tree's build; nothing is left to release here
  if (!t->bstreams.empty())
    return fail(HSSC_ERR_INVALID_ARG, "the tree's interpolation operators are still pending");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != t->device)
    return fail(HSSC_ERR_INVALID_ARG, "the current device is not the inverse's device");
  void* base = nullptr;
  if (int s = scratch_address(dev, &base)) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool fp64 = inv->prec == HSSC_PREC_FP64;
  return guarded([&]() {
    int s;
    if (t->cx)
      s = fp64 ? applyinv_borrowed<zdouble>(inv, dX, dY, m, st, base, kNoallocScratchBytes)
               : applyinv_borrowed<zfloat>(inv, dX, dY, m, st, base, kNoallocScratchBytes);
    else
      s = fp64 ? applyinv_borrowed<double>(inv, dX, dY, m, st, base, kNoallocScratchBytes)
               : applyinv_borrowed<float>(inv, dX, dY, m, st, base, kNoallocScratchBytes);
    if (s == HSSC_OK) HIPCHK(hipGetLastError());
    return s;
  });
}

int hssc_hss2d_apply_device(const hssc_operator* op, const void* dX, void* dY, int64_t m,
                            void* stream) {
  if (!op || !dX || !dY || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid apply arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return guarded([&]() {
    apply_dispatch(op, dX, dY, m, st);
    HIPCHK(hipGetLastError());
    return HSSC_OK;
  });
}

// X, Y: N x m doubles, or N x m interleaved complex doubles for a complex kernel
static int host_wrap(int64_t N, int64_t m, bool cx, const double* X, double* Y,
                     const std::function<int(const void*, void*)>& f) {
  const size_t bytes = size_t(N * m) * (cx ? sizeof(zdouble) : sizeof(double));
  return guarded([&]() {
    void* dX = dalloc<char>(bytes);
    void* dY = dalloc<char>(bytes);
    HIPCHK(hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice));
    int s = f(dX, dY);
    if (s == HSSC_OK) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipMemcpy(Y, dY, bytes, hipMemcpyDeviceToHost));
    }
    dfree(dX), dfree(dY);
    return s;
  });
}

int hssc_hss2d_applyinv(const hssc_inverse* inv, const double* X, double* Y, int64_t m) {
  if (!inv || !X || !Y || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid applyinv arguments");
  return host_wrap(inv->tree->N, m, inv->tree->cx, X, Y, [&](const void* dX, void* dY) {
    return hssc_hss2d_applyinv_device(inv, dX, dY, m, nullptr);
  });
}

int hssc_hss2d_apply(const hssc_operator* op, const double* X, double* Y, int64_t m) {
  if (!op || !X || !Y || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid apply arguments");
  return host_wrap(op->tree->N, m, op->tree->cx, X, Y, [&](const void* dX, void* dY) {
    return hssc_hss2d_apply_device(op, dX, dY, m, nullptr);
  });
}

int hssc_hss2d_solve_device(const hssc_inverse* inv, const hssc_operator* op, const void* dB,
                            void* dY, int64_t m, double tol, int32_t maxit, int32_t* iters,
                            double* relres, void* stream) {
  // Common/Iterative_Refinement.m:20-35 with A = the FP64 compressed operator.
  if (!inv || !op || !dB || !dY || m < 1 || maxit < 1 || inv->tree != op->tree)
    return fail(HSSC_ERR_INVALID_ARG, "invalid solve arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return guarded([&]() {
    if (inv->tree->cx)
      return solve_impl<zdouble>(inv, op, static_cast<const zdouble*>(dB), static_cast<zdouble*>(dY),
                                 m, tol, maxit, iters, relres, st);
    return solve_impl<double>(inv, op, static_cast<const double*>(dB), static_cast<double*>(dY), m,
                              tol, maxit, iters, relres, st);
  });
}

int hssc_hss2d_solve_exact_device(const hssc_inverse* inv, const hssc_grid_op* A,
                                  const void* dB, void* dY, int64_t m, double tol,
                                  int32_t maxit, int32_t* iters, double* relres, void* stream) {
  // Common/Iterative_Refinement.m:20-35 with A = the exact grid operator.
  if (!inv || !A || !dB || !dY || m < 1 || maxit < 1)
    return fail(HSSC_ERR_INVALID_ARG, "invalid solve arguments");
  if (grid_op_n(A) != inv->tree->N || grid_op_complex(A) != inv->tree->cx)
    return fail(HSSC_ERR_INVALID_ARG, "the grid operator and the inverse belong to other grids");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return guarded([&]() {
    if (inv->tree->cx)
      return solve_exact_impl<zdouble>(inv, A, static_cast<const zdouble*>(dB),
                                       static_cast<zdouble*>(dY), m, tol, maxit, iters, relres, st);
    return solve_exact_impl<double>(inv, A, static_cast<const double*>(dB),
                                    static_cast<double*>(dY), m, tol, maxit, iters, relres, st);
  });
}

}  // extern "C"

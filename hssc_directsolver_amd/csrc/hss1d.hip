// HSS1D compress / invert / apply for second-kind boundary integral equations on curves
// (HSS1D/HSS1D_compress_nsym.m 'green', HSS1D_invert_nsym.m, HSS1D_applyinv_nsym.m,
// HSS1D_apply_nsym.m).
//
// Ranks come from adaptive IDs, so every level is a ragged batch: per-box sizes and offsets
// live in device arrays and each kernel runs one workgroup per box (the boxes of one level
// are independent).  The host builds the binary tree, neighbour lists and proxy circles
// (integer / geometry work) exactly as the reference does; the GPU generates every kernel
// block, runs the interpolative decompositions, the inversion and the sweeps.
//
// Operators (hssc_hss1d.kind):
//   0  A = -1/2 I + D, D the Laplace double layer with per-point normals, weights and
//      curvature (diagonal -kappa w / 4pi); proxy sources carry log charges (1/2pi) log|x - z|;
//   1  Kernel_Eval(X, X, params) on arbitrary points (Common/Kernel_Eval.m, real kernels),
//      proxy rows and columns by the same kernel ('green' as the reference runs it);
//   2  a given dense matrix ('brute', HSS1D_compress_nsym.m:158-181), no geometry.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "pool.h"
#include "internal.h"
#include "kernels.h"
#include "stats.h"

namespace hssc {
k::KernelSpec kernel_spec(const hssc_params& p);
namespace {

struct HipErr {
  int status;
};

#define H1CHK(expr)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fail(e_ == hipErrorOutOfMemory ? HSSC_ERR_OUT_OF_MEMORY : HSSC_ERR_HIP,              \
           std::string(#expr) + ": " + hipGetErrorString(e_));                             \
      throw HipErr{e_ == hipErrorOutOfMemory ? HSSC_ERR_OUT_OF_MEMORY : HSSC_ERR_HIP};     \
    }                                                                                      \
  } while (0)

template <typename T>
T* dmalloc(size_t n) {
  void* p = nullptr;
  H1CHK(pool::alloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

template <typename T>
T* dput(const std::vector<T>& h) {
  T* d = dmalloc<T>(h.size());
  if (!h.empty()) H1CHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

template <typename T>
std::vector<T> dget(const T* d, size_t n) {
  std::vector<T> h(n);
  if (n) H1CHK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

void dfree1(void* p) { pool::release(p); }

// Runs f(i) for i in [0, n) on up to 16 host threads (independent iterations).
template <class F>
void par_for(int n, F&& f) {
  const int nt = std::max(1, std::min({n / 64 + 1, int(std::thread::hardware_concurrency()), 16}));
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (int i = t; i < n; i += nt) f(i);
    });
  for (auto& th : pool) th.join();
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kRadiusRel = 1.75;  // HSS1D/HSS1D_compress_nsym.m:51

// ------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------

struct CurveDev {
  const double *x, *y, *nx, *ny, *w, *kap;
};

// Where the entries of an operator come from (hssc_hss1d.kind).
struct Gen1D {
  int kind;           // 0 double layer, 1 Kernel_Eval on points, 2 dense matrix
  CurveDev C;         // points (kinds 0, 1), normals / weights / curvature (kind 0)
  k::KernelSpec ks;   // kind 1
  const double* wb;   // kind 1: Kernel_Eval's row factor b at every point
  const double* wc;   // kind 1: column factor c at every point
  const double* K;    // kind 2: N x N, column-major, leading dimension ldk
  int64_t ldk;
};

// Double layer at target (tx,ty) from curve source j; tgt = index of the target if it is a
// curve point (self term), -1 otherwise.
__device__ __forceinline__ double dl_entry(const CurveDev& C, double tx, double ty, int tgt, int j) {
  if (tgt == j) return -0.5 - C.kap[j] * C.w[j] / (4 * kPi);
  const double dx = tx - C.x[j], dy = ty - C.y[j];
  const double r2 = dx * dx + dy * dy;
  return (dx * C.nx[j] + dy * C.ny[j]) / r2 / (2 * kPi) * C.w[j];
}

// Ragged block generation.  Matrix b is m x n at A + off[b] (leading dimension m).
//   mode 0 (A_out rows x sources): rows r < mc[b] are curve targets ridx[ro[b]+r], the other
//          rows proxy targets (px, py, weight pw at pofs[b] + r - mc); columns are curve
//          sources cidx.
//   mode 1 (A_in^T): rows r < mc[b] are curve sources ridx[..], the other rows proxy
//          sources (log charges for the double layer); columns are curve targets cidx.
// Kernel_Eval entries take b at the target and c at the source (Common/Kernel_Eval.m:35);
// a proxy point's own factor is pw.
__global__ void __launch_bounds__(256) gen1d_kernel(
    Gen1D G, int mode, const int64_t* __restrict__ off, const int* __restrict__ mr,
    const int* __restrict__ nc, const int* __restrict__ mcv, const int* __restrict__ ridx,
    const int64_t* __restrict__ ro, const int* __restrict__ cidx, const int64_t* __restrict__ co,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pw,
    const int64_t* __restrict__ pofs, double* __restrict__ A) {
  const int b = blockIdx.y;
  const int m = mr[b], n = nc[b], mc = mcv[b];
  const CurveDev& C = G.C;
  double* M = A + off[b];
  for (int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x; idx < int64_t(m) * n;
       idx += int64_t(gridDim.x) * 256) {
    const int r = int(idx % m), c = int(idx / m);
    const int cj = cidx[co[b] + c];
    double v;
    if (G.kind == 2) {
      const int ri = ridx[ro[b] + r];  // dense: no proxy rows (mc == m)
      v = mode == 0 ? G.K[ri + int64_t(cj) * G.ldk] : G.K[cj + int64_t(ri) * G.ldk];
    } else if (mode == 0) {
      if (r < mc) {
        const int ti = ridx[ro[b] + r];
        v = G.kind == 0 ? dl_entry(C, C.x[ti], C.y[ti], ti, cj)
                        : k::kernel_entry<double>(G.ks, C.x[ti], C.y[ti], G.wb[ti], C.x[cj],
                                                  C.y[cj], G.wc[cj]);
      } else {
        const int64_t q = pofs[b] + r - mc;
        v = G.kind == 0 ? dl_entry(C, px[q], py[q], -1, cj)
                        : k::kernel_entry<double>(G.ks, px[q], py[q], pw[q], C.x[cj], C.y[cj],
                                                  G.wc[cj]);
      }
    } else {
      if (r < mc) {
        const int sj = ridx[ro[b] + r];
        v = G.kind == 0 ? dl_entry(C, C.x[cj], C.y[cj], cj, sj)
                        : k::kernel_entry<double>(G.ks, C.x[cj], C.y[cj], G.wb[cj], C.x[sj],
                                                  C.y[sj], G.wc[sj]);
      } else {
        const int64_t q = pofs[b] + r - mc;
        if (G.kind == 0) {
          const double dx = C.x[cj] - px[q], dy = C.y[cj] - py[q];
          v = log(dx * dx + dy * dy) / (4 * kPi);
        } else {
          v = k::kernel_entry<double>(G.ks, C.x[cj], C.y[cj], G.wb[cj], px[q], py[q], pw[q]);
        }
      }
    }
    M[idx] = v;
  }
}

// Per-box telescoping inverse (HSS1D/HSS1D_invert_nsym.m:44-83), one workgroup per box,
// all blocks in global scratch (they are at most 128 x 128 and stay in L2).
struct BoxInv {
  int n, k;
  int64_t F;      // F (n x n) -> Finv in place (assembled by the kernel)
  int64_t Jup, Jdn, Tup, Tdn;  // permutations (ints) and T (k x (n-k))
  int64_t E;      // E (k x k)
  int64_t RF;     // R Finv (k x n)
  int64_t FL;     // Finv L (n x k)
  int64_t Linv, Rinv, Dinv;
  // assembly: leaf -> Mself offset; else children E offsets and M_sib offsets
  int leaf;
  int64_t Mself, E1, E2, S1, S2;
  int k1, k2;
};

template <typename T>
__device__ void gj_inplace(T* A, int n, int* piv, T* f, int* sp) {
  // Gauss-Jordan with partial pivoting on an n x n column-major matrix in global memory
  // (one workgroup); swaps undone by columns at the end.
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int c = 0; c < n; ++c) {
    if (tid < 64) {
      T best = -1;
      int bi = c;
      for (int r = c + tid; r < n; r += 64) {
        const T v = fabs(A[r + c * n]);
        if (v > best) best = v, bi = r;
      }
      for (int o = 32; o > 0; o >>= 1) {
        const T ob = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ob > best || (ob == best && oi < bi)) best = ob, bi = oi;
      }
      if (tid == 0) *sp = bi, piv[c] = bi;
    }
    __syncthreads();
    const int p = *sp;
    if (p != c)
      for (int j = tid; j < n; j += nt) {
        const T t = A[c + j * n];
        A[c + j * n] = A[p + j * n];
        A[p + j * n] = t;
      }
    __syncthreads();
    for (int i = tid; i < n; i += nt) f[i] = A[i + c * n];
    __syncthreads();
    const T inv = T(1) / f[c];
    for (int j = tid; j < n; j += nt) A[c + j * n] = (j == c) ? inv : A[c + j * n] * inv;
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += nt) {
      const int i = idx % n, j = idx / n;
      if (i == c) continue;
      if (j == c) A[i + j * n] = -f[i] * inv;
      else A[i + j * n] -= f[i] * A[c + j * n];
    }
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; --c) {
    const int p = piv[c];
    if (p != c) {
      for (int i = tid; i < n; i += nt) {
        const T t = A[i + c * n];
        A[i + c * n] = A[i + p * n];
        A[i + p * n] = t;
      }
      __syncthreads();
    }
  }
}

// Blocks of the inverse (F, E, R Finv, Finv L, Linv, Rinv, Dinv) live in W (element type T:
// double for the FP64 inverse, float for the FP32 one), at their arena offsets minus wofs;
// the compressed operator (M_self, M_sib, T, permutations) is read from S (double) and I.
template <typename T>
__global__ void __launch_bounds__(256) hss1d_invert_kernel(const BoxInv* __restrict__ boxes,
                                                           const double* __restrict__ S,
                                                           T* __restrict__ W, int64_t wofs,
                                                           const int* __restrict__ I) {
  __shared__ T f[256];
  __shared__ int piv[256];
  __shared__ int sp;
  const BoxInv B = boxes[blockIdx.x];
  const int n = B.n, k = B.k, tid = threadIdx.x, nt = blockDim.x;
  W -= wofs;  // W[o] for arena offsets o >= wofs
  T* F = W + B.F;
  // assemble F
  for (int idx = tid; idx < n * n; idx += nt) {
    const int i = idx % n, j = idx / n;
    T v;
    if (B.leaf) {
      v = T(S[B.Mself + idx]);
    } else if (i < B.k1 && j < B.k1) {
      v = W[B.E1 + i + j * B.k1];
    } else if (i >= B.k1 && j >= B.k1) {
      v = W[B.E2 + (i - B.k1) + (j - B.k1) * B.k2];
    } else if (i < B.k1) {
      v = T(S[B.S1 + i + (j - B.k1) * B.k1]);  // M_sib(c1): k1 x k2
    } else {
      v = T(S[B.S2 + (i - B.k1) + j * B.k2]);  // M_sib(c2): k2 x k1
    }
    F[idx] = v;
  }
  __syncthreads();
  gj_inplace<T>(F, n, piv, f, &sp);
  __syncthreads();
  const int* Ju = I + B.Jup;
  const int* Jd = I + B.Jdn;
  const double* Tu = S + B.Tup;  // k x (n-k)
  const double* Td = S + B.Tdn;
  T* RF = W + B.RF;
  T* FL = W + B.FL;
  // RF = R Finv: row i = Finv(Ju[i], :) + sum_c Tu(i, c) Finv(Ju[k+c], :)
  for (int idx = tid; idx < k * n; idx += nt) {
    const int i = idx % k, j = idx / k;
    T s = F[Ju[i] + j * n];
    for (int c = 0; c < n - k; ++c) s += T(Tu[i + c * k]) * F[Ju[k + c] + j * n];
    RF[idx] = s;
  }
  // FL = Finv L: column j = Finv(:, Jd[j]) + sum_c Finv(:, Jd[k+c]) Td(j, c)
  for (int idx = tid; idx < n * k; idx += nt) {
    const int i = idx % n, j = idx / n;
    T s = F[i + Jd[j] * n];
    for (int c = 0; c < n - k; ++c) s += F[i + Jd[k + c] * n] * T(Td[j + c * k]);
    FL[idx] = s;
  }
  __syncthreads();
  // E^-1 = RF L = R Finv L
  T* E = W + B.E;
  for (int idx = tid; idx < k * k; idx += nt) {
    const int i = idx % k, j = idx / k;
    T s = RF[i + Jd[j] * k];
    for (int c = 0; c < n - k; ++c) s += RF[i + Jd[k + c] * k] * T(Td[j + c * k]);
    E[idx] = s;
  }
  __syncthreads();
  gj_inplace<T>(E, k, piv, f, &sp);
  __syncthreads();
  T* Li = W + B.Linv;
  T* Ri = W + B.Rinv;
  T* Di = W + B.Dinv;
  for (int idx = tid; idx < n * k; idx += nt) {
    const int i = idx % n, j = idx / n;
    T s = 0;
    for (int q = 0; q < k; ++q) s += FL[i + q * n] * E[q + j * k];
    Li[idx] = s;  // Linv = Finv L E
  }
  for (int idx = tid; idx < k * n; idx += nt) {
    const int i = idx % k, j = idx / k;
    T s = 0;
    for (int q = 0; q < k; ++q) s += E[i + q * k] * RF[q + j * k];
    Ri[idx] = s;  // Rinv = E R Finv
  }
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += nt) {
    const int i = idx % n, j = idx / n;
    T s = F[idx];
    for (int q = 0; q < k; ++q) s -= Li[i + q * n] * RF[q + j * k];
    Di[idx] = s;  // Dinv = Finv - Linv R Finv
  }
}

// The top block of each problem: inv([E2 M_sib2; M_sib3 E3]) for the root's children
// (HSS1D_invert_nsym.m:88-91), one workgroup per problem.
struct TopOp {
  int64_t FT, E2, E3, S2, S3;
  int k2, k3;
};

template <typename T>
__global__ void __launch_bounds__(256) hss1d_top_kernel(const TopOp* __restrict__ ops,
                                                        const double* __restrict__ S,
                                                        T* __restrict__ W, int64_t wofs) {
  __shared__ T f[256];
  __shared__ int piv[256];
  __shared__ int sp;
  const TopOp O = ops[blockIdx.x];
  const int64_t FT = O.FT, E2 = O.E2, E3 = O.E3, S2 = O.S2, S3 = O.S3;
  const int k2 = O.k2, k3 = O.k3;
  const int n = k2 + k3, tid = threadIdx.x;
  W -= wofs;
  T* F = W + FT;
  for (int idx = tid; idx < n * n; idx += blockDim.x) {
    const int i = idx % n, j = idx / n;
    T v;
    if (i < k2 && j < k2) v = W[E2 + i + j * k2];
    else if (i >= k2 && j >= k2) v = W[E3 + (i - k2) + (j - k2) * k3];
    else if (i < k2) v = T(S[S2 + i + (j - k2) * k2]);
    else v = T(S[S3 + (i - k2) + j * k3]);
    F[idx] = v;
  }
  __syncthreads();
  gj_inplace<T>(F, n, piv, f, &sp);
}

// One level of a sweep: out_box = sum_t M_t x_t (up to three terms), one workgroup per
// box.  Term t reads M_t(rofs + i, c) = S[mat + rofs + i + c*ld]; its input is either the
// global vector at gofs (src_global) or the concatenation V[in .. in+len1) ++ V[in2 ..).
// The output goes to V[out ..] (rows >= split to V[out2 ..] when out2 >= 0) or to the
// global vector at gofs (scatter).
struct SweepTerm {
  int64_t mat;
  int ld, rofs, cols, len1, src_global;
  int64_t in, in2;
};
struct SweepOp {
  int rows, nterm;
  SweepTerm t[3];
  int64_t out, out2;
  int split, scatter;
  int64_t gofs;
};

// The matrices come from one arena (Tm: double, or float for the FP32 inverse) at offset -
// sofs; vectors and accumulation stay in double.  blockIdx.y is the right-hand side: column
// j of the global vectors starts at j * ldg, of the level vectors at j * ldv.
template <typename Tm>
__global__ void __launch_bounds__(256) hss1d_sweep_kernel(const SweepOp* __restrict__ ops,
                                                          const Tm* __restrict__ S, int64_t sofs,
                                                          double* __restrict__ V,
                                                          const double* __restrict__ gin,
                                                          double* __restrict__ gout, int64_t ldg,
                                                          int64_t ldv) {
  const SweepOp O = ops[blockIdx.x];
  S -= sofs;
  V += int64_t(blockIdx.y) * ldv;
  gin += int64_t(blockIdx.y) * ldg;
  gout += int64_t(blockIdx.y) * ldg;
  for (int i = threadIdx.x; i < O.rows; i += blockDim.x) {
    double s = 0;
    for (int t = 0; t < O.nterm; ++t) {
      const SweepTerm& T = O.t[t];
      const Tm* M = S + T.mat + T.rofs + i;
      for (int c = 0; c < T.cols; ++c) {
        double x;
        if (T.src_global) x = gin[O.gofs + c];
        else if (c < T.len1) x = V[T.in + c];
        else x = V[T.in2 + c - T.len1];
        s += double(M[int64_t(c) * T.ld]) * x;
      }
    }
    if (O.scatter) gout[O.gofs + i] = s;
    else if (O.out2 >= 0 && i >= O.split) V[O.out2 + i - O.split] = s;
    else V[O.out + i] = s;
  }
}

// One box of the compressed operator laid out in a level's block arena: T_up and T_dn
// (k x (n-k)), R = [I T_up] with columns permuted by J_up (k x n) and L = [I; T_dn^T] with
// rows permuted by J_dn (n x k).  J_up and J_dn are permutations, so every entry of R and L
// is written once.
struct PackOp {
  const double *tu, *td;
  int k, n;
  int64_t otu, otd, orr, ol;
  int64_t pu, pd;  // offsets of J_up, J_dn in the permutation array
};

__global__ void __launch_bounds__(256) hss1d_pack_kernel(const PackOp* __restrict__ ops,
                                                         const int* __restrict__ J,
                                                         double* __restrict__ S) {
  const PackOp O = ops[blockIdx.x];
  const int k = O.k, n = O.n, r = n - k, tid = threadIdx.x;
  const int* Ju = J + O.pu;
  const int* Jd = J + O.pd;
  for (int idx = tid; idx < k * r; idx += 256) {
    S[O.otu + idx] = O.tu[idx];
    S[O.otd + idx] = O.td[idx];
  }
  for (int idx = tid; idx < k * n; idx += 256) {
    const int i = idx % k, c = idx / k;  // R(i, Ju[c]) and L(Jd[c], i)
    const double ru = c < k ? (c == i ? 1.0 : 0.0) : O.tu[i + int64_t(c - k) * k];
    const double ld = c < k ? (c == i ? 1.0 : 0.0) : O.td[i + int64_t(c - k) * k];
    S[O.orr + i + int64_t(Ju[c]) * k] = ru;
    S[O.ol + Jd[c] + int64_t(i) * n] = ld;
  }
}

}  // namespace
}  // namespace hssc

using namespace hssc;

// ---------------------------------------------------------------------------------------
// Host object
// ---------------------------------------------------------------------------------------

struct hssc_hss1d {
  int64_t N = 0;
  double acc = 0;
  int nproxy = 0;
  int kind = 0;                 // 0 double layer, 1 Kernel_Eval on points, 2 dense matrix
  hssc_params prm{};            // kind 1
  std::vector<double> X, Y, NX, NY, W, KAP;
  std::vector<double> WB, WC;   // kind 1: Kernel_Eval's b and c at every point
  double *dwb = nullptr, *dwc = nullptr;
  const double* dK = nullptr;   // kind 2: the caller's matrix on the device (compression only)
  int64_t ldk = 0;
  // tree (1-based boxes, index 0 unused)
  std::vector<int> level, parent, c1, c2, start, count;
  int nboxes = 0, depth = 0;
  std::vector<std::vector<int>> nei;
  // compression
  std::vector<std::vector<int>> isk_up, isk_dn, J_up, J_dn;
  std::vector<int> nskel, kskel;
  // device scratch: all matrices (T, M_self, M_sib, inverse blocks) in one arena, all ints
  // (permutations) in another; offsets per box
  std::vector<int64_t> oTup, oTdn, oJup, oJdn, oMself, oMsib, oR, oL;
  int64_t arena_n = 0, iarena_n = 0;
  double* S = nullptr;
  int* I = nullptr;
  double *dx = nullptr, *dy = nullptr, *dnx = nullptr, *dny = nullptr, *dw = nullptr,
         *dkap = nullptr;
  // inverse: block offsets (laid out once, beyond the compressed operator in the arena
  // offsets [inv_lo, inv_hi)); the FP64 blocks live in S, the FP32 ones in W32
  bool inverted = false, inverted32 = false, laid_out = false;
  std::vector<int64_t> oE, oLinv, oRinv, oDinv;
  std::vector<BoxInv> inv_boxes;
  std::vector<TopOp> inv_tops;
  int64_t inv_lo = 0, inv_hi = 0;
  float* W32 = nullptr;
  // per-box vector offsets for the sweeps
  std::vector<int64_t> vUp, vDn;
  int64_t vlen = 0;
  // forest of independent problems: root box of each problem (level 0), the root of every
  // box, the circumradius of each problem's curve and each problem's top inverse block
  std::vector<int> roots, rootof;
  std::vector<double> Rroot;
  std::vector<int64_t> oFTinvs;
  // rank-predicted sketching of the IDs: per-box features, counters
  std::vector<std::array<float, 8>> feat;
  int64_t id_boxes = 0, id_fallbacks = 0, id_sketch_rows = 0, id_full_rows = 0;
  ~hssc_hss1d() {
    dfree1(S), dfree1(I), dfree1(dx), dfree1(dy), dfree1(dnx), dfree1(dny), dfree1(dw),
        dfree1(dkap), dfree1(W32), dfree1(dwb), dfree1(dwc);
  }
  bool leaf(int b) const { return c1[b] <= 0 && c2[b] <= 0; }
  bool root(int b) const { return parent[b] == 0; }
  Gen1D gen() const {
    Gen1D g{};
    g.kind = kind;
    g.C = CurveDev{dx, dy, dnx, dny, dw, dkap};
    if (kind == 1) g.ks = kernel_spec(prm);
    g.wb = dwb, g.wc = dwc, g.K = dK, g.ldk = ldk;
    return g;
  }
};

namespace hssc {
namespace {

// LOCAL_get_tree (HSS1D_compress_nsym.m:277-334) for every problem of the batch: the roots
// (boxes 1..P, problem p owning points [off_p, off_p + N_p)) are halved breadth first.
void build_tree1d(hssc_hss1d* h, int nbox_max, const std::vector<int64_t>& Ns) {
  auto push = [&](int lev, int par, int st, int cnt) {
    h->level.push_back(lev), h->parent.push_back(par), h->c1.push_back(-1),
        h->c2.push_back(-1), h->start.push_back(st), h->count.push_back(cnt);
  };
  h->level.clear(), h->parent.clear(), h->c1.clear(), h->c2.clear(), h->start.clear(),
      h->count.clear();
  push(0, 0, 0, 0);  // index 0 unused
  int64_t o = 0;
  h->roots.clear();
  for (int64_t n : Ns) {
    push(0, 0, int(o), int(n));
    h->roots.push_back(int(h->level.size()) - 1);
    o += n;
  }
  int first, last = 0, created = int(Ns.size()), lev = 0;
  while (created > 0) {
    first = last + 1;
    last = last + created;
    created = 0;
    for (int b = first; b <= last; ++b) {
      const int nb = h->count[b];
      if (nb > nbox_max) {
        const int nh = (nb + 1) / 2;
        const int s1 = last + created + 1, s2 = last + created + 2;
        h->c1[b] = s1, h->c2[b] = s2;
        push(lev + 1, b, h->start[b], nh);
        push(lev + 1, b, h->start[b] + nh, nb - nh);
        created += 2;
      }
    }
    ++lev;
  }
  h->nboxes = last;
  h->depth = 0;
  for (int b = 1; b <= last; ++b) h->depth = std::max(h->depth, h->level[b]);
  h->rootof.assign(last + 1, 0);
  for (int b = 1; b <= last; ++b) h->rootof[b] = h->parent[b] == 0 ? b : h->rootof[h->parent[b]];
}

// LOCAL_get_circum_circle (:339-359) on a list of point indices.
void circum(const hssc_hss1d* h, const std::vector<int>& idx, double* cx, double* cy, double* R) {
  const size_t n = idx.size();
  double xc = 0.5 * (h->X[idx[0]] + h->X[idx[n - 1]]);
  double yc = 0.5 * (h->Y[idx[0]] + h->Y[idx[n - 1]]);
  double m = 0;
  for (int i : idx) m = std::max(m, (h->X[i] - xc) * (h->X[i] - xc) + (h->Y[i] - yc) * (h->Y[i] - yc));
  double r = std::sqrt(m);
  const double ex = h->X[idx[0]] - h->X[idx[n - 1]], ey = h->Y[idx[0]] - h->Y[idx[n - 1]];
  if (1.2 * r * r > ex * ex + ey * ey) {
    xc = 0, yc = 0;
    for (int i : idx) xc += h->X[i], yc += h->Y[i];
    xc *= 1.0 / double(n), yc *= 1.0 / double(n);
    m = 0;
    for (int i : idx) m = std::max(m, (h->X[i] - xc) * (h->X[i] - xc) + (h->Y[i] - yc) * (h->Y[i] - yc));
    r = std::sqrt(m);
  }
  *cx = xc, *cy = yc, *R = r;
}

std::vector<int> range_of(const hssc_hss1d* h, int b) {
  std::vector<int> v(h->count[b]);
  for (int i = 0; i < h->count[b]; ++i) v[i] = h->start[b] + i;
  return v;
}

// LOCAL_construct_potential_neighborlist (:376-412)
void neighbours(hssc_hss1d* h) {
  const int nb = h->nboxes;
  h->nei.assign(nb + 1, {});
  for (int r : h->roots) h->nei[r] = {r};
  for (int b = 1; b <= nb; ++b) {
    if (h->root(b)) continue;
    double xc, yc, R;
    circum(h, range_of(h, b), &xc, &yc, &R);
    std::vector<int> pot;
    for (int j : h->nei[h->parent[b]]) {
      if (h->c1[j] > 0) pot.push_back(h->c1[j]);
      if (h->c2[j] > 0) pot.push_back(h->c2[j]);
    }
    for (int j : pot) {
      double mn = 1e300;
      for (int i = h->start[j]; i < h->start[j] + h->count[j]; ++i)
        mn = std::min(mn, (h->X[i] - xc) * (h->X[i] - xc) + (h->Y[i] - yc) * (h->Y[i] - yc));
      if (mn < kRadiusRel * kRadiusRel * R * R) h->nei[b].push_back(j);
    }
  }
  for (int b = 1; b <= nb; ++b) {
    if (h->root(b)) continue;
    auto& v = h->nei[b];
    v.erase(std::remove(v.begin(), v.end(), b), v.end());
    std::sort(v.begin(), v.end());
  }
}

// Appends a block to the double arena (host-side bookkeeping of offsets).
int64_t take(int64_t* top, int64_t n) {
  const int64_t o = *top;
  *top += std::max<int64_t>(n, 0);
  return o;
}

// Grows a device arena, keeping its contents.
template <typename T>
void grow(T** p, int64_t* cap, int64_t need) {
  if (need <= *cap) return;
  int64_t nc = std::max<int64_t>(need, *cap * 2);
  T* q = dmalloc<T>(nc);
  if (*p && *cap) H1CHK(hipMemcpy(q, *p, *cap * sizeof(T), hipMemcpyDeviceToDevice));
  dfree1(*p);
  *p = q;
  *cap = nc;
}

struct Ragged {
  std::vector<int64_t> off;  // matrix offsets (doubles)
  std::vector<int> m, n, mc;
  std::vector<int> ridx, cidx;
  std::vector<int64_t> ro, co, pofs;
  std::vector<double> px, py, pw;
  int64_t total = 0;
  // pw: the proxy points' own Kernel_Eval factors (b as targets, c as sources); empty = 1
  void add(const std::vector<int>& rows, const std::vector<double>& zx,
           const std::vector<double>& zy, const std::vector<int>& cols,
           const std::vector<double>& zw = {}) {
    off.push_back(total);
    const int mm = int(rows.size() + zx.size()), nn = int(cols.size());
    m.push_back(mm), n.push_back(nn), mc.push_back(int(rows.size()));
    ro.push_back(int64_t(ridx.size())), co.push_back(int64_t(cidx.size()));
    pofs.push_back(int64_t(px.size()));
    ridx.insert(ridx.end(), rows.begin(), rows.end());
    cidx.insert(cidx.end(), cols.begin(), cols.end());
    px.insert(px.end(), zx.begin(), zx.end());
    py.insert(py.end(), zy.begin(), zy.end());
    if (zw.empty()) pw.insert(pw.end(), zx.size(), 1.0);
    else pw.insert(pw.end(), zw.begin(), zw.end());
    total += int64_t(mm) * nn;
  }
};

// Generates ragged blocks on the device; returns the device matrix array (caller frees).
double* gen_ragged(const hssc_hss1d* h, const Ragged& R, int mode, std::vector<void*>* tmp) {
  auto keep = [&](void* p) { tmp->push_back(p); return p; };
  const int nm = int(R.off.size());
  int64_t* doff = static_cast<int64_t*>(keep(dput(R.off)));
  int* dm = static_cast<int*>(keep(dput(R.m)));
  int* dn = static_cast<int*>(keep(dput(R.n)));
  int* dmc = static_cast<int*>(keep(dput(R.mc)));
  int* dri = static_cast<int*>(keep(dput(R.ridx)));
  int* dci = static_cast<int*>(keep(dput(R.cidx)));
  int64_t* dro = static_cast<int64_t*>(keep(dput(R.ro)));
  int64_t* dco = static_cast<int64_t*>(keep(dput(R.co)));
  int64_t* dpo = static_cast<int64_t*>(keep(dput(R.pofs)));
  double* dpx = static_cast<double*>(keep(dput(R.px)));
  double* dpy = static_cast<double*>(keep(dput(R.py)));
  double* dpw = static_cast<double*>(keep(dput(R.pw)));
  double* A = dmalloc<double>(R.total);
  if (nm > 0) {
    int64_t mx = 1;
    for (int i = 0; i < nm; ++i) mx = std::max<int64_t>(mx, int64_t(R.m[i]) * R.n[i]);
    dim3 grid(unsigned(std::min<int64_t>((mx + 255) / 256, 64)), nm);
    HSSC_TIMED("gen1d_f64", 0, double(R.total), double(R.total) * 8,
               hipLaunchKernelGGL(gen1d_kernel, grid, dim3(256), 0, 0, h->gen(), mode, doff, dm,
                                  dn, dmc, dri, dro, dci, dco, dpx, dpy, dpw, dpo, A));
  }
  return A;
}

// Result of a batched ID: ranks and permutations on the host, each block's interpolation
// matrix T on the device (in buffers the result keeps alive).
struct IdResult {
  std::vector<int> k;
  std::vector<std::vector<int>> perm;
  std::vector<const double*> T;  // device: k x (n-k) column-major (leading dimension k)
  std::vector<std::shared_ptr<void>> bufs;  // device buffers holding T
};

std::shared_ptr<void> dev_owner(void* p) { return std::shared_ptr<void>(p, dfree1); }

IdResult run_id(const Ragged& R, double* A, double acc, const std::vector<int>* kfix,
                uint64_t seed) {
  const int nm = int(R.off.size());
  IdResult out;
  out.k.resize(nm), out.perm.resize(nm), out.T.resize(nm);
  if (nm == 0) return out;
  std::vector<int64_t> poff(nm), toff(nm), woff(nm);
  int64_t pt = 0, tt = 0, wt = 0;
  for (int i = 0; i < nm; ++i) {
    poff[i] = pt, pt += R.n[i];
    toff[i] = tt, tt += int64_t(R.n[i]) * R.n[i] / 4 + R.n[i];  // k (n - k) <= n^2 / 4
    woff[i] = wt, wt += 6 * int64_t(R.n[i]) + R.m[i];
  }
  std::vector<void*> tmp;
  double* Ac = dmalloc<double>(R.total);
  tmp.push_back(Ac);
  H1CHK(hipMemcpy(Ac, A, R.total * 8, hipMemcpyDeviceToDevice));
  int64_t* doff = dput(R.off);
  int* dm = dput(R.m);
  int* dn = dput(R.n);
  int64_t* dpo = dput(poff);
  int64_t* dto = dput(toff);
  int64_t* dwo = dput(woff);
  int* dperm = dmalloc<int>(pt);
  int* dk = dmalloc<int>(nm);
  double* dT = dmalloc<double>(tt);
  double* dwork = dmalloc<double>(wt);
  int* dkf = kfix ? dput(*kfix) : nullptr;
  tmp.insert(tmp.end(), {doff, dm, dn, dpo, dto, dwo, dperm, dk, dT, dwork, dkf});
  k::cpqr_id<double>(0, nm, doff, dm, dn, A, Ac, kfix ? -1.0 : acc, dkf, dperm, dpo, dk, dT, dto,
                     dwork, dwo, seed);
  H1CHK(hipDeviceSynchronize());
  H1CHK(hipGetLastError());
  std::vector<int> hk = dget(dk, nm);
  std::vector<int> hp = dget(dperm, pt);
  for (int i = 0; i < nm; ++i) {
    out.k[i] = hk[i];
    out.perm[i].assign(hp.begin() + poff[i], hp.begin() + poff[i] + R.n[i]);
    out.T[i] = dT + toff[i];
  }
  tmp.erase(std::find(tmp.begin(), tmp.end(), static_cast<void*>(dT)));
  out.bufs.push_back(dev_owner(dT));
  // restore A from the copy so a second pass can reuse it
  H1CHK(hipMemcpy(A, Ac, R.total * 8, hipMemcpyDeviceToDevice));
  for (void* p : tmp) dfree1(p);
  return out;
}

// ---------------------------------------------------------------------------------------
// Rank-predicted sketching (randomized ID with the predicted rank as the sample count)
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Y_b (s_b x n_b) = G_b A_b with G_b a random sign matrix (s_b x m_b).
__global__ void __launch_bounds__(256) sketch1d_kernel(
    const int64_t* __restrict__ off, const int* __restrict__ mr, const int* __restrict__ nc,
    const double* __restrict__ A, const int64_t* __restrict__ yoff, const int* __restrict__ sr,
    double* __restrict__ Y, uint64_t seed) {
  const int b = blockIdx.y;
  const int m = mr[b], n = nc[b], s = sr[b];
  const double* M = A + off[b];
  const uint64_t key = seed ^ (uint64_t(b) << 24);
  for (int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x; idx < int64_t(s) * n;
       idx += int64_t(gridDim.x) * 256) {
    const int i = int(idx % s), c = int(idx / s);
    double acc = 0;
    for (int r = 0; r < m; ++r) {
      const uint64_t h = mix64(key + 0x9E3779B97F4A7C15ull * (uint64_t(i) * uint64_t(m) + r + 1));
      acc += (h & 1ull) ? M[r + int64_t(c) * m] : -M[r + int64_t(c) * m];
    }
    Y[yoff[b] + idx] = acc;
  }
}

// Acceptance of a sketched ID on the full block A_b (m x n): with a random unit x over the
// residual columns and u = T x, e = |A_s u - A_r x|; with a random unit g over all columns,
// a = |A g|.  Accepted when the sketch was not saturated (k < s, or s covers min(m,n)) and
// e <= 10 acc a.  One workgroup per block; scratch: m doubles per block at wk + woff[b].
__global__ void __launch_bounds__(256) verify1d_kernel(
    const int64_t* __restrict__ off, const int* __restrict__ mr, const int* __restrict__ nc,
    const int* __restrict__ sr, const double* __restrict__ A, const int* __restrict__ perm,
    const int64_t* __restrict__ poff, const int* __restrict__ kk,
    const double* const* __restrict__ Tp, double* __restrict__ wk, const int64_t* __restrict__ woff,
    double acc, uint64_t seed, int* __restrict__ ok) {
  __shared__ double red[4];
  __shared__ double xs[256], us[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int m = mr[b], n = nc[b], s = sr[b], k = kk[b], r_ = n - k;
  const double* M = A + off[b];
  const int* P = perm + poff[b];
  const double* Tb = Tp[b];
  double* y = wk + woff[b];
  auto bsum = [&](double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
  };
  auto gauss = [&](uint64_t key, uint64_t i) {
    const double u1 = fmax(double(mix64(key + 0x9E3779B97F4A7C15ull * (2 * i + 1)) >> 11) *
                               (1.0 / 9007199254740992.0), 1e-300);
    const double u2 = double(mix64(key + 0x9E3779B97F4A7C15ull * (2 * i + 2)) >> 11) *
                      (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  };
  const uint64_t key = seed ^ (uint64_t(b) << 24) ^ 0x7777ull;
  // x over the residual columns (n <= 256 here), u = T x
  double v = 0;
  if (tid < r_) v = xs[tid] = gauss(key, tid);
  const double xn = sqrt(bsum(v * v));
  if (tid < r_) xs[tid] /= xn;
  __syncthreads();
  if (tid < k) {
    double t = 0;
    for (int c = 0; c < r_; ++c) t += Tb[tid + int64_t(c) * k] * xs[c];
    us[tid] = t;
  }
  __syncthreads();
  double e2 = 0;
  for (int r = tid; r < m; r += 256) {
    double t = 0;
    for (int i = 0; i < k; ++i) t += M[r + int64_t(P[i]) * m] * us[i];
    for (int c = 0; c < r_; ++c) t -= M[r + int64_t(P[k + c]) * m] * xs[c];
    e2 += t * t;
  }
  e2 = bsum(e2);
  // g over all columns, a = |A g|
  v = 0;
  if (tid < n) v = xs[tid] = gauss(key ^ 0x5555ull, tid);
  const double gn = sqrt(bsum(v * v));
  if (tid < n) xs[tid] /= gn;
  __syncthreads();
  double a2 = 0;
  for (int r = tid; r < m; r += 256) {
    double t = 0;
    for (int c = 0; c < n; ++c) t += M[r + int64_t(c) * m] * xs[c];
    a2 += t * t;
  }
  a2 = bsum(a2);
  if (tid == 0) {
    const int full = (m < n ? m : n);
    const bool saturated = (k >= s) && (s < full);
    ok[b] = (!saturated && sqrt(e2) <= 10.0 * acc * sqrt(a2)) ? 1 : 0;
    (void)y;
  }
}

// Randomized ID of every block of R with s[i] sketch rows: CPQR-ID (acc mode) of the sketch
// Y = G A, then the acceptance test on A.  ok[i] = 0 marks blocks to redo with the full ID.
IdResult run_id_sketched(const Ragged& R, double* A, double acc, const std::vector<int>& s,
                         uint64_t seed, std::vector<int>* ok,
                         const std::vector<int>* kfix = nullptr) {
  const int nm = int(R.off.size());
  IdResult out;
  ok->assign(nm, 0);
  if (nm == 0) return out;
  Ragged Y;
  int64_t ytot = 0;
  std::vector<int64_t> yoff(nm);
  for (int i = 0; i < nm; ++i) {
    yoff[i] = ytot;
    Y.off.push_back(ytot), Y.m.push_back(s[i]), Y.n.push_back(R.n[i]);
    ytot += int64_t(s[i]) * R.n[i];
  }
  Y.total = ytot;
  std::vector<void*> tmp;
  double* dY = dmalloc<double>(ytot);
  tmp.push_back(dY);
  int64_t* doff = dput(R.off);
  int* dm = dput(R.m);
  int* dn = dput(R.n);
  int* ds = dput(s);
  int64_t* dyo = dput(yoff);
  tmp.insert(tmp.end(), {doff, dm, dn, ds, dyo});
  int64_t mx = 1;
  for (int i = 0; i < nm; ++i) mx = std::max<int64_t>(mx, int64_t(s[i]) * R.n[i]);
  dim3 grid(unsigned(std::min<int64_t>((mx + 255) / 256, 64)), nm);
  HSSC_TIMED("sketch1d_f64", 0, 0.0, 0.0,
             hipLaunchKernelGGL(sketch1d_kernel, grid, dim3(256), 0, 0, doff, dm, dn, A, dyo, ds,
                                dY, seed));
  out = run_id(Y, dY, acc, kfix, seed ^ 0x3141ull);
  // acceptance on the full blocks
  std::vector<int64_t> poff(nm), woff(nm);
  std::vector<int> kk(nm), pflat;
  int64_t wt = 0;
  for (int i = 0; i < nm; ++i) {
    poff[i] = int64_t(pflat.size());
    pflat.insert(pflat.end(), out.perm[i].begin(), out.perm[i].end());
    kk[i] = out.k[i];
    woff[i] = wt, wt += R.m[i];
  }
  int* dp = dput(pflat);
  int64_t* dpo = dput(poff);
  int* dk = dput(kk);
  const double** dTp = dput(out.T);
  double* dw = dmalloc<double>(wt);
  int64_t* dwo = dput(woff);
  int* dok = dmalloc<int>(nm);
  tmp.insert(tmp.end(), {dp, dpo, dk, (void*)dTp, dw, dwo, dok});
  hipLaunchKernelGGL(verify1d_kernel, dim3(nm), dim3(256), 0, 0, doff, dm, dn, ds, A, dp, dpo,
                     dk, dTp, dw, dwo, acc, seed, dok);
  H1CHK(hipDeviceSynchronize());
  H1CHK(hipGetLastError());
  *ok = dget(dok, nm);
  for (void* p : tmp) dfree1(p);
  return out;
}

// A view of some blocks of R (same device storage).
Ragged subset(const Ragged& R, const std::vector<int>& idx) {
  Ragged S;
  for (int i : idx) S.off.push_back(R.off[i]), S.m.push_back(R.m[i]), S.n.push_back(R.n[i]);
  S.total = R.total;
  return S;
}

// ID of every block: blocks with s[i] > 0 try the sketched ID first; the others and the
// rejected sketches run the full CPQR ID.  Adds to the counters of h.
IdResult run_id_mixed(hssc_hss1d* h, const Ragged& R, double* A, double acc,
                      const std::vector<int>& s, uint64_t seed,
                      const std::vector<int>* kfix = nullptr) {
  const int nm = int(R.off.size());
  IdResult out;
  out.k.resize(nm), out.perm.resize(nm), out.T.resize(nm);
  std::vector<int> sk, full;
  for (int i = 0; i < nm; ++i) (s[i] > 0 ? sk : full).push_back(i);
  if (!sk.empty()) {
    std::vector<int> ss, ok, kf;
    for (int i : sk) ss.push_back(s[i]);
    if (kfix)
      for (int i : sk) kf.push_back((*kfix)[i]);
    IdResult r = run_id_sketched(subset(R, sk), A, acc, ss, seed, &ok, kfix ? &kf : nullptr);
    out.bufs.insert(out.bufs.end(), r.bufs.begin(), r.bufs.end());
    for (size_t j = 0; j < sk.size(); ++j) {
      h->id_sketch_rows += ss[j];
      if (ok[j]) {
        const int i = sk[j];
        out.k[i] = r.k[j], out.perm[i] = std::move(r.perm[j]), out.T[i] = r.T[j];
      } else {
        full.push_back(sk[j]);
        ++h->id_fallbacks;
      }
    }
  }
  if (!full.empty()) {
    std::sort(full.begin(), full.end());
    std::vector<int> kf;
    if (kfix)
      for (int i : full) kf.push_back((*kfix)[i]);
    IdResult r = run_id(subset(R, full), A, acc, kfix ? &kf : nullptr, seed);
    for (size_t j = 0; j < full.size(); ++j) {
      const int i = full[j];
      h->id_full_rows += R.m[i];
      out.k[i] = r.k[j], out.perm[i] = std::move(r.perm[j]), out.T[i] = r.T[j];
    }
    out.bufs.insert(out.bufs.end(), r.bufs.begin(), r.bufs.end());
  }
  if (!kfix) h->id_boxes += nm;
  return out;
}

// Sketch rows of a box from its predicted rank: the rank plus oversampling, rounded up to
// the 16-row MFMA granule, capped by the block's size.
int sketch_rows(double pred, int oversample, int m, int n) {
  const int want = int(std::ceil(std::max(pred, 1.0))) + oversample;
  const int bucket = ((want + 15) / 16) * 16;
  return std::min(bucket, std::min(m, n));
}


}  // namespace

// HSS2D/predict_rank.m:5-23: rank ~ m_p log10(n) - b_p (r0_p below n = 150) for
// p = ceil(-log10 acc) clamped to [7, 12].
double predict_rank_table(int n, double acc) {
  static const int mt[6] = {96, 138, 184, 275, 526, 487};
  static const int bt[6] = {138, 215, 300, 496, 1074, 954};
  static const int r0[6] = {86, 104, 116, 130, 137, 140};
  int p = int(std::ceil(-std::log10(acc)));
  p = std::min(12, std::max(7, p));
  const int i = p - 7;
  if (n < 150) return std::min(n, r0[i]);
  return std::min<double>(n, std::ceil(mt[i] * std::log10(double(n)) - bt[i]));
}
}  // namespace hssc

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

namespace {
// What an HSS1D compression reads (hssc_hss1d.kind): curves with normals, weights and
// curvature (0), points with Kernel_Eval's params (1), or a dense matrix (2).
struct Source1D {
  int kind = 0;
  const double *Xc = nullptr, *nrm = nullptr, *w = nullptr, *kappa = nullptr;  // 0; Xc for 1
  const hssc_params* p = nullptr;                                              // 1
  const double* K = nullptr;                                                   // 2 (host)
  int64_t ldk = 0;
};

int hss1d_compress_impl(const Source1D& src, const std::vector<int64_t>& Ns, int32_t nbox_max,
                        double acc, int32_t nproxy, int rank_mode, hssc_mlp* mlp, int oversample,
                        hssc_hss1d** out) {
  int64_t N = 0;
  for (int64_t n : Ns) {
    if (n < 4) return fail(HSSC_ERR_INVALID_ARG, "every curve needs at least 4 nodes");
    N += n;
  }
  const int kind = src.kind;
  const bool geo = kind != 2;  // points, neighbour lists and proxy circles
  if (!out || Ns.empty() || nbox_max < 2 || !(acc > 0 && acc < 1) || (geo && nproxy < 4) ||
      (kind == 0 && (!src.Xc || !src.nrm || !src.w || !src.kappa)) ||
      (kind == 1 && (!src.Xc || !src.p)) || (kind == 2 && (!src.K || src.ldk < N)) ||
      kind < 0 || kind > 2)
    return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d_compress arguments");
  if (kind == 1) {
    if (int st = check_params(*src.p)) return st;
    if (src.p->flag_pot != HSSC_POT_SL_L_2D)
      return fail(HSSC_ERR_UNSUPPORTED, "HSS1D operators are real: Kernel_Eval needs SL_L_2D");
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(HSSC_ERR_HIP, "no HIP device is available");
  hssc_hss1d* h = nullptr;
  double* dKown = nullptr;
  struct KFree {
    double** p;
    ~KFree() { dfree1(*p); }
  } kfree{&dKown};
  try {
    h = new hssc_hss1d();
    h->N = N, h->acc = acc, h->nproxy = geo ? nproxy : 0, h->kind = kind;
    if (kind == 0) {
      const double *Xc = src.Xc, *nrm = src.nrm, *w = src.w, *kappa = src.kappa;
      h->X.assign(Xc, Xc + N), h->Y.assign(Xc + N, Xc + 2 * N);
      h->NX.assign(nrm, nrm + N), h->NY.assign(nrm + N, nrm + 2 * N);
      h->W.assign(w, w + N), h->KAP.assign(kappa, kappa + N);
    } else {
      if (kind == 1) h->X.assign(src.Xc, src.Xc + N), h->Y.assign(src.Xc + N, src.Xc + 2 * N);
      else h->X.assign(N, 0.0), h->Y.assign(N, 0.0);
      h->NX.assign(N, 0.0), h->NY.assign(N, 0.0), h->W.assign(N, 0.0), h->KAP.assign(N, 0.0);
    }
    h->dx = dput(h->X), h->dy = dput(h->Y), h->dnx = dput(h->NX), h->dny = dput(h->NY);
    h->dw = dput(h->W), h->dkap = dput(h->KAP);
    if (kind == 1) {
      h->prm = *src.p;
      h->WB.resize(N), h->WC.resize(N);
      for (int64_t i = 0; i < N; ++i) {
        h->WB[i] = point_weight(h->prm, h->X[i], h->Y[i], true);
        h->WC[i] = point_weight(h->prm, h->X[i], h->Y[i], false);
      }
      h->dwb = dput(h->WB), h->dwc = dput(h->WC);
    }
    if (kind == 2) {
      // the caller's matrix on the device for the compression (leading dimension N)
      dKown = dmalloc<double>(size_t(N) * N);
      H1CHK(hipMemcpy2D(dKown, N * 8, src.K, src.ldk * 8, N * 8, N, hipMemcpyHostToDevice));
      h->dK = dKown, h->ldk = N;
    }
    build_tree1d(h, nbox_max, Ns);
    for (int r : h->roots)
      if (h->leaf(r)) throw HipErr{fail(HSSC_ERR_INVALID_ARG, "tree needs at least 2 levels")};
    if (geo) neighbours(h);
    const int nb = h->nboxes;
    h->Rroot.assign(nb + 1, 1.0);
    for (int r : h->roots) {
      double cx, cy;
      if (geo) circum(h, range_of(h, r), &cx, &cy, &h->Rroot[r]);
    }
    // 'brute': per level, the skeletons of its boxes in the order the reference compresses
    // them (nboxes down to 2), I_src{lev,1:2} of HSS1D_compress_nsym.m:210-214
    std::vector<std::vector<int>> isrc_up(h->depth + 2), isrc_dn(h->depth + 2);
    h->feat.assign(nb + 1, {});
    h->isk_up.assign(nb + 1, {}), h->isk_dn.assign(nb + 1, {});
    h->J_up.assign(nb + 1, {}), h->J_dn.assign(nb + 1, {});
    h->nskel.assign(nb + 1, 0), h->kskel.assign(nb + 1, 0);
    h->oTup.assign(nb + 1, -1), h->oTdn.assign(nb + 1, -1), h->oJup.assign(nb + 1, -1),
        h->oJdn.assign(nb + 1, -1), h->oMself.assign(nb + 1, -1), h->oMsib.assign(nb + 1, -1);
    h->oR.assign(nb + 1, -1), h->oL.assign(nb + 1, -1);
    std::vector<int> iarena;            // permutations of every box (host, uploaded at the end)
    std::vector<double*> level_blocks;  // per level: T, R, L blocks packed on the device
    std::vector<int64_t> level_sizes;
    std::vector<int> box_block_level(nb + 1, -1);
    struct BlocksFree {
      std::vector<double*>* v;
      ~BlocksFree() {
        for (double* p : *v) dfree1(p);
      }
    } blocks_free{&level_blocks};
    // compress level by level, deepest first (HSS1D_compress_nsym.m:84-216)
    for (int lev = h->depth; lev >= 1; --lev) {
      std::vector<int> boxes;
      for (int b = 1; b <= nb; ++b)
        if (h->level[b] == lev) boxes.push_back(b);
      Ragged Rout, Rin;
      const int nq0 = int(boxes.size());
      std::vector<std::vector<int>> outs(nq0), inns(nq0), oins(nq0), iins(nq0);
      std::vector<std::vector<double>> zxs(nq0), zys(nq0);
      std::vector<double> Rbox(nq0);
      // per box (independent, threaded): the skeleton candidates, the proxy circle and the
      // near points inside it, and the box's geometric features
      par_for(nq0, [&](int q) {
        const int b = boxes[q];
        std::vector<int> out, inn;
        if (h->leaf(b)) {
          out = range_of(h, b);
          inn = out;
        } else {
          out = h->isk_up[h->c1[b]];
          out.insert(out.end(), h->isk_up[h->c2[b]].begin(), h->isk_up[h->c2[b]].end());
          inn = h->isk_dn[h->c1[b]];
          inn.insert(inn.end(), h->isk_dn[h->c2[b]].begin(), h->isk_dn[h->c2[b]].end());
        }
        if (!geo) {
          // HSS1D_compress_nsym.m:158-181: A_out = K(ind_offd_in, indskel_out),
          // A_in = K(indskel_in, ind_offd_out); a leaf's rows are every index outside it, a
          // parent's the level below's skeletons outside it
          const int lo = h->start[b], hi = h->start[b] + h->count[b];
          std::vector<int>& oin = oins[q];
          std::vector<int>& iin = iins[q];
          if (h->leaf(b)) {
            for (int i = 0; i < int(N); ++i)
              if (i < lo || i >= hi) oin.push_back(i);
            iin = oin;
          } else {
            for (int i : isrc_up[lev + 1])
              if (i < lo || i >= hi) oin.push_back(i);
            for (int i : isrc_dn[lev + 1])
              if (i < lo || i >= hi) iin.push_back(i);
          }
          auto& f = h->feat[b];
          f = {};
          f[0] = float(lev);
          f[1] = float(std::log2(double(h->count[b])));
          f[2] = float(std::log2(double(out.size())));
          f[6] = float(std::log2(double(std::max<size_t>(iin.size(), 1))));
          f[7] = float(-std::log10(acc));
          outs[q].swap(out), inns[q].swap(inn);
          Rbox[q] = 0;
          return;
        }
        std::vector<int> uni = out;
        uni.insert(uni.end(), inn.begin(), inn.end());
        std::sort(uni.begin(), uni.end());
        double xc, yc, R;
        circum(h, uni, &xc, &yc, &R);
        const double rp = kRadiusRel * R;
        std::vector<double>& zx = zxs[q];
        std::vector<double>& zy = zys[q];
        zx.resize(nproxy), zy.resize(nproxy);
        for (int t = 0; t < nproxy; ++t) {
          // linspace(0, 2pi(1-1/n), n)
          const double th = (nproxy > 1) ? (2 * kPi * (1 - 1.0 / nproxy)) * t / (nproxy - 1) : 0;
          zx[t] = xc + rp * std::cos(th), zy[t] = yc + rp * std::sin(th);
        }
        std::vector<int> omay, imay;
        for (int j : h->nei[b]) {
          if (h->leaf(j)) {
            auto r = range_of(h, j);
            omay.insert(omay.end(), r.begin(), r.end());
            imay.insert(imay.end(), r.begin(), r.end());
          } else {
            for (int sb : {h->c1[j], h->c2[j]}) {
              omay.insert(omay.end(), h->isk_up[sb].begin(), h->isk_up[sb].end());
              imay.insert(imay.end(), h->isk_dn[sb].begin(), h->isk_dn[sb].end());
            }
          }
        }
        std::vector<int>& oin = oins[q];
        std::vector<int>& iin = iins[q];
        for (int i : omay)
          if ((h->X[i] - xc) * (h->X[i] - xc) + (h->Y[i] - yc) * (h->Y[i] - yc) < rp * rp)
            oin.push_back(i);
        for (int i : imay)
          if ((h->X[i] - xc) * (h->X[i] - xc) + (h->Y[i] - yc) * (h->Y[i] - yc) < rp * rp)
            iin.push_back(i);
        // geometric features of the box (inputs of the rank predictor)
        double kmax = 0, ksum = 0;
        for (int i = h->start[b]; i < h->start[b] + h->count[b]; ++i)
          kmax = std::max(kmax, std::fabs(h->KAP[i])), ksum += std::fabs(h->KAP[i]);
        auto& f = h->feat[b];
        f[0] = float(lev);
        f[1] = float(std::log2(double(h->count[b])));
        f[2] = float(std::log2(double(out.size())));
        f[3] = float(std::log2(R / h->Rroot[h->rootof[b]]));
        f[4] = float(kmax * R);
        f[5] = float(ksum / h->count[b] * R);
        f[6] = float(std::log2(double(iin.size() + size_t(nproxy))));  // rows of A_out
        f[7] = float(-std::log10(acc));
        outs[q].swap(out), inns[q].swap(inn);
        Rbox[q] = R;
      });
      for (int q = 0; q < nq0; ++q) {
        // Kernel_Eval factors of the proxy points: b as targets of A_out, c as sources of A_in
        std::vector<double> wbz, wcz;
        if (kind == 1)
          for (size_t t = 0; t < zxs[q].size(); ++t) {
            wbz.push_back(point_weight(h->prm, zxs[q][t], zys[q][t], true));
            wcz.push_back(point_weight(h->prm, zxs[q][t], zys[q][t], false));
          }
        Rout.add(iins[q], zxs[q], zys[q], outs[q], wbz);  // A_out = [K(X_in_inside, X_out); K(Zp, X_out)]
        Rin.add(oins[q], zxs[q], zys[q], inns[q], wcz);   // A_in^T = [K(X_in, X_out_inside), K(X_in, Zp)]^T
      }
      std::vector<void*> tmp;
      double* Aout = gen_ragged(h, Rout, 0, &tmp);
      double* Ain = gen_ragged(h, Rin, 1, &tmp);
      tmp.push_back(Aout), tmp.push_back(Ain);
      const uint64_t seed = 0x1234ull + uint64_t(lev) * 7919ull;
      const int nq = int(boxes.size());
      // sketch rows per box from the predicted rank (0 = full ID)
      std::vector<int> s_out(nq, 0), s_in(nq, 0);
      if (rank_mode != 0) {
        std::vector<double> pred(nq);
        if (rank_mode == 1) {
          for (int q = 0; q < nq; ++q) pred[q] = predict_rank_table(int(outs[q].size()), acc);
        } else {
          std::vector<float> Xf(size_t(nq) * 8), y(nq);
          for (int q = 0; q < nq; ++q)
            for (int c = 0; c < 8; ++c) Xf[q + size_t(c) * nq] = h->feat[boxes[q]][c];
          if (int st = hssc_mlp_predict(mlp, Xf.data(), nq, y.data())) throw HipErr{st};
          for (int q = 0; q < nq; ++q) pred[q] = std::exp2(double(y[q]));
        }
        for (int q = 0; q < nq; ++q) {
          if (Rout.n[q] <= 256) {
            s_out[q] = sketch_rows(pred[q], oversample, Rout.m[q], Rout.n[q]);
            if (s_out[q] >= std::min(Rout.m[q], Rout.n[q])) s_out[q] = 0;
          }
          if (Rin.n[q] <= 256) {
            s_in[q] = sketch_rows(pred[q], oversample, Rin.m[q], Rin.n[q]);
            if (s_in[q] >= std::min(Rin.m[q], Rin.n[q])) s_in[q] = 0;
          }
        }
      }
      IdResult io = run_id_mixed(h, Rout, Aout, acc, s_out, seed);
      IdResult ii = run_id_mixed(h, Rin, Ain, acc, s_in, seed ^ 0xABCDull);
      // rank equalisation (HSS1D_compress_nsym.m:190-196): the smaller side is redone with
      // the fixed rank k = max(k_out, k_in)
      std::vector<int> kf(nq);
      std::vector<int> redo_out, redo_in;
      for (int q = 0; q < nq; ++q) {
        kf[q] = std::max(io.k[q], ii.k[q]);
        if (kf[q] != io.k[q]) redo_out.push_back(q);
        if (kf[q] != ii.k[q]) redo_in.push_back(q);
      }
      auto redo = [&](const Ragged& Rr, double* Ar, const std::vector<int>& sq,
                      const std::vector<int>& idx, IdResult* res, uint64_t sd) {
        if (idx.empty()) return;
        std::vector<int> ks, ss;
        for (int q : idx) {
          ks.push_back(kf[q]);
          int sr = 0;
          if (sq[q] > 0) {
            sr = sketch_rows(double(kf[q]), oversample, Rr.m[q], Rr.n[q]);
            if (sr >= std::min(Rr.m[q], Rr.n[q]) || sr <= kf[q]) sr = 0;
          }
          ss.push_back(sr);
        }
        IdResult f = run_id_mixed(h, subset(Rr, idx), Ar, acc, ss, sd, &ks);
        res->bufs.insert(res->bufs.end(), f.bufs.begin(), f.bufs.end());
        for (size_t j = 0; j < idx.size(); ++j) {
          const int q = idx[j];
          res->k[q] = f.k[j], res->perm[q] = std::move(f.perm[j]), res->T[q] = f.T[j];
        }
      };
      redo(Rout, Aout, s_out, redo_out, &io, seed);
      redo(Rin, Ain, s_in, redo_in, &ii, seed ^ 0xABCDull);
      for (void* p : tmp) dfree1(p);
      // per box: skeleton index sets and permutations (host); T blocks and the explicit R, L
      // packed on the device into this level's block arena
      std::vector<int64_t> aoff(nq + 1, 0), ioff(nq + 1, int64_t(iarena.size()));
      for (int q = 0; q < nq; ++q) {
        const int64_t n = int64_t(outs[q].size()), k = kf[q];
        aoff[q + 1] = aoff[q] + 2 * k * (n - k) + 2 * k * n;
        ioff[q + 1] = ioff[q] + int64_t(io.perm[q].size()) + int64_t(ii.perm[q].size());
      }
      iarena.resize(ioff[nq]);
      std::vector<PackOp> pk(nq);
      const int lvl = int(level_blocks.size());
      par_for(nq, [&](int q) {
        const int b = boxes[q], k = kf[q];
        const int n = int(outs[q].size());
        h->nskel[b] = n;
        h->kskel[b] = k;
        h->J_up[b] = io.perm[q];
        h->J_dn[b] = ii.perm[q];
        h->isk_up[b].resize(k), h->isk_dn[b].resize(k);
        for (int i = 0; i < k; ++i)
          h->isk_up[b][i] = outs[q][io.perm[q][i]], h->isk_dn[b][i] = inns[q][ii.perm[q][i]];
        int64_t io_ = ioff[q];
        h->oJup[b] = io_;
        std::copy(io.perm[q].begin(), io.perm[q].end(), iarena.begin() + io_);
        io_ += int64_t(io.perm[q].size());
        h->oJdn[b] = io_;
        std::copy(ii.perm[q].begin(), ii.perm[q].end(), iarena.begin() + io_);
        PackOp& P = pk[q];
        P.tu = io.T[q], P.td = ii.T[q], P.k = k, P.n = n;
        P.otu = aoff[q], P.otd = P.otu + int64_t(k) * (n - k);
        P.orr = P.otd + int64_t(k) * (n - k), P.ol = P.orr + int64_t(k) * n;
        P.pu = h->oJup[b], P.pd = h->oJdn[b];
        // level-local offsets, shifted to the arena once every level is laid out
        h->oTup[b] = P.otu, h->oTdn[b] = P.otd, h->oR[b] = P.orr, h->oL[b] = P.ol;
        box_block_level[b] = lvl;
      });
      {
        // this level's permutations are the tail of iarena: upload them with the ops
        int* dJ = dput(std::vector<int>(iarena.begin(), iarena.end()));
        double* blk = dmalloc<double>(std::max<int64_t>(aoff[nq], 1));
        PackOp* dpk = dput(pk);
        if (nq > 0)
          hipLaunchKernelGGL(hss1d_pack_kernel, dim3(nq), dim3(256), 0, 0, dpk, dJ, blk);
        H1CHK(hipDeviceSynchronize());
        H1CHK(hipGetLastError());
        dfree1(dJ), dfree1(dpk);
        level_blocks.push_back(blk);
        level_sizes.push_back(aoff[nq]);
      }
      if (!geo)
        for (int q = nq - 1; q >= 0; --q) {
          const int b = boxes[q];
          isrc_up[lev].insert(isrc_up[lev].end(), h->isk_up[b].begin(), h->isk_up[b].end());
          isrc_dn[lev].insert(isrc_dn[lev].end(), h->isk_dn[b].begin(), h->isk_dn[b].end());
        }
    }
    // M_self at leaves, M_sib at children of every parent (HSS1D_compress_nsym.m:220-271)
    Ragged Rm;
    std::vector<std::pair<int, int>> which;  // (box, 0 = self / 1 = sib)
    for (int b = nb; b >= 1; --b) {
      if (h->leaf(b)) {
        auto r = range_of(h, b);
        Rm.add(r, {}, {}, r);
        which.push_back({b, 0});
      } else {
        Rm.add(h->isk_dn[h->c1[b]], {}, {}, h->isk_up[h->c2[b]]);
        which.push_back({h->c1[b], 1});
        Rm.add(h->isk_dn[h->c2[b]], {}, {}, h->isk_up[h->c1[b]]);
        which.push_back({h->c2[b], 1});
      }
    }
    std::vector<void*> tmp;
    double* Mblk = gen_ragged(h, Rm, 0, &tmp);
    tmp.push_back(Mblk);
    // arena: the levels' packed blocks, then the M blocks
    std::vector<int64_t> lbase(level_blocks.size());
    int64_t base = 0;
    for (size_t l = 0; l < level_blocks.size(); ++l) lbase[l] = base, base += level_sizes[l];
    for (int b = 1; b <= nb; ++b) {
      const int l = box_block_level[b];
      if (l < 0) continue;
      h->oTup[b] += lbase[l], h->oTdn[b] += lbase[l], h->oR[b] += lbase[l], h->oL[b] += lbase[l];
    }
    for (size_t q = 0; q < which.size(); ++q) {
      const int b = which[q].first;
      (which[q].second == 0 ? h->oMself[b] : h->oMsib[b]) = base + Rm.off[q];
    }
    h->arena_n = base + Rm.total;
    h->S = dmalloc<double>(h->arena_n);
    for (size_t l = 0; l < level_blocks.size(); ++l)
      if (level_sizes[l] > 0)
        H1CHK(hipMemcpy(h->S + lbase[l], level_blocks[l], level_sizes[l] * 8,
                        hipMemcpyDeviceToDevice));
    H1CHK(hipMemcpy(h->S + base, Mblk, Rm.total * 8, hipMemcpyDeviceToDevice));
    h->iarena_n = int64_t(iarena.size());
    h->I = dput(iarena);
    h->vUp.assign(nb + 1, 0), h->vDn.assign(nb + 1, 0);
    int64_t vv = 0;
    for (int b = 1; b <= nb; ++b)
      if (!h->root(b)) h->vUp[b] = vv, vv += h->kskel[b];
    for (int b = 1; b <= nb; ++b)
      if (!h->root(b)) h->vDn[b] = vv, vv += h->kskel[b];
    h->vlen = vv;
    H1CHK(hipDeviceSynchronize());
    for (void* p : tmp) dfree1(p);
    h->dK = nullptr;  // the matrix copy is released on return
    *out = h;
    return HSSC_OK;
  } catch (const HipErr& e) {
    delete h;
    return e.status;
  } catch (const std::bad_alloc&) {
    delete h;
    return fail(HSSC_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}
}  // namespace

namespace {
Source1D curves(const double* Xc, const double* nrm, const double* w, const double* kappa) {
  Source1D s;
  s.kind = 0, s.Xc = Xc, s.nrm = nrm, s.w = w, s.kappa = kappa;
  return s;
}
}  // namespace

extern "C" {

int hssc_hss1d_compress(const double* Xc, const double* nrm, const double* w, const double* kappa,
                        int64_t N, int32_t nbox_max, double acc, int32_t nproxy,
                        hssc_hss1d** out) {
  return hss1d_compress_impl(curves(Xc, nrm, w, kappa), {N}, nbox_max, acc, nproxy, 0, nullptr,
                             0, out);
}

int hssc_hss1d_compress_kernel(const hssc_params* p, const double* X, int64_t N,
                               int32_t nbox_max, double acc, int32_t nproxy, hssc_hss1d** out) {
  Source1D s;
  s.kind = 1, s.Xc = X, s.p = p;
  return hss1d_compress_impl(s, {N}, nbox_max, acc, nproxy, 0, nullptr, 0, out);
}

int hssc_hss1d_compress_dense(const double* K, int64_t N, int64_t ldk, int32_t nbox_max,
                              double acc, hssc_hss1d** out) {
  Source1D s;
  s.kind = 2, s.K = K, s.ldk = ldk;
  return hss1d_compress_impl(s, {N}, nbox_max, acc, 0, 0, nullptr, 0, out);
}

int hssc_hss1d_compress_ranked(const double* Xc, const double* nrm, const double* w,
                               const double* kappa, int64_t N, int32_t nbox_max, double acc,
                               int32_t nproxy, int32_t rank_mode, hssc_mlp* mlp,
                               int32_t oversample, hssc_hss1d** out) {
  if (rank_mode < 0 || rank_mode > 2 || (rank_mode == 2 && !mlp) || oversample < 0)
    return fail(HSSC_ERR_INVALID_ARG, "invalid rank prediction arguments");
  return hss1d_compress_impl(curves(Xc, nrm, w, kappa), {N}, nbox_max, acc, nproxy, rank_mode,
                             mlp, oversample, out);
}

int hssc_hss1d_compress_batched(int32_t nprob, const int64_t* Ns, const double* Xc,
                                const double* nrm, const double* w, const double* kappa,
                                int32_t nbox_max, double acc, int32_t nproxy, int32_t rank_mode,
                                hssc_mlp* mlp, int32_t oversample, hssc_hss1d** out) {
  if (nprob < 1 || !Ns || rank_mode < 0 || rank_mode > 2 || (rank_mode == 2 && !mlp) ||
      oversample < 0)
    return fail(HSSC_ERR_INVALID_ARG, "invalid batched hss1d_compress arguments");
  return hss1d_compress_impl(curves(Xc, nrm, w, kappa), std::vector<int64_t>(Ns, Ns + nprob),
                             nbox_max, acc, nproxy, rank_mode, mlp, oversample, out);
}

int hssc_hss1d_box_features(const hssc_hss1d* h, int32_t box, float* f) {
  if (!h || !f || box < 1 || box > h->nboxes || h->root(box))
    return fail(HSSC_ERR_INVALID_ARG, "bad box (roots have no features)");
  for (int c = 0; c < 8; ++c) f[c] = h->feat[box][c];
  return HSSC_OK;
}

int hssc_hss1d_id_stats(const hssc_hss1d* h, int64_t* boxes, int64_t* fallbacks,
                        int64_t* sketch_rows, int64_t* full_rows) {
  if (!h) return fail(HSSC_ERR_INVALID_ARG, "null handle");
  if (boxes) *boxes = h->id_boxes;
  if (fallbacks) *fallbacks = h->id_fallbacks;
  if (sketch_rows) *sketch_rows = h->id_sketch_rows;
  if (full_rows) *full_rows = h->id_full_rows;
  return HSSC_OK;
}

}  // extern "C"

namespace {

// Offsets of every inverse block (HSS1D_invert_nsym.m:44-91), computed once per operator.
void inverse_layout(hssc_hss1d* h) {
  if (h->laid_out) return;
  const int nb = h->nboxes;
  for (int b = 1; b <= nb; ++b)
    if (!h->root(b) && h->nskel[b] > 256)
      throw HipErr{fail(HSSC_ERR_RANK_OVERFLOW, "HSS1D block larger than 256")};
  int64_t top = h->arena_n;
  h->inv_lo = top;
  h->oE.assign(nb + 1, -1), h->oLinv.assign(nb + 1, -1), h->oRinv.assign(nb + 1, -1),
      h->oDinv.assign(nb + 1, -1);
  std::vector<BoxInv> all(nb + 1);
  for (int b = 1; b <= nb; ++b) {
    if (h->root(b)) continue;
    const int n = h->nskel[b], k = h->kskel[b];
    BoxInv B{};
    B.n = n, B.k = k;
    B.F = take(&top, int64_t(n) * n);
    B.Jup = h->oJup[b], B.Jdn = h->oJdn[b], B.Tup = h->oTup[b], B.Tdn = h->oTdn[b];
    B.E = h->oE[b] = take(&top, int64_t(k) * k);
    B.RF = take(&top, int64_t(k) * n);
    B.FL = take(&top, int64_t(n) * k);
    B.Linv = h->oLinv[b] = take(&top, int64_t(n) * k);
    B.Rinv = h->oRinv[b] = take(&top, int64_t(k) * n);
    B.Dinv = h->oDinv[b] = take(&top, int64_t(n) * n);
    B.leaf = h->leaf(b) ? 1 : 0;
    B.Mself = h->oMself[b];
    all[b] = B;
  }
  for (int b = 1; b <= nb; ++b)
    if (!h->root(b) && !h->leaf(b)) {
      all[b].E1 = h->oE[h->c1[b]], all[b].E2 = h->oE[h->c2[b]];
      all[b].S1 = h->oMsib[h->c1[b]], all[b].S2 = h->oMsib[h->c2[b]];
      all[b].k1 = h->kskel[h->c1[b]], all[b].k2 = h->kskel[h->c2[b]];
    }
  std::vector<TopOp> tops;
  h->oFTinvs.assign(nb + 1, -1);
  for (int r : h->roots) {
    const int a = h->c1[r], c = h->c2[r];
    const int k2 = h->kskel[a], k3 = h->kskel[c];
    if (k2 + k3 > 256) throw HipErr{fail(HSSC_ERR_RANK_OVERFLOW, "HSS1D top block larger than 256")};
    TopOp O{};
    O.FT = h->oFTinvs[r] = take(&top, int64_t(k2 + k3) * (k2 + k3));
    O.k2 = k2, O.k3 = k3;
    O.E2 = h->oE[a], O.E3 = h->oE[c], O.S2 = h->oMsib[a], O.S3 = h->oMsib[c];
    tops.push_back(O);
  }
  h->inv_boxes.swap(all);
  h->inv_tops.swap(tops);
  h->inv_hi = top;
  h->laid_out = true;
}

// The inverse in precision T into W (arena offsets minus wofs): one launch per level
// (deepest first), then the top blocks.
template <typename T>
void invert_run(hssc_hss1d* h, T* W, int64_t wofs) {
  const int nb = h->nboxes;
  for (int lev = h->depth; lev >= 1; --lev) {
    std::vector<BoxInv> ops;
    for (int b = 1; b <= nb; ++b)
      if (h->level[b] == lev) ops.push_back(h->inv_boxes[b]);
    if (ops.empty()) continue;
    BoxInv* d = dput(ops);
    HSSC_TIMED(sizeof(T) == 8 ? "hss1d_invert_f64" : "hss1d_invert_f32", 0, 0.0, 0.0,
               hipLaunchKernelGGL(hss1d_invert_kernel<T>, dim3(ops.size()), dim3(256), 0, 0, d,
                                  h->S, W, wofs, h->I));
    H1CHK(hipDeviceSynchronize());
    dfree1(d);
  }
  TopOp* dtops = dput(h->inv_tops);
  hipLaunchKernelGGL(hss1d_top_kernel<T>, dim3(h->inv_tops.size()), dim3(256), 0, 0, dtops, h->S,
                     W, wofs);
  H1CHK(hipDeviceSynchronize());
  dfree1(dtops);
  H1CHK(hipGetLastError());
}

}  // namespace

extern "C" {

int hssc_hss1d_invert_prec(hssc_hss1d* h, int32_t precision) {
  if (!h) return fail(HSSC_ERR_INVALID_ARG, "null handle");
  if (precision != HSSC_PREC_FP64 && precision != HSSC_PREC_FP32)
    return fail(HSSC_ERR_INVALID_ARG, "precision must be HSSC_PREC_FP64 or HSSC_PREC_FP32");
  try {
    inverse_layout(h);
    if (precision == HSSC_PREC_FP64) {
      grow(&h->S, &h->arena_n, h->inv_hi);
      invert_run<double>(h, h->S, 0);
      h->inverted = true;
    } else {
      if (!h->W32) h->W32 = dmalloc<float>(h->inv_hi - h->inv_lo);
      invert_run<float>(h, h->W32, h->inv_lo);
      h->inverted32 = true;
    }
    return HSSC_OK;
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_invert(hssc_hss1d* h) { return hssc_hss1d_invert_prec(h, HSSC_PREC_FP64); }

static SweepTerm term(int64_t mat, int ld, int cols, int64_t in, int len1, int64_t in2,
                      int rofs = 0, int src_global = 0) {
  SweepTerm T{};
  T.mat = mat, T.ld = ld, T.cols = cols, T.in = in, T.len1 = len1, T.in2 = in2, T.rofs = rofs;
  T.src_global = src_global;
  return T;
}

static int sweep1d(const hssc_hss1d* h, const double* dIn, double* dOut, bool inverse,
                   bool fp32 = false, int64_t m = 1) {
  // inverse: HSS1D/HSS1D_applyinv_nsym.m:43-86; forward: HSS1D/HSS1D_apply_nsym.m:38-101.
  // One launch per level and direction for all m right-hand sides (N x m column-major).
  // fp32: the inverse blocks of the FP32 inverse.
  const int nb = h->nboxes;
  const int64_t ldv = h->vlen + 1;
  double* V = dmalloc<double>(ldv * m);
  std::vector<void*> tmp{V};
  auto launch = [&](const std::vector<SweepOp>& ops) {
    if (ops.empty()) return;
    SweepOp* d = dput(ops);
    tmp.push_back(d);
    const dim3 grid(unsigned(ops.size()), unsigned(m));
    if (fp32) {
      HSSC_TIMED("hss1d_sweep_f32", 0, 0.0, 0.0,
                 hipLaunchKernelGGL(hss1d_sweep_kernel<float>, grid, dim3(256), 0, 0, d,
                                    (const float*)h->W32, h->inv_lo, V, dIn, dOut, h->N, ldv));
    } else {
      HSSC_TIMED("hss1d_sweep_f64", 0, 0.0, 0.0,
                 hipLaunchKernelGGL(hss1d_sweep_kernel<double>, grid, dim3(256), 0, 0, d,
                                    (const double*)h->S, int64_t(0), V, dIn, dOut, h->N, ldv));
    }
  };
  auto up_op = [&](int b, int64_t mat) {
    // up[b] = M [f_b or (up c1; up c2)]  with M = Rinv (inverse) or R (forward)
    SweepOp O{};
    O.rows = h->kskel[b], O.nterm = 1, O.out = h->vUp[b], O.out2 = -1;
    if (h->leaf(b)) O.gofs = h->start[b], O.t[0] = term(mat, h->kskel[b], h->nskel[b], 0, 0, 0, 0, 1);
    else
      O.t[0] = term(mat, h->kskel[b], h->nskel[b], h->vUp[h->c1[b]], h->kskel[h->c1[b]],
                    h->vUp[h->c2[b]]);
    return O;
  };
  for (int lev = h->depth; lev >= 1; --lev) {
    std::vector<SweepOp> ops;
    for (int b = 1; b <= nb; ++b)
      if (h->level[b] == lev) ops.push_back(up_op(b, inverse ? h->oRinv[b] : h->oR[b]));
    launch(ops);
  }
  if (inverse) {
    // per problem: [dn2; dn3] = FTinv [up2; up3] for the root's children 2, 3
    std::vector<SweepOp> tops;
    for (int r : h->roots) {
      const int a = h->c1[r], c = h->c2[r], k2 = h->kskel[a], k3 = h->kskel[c];
      SweepOp O{};
      O.rows = k2 + k3, O.nterm = 1, O.out = h->vDn[a], O.out2 = h->vDn[c], O.split = k2;
      O.t[0] = term(h->oFTinvs[r], k2 + k3, k2 + k3, h->vUp[a], k2, h->vUp[c]);
      tops.push_back(O);
    }
    launch(tops);
    for (int lev = 1; lev <= h->depth; ++lev) {
      std::vector<SweepOp> ops;
      for (int b = 1; b <= nb; ++b) {
        if (h->level[b] != lev) continue;
        const int n = h->nskel[b], k = h->kskel[b];
        SweepOp O{};
        O.rows = n, O.nterm = 2;
        O.t[1] = term(h->oLinv[b], n, k, h->vDn[b], k, 0);
        if (h->leaf(b)) {
          O.scatter = 1, O.gofs = h->start[b];
          O.t[0] = term(h->oDinv[b], n, n, 0, 0, 0, 0, 1);
        } else {
          O.t[0] = term(h->oDinv[b], n, n, h->vUp[h->c1[b]], h->kskel[h->c1[b]], h->vUp[h->c2[b]]);
          O.out = h->vDn[h->c1[b]], O.out2 = h->vDn[h->c2[b]], O.split = h->kskel[h->c1[b]];
        }
        ops.push_back(O);
      }
      launch(ops);
    }
  } else {
    // per problem: dn2 = M_sib2 up3, dn3 = M_sib3 up2 for the root's children 2, 3
    std::vector<SweepOp> tops;
    for (int r : h->roots) {
      const int a = h->c1[r], c = h->c2[r], k2 = h->kskel[a], k3 = h->kskel[c];
      SweepOp A2{}, A3{};
      A2.rows = k2, A2.nterm = 1, A2.out = h->vDn[a], A2.out2 = -1;
      A2.t[0] = term(h->oMsib[a], k2, k3, h->vUp[c], k3, 0);
      A3.rows = k3, A3.nterm = 1, A3.out = h->vDn[c], A3.out2 = -1;
      A3.t[0] = term(h->oMsib[c], k3, k2, h->vUp[a], k2, 0);
      tops.push_back(A2), tops.push_back(A3);
    }
    launch(tops);
    for (int lev = 1; lev <= h->depth; ++lev) {
      std::vector<SweepOp> ops;
      for (int b = 1; b <= nb; ++b) {
        if (h->level[b] != lev) continue;
        const int n = h->nskel[b], k = h->kskel[b];
        if (h->leaf(b)) {
          // u_b = L dn_b + M_self q_b
          SweepOp O{};
          O.rows = n, O.nterm = 2, O.scatter = 1, O.gofs = h->start[b];
          O.t[0] = term(h->oMself[b], n, n, 0, 0, 0, 0, 1);
          O.t[1] = term(h->oL[b], n, k, h->vDn[b], k, 0);
          ops.push_back(O);
        } else {
          // dn c1 = (L dn)(0:n1) + M_sib(c1) up c2 ; dn c2 = (L dn)(n1:) + M_sib(c2) up c1
          const int c1 = h->c1[b], c2 = h->c2[b], n1 = h->kskel[c1], n2 = h->kskel[c2];
          SweepOp O1{}, O2{};
          O1.rows = n1, O1.nterm = 2, O1.out = h->vDn[c1], O1.out2 = -1;
          O1.t[0] = term(h->oL[b], n, k, h->vDn[b], k, 0, 0);
          O1.t[1] = term(h->oMsib[c1], n1, n2, h->vUp[c2], n2, 0);
          O2.rows = n2, O2.nterm = 2, O2.out = h->vDn[c2], O2.out2 = -1;
          O2.t[0] = term(h->oL[b], n, k, h->vDn[b], k, 0, n1);
          O2.t[1] = term(h->oMsib[c2], n2, n1, h->vUp[c1], n1, 0);
          ops.push_back(O1), ops.push_back(O2);
        }
      }
      launch(ops);
    }
  }
  H1CHK(hipDeviceSynchronize());
  H1CHK(hipGetLastError());
  for (void* p : tmp) dfree1(p);
  return HSSC_OK;
}

int hssc_hss1d_applyinv_device(const hssc_hss1d* h, const void* dF, void* dX, int64_t m) {
  if (!h || !dF || !dX || m < 1 || m > 65535)
    return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d applyinv arguments");
  if (!h->inverted) return fail(HSSC_ERR_INVALID_ARG, "hss1d operator is not inverted");
  try {
    return sweep1d(h, static_cast<const double*>(dF), static_cast<double*>(dX), true, false, m);
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_solve_device(const hssc_hss1d* h, const void* dF, void* dX, double tol,
                            int32_t maxit, int32_t* iters, double* relres) {
  // Common/Iterative_Refinement.m:20-35 with the FP32 inverse (the FP64 one when no FP32
  // inverse exists) as the approximate solver and the FP64 compressed operator as A.
  if (!h || !dF || !dX || maxit < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d solve arguments");
  const bool fp32 = h->inverted32;
  if (!fp32 && !h->inverted) return fail(HSSC_ERR_INVALID_ARG, "hss1d operator is not inverted");
  try {
    const int64_t N = h->N;
    const double* F = static_cast<const double*>(dF);
    double* X = static_cast<double*>(dX);
    double* Rv = dmalloc<double>(N);
    double* D = dmalloc<double>(N);
    double* part = dmalloc<double>(256);
    std::vector<void*> tmp{Rv, D, part};
    auto norm = [&](const double* v) {
      k::sumsq<double>(0, N, v, part, 256);
      double ph[256];
      H1CHK(hipMemcpy(ph, part, sizeof(ph), hipMemcpyDeviceToHost));
      double ss = 0;
      for (double x : ph) ss += x;
      return std::sqrt(ss);
    };
    const double nf = norm(F);
    int s = sweep1d(h, F, X, true, fp32);
    int it = 1;
    double rel = 0;
    while (s == HSSC_OK) {
      s = sweep1d(h, X, Rv, false);  // R = A X
      if (s != HSSC_OK) break;
      k::axpby<double, double>(0, N, 1.0, F, -1.0, Rv);  // R = F - A X
      rel = norm(Rv) / (nf > 0 ? nf : 1.0);
      if (rel <= tol || it >= maxit) break;
      s = sweep1d(h, Rv, D, true, fp32);
      if (s != HSSC_OK) break;
      k::axpby<double, double>(0, N, 1.0, D, 1.0, X);  // X += A^-1 R
      ++it;
    }
    H1CHK(hipDeviceSynchronize());
    for (void* p : tmp) dfree1(p);
    if (s != HSSC_OK) return s;
    if (iters) *iters = it;
    if (relres) *relres = rel;
    return rel <= tol ? HSSC_OK : fail(HSSC_ERR_NOT_CONVERGED, "hss1d refinement did not reach tol");
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_solve(const hssc_hss1d* h, const double* F, double* X, double tol, int32_t maxit,
                     int32_t* iters, double* relres) {
  if (!h || !F || !X) return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d solve arguments");
  try {
    double* dF = dmalloc<double>(h->N);
    double* dX = dmalloc<double>(h->N);
    H1CHK(hipMemcpy(dF, F, h->N * 8, hipMemcpyHostToDevice));
    const int s = hssc_hss1d_solve_device(h, dF, dX, tol, maxit, iters, relres);
    if (s == HSSC_OK || s == HSSC_ERR_NOT_CONVERGED)
      H1CHK(hipMemcpy(X, dX, h->N * 8, hipMemcpyDeviceToHost));
    dfree1(dF), dfree1(dX);
    return s;
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_applyinv(const hssc_hss1d* h, const double* F, double* X, int64_t m) {
  if (!h || !F || !X || m < 1 || m > 65535)
    return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d applyinv arguments");
  try {
    const int64_t len = h->N * m;
    double* dF = dmalloc<double>(len);
    double* dX = dmalloc<double>(len);
    H1CHK(hipMemcpy(dF, F, len * 8, hipMemcpyHostToDevice));
    int s = hssc_hss1d_applyinv_device(h, dF, dX, m);
    if (s == HSSC_OK) H1CHK(hipMemcpy(X, dX, len * 8, hipMemcpyDeviceToHost));
    dfree1(dF), dfree1(dX);
    return s;
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_apply_device(const hssc_hss1d* h, const void* dQ, void* dU, int64_t m) {
  if (!h || !dQ || !dU || m < 1 || m > 65535)
    return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d apply arguments");
  try {
    return sweep1d(h, static_cast<const double*>(dQ), static_cast<double*>(dU), false, false, m);
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_apply(const hssc_hss1d* h, const double* Q, double* U, int64_t m) {
  if (!h || !Q || !U || m < 1 || m > 65535)
    return fail(HSSC_ERR_INVALID_ARG, "invalid hss1d apply arguments");
  try {
    const int64_t len = h->N * m;
    double* dQ = dmalloc<double>(len);
    double* dU = dmalloc<double>(len);
    H1CHK(hipMemcpy(dQ, Q, len * 8, hipMemcpyHostToDevice));
    int s = hssc_hss1d_apply_device(h, dQ, dU, m);
    if (s == HSSC_OK) H1CHK(hipMemcpy(U, dU, len * 8, hipMemcpyDeviceToHost));
    dfree1(dQ), dfree1(dU);
    return s;
  } catch (const HipErr& e) {
    return e.status;
  }
}

int hssc_hss1d_info(const hssc_hss1d* h, int32_t* nboxes, int32_t* depth) {
  if (!h) return fail(HSSC_ERR_INVALID_ARG, "null handle");
  if (nboxes) *nboxes = h->nboxes;
  if (depth) *depth = h->depth;
  return HSSC_OK;
}

int hssc_hss1d_box(const hssc_hss1d* h, int32_t box, int32_t* level, int32_t* nskel,
                   int32_t* kskel, int64_t* isk_up, int64_t* isk_dn) {
  if (!h || box < 1 || box > h->nboxes) return fail(HSSC_ERR_INVALID_ARG, "box out of range");
  if (level) *level = h->level[box];
  if (nskel) *nskel = h->nskel[box];
  if (kskel) *kskel = h->kskel[box];
  if (isk_up) std::copy(h->isk_up[box].begin(), h->isk_up[box].end(), isk_up);
  if (isk_dn) std::copy(h->isk_dn[box].begin(), h->isk_dn[box].end(), isk_dn);
  return HSSC_OK;
}

void hssc_hss1d_destroy(hssc_hss1d* h) { delete h; }

}  // extern "C"

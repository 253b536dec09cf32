// Host-side launchers of the HIP kernels (kernels.hip).  All launches are asynchronous on
// the given stream; arrays are column-major; s* arguments are per-batch element strides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hssc.h"
#include "scalar.h"

namespace hssc {
namespace k {

// The kernel of Common/Kernel_Eval.m:29-96 as evaluated on the GPU (2D plane, per-point
// weights b(x) = c(x) passed with the points).
struct KernelSpec {
  int pot = HSSC_POT_SL_L_2D;  // HSSC_POT_*
  double scale = 0;            // SL_L_2D: h^2 / 2 (times log r^2); SL_H_2D: h^2 k^2 / 4
  double kh = 0;               // wavenumber
  double wre = 0, wim = 0;     // SL_H_2D diagonal w = 1 + k^2 (i/4) corr0
  double lshift = 0;           // SL_L_2D: added to log r^2 (log L^2 of a length scale L)
};

// One entry of Common/Kernel_Eval.m:29-96 (2D plane):
// SL_L_2D: A(r,c) = (den==0) + h^2 * 0.5 * w_r * w_c * log(den),  den = |x_r - x_c|^2
//          (the log term vanishes at den == 0: log(den + (den==0)) = 0 in the reference);
//          with a kernel scale s and length scale L, s (L h)^2 w_r w_c log(L r), i.e. scale
//          0.5 s L^2 h^2 and log(den) + log L^2.
// SL_H_2D: A(r,c) = w (den==0) + h^2 k^2 (i/4) w_r w_c H0(k sqrt(den)), w = 1 + k^2 (i/4)
//          corr0 (:44-49; b and c are not applied on the diagonal).  With H0 = J0 + i Y0,
//          (i c)(J0 + i Y0) = -c Y0 + i c J0.
template <typename T>
__device__ __forceinline__ T kernel_entry(const KernelSpec& ks, double xr, double yr, double wr,
                                          double xc, double yc, double wc) {
  const double dx = xr - xc, dy = yr - yc;
  const double den = dx * dx + dy * dy;
  if constexpr (is_complex_v<T>) {
    using R = real_t<T>;
    if (ks.pot == HSSC_POT_SL_H_2D) {
      if (den == 0.0) return T(R(ks.wre), R(ks.wim));
      const double z = ks.kh * sqrt(den), c = ks.scale * wr * wc;
      return T(R(-c * y0(z)), R(c * j0(z)));
    }
    if (den == 0.0) return T(R(1));
    return T(R(ks.scale * wr * wc * (log(den) + ks.lshift)));
  } else {
    if (den == 0.0) return T(1.0);
    return T(ks.scale * wr * wc * (log(den) + ks.lshift));
  }
}

// A(r,c) = K(R_r, C_c) with per-point weights (real T for SL_L_2D, complex for SL_H_2D).
template <typename T>
void gen_block(hipStream_t st, int64_t m, int64_t n, int64_t batch, const double* Rx,
               const double* Ry, const double* Rw, int64_t sR, const double* Cx, const double* Cy,
               const double* Cw, int64_t sC, T* A, int64_t lda, int64_t sA, const KernelSpec& ks);

// Parent F = [E1 K12; K12^T E2] assembled directly in the parent's J order.
template <typename T>
void gen_parent(hipStream_t st, int64_t n, int64_t kc, int64_t batch, const int32_t* J,
                int64_t sJ, const double* Px, const double* Py, const double* Pw, int64_t sP,
                const T* Ech, int64_t sEbox, int64_t sEparent, T* F, int64_t sF,
                const KernelSpec& ks);

template <typename T>
void gemm(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
          const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta, T* C,
          int64_t ldc, int64_t sC, int64_t batch);

// C = alpha A^H B + beta C (A^T for real data)
template <typename T>
void gemm_h(hipStream_t st, int64_t m, int64_t n, int64_t kdim, T alpha, const T* A, int64_t lda,
            int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta, T* C, int64_t ldc, int64_t sC,
            int64_t batch);

template <typename T>
void gemm_skip(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
               const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta,
               T* C, int64_t ldc, int64_t sC, int64_t batch, int64_t rtiles, int64_t ctiles,
               int64_t rs0, int64_t rsn, int64_t cs0, int64_t csn, int tri = 0);

// Lower 64 x 64 tiles of a symmetric matrix from the upper ones (A(i, j) = A(j, i) for the
// tiles below the diagonal; the diagonal tiles are left as they are).
template <typename T>
void sym_lower_fill(hipStream_t st, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch);

// C = A^H B with the long K dimension split in chunks of kchunk rows (deterministic sum);
// the split needs batch * ceil(kdim/kchunk) * m * n elements of part, whose capacity is
// part_cap elements: a product whose partial sums would not fit is computed unsplit.
template <typename T>
void gemm_tn_splitk(hipStream_t st, int64_t m, int64_t n, int64_t kdim, const T* A, int64_t lda,
                    int64_t sA, const T* B, int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC,
                    int64_t batch, T* part, int64_t kchunk, int64_t part_cap);
// W2 = Ts^H (A^H B) for a sub-panel update (m <= 32 rows): when the product is split along K,
// the reduction of the chunks and the product with the triangular factor Ts (m x m) run as one
// launch; otherwise W1 = A^H B and W2 = Ts^H W1 as two products.  W1, W2: m x n, ld m.
template <typename T>
void gemm_tn_splitk_th(hipStream_t st, int64_t m, int64_t n, int64_t kdim, const T* A,
                       int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, const T* Ts,
                       int64_t ldt, int64_t sT, T* W1, T* W2, int64_t sW, int64_t batch, T* part,
                       int64_t kchunk, int64_t part_cap);
// Elements of the split-K buffer of one level's least squares (per matrix): the outer-panel
// products (jbo x nc, chunks of kchunk rows), the sub-panel products (jb x nin with
// jb <= kQrNb and nin < n, chunks of kchunk_thin rows) and the panel Gram matrices
// (nbo x nbo, chunks of kchunk_thin rows).
int64_t ls_part_elems(int64_t M, int64_t n, int64_t nbo, int64_t kchunk, int64_t kchunk_thin);

// In-place inverse of a batch of n x n matrices.  aux (optional) is a second stream on
// which the blocked sweep inverts the next diagonal block while st finishes the current
// step; all work is ordered back onto st before the call returns.
// work: invert_ws_elems<T>(n, batch) elements (none when one kernel inverts the matrices).
// sym: the matrices are symmetric (A^T = A, complex symmetric for complex T); complex ones
// are then swept on their upper tiles only.  flag (optional, device int): set to nonzero when
// a pivot of a diagonal block is not finite or at most n eps times the block's largest entry.
template <typename T>
void invert_batched(hipStream_t st, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch,
                    hipStream_t aux, T* work, bool sym, int* flag = nullptr);
template <typename T>
int64_t invert_ws_elems(int64_t n, int64_t batch);

constexpr int kQrNb = 32;  // QR panel width (= QR_NB in kernels.hip)

// One 32-column panel of the blocked Householder QR: shifted CholeskyQR3 with Householder
// reconstruction.  Writes the panel's R rows into A, the unit lower trapezoidal V (rows
// j0..M-1, ld ldv) and the compact-WY factor Tf (32 x 32).  The Gram matrices, Cholesky
// factors, Q_1 and the reconstruction are computed in FP64 also for FP32 data (whose panels
// can be too ill-conditioned for CholeskyQR in FP32 arithmetic); only A, V and Tf are stored
// in T.  ws: batch * cqr_ws_elems(M) FP64 elements of the field, laid out per matrix as
// Racc | Qtop | two sets of ceil(M / kCqrRows) partial Gram blocks | Q_1 (M x 32).
constexpr int kCqrRows = 256;
constexpr int64_t cqr_row_blocks(int64_t M) { return (M + kCqrRows - 1) / kCqrRows; }
constexpr int64_t cqr_ws_elems(int64_t M) {
  return (2 + 2 * cqr_row_blocks(M)) * kQrNb * kQrNb + M * kQrNb;
}
template <typename T>
void cqr_panel(hipStream_t st, int64_t M, int64_t j0, int jb, T* A, int64_t lda, int64_t sA,
               T* V, int64_t ldv, int64_t sV, T* Tf, int64_t sT, wide_t<T>* ws, int64_t batch);
// T (n x n, upper) of H_0 ... H_{ns-1} from the 32 x 32 factors Ts (stride 32*32 per block,
// sTs per matrix) and the Gram matrix GV = V^H V (n x n); X: n * 32 scratch per matrix.
template <typename T>
void t_assemble(hipStream_t st, int n, const T* Ts, int64_t sTs, const T* GV, int64_t ldg,
                int64_t sG, T* To, int64_t ldt, int64_t sT, T* X, int64_t sX, int64_t batch);
template <typename T>
void zero2d(hipStream_t st, int64_t m, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch);
// Regularisation rows below the Mq x n block of each matrix: A(Mq + i, i) = lambda (i < k),
// lambda = eta times the RMS norm of the block's first min(k, 32) columns (the other entries
// of the k appended rows must already be zero).
template <typename T>
void tikhonov_rows(hipStream_t st, int64_t Mq, int64_t k, T* A, int64_t lda, int64_t sA,
                   int64_t batch, double eta);
template <typename T>
void trinv_upper(hipStream_t st, int64_t p, int nb, const T* R, int64_t ldr, int64_t sR, T* W,
                 int64_t sW, int64_t batch);
// Every 32 x 32 diagonal block of the leading kk x kk upper triangle, in one launch: block q
// (rows [32 q, min(32 q + 32, kk))) into W + e sW + q 32^2 (ld its size); sW >= ceil(kk / 32) 32^2.
template <typename T>
void trinv_upper_all(hipStream_t st, int64_t kk, const T* R, int64_t ldr, int64_t sR, T* W,
                     int64_t sW, int64_t batch);

// y = alpha op(A) x + gamma z ; work needs batch * ceil(inner/512) * out elements.
template <typename T>
void gemv(hipStream_t st, bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
          int64_t sA, const T* x, int64_t sx, T gamma, const T* z, int64_t sz, T* y, int64_t sy,
          int64_t batch, T* work);
constexpr int64_t kGemvChunk = 512;

// y = alpha A x + gamma z for symmetric A (A^T = A), read from its upper triangle only;
// work needs symv_work_elems(n, batch) elements.
template <typename T>
void symv_upper(hipStream_t st, int64_t n, T alpha, const T* A, int64_t lda, int64_t sA,
                const T* x, int64_t sx, T gamma, const T* z, int64_t sz, T* y, int64_t sy,
                int64_t batch, T* work);
int64_t symv_work_elems(int64_t n, int64_t batch);

template <typename TS, typename TD>
void gather(hipStream_t st, int64_t n, int64_t m, const int32_t* idx, int64_t sidx, const TS* src,
            int64_t ld_src, int64_t s_src, TD* dst, int64_t ld_dst, int64_t s_dst, int64_t batch);
template <typename TS, typename TD>
void scatter(hipStream_t st, int64_t n, int64_t m, const int32_t* idx, int64_t sidx, const TS* src,
             int64_t ld_src, int64_t s_src, TD* dst, int64_t ld_dst, int64_t s_dst, int64_t batch,
             bool accumulate);
template <typename TX, typename TY>
void axpby(hipStream_t st, int64_t len, TY a, const TX* x, TY b, TY* y);
template <typename TS, typename TD>
void copy2d(hipStream_t st, int64_t m, int64_t n, const TS* src, int64_t lds, int64_t ss, TD* dst,
            int64_t ldd, int64_t sd, int64_t batch);
template <typename T>
void add2d(hipStream_t st, int64_t m, int64_t n, T a, const T* s1, int64_t ld1, int64_t ss1, T b,
           const T* s2, int64_t ld2, int64_t ss2, T* dst, int64_t ldd, int64_t sd, int64_t batch);
template <typename T>
void sumsq(hipStream_t st, int64_t len, const T* x, double* part, int nblocks);

// Batched ragged interpolative decomposition (id.hip).  Work per matrix: 5 n + m; after the
// call work[woff[b] + n + j] holds |R_jj| (real part) for j < min(m, n).  mmax and nmax (> 0
// when every block is known to be at most mmax x nmax) select the register-resident kernel for
// complex blocks that fit it; a block larger than those maxima gets rank -1.  kpin[b] (when
// given) leading columns of block b are factored in order before column pivoting starts (the
// global-memory kernel runs then).
template <typename T>
void cpqr_id(hipStream_t st, int nmat, const int64_t* off, const int* mr, const int* nc, T* A,
             const T* Acopy, double acc, const int* kfix, int* perm, const int64_t* poff, int* kout,
             T* Tout, const int64_t* toff, T* work, const int64_t* woff, uint64_t seed,
             int64_t mmax = 0, int64_t nmax = 0, const int* kpin = nullptr);

// Acceptance of IDs computed on sketches with sr[b] rows, checked on the full blocks (n <= 256):
// ok[b] = 1 when |A_s T x - A_r x| <= 10 acc |A g| for random unit x, g (Common/ID.m:55-59)
// and the sketch was not saturated.  T of block b at Tout + toff[b], leading dimension kk[b].
template <typename T>
void id_verify(hipStream_t st, int nmat, const int64_t* off, const int* mr, const int* nc,
               const int* sr, const T* A, const int* perm, const int64_t* poff, const int* kk,
               const T* Tout, const int64_t* toff, double acc, uint64_t seed, int* ok);

// D_j (rows[j] x n, ld rows[j], at D + doff[j]) = the first rows[j] rows of S_j (ld lds, at
// S + soff[j]), j < nj: ragged sketches out of a uniform batch.
template <typename T>
void pack_rows(hipStream_t st, int nj, const int64_t* soff, int64_t lds, const int64_t* doff,
               const int* rows, int64_t n, const T* S, T* D);

// G[i] ~ N(0,1), deterministic in (seed, base + i).
void randn(hipStream_t st, int64_t len, uint64_t seed, uint64_t base, double* G);

}  // namespace k
}  // namespace hssc

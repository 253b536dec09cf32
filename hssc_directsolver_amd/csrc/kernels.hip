// HIP kernels for the HSS compress / invert / apply path on CDNA4 (gfx950).
//
// All matrices are column-major.  Batched kernels take a batch index in blockIdx.z (or
// blockIdx.y) and element strides per batch; a stride of 0 shares one operand across the
// batch (translation-invariant levels).  Kernels are templated on the working scalar
// (double for the FP64 path, float for the FP32 factorization).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <climits>

#include "kernels.h"
#include "stats.h"

namespace hssc {
namespace k {

template <typename T>
constexpr const char* tname(const char* f64, const char* f32) {
  return sizeof(T) == 8 ? f64 : f32;
}

// ------------------------------------------------------------------------------------
// Kernel-block generation (Common/Kernel_Eval.m:29-96, 2D Laplace single layer)
// ------------------------------------------------------------------------------------
// A(r,c) = (den==0) + h^2 * 0.5 * w_r * w_c * log(den),  den = |x_r - x_c|^2,
// with the log term dropped when den == 0 (log(den + (den==0)) = 0 in the reference).

template <typename T>
__device__ __forceinline__ T laplace_entry(double xr, double yr, double wr, double xc, double yc,
                                           double wc, double hh2) {
  const double dx = xr - xc, dy = yr - yc;
  const double den = dx * dx + dy * dy;
  if (den == 0.0) return T(1.0);
  return T(hh2 * wr * wc * log(den));
}

// One thread per entry; threadIdx.x runs along rows so column-major stores coalesce.
template <typename T>
__global__ void __launch_bounds__(256) gen_block_kernel(
    int64_t m, int64_t n, const double* __restrict__ Rx, const double* __restrict__ Ry,
    const double* __restrict__ Rw, int64_t sR, const double* __restrict__ Cx,
    const double* __restrict__ Cy, const double* __restrict__ Cw, int64_t sC, T* __restrict__ A,
    int64_t lda, int64_t sA, double hh2) {
  const int64_t e = blockIdx.z;
  const int64_t r = int64_t(blockIdx.x) * 64 + (threadIdx.x & 63);
  const int64_t c0 = int64_t(blockIdx.y) * 16 + (threadIdx.x >> 6) * 4;
  if (r >= m) return;
  Rx += e * sR, Ry += e * sR, Rw += e * sR;
  Cx += e * sC, Cy += e * sC, Cw += e * sC;
  A += e * sA;
  const double xr = Rx[r], yr = Ry[r], wr = Rw[r];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t c = c0 + j;
    if (c < n) A[r + c * lda] = laplace_entry<T>(xr, yr, wr, Cx[c], Cy[c], Cw[c], hh2);
  }
}

template <typename T>
void gen_block(hipStream_t st, int64_t m, int64_t n, int64_t batch, const double* Rx,
               const double* Ry, const double* Rw, int64_t sR, const double* Cx, const double* Cy,
               const double* Cw, int64_t sC, T* A, int64_t lda, int64_t sA, double h) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((m + 63) / 64, (n + 15) / 16, batch);
  const double ent = double(m) * n * batch;
  HSSC_TIMED(tname<T>("gen_block_f64", "gen_block_f32"), st, ent, ent * sizeof(T),
             hipLaunchKernelGGL(gen_block_kernel<T>, grid, dim3(256), 0, st, m, n, Rx, Ry, Rw, sR,
                                Cx, Cy, Cw, sC, A, lda, sA, 0.5 * h * h));
}

// Parent block F = [E1 K12; K12^T E2] in the parent's J order
// (HSS2D/HSS2D_build_Finv_fsym.m:93-101):  FJ(a,c) = F(J[a], J[c]).  Natural index
// i < kc belongs to child 1 (its skeleton position i), i >= kc to child 2.
template <typename T>
__global__ void __launch_bounds__(256) gen_parent_kernel(
    int64_t n, int64_t kc, const int32_t* __restrict__ J, int64_t sJ,
    const double* __restrict__ Px, const double* __restrict__ Py, const double* __restrict__ Pw,
    int64_t sP, const T* __restrict__ Ech, int64_t sEbox, int64_t sEkey_per_parent, T* __restrict__ F,
    int64_t sF, double hh2) {
  const int64_t e = blockIdx.z;
  const int64_t a = int64_t(blockIdx.x) * 64 + (threadIdx.x & 63);
  const int64_t c0 = int64_t(blockIdx.y) * 16 + (threadIdx.x >> 6) * 4;
  if (a >= n) return;
  J += e * sJ;
  Px += e * sP, Py += e * sP, Pw += e * sP;
  F += e * sF;
  // children E matrices: child s of parent e at Ech + e*sEkey_per_parent + s*sEbox
  const T* E1 = Ech + e * sEkey_per_parent;
  const T* E2 = E1 + sEbox;
  const int32_t i = J[a];
  const double xa = Px[a], ya = Py[a], wa = Pw[a];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int64_t c = c0 + jj;
    if (c >= n) break;
    const int32_t j = J[c];
    T v;
    if (i < kc && j < kc) {
      v = E1[i + int64_t(j) * kc];
    } else if (i >= kc && j >= kc) {
      v = E2[(i - kc) + int64_t(j - kc) * kc];
    } else {
      v = laplace_entry<T>(xa, ya, wa, Px[c], Py[c], Pw[c], hh2);
    }
    F[a + c * n] = v;
  }
}

template <typename T>
void gen_parent(hipStream_t st, int64_t n, int64_t kc, int64_t batch, const int32_t* J,
                int64_t sJ, const double* Px, const double* Py, const double* Pw, int64_t sP,
                const T* Ech, int64_t sEbox, int64_t sEparent, T* F, int64_t sF, double h) {
  dim3 grid((n + 63) / 64, (n + 15) / 16, batch);
  const double ent = double(n) * n * batch;
  HSSC_TIMED(tname<T>("gen_parent_f64", "gen_parent_f32"), st, ent, ent * sizeof(T),
             hipLaunchKernelGGL(gen_parent_kernel<T>, grid, dim3(256), 0, st, n, kc, J, sJ, Px, Py,
                                Pw, sP, Ech, sEbox, sEparent, F, sF, 0.5 * h * h));
}

// ------------------------------------------------------------------------------------
// Batched GEMM on MFMA:  C = alpha * op(A) * op(B) + beta * C
// ------------------------------------------------------------------------------------
// 64x64 output tile per 256-thread workgroup, K staged through LDS in steps of 16.
// Each wave owns a 32x32 quarter = 2x2 MFMA 16x16 tiles.  The MFMA is issued as
// D = B^T A^T so that a lane's 4 accumulators run down one column of C with lane%16
// indexing rows: column-major stores of 16 consecutive rows coalesce.

// Operand layout (both types): lane l supplies A(l%16, l/16) and B(l/16, l%16).
// Result layout, measured on gfx950: item r of lane l holds D(row(l,r), l%16) with
//   f64 16x16x4: row = l/16 + 4r        f32 16x16x4: row = 4*(l/16) + r.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int GB_M = 64, GB_N = 64, GB_K = 16;

// Tile origins at or beyond rs0 (cs0) are shifted by rsn (csn): the blocked Gauss-Jordan
// steps use this to run over the index set [0, p) U [p+nb, n) with 64-aligned p.
template <typename T, bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_kernel(int64_t m, int64_t n, int64_t kdim, T alpha,
                                                  const T* A, int64_t lda, int64_t sA,
                                                  const T* B, int64_t ldb, int64_t sB,
                                                  T beta, T* C, int64_t ldc,
                                                  int64_t sC, int64_t rs0, int64_t rsn, int64_t cs0,
                                                  int64_t csn) {
  using v4 = typename Mfma<T>::v4;
  __shared__ T As[GB_K][GB_M + 1];  // As[kk][i] = op(A)(i0+i, k0+kk)
  __shared__ T Bs[GB_K][GB_N + 1];  // Bs[kk][j] = op(B)(k0+kk, j0+j)
  const int64_t e = blockIdx.z;
  A += e * sA;
  B += e * sB;
  C += e * sC;
  int64_t i0 = int64_t(blockIdx.x) * GB_M, j0 = int64_t(blockIdx.y) * GB_N;
  if (i0 >= rs0) i0 += rsn;
  if (j0 >= cs0) j0 += csn;
  if (i0 >= m || j0 >= n) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  v4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = v4{0, 0, 0, 0};

  for (int64_t k0 = 0; k0 < kdim; k0 += GB_K) {
    // Load op(A) tile (64 x 16): 1024 elements, 4 per thread.
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int idx = tid + t * 256;
      int ii, kk;
      if (!TA) {
        ii = idx & 63, kk = idx >> 6;  // contiguous along i
      } else {
        kk = idx & 15, ii = idx >> 4;  // contiguous along k
      }
      const int64_t gi = i0 + ii, gk = k0 + kk;
      T v = T(0);
      if (gi < m && gk < kdim) v = TA ? A[gk + gi * lda] : A[gi + gk * lda];
      As[kk][ii] = v;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int idx = tid + t * 256;
      int jj, kk;
      if (!TB) {
        kk = idx & 15, jj = idx >> 4;  // contiguous along k
      } else {
        jj = idx & 63, kk = idx >> 6;  // contiguous along j
      }
      const int64_t gj = j0 + jj, gk = k0 + kk;
      T v = T(0);
      if (gj < n && gk < kdim) v = TB ? B[gj + gk * ldb] : B[gk + gj * ldb];
      Bs[kk][jj] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      const int kq = kk + (lane >> 4);
      T av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = As[kq][wi + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = Bs[kq][wj + y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = Mfma<T>::mma(bv[y], av[x], acc[x][y]);
    }
    __syncthreads();
  }
  // D(jrow, icol) = sum_k B(k, j) A(i, k): the D row of item r is the C column, the D
  // column lane%16 is the C row.
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = i0 + wi + x * 16 + (lane & 15);
        const int64_t gj = j0 + wj + y * 16 + Mfma<T>::row(lane, r);
        if (gi < m && gj < n) {
          T* c = C + gi + gj * ldc;
          const T v = alpha * acc[x][y][r];
          *c = (beta == T(0)) ? v : v + beta * *c;
        }
      }
}

template <typename T>
void gemm_skip(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
               const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta,
               T* C, int64_t ldc, int64_t sC, int64_t batch, int64_t rtiles, int64_t ctiles,
               int64_t rs0, int64_t rsn, int64_t cs0, int64_t csn) {
  if (m <= 0 || n <= 0 || batch <= 0 || rtiles <= 0 || ctiles <= 0) return;
  dim3 grid(rtiles, ctiles, batch);
  // algorithmic work: rows / cols actually produced (the skipped block excluded)
  const double rows = double(rs0 < m ? m - rsn : m), cols = double(cs0 < n ? n - csn : n);
  const double fl = 2.0 * rows * cols * double(kdim) * double(batch);
  const double by = (rows * kdim + kdim * cols + 2 * rows * cols) * sizeof(T) * double(batch);
#define HSSC_GEMM_LAUNCH(TA_, TB_)                                                              \
  HSSC_TIMED(tname<T>("gemm_f64", "gemm_f32"), st, fl, by,                                      \
             hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_>), grid, dim3(256), 0, st, m, n, kdim,  \
                                alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, rs0, rsn, cs0, \
                                csn))
  if (!ta && !tb) HSSC_GEMM_LAUNCH(false, false);
  else if (ta && !tb) HSSC_GEMM_LAUNCH(true, false);
  else if (!ta && tb) HSSC_GEMM_LAUNCH(false, true);
  else HSSC_GEMM_LAUNCH(true, true);
#undef HSSC_GEMM_LAUNCH
}

template <typename T>
void gemm(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
          const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta, T* C,
          int64_t ldc, int64_t sC, int64_t batch) {
  const int64_t big = INT64_MAX;
  gemm_skip<T>(st, ta, tb, m, n, kdim, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
               (m + GB_M - 1) / GB_M, (n + GB_N - 1) / GB_N, big, 0, big, 0);
}

// ------------------------------------------------------------------------------------
// Gauss-Jordan inversion (MATLAB inv in HSS2D_invert_fsym.m:108, build_Finv_fsym.m:101,
// build_E_fsym.m:51,64)
// ------------------------------------------------------------------------------------
// In-place GJ sweep with partial (row) pivoting.  After pivot c: A(c,c) = 1/p,
// A(c,j) /= p, A(i,j) -= A(i,c)_old * A(c,j), A(i,c) = -A(i,c)_old / p.  Row swaps are
// undone at the end by swapping columns in reverse order.  The matrix lives in LDS.

template <typename T>
__device__ void gj_lds(T* S, int n, int ld, int* piv, T* f, int tid, int nthr) {
  const int lane = tid & 63, wave = tid >> 6;
  __shared__ int s_p;
  for (int c = 0; c < n; ++c) {
    if (wave == 0) {
      // argmax |S(r,c)| over r >= c; ties -> lowest row (LAPACK idamax convention)
      T best = T(-1);
      int bi = c;
      for (int r = c + lane; r < n; r += 64) {
        const T v = fabs(S[r + c * ld]);
        if (v > best) best = v, bi = r;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const T ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) best = ob, bi = oi;
      }
      if (lane == 0) s_p = bi, piv[c] = bi;
    }
    __syncthreads();
    const int p = s_p;
    if (p != c) {
      for (int j = tid; j < n; j += nthr) {
        const T t = S[c + j * ld];
        S[c + j * ld] = S[p + j * ld];
        S[p + j * ld] = t;
      }
      __syncthreads();
    }
    for (int i = tid; i < n; i += nthr) f[i] = S[i + c * ld];
    __syncthreads();
    const T inv = T(1) / f[c];
    for (int j = tid; j < n; j += nthr) S[c + j * ld] = (j == c) ? inv : S[c + j * ld] * inv;
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += nthr) {
      const int i = idx % n, j = idx / n;
      if (i == c) continue;
      if (j == c) S[i + j * ld] = -f[i] * inv;
      else S[i + j * ld] -= f[i] * S[c + j * ld];
    }
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; --c) {
    const int p = piv[c];
    if (p != c) {
      for (int i = tid; i < n; i += nthr) {
        const T t = S[i + c * ld];
        S[i + c * ld] = S[i + p * ld];
        S[i + p * ld] = t;
      }
      __syncthreads();
    }
  }
}

// Small matrices: one workgroup per matrix, whole matrix in dynamic LDS.
template <typename T>
__global__ void __launch_bounds__(256) gj_small_kernel(int n, T* __restrict__ A, int64_t lda,
                                                       int64_t sA) {
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);
  const int ld = n + 1;
  T* f = S + int64_t(ld) * n;
  int* piv = reinterpret_cast<int*>(f + n);
  A += int64_t(blockIdx.x) * sA;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;
    S[i + j * ld] = A[i + j * lda];
  }
  __syncthreads();
  gj_lds<T>(S, n, ld, piv, f, tid, 256);
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;
    A[i + j * lda] = S[i + j * ld];
  }
}

template <typename T>
size_t gj_small_lds(int n) {
  return sizeof(T) * (size_t(n + 1) * n + n) + sizeof(int) * n;
}

// Largest dynamic LDS a workgroup may request on the current device (160 KiB on gfx950).
static int max_lds_per_block() {
  static int v = -1;
  if (v < 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
  }
  return v;
}

template <typename T>
bool gj_small(hipStream_t st, int n, T* A, int64_t lda, int64_t sA, int64_t batch) {
  if (batch <= 0 || n <= 0) return true;
  const size_t lds = gj_small_lds<T>(n);
  if (lds > size_t(max_lds_per_block())) return false;
  if (lds > 65536) {
    static size_t granted = 65536;
    if (lds > granted) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(&gj_small_kernel<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
        return false;
      granted = lds;
    }
  }
  const double fl = 2.0 * n * double(n) * n * batch;
  HSSC_TIMED(tname<T>("gj_small_f64", "gj_small_f32"), st, fl, 2.0 * n * n * sizeof(T) * batch,
             hipLaunchKernelGGL(gj_small_kernel<T>, dim3(batch), dim3(256), lds, st, n, A, lda,
                                sA));
  return true;
}

// Blocked GJ without pivoting across diagonal blocks (pivoting inside each nb x nb block).
// Step p with block rows/cols B = [p, p+nb):
//   (1) A_BB <- inv(A_BB)                                   gj_diag_kernel
//   (2) A_XB <- -A_XB * A_BB          (rows X outside B)    gj_col_kernel
//   (3) A_XY <- A_XY + A_XB * A_BY    (X, Y outside B)      gj_rank_kernel
//   (4) A_BY <- A_BB * A_BY           (cols Y outside B)    gj_row_kernel
// which is the block sweep A11' = A11^-1, A12' = A11^-1 A12, A21' = -A21 A11^-1,
// A22' = A22 - A21 A11^-1 A12.
constexpr int GJ_NB = 64;

template <typename T>
__global__ void __launch_bounds__(256) gj_diag_kernel(int64_t p, int nb, T* __restrict__ A,
                                                      int64_t lda, int64_t sA) {
  __shared__ T S[GJ_NB * (GJ_NB + 1)];
  __shared__ T f[GJ_NB];
  __shared__ int piv[GJ_NB];
  A += int64_t(blockIdx.x) * sA + p + p * lda;
  const int tid = threadIdx.x, ld = GJ_NB + 1;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int i = idx % nb, j = idx / nb;
    S[i + j * ld] = A[i + j * lda];
  }
  __syncthreads();
  gj_lds<T>(S, nb, ld, piv, f, tid, 256);
  __syncthreads();
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int i = idx % nb, j = idx / nb;
    A[i + j * lda] = S[i + j * ld];
  }
}

template <typename T>
void gj_blocked(hipStream_t st, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch) {
  // p advances in steps of 64 = the GEMM tile, so no 64-row tile straddles block B; the
  // tile-skip ranges of gemm_skip enumerate the tiles of [0,p) U [p+nb,n).  Steps (2) and
  // (4) write in place: each output tile is produced by one workgroup that reads all of
  // its inputs (K = nb <= 64 fits one tile) before it stores.
  static_assert(GJ_NB == GB_M && GJ_NB == GB_N, "block size must equal the GEMM tile");
  for (int64_t p = 0; p < n; p += GJ_NB) {
    const int nb = int(p + GJ_NB <= n ? GJ_NB : n - p);
    const int64_t nout = n - nb;
    HSSC_TIMED(tname<T>("gj_diag_f64", "gj_diag_f32"), st, 2.0 * nb * nb * nb * batch,
               2.0 * nb * nb * sizeof(T) * batch,
               hipLaunchKernelGGL(gj_diag_kernel<T>, dim3(batch), dim3(256), 0, st, p, nb, A, lda,
                                  sA));
    if (nout <= 0) continue;
    const int64_t tiles = (nout + 63) / 64;
    T* Apb = A + p;             // rows B
    T* Abp = A + p * lda;       // cols B
    T* Abb = A + p + p * lda;   // block B
    // (2) A_XB = -A_XB * A_BB  (m = n rows with skip, n = nb cols)
    gemm_skip<T>(st, false, false, n, nb, nb, T(-1), A + p * lda, lda, sA, Abb, lda, sA, T(0),
                 Abp, lda, sA, batch, tiles, 1, p, nb, INT64_MAX, 0);
    // (3) A_XY += A_XB * A_BY
    gemm_skip<T>(st, false, false, n, n, nb, T(1), Abp, lda, sA, Apb, lda, sA, T(1), A, lda, sA,
                 batch, tiles, tiles, p, nb, p, nb);
    // (4) A_BY = A_BB * A_BY
    gemm_skip<T>(st, false, false, nb, n, nb, T(1), Abb, lda, sA, Apb, lda, sA, T(0), Apb, lda,
                 sA, batch, 1, tiles, INT64_MAX, 0, p, nb);
  }
}

template <typename T>
void invert_batched(hipStream_t st, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch) {
  if (n <= 0 || batch <= 0) return;
  if (n <= 128 && gj_small<T>(st, int(n), A, lda, sA, batch)) return;
  gj_blocked<T>(st, n, A, lda, sA, batch);
}

// ------------------------------------------------------------------------------------
// Householder QR least squares (MATLAB K11\K21 in HSS2D/interpolation_operator.m:66)
// ------------------------------------------------------------------------------------
// Panel kernel: one workgroup factors columns [j0, j0+jb) of A (rows j0..M-1), streaming
// the panel through L2.  Reflector i: H_i = I - tau_i v_i v_i^T with v_i(0) = 1 stored
// below the diagonal, beta on the diagonal (LAPACK dgeqr2 convention, beta = -sign(a) *
// norm).  It then forms the compact-WY triangular factor Tf (jb x jb, upper) with
// H_0 ... H_{jb-1} = I - V Tf V^T (dlarft, forward, columnwise).

template <typename T>
__device__ T block_sum(T v, T* red, int tid) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  T s = T(0);
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

constexpr int QR_NB = 32;

template <typename T>
__global__ void __launch_bounds__(1024) qr_panel_kernel(int64_t M, int64_t j0, int jb,
                                                        T* __restrict__ A, int64_t lda, int64_t sA,
                                                        T* __restrict__ Tf, int64_t sT) {
  __shared__ T red[16 * QR_NB];
  __shared__ T sdot[QR_NB];
  __shared__ T stau[QR_NB];
  __shared__ T G[QR_NB][QR_NB + 1];
  A += int64_t(blockIdx.x) * sA;
  Tf += int64_t(blockIdx.x) * sT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nthr = blockDim.x;
  const int nw = nthr >> 6;
  for (int c = 0; c < jb; ++c) {
    const int64_t col = j0 + c;
    const int64_t r0 = j0 + c;  // diagonal row
    T* a = A + col * lda;
    // norm^2 of a(r0+1 .. M-1)
    T s = T(0);
    for (int64_t r = r0 + 1 + tid; r < M; r += nthr) s += a[r] * a[r];
    s = block_sum<T>(s, red, tid);
    const T alpha = a[r0];
    T beta, tau, scal;
    if (s == T(0)) {
      tau = T(0), beta = alpha, scal = T(0);
    } else {
      T nrm = sqrt(alpha * alpha + s);
      beta = alpha >= T(0) ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    __syncthreads();
    for (int64_t r = r0 + 1 + tid; r < M; r += nthr) a[r] *= scal;
    if (tid == 0) {
      a[r0] = beta;
      stau[c] = tau;
    }
    __syncthreads();
    // apply H_c to the remaining panel columns: w_j = v^T a_j, a_j -= tau v w_j
    const int nrest = jb - c - 1;
    if (nrest > 0 && tau != T(0)) {
      T part[QR_NB];
#pragma unroll
      for (int j = 0; j < QR_NB; ++j) part[j] = T(0);
      for (int64_t r = r0 + tid; r < M; r += nthr) {
        const T v = (r == r0) ? T(1) : a[r];
#pragma unroll
        for (int j = 0; j < QR_NB; ++j)
          if (j < nrest) part[j] += v * A[r + (col + 1 + j) * lda];
      }
#pragma unroll
      for (int j = 0; j < QR_NB; ++j) {
        T v = part[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0 && j < nrest) red[wave * QR_NB + j] = v;
      }
      __syncthreads();
      if (tid < nrest) {
        T t = T(0);
        for (int w = 0; w < nw; ++w) t += red[w * QR_NB + tid];
        sdot[tid] = t * tau;
      }
      __syncthreads();
      for (int64_t r = r0 + tid; r < M; r += nthr) {
        const T v = (r == r0) ? T(1) : a[r];
        for (int j = 0; j < nrest; ++j) A[r + (col + 1 + j) * lda] -= v * sdot[j];
      }
      __syncthreads();
    }
  }
  // Gram matrix G = V^T V (unit lower trapezoidal V, rows j0..M-1).
  {
    for (int idx = tid; idx < jb * jb; idx += nthr) {
      const int i = idx % jb, j = idx / jb;
      G[i][j] = T(0);
    }
    __syncthreads();
    // each wave accumulates a subset of (i, j>i) pairs over all rows
    for (int pair = wave; pair < jb * jb; pair += nw) {
      const int i = pair % jb, j = pair / jb;
      if (j <= i) continue;
      T s = T(0);
      // v_i(r) for r >= j0+i with v_i(j0+i) = 1; v_j(r) zero for r < j0+j, one at j0+j
      for (int64_t r = j0 + j + lane; r < M; r += 64) {
        const T vi = A[r + (j0 + i) * lda];  // r > j0+i always here
        const T vj = (r == j0 + j) ? T(1) : A[r + (j0 + j) * lda];
        s += vi * vj;
      }
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) G[i][j] = s;
    }
    __syncthreads();
    // Tf(0:i, i) = -tau_i * Tf(0:i, 0:i) * G(0:i, i);  Tf(i,i) = tau_i
    if (tid == 0) {
      for (int i = 0; i < jb; ++i) {
        for (int r = 0; r < i; ++r) {
          T t = T(0);
          for (int q = r; q < i; ++q) t += Tf[r + q * QR_NB] * G[q][i];
          red[r] = t;
        }
        for (int r = 0; r < i; ++r) Tf[r + i * QR_NB] = -stau[i] * red[r];
        Tf[i + i * QR_NB] = stau[i];
        for (int r = i + 1; r < QR_NB; ++r) Tf[r + i * QR_NB] = T(0);
      }
      for (int i = jb; i < QR_NB; ++i)
        for (int r = 0; r < QR_NB; ++r) Tf[r + i * QR_NB] = T(0);
    }
  }
}

// Copies V (unit lower trapezoidal, rows j0..M-1, jb columns) into a dense workspace
// so the trailing update can use the plain GEMM kernel.
template <typename T>
__global__ void __launch_bounds__(256) qr_copyV_kernel(int64_t M, int64_t j0, int jb,
                                                       const T* __restrict__ A, int64_t lda,
                                                       int64_t sA, T* __restrict__ V, int64_t ldv,
                                                       int64_t sV) {
  const int64_t e = blockIdx.z;
  A += e * sA;
  V += e * sV;
  const int64_t r = j0 + int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (r >= M) return;
  for (int c = 0; c < jb; ++c) {
    const int64_t rr = r - j0;
    T v;
    if (rr < c) v = T(0);
    else if (rr == c) v = T(1);
    else v = A[r + (j0 + c) * lda];
    V[rr + c * ldv] = v;
  }
}

// Triangular solve helper: in-place inverse of an upper-triangular nb x nb diagonal block
// (used by the blocked back substitution for T = R11^-1 C).
template <typename T>
__global__ void __launch_bounds__(256) trinv_upper_kernel(int64_t p, int nb, const T* __restrict__ R,
                                                          int64_t ldr, int64_t sR,
                                                          T* __restrict__ W, int64_t sW) {
  __shared__ T S[QR_NB][QR_NB + 1];
  __shared__ T X[QR_NB][QR_NB + 1];
  R += int64_t(blockIdx.x) * sR;
  W += int64_t(blockIdx.x) * sW;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int i = idx % nb, j = idx / nb;
    S[i][j] = (i <= j) ? R[(p + i) + (p + j) * ldr] : T(0);
    X[i][j] = T(0);
  }
  __syncthreads();
  // column j of X = inverse: back substitution, one thread per column
  if (tid < nb) {
    const int j = tid;
    for (int i = j; i >= 0; --i) {
      T s = (i == j) ? T(1) : T(0);
      for (int t = i + 1; t <= j; ++t) s -= S[i][t] * X[t][j];
      X[i][j] = s / S[i][i];
    }
  }
  __syncthreads();
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int i = idx % nb, j = idx / nb;
    W[i + j * nb] = X[i][j];
  }
}

template <typename T>
void qr_panel(hipStream_t st, int64_t M, int64_t j0, int jb, T* A, int64_t lda, int64_t sA,
              T* Tf, int64_t sT, int64_t batch) {
  const double rows = double(M - j0);
  HSSC_TIMED(tname<T>("qr_panel_f64", "qr_panel_f32"), st, 4.0 * rows * jb * jb * batch,
             2.0 * rows * jb * sizeof(T) * batch,
             hipLaunchKernelGGL(qr_panel_kernel<T>, dim3(batch), dim3(1024), 0, st, M, j0, jb, A,
                                lda, sA, Tf, sT));
}

template <typename T>
void qr_copyV(hipStream_t st, int64_t M, int64_t j0, int jb, const T* A, int64_t lda, int64_t sA,
              T* V, int64_t ldv, int64_t sV, int64_t batch) {
  dim3 grid((M - j0 + 255) / 256, 1, batch);
  hipLaunchKernelGGL(qr_copyV_kernel<T>, grid, dim3(256), 0, st, M, j0, jb, A, lda, sA, V, ldv,
                     sV);
}

template <typename T>
void trinv_upper(hipStream_t st, int64_t p, int nb, const T* R, int64_t ldr, int64_t sR, T* W,
                 int64_t sW, int64_t batch) {
  hipLaunchKernelGGL(trinv_upper_kernel<T>, dim3(batch), dim3(256), 0, st, p, nb, R, ldr, sR, W,
                     sW);
}

// ------------------------------------------------------------------------------------
// Batched GEMV (solve sweeps of HSS2D_applyinv_*.m / HSS2D_apply_*.m, one RHS)
// ------------------------------------------------------------------------------------
// y = alpha * op(A) x + gamma * z   (z may alias nothing; gamma = 0 ignores it)
// Two phases for determinism: partial sums per column (or row) chunk, then a reduce.

constexpr int GV_ROWS = 256, GV_CHUNK = 512;

template <typename T>
__global__ void __launch_bounds__(256) gemv_n_partial(int64_t m, int64_t n, const T* __restrict__ A,
                                                      int64_t lda, int64_t sA,
                                                      const T* __restrict__ x, int64_t sx,
                                                      T* __restrict__ part, int64_t nsplit) {
  __shared__ T xs[GV_CHUNK];
  const int64_t e = blockIdx.z;
  A += e * sA;
  x += e * sx;
  const int64_t r = int64_t(blockIdx.x) * GV_ROWS + threadIdx.x;
  const int64_t c0 = int64_t(blockIdx.y) * GV_CHUNK;
  const int64_t c1 = (c0 + GV_CHUNK < n) ? c0 + GV_CHUNK : n;
  for (int64_t c = c0 + threadIdx.x; c < c1; c += 256) xs[c - c0] = x[c];
  __syncthreads();
  if (r >= m) return;
  T acc0 = T(0), acc1 = T(0);
  const T* a = A + r;
  int64_t c = c0;
  for (; c + 1 < c1; c += 2) {
    acc0 += a[c * lda] * xs[c - c0];
    acc1 += a[(c + 1) * lda] * xs[c + 1 - c0];
  }
  if (c < c1) acc0 += a[c * lda] * xs[c - c0];
  part[(e * nsplit + blockIdx.y) * m + r] = acc0 + acc1;
}

template <typename T>
__global__ void __launch_bounds__(256) gemv_t_partial(int64_t m, int64_t n, const T* __restrict__ A,
                                                      int64_t lda, int64_t sA,
                                                      const T* __restrict__ x, int64_t sx,
                                                      T* __restrict__ part, int64_t nsplit) {
  // y_j = sum_i A(i,j) x_i : one wave per column j, rows split in chunks along blockIdx.y
  const int64_t e = blockIdx.z;
  A += e * sA;
  x += e * sx;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t j = int64_t(blockIdx.x) * 4 + wave;
  const int64_t r0 = int64_t(blockIdx.y) * GV_CHUNK;
  const int64_t r1 = (r0 + GV_CHUNK < m) ? r0 + GV_CHUNK : m;
  if (j >= n) return;
  T acc = T(0);
  for (int64_t r = r0 + lane; r < r1; r += 64) acc += A[r + j * lda] * x[r];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) part[(e * nsplit + blockIdx.y) * n + j] = acc;
}

template <typename T>
__global__ void __launch_bounds__(256) gemv_reduce(int64_t len, int64_t nsplit,
                                                   const T* __restrict__ part, T alpha,
                                                   const T* __restrict__ z, int64_t sz, T gamma,
                                                   T* __restrict__ y, int64_t sy) {
  const int64_t e = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= len) return;
  T s = T(0);
  for (int64_t q = 0; q < nsplit; ++q) s += part[(e * nsplit + q) * len + i];
  T v = alpha * s;
  if (gamma != T(0)) v += gamma * z[e * sz + i];
  y[e * sy + i] = v;
}

template <typename T>
void gemv(hipStream_t st, bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
          int64_t sA, const T* x, int64_t sx, T gamma, const T* z, int64_t sz, T* y, int64_t sy,
          int64_t batch, T* work) {
  if (batch <= 0) return;
  const int64_t out = trans ? n : m;
  if (out <= 0) return;
  const int64_t inner = trans ? m : n;
  if (inner <= 0) {
    // y = gamma z
    hipLaunchKernelGGL(gemv_reduce<T>, dim3((out + 255) / 256, batch), dim3(256), 0, st, out,
                       int64_t(0), work, alpha, z, sz, gamma, y, sy);
    return;
  }
  const int64_t nsplit = (inner + GV_CHUNK - 1) / GV_CHUNK;
  const double fl = 2.0 * m * double(n) * batch;
  const double by = (double(m) * n + m + n) * sizeof(T) * batch;
  if (!trans) {
    dim3 grid((m + GV_ROWS - 1) / GV_ROWS, nsplit, batch);
    HSSC_TIMED(tname<T>("gemv_f64", "gemv_f32"), st, fl, by,
               hipLaunchKernelGGL(gemv_n_partial<T>, grid, dim3(256), 0, st, m, n, A, lda, sA, x,
                                  sx, work, nsplit));
  } else {
    dim3 grid((n + 3) / 4, nsplit, batch);
    HSSC_TIMED(tname<T>("gemvt_f64", "gemvt_f32"), st, fl, by,
               hipLaunchKernelGGL(gemv_t_partial<T>, grid, dim3(256), 0, st, m, n, A, lda, sA, x,
                                  sx, work, nsplit));
  }
  hipLaunchKernelGGL(gemv_reduce<T>, dim3((out + 255) / 256, batch), dim3(256), 0, st, out, nsplit,
                     work, alpha, z, sz, gamma, y, sy);
}

// ------------------------------------------------------------------------------------
// Gathers / scatters / small elementwise kernels for the tree sweeps
// ------------------------------------------------------------------------------------

// dst[e][a + c*ld_dst] = src[idx_e[a] + c*ld_src + e*s_src]  (a < n, c < m)
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) gather_kernel(int64_t n, int64_t m, const int32_t* __restrict__ idx,
                                                     int64_t sidx, const TS* __restrict__ src,
                                                     int64_t ld_src, int64_t s_src,
                                                     TD* __restrict__ dst, int64_t ld_dst,
                                                     int64_t s_dst) {
  const int64_t e = blockIdx.z, c = blockIdx.y;
  const int64_t a = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (a >= n) return;
  dst[e * s_dst + c * ld_dst + a] = TD(src[e * s_src + c * ld_src + idx[e * sidx + a]]);
}

template <typename TS, typename TD>
void gather(hipStream_t st, int64_t n, int64_t m, const int32_t* idx, int64_t sidx, const TS* src,
            int64_t ld_src, int64_t s_src, TD* dst, int64_t ld_dst, int64_t s_dst, int64_t batch) {
  if (n <= 0 || m <= 0 || batch <= 0) return;
  dim3 grid((n + 255) / 256, m, batch);
  hipLaunchKernelGGL((gather_kernel<TS, TD>), grid, dim3(256), 0, st, n, m, idx, sidx, src, ld_src,
                     s_src, dst, ld_dst, s_dst);
}

// dst[idx_e[a] + c*ld_dst + e*s_dst] = src[e][a + c*ld_src]  (+= when accumulate)
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) scatter_kernel(int64_t n, int64_t m, const int32_t* __restrict__ idx,
                                                      int64_t sidx, const TS* __restrict__ src,
                                                      int64_t ld_src, int64_t s_src,
                                                      TD* __restrict__ dst, int64_t ld_dst,
                                                      int64_t s_dst, int accumulate) {
  const int64_t e = blockIdx.z, c = blockIdx.y;
  const int64_t a = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (a >= n) return;
  TD* d = dst + e * s_dst + c * ld_dst + idx[e * sidx + a];
  const TD v = TD(src[e * s_src + c * ld_src + a]);
  *d = accumulate ? *d + v : v;
}

template <typename TS, typename TD>
void scatter(hipStream_t st, int64_t n, int64_t m, const int32_t* idx, int64_t sidx, const TS* src,
             int64_t ld_src, int64_t s_src, TD* dst, int64_t ld_dst, int64_t s_dst, int64_t batch,
             bool accumulate) {
  if (n <= 0 || m <= 0 || batch <= 0) return;
  dim3 grid((n + 255) / 256, m, batch);
  hipLaunchKernelGGL((scatter_kernel<TS, TD>), grid, dim3(256), 0, st, n, m, idx, sidx, src,
                     ld_src, s_src, dst, ld_dst, s_dst, accumulate ? 1 : 0);
}

// y = a*x + b*y over a flat range (with type conversion from x)
template <typename TX, typename TY>
__global__ void __launch_bounds__(256) axpby_kernel(int64_t len, TY a, const TX* __restrict__ x, TY b,
                                                    TY* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= len) return;
  const TY xv = TY(x[i]);
  y[i] = (b == TY(0)) ? a * xv : a * xv + b * y[i];
}

template <typename TX, typename TY>
void axpby(hipStream_t st, int64_t len, TY a, const TX* x, TY b, TY* y) {
  if (len <= 0) return;
  hipLaunchKernelGGL((axpby_kernel<TX, TY>), dim3((len + 255) / 256), dim3(256), 0, st, len, a, x,
                     b, y);
}

// Strided 2D copy with conversion: dst[e][i + j*ldd] = src[e][i + j*lds]
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) copy2d_kernel(int64_t m, int64_t n, const TS* __restrict__ src,
                                                     int64_t lds, int64_t ss, TD* __restrict__ dst,
                                                     int64_t ldd, int64_t sd) {
  const int64_t e = blockIdx.z;
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= m) return;
  dst[e * sd + i + j * ldd] = TD(src[e * ss + i + j * lds]);
}

template <typename TS, typename TD>
void copy2d(hipStream_t st, int64_t m, int64_t n, const TS* src, int64_t lds, int64_t ss, TD* dst,
            int64_t ldd, int64_t sd, int64_t batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((m + 255) / 256, n, batch);
  hipLaunchKernelGGL((copy2d_kernel<TS, TD>), grid, dim3(256), 0, st, m, n, src, lds, ss, dst, ldd,
                     sd);
}

// dst[e][i + j*ldd] = a * s1[e][i + j*ld1] + b * s2[e][i + j*ld2]   (b == 0 ignores s2)
template <typename T>
__global__ void __launch_bounds__(256) add2d_kernel(int64_t m, int64_t n, T a, const T* s1, int64_t ld1,
                                                    int64_t ss1, T b, const T* s2, int64_t ld2,
                                                    int64_t ss2, T* dst, int64_t ldd, int64_t sd) {
  const int64_t e = blockIdx.z;
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= m) return;
  T v = a * s1[e * ss1 + i + j * ld1];
  if (b != T(0)) v += b * s2[e * ss2 + i + j * ld2];
  dst[e * sd + i + j * ldd] = v;
}

template <typename T>
void add2d(hipStream_t st, int64_t m, int64_t n, T a, const T* s1, int64_t ld1, int64_t ss1, T b,
           const T* s2, int64_t ld2, int64_t ss2, T* dst, int64_t ldd, int64_t sd, int64_t batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((m + 255) / 256, n, batch);
  hipLaunchKernelGGL(add2d_kernel<T>, grid, dim3(256), 0, st, m, n, a, s1, ld1, ss1, b, s2, ld2, ss2,
                     dst, ldd, sd);
}

// Sum of squares (for residual norms), deterministic two-phase.
template <typename T>
__global__ void __launch_bounds__(256) sumsq_partial(int64_t len, const T* __restrict__ x,
                                                     double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += int64_t(gridDim.x) * 256) {
    const double v = double(x[i]);
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

template <typename T>
void sumsq(hipStream_t st, int64_t len, const T* x, double* part, int nblocks) {
  hipLaunchKernelGGL(sumsq_partial<T>, dim3(nblocks), dim3(256), 0, st, len, x, part);
}

// ------------------------------------------------------------------------------------
// Explicit instantiations
// ------------------------------------------------------------------------------------
#define HSSC_INST(T)                                                                            \
  template void gen_block<T>(hipStream_t, int64_t, int64_t, int64_t, const double*, const double*, \
                             const double*, int64_t, const double*, const double*, const double*,  \
                             int64_t, T*, int64_t, int64_t, double);                               \
  template void gen_parent<T>(hipStream_t, int64_t, int64_t, int64_t, const int32_t*, int64_t,     \
                              const double*, const double*, const double*, int64_t, const T*,      \
                              int64_t, int64_t, T*, int64_t, double);                              \
  template void gemm<T>(hipStream_t, bool, bool, int64_t, int64_t, int64_t, T, const T*, int64_t,  \
                        int64_t, const T*, int64_t, int64_t, T, T*, int64_t, int64_t, int64_t);    \
  template void gemm_skip<T>(hipStream_t, bool, bool, int64_t, int64_t, int64_t, T, const T*,    \
                             int64_t, int64_t, const T*, int64_t, int64_t, T, T*, int64_t, int64_t, \
                             int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t);        \
  template void invert_batched<T>(hipStream_t, int64_t, T*, int64_t, int64_t, int64_t);            \
  template void qr_panel<T>(hipStream_t, int64_t, int64_t, int, T*, int64_t, int64_t, T*, int64_t, \
                            int64_t);                                                              \
  template void qr_copyV<T>(hipStream_t, int64_t, int64_t, int, const T*, int64_t, int64_t, T*,    \
                            int64_t, int64_t, int64_t);                                            \
  template void trinv_upper<T>(hipStream_t, int64_t, int, const T*, int64_t, int64_t, T*, int64_t, \
                               int64_t);                                                           \
  template void gemv<T>(hipStream_t, bool, int64_t, int64_t, T, const T*, int64_t, int64_t,        \
                        const T*, int64_t, T, const T*, int64_t, T*, int64_t, int64_t, T*);        \
  template void sumsq<T>(hipStream_t, int64_t, const T*, double*, int);                           \
  template void add2d<T>(hipStream_t, int64_t, int64_t, T, const T*, int64_t, int64_t, T, const T*, \
                         int64_t, int64_t, T*, int64_t, int64_t, int64_t);

HSSC_INST(double)
HSSC_INST(float)

#define HSSC_INST2(TS, TD)                                                                        \
  template void gather<TS, TD>(hipStream_t, int64_t, int64_t, const int32_t*, int64_t, const TS*,  \
                               int64_t, int64_t, TD*, int64_t, int64_t, int64_t);                  \
  template void scatter<TS, TD>(hipStream_t, int64_t, int64_t, const int32_t*, int64_t, const TS*, \
                                int64_t, int64_t, TD*, int64_t, int64_t, int64_t, bool);           \
  template void axpby<TS, TD>(hipStream_t, int64_t, TD, const TS*, TD, TD*);                       \
  template void copy2d<TS, TD>(hipStream_t, int64_t, int64_t, const TS*, int64_t, int64_t, TD*,    \
                               int64_t, int64_t, int64_t);

HSSC_INST2(double, double)
HSSC_INST2(float, float)
HSSC_INST2(double, float)
HSSC_INST2(float, double)

}  // namespace k
}  // namespace hssc

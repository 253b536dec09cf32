// HIP kernels for the HSS compress / invert / apply path on CDNA4 (gfx950).
//
// All matrices are column-major.  Batched kernels take a batch index in blockIdx.z (or
// blockIdx.y) and element strides per batch; a stride of 0 shares one operand across the
// batch (translation-invariant levels).  Kernels are templated on the working scalar:
// double / float for the real kernels (FP64 path, FP32 factorization) and interleaved
// complex cplx<double> / cplx<float> for the Helmholtz kernel.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <limits>
#include <cstdint>
#include <string>

#include "kernels.h"
#include "stats.h"

namespace hssc {
namespace k {

template <typename T>
constexpr const char* type_suffix() {
  return is_complex_v<T> ? (sizeof(real_t<T>) == 8 ? "_c128" : "_c64")
                         : (sizeof(T) == 8 ? "_f64" : "_f32");
}
// Instrumentation label of a kernel class, e.g. "gemm_f64", "gemm_c64".
template <typename T>
std::string lbl(const char* base) {
  return std::string(base) + type_suffix<T>();
}

// ------------------------------------------------------------------------------------
// Kernel-block generation (Common/Kernel_Eval.m:29-96, 2D plane)
// ------------------------------------------------------------------------------------
// The kernel entries come from kernel_entry (kernels.h).

// One thread per entry; threadIdx.x runs along rows so column-major stores coalesce.
template <typename T>
__global__ void __launch_bounds__(256) gen_block_kernel(
    int64_t m, int64_t n, const double* __restrict__ Rx, const double* __restrict__ Ry,
    const double* __restrict__ Rw, int64_t sR, const double* __restrict__ Cx,
    const double* __restrict__ Cy, const double* __restrict__ Cw, int64_t sC, T* __restrict__ A,
    int64_t lda, int64_t sA, KernelSpec ks) {
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
    if (c < n) A[r + c * lda] = kernel_entry<T>(ks, xr, yr, wr, Cx[c], Cy[c], Cw[c]);
  }
}

template <typename T>
void gen_block(hipStream_t st, int64_t m, int64_t n, int64_t batch, const double* Rx,
               const double* Ry, const double* Rw, int64_t sR, const double* Cx, const double* Cy,
               const double* Cw, int64_t sC, T* A, int64_t lda, int64_t sA, const KernelSpec& ks) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid((m + 63) / 64, (n + 15) / 16, batch);
  const double ent = double(m) * n * batch;
  HSSC_TIMED(lbl<T>("gen_block").c_str(), st, ent, ent * sizeof(T),
             hipLaunchKernelGGL(gen_block_kernel<T>, grid, dim3(256), 0, st, m, n, Rx, Ry, Rw, sR,
                                Cx, Cy, Cw, sC, A, lda, sA, ks));
}

// Parent block F = [E1 K12; K12^T E2] in the parent's J order
// (HSS2D/HSS2D_build_Finv_fsym.m:93-101):  FJ(a,c) = F(J[a], J[c]).  Natural index
// i < kc belongs to child 1 (its skeleton position i), i >= kc to child 2.
template <typename T>
__global__ void __launch_bounds__(256) gen_parent_kernel(
    int64_t n, int64_t kc, const int32_t* __restrict__ J, int64_t sJ,
    const double* __restrict__ Px, const double* __restrict__ Py, const double* __restrict__ Pw,
    int64_t sP, const T* __restrict__ Ech, int64_t sEbox, int64_t sEkey_per_parent,
    T* __restrict__ F, int64_t sF, KernelSpec ks) {
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
      v = kernel_entry<T>(ks, xa, ya, wa, Px[c], Py[c], Pw[c]);
    }
    F[a + c * n] = v;
  }
}

template <typename T>
void gen_parent(hipStream_t st, int64_t n, int64_t kc, int64_t batch, const int32_t* J,
                int64_t sJ, const double* Px, const double* Py, const double* Pw, int64_t sP,
                const T* Ech, int64_t sEbox, int64_t sEparent, T* F, int64_t sF,
                const KernelSpec& ks) {
  dim3 grid((n + 63) / 64, (n + 15) / 16, batch);
  const double ent = double(n) * n * batch;
  HSSC_TIMED(lbl<T>("gen_parent").c_str(), st, ent, ent * sizeof(T),
             hipLaunchKernelGGL(gen_parent_kernel<T>, grid, dim3(256), 0, st, n, kc, J, sJ, Px, Py,
                                Pw, sP, Ech, sEbox, sEparent, F, sF, ks));
}

// ------------------------------------------------------------------------------------
// Batched GEMM on MFMA:  C = alpha * op(A) * op(B) + beta * C
// ------------------------------------------------------------------------------------
// 64x64 output tile per 256-thread workgroup, K staged through LDS in steps of 16.
// Each wave owns a 32x32 quarter = 2x2 MFMA 16x16 tiles.  The MFMA is issued as
// D = B^T A^T so that a lane's 4 accumulators run down one column of C with lane%16
// indexing rows: column-major stores of 16 consecutive rows coalesce.
// Complex operands are split into real and imaginary LDS planes; each k-step issues four
// real MFMAs per tile for complex double and three (Gauss's product) for complex float.

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
// CA conjugates op(A) (complex only: C = A^H B for the Householder updates).
// PRE reads the C tile (beta != 0) before the K loop, so that its latency hides behind the
// products: worth it for the short K of the rank-32 and rank-64 updates, where the loop runs
// two to four tiles; for longer K the extra registers cost more than the latency saved.
template <typename T, bool TA, bool CA, bool TB, bool PRE>
__global__ void __launch_bounds__(256) gemm_kernel(int64_t m, int64_t n, int64_t kdim, T alpha,
                                                  const T* A, int64_t lda, int64_t sA,
                                                  const T* B, int64_t ldb, int64_t sB,
                                                  T beta, T* C, int64_t ldc,
                                                  int64_t sC, int64_t rs0, int64_t rsn, int64_t cs0,
                                                  int64_t csn, int64_t ksplit, int64_t kchunk,
                                                  int tri) {
  using R = real_t<T>;
  constexpr bool CX = is_complex_v<T>;
  constexpr int NP = CX ? 2 : 1;  // LDS planes per operand
  using v4 = typename Mfma<R>::v4;
  // two buffers of As[p][kk][i] = plane p of op(A)(i0+i, k0+kk) and
  // Bs[p][kk][j] = plane p of op(B)(k0+kk, j0+j)
  __shared__ R As[2][NP][GB_K][GB_M + 1];
  __shared__ R Bs[2][NP][GB_K][GB_N + 1];
  // blockIdx.z = e * ksplit + q: batch element e, K chunk q (rows [q*kchunk, ...) of op(B)
  // and columns of op(A)); chunk q writes its own output slice C + (e*ksplit + q)*sC.
  const int64_t z = blockIdx.z;
  const int64_t e = z / ksplit, q = z % ksplit;
  A += e * sA;
  B += e * sB;
  C += z * sC;
  if (ksplit > 1) {
    const int64_t k0q = q * kchunk;
    A += TA ? k0q : k0q * lda;
    B += TB ? k0q * ldb : k0q;
    kdim = (kdim - k0q < kchunk) ? kdim - k0q : kchunk;
  }
  int64_t i0 = int64_t(blockIdx.x) * GB_M, j0 = int64_t(blockIdx.y) * GB_N;
  if (i0 >= rs0) i0 += rsn;
  if (j0 >= cs0) j0 += csn;
  if (i0 >= m || j0 >= n) return;
  if (tri && i0 > j0) return;  // upper tiles only (symmetric updates)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  // complex float: three real products per tile (Gauss): P0 = re re, P1 = im im,
  // P2 = (re + im)(re + im); Re = P0 - P1, Im = P2 - P0 - P1.  Complex double keeps the four
  // products (re re - im im, re im + im re): faster there, and the FP64 path's rounding stays
  // that of the plain product.
  constexpr bool G3 = CX && sizeof(R) == 4;
  constexpr int NACC = CX ? (G3 ? 3 : 2) : 1;
  v4 acc[NACC][2][2];
#pragma unroll
  for (int p = 0; p < NACC; ++p)
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) acc[p][x][y] = v4{0, 0, 0, 0};

  // Software pipeline: the next K tile is loaded into registers while the MFMAs run on the
  // current LDS buffer, then written to the other buffer (one barrier per tile).
  T ra[4], rb[4];
  auto gload = [&](int64_t k0) {
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
      ra[t] = (gi < m && gk < kdim) ? (TA ? A[gk + gi * lda] : A[gi + gk * lda]) : T(0);
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
      rb[t] = (gj < n && gk < kdim) ? (TB ? B[gj + gk * ldb] : B[gk + gj * ldb]) : T(0);
    }
  };
  auto lstore = [&](int bf) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int idx = tid + t * 256;
      const int ii = !TA ? (idx & 63) : (idx >> 4), kk = !TA ? (idx >> 6) : (idx & 15);
      As[bf][0][kk][ii] = re_s(ra[t]);
      if constexpr (CX) As[bf][NP - 1][kk][ii] = CA ? -im_s(ra[t]) : im_s(ra[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int idx = tid + t * 256;
      const int jj = !TB ? (idx >> 4) : (idx & 63), kk = !TB ? (idx & 15) : (idx >> 6);
      Bs[bf][0][kk][jj] = re_s(rb[t]);
      if constexpr (CX) Bs[bf][NP - 1][kk][jj] = im_s(rb[t]);
    }
  };
  gload(0);
  T cpre[PRE ? 2 : 1][PRE ? 2 : 1][PRE ? 4 : 1];
  if constexpr (PRE) {
    if (beta != T(0)) {
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t gi = i0 + wi + x * 16 + (lane & 15);
            const int64_t gj = j0 + wj + y * 16 + Mfma<R>::row(lane, r);
            cpre[x][y][r] = (gi < m && gj < n) ? C[gi + gj * ldc] : T(0);
          }
    }
  }
  lstore(0);
  __syncthreads();
  int cur = 0;
  for (int64_t k0 = 0; k0 < kdim; k0 += GB_K) {
    const bool more = k0 + GB_K < kdim;
    if (more) gload(k0 + GB_K);
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      const int kq = kk + (lane >> 4);
      R av[NP][2], bv[NP][2];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int x = 0; x < 2; ++x) av[p][x] = As[cur][p][kq][wi + x * 16 + (lane & 15)];
#pragma unroll
        for (int y = 0; y < 2; ++y) bv[p][y] = Bs[cur][p][kq][wj + y * 16 + (lane & 15)];
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[0][x][y] = Mfma<R>::mma(bv[0][y], av[0][x], acc[0][x][y]);
      if constexpr (G3) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
            acc[1][x][y] = Mfma<R>::mma(bv[NP - 1][y], av[NP - 1][x], acc[1][x][y]);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
            acc[NACC - 1][x][y] = Mfma<R>::mma(bv[0][y] + bv[NP - 1][y], av[0][x] + av[NP - 1][x],
                                               acc[NACC - 1][x][y]);
      } else if constexpr (CX) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            acc[0][x][y] = Mfma<R>::mma(-bv[NP - 1][y], av[NP - 1][x], acc[0][x][y]);
            acc[1][x][y] = Mfma<R>::mma(bv[NP - 1][y], av[0][x], acc[1][x][y]);
            acc[1][x][y] = Mfma<R>::mma(bv[0][y], av[NP - 1][x], acc[1][x][y]);
          }
      }
    }
    if (more) lstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
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
        const int64_t gj = j0 + wj + y * 16 + Mfma<R>::row(lane, r);
        if (gi < m && gj < n) {
          T* c = C + gi + gj * ldc;
          T a;
          if constexpr (G3)
            a = T(acc[0][x][y][r] - acc[1][x][y][r],
                  acc[NACC - 1][x][y][r] - acc[0][x][y][r] - acc[1][x][y][r]);
          else if constexpr (CX)
            a = T(acc[0][x][y][r], acc[1][x][y][r]);
          else
            a = acc[0][x][y][r];
          const T v = alpha * a;
          if constexpr (PRE) *c = (beta == T(0)) ? v : v + beta * cpre[x][y][r];
          else *c = (beta == T(0)) ? v : v + beta * *c;
        }
      }
}

template <typename T>
void gemm_skip_split(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
                     const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB,
                     T beta, T* C, int64_t ldc, int64_t sC, int64_t batch, int64_t rtiles,
                     int64_t ctiles, int64_t rs0, int64_t rsn, int64_t cs0, int64_t csn,
                     int64_t ksplit, int64_t kchunk, bool ca = false, int tri = 0);

template <typename T>
void gemm_skip(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
               const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta,
               T* C, int64_t ldc, int64_t sC, int64_t batch, int64_t rtiles, int64_t ctiles,
               int64_t rs0, int64_t rsn, int64_t cs0, int64_t csn, int tri) {
  gemm_skip_split<T>(st, ta, tb, m, n, kdim, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                     rtiles, ctiles, rs0, rsn, cs0, csn, 1, kdim, false, tri);
}

template <typename T>
void gemm_skip_split(hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t kdim, T alpha,
                     const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB,
                     T beta, T* C, int64_t ldc, int64_t sC, int64_t batch, int64_t rtiles,
                     int64_t ctiles, int64_t rs0, int64_t rsn, int64_t cs0, int64_t csn,
                     int64_t ksplit, int64_t kchunk, bool ca, int tri) {
  if (m <= 0 || n <= 0 || batch <= 0 || rtiles <= 0 || ctiles <= 0) return;
  dim3 grid(rtiles, ctiles, batch * ksplit);
  // algorithmic work: rows / cols actually produced (the skipped block excluded; about half
  // the tiles with tri); a complex multiply-add is 8 real flops
  const double rows = double(rs0 < m ? m - rsn : m), cols = double(cs0 < n ? n - csn : n);
  const double part = tri ? 0.5 : 1.0;
  const double fl =
      part * (is_complex_v<T> ? 8.0 : 2.0) * rows * cols * double(kdim) * double(batch);
  const double by =
      (rows * kdim + kdim * cols + part * 2 * rows * cols) * sizeof(T) * double(batch);
  const bool conjA = is_complex_v<T> && ca;
  // label per kernel variant: gemm_<type>_<op(A)><op(B)>, op = n, t, or h (conjugate transpose)
  const std::string name = lbl<T>("gemm") + "_" + (ta ? (conjA ? "h" : "t") : (conjA ? "c" : "n")) +
                           (tb ? "t" : "n");
  // the C tile read ahead for short K (at most four K tiles per output), single precision
  // only: with FP64 entries the extra registers halve the waves per SIMD
  constexpr bool kPre = sizeof(real_t<T>) == 4;
  const bool pre = kPre && beta != T(0) && (ksplit > 1 ? kchunk : kdim) <= 4 * GB_K;
#define HSSC_GEMM_LAUNCH(TA_, CA_, TB_)                                                          \
  HSSC_TIMED(name.c_str(), st, fl, by,                                                           \
             if (pre) hipLaunchKernelGGL((gemm_kernel<T, TA_, CA_, TB_, kPre>), grid, dim3(256),  \
                                         0, st, m, n, kdim, alpha, A, lda, sA, B, ldb, sB, beta,  \
                                         C, ldc, sC, rs0, rsn, cs0, csn, ksplit, kchunk, tri);    \
             else hipLaunchKernelGGL((gemm_kernel<T, TA_, CA_, TB_, false>), grid, dim3(256), 0,  \
                                     st, m, n, kdim, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, \
                                     sC, rs0, rsn, cs0, csn, ksplit, kchunk, tri))
  if (conjA) {
    if (ta && !tb) HSSC_GEMM_LAUNCH(true, true, false);
    else if (ta && tb) HSSC_GEMM_LAUNCH(true, true, true);
    else if (!tb) HSSC_GEMM_LAUNCH(false, true, false);
    else HSSC_GEMM_LAUNCH(false, true, true);
  } else {
    if (!ta && !tb) HSSC_GEMM_LAUNCH(false, false, false);
    else if (ta && !tb) HSSC_GEMM_LAUNCH(true, false, false);
    else if (!ta && tb) HSSC_GEMM_LAUNCH(false, false, true);
    else HSSC_GEMM_LAUNCH(true, false, true);
  }
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

template <typename T>
void gemm_h(hipStream_t st, int64_t m, int64_t n, int64_t kdim, T alpha, const T* A, int64_t lda,
            int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta, T* C, int64_t ldc, int64_t sC,
            int64_t batch) {
  const int64_t big = INT64_MAX;
  gemm_skip_split<T>(st, true, false, m, n, kdim, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                     batch, (m + GB_M - 1) / GB_M, (n + GB_N - 1) / GB_N, big, 0, big, 0, 1, kdim,
                     true);
}

// ------------------------------------------------------------------------------------
// Gauss-Jordan inversion (MATLAB inv in HSS2D_invert_fsym.m:108, build_Finv_fsym.m:101,
// build_E_fsym.m:51,64)
// ------------------------------------------------------------------------------------
// In-place GJ sweep with partial (row) pivoting: after pivot c, A(c,c) = 1/p,
// A(c,j) /= p, A(i,j) -= A(i,c)_old * A(c,j), A(i,c) = -A(i,c)_old / p.  Row swaps are
// undone at the end by permuting columns.  Pivot magnitude |re| + |im| (LAPACK izamax).
//
// Register-resident Gauss-Jordan for n <= 64*S (n <= 4*CB): 256 threads, lanes are rows
// (row = lane + 64*s), wave w owns columns [w*CB, (w+1)*CB).  Per pivot column c:
//   * the wave owning column c finds the pivot row by a shuffle argmax over rows >= c and
//     publishes it, with the old column c, through double-buffered LDS;
//   * one barrier;
//   * every wave reads the old rows c and p of its own columns by lane shuffles, applies
//     the swap logically and performs the sweep on its registers.
// Column unscrambling (swaps undone in reverse order) is folded into the final store.
// Singularity: a pivot that is not finite, or whose magnitude is at most n * eps times the
// block's largest entry (no correct digit left in the inverse), sets *flag (when given).
template <typename T, int S, int CB>
__global__ void __launch_bounds__(256) gj_reg_kernel(int n, T* __restrict__ A, int64_t lda,
                                                     int64_t sA, int64_t off, int* flag) {
  using R = real_t<T>;
  __shared__ T colbuf[2][64 * S];
  __shared__ int spiv[2];
  __shared__ int pivs[64 * S];
  __shared__ int dest[64 * S];
  __shared__ R wmax[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  A += int64_t(blockIdx.x) * sA + off;
  T v[S][CB];
  R amax = 0;
#pragma unroll
  for (int s = 0; s < S; ++s)
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      const int i = lane + 64 * s, j = wave * CB + t;
      v[s][t] = (i < n && j < n) ? A[i + int64_t(j) * lda] : T(0);
      amax = fmax(amax, abs1(v[s][t]));
    }
  for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o));
  if (lane == 0) wmax[wave] = amax;
  // (wmax is read after the first barrier of the sweep)
  const R tiny_rel = R(n) * std::numeric_limits<R>::epsilon();
  auto publish = [&](int c) {
    // called by the wave owning column c.  Candidates are ranked by a 64-bit key: the
    // single-precision magnitude (non-negative floats order like their bit patterns) above
    // the complemented row, so the largest magnitude wins and ties go to the lowest row.
    const int tc = c - wave * CB;
    unsigned long long best = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      T x = T(0);
#pragma unroll
      for (int t = 0; t < CB; ++t)
        if (t == tc) x = v[s][t];
      const int i = lane + 64 * s;
      if (i < n) colbuf[c & 1][i] = x;
      if (i >= c && i < n) {
        const unsigned long long key =
            (static_cast<unsigned long long>(__float_as_uint(float(abs1(x)))) << 32) |
            static_cast<unsigned>(0xffffffffu - unsigned(i));
        if (key > best) best = key;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long ob = __shfl_xor(best, o);
      if (ob > best) best = ob;
    }
    const int bi = int(0xffffffffu - unsigned(best & 0xffffffffu));
    if (lane == 0) spiv[c & 1] = bi, pivs[c] = bi;
  };
  if (wave == 0) publish(0);
  for (int c = 0; c < n; ++c) {
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(spiv[c & 1]);
    const T* cb = colbuf[c & 1];
    const T piv = cb[p];
    const T inv = T(1) / piv;
    if (flag && threadIdx.x == 0) {
      const R a = abs1(piv);
      const R bmax = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
      if (!(a > tiny_rel * bmax) || !(a <= std::numeric_limits<R>::max())) atomicOr(flag, 1);
    }
    // old rows c and p of this wave's columns
    T rc[CB], rp[CB];
    const int sc = c >> 6, lc = c & 63, sp = p >> 6, lp = p & 63;
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      T xc = v[0][t], xp = v[0][t];
#pragma unroll
      for (int s = 1; s < S; ++s) {
        if (s == sc) xc = v[s][t];
        if (s == sp) xp = v[s][t];
      }
      rc[t] = readlane_s(xc, lc);
      rp[t] = readlane_s(xp, lp);
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int i = lane + 64 * s;
      const T f = (i == c) ? cb[p] : (i == p) ? cb[c] : cb[i < n ? i : 0];
#pragma unroll
      for (int t = 0; t < CB; ++t) {
        const int j = wave * CB + t;
        T old = (i == c) ? rp[t] : (i == p) ? rc[t] : v[s][t];
        T nv;
        if (i == c) nv = (j == c) ? inv : rp[t] * inv;
        else nv = (j == c) ? -(f * inv) : old - f * (rp[t] * inv);
        v[s][t] = nv;
      }
    }
    if (c + 1 < n && wave == (c + 1) / CB) publish(c + 1);
  }
  __syncthreads();
  // final column j of the inverse is column sigma(j) of the swept matrix, where sigma
  // applies the swaps (c, pivs[c]) for c = n-1 .. 0; each column is stored at
  // dest = sigma^-1(its index).
  if (threadIdx.x == 0) {
    int* sig = dest;
    for (int j = 0; j < n; ++j) sig[j] = j;
    for (int c = n - 1; c >= 0; --c) {
      const int q = pivs[c];
      const int a = sig[c];
      sig[c] = sig[q];
      sig[q] = a;
    }
    for (int j = 0; j < n; ++j) pivs[sig[j]] = j;
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s)
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      const int i = lane + 64 * s, j = wave * CB + t;
      if (i < n && j < n) A[i + int64_t(pivs[j]) * lda] = v[s][t];
    }
}

// Reciprocal of a pivot: the hardware estimate (v_rcp) refined by Newton's iteration
// r <- r + r (1 - x r), once in single precision, twice in double (each step squares the
// relative error); complex: the exponent of max(|re|, |im|) is scaled out exactly first, so
// |z|^2 neither overflows nor underflows.  Zero, infinite or NaN pivots give non-finite
// results (flagged by the sweep).
__device__ inline float rcp_nr(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ inline double recip_s(double x) { return rcp_nr(x); }
__device__ inline float recip_s(float x) { return rcp_nr(x); }
template <typename R>
__device__ inline cplx<R> recip_s(const cplx<R>& z) {
  int e;
  (void)frexp(fmax(fabs(z.re), fabs(z.im)), &e);
  const R a = ldexp(z.re, -e), b = ldexp(z.im, -e);
  const R r = rcp_nr(a * a + b * b);
  return cplx<R>(ldexp(a * r, -e), ldexp(-b * r, -e));
}
__device__ inline double fnma_s(double g, double b, double c) { return fma(-g, b, c); }
__device__ inline float fnma_s(float g, float b, float c) { return fmaf(-g, b, c); }
template <typename R>
__device__ inline cplx<R> fnma_s(const cplx<R>& g, const cplx<R>& b, const cplx<R>& c) {
  return cplx<R>(fma(-g.re, b.re, fma(g.im, b.im, c.re)), fma(-g.re, b.im, fma(-g.im, b.re, c.im)));
}

// Largest value over the wave: DPP within each row of 16 lanes (pairs, quads, the two quads
// of a group of eight through the half-row mirror, the two halves of the row through the row
// mirror), then the four rows through v_readlane.
__device__ inline float wave_max_f(float v) {
  v = fmaxf(v, dpp_move<0xB1>(v));
  v = fmaxf(v, dpp_move<0x4E>(v));
  v = fmaxf(v, dpp_move<0x141>(v));
  v = fmaxf(v, dpp_move<0x140>(v));
  const float r0 = readlane_s(v, 0), r1 = readlane_s(v, 16), r2 = readlane_s(v, 32),
              r3 = readlane_s(v, 48);
  return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// Gauss-Jordan inversion of an n x n block (n <= 64) per workgroup, in registers and without
// row interchanges.  Lanes are rows; wave w of NW holds columns [CB w, CB w + CB),
// CB = 64 / NW.  Step c pivots column c on the unused row p of largest |re| + |im| (LAPACK
// izamax's magnitude, compared in single precision; ties to the lowest row) and sweeps: row p
// becomes (row p) / a_pc with 1 / a_pc in column c, every other row i subtracts
// (a_ic / a_pc) (row p) and takes -a_ic / a_pc in column c.  Once every column is swept, row
// i of the result is row c(i) of the inverse with its columns in pivot-row order:
// inv(c(i), p_j) = B(i, j), c(i) the column pivoted on row i and p_j the pivot row of column
// j; the stores apply the permutation.
// Per step: one barrier; the wave that owns column c + 1 updates it first, writes it to LDS
// and finds its pivot row (a DPP wave max and a ballot) and the pivot's reciprocal, which
// the others read right after the barrier; the pivot row's entries come from lane p through
// scalar registers.  Singularity: a pivot that is not finite, or whose magnitude is at most
// n eps times the block's largest entry, sets *flag (when given).
template <typename T, int NW>
__global__ void __launch_bounds__(NW * 64) gj64_kernel(int n, T* __restrict__ A, int64_t lda,
                                                       int64_t sA, int64_t off, int* flag) {
  constexpr int CB = 64 / NW;  // columns per wave
  using R = real_t<T>;
  __shared__ T colbuf[2][64];
  __shared__ int spiv[2];
  __shared__ T sinv[2];
  __shared__ int prow_of[64];  // pivot row of column c
  __shared__ int col_of[64];   // column pivoted on row i
  __shared__ R wmax[NW];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  A += int64_t(blockIdx.x) * sA + off;
  T v[CB];
  R amax = R(0);
#pragma unroll
  for (int t = 0; t < CB; ++t) {
    const int j = wave * CB + t;
    v[t] = (lane < n && j < n) ? A[lane + int64_t(j) * lda] : T(0);
    amax = fmax(amax, abs1(v[t]));
  }
  for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o));
  if (lane == 0) wmax[wave] = amax;
  bool used = lane >= n;  // rows past the block never pivot
  const R tiny_rel = R(n) * std::numeric_limits<R>::epsilon();
  // the entry of column c (uniform c) of this lane's row: a switch keeps v in registers
  auto pick = [&](int c) {
    T x = v[CB - 1];
    switch (c % CB) {
#define HSSC_GJ_CASE(t) \
  case t:               \
    if (t < CB) x = v[t < CB ? t : 0]; \
    break;
      HSSC_GJ_CASE(0) HSSC_GJ_CASE(1) HSSC_GJ_CASE(2) HSSC_GJ_CASE(3) HSSC_GJ_CASE(4)
      HSSC_GJ_CASE(5) HSSC_GJ_CASE(6) HSSC_GJ_CASE(7) HSSC_GJ_CASE(8) HSSC_GJ_CASE(9)
      HSSC_GJ_CASE(10) HSSC_GJ_CASE(11) HSSC_GJ_CASE(12) HSSC_GJ_CASE(13) HSSC_GJ_CASE(14)
#undef HSSC_GJ_CASE
      default: break;
    }
    return x;
  };
  // column x into LDS, its pivot row and the pivot's reciprocal (by the wave owning the
  // column, so that the others read both right after the barrier)
  auto search = [&](const T& x, bool done, int buf) {
    colbuf[buf][lane] = x;
    const float mag = done ? -1.0f : float(abs1(x));
    const float mx = wave_max_f(mag);
    unsigned long long ball = __ballot(!done && mag == mx);
    if (!ball) ball = __ballot(!done);  // every candidate is NaN: flagged below
    const int p = __builtin_ctzll(ball);
    const T piv = readlane_s(x, p);
    if (lane == 0) {
      spiv[buf] = p;
      sinv[buf] = recip_s(piv);
      const R a = abs1(piv);
      R bmax = wmax[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) bmax = fmax(bmax, wmax[w]);
      if (flag && (!(a > tiny_rel * bmax) || !(a <= std::numeric_limits<R>::max())))
        atomicOr(flag, 1);
    }
  };
  __syncthreads();  // wmax
  if (wave == 0) search(pick(0), used, 0);
  for (int c = 0; c < n; ++c) {
    const int buf = c & 1;
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(spiv[buf]);
    const T inv = sinv[buf];
    const T f = colbuf[buf][lane];
    if (threadIdx.x == 0) prow_of[c] = p, col_of[p] = c;
    const bool isp = lane == p;
    // row i: v - g (row p) with g = a_ic / a_pc; row p: 0 - (-1 / a_pc) (row p)
    const T g = isp ? -inv : f * inv;
    // the next column first (the same formula as the sweep below, so the same bits), and
    // its pivot search, which the next step waits for
    if (c + 1 < n && wave == (c + 1) / CB) {
      const T vo = pick(c + 1);
      search(fnma_s(g, readlane_s(vo, p), isp ? T(0) : vo), used || isp, buf ^ 1);
    }
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      const T pr = readlane_s(v[t], p);
      v[t] = fnma_s(g, pr, isp ? T(0) : v[t]);
    }
    if (wave == c / CB) {  // column c: -a_ic / a_pc, and 1 / a_pc on row p (= -g there)
      switch (c % CB) {
#define HSSC_GJ_CASE(t) \
  case t:               \
    if (t < CB) v[t < CB ? t : 0] = -g; \
    break;
        HSSC_GJ_CASE(0) HSSC_GJ_CASE(1) HSSC_GJ_CASE(2) HSSC_GJ_CASE(3) HSSC_GJ_CASE(4)
        HSSC_GJ_CASE(5) HSSC_GJ_CASE(6) HSSC_GJ_CASE(7) HSSC_GJ_CASE(8) HSSC_GJ_CASE(9)
        HSSC_GJ_CASE(10) HSSC_GJ_CASE(11) HSSC_GJ_CASE(12) HSSC_GJ_CASE(13) HSSC_GJ_CASE(14)
#undef HSSC_GJ_CASE
        default: v[CB - 1] = -g;
      }
    }
    used = used || isp;
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < CB; ++t) {
    const int j = wave * CB + t;
    if (lane < n && j < n) A[col_of[lane] + int64_t(prow_of[j]) * lda] = v[t];
  }
}

// Waves per block of the register Gauss-Jordan kernel (8 columns each).  One 64 x 64 block:
// 62 / 52 / 47 us (complex64) and 75 / 63 / 62 us (complex128) with 4 / 8 / 16 waves; 4096
// blocks in one launch: 448 / 430 / 524 us and 700 / 673 / 776 us.
constexpr int kGjWaves = 8;

template <typename T>
void gj_reg(hipStream_t st, int n, T* A, int64_t lda, int64_t sA, int64_t off, int64_t batch,
            const std::string& label, int* flag) {
  const double fl = (is_complex_v<T> ? 8.0 : 2.0) * n * double(n) * n * batch;
  if (n <= 64) {
    HSSC_TIMED(label.c_str(), st, fl, 2.0 * n * n * sizeof(T) * batch,
               hipLaunchKernelGGL((gj64_kernel<T, kGjWaves>), dim3(batch), dim3(kGjWaves * 64), 0,
                                  st, n, A, lda, sA, off, flag));
  } else if constexpr (!is_complex_v<T>) {
    HSSC_TIMED(label.c_str(), st, fl, 2.0 * n * n * sizeof(T) * batch,
               hipLaunchKernelGGL((gj_reg_kernel<T, 2, 32>), dim3(batch), dim3(256), 0, st, n, A,
                                  lda, sA, off, flag));
  }
}

// Largest n one register-resident kernel inverts: 128 for real entries, 64 for complex ones.
// The blocked sweep uses blocks of that size.  (Complex float would fit 128 in registers, but
// on the Helmholtz F blocks the sweep with 128 blocks, pivoting inside each block only, left
// the FP32 inverse visibly less accurate: eight refinement steps instead of four.)
template <typename T>
constexpr int gj_reg_max() {
  return is_complex_v<T> ? 64 : 128;
}

// Blocked GJ without pivoting across diagonal blocks (pivoting inside each nb x nb block).
// Step p with block rows/cols B = [p, p+nb), nb = gj_reg_max (128, or 64 for complex double):
//   (1) A_BB <- inv(A_BB)                                   gj_reg on the block
//   (2) W_X  <- -A_XB * A_BB          (rows X outside B)    gemm_skip into workspace
//   (3) A_XY <- A_XY + W_X * A_BY     (X, Y outside B)      gemm_skip
//       A_XB <- W_X
//   (4) W_Y  <- A_BB * A_BY           (cols Y outside B)    gemm_skip into workspace
//       A_BY <- W_Y
// which is the block sweep A11' = A11^-1, A12' = A11^-1 A12, A21' = -A21 A11^-1,
// A22' = A22 - A21 A11^-1 A12.  With 128 blocks, steps (2) and (4) go through the workspace
// because an output tile (64 wide) covers only part of the block its inputs span; 64 blocks
// work in place.  Rank-128 updates keep the trailing GEMM, which streams all of A once per
// step, at half the traffic per flop of rank-64 ones.
template <typename T>
void gj_blocked(hipStream_t st, hipStream_t aux, int64_t n, T* A, int64_t lda, int64_t sA,
                int64_t batch, T* work, int* flag) {
  // p advances in multiples of 64 = the GEMM tile, so no 64-row tile straddles block B; the
  // tile-skip ranges of gemm_skip enumerate the tiles of [0,p) U [p+nb,n).
  // Look-ahead (when an aux stream is given): step (3) first updates the row block B' of
  // the next step; the inverse of A_B'B' then runs on the aux stream while the rest of (3)
  // and (4) run on st.  Neither touches A_B'B'; step p+1 waits for the inverse.
  constexpr int NB = gj_reg_max<T>();
  static_assert(NB % GB_M == 0 && GB_M == GB_N, "blocks of whole GEMM tiles");
  const std::string diag = lbl<T>("gj_diag");
  constexpr bool in_place = NB == GB_M;  // one-tile blocks need no workspace
  const int64_t sW = 2 * n * NB;
  T* Wx = work;            // n x nb, ld n (rows X)
  T* Wy = work + n * NB;   // nb x n, ld NB (cols Y)
  hipEvent_t upd = nullptr, inv = nullptr;
  if (aux) {
    if (hipEventCreateWithFlags(&upd, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&inv, hipEventDisableTiming) != hipSuccess) {
      if (upd) (void)hipEventDestroy(upd);
      upd = inv = nullptr;
      aux = nullptr;
    }
  }
  const int64_t big = INT64_MAX;
  gj_reg<T>(st, int(n < NB ? n : NB), A, lda, sA, 0, batch, diag, flag);
  for (int64_t p = 0; p < n; p += NB) {
    const int nb = int(p + NB <= n ? NB : n - p);
    const int64_t nout = n - nb;
    if (p > 0 && aux) (void)hipStreamWaitEvent(st, inv, 0);  // inverse of A_BB (aux stream)
    if (nout <= 0) continue;
    const int64_t tiles = (nout + GB_M - 1) / GB_M, btiles = (nb + GB_M - 1) / GB_M;
    T* Apb = A + p;             // rows B
    T* Abp = A + p * lda;       // cols B
    T* Abb = A + p + p * lda;   // block B
    // (2) W_X = -A_XB * A_BB (in place when the block is one tile wide: each output tile is
    // then produced by one workgroup that reads all of its inputs before it stores)
    T* X = in_place ? Abp : Wx;
    const int64_t ldx = in_place ? lda : n, sX = in_place ? sA : sW;
    gemm_skip<T>(st, false, false, n, nb, nb, T(-1), Abp, lda, sA, Abb, lda, sA, T(0), X, ldx, sX,
                 batch, tiles, btiles, p, nb, big, 0);
    // (3) A_XY += W_X * A_BY, the next row block B' = [p+nb, p+nb+nb2) first
    const int64_t q = p + nb;
    const int nb2 = int(q < n ? (q + NB <= n ? NB : n - q) : 0);
    if (nb2 > 0) {
      gemm_skip<T>(st, false, false, q + nb2, n, nb, T(1), X, ldx, sX, Apb, lda, sA, T(1), A, lda,
                   sA, batch, (nb2 + GB_M - 1) / GB_M, tiles, 0, q, p, nb);
      if (aux) {
        (void)hipEventRecord(upd, st);
        (void)hipStreamWaitEvent(aux, upd, 0);
        gj_reg<T>(aux, nb2, A, lda, sA, q + q * lda, batch, diag, flag);
        (void)hipEventRecord(inv, aux);
      }
    }
    const int64_t rest = nout - nb2;  // rows [0,p) U [q+nb2,n)
    if (rest > 0)
      gemm_skip<T>(st, false, false, n, n, nb, T(1), X, ldx, sX, Apb, lda, sA, T(1), A, lda, sA,
                   batch, (rest + GB_M - 1) / GB_M, tiles, p, nb + nb2, p, nb);
    if (in_place) {
      // (4) A_BY = A_BB * A_BY
      gemm_skip<T>(st, false, false, nb, n, nb, T(1), Abb, lda, sA, Apb, lda, sA, T(0), Apb, lda,
                   sA, batch, btiles, tiles, big, 0, p, nb);
    } else {
      copy2d<T, T>(st, p, nb, Wx, n, sW, Abp, lda, sA, batch);
      copy2d<T, T>(st, n - q, nb, Wx + q, n, sW, Abp + q, lda, sA, batch);
      // (4) W_Y = A_BB * A_BY
      gemm_skip<T>(st, false, false, nb, n, nb, T(1), Abb, lda, sA, Apb, lda, sA, T(0), Wy, NB,
                   sW, batch, btiles, tiles, big, 0, p, nb);
      copy2d<T, T>(st, nb, p, Wy, NB, sW, Apb, lda, sA, batch);
      copy2d<T, T>(st, nb, n - q, Wy + q * NB, NB, sW, Apb + q * lda, lda, sA, batch);
    }
    if (nb2 > 0 && !aux) gj_reg<T>(st, nb2, A, lda, sA, q + q * lda, batch, diag, flag);
  }
  if (upd) (void)hipEventDestroy(upd);
  if (inv) (void)hipEventDestroy(inv);
}

// A <- (A + A^T) / 2 in place (32 x 32 tile pairs through LDS).  A matrix that is symmetric
// only up to rounding (E^-1 = Phi_ss + T Phi_rs + (Phi_sr + T Phi_rr) T^T forms its two
// off-diagonal halves by different products) is replaced by its symmetric part before the
// upper-tile sweep: that drops the antisymmetric rounding to first order, as the reference's
// symmetrised inverse (Ai + Ai.')/2 does (HSS2D_build_E_fsym.m:255-259).
template <typename T>
__global__ void __launch_bounds__(256) sym_average_kernel(int64_t n, T* __restrict__ A,
                                                          int64_t lda, int64_t sA) {
  __shared__ T X[32][33];
  __shared__ T Y[32][33];
  const int64_t ti = blockIdx.x, tj = blockIdx.y;
  if (ti > tj) return;
  A += int64_t(blockIdx.z) * sA;
  const int64_t i0 = ti * 32, j0 = tj * 32;
  for (int idx = threadIdx.x; idx < 32 * 32; idx += 256) {
    const int r = idx % 32, c = idx / 32;
    if (i0 + r < n && j0 + c < n) X[r][c] = A[(i0 + r) + (j0 + c) * lda];
    if (j0 + r < n && i0 + c < n) Y[r][c] = A[(j0 + r) + (i0 + c) * lda];
  }
  __syncthreads();
  const T half = T(real_t<T>(0.5));
  for (int idx = threadIdx.x; idx < 32 * 32; idx += 256) {
    const int r = idx % 32, c = idx / 32;  // A(i0 + r, j0 + c) and A(j0 + c, i0 + r)
    if (i0 + r < n && j0 + c < n) {
      const T v = (X[r][c] + Y[c][r]) * half;
      A[(i0 + r) + (j0 + c) * lda] = v;
      A[(j0 + c) + (i0 + r) * lda] = v;
    }
  }
}

// A(i, j) = A(j, i) below the diagonal tiles (64 x 64 tiles, row tile > column tile), through
// an LDS tile: the lower triangle of a symmetric matrix kept in its upper tiles.
template <typename T>
__global__ void __launch_bounds__(256) sym_lower_fill_kernel(int64_t n, T* __restrict__ A,
                                                             int64_t lda, int64_t sA) {
  __shared__ T tile[64][65];
  const int64_t ti = blockIdx.x, tj = blockIdx.y;  // destination tile
  if (ti <= tj) return;
  A += int64_t(blockIdx.z) * sA;
  const int64_t i0 = ti * 64, j0 = tj * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int r = idx % 64, c = idx / 64;  // source A(j0 + r, i0 + c)
    if (j0 + r < n && i0 + c < n) tile[r][c] = A[(j0 + r) + (i0 + c) * lda];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int r = idx % 64, c = idx / 64;  // destination A(i0 + r, j0 + c)
    if (i0 + r < n && j0 + c < n) A[(i0 + r) + (j0 + c) * lda] = tile[c][r];
  }
}

template <typename T>
void sym_lower_fill(hipStream_t st, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch) {
  if (n <= 0 || batch <= 0) return;
  const int64_t nt = (n + 63) / 64;
  hipLaunchKernelGGL(sym_lower_fill_kernel<T>, dim3(nt, nt, batch), dim3(256), 0, st, n, A, lda,
                     sA);
}

// Blocked GJ of a symmetric matrix (complex symmetric: A^T = A, no conjugate) on its upper
// tiles.  With P the blocks already swept and U the others, the swept matrix keeps
// M_PP = M_PP^T, M_UU = M_UU^T and M_UP = -M_PU^T (A11^-1 and the Schur complement are
// symmetric; A21' = -A21 A11^-1 = -(A11^-1 A12)^T).  So every operand of step B can be read
// from the upper tiles, and only they are updated: for X, Y outside B, X <= Y,
//   X, Y < B:  A_XY += W_X (-A_YB^T)        W_X = -A_XB A_BB^-1   (new A_XB)
//   X < B < Y: A_XY += W_X A_BY
//   B < X, Y:  A_XY += (-A_BX^T) W_Y        W_Y = A_BB^-1 A_BY     (new A_BY)
// with the old A_YB, A_BX, A_BY (W_X and W_Y are written back afterwards).  Half the trailing
// tiles of the general sweep, which streams the whole matrix once per block; the lower
// triangle is filled at the end.  Diagonal tiles are updated whole, so the 64 x 64 diagonal
// blocks inverted by gj_reg are complete.
// NB = 128 sweeps two 64-column blocks at once: the 128 x 128 diagonal block is inverted by
// the same sweep with NB = 64 (so the pivots are chosen inside the same 64 x 64 blocks as
// there; sweeping B1 and then B2 equals sweeping B1 u B2), and the rest of the matrix takes
// rank-128 updates: half the steps, launches and trailing-matrix traffic of the 64-wide sweep.
template <typename T, int NB = GB_M>
void gj_blocked_sym(hipStream_t st, hipStream_t aux, int64_t n, T* A, int64_t lda, int64_t sA,
                    int64_t batch, T* work, int* flag) {
  static_assert(NB == GB_M || NB == 2 * GB_M, "blocks of one or two GEMM tiles");
  const std::string diag = lbl<T>("gj_diag");
  const int64_t sW = 2 * n * NB;
  T* Wx = work;           // rows X < B: p x nb, ld n
  T* Wy = work + n * NB;  // cols Y > B: nb x (n - q), ld NB
  T* Wd = work + sW * batch;  // NB = 128: workspace of the diagonal blocks' own sweeps
  hipEvent_t upd = nullptr, inv = nullptr;
  if (aux) {
    if (hipEventCreateWithFlags(&upd, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&inv, hipEventDisableTiming) != hipSuccess) {
      if (upd) (void)hipEventDestroy(upd);
      upd = inv = nullptr;
      aux = nullptr;
    }
  }
  const int64_t big = INT64_MAX;
  auto tl = [](int64_t x) { return (x + GB_M - 1) / GB_M; };
  // the pivoted inverse of a diagonal block is symmetric only to rounding; the sweep reads it
  // on both sides (A_BB^-1 in the new A_XB and, through the upper tiles, A_BB^-T), so it is
  // replaced by its symmetric part
  auto diag_inverse = [&](hipStream_t s, int nbk, int64_t off) {
    if (nbk > GB_M) {  // NB = 128: the 64-wide sweep of the block, on this stream
      if constexpr (NB > GB_M)
        gj_blocked_sym<T, GB_M>(s, nullptr, nbk, A + off, lda, sA, batch, Wd, flag);
      return;
    }
    gj_reg<T>(s, nbk, A, lda, sA, off, batch, diag, flag);
    hipLaunchKernelGGL(sym_average_kernel<T>, dim3((nbk + 31) / 32, (nbk + 31) / 32, batch),
                       dim3(256), 0, s, int64_t(nbk), A + off, lda, sA);
  };
  diag_inverse(st, int(n < NB ? n : NB), 0);
  for (int64_t p = 0; p < n; p += NB) {
    const int nb = int(p + NB <= n ? NB : n - p);
    const int64_t q = p + nb, nr = n - q;
    if (p > 0 && aux) (void)hipStreamWaitEvent(st, inv, 0);  // inverse of A_BB (aux stream)
    if (n - nb <= 0) continue;
    T* Abb = A + p + p * lda;
    T* Abq = A + p + q * lda;  // A(B, q:n)
    if (p > 0)
      gemm_skip<T>(st, false, false, p, nb, nb, T(-1), A + p * lda, lda, sA, Abb, lda, sA, T(0),
                   Wx, n, sW, batch, tl(p), tl(nb), big, 0, big, 0);
    if (nr > 0)
      gemm_skip<T>(st, false, false, nb, nr, nb, T(1), Abb, lda, sA, Abq, lda, sA, T(0), Wy, NB,
                   sW, batch, tl(nb), tl(nr), big, 0, big, 0);
    // B < X <= Y: the next block row B' first, so that its inverse can start on aux
    const int nb2 = int(q < n ? (q + NB <= n ? NB : n - q) : 0);
    if (nb2 > 0) {
      gemm_skip<T>(st, true, false, nb2, nr, nb, T(-1), Abq, lda, sA, Wy, NB, sW, T(1),
                   A + q + q * lda, lda, sA, batch, tl(nb2), tl(nr), big, 0, big, 0, 1);
      if (aux) {
        (void)hipEventRecord(upd, st);
        (void)hipStreamWaitEvent(aux, upd, 0);
        diag_inverse(aux, nb2, q + q * lda);
        (void)hipEventRecord(inv, aux);
      }
    }
    const int64_t q2 = q + nb2, nr2 = n - q2;
    if (nr2 > 0)
      gemm_skip<T>(st, true, false, nr2, nr2, nb, T(-1), A + p + q2 * lda, lda, sA,
                   Wy + nb2 * NB, NB, sW, T(1), A + q2 + q2 * lda, lda, sA, batch, tl(nr2),
                   tl(nr2), big, 0, big, 0, 1);
    if (p > 0) {
      // X <= Y < B
      gemm_skip<T>(st, false, true, p, p, nb, T(-1), Wx, n, sW, A + p * lda, lda, sA, T(1), A,
                   lda, sA, batch, tl(p), tl(p), big, 0, big, 0, 1);
      // X < B < Y
      if (nr > 0)
        gemm_skip<T>(st, false, false, p, nr, nb, T(1), Wx, n, sW, Abq, lda, sA, T(1),
                     A + q * lda, lda, sA, batch, tl(p), tl(nr), big, 0, big, 0);
      copy2d<T, T>(st, p, nb, Wx, n, sW, A + p * lda, lda, sA, batch);
    }
    if (nr > 0) copy2d<T, T>(st, nb, nr, Wy, NB, sW, Abq, lda, sA, batch);
    if (nb2 > 0 && !aux) diag_inverse(st, nb2, q + q * lda);
  }
  const int64_t nt = tl(n);
  hipLaunchKernelGGL(sym_lower_fill_kernel<T>, dim3(nt, nt, batch), dim3(256), 0, st, n, A, lda,
                     sA);
  // the diagonal tiles were updated whole, their two halves by different products: the
  // symmetric part makes the inverse exactly symmetric (the filled tiles are left unchanged,
  // (x + x) / 2 = x)
  const int64_t nt32 = (n + 31) / 32;
  hipLaunchKernelGGL(sym_average_kernel<T>, dim3(nt32, nt32, batch), dim3(256), 0, st, n, A, lda,
                     sA);
  if (upd) (void)hipEventDestroy(upd);
  if (inv) (void)hipEventDestroy(inv);
}

// The symmetric sweep serves complex matrices: 64-wide blocks (the GEMM tile), 128-wide ones
// from kWideSweep columns on, where the rank-128 updates of the large matrices pay for the
// diagonal blocks' own sweeps; real matrices use the general sweep with 128-wide blocks.
constexpr int64_t kWideSweep = 2048;
template <typename T>
int64_t invert_ws_elems(int64_t n, int64_t batch) {
  if (n <= gj_reg_max<T>()) return 0;
  if (is_complex_v<T> && n >= kWideSweep)
    return (2 * n * 2 * GB_M + 2 * 2 * GB_M * GB_M) * batch;
  return 2 * n * gj_reg_max<T>() * batch;
}

template <typename T>
void invert_batched(hipStream_t st, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch,
                    hipStream_t aux, T* work, bool sym, int* flag) {
  if (n <= 0 || batch <= 0) return;
  if (n <= gj_reg_max<T>()) {
    gj_reg<T>(st, int(n), A, lda, sA, 0, batch, lbl<T>("gj_small"), flag);
    if (sym && is_complex_v<T>) {  // the pivoted inverse is symmetric only to rounding
      const int64_t nt = (n + 31) / 32;
      hipLaunchKernelGGL(sym_average_kernel<T>, dim3(nt, nt, batch), dim3(256), 0, st, n, A, lda,
                         sA);
    }
    return;
  }
  if (sym && is_complex_v<T>) {
    const int64_t nt = (n + 31) / 32;
    hipLaunchKernelGGL(sym_average_kernel<T>, dim3(nt, nt, batch), dim3(256), 0, st, n, A, lda,
                       sA);
    if (n >= kWideSweep) gj_blocked_sym<T, 2 * GB_M>(st, aux, n, A, lda, sA, batch, work, flag);
    else gj_blocked_sym<T, GB_M>(st, aux, n, A, lda, sA, batch, work, flag);
  } else {
    gj_blocked<T>(st, aux, n, A, lda, sA, batch, work, flag);
  }
}

// ------------------------------------------------------------------------------------
// Householder QR least squares (MATLAB K11\K21 in HSS2D/interpolation_operator.m:66)
// ------------------------------------------------------------------------------------
// The interpolation least squares run as blocked Householder QR: each 32-column panel is
// factored by the row-parallel CholeskyQR3 kernel below and brought back to Householder form
// (V unit lower trapezoidal, the compact-WY factor Tf with H_0 ... H_{jb-1} = I - V Tf V^H,
// as LAPACK's xgeqrf / xlarft produce it, up to unit-modulus scalings of the rows of R); the
// trailing updates are GEMMs.

constexpr int QR_NB = 32;  // panel width (the trailing update runs as GEMMs)

// Triangular solve helper: inverses of the upper-triangular diagonal blocks of R (used by the
// blocked back substitution for T = R11^-1 C).  Workgroup (e, q) inverts the block of rows and
// columns [p, p + nb), p = p0 + 32 q, nb = min(32, pend - p), of matrix e into
// W + e sW + q 32^2 (nb x nb, ld nb).
template <typename T>
__global__ void __launch_bounds__(256) trinv_upper_kernel(int64_t p0, int64_t pend,
                                                          const T* __restrict__ R, int64_t ldr,
                                                          int64_t sR, T* __restrict__ W,
                                                          int64_t sW) {
  __shared__ T S[QR_NB][QR_NB + 1];
  __shared__ T X[QR_NB][QR_NB + 1];
  const int64_t p = p0 + int64_t(blockIdx.y) * QR_NB;
  const int nb = int(pend - p < QR_NB ? pend - p : QR_NB);
  R += int64_t(blockIdx.x) * sR;
  W += int64_t(blockIdx.x) * sW + int64_t(blockIdx.y) * QR_NB * QR_NB;
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
void trinv_upper(hipStream_t st, int64_t p, int nb, const T* R, int64_t ldr, int64_t sR, T* W,
                 int64_t sW, int64_t batch) {
  hipLaunchKernelGGL(trinv_upper_kernel<T>, dim3(batch, 1), dim3(256), 0, st, p, p + nb, R, ldr,
                     sR, W, sW);
}

template <typename T>
void trinv_upper_all(hipStream_t st, int64_t kk, const T* R, int64_t ldr, int64_t sR, T* W,
                     int64_t sW, int64_t batch) {
  if (kk <= 0) return;
  hipLaunchKernelGGL(trinv_upper_kernel<T>, dim3(batch, (kk + QR_NB - 1) / QR_NB), dim3(256), 0,
                     st, int64_t(0), kk, R, ldr, sR, W, sW);
}

// ------------------------------------------------------------------------------------
// Panels: shifted CholeskyQR3 and Householder reconstruction
// ------------------------------------------------------------------------------------
// A 32-column panel (Mp = M - j0 rows) is factored by shifted CholeskyQR3, which keeps every
// step row-parallel (a one-workgroup Householder panel streams the panel through L2 once per
// column on one CU, a latency chain of 32 reductions): (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput. 42,
// 2020): Q_1 = A R_1^-1, Q_2 = Q_1 R_2^-1, Q = Q_2 R_3^-1 with R_p the Cholesky factor of
// the Gram matrix of the previous Q (the first one shifted by 11 (Mp jb + jb (jb + 1)) u
// ||A||_F^2), R = R_3 R_2 R_1.  The orthonormal Q is turned back into Householder form
// (Ballard, Demmel, Grigori, Jacquelin, Knight, Nguyen, IPDPS 2014): the LU factorisation
// Q_top - S = L U, no pivoting, with the unit-modulus diagonal S chosen during the
// elimination (s_k = -c / |c| for the pivot candidate c, so |pivot| >= 1), gives
// V = [L; Q_below U^-1], Tf = -U S^H L^-H and the panel's R rows S R, with
// I - V Tf V^H = H_0 ... H_{jb-1} unitary and its first jb columns Q S^H: what the
// Householder panel produces up to unit-modulus row scalings of R, so the trailing update
// and the back substitution are shared.
//
// Four launches of one kernel, 256 rows per workgroup.  Pass p > 0 first forms the
// Cholesky factor of the summed partial Gram matrices of pass p - 1 (every workgroup
// redundantly: 32 x 32, one barrier per elimination step), then applies it to its rows by
// forward substitution; passes 0-2 leave the partial Gram matrices of their rows for the
// next pass (two alternating buffers).  Workgroup 0 accumulates R and, in pass 3, writes Tf,
// the R rows of A and the top rows of V.

constexpr int CQR_ROWS = kCqrRows;                     // rows per workgroup
constexpr int CQR_G = QR_NB * QR_NB;                    // one NB x NB block
constexpr int CQR_NE = CQR_G / CQR_ROWS;                // NB x NB entries per thread
static_assert(QR_NB == kQrNb, "panel width");
static_assert(CQR_ROWS == 256 && CQR_G % CQR_ROWS == 0, "entry split");

// workspace offsets (per matrix): Racc | Qtop | partial Gram A | partial Gram B | Q_1; the
// second partial set starts at part1 = CQR_PART0 + cqr_row_blocks(M) * CQR_G
constexpr int64_t CQR_RACC = 0, CQR_QTOP = CQR_G, CQR_PART0 = 2 * CQR_G;

template <typename T, typename XT, typename QT>
__global__ void __launch_bounds__(CQR_ROWS, 2) cqr_pass_kernel(
    int pass, int64_t Mp, int jb, int P, int64_t part1, const XT* X, int64_t ldx, int64_t sX,
    QT* Qo, int64_t ldq, int64_t sQ, wide_t<T>* __restrict__ ws, int64_t sW, T* V, int64_t ldv,
    int64_t sV, T* __restrict__ A, int64_t lda, int64_t sA, int64_t j0, T* __restrict__ Tfo,
    int64_t sT) {
  using C = wide_t<T>;  // arithmetic type: FP64 whatever the storage precision
  using R = real_t<C>;
  __shared__ C Rm[QR_NB][QR_NB + 1];       // Cholesky factor of this pass
  __shared__ C Wb[2 * QR_NB][QR_NB + 1];   // Gram staging (64 rows); pass 3: LU | scratch
  __shared__ C sdiag[QR_NB];               // S of the reconstruction
  __shared__ R rdi[QR_NB];                 // 1 / R_cc
  __shared__ R rdg[QR_NB];                 // R_cc
  __shared__ C udi[QR_NB];                 // 1 / U_cc
  const int64_t e = blockIdx.y;
  const int blk = blockIdx.x, tid = threadIdx.x;
  const bool wg0 = blk == 0;
  ws += e * sW;
  X += e * sX;
  const C* pin = ws + ((pass & 1) ? CQR_PART0 : part1);   // partial Grams of pass - 1
  C* pout = ws + ((pass & 1) ? part1 : CQR_PART0);        // partial Grams of this pass
  if (pass > 0) {
    // R = chol(sum of the partial Gram matrices (+ shift)); the partials are loaded eight at
    // a time and summed in block order
    C sum[CQR_NE];
#pragma unroll
    for (int m = 0; m < CQR_NE; ++m) sum[m] = C(0);
    for (int q0 = 0; q0 < P; q0 += 8) {
      C v[8][CQR_NE];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int m = 0; m < CQR_NE; ++m)
          v[u][m] = q0 + u < P ? pin[int64_t(q0 + u) * CQR_G + tid + m * CQR_ROWS] : C(0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int m = 0; m < CQR_NE; ++m) sum[m] += v[u][m];
    }
#pragma unroll
    for (int m = 0; m < CQR_NE; ++m) {
      const int ent = tid + m * CQR_ROWS;
      Rm[ent % QR_NB][ent / QR_NB] = sum[m];
    }
    __syncthreads();
    if (pass == 1) {
      R tr = R(0);
      for (int i = 0; i < jb; ++i) tr += re_s(Rm[i][i]);
      const R u = std::numeric_limits<R>::epsilon() / 2;
      const R shift = R(11) * (R(Mp) * R(jb) + R(jb) * R(jb + 1)) * u * tr;
      __syncthreads();
      if (tid < jb) Rm[tid][tid] += C(shift);
      __syncthreads();
    }
    // Elimination on the unscaled rows: step k subtracts conj(U_ki) U_kj / U_kk from the
    // trailing entries (U_kk = R_kk^2 is final from step k on); one barrier per step, and
    // each thread's operands are read together before its updates.  The rows are scaled
    // by 1 / R_kk afterwards.
    for (int k = 0; k < jb; ++k) {
      const R d2 = re_s(Rm[k][k]);
      C ua[CQR_NE], ub[CQR_NE], uw[CQR_NE];
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
        ua[m] = Rm[k][i], ub[m] = Rm[k][j], uw[m] = Rm[i][j];
      }
      const R inv2 = d2 > R(0) ? R(1) / d2 : std::numeric_limits<R>::infinity();
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
        if (i > k && i <= j && j < jb) Rm[i][j] = uw[m] - scale_s(conj_s(ua[m]) * ub[m], inv2);
      }
      __syncthreads();
    }
    // R_kk, and 1 / R_kk (0 past the panel: the row solves below run over all QR_NB
    // columns without branches, the padded columns of x and R being zero)
    if (tid < QR_NB) {
      R d = R(0), rd = R(0);
      if (tid < jb) {
        const R d2 = re_s(Rm[tid][tid]);
        d = d2 > R(0) ? sqrt(d2) : std::numeric_limits<R>::min();
        rd = R(1) / d;
      }
      rdg[tid] = d, rdi[tid] = rd;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CQR_NE; ++m) {
      const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
      if (i <= j && j < jb) Rm[i][j] = i == j ? C(rdg[i]) : scale_s(Rm[i][j], rdi[i]);
    }
    __syncthreads();
    if (wg0) {
      // Racc <- R Racc (Racc = R_1 after pass 1)
      C* Racc = ws + CQR_RACC;
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
        Wb[QR_NB + i][j] = pass == 1 ? (i == j ? C(1) : C(0)) : Racc[ent];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        // lanes run along j here (Rm[i][t] read by half a wave at once, Wb rows contiguous);
        // blocks of eight terms loaded together (ascending t, as one sum over [i, j])
        const int e2 = tid + m * CQR_ROWS, i = e2 / QR_NB, j = e2 % QR_NB;
        const int ent = i + j * QR_NB;
        C sum = C(0);
        if (i <= j && j < jb) {
          for (int t0 = i & ~7; t0 <= j; t0 += 8) {
            C av[8], bv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) av[u] = Rm[i][t0 + u], bv[u] = Wb[QR_NB + t0 + u][j];
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (t0 + u >= i && t0 + u <= j) sum += av[u] * bv[u];
          }
        }
        Racc[ent] = sum;
      }
    }
    __syncthreads();
  }
  if (pass == 3) {
    // Q_top = Q2_top R^-1 (rows by forward substitution), then LU of Q_top - S
    const C* Qtop = ws + CQR_QTOP;
#pragma unroll
    for (int m = 0; m < CQR_NE; ++m) {
      const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
      Wb[i][j] = (i < jb && j < jb) ? Qtop[ent] : C(0);
    }
    __syncthreads();
    {
      // forward substitution of row i on 8 threads (t = s + 8 u, u < 4): a 32-step chain of
      // short partial sums and a 3-step lane reduction
      const int i = tid >> 3, sl = tid & 7;
      C w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) w[u] = Wb[i][sl + 8 * u];
      for (int c = 0; c < jb; ++c) {
        C part = C(0);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (sl + 8 * u < c) part += w[u] * Rm[sl + 8 * u][c];
        part = sum8_lanes(part);
        if (sl == (c & 7)) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (u == (c >> 3)) w[u] = (w[u] - part) * C(rdi[c]);
        }
      }
      if (i < jb) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (sl + 8 * u < jb) Wb[i][sl + 8 * u] = w[u];
      }
    }
    __syncthreads();
    // Step k eliminates with the unscaled column k; the columns of L (column k times
    // 1 / U_kk) and the pivots U_kk = c_k - s_k are written afterwards.  |U_kk| = |c_k| + 1.
    for (int k = 0; k < jb; ++k) {
      const C c = Wb[k][k];
      C la[CQR_NE], lb[CQR_NE], lw[CQR_NE];
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
        la[m] = Wb[i][k], lb[m] = Wb[k][j], lw[m] = Wb[i][j];
      }
      const R a = sqrt(abs2(c));
      const C sk = a > R(0) ? scale_s(c, R(-1) / a) : C(-1);
      const C piv = c - sk;
      const C pinv = scale_s(conj_s(piv), R(1) / abs2(piv));
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
        if (i > k && j > k && i < jb && j < jb) Wb[i][j] = lw[m] - la[m] * pinv * lb[m];
      }
      if (tid == k) sdiag[k] = sk, udi[k] = pinv;
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < CQR_NE; ++m) {
      const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
      if (i < jb && j < jb) {
        if (i > j) Wb[i][j] = Wb[i][j] * udi[j];
        else if (i == j) Wb[i][i] = Wb[i][i] - sdiag[i];
      }
    }
    if (tid >= jb && tid < QR_NB) udi[tid] = C(0);
    __syncthreads();
  }
  // rows of this workgroup: x <- x R^-1 (pass 3: then x <- x U^-1)
  const int64_t r = int64_t(blk) * CQR_ROWS + tid;
  const bool live = r < Mp && !(pass == 3 && r < jb);
  C x[QR_NB];
#pragma unroll
  for (int c = 0; c < QR_NB; ++c) x[c] = (live && c < jb) ? C(X[r + int64_t(c) * ldx]) : C(0);
  // Row substitutions over all QR_NB columns, with no branches (one basic block)
  if (pass > 0) {
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) {
      C sum = x[c];
#pragma unroll
      for (int i = 0; i < c; ++i) sum -= x[i] * Rm[i][c];
      x[c] = scale_s(sum, rdi[c]);
    }
  }
  if (pass == 3) {
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) {
      C sum = x[c];
#pragma unroll
      for (int i = 0; i < c; ++i) sum -= x[i] * Wb[i][c];
      x[c] = sum * udi[c];
    }
  }
  if (Qo && live) {
    Qo += e * sQ;
#pragma unroll
    for (int c = 0; c < QR_NB; ++c)
      if (c < jb) Qo[r + int64_t(c) * ldq] = QT(x[c]);
  }
  if (pass == 2 && wg0 && tid < jb) {
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) ws[CQR_QTOP + tid + c * QR_NB] = c < jb ? x[c] : C(0);
  }
  if (pass < 3) {
    // Partial Gram matrix X^H X of this workgroup's rows on MFMA.  The rows are staged
    // through LDS 64 at a time as real and imaginary planes (in Wb's storage); wave w
    // accumulates the 16 x 16 tile (I0, J0) = 16 (w & 1, w >> 1).  Operand a of lane l is
    // conj X(t, I0 + l%16), operand b is X(t, J0 + l%16), t = 4 kq + l/16; item r of lane l
    // is G(I0 + row(l, r), J0 + l%16).
    using v4 = typename Mfma<R>::v4;
    constexpr bool CX = is_complex_v<C>;
    constexpr int LDP = QR_NB + 1;
    R* Sr = reinterpret_cast<R*>(&Wb[0][0]);
    R* Si = Sr + 64 * LDP;
    const int wave = tid >> 6, lane = tid & 63;
    const int I0 = (wave & 1) * 16, J0 = (wave >> 1) * 16, lc = lane & 15, lk = lane >> 4;
    v4 gr = {0, 0, 0, 0}, gi = {0, 0, 0, 0};
    for (int w = 0; w < CQR_ROWS / 64; ++w) {
      __syncthreads();
      if (wave == w) {
#pragma unroll
        for (int c = 0; c < QR_NB; ++c) {
          Sr[lane * LDP + c] = re_s(x[c]);
          if constexpr (CX) Si[lane * LDP + c] = im_s(x[c]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int kq = 0; kq < 16; ++kq) {
        const int t = 4 * kq + lk;
        const R ar = Sr[t * LDP + I0 + lc], br = Sr[t * LDP + J0 + lc];
        gr = Mfma<R>::mma(ar, br, gr);
        if constexpr (CX) {
          const R ai = Si[t * LDP + I0 + lc], bi = Si[t * LDP + J0 + lc];
          gr = Mfma<R>::mma(ai, bi, gr);
          gi = Mfma<R>::mma(ar, bi, gi);
          gi = Mfma<R>::mma(-ai, br, gi);
        }
      }
    }
    C* pp = pout + int64_t(blk) * CQR_G;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = I0 + Mfma<R>::row(lane, r), j = J0 + lc;
      if constexpr (CX) pp[i + j * QR_NB] = C(gr[r], gi[r]);
      else pp[i + j * QR_NB] = C(gr[r]);
    }
    return;
  }
  if (!wg0) return;
  // pass 3, workgroup 0: Tf = -U S^H L^-H, the panel's R rows S Racc, the top rows of V = L
  C* Racc = ws + CQR_RACC;
  Tfo += e * sT;
  A += e * sA;
  V += e * sV;
#pragma unroll
  for (int m = 0; m < CQR_NE; ++m) {
    const int ent = tid + m * CQR_ROWS, i = ent % QR_NB, j = ent / QR_NB;
    if (i < jb && j < jb) {
      A[(j0 + i) + (j0 + j) * lda] = T(i <= j ? sdiag[i] * Racc[ent] : C(0));
      V[i + int64_t(j) * ldv] = T(i > j ? Wb[i][j] : (i == j ? C(1) : C(0)));
    }
  }
  {
    // row i of Tf: Tf L^H = B with B = -U S^H (upper); unit upper L^H, ascending columns,
    // on 8 threads per row (t = s + 8 u) as the forward substitution above
    const int i = tid >> 3, sl = tid & 7;
    C tr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) tr[u] = C(0);
    for (int j = 0; j < QR_NB; ++j) {
      C part = C(0);
      if (j < jb) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (sl + 8 * u < j) part += tr[u] * conj_s(Wb[j][sl + 8 * u]);
      }
      part = sum8_lanes(part);
      if (sl == (j & 7)) {
        const C b = (i <= j && j < jb) ? -(Wb[i][j] * conj_s(sdiag[j])) : C(0);
        const C v = (i < jb && j < jb) ? b - part : C(0);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (u == (j >> 3)) tr[u] = v;
        Tfo[i + j * QR_NB] = T(v);
      }
    }
  }
}

template <typename T>
void cqr_panel(hipStream_t st, int64_t M, int64_t j0, int jb, T* A, int64_t lda, int64_t sA,
               T* V, int64_t ldv, int64_t sV, T* Tf, int64_t sT, wide_t<T>* ws, int64_t batch) {
  using C = wide_t<T>;
  const int64_t Mp = M - j0, sW = cqr_ws_elems(M), Pm = cqr_row_blocks(M);
  if (Mp <= 0 || jb <= 0 || batch <= 0) return;
  const int P = int((Mp + CQR_ROWS - 1) / CQR_ROWS);
  const int64_t part1 = CQR_PART0 + Pm * CQR_G;
  C* Q1 = ws + part1 + Pm * CQR_G;
  const T* Ap = A + j0 + j0 * lda;
  const dim3 g(P, batch), b(CQR_ROWS);
  const double fl = (is_complex_v<T> ? 4.0 : 1.0) * 2.0 * 5.0 * double(Mp) * jb * jb * batch;
  const double by = 7.0 * double(Mp) * jb * sizeof(T) * batch;
  HSSC_TIMED(lbl<T>("cqr_panel").c_str(), st, fl, by, {
    // pass 0: Gram of the panel; 1: Q_1 -> workspace; 2: Q_2 -> V; 3: V <- Householder form
    hipLaunchKernelGGL((cqr_pass_kernel<T, T, T>), g, b, 0, st, 0, Mp, jb, P, part1, Ap, lda, sA,
                       (T*)nullptr, int64_t(0), int64_t(0), ws, sW, V, ldv, sV, A, lda, sA, j0, Tf,
                       sT);
    hipLaunchKernelGGL((cqr_pass_kernel<T, T, C>), g, b, 0, st, 1, Mp, jb, P, part1, Ap, lda, sA,
                       Q1, Mp, sW, ws, sW, V, ldv, sV, A, lda, sA, j0, Tf, sT);
    hipLaunchKernelGGL((cqr_pass_kernel<T, C, T>), g, b, 0, st, 2, Mp, jb, P, part1,
                       (const C*)Q1, Mp, sW, V, ldv, sV, ws, sW, V, ldv, sV, A, lda, sA, j0, Tf, sT);
    hipLaunchKernelGGL((cqr_pass_kernel<T, T, T>), g, b, 0, st, 3, Mp, jb, P, part1,
                       (const T*)V, ldv, sV, V, ldv, sV, ws, sW, V, ldv, sV, A, lda, sA, j0, Tf, sT);
  });
}

// Triangular factor of a product of block reflectors: H = H_0 H_1 ... H_{ns-1}, H_s =
// I - V_s Ts_s V_s^H with V = [V_0 ... V_{ns-1}] (n columns, blocks of 32) gives
// H = I - V T V^H with T upper block triangular, diagonal blocks Ts_s and
// T(0:c0, c0:c0+w) = -T(0:c0, 0:c0) (V^H V)(0:c0, c0:c0+w) Ts_s for the block at c0 = 32 s
// (LAPACK xlarft's recurrence, one block at a time).  One workgroup per matrix; X is scratch
// of n * 32 elements per matrix.
template <typename T>
__global__ void __launch_bounds__(256) t_assemble_kernel(int n, const T* __restrict__ Ts,
                                                         int64_t sTs, const T* __restrict__ GV,
                                                         int64_t ldg, int64_t sG,
                                                         T* __restrict__ To, int64_t ldt,
                                                         int64_t sT, T* __restrict__ X,
                                                         int64_t sX) {
  __shared__ T Ta[QR_NB][QR_NB + 1];
  __shared__ T Tb[QR_NB][QR_NB + 1];
  const int64_t e = blockIdx.x;
  const int tid = threadIdx.x, ti = tid % QR_NB, tj = tid / QR_NB;  // entries (ti, tj + 8 m)
  Ts += e * sTs;
  GV += e * sG;
  To += e * sT;
  X += e * sX;
  const int ns = (n + QR_NB - 1) / QR_NB;
  // diagonal blocks Ts_s, zeros below
  for (int idx = tid; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;
    To[i + int64_t(j) * ldt] =
        (i / QR_NB == j / QR_NB) ? Ts[(i / QR_NB) * CQR_G + (i % QR_NB) + (j % QR_NB) * QR_NB]
                                 : T(0);
  }
  for (int sb = 1; sb < ns; ++sb) {
    const int c0 = sb * QR_NB, w = n - c0 < QR_NB ? n - c0 : QR_NB;
    // Y_q = GV(q block, c0 block) Ts_sb for q < sb, into X (32 x 32 per q)
    __syncthreads();
    for (int idx = tid; idx < CQR_G; idx += 256) {
      const int i = idx % QR_NB, j = idx / QR_NB;
      Ta[i][j] = (i < w && j < w) ? Ts[sb * CQR_G + idx] : T(0);
    }
    for (int q = 0; q < sb; ++q) {
      for (int idx = tid; idx < CQR_G; idx += 256) {
        const int i = idx % QR_NB, j = idx / QR_NB;
        Tb[i][j] = j < w ? GV[(q * QR_NB + i) + int64_t(c0 + j) * ldg] : T(0);
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int j = tj + 8 * m;
        T sum = T(0);
        for (int t = 0; t < QR_NB; ++t) sum += Tb[ti][t] * Ta[t][j];
        X[q * CQR_G + ti + j * QR_NB] = sum;
      }
      __syncthreads();
    }
    // T(r block, c0 block) = -sum_{q = r}^{sb - 1} T(r, q) Y_q
    for (int r = 0; r < sb; ++r) {
      T acc[CQR_NE];
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) acc[m] = T(0);
      for (int q = r; q < sb; ++q) {
        for (int idx = tid; idx < CQR_G; idx += 256) {
          const int i = idx % QR_NB, j = idx / QR_NB;
          Tb[i][j] = To[(r * QR_NB + i) + int64_t(q * QR_NB + j) * ldt];
          Ta[i][j] = X[q * CQR_G + idx];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < CQR_NE; ++m) {
          const int j = tj + 8 * m;
          for (int t = 0; t < QR_NB; ++t) acc[m] += Tb[ti][t] * Ta[t][j];
        }
        __syncthreads();
      }
#pragma unroll
      for (int m = 0; m < CQR_NE; ++m) {
        const int j = tj + 8 * m;
        if (j < w) To[(r * QR_NB + ti) + int64_t(c0 + j) * ldt] = -acc[m];
      }
    }
  }
}

template <typename T>
void t_assemble(hipStream_t st, int n, const T* Ts, int64_t sTs, const T* GV, int64_t ldg,
                int64_t sG, T* To, int64_t ldt, int64_t sT, T* X, int64_t sX, int64_t batch) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(t_assemble_kernel<T>, dim3(batch), dim3(256), 0, st, n, Ts, sTs, GV, ldg, sG,
                     To, ldt, sT, X, sX);
}

template <typename T>
__global__ void __launch_bounds__(256) zero2d_kernel(int64_t m, int64_t n, T* A, int64_t lda,
                                                     int64_t sA) {
  const int64_t e = blockIdx.z;
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x, j = blockIdx.y;
  if (i < m) A[e * sA + i + j * lda] = T(0);
}

template <typename T>
void zero2d(hipStream_t st, int64_t m, int64_t n, T* A, int64_t lda, int64_t sA, int64_t batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(zero2d_kernel<T>, dim3((m + 255) / 256, n, batch), dim3(256), 0, st, m, n, A,
                     lda, sA);
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
  // eight column loads issued before their products (memory-level parallelism per thread)
  T acc0 = T(0), acc1 = T(0);
  const T* a = A + r;
  int64_t c = c0;
  for (; c + 7 < c1; c += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = a[(c + u) * lda];
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      acc0 += v[u] * xs[c + u - c0];
      acc1 += v[u + 1] * xs[c + u + 1 - c0];
    }
  }
  for (; c < c1; ++c) acc0 += a[c * lda] * xs[c - c0];
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
  // the chunk's eight loads of this lane issued before their products
  const T* a = A + j * lda;
  T acc = T(0), acc1 = T(0);
  int64_t r = r0 + lane;
  if constexpr (sizeof(T) == 8) {
    // 8-byte elements (complex64): each lane takes two adjacent rows in one 16-byte load
    using V2 = float4;
    if (r0 + GV_CHUNK <= m && ((reinterpret_cast<uintptr_t>(a + r0) |
                                reinterpret_cast<uintptr_t>(x + r0)) & 15) == 0) {
      const V2* av = reinterpret_cast<const V2*>(a + r0) + lane;
      const V2* xv = reinterpret_cast<const V2*>(x + r0) + lane;
      V2 v[GV_CHUNK / 128];
#pragma unroll
      for (int u = 0; u < GV_CHUNK / 128; ++u) v[u] = av[u * 64];
#pragma unroll
      for (int u = 0; u < GV_CHUNK / 128; ++u) {
        const V2 w = xv[u * 64];
        T a0, a1, x0, x1;
        static_assert(sizeof(V2) == 2 * sizeof(T), "two elements per vector");
        __builtin_memcpy(&a0, &v[u].x, sizeof(T));
        __builtin_memcpy(&a1, &v[u].z, sizeof(T));
        __builtin_memcpy(&x0, &w.x, sizeof(T));
        __builtin_memcpy(&x1, &w.z, sizeof(T));
        acc += a0 * x0;
        acc1 += a1 * x1;
      }
      acc += acc1;
      for (int off = 32; off > 0; off >>= 1) acc += shfl_xor_s(acc, off);
      if (lane == 0) part[(e * nsplit + blockIdx.y) * n + j] = acc;
      return;
    }
  }
  if (r0 + GV_CHUNK <= m) {
    T v[GV_CHUNK / 64];
#pragma unroll
    for (int u = 0; u < GV_CHUNK / 64; ++u) v[u] = a[r + u * 64];
#pragma unroll
    for (int u = 0; u < GV_CHUNK / 64; u += 2) {
      acc += v[u] * x[r + u * 64];
      acc1 += v[u + 1] * x[r + (u + 1) * 64];
    }
  } else {
    for (; r < r1; r += 64) acc += a[r] * x[r];
  }
  acc += acc1;
  for (int off = 32; off > 0; off >>= 1) acc += shfl_xor_s(acc, off);
  if (lane == 0) part[(e * nsplit + blockIdx.y) * n + j] = acc;
}

// y = alpha A x + gamma z for a symmetric A (A^T = A, no conjugate) read from its upper
// triangle only, half the bytes of a GEMV: each stored A(r, j), r <= j, gives A(r, j) x_j
// to y_r and, off the diagonal, A(r, j) x_r to y_j.  Block (cb, rc) takes the columns
// [64 cb, 64 cb + 64) and rows [256 rc, 256 rc + 256) of one matrix and leaves row partials
// rp[e][cb][r] (its columns' part of y_r) and column partials cp[e][rc][j] (its rows' part of
// y_j, a wave sum per column); a tile below the diagonal reads nothing and leaves zeros.
constexpr int SV_CB = 64, SV_RC = 256;

template <typename T>
__global__ void __launch_bounds__(256) symv_upper_partial(int64_t n, const T* __restrict__ A,
                                                          int64_t lda, int64_t sA,
                                                          const T* __restrict__ x, int64_t sx,
                                                          T* __restrict__ rp,
                                                          T* __restrict__ cp) {
  __shared__ T red[4][SV_RC];
  const int64_t e = blockIdx.z;
  const int cb = blockIdx.x, rc = blockIdx.y, ncb = gridDim.x, nrc = gridDim.y;
  A += e * sA;
  x += e * sx;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t j0 = int64_t(cb) * SV_CB, r0 = int64_t(rc) * SV_RC;
  const bool active = r0 < j0 + SV_CB;  // the tile meets the upper triangle
  constexpr int RU = SV_RC / 64;
  T ra[RU], xr[RU];
#pragma unroll
  for (int u = 0; u < RU; ++u) {
    const int64_t r = r0 + lane + 64 * u;
    ra[u] = T(0);
    xr[u] = (active && r < n) ? x[r] : T(0);
  }
  T* cpo = cp + (e * nrc + rc) * n;
  for (int jj = 0; jj < SV_CB / 4; ++jj) {
    const int64_t j = j0 + wave * (SV_CB / 4) + jj;  // wave-uniform
    if (j >= n) break;
    T ca = T(0);
    if (active) {
      const T xj = x[j];
      const T* col = A + j * lda;
      T v[RU];
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int64_t r = r0 + lane + 64 * u;
        v[u] = r <= j ? col[r] : T(0);
      }
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int64_t r = r0 + lane + 64 * u;
        ra[u] += v[u] * xj;
        if (r < j) ca += v[u] * xr[u];
      }
      for (int off = 32; off > 0; off >>= 1) ca += shfl_xor_s(ca, off);
    }
    if (lane == 0) cpo[j] = ca;
  }
#pragma unroll
  for (int u = 0; u < RU; ++u) red[wave][lane + 64 * u] = ra[u];
  __syncthreads();
  T* rpo = rp + (e * ncb + cb) * n;
  const int64_t r = r0 + threadIdx.x;
  if (r < n) rpo[r] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                      (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

template <typename T>
__global__ void __launch_bounds__(256) symv_reduce(int64_t n, int ncb, int nrc,
                                                   const T* __restrict__ rp,
                                                   const T* __restrict__ cp, T alpha,
                                                   const T* __restrict__ z, int64_t sz, T gamma,
                                                   T* __restrict__ y, int64_t sy) {
  const int64_t e = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  T s = T(0);
  for (int q = 0; q < ncb; ++q) s += rp[(e * ncb + q) * n + i];
  for (int q = 0; q < nrc; ++q) s += cp[(e * nrc + q) * n + i];
  T v = alpha * s;
  if (gamma != T(0)) v += gamma * z[e * sz + i];
  y[e * sy + i] = v;
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
  const double fl = (is_complex_v<T> ? 8.0 : 2.0) * m * double(n) * batch;
  const double by = (double(m) * n + m + n) * sizeof(T) * batch;
  if (!trans) {
    dim3 grid((m + GV_ROWS - 1) / GV_ROWS, nsplit, batch);
    HSSC_TIMED(lbl<T>("gemv").c_str(), st, fl, by,
               hipLaunchKernelGGL(gemv_n_partial<T>, grid, dim3(256), 0, st, m, n, A, lda, sA, x,
                                  sx, work, nsplit));
  } else {
    dim3 grid((n + 3) / 4, nsplit, batch);
    HSSC_TIMED(lbl<T>("gemvt").c_str(), st, fl, by,
               hipLaunchKernelGGL(gemv_t_partial<T>, grid, dim3(256), 0, st, m, n, A, lda, sA, x,
                                  sx, work, nsplit));
  }
  hipLaunchKernelGGL(gemv_reduce<T>, dim3((out + 255) / 256, batch), dim3(256), 0, st, out, nsplit,
                     work, alpha, z, sz, gamma, y, sy);
}

int64_t symv_work_elems(int64_t n, int64_t batch) {
  return batch * ((n + SV_CB - 1) / SV_CB + (n + SV_RC - 1) / SV_RC) * n;
}

template <typename T>
void symv_upper(hipStream_t st, int64_t n, T alpha, const T* A, int64_t lda, int64_t sA,
                const T* x, int64_t sx, T gamma, const T* z, int64_t sz, T* y, int64_t sy,
                int64_t batch, T* work) {
  if (batch <= 0 || n <= 0) return;
  const int ncb = int((n + SV_CB - 1) / SV_CB), nrc = int((n + SV_RC - 1) / SV_RC);
  T* rp = work;
  T* cp = work + batch * ncb * n;
  const double fl = (is_complex_v<T> ? 8.0 : 2.0) * double(n) * n * batch;
  const double by = (double(n) * (n + 1) / 2 + 2 * n) * sizeof(T) * batch;
  HSSC_TIMED(lbl<T>("symv").c_str(), st, fl, by,
             hipLaunchKernelGGL(symv_upper_partial<T>, dim3(ncb, nrc, batch), dim3(256), 0, st, n,
                                A, lda, sA, x, sx, rp, cp));
  hipLaunchKernelGGL(symv_reduce<T>, dim3((n + 255) / 256, batch), dim3(256), 0, st, n, ncb, nrc,
                     rp, cp, alpha, z, sz, gamma, y, sy);
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

// out[e][i] = sum_{q < nsplit} part[e*nsplit + q][i]   (split-K reduction, fixed order)
template <typename T>
__global__ void __launch_bounds__(256) sum_split_kernel(int64_t len, int64_t nsplit,
                                                        const T* __restrict__ part,
                                                        T* __restrict__ out, int64_t sout) {
  const int64_t e = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= len) return;
  const T* pe = part + e * nsplit * len + i;
  T s = T(0);
  int64_t q = 0;
  for (; q + 8 <= nsplit; q += 8) {  // eight loads in flight, summed in chunk order
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = pe[(q + u) * len];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; q < nsplit; ++q) s += pe[q * len];
  out[e * sout + i] = s;
}

// W2[e] = Ts[e]^H W1[e], W1[e] = sum_{q < nsplit} part[e*nsplit + q] (m x n, m <= QR_NB, ld
// m): the split-K reduction of a sub-panel's V^H C (the chunks summed in the order of
// sum_split_kernel) and its product with the sub-panel's triangular factor, in one launch.
// Sixteen columns per workgroup; Ts and the reduced columns are staged in LDS.
constexpr int kThCols = 16;
template <typename T>
__global__ void __launch_bounds__(256) sum_split_th_kernel(int m, int64_t n, int64_t nsplit,
                                                           const T* __restrict__ part,
                                                           const T* __restrict__ Ts,
                                                           int64_t ldt, int64_t sT,
                                                           T* __restrict__ W2, int64_t sW2) {
  __shared__ T w1[QR_NB][kThCols + 1];
  __shared__ T th[QR_NB][QR_NB + 1];
  const int64_t e = blockIdx.y;
  const int64_t c0 = int64_t(blockIdx.x) * kThCols;
  const int tid = threadIdx.x;
  const int64_t len = int64_t(m) * n;
  const T* pe = part + e * nsplit * len;
  Ts += e * sT;
  for (int idx = tid; idx < m * m; idx += 256) th[idx % m][idx / m] = Ts[idx % m + (idx / m) * ldt];
  for (int idx = tid; idx < m * kThCols; idx += 256) {
    const int i = idx % m, j = idx / m;
    T s = T(0);
    if (c0 + j < n) {
      const T* pp = pe + i + (c0 + j) * m;
      int64_t q = 0;
      for (; q + 8 <= nsplit; q += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pp[(q + u) * len];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
      }
      for (; q < nsplit; ++q) s += pp[q * len];
    }
    w1[i][j] = s;
  }
  __syncthreads();
  for (int idx = tid; idx < m * kThCols; idx += 256) {
    const int i = idx % m, j = idx / m;
    if (c0 + j >= n) continue;
    T s = T(0);
    for (int t = 0; t < m; ++t) s += conj_s(th[t][i]) * w1[t][j];
    W2[e * sW2 + i + (c0 + j) * m] = s;
  }
}

// Whether gemm_tn_splitk cuts the K range (see there).
inline bool splitk_pays(int64_t m, int64_t n, int64_t kdim, int64_t batch, int64_t kchunk,
                        int64_t part_cap) {
  const int64_t nsplit = (kdim + kchunk - 1) / kchunk;
  // split only while the unsplit grid leaves CUs idle: with many matrices (the small boxes of
  // the low levels) the partial products and their sum only add workgroups and a launch; and
  // only when the partial products fit the caller's buffer
  const int64_t tiles = ((m + GB_M - 1) / GB_M) * ((n + GB_N - 1) / GB_N) * batch;
  return nsplit > 1 && tiles < 512 && batch * nsplit * m * n <= part_cap;
}

// The chunks of A^H B (K cut in chunks of kchunk rows) into part, ld m.
template <typename T>
void splitk_chunks(hipStream_t st, int64_t m, int64_t n, int64_t kdim, const T* A, int64_t lda,
                   int64_t sA, const T* B, int64_t ldb, int64_t sB, int64_t batch, T* part,
                   int64_t kchunk) {
  const int64_t nsplit = (kdim + kchunk - 1) / kchunk;
  const int64_t big = INT64_MAX;
  gemm_skip_split<T>(st, true, false, m, n, kdim, T(1), A, lda, sA, B, ldb, sB, T(0), part, m,
                     m * n, batch, (m + GB_M - 1) / GB_M, (n + GB_N - 1) / GB_N, big, 0, big, 0,
                     nsplit, kchunk, true);
}

// C (m x n) = A^H B for tall A (K x m) and B (K x n) with a long K: the K range is cut in
// nsplit chunks computed as independent GEMMs into `part`, then summed in a fixed order.
// (A^H = A^T for real data.)
template <typename T>
void gemm_tn_splitk(hipStream_t st, int64_t m, int64_t n, int64_t kdim, const T* A, int64_t lda,
                    int64_t sA, const T* B, int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC,
                    int64_t batch, T* part, int64_t kchunk, int64_t part_cap) {
  if (ldc != m || !splitk_pays(m, n, kdim, batch, kchunk, part_cap)) {
    gemm_h<T>(st, m, n, kdim, T(1), A, lda, sA, B, ldb, sB, T(0), C, ldc, sC, batch);
    return;
  }
  const int64_t nsplit = (kdim + kchunk - 1) / kchunk, len = m * n;
  splitk_chunks<T>(st, m, n, kdim, A, lda, sA, B, ldb, sB, batch, part, kchunk);
  hipLaunchKernelGGL(sum_split_kernel<T>, dim3((len + 255) / 256, batch), dim3(256), 0, st, len,
                     nsplit, part, C, sC);
}

template <typename T>
void gemm_tn_splitk_th(hipStream_t st, int64_t m, int64_t n, int64_t kdim, const T* A,
                       int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB, const T* Ts,
                       int64_t ldt, int64_t sT, T* W1, T* W2, int64_t sW, int64_t batch, T* part,
                       int64_t kchunk, int64_t part_cap) {
  if (m > QR_NB || !splitk_pays(m, n, kdim, batch, kchunk, part_cap)) {
    gemm_tn_splitk<T>(st, m, n, kdim, A, lda, sA, B, ldb, sB, W1, m, sW, batch, part, kchunk,
                      part_cap);
    gemm_h<T>(st, m, n, m, T(1), Ts, ldt, sT, W1, m, sW, T(0), W2, m, sW, batch);
    return;
  }
  const int64_t nsplit = (kdim + kchunk - 1) / kchunk;
  splitk_chunks<T>(st, m, n, kdim, A, lda, sA, B, ldb, sB, batch, part, kchunk);
  hipLaunchKernelGGL(sum_split_th_kernel<T>, dim3((n + kThCols - 1) / kThCols, batch), dim3(256),
                     0, st, int(m), n, nsplit, part, Ts, ldt, sT, W2, sW);
}

int64_t ls_part_elems(int64_t M, int64_t n, int64_t nbo, int64_t kchunk, int64_t kchunk_thin) {
  const int64_t nsplit = (M + kchunk - 1) / kchunk;
  const int64_t nsplit_thin = (M + kchunk_thin - 1) / kchunk_thin;
  return std::max({nsplit * nbo * n, nsplit_thin * int64_t(kQrNb) * n, nsplit_thin * nbo * nbo});
}

// Regularisation rows of the least squares of one box (one workgroup per box): below the
// Mq x n block, A(Mq + i, i) = lambda for i < k (the rows themselves are zero), with lambda =
// eta times the RMS norm of the block's first min(k, 32) columns.
template <typename T>
__global__ void __launch_bounds__(256) tikhonov_rows_kernel(int64_t Mq, int64_t k,
                                                            T* __restrict__ A, int64_t lda,
                                                            int64_t sA, double eta) {
  __shared__ double red[4];
  A += int64_t(blockIdx.x) * sA;
  const int64_t c = k < 32 ? k : 32;
  double s = 0;
  for (int64_t idx = threadIdx.x; idx < Mq * c; idx += 256)
    s += double(abs2(A[(idx % Mq) + (idx / Mq) * lda]));
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  const double lam = eta * sqrt((red[0] + red[1] + red[2] + red[3]) / double(c > 0 ? c : 1));
  for (int64_t i = threadIdx.x; i < k; i += 256) A[Mq + i + i * lda] = T(real_t<T>(lam));
}

template <typename T>
void tikhonov_rows(hipStream_t st, int64_t Mq, int64_t k, T* A, int64_t lda, int64_t sA,
                   int64_t batch, double eta) {
  if (k <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(tikhonov_rows_kernel<T>, dim3(batch), dim3(256), 0, st, Mq, k, A, lda, sA,
                     eta);
}

// Sum of squares (for residual norms), deterministic two-phase.
template <typename T>
__global__ void __launch_bounds__(256) sumsq_partial(int64_t len, const T* __restrict__ x,
                                                     double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += int64_t(gridDim.x) * 256) {
    s += double(abs2(x[i]));
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
                             int64_t, T*, int64_t, int64_t, const KernelSpec&);                    \
  template void gen_parent<T>(hipStream_t, int64_t, int64_t, int64_t, const int32_t*, int64_t,     \
                              const double*, const double*, const double*, int64_t, const T*,      \
                              int64_t, int64_t, T*, int64_t, const KernelSpec&);                   \
  template void gemm<T>(hipStream_t, bool, bool, int64_t, int64_t, int64_t, T, const T*, int64_t,  \
                        int64_t, const T*, int64_t, int64_t, T, T*, int64_t, int64_t, int64_t);    \
  template void gemm_h<T>(hipStream_t, int64_t, int64_t, int64_t, T, const T*, int64_t, int64_t,   \
                          const T*, int64_t, int64_t, T, T*, int64_t, int64_t, int64_t);           \
  template void gemm_skip<T>(hipStream_t, bool, bool, int64_t, int64_t, int64_t, T, const T*,    \
                             int64_t, int64_t, const T*, int64_t, int64_t, T, T*, int64_t, int64_t, \
                             int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int);   \
  template void invert_batched<T>(hipStream_t, int64_t, T*, int64_t, int64_t, int64_t,           \
                                  hipStream_t, T*, bool, int*);                                    \
  template int64_t invert_ws_elems<T>(int64_t, int64_t);                                           \
  template void symv_upper<T>(hipStream_t, int64_t, T, const T*, int64_t, int64_t, const T*,       \
                              int64_t, T, const T*, int64_t, T*, int64_t, int64_t, T*);            \
  template void sym_lower_fill<T>(hipStream_t, int64_t, T*, int64_t, int64_t, int64_t);            \
  template void trinv_upper<T>(hipStream_t, int64_t, int, const T*, int64_t, int64_t, T*, int64_t, \
                               int64_t);                                                           \
  template void trinv_upper_all<T>(hipStream_t, int64_t, const T*, int64_t, int64_t, T*, int64_t,  \
                                   int64_t);                                                       \
  template void cqr_panel<T>(hipStream_t, int64_t, int64_t, int, T*, int64_t, int64_t, T*, int64_t, \
                             int64_t, T*, int64_t, wide_t<T>*, int64_t);                           \
  template void t_assemble<T>(hipStream_t, int, const T*, int64_t, const T*, int64_t, int64_t, T*,  \
                              int64_t, int64_t, T*, int64_t, int64_t);                             \
  template void zero2d<T>(hipStream_t, int64_t, int64_t, T*, int64_t, int64_t, int64_t);            \
  template void tikhonov_rows<T>(hipStream_t, int64_t, int64_t, T*, int64_t, int64_t, int64_t,     \
                                 double);                                                          \
  template void gemv<T>(hipStream_t, bool, int64_t, int64_t, T, const T*, int64_t, int64_t,        \
                        const T*, int64_t, T, const T*, int64_t, T*, int64_t, int64_t, T*);        \
  template void sumsq<T>(hipStream_t, int64_t, const T*, double*, int);                           \
  template void gemm_tn_splitk<T>(hipStream_t, int64_t, int64_t, int64_t, const T*, int64_t,       \
                                  int64_t, const T*, int64_t, int64_t, T*, int64_t, int64_t,       \
                                  int64_t, T*, int64_t, int64_t);                                  \
  template void gemm_tn_splitk_th<T>(hipStream_t, int64_t, int64_t, int64_t, const T*, int64_t,    \
                                     int64_t, const T*, int64_t, int64_t, const T*, int64_t,       \
                                     int64_t, T*, T*, int64_t, int64_t, T*, int64_t, int64_t);     \
  template void add2d<T>(hipStream_t, int64_t, int64_t, T, const T*, int64_t, int64_t, T, const T*, \
                         int64_t, int64_t, T*, int64_t, int64_t, int64_t);

HSSC_INST(double)
HSSC_INST(float)
HSSC_INST(zdouble)
HSSC_INST(zfloat)

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
HSSC_INST2(zdouble, zdouble)
HSSC_INST2(zfloat, zfloat)
HSSC_INST2(zdouble, zfloat)
HSSC_INST2(zfloat, zdouble)

}  // namespace k
}  // namespace hssc

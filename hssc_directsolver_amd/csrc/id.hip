// Interpolative decomposition on the GPU (Common/ID.m): column-pivoted Householder QR of a
// batch of ragged matrices, one workgroup per matrix, the reference's rank rule and its
// randomized acceptance check, and the interpolation matrix T = R11^-1 R12.
//
// Matrix b is m_b x n_b, column-major with leading dimension m_b, at A + off[b].  The
// matrix is overwritten by its QR factors; an untouched copy (Acopy, same layout) feeds the
// randomized check.  Outputs per matrix: the column permutation (perm, n_b entries, skeleton
// first), the rank k_b and T (k_b x (n_b - k_b), leading dimension k_b) at Tout + toff[b].

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"
#include "stats.h"

namespace hssc {
namespace k {

namespace {

constexpr int ID_NT = 256;

template <typename T>
__device__ __forceinline__ T wsum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide sum over ID_NT threads (every thread gets the result).
template <typename T>
__device__ T bsum(T v, T* red) {
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = T(0);
  for (int w = 0; w < ID_NT / 64; ++w) s += red[w];
  return s;
}

// Deterministic uniform [0,1) numbers from (seed, counter): a 64-bit mix (splitmix64).
__device__ __forceinline__ double urand(uint64_t seed, uint64_t i) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return double(z >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

// mode: acc > 0 -> rank by |R_ii| > acc |R_11| plus the randomized check (ID.m:40-65);
//       acc <= 0 -> fixed rank kfix[b] (ID.m:106-113, T by back substitution).
template <typename T>
__global__ void __launch_bounds__(ID_NT) cpqr_id_kernel(
    int nmat, const int64_t* __restrict__ off, const int* __restrict__ mr, const int* __restrict__ nc,
    T* __restrict__ A, const T* __restrict__ Acopy, double acc, const int* __restrict__ kfix,
    int* __restrict__ perm, const int64_t* __restrict__ poff, int* __restrict__ kout,
    T* __restrict__ Tout, const int64_t* __restrict__ toff, T* __restrict__ work,
    const int64_t* __restrict__ woff, uint64_t seed) {
  __shared__ T red[ID_NT / 64];
  __shared__ T sbest[ID_NT / 64];
  __shared__ int sidx[ID_NT / 64];
  __shared__ int s_k;
  const int b = blockIdx.x;
  if (b >= nmat) return;
  const int m = mr[b], n = nc[b];
  T* M = A + off[b];
  const T* Mc = Acopy + off[b];
  int* P = perm + poff[b];
  T* wk = work + woff[b];        // n norms^2, then n (rdiag) , then scratch
  T* nrm = wk;
  T* rdg = wk + n;
  T* tau = wk + 2 * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kmax = m < n ? m : n;
  for (int c = tid; c < n; c += ID_NT) {
    P[c] = c;
    T s = T(0);
    for (int r = 0; r < m; ++r) s += M[r + int64_t(c) * m] * M[r + int64_t(c) * m];
    nrm[c] = s;
  }
  __syncthreads();
  for (int j = 0; j < kmax; ++j) {
    // pivot: largest remaining column norm (ties -> lowest index)
    T best = T(-1);
    int bi = j;
    for (int c = j + tid; c < n; c += ID_NT)
      if (nrm[c] > best) best = nrm[c], bi = c;
    for (int o = 32; o > 0; o >>= 1) {
      const T ob = __shfl_xor(best, o);
      const int oi = __shfl_xor(bi, o);
      if (ob > best || (ob == best && oi < bi)) best = ob, bi = oi;
    }
    if (lane == 0) sbest[wave] = best, sidx[wave] = bi;
    __syncthreads();
    best = sbest[0], bi = sidx[0];
    for (int w = 1; w < ID_NT / 64; ++w)
      if (sbest[w] > best || (sbest[w] == best && sidx[w] < bi)) best = sbest[w], bi = sidx[w];
    const int p = bi;
    __syncthreads();
    if (p != j) {
      for (int r = tid; r < m; r += ID_NT) {
        const T t = M[r + int64_t(j) * m];
        M[r + int64_t(j) * m] = M[r + int64_t(p) * m];
        M[r + int64_t(p) * m] = t;
      }
      if (tid == 0) {
        const T t = nrm[j];
        nrm[j] = nrm[p], nrm[p] = t;
        const int q = P[j];
        P[j] = P[p], P[p] = q;
      }
    }
    __syncthreads();
    // Householder reflector for column j (rows j..m-1)
    T* a = M + int64_t(j) * m;
    T sg = T(0);
    for (int r = j + 1 + tid; r < m; r += ID_NT) sg += a[r] * a[r];
    sg = bsum(sg, red);
    const T alpha = a[j];
    T beta, tv, scal;
    if (sg == T(0)) {
      tv = T(0), beta = alpha, scal = T(0);
    } else {
      const T nr = sqrt(alpha * alpha + sg);
      beta = alpha >= T(0) ? -nr : nr;
      tv = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    __syncthreads();
    for (int r = j + 1 + tid; r < m; r += ID_NT) a[r] *= scal;
    if (tid == 0) a[j] = beta, rdg[j] = fabs(beta), tau[j] = tv;
    __syncthreads();
    // apply to columns c > j (one wave per column) and recompute their trailing norms
    for (int c = j + 1 + wave; c < n; c += ID_NT / 64) {
      T* col = M + int64_t(c) * m;
      T s = T(0);
      for (int r = j + lane; r < m; r += 64) s += (r == j ? T(1) : a[r]) * col[r];
      s = wsum(s) * tv;
      T q = T(0);
      for (int r = j + lane; r < m; r += 64) {
        const T v = col[r] - (r == j ? T(1) : a[r]) * s;
        col[r] = v;
        if (r > j) q += v * v;
      }
      q = wsum(q);
      if (lane == 0) nrm[c] = q;
    }
    __syncthreads();
  }
  // rank
  if (tid == 0) {
    int k;
    if (acc > 0) {
      const T thr = kmax > 0 ? rdg[0] * T(acc) : T(0);
      k = 0;
      for (int i = 0; i < kmax; ++i) k += rdg[i] > thr ? 1 : 0;
    } else {
      k = kfix[b] < kmax ? kfix[b] : kmax;
    }
    s_k = k;
  }
  __syncthreads();
  T* To = Tout + toff[b];
  T thr = kmax > 0 ? rdg[0] * T(acc) : T(0);
  for (int round = 0;; ++round) {
    const int k = s_k;
    // T = R11^-1 R12 by back substitution, one thread per column of R12
    for (int c = tid; c < n - k; c += ID_NT) {
      T* tc = To + int64_t(c) * k;
      for (int i = k - 1; i >= 0; --i) {
        T s = M[i + int64_t(k + c) * m];
        for (int q = i + 1; q < k; ++q) s -= M[i + int64_t(q) * m] * tc[q];
        const T d = M[i + int64_t(i) * m];
        tc[i] = (d != T(0)) ? s / d : T(0);  // a zero pivot (rank forced above the
                                             // numerical rank) contributes nothing
      }
    }
    __syncthreads();
    if (acc <= 0 || k >= kmax || k == n) break;
    // randomized check (ID.m:55-59,68-86): err = sqrt(|T^T A_s^T y - A_r^T y|) with
    // y = A_s T x - A_r x for a random unit x; nrm = sqrt(|A_r^T A_r x|)
    const int r_ = n - k;
    T* x = tau + n;        // r_ entries
    T* y = x + n;          // m entries
    T* z = y + m;          // n entries (A_s^T y etc.)
    T xs = T(0);
    for (int i = tid; i < r_; i += ID_NT) {
      const T v = T(urand(seed ^ (uint64_t(b) << 20), uint64_t(round) * 4096 + i));
      x[i] = v;
      xs += v * v;
    }
    xs = bsum(xs, red);
    const T xinv = T(1) / sqrt(xs);
    for (int i = tid; i < r_; i += ID_NT) x[i] *= xinv;
    __syncthreads();
    // u = T x (k), stored in z
    for (int i = tid; i < k; i += ID_NT) {
      T s = T(0);
      for (int c = 0; c < r_; ++c) s += To[i + int64_t(c) * k] * x[c];
      z[i] = s;
    }
    __syncthreads();
    // y = A_s u - A_r x  (original columns P[..])
    for (int rr = tid; rr < m; rr += ID_NT) {
      T s = T(0);
      for (int i = 0; i < k; ++i) s += Mc[rr + int64_t(P[i]) * m] * z[i];
      for (int c = 0; c < r_; ++c) s -= Mc[rr + int64_t(P[k + c]) * m] * x[c];
      y[rr] = s;
    }
    __syncthreads();
    // e = T^T (A_s^T y) - A_r^T y ; |e|
    for (int i = tid; i < k; i += ID_NT) {
      T s = T(0);
      for (int rr = 0; rr < m; ++rr) s += Mc[rr + int64_t(P[i]) * m] * y[rr];
      z[i] = s;
    }
    __syncthreads();
    T e2 = T(0);
    for (int c = tid; c < r_; c += ID_NT) {
      T s = T(0);
      for (int i = 0; i < k; ++i) s += To[i + int64_t(c) * k] * z[i];
      T t = T(0);
      for (int rr = 0; rr < m; ++rr) t += Mc[rr + int64_t(P[k + c]) * m] * y[rr];
      s -= t;
      e2 += s * s;
    }
    e2 = bsum(e2, red);
    // n2 = |A_r^T A_r x| with a fresh random x
    T xs2 = T(0);
    for (int i = tid; i < r_; i += ID_NT) {
      const T v = T(urand(seed ^ (uint64_t(b) << 20) ^ 0x5555ull, uint64_t(round) * 4096 + i));
      x[i] = v;
      xs2 += v * v;
    }
    xs2 = bsum(xs2, red);
    const T xinv2 = T(1) / sqrt(xs2);
    for (int i = tid; i < r_; i += ID_NT) x[i] *= xinv2;
    __syncthreads();
    for (int rr = tid; rr < m; rr += ID_NT) {
      T s = T(0);
      for (int c = 0; c < r_; ++c) s += Mc[rr + int64_t(P[k + c]) * m] * x[c];
      y[rr] = s;
    }
    __syncthreads();
    T n2 = T(0);
    for (int c = tid; c < r_; c += ID_NT) {
      T t = T(0);
      for (int rr = 0; rr < m; ++rr) t += Mc[rr + int64_t(P[k + c]) * m] * y[rr];
      n2 += t * t;
    }
    n2 = bsum(n2, red);
    const T err = sqrt(sqrt(e2)), nrmr = sqrt(sqrt(n2));
    if (!(err > nrmr * T(acc))) break;
    // relax the threshold (Rmx = 0.75 Rmx) and recount
    thr *= T(0.75);
    __syncthreads();
    if (tid == 0) {
      int kk = 0;
      for (int i = 0; i < kmax; ++i) kk += rdg[i] > thr ? 1 : 0;
      s_k = kk;
    }
    __syncthreads();
  }
  if (tid == 0) kout[b] = s_k;
}

template <typename T>
void cpqr_id(hipStream_t st, int nmat, const int64_t* off, const int* mr, const int* nc, T* A,
             const T* Acopy, double acc, const int* kfix, int* perm, const int64_t* poff, int* kout,
             T* Tout, const int64_t* toff, T* work, const int64_t* woff, uint64_t seed) {
  if (nmat <= 0) return;
  HSSC_TIMED(sizeof(T) == 8 ? "cpqr_id_f64" : "cpqr_id_f32", st, 0.0, 0.0,
             hipLaunchKernelGGL(cpqr_id_kernel<T>, dim3(nmat), dim3(ID_NT), 0, st, nmat, off, mr,
                                nc, A, Acopy, acc, kfix, perm, poff, kout, Tout, toff, work, woff,
                                seed));
}

template void cpqr_id<double>(hipStream_t, int, const int64_t*, const int*, const int*, double*,
                              const double*, double, const int*, int*, const int64_t*, int*,
                              double*, const int64_t*, double*, const int64_t*, uint64_t);

}  // namespace k
}  // namespace hssc

// Interpolative decomposition on the GPU (Common/ID.m): column-pivoted Householder QR of a
// batch of ragged matrices, one workgroup per matrix, the reference's rank rule and its
// randomized acceptance check, and the interpolation matrix T = R11^-1 R12.
//
// Matrix b is m_b x n_b, column-major with leading dimension m_b, at A + off[b].  The
// matrix is overwritten by its QR factors; an untouched copy (Acopy, same layout) feeds the
// randomized check.  Outputs per matrix: the column permutation (perm, n_b entries, skeleton
// first), the rank k_b and T (k_b x (n_b - k_b), leading dimension k_b) at Tout + toff[b].

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>

#include "internal.h"

#include "kernels.h"
#include "stats.h"

namespace hssc {
namespace k {

namespace {

constexpr int ID_NT = 256;

template <typename T>
__device__ __forceinline__ T wsum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += shfl_xor_s(v, o);
  return v;
}

// Block-wide sum over NT threads (every thread gets the result).
template <int NT, typename T>
__device__ T bsum(T v, T* red) {
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = T(0);
  for (int w = 0; w < NT / 64; ++w) s += red[w];
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

// Rank, interpolation matrix and randomized acceptance check of one block (ID.m:40-113),
// run by the NT threads of its workgroup once the block's column-pivoted R factor is in M
// (columns in pivot order, ld m), its |R_ii| in rdg and the pivot order in P.  scratch holds
// 2 n + m entries; s_k is a shared int.
template <int NT, typename T>
__device__ void id_tail(int b, int m, int n, const T* M, const T* Mc, const int* P, const T* rdg,
                        T* scratch, double acc, const int* kfix, int* kout, T* To, uint64_t seed,
                        real_t<T>* redr, int* s_k) {
  using R = real_t<T>;
  const int tid = threadIdx.x;
  const int kmax = m < n ? m : n;
  if (tid == 0) {
    int k;
    if (acc > 0) {
      const R thr = kmax > 0 ? re_s(rdg[0]) * R(acc) : R(0);
      k = 0;
      for (int i = 0; i < kmax; ++i) k += re_s(rdg[i]) > thr ? 1 : 0;
    } else {
      k = kfix[b] < kmax ? kfix[b] : kmax;
    }
    *s_k = k;
  }
  __syncthreads();
  R thr = kmax > 0 ? re_s(rdg[0]) * R(acc) : R(0);
  for (int round = 0;; ++round) {
    const int k = *s_k;
    // T = R11^-1 R12 by back substitution, one thread per column of R12
    for (int c = tid; c < n - k; c += NT) {
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
    // randomized check (ID.m:55-59,68-86): err = sqrt(|T^H A_s^H y - A_r^H y|) with
    // y = A_s T x - A_r x for a random unit x; nrm = sqrt(|A_r^H A_r x|)
    const int r_ = n - k;
    T* x = scratch;        // r_ entries
    T* y = x + n;          // m entries
    T* z = y + m;          // n entries (A_s^H y etc.)
    R xs = R(0);
    for (int i = tid; i < r_; i += NT) {
      const R v = R(urand(seed ^ (uint64_t(b) << 20), uint64_t(round) * 4096 + i));
      x[i] = T(v);
      xs += v * v;
    }
    xs = bsum<NT>(xs, redr);
    const R xinv = R(1) / sqrt(xs);
    for (int i = tid; i < r_; i += NT) x[i] *= T(xinv);
    __syncthreads();
    // u = T x (k), stored in z
    for (int i = tid; i < k; i += NT) {
      T s = T(0);
      for (int c = 0; c < r_; ++c) s += To[i + int64_t(c) * k] * x[c];
      z[i] = s;
    }
    __syncthreads();
    // y = A_s u - A_r x  (original columns P[..])
    for (int rr = tid; rr < m; rr += NT) {
      T s = T(0);
      for (int i = 0; i < k; ++i) s += Mc[rr + int64_t(P[i]) * m] * z[i];
      for (int c = 0; c < r_; ++c) s -= Mc[rr + int64_t(P[k + c]) * m] * x[c];
      y[rr] = s;
    }
    __syncthreads();
    // e = T^H (A_s^H y) - A_r^H y ; |e|
    for (int i = tid; i < k; i += NT) {
      T s = T(0);
      for (int rr = 0; rr < m; ++rr) s += conj_s(Mc[rr + int64_t(P[i]) * m]) * y[rr];
      z[i] = s;
    }
    __syncthreads();
    R e2 = R(0);
    for (int c = tid; c < r_; c += NT) {
      T s = T(0);
      for (int i = 0; i < k; ++i) s += conj_s(To[i + int64_t(c) * k]) * z[i];
      T t = T(0);
      for (int rr = 0; rr < m; ++rr) t += conj_s(Mc[rr + int64_t(P[k + c]) * m]) * y[rr];
      s -= t;
      e2 += abs2(s);
    }
    e2 = bsum<NT>(e2, redr);
    // n2 = |A_r^H A_r x| with a fresh random x
    R xs2 = R(0);
    for (int i = tid; i < r_; i += NT) {
      const R v = R(urand(seed ^ (uint64_t(b) << 20) ^ 0x5555ull, uint64_t(round) * 4096 + i));
      x[i] = T(v);
      xs2 += v * v;
    }
    xs2 = bsum<NT>(xs2, redr);
    const R xinv2 = R(1) / sqrt(xs2);
    for (int i = tid; i < r_; i += NT) x[i] *= T(xinv2);
    __syncthreads();
    for (int rr = tid; rr < m; rr += NT) {
      T s = T(0);
      for (int c = 0; c < r_; ++c) s += Mc[rr + int64_t(P[k + c]) * m] * x[c];
      y[rr] = s;
    }
    __syncthreads();
    R n2 = R(0);
    for (int c = tid; c < r_; c += NT) {
      T t = T(0);
      for (int rr = 0; rr < m; ++rr) t += conj_s(Mc[rr + int64_t(P[k + c]) * m]) * y[rr];
      n2 += abs2(t);
    }
    n2 = bsum<NT>(n2, redr);
    const R err = sqrt(sqrt(e2)), nrmr = sqrt(sqrt(n2));
    if (!(err > nrmr * R(acc))) break;
    // relax the threshold (Rmx = 0.75 Rmx) and recount; a count that does not grow would
    // repeat the same check, so keep relaxing until it grows, and stop once it cannot
    // (every thread runs the same loop on the same data)
    int kk = k;
    for (int g = 0; g < 4096 && kk == k && thr > R(0); ++g) {
      thr *= R(0.75);
      kk = 0;
      for (int i = 0; i < kmax; ++i) kk += re_s(rdg[i]) > thr ? 1 : 0;
    }
    __syncthreads();
    if (tid == 0) *s_k = kk;
    __syncthreads();
    if (kk == k) break;
  }
  if (tid == 0) kout[b] = *s_k;
}

// id_tail for a workgroup that holds the block's R in LDS: R(0:q, q) of every pivot position
// q < min(m, n) packed at Rl + q (q + 1) / 2 (k <= 128).  The same quantities, with the
// substitution done column by column by one wave per column of R12 (the residual in the
// lanes) and every product over the rows of the block spread over the lanes of a wave, so
// that no thread walks a column of the block on its own.
template <int NT, typename T>
__device__ void id_tail_wave(int b, int m, int n, const T* M, const T* Mc, const int* P,
                             const T* rdg, T* scratch, double acc, const int* kfix, int* kout,
                             T* To, uint64_t seed, real_t<T>* redr, int* s_k, const T* Rl) {
  using R = real_t<T>;
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kmax = m < n ? m : n;
  auto rget = [&](int i, int q) -> T {
    return q < kmax ? Rl[q * (q + 1) / 2 + i] : M[i + int64_t(q) * m];
  };
  if (tid == 0) {
    int k;
    if (acc > 0) {
      const R thr = kmax > 0 ? re_s(rdg[0]) * R(acc) : R(0);
      k = 0;
      for (int i = 0; i < kmax; ++i) k += re_s(rdg[i]) > thr ? 1 : 0;
    } else {
      k = kfix[b] < kmax ? kfix[b] : kmax;
    }
    *s_k = k;
  }
  __syncthreads();
  R thr = kmax > 0 ? re_s(rdg[0]) * R(acc) : R(0);
  for (int round = 0;; ++round) {
    const int k = *s_k;
    // T = R11^-1 R12: t_i = r_i / R_ii, then r(0:i-1) -= R(0:i-1, i) t_i
    for (int c = wave; c < n - k; c += NW) {
      T r0 = lane < k ? rget(lane, k + c) : T(0);
      T r1 = lane + 64 < k ? rget(lane + 64, k + c) : T(0);
      T* tc = To + int64_t(c) * k;
      for (int i = k - 1; i >= 0; --i) {
        const T ri = readlane_s(i < 64 ? r0 : r1, i & 63);
        const T d = rget(i, i);
        const T ti = (d != T(0)) ? ri / d : T(0);  // a zero pivot contributes nothing
        if (lane == 0) tc[i] = ti;
        if (lane < i) r0 -= rget(lane, i) * ti;
        if (lane + 64 < i) r1 -= rget(lane + 64, i) * ti;
      }
    }
    __syncthreads();
    if (acc <= 0 || k >= kmax || k == n) break;
    // randomized check (ID.m:55-59,68-86), as in id_tail
    const int r_ = n - k;
    T* x = scratch;        // r_ entries
    T* y = x + n;          // m entries
    T* z = y + m;          // n entries
    R xs = R(0);
    for (int i = tid; i < r_; i += NT) {
      const R v = R(urand(seed ^ (uint64_t(b) << 20), uint64_t(round) * 4096 + i));
      x[i] = T(v);
      xs += v * v;
    }
    xs = bsum<NT>(xs, redr);
    const R xinv = R(1) / sqrt(xs);
    for (int i = tid; i < r_; i += NT) x[i] *= T(xinv);
    __syncthreads();
    for (int i = tid; i < k; i += NT) {  // u = T x
      T s = T(0);
      for (int c = 0; c < r_; ++c) s += To[i + int64_t(c) * k] * x[c];
      z[i] = s;
    }
    __syncthreads();
    for (int rr = tid; rr < m; rr += NT) {  // y = A_s u - A_r x
      T s = T(0);
#pragma unroll 4
      for (int i = 0; i < k; ++i) s += Mc[rr + int64_t(P[i]) * m] * z[i];
#pragma unroll 4
      for (int c = 0; c < r_; ++c) s -= Mc[rr + int64_t(P[k + c]) * m] * x[c];
      y[rr] = s;
    }
    __syncthreads();
    for (int i = wave; i < k; i += NW) {  // z = A_s^H y
      const T* col = Mc + int64_t(P[i]) * m;
      T s = T(0);
      for (int rr = lane; rr < m; rr += 64) s += conj_s(col[rr]) * y[rr];
      s = wsum(s);
      if (lane == 0) z[i] = s;
    }
    __syncthreads();
    R e2 = R(0);
    for (int c = wave; c < r_; c += NW) {  // e = T^H z - A_r^H y
      const T* col = Mc + int64_t(P[k + c]) * m;
      T s = T(0);
      for (int i = lane; i < k; i += 64) s += conj_s(To[i + int64_t(c) * k]) * z[i];
      for (int rr = lane; rr < m; rr += 64) s -= conj_s(col[rr]) * y[rr];
      s = wsum(s);
      if (lane == 0) e2 += abs2(s);
    }
    e2 = bsum<NT>(e2, redr);
    R xs2 = R(0);
    for (int i = tid; i < r_; i += NT) {
      const R v = R(urand(seed ^ (uint64_t(b) << 20) ^ 0x5555ull, uint64_t(round) * 4096 + i));
      x[i] = T(v);
      xs2 += v * v;
    }
    xs2 = bsum<NT>(xs2, redr);
    const R xinv2 = R(1) / sqrt(xs2);
    for (int i = tid; i < r_; i += NT) x[i] *= T(xinv2);
    __syncthreads();
    for (int rr = tid; rr < m; rr += NT) {
      T s = T(0);
#pragma unroll 4
      for (int c = 0; c < r_; ++c) s += Mc[rr + int64_t(P[k + c]) * m] * x[c];
      y[rr] = s;
    }
    __syncthreads();
    R n2 = R(0);
    for (int c = wave; c < r_; c += NW) {
      const T* col = Mc + int64_t(P[k + c]) * m;
      T t = T(0);
      for (int rr = lane; rr < m; rr += 64) t += conj_s(col[rr]) * y[rr];
      t = wsum(t);
      if (lane == 0) n2 += abs2(t);
    }
    n2 = bsum<NT>(n2, redr);
    const R err = sqrt(sqrt(e2)), nrmr = sqrt(sqrt(n2));
    if (!(err > nrmr * R(acc))) break;
    int kk = k;
    for (int g = 0; g < 4096 && kk == k && thr > R(0); ++g) {
      thr *= R(0.75);
      kk = 0;
      for (int i = 0; i < kmax; ++i) kk += re_s(rdg[i]) > thr ? 1 : 0;
    }
    __syncthreads();
    if (tid == 0) *s_k = kk;
    __syncthreads();
    if (kk == k) break;
  }
  if (tid == 0) kout[b] = *s_k;
}

// mode: acc > 0 -> rank by |R_ii| > acc |R_11| plus the randomized check (ID.m:40-65);
//       acc <= 0 -> fixed rank kfix[b] (ID.m:106-113, T by back substitution).
// Complex data: Householder reflectors H = I - tau v v^H applied as H^H (xgeqp3
// conventions, real R diagonal); the check uses conjugate transposes as ID.m's ' does.
template <typename T>
__global__ void __launch_bounds__(ID_NT) cpqr_id_kernel(
    int nmat, const int64_t* __restrict__ off, const int* __restrict__ mr, const int* __restrict__ nc,
    T* __restrict__ A, const T* __restrict__ Acopy, double acc, const int* __restrict__ kfix,
    int* __restrict__ perm, const int64_t* __restrict__ poff, int* __restrict__ kout,
    T* __restrict__ Tout, const int64_t* __restrict__ toff, T* __restrict__ work,
    const int64_t* __restrict__ woff, uint64_t seed, const int* __restrict__ kpin) {
  using R = real_t<T>;
  __shared__ R redr[ID_NT / 64];
  __shared__ R sbest[ID_NT / 64];
  __shared__ int sidx[ID_NT / 64];
  __shared__ int s_k;
  const int b = blockIdx.x;
  if (b >= nmat) return;
  const int m = mr[b], n = nc[b];
  const int pin = kpin ? kpin[b] : 0;  // leading columns taken in order, without pivoting
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
    R s = R(0);
    for (int r = 0; r < m; ++r) s += abs2(M[r + int64_t(c) * m]);
    nrm[c] = T(s);
  }
  __syncthreads();
  for (int j = 0; j < kmax; ++j) {
    // pivot: largest remaining column norm (ties -> lowest index); pinned columns in order
    R best = R(-1);
    int bi = j;
    if (j >= pin) {  // uniform over the workgroup
      for (int c = j + tid; c < n; c += ID_NT)
        if (re_s(nrm[c]) > best) best = re_s(nrm[c]), bi = c;
      for (int o = 32; o > 0; o >>= 1) {
        const R ob = shfl_xor_s(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ob > best || (ob == best && oi < bi)) best = ob, bi = oi;
      }
      if (lane == 0) sbest[wave] = best, sidx[wave] = bi;
      __syncthreads();
      best = sbest[0], bi = sidx[0];
      for (int w = 1; w < ID_NT / 64; ++w)
        if (sbest[w] > best || (sbest[w] == best && sidx[w] < bi)) best = sbest[w], bi = sidx[w];
    }
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
    R sg = R(0);
    for (int r = j + 1 + tid; r < m; r += ID_NT) sg += abs2(a[r]);
    sg = bsum<ID_NT>(sg, redr);
    const T alpha = a[j];
    T tv, scal;
    R beta;
    if (sg == R(0) && im_s(alpha) == R(0)) {
      tv = T(0), beta = re_s(alpha), scal = T(0);
    } else {
      const R nr = sqrt(abs2(alpha) + sg);
      beta = re_s(alpha) >= R(0) ? -nr : nr;
      tv = (T(beta) - alpha) / T(beta);
      scal = T(1) / (alpha - T(beta));
    }
    __syncthreads();
    for (int r = j + 1 + tid; r < m; r += ID_NT) a[r] *= scal;
    if (tid == 0) a[j] = T(beta), rdg[j] = T(fabs(beta)), tau[j] = tv;
    __syncthreads();
    // apply H^H to columns c > j (one wave per column) and recompute their trailing norms
    for (int c = j + 1 + wave; c < n; c += ID_NT / 64) {
      T* col = M + int64_t(c) * m;
      T s = T(0);
      for (int r = j + lane; r < m; r += 64) s += conj_s(r == j ? T(1) : a[r]) * col[r];
      s = wsum(s) * conj_s(tv);
      R q = R(0);
      for (int r = j + lane; r < m; r += 64) {
        const T v = col[r] - (r == j ? T(1) : a[r]) * s;
        col[r] = v;
        if (r > j) q += abs2(v);
      }
      q = wsum(q);
      if (lane == 0) nrm[c] = T(q);
    }
    __syncthreads();
  }
  id_tail<ID_NT>(b, m, n, M, Mc, P, rdg, wk + 3 * n, acc, kfix, kout, Tout + toff[b], seed, redr,
                 &s_k);
}

// Register-resident form of cpqr_id_kernel for blocks of up to NW RW rows and 64 CL
// columns (complex 256 x 120: the skeleton IDs of the bottom two levels).  Lanes own
// columns (c = lane + 64 s) and waves own rows (r = i NW + wave): row i of a wave sits in
// registers for i < RREG and in LDS above, so a column's dot product is a sum inside the lane
// plus an NW-way sum through LDS, with no cross-lane reductions.  Columns never move: the
// pivot order is kept in LDS and only R is formed (the reflectors are dropped), so the block
// is read once and R written once, where the global-memory kernel streams the trailing block
// through HBM once per column.  Same pivot rule (largest trailing norm, ties to the lowest
// position), reflector convention and rank rules as cpqr_id_kernel; sums are reduced in
// another order.  Per step: every wave picks the pivot from the NW partial norms of each
// column; the pivot column is gathered in LDS; every wave forms the same reflector from it,
// takes its partial dot products, and after their NW-way sum applies the reflector to its
// rows and leaves the partial norms below row j.
constexpr int IDR_NW = 8, IDR_RW = 32, IDR_RREG = 24, IDR_CL = 2;
// blocks of at most 192 x 64 (the leaf level): one column per lane, every row in registers,
// two workgroups per CU (four waves per SIMD)
constexpr int IDS_RW = 24, IDS_CL = 1;

template <typename T, int NW, int RW, int RREG, int CL, int OCC>
__global__ void __launch_bounds__(NW * 64, OCC) cpqr_id_reg_kernel(
    int nmat, const int64_t* __restrict__ off, const int* __restrict__ mr, const int* __restrict__ nc,
    T* __restrict__ A, const T* __restrict__ Acopy, double acc, const int* __restrict__ kfix,
    int* __restrict__ perm, const int64_t* __restrict__ poff, int* __restrict__ kout,
    T* __restrict__ Tout, const int64_t* __restrict__ toff, T* __restrict__ work,
    const int64_t* __restrict__ woff, uint64_t seed, int nld) {
  using R = real_t<T>;
  constexpr int NT = NW * 64, NCAP = 64 * CL, MCAP = NW * RW;
  // dynamic LDS, nld columns (>= every n): [part NW x nld T][part2 NW x nld R]
  // [rows RREG..RW-1 of every wave: (RW - RREG) NW x nld T, only when some m > NW RREG];
  // after the factorization the packed R of the tail (min(m, n) (min(m, n) + 1) / 2 T)
  extern __shared__ __align__(16) unsigned char idr_lds[];
  __shared__ T colbuf[MCAP];
  __shared__ R rd[NCAP];
  __shared__ int at[NCAP], cur[NCAP];  // column at each position / position of each column
  __shared__ R redr[NW];
  __shared__ int s_k;
  const int b = blockIdx.x;
  if (b >= nmat) return;
  const int m = mr[b], n = nc[b];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
  if (m > MCAP || n > NCAP || n > nld || m < 0 || n < 0) {  // the launcher checks the maxima
    if (tid == 0) kout[b] = -1;
    return;
  }
  T* M = A + off[b];
  const T* Mc = Acopy + off[b];
  int* P = perm + poff[b];
  T* wk = work + woff[b];
  T* part = reinterpret_cast<T*>(idr_lds);
  R* part2 = reinterpret_cast<R*>(part + NW * nld);
  T* la = reinterpret_cast<T*>(part2 + NW * nld);
  const bool lrow = m > NW * RREG;
  auto lidx = [&](int i, int c) { return ((i - RREG) * NW + wave) * nld + c; };
  T a[RREG][CL];
  uint32_t dn = 0;  // bit s: column lane + 64 s is pivoted (or absent)
#pragma unroll
  for (int s = 0; s < CL; ++s) {
    const int c = lane + 64 * s;
    R q = R(0);
    if (c >= n) dn |= 1u << s;
    const T* src = M + int64_t(c < n ? c : 0) * m + wave;  // rows wave + NW i
#pragma unroll
    for (int i = 0; i < RREG; ++i) {
      const int r = i * NW + wave;
      a[i][s] = (c < n && r < m) ? src[i * NW] : T(0);
      q += abs2(a[i][s]);
    }
    if (lrow && c < n)
      for (int i = RREG; i < RW; ++i) {
        const int r = i * NW + wave;
        const T v = r < m ? M[r + int64_t(c) * m] : T(0);
        la[lidx(i, c)] = v;
        q += abs2(v);
      }
    if (c < n) part2[wave * nld + c] = q, at[c] = c, cur[c] = c;
  }
  __syncthreads();
  const int kmax = m < n ? m : n;
  for (int j = 0; j < kmax; ++j) {
    // pivot, chosen by every wave alike: the largest remaining norm; norms within a relative
    // 1e-13 of it count as tied (mirror-symmetric points give equal norms up to rounding)
    // and the tie goes to the lowest current position, the positions being those the
    // column swaps of cpqr_id_kernel would give, so that both kernels break ties alike
    R best = R(-1), nv[CL];
#pragma unroll
    for (int s = 0; s < CL; ++s) {
      const int c = lane + 64 * s;
      nv[s] = R(-1);
      if ((dn >> s) & 1u) continue;
      R v = R(0);
      for (int w = 0; w < NW; ++w) v += part2[w * nld + c];
      if (v != v) v = R(0);
      nv[s] = v;
      best = v > best ? v : best;
    }
    for (int o = 32; o > 0; o >>= 1) {
      const R ob = shfl_xor_s(best, o);
      best = ob > best ? ob : best;
    }
    const R lim = best * R(1 - 1e-13);  // best >= 0: n - j >= 1 columns remain
    int bp = INT_MAX;
#pragma unroll
    for (int s = 0; s < CL; ++s) {
      const int c = lane + 64 * s;
      if (!((dn >> s) & 1u) && nv[s] >= lim && cur[c] < bp) bp = cur[c];
    }
    for (int o = 32; o > 0; o >>= 1) {
      const int ob = __shfl_xor(bp, o);
      bp = ob < bp ? ob : bp;
    }
    const int p = at[__builtin_amdgcn_readfirstlane(bp)];
    const int sp = p >> 6;
    const bool own = lane == (p & 63);
    if (own) dn |= 1u << sp;
    // gather the pivot column (rows of this wave; absent rows are zero)
    if (own) {
#pragma unroll
      for (int i = 0; i < RREG; ++i) {
        const int r = i * NW + wave;
        T v = T(0);
#pragma unroll
        for (int s = 0; s < CL; ++s)
          if (s == sp) v = a[i][s];
        colbuf[r] = v;
      }
      for (int i = RREG; i < RW; ++i) colbuf[i * NW + wave] = lrow ? la[lidx(i, p)] : T(0);
    }
    // every wave has read at[] and cur[] for this step before they change: a wave that read
    // them mid-swap would take another pivot column than the rest
    __syncthreads();
    if (tid == 0) {  // swap positions j and cur[p]
      const int cj = at[j], pp = cur[p];
      at[pp] = cj, cur[cj] = pp, at[j] = p, cur[p] = j;
    }
    __syncthreads();
    // the reflector, formed alike by every wave: v = [1; colbuf(j+1:m) scal] from row j
    R sg = R(0);
    for (int r = j + 1 + lane; r < m; r += 64) sg += abs2(colbuf[r]);
    sg = wsum(sg);
    const T alpha = colbuf[j];
    T tv, scal;
    R beta;
    if (sg == R(0) && im_s(alpha) == R(0)) {
      tv = T(0), beta = re_s(alpha), scal = T(0);
    } else {
      const R nr = sqrt(abs2(alpha) + sg);
      beta = re_s(alpha) >= R(0) ? -nr : nr;
      tv = (T(beta) - alpha) / T(beta);
      scal = T(1) / (alpha - T(beta));
    }
    if (tid == 0) rd[j] = fabs(beta);
    const T ctv = conj_s(tv);
    // the pivot column becomes R(0:j, p)
    if (own) {
#pragma unroll
      for (int i = 0; i < RREG; ++i) {
        const int r = i * NW + wave;
        if (r >= j) {
          const T w = r == j ? T(beta) : T(0);
#pragma unroll
          for (int s = 0; s < CL; ++s)
            if (s == sp) a[i][s] = w;
        }
      }
      if (lrow)
        for (int i = RREG; i < RW; ++i) {
          const int r = i * NW + wave;
          if (r >= j) la[lidx(i, p)] = r == j ? T(beta) : T(0);
        }
    }
    // v^H a = a_j + conj(scal) sum_{r > j} conj(colbuf_r) a_r: this wave's part (rows >= m
    // hold zeros), one column slot at a time (fewer live registers than all slots at once)
    const T cscal = conj_s(scal);
    // the pivot column's entry of this wave's row i: with one column slot, lane i holds it
    // and v_readlane broadcasts it (no LDS latency per row); with more, the registers are
    // taken by the block and it is read from LDS
    const T cbm = (CL == 1 && lane < RW) ? colbuf[lane * NW + wave] : T(0);
    auto cbr = [&](int i) -> T {
      if constexpr (CL == 1)
        return readlane_s(cbm, i);
      else
        return colbuf[i * NW + wave];
    };
    T d[CL], aj[CL];
#pragma unroll
    for (int s = 0; s < CL; ++s) d[s] = T(0), aj[s] = T(0);
#pragma unroll
    for (int i = 0; i < RREG; ++i) {
      const int r = i * NW + wave;
      if (r > j) {
        const T cb = conj_s(cbr(i));
#pragma unroll
        for (int s = 0; s < CL; ++s) d[s] += cb * a[i][s];
      } else if (r == j) {
#pragma unroll
        for (int s = 0; s < CL; ++s) aj[s] = a[i][s];
      }
    }
    if (lrow)
      for (int i = RREG; i < RW; ++i) {
        const int r = i * NW + wave;
        if (r > j) {
          const T cb = conj_s(cbr(i));
#pragma unroll
          for (int s = 0; s < CL; ++s)
            if (lane + 64 * s < n) d[s] += cb * la[lidx(i, lane + 64 * s)];
        } else if (r == j) {
#pragma unroll
          for (int s = 0; s < CL; ++s)
            if (lane + 64 * s < n) aj[s] = la[lidx(i, lane + 64 * s)];
        }
      }
#pragma unroll
    for (int s = 0; s < CL; ++s)
      if (lane + 64 * s < n) part[wave * nld + lane + 64 * s] = d[s] * cscal + aj[s];
    __syncthreads();
    // apply H^H: a_j -= sc, a_r -= colbuf_r (scal sc) below, sc = conj(tau) v^H a (zero for
    // pivoted columns, which stay as they are); leave the partial norms below row j
    T sc[CL], vs[CL];
    R q[CL];
#pragma unroll
    for (int s = 0; s < CL; ++s) {
      const int c = lane + 64 * s;
      T t = T(0);
      if (!((dn >> s) & 1u))
        for (int w = 0; w < NW; ++w) t += part[w * nld + c];
      sc[s] = t * ctv;
      vs[s] = scal * sc[s];
      q[s] = R(0);
    }
#pragma unroll
    for (int i = 0; i < RREG; ++i) {
      const int r = i * NW + wave;
      if (r > j) {
        const T cb = cbr(i);
#pragma unroll
        for (int s = 0; s < CL; ++s) {
          a[i][s] -= cb * vs[s];
          q[s] += abs2(a[i][s]);
        }
      } else if (r == j) {
#pragma unroll
        for (int s = 0; s < CL; ++s) a[i][s] -= sc[s];
      }
    }
    if (lrow)
      for (int i = RREG; i < RW; ++i) {
        const int r = i * NW + wave;
        if (r < j) continue;
        const T cb = cbr(i);
#pragma unroll
        for (int s = 0; s < CL; ++s) {
          const int c = lane + 64 * s;
          if (c < n) {
            const T x = la[lidx(i, c)] - (r == j ? sc[s] : cb * vs[s]);
            la[lidx(i, c)] = x;
            if (r > j) q[s] += abs2(x);
          }
        }
      }
#pragma unroll
    for (int s = 0; s < CL; ++s)
      if (lane + 64 * s < n) part2[wave * nld + lane + 64 * s] = q[s];
    __syncthreads();
  }
  // R: rows 0..min(position, kmax - 1) of every column at its position (columns never
  // pivoted, n > m, stay where the swaps left them)
#pragma unroll
  for (int s = 0; s < CL; ++s) {
    const int c = lane + 64 * s;
    if (c < n) {
      const int q = cur[c];
      const int rlast = q < kmax ? q : kmax - 1;
      T* dst = M + int64_t(q) * m + wave;  // rows wave + NW i
#pragma unroll
      for (int i = 0; i < RREG; ++i)
        if (i * NW + wave <= rlast) dst[i * NW] = a[i][s];
      if (lrow)
        for (int i = RREG; i < RW; ++i)
          if (i * NW + wave <= rlast) dst[i * NW] = la[lidx(i, c)];
    }
  }
  T* rdg = wk + n;
  for (int i = tid; i < n; i += NT) P[i] = at[i];
  for (int i = tid; i < kmax; i += NT) rdg[i] = T(rd[i]);
  __syncthreads();
  // R(0:q, q) of the pivot positions, packed in the (now free) LDS for the tail
  T* Rl = reinterpret_cast<T*>(idr_lds);
  for (int q = wave; q < kmax; q += NW) {
    const T* src = M + int64_t(q) * m;
    T* dst = Rl + q * (q + 1) / 2;
    for (int i = lane; i <= q; i += 64) dst[i] = src[i];
  }
  __syncthreads();
  id_tail_wave<NT>(b, m, n, M, Mc, P, rdg, wk + 3 * n, acc, kfix, kout, Tout + toff[b], seed,
                   redr, &s_k, Rl);
}

// Acceptance test of IDs computed on sketches, on the full blocks A_b (m x n, column-major at
// A + off[b], n <= 256), the check of Common/ID.m:55-59 with one random vector: x a random
// unit vector over the residual columns and u = T x, e = |A_s u - A_r x| (A_s the k skeleton
// columns, A_r the others, in the order perm); g a random unit vector over all columns,
// a = |A g|.  ok[b] = 1 when e <= 10 acc a and the sketch was not saturated (k < s, unless
// the s sketch rows already cover min(m, n)).  One workgroup per block.
template <typename T>
__global__ void __launch_bounds__(256) id_verify_kernel(
    const int64_t* __restrict__ off, const int* __restrict__ mr, const int* __restrict__ nc,
    const int* __restrict__ sr, const T* __restrict__ A, const int* __restrict__ perm,
    const int64_t* __restrict__ poff, const int* __restrict__ kk, const T* __restrict__ Tout,
    const int64_t* __restrict__ toff, double acc, uint64_t seed, int* __restrict__ ok) {
  using R = real_t<T>;
  __shared__ R red[4];
  __shared__ T xs[256], us[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int m = mr[b], n = nc[b], s = sr[b], k = kk[b], r_ = n - k;
  const T* M = A + off[b];
  const int* P = perm + poff[b];
  const T* Tb = Tout + toff[b];
  auto bsum = [&](R v) {
    for (int o = 32; o > 0; o >>= 1) v += shfl_xor_s(v, o);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
  };
  auto gauss = [&](uint64_t key, uint64_t i) {  // Box-Muller
    const double u1 = fmax(urand(key, 2 * i), 1e-300), u2 = urand(key, 2 * i + 1);
    return R(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  };
  const uint64_t key = seed ^ (uint64_t(b) << 24) ^ 0x7777ull;
  R v = 0;
  if (tid < r_) {
    const R g = gauss(key, tid);
    xs[tid] = T(g), v = g * g;
  }
  const R xn = sqrt(bsum(v));
  if (tid < r_) xs[tid] = xs[tid] * T(R(1) / xn);
  __syncthreads();
  if (tid < k) {
    T t = T(0);
    for (int c = 0; c < r_; ++c) t += Tb[tid + int64_t(c) * k] * xs[c];
    us[tid] = t;
  }
  __syncthreads();
  R e2 = 0;
  for (int r = tid; r < m; r += 256) {
    T t = T(0);
    for (int i = 0; i < k; ++i) t += M[r + int64_t(P[i]) * m] * us[i];
    for (int c = 0; c < r_; ++c) t -= M[r + int64_t(P[k + c]) * m] * xs[c];
    e2 += abs2(t);
  }
  e2 = bsum(e2);
  __syncthreads();
  v = 0;
  if (tid < n) {
    const R g = gauss(key ^ 0x5555ull, tid);
    xs[tid] = T(g), v = g * g;
  }
  const R gn = sqrt(bsum(v));
  if (tid < n) xs[tid] = xs[tid] * T(R(1) / gn);
  __syncthreads();
  R a2 = 0;
  for (int r = tid; r < m; r += 256) {
    T t = T(0);
    for (int c = 0; c < n; ++c) t += M[r + int64_t(c) * m] * xs[c];
    a2 += abs2(t);
  }
  a2 = bsum(a2);
  if (tid == 0) {
    const int full = m < n ? m : n;
    const bool saturated = k >= s && s < full;
    ok[b] = (n <= 256 && !saturated && sqrt(double(e2)) <= 10.0 * acc * sqrt(double(a2))) ? 1 : 0;
  }
}

template <typename T>
void id_verify(hipStream_t st, int nmat, const int64_t* off, const int* mr, const int* nc,
               const int* sr, const T* A, const int* perm, const int64_t* poff, const int* kk,
               const T* Tout, const int64_t* toff, double acc, uint64_t seed, int* ok) {
  if (nmat <= 0) return;
  hipLaunchKernelGGL(id_verify_kernel<T>, dim3(nmat), dim3(256), 0, st, off, mr, nc, sr, A, perm,
                     poff, kk, Tout, toff, acc, seed, ok);
}
template void id_verify<zdouble>(hipStream_t, int, const int64_t*, const int*, const int*,
                                 const int*, const zdouble*, const int*, const int64_t*,
                                 const int*, const zdouble*, const int64_t*, double, uint64_t,
                                 int*);

// D_j (rows[j] x n, ld rows[j], at doff[j]) = the first rows[j] rows of S_j (ld lds, at
// soff[j]).  One workgroup per block.
template <typename T>
__global__ void __launch_bounds__(256) pack_rows_kernel(
    const int64_t* __restrict__ soff, int64_t lds, const int64_t* __restrict__ doff,
    const int* __restrict__ rows, int64_t n, const T* __restrict__ S, T* __restrict__ D) {
  const int j = blockIdx.x;
  const int64_t r = rows[j], len = r * n;
  const T* s = S + soff[j];
  T* d = D + doff[j];
  for (int64_t e = threadIdx.x; e < len; e += 256) d[e] = s[(e % r) + (e / r) * lds];
}

template <typename T>
void pack_rows(hipStream_t st, int nj, const int64_t* soff, int64_t lds, const int64_t* doff,
               const int* rows, int64_t n, const T* S, T* D) {
  if (nj <= 0) return;
  hipLaunchKernelGGL(pack_rows_kernel<T>, dim3(nj), dim3(256), 0, st, soff, lds, doff, rows, n, S,
                     D);
}
template void pack_rows<zdouble>(hipStream_t, int, const int64_t*, int64_t, const int64_t*,
                                 const int*, int64_t, const zdouble*, zdouble*);

template <typename T>
void cpqr_id(hipStream_t st, int nmat, const int64_t* off, const int* mr, const int* nc, T* A,
             const T* Acopy, double acc, const int* kfix, int* perm, const int64_t* poff, int* kout,
             T* Tout, const int64_t* toff, T* work, const int64_t* woff, uint64_t seed,
             int64_t mmax, int64_t nmax, const int* kpin) {
  if (nmat <= 0) return;
  // The register-resident kernels run when HSSC_ID_KERNEL=register (read at every call).
  // They give IDs of the same quality 1.7x faster, but their sums round differently, and the
  // pivots past the numerical rank (trailing norms near acc^2 |A|^2, rounding noise) then
  // differ from the global-memory kernel's; on some trees (the 64^2 Lippmann-Schwinger grid)
  // the skeletons chosen that way condition the upper levels worse (complex64 refinement
  // stalls), so the global-memory kernel stays the default.
  const char* sel = std::getenv("HSSC_ID_KERNEL");
  const bool use_reg = sel && std::strcmp(sel, "register") == 0 && !kpin;
  if constexpr (is_complex_v<T>) {
    if (!use_reg) mmax = 0;
    if (mmax > 0 && nmax > 0 && mmax <= IDR_NW * IDS_RW && nmax <= 64 * IDS_CL) {
      const size_t kcap = size_t(std::min(mmax, nmax));
      const size_t lds = std::max(size_t(nmax) * IDR_NW * (sizeof(T) + sizeof(real_t<T>)),
                                  kcap * (kcap + 1) / 2 * sizeof(T));
      if (lds <= 64 * 1024) {
        HSSC_TIMED("cpqr_id_reg_c128", st, 0.0, 0.0,
                   hipLaunchKernelGGL((cpqr_id_reg_kernel<T, IDR_NW, IDS_RW, IDS_RW, IDS_CL, 4>),
                                      dim3(nmat), dim3(IDR_NW * 64), lds, st, nmat, off, mr, nc, A,
                                      Acopy, acc, kfix, perm, poff, kout, Tout, toff, work, woff,
                                      seed, int(nmax)));
        return;
      }
    }
    if (mmax > 0 && nmax > 0 && mmax <= IDR_NW * IDR_RW && nmax <= 64 * IDR_CL) {
      auto* kern = cpqr_id_reg_kernel<T, IDR_NW, IDR_RW, IDR_RREG, IDR_CL, 2>;
      const size_t lrows = mmax > IDR_NW * IDR_RREG ? size_t(IDR_RW - IDR_RREG) * IDR_NW : 0;
      const size_t kcap = size_t(std::min(mmax, nmax));
      const size_t lds = std::max(
          size_t(nmax) * (IDR_NW * (sizeof(T) + sizeof(real_t<T>)) + lrows * sizeof(T)),
          kcap * (kcap + 1) / 2 * sizeof(T));  // the factorization's, then packed R
      // dynamic LDS above 64 KiB must be allowed first; when the device refuses (static plus
      // dynamic above its LDS), the global-memory kernel runs instead
      static size_t allowed = 64 * 1024;
      bool fits = lds <= allowed;
      if (!fits) {
        fits = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) == hipSuccess;
        if (fits)
          allowed = lds;
        else
          (void)hipGetLastError();
      }
      if (fits) {
        HSSC_TIMED("cpqr_id_reg_c128", st, 0.0, 0.0,
                   hipLaunchKernelGGL(kern, dim3(nmat), dim3(IDR_NW * 64), lds, st, nmat, off, mr,
                                      nc, A, Acopy, acc, kfix, perm, poff, kout, Tout, toff, work,
                                      woff, seed, int(nmax)));
        return;
      }
    }
  }
  HSSC_TIMED(is_complex_v<T> ? "cpqr_id_c128" : "cpqr_id_f64", st, 0.0, 0.0,
             hipLaunchKernelGGL(cpqr_id_kernel<T>, dim3(nmat), dim3(ID_NT), 0, st, nmat, off, mr,
                                nc, A, Acopy, acc, kfix, perm, poff, kout, Tout, toff, work, woff,
                                seed, kpin));
}

template void cpqr_id<double>(hipStream_t, int, const int64_t*, const int*, const int*, double*,
                              const double*, double, const int*, int*, const int64_t*, int*,
                              double*, const int64_t*, double*, const int64_t*, uint64_t,
                              int64_t, int64_t, const int*);
template void cpqr_id<zdouble>(hipStream_t, int, const int64_t*, const int*, const int*, zdouble*,
                               const zdouble*, double, const int*, int*, const int64_t*, int*,
                               zdouble*, const int64_t*, zdouble*, const int64_t*, uint64_t,
                               int64_t, int64_t, const int*);

// Standard normal numbers (Box-Muller on the counter-based uniforms): G[i], i < len,
// element i drawn from (seed, base + i).
__global__ void __launch_bounds__(256) randn_kernel(int64_t len, uint64_t seed, uint64_t base,
                                                    double* __restrict__ G) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= len) return;
  const uint64_t c = base + uint64_t(i);
  double u1 = urand(seed, 2 * c), u2 = urand(seed, 2 * c + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  G[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void randn(hipStream_t st, int64_t len, uint64_t seed, uint64_t base, double* G) {
  if (len <= 0) return;
  hipLaunchKernelGGL(randn_kernel, dim3((len + 255) / 256), dim3(256), 0, st, len, seed, base, G);
}

}  // namespace k

namespace {

struct IdErr {
  int status;
};

#define IDCHK(expr)                                                                 \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      throw IdErr{fail(e_ == hipErrorOutOfMemory ? HSSC_ERR_OUT_OF_MEMORY : HSSC_ERR_HIP, \
                       std::string(#expr) + ": " + hipGetErrorString(e_))};        \
    }                                                                               \
  } while (0)

template <typename T>
T* idalloc(size_t n) {
  void* p = nullptr;
  IDCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

// One matrix m x n (column-major, device, ld m) through the batched kernel.  mode acc > 0:
// adaptive rank; acc <= 0: fixed rank kfix.  Returns rank, permutation, T (k x (n-k)) and
// the |R_ii| of the pivoted QR.
struct IdOut {
  int k = 0;
  std::vector<int> perm;
  std::vector<double> T, rdiag;
};

IdOut id_device(double* dA, int m, int n, double acc, int kfix, uint64_t seed) {
  std::vector<void*> tmp;
  auto al = [&](size_t bytes) {
    void* p = nullptr;
    IDCHK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
    tmp.push_back(p);
    return p;
  };
  IdOut out;
  try {
    const int64_t zero = 0;
    int64_t* doff = static_cast<int64_t*>(al(8));
    int* dm = static_cast<int*>(al(4));
    int* dn = static_cast<int*>(al(4));
    int* dkf = static_cast<int*>(al(4));
    int* dk = static_cast<int*>(al(4));
    int* dperm = static_cast<int*>(al(size_t(n) * 4));
    double* dT = static_cast<double*>(al(size_t(n) * n * 8));
    double* dW = static_cast<double*>(al((6 * size_t(n) + m) * 8));
    double* dC = static_cast<double*>(al(size_t(m) * n * 8));
    IDCHK(hipMemcpy(doff, &zero, 8, hipMemcpyHostToDevice));
    IDCHK(hipMemcpy(dm, &m, 4, hipMemcpyHostToDevice));
    IDCHK(hipMemcpy(dn, &n, 4, hipMemcpyHostToDevice));
    IDCHK(hipMemcpy(dkf, &kfix, 4, hipMemcpyHostToDevice));
    IDCHK(hipMemcpy(dC, dA, size_t(m) * n * 8, hipMemcpyDeviceToDevice));
    k::cpqr_id<double>(0, 1, doff, dm, dn, dA, dC, acc > 0 ? acc : -1.0, dkf, dperm, doff, dk, dT,
                       doff, dW, doff, seed);
    IDCHK(hipDeviceSynchronize());
    IDCHK(hipGetLastError());
    IDCHK(hipMemcpy(&out.k, dk, 4, hipMemcpyDeviceToHost));
    out.perm.resize(n);
    IDCHK(hipMemcpy(out.perm.data(), dperm, size_t(n) * 4, hipMemcpyDeviceToHost));
    out.T.resize(size_t(out.k) * (n - out.k));
    if (!out.T.empty())
      IDCHK(hipMemcpy(out.T.data(), dT, out.T.size() * 8, hipMemcpyDeviceToHost));
    const int kmax = std::min(m, n);
    out.rdiag.resize(kmax);
    if (kmax) IDCHK(hipMemcpy(out.rdiag.data(), dW + n, size_t(kmax) * 8, hipMemcpyDeviceToHost));
  } catch (...) {
    for (void* p : tmp) (void)hipFree(p);
    throw;
  }
  for (void* p : tmp) (void)hipFree(p);
  return out;
}

double* upload_matrix(const double* A, int64_t m, int64_t n, int64_t lda) {
  double* d = idalloc<double>(size_t(m) * n);
  IDCHK(hipMemcpy2D(d, size_t(m) * 8, A, size_t(lda) * 8, size_t(m) * 8, n, hipMemcpyHostToDevice));
  return d;
}

}  // namespace
}  // namespace hssc

using namespace hssc;

extern "C" {

int hssc_id(const double* A, int64_t m, int64_t n, int64_t lda, double par, int64_t* perm,
            double* T, int64_t* krank) {
  // Common/ID.m:1 ([T,I] = ID(A,par)): par < 1 is the accuracy, par >= 1 the fixed rank.
  if (!A || !perm || !T || !krank || m < 1 || n < 1 || lda < m || m > INT32_MAX || n > 65536)
    return fail(HSSC_ERR_INVALID_ARG, "invalid ID arguments");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(HSSC_ERR_HIP, "no HIP device is available");
  double* dA = nullptr;
  try {
    dA = upload_matrix(A, m, n, lda);
    const bool fixed = par >= 1;
    IdOut o = id_device(dA, int(m), int(n), fixed ? -1.0 : par, fixed ? int(par) : 0, 0x1D5EEDull);
    (void)hipFree(dA);
    *krank = o.k;
    for (int64_t i = 0; i < n; ++i) perm[i] = o.perm[i];
    std::copy(o.T.begin(), o.T.end(), T);
    return HSSC_OK;
  } catch (const IdErr& e) {
    if (dA) (void)hipFree(dA);
    return e.status;
  }
}

int hssc_lrid_rand(const double* A, int64_t m, int64_t n, int64_t lda, double acc, int64_t k0,
                   int64_t Cinc, uint64_t seed, int64_t* krank, int64_t* J, double* U, double* V) {
  // Common/LRID_rand.m:10-40 with dense matvecs: Y = A^T G (Gaussian G, QR_findrank.m
  // with the sketch grown by C columns while the last pivoted-QR diagonal of Y exceeds acc),
  // then the fixed-rank ID of Z = Y(:,1:k)^T (ID_rand.m:41-49); A ~ U V^T with
  // U = A(:, J(1:k)) and V(J,:) = [I T]^T.  U (m x k) and V (n x k) need min(m,n) columns.
  if (!A || !krank || !J || !U || !V || m < 1 || n < 1 || lda < m || !(acc > 0) || k0 < 1 ||
      Cinc < 1 || n > 65536)
    return fail(HSSC_ERR_INVALID_ARG, "invalid LRID_rand arguments");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(HSSC_ERR_HIP, "no HIP device is available");
  std::vector<void*> tmp;
  try {
    double* dA = upload_matrix(A, m, n, lda);
    tmp.push_back(dA);
    const int64_t kmax = std::min(m, n);
    int64_t kk = std::min(k0, kmax);
    double* G = idalloc<double>(size_t(m) * kmax);
    double* Y = idalloc<double>(size_t(n) * kmax);
    double* Yc = idalloc<double>(size_t(n) * kmax);
    tmp.push_back(G), tmp.push_back(Y), tmp.push_back(Yc);
    int64_t have = 0, krk = 0;
    for (;;) {
      // extend the sketch to kk columns: G(:, have:kk) ~ N(0,1), Y(:, have:kk) = A^T G
      k::randn(0, m * (kk - have), seed, uint64_t(m) * have, G + m * have);
      k::gemm<double>(0, true, false, n, kk - have, m, 1.0, dA, m, 0, G + m * have, m, 0, 0.0,
                      Y + n * have, n, 0, 1);
      have = kk;
      IDCHK(hipMemcpy(Yc, Y, size_t(n) * kk * 8, hipMemcpyDeviceToDevice));
      IdOut q = id_device(Yc, int(n), int(kk), -1.0, int(kk), seed ^ 0x51ull);
      krk = 0;
      for (double d : q.rdiag) krk += d > acc ? 1 : 0;
      const double last = q.rdiag.empty() ? 0.0 : q.rdiag.back();
      if (!(last > acc) || kk >= kmax) break;
      kk = std::min(kk + Cinc, kmax);
    }
    if (krk == 0) {
      *krank = 0;
      return HSSC_OK;
    }
    // Z = Y(:, 1:krk)^T  (krk x n), fixed-rank ID
    std::vector<double> hY(size_t(n) * krk), hZ(size_t(krk) * n);
    IDCHK(hipMemcpy(hY.data(), Y, hY.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < krk; ++j) hZ[j + i * krk] = hY[i + j * n];
    double* dZ = idalloc<double>(hZ.size());
    tmp.push_back(dZ);
    IDCHK(hipMemcpy(dZ, hZ.data(), hZ.size() * 8, hipMemcpyHostToDevice));
    IdOut z = id_device(dZ, int(krk), int(n), -1.0, int(krk), seed ^ 0x77ull);
    const int64_t k = z.k;
    *krank = k;
    for (int64_t i = 0; i < n; ++i) J[i] = z.perm[i];
    for (int64_t j = 0; j < k; ++j)
      for (int64_t i = 0; i < m; ++i) U[i + j * m] = A[i + int64_t(z.perm[j]) * lda];
    for (int64_t i = 0; i < n * k; ++i) V[i] = 0;
    for (int64_t j = 0; j < k; ++j) {
      V[z.perm[j] + j * n] = 1.0;
      for (int64_t c = 0; c < n - k; ++c) V[z.perm[k + c] + j * n] = z.T[j + c * k];
    }
    for (void* p : tmp) (void)hipFree(p);
    return HSSC_OK;
  } catch (const IdErr& e) {
    for (void* p : tmp) (void)hipFree(p);
    return e.status;
  }
}

}  // extern "C"

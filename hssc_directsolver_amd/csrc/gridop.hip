// The exact operator of a uniform grid, A = K(X_source, X_source) of Common/Kernel_Eval.m,
// applied in O(N log N) by FFT convolution.
//
// On the grid of Common/HSS_tree_parameters.m:144-160 (G x G cell-centred points, spacing h)
// the kernel depends on the offset of two points only, so
//     (A x)_i = d x_i + b_i sum_{j != i} K0(p_i - p_j) c_j x_j,
// with d the diagonal (1, or w = 1 + k^2 (i/4) corr0), b and c the per-point weights and K0
// the kernel entry without weights.  The sum is a 2D linear convolution: x is placed (times
// c) in an L x L array, L >= 2G - 1 a power of two, and multiplied in Fourier space by the
// transform of K0 wrapped circularly onto the same array (offsets -(G-1)..G-1).
//
// The 2D transform is row FFTs in LDS (radix-2 Stockham, one workgroup per row), a tiled
// transpose, and row FFTs again, so the spectrum lives transposed ([kx][ky]); the inverse
// runs the same steps backwards.  Rows that are zero padding are never transformed.

#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "gridop.h"
#include "internal.h"
#include "kernels.h"
#include "pool.h"
#include "stats.h"

namespace hssc {
k::KernelSpec kernel_spec(const hssc_params& p);
}

struct hssc_grid_op {
  hssc_params p{};
  int device = 0;
  int64_t N = 0;
  int G = 0, L = 0, Ng = 0, np = 0;
  bool cx = false;
  double dre = 1, dim = 0;        // diagonal
  double* b = nullptr;            // N row weights (X_source order)
  double* c = nullptr;            // N column weights
  hssc::zdouble* khat = nullptr;  // L x L spectrum of K0, [kx][ky], times 1 / L^2
  hssc::zdouble* tw = nullptr;    // L / 2 twiddles exp(-2 pi i k / L)
  hssc::zdouble* U = nullptr;     // L x L work arrays
  hssc::zdouble* V = nullptr;
  uint64_t uid = 0;               // unique per operator of the process
  ~hssc_grid_op() {
    (void)hipDeviceSynchronize();
    for (void* q : {(void*)b, (void*)c, (void*)khat, (void*)tw, (void*)U, (void*)V})
      hssc::pool::release(q);
  }
};

namespace hssc {
namespace {

struct GridError {
  int status;
};

#define GCHK(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw GridError{fail(e_ == hipErrorOutOfMemory ? HSSC_ERR_OUT_OF_MEMORY : HSSC_ERR_HIP, \
                           std::string(#expr) + ": " + hipGetErrorString(e_))};           \
  } while (0)

template <typename T>
T* galloc(size_t count) {
  void* p = nullptr;
  GCHK(pool::alloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

constexpr int FFT_T = 256;

// In-place FFT of rows [0, nrows) of A (length L, row stride ld): radix-2 Stockham
// autosort through two LDS buffers (stage with sub-transform size ns: butterfly j pairs
// elements j and j + L/2, twiddle exp(-+2 pi i (j mod ns) / (2 ns)), outputs at
// 2 ns floor(j / ns) + (j mod ns) and + ns), natural order in and out.
template <bool INV>
__global__ void __launch_bounds__(FFT_T) fft_rows_kernel(int L, zdouble* __restrict__ A, int64_t ld,
                                                         const zdouble* __restrict__ tw) {
  extern __shared__ zdouble lds[];
  zdouble* src = lds;
  zdouble* dst = lds + L;
  zdouble* row = A + int64_t(blockIdx.x) * ld;
  const int tid = threadIdx.x, half = L >> 1;
  for (int i = tid; i < L; i += FFT_T) src[i] = row[i];
  __syncthreads();
  for (int ns = 1; ns < L; ns <<= 1) {
    const int tstep = half / ns;  // table stride: exp(-2 pi i k / (2 ns)) = tw[k * L / (2 ns)]
    for (int j = tid; j < half; j += FFT_T) {
      const int kk = j & (ns - 1);
      zdouble w = tw[kk * tstep];
      if (INV) w.im = -w.im;
      const zdouble a0 = src[j];
      const zdouble a1 = src[j + half] * w;
      const int d = ((j - kk) << 1) + kk;
      dst[d] = a0 + a1;
      dst[d + ns] = a0 - a1;
    }
    __syncthreads();
    zdouble* t = src;
    src = dst;
    dst = t;
  }
  for (int i = tid; i < L; i += FFT_T) row[i] = src[i];
}

// D[c][r] = S[r][c] for r < R, c < C (rows r >= rvalid of S read as zeros); 32 x 32 tiles.
__global__ void __launch_bounds__(256) transpose_kernel(int64_t R, int64_t C, int64_t rvalid,
                                                        const zdouble* __restrict__ S, int64_t lds,
                                                        zdouble* __restrict__ D, int64_t ldd) {
  __shared__ zdouble tile[32][33];
  const int64_t r0 = int64_t(blockIdx.y) * 32, c0 = int64_t(blockIdx.x) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32 per pass
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < R && c < C) tile[i][tx] = r < rvalid ? S[r * lds + c] : zdouble(0.0);
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int64_t c = c0 + i, r = r0 + tx;
    if (r < R && c < C) D[c * ldd + r] = tile[tx][i];
  }
}

// Grid position (ix, iy) of X_source point g (panel-major, y fastest) and back.
__device__ __forceinline__ int64_t grid_index(int ix, int iy, int Ng, int np) {
  const int px = ix / Ng, jx = ix - px * Ng, py = iy / Ng, jy = iy - py * Ng;
  return (int64_t(py) + int64_t(px) * np) * Ng * Ng + int64_t(jx) * Ng + jy;
}

// U[iy][ix] = c_g x_g for ix < G, 0 for G <= ix < L (rows iy < G).
template <typename S>
__global__ void __launch_bounds__(256) pack_kernel(int G, int L, int Ng, int np,
                                                   const S* __restrict__ x,
                                                   const double* __restrict__ c,
                                                   zdouble* __restrict__ U) {
  const int ix = blockIdx.x * 256 + threadIdx.x, iy = blockIdx.y;
  if (ix >= L) return;
  zdouble v(0.0);
  if (ix < G) {
    const int64_t g = grid_index(ix, iy, Ng, np);
    v = zdouble(x[g]) * zdouble(c[g]);
  }
  U[int64_t(iy) * L + ix] = v;
}

__device__ __forceinline__ void store_out(double* y, int64_t i, zdouble v) { y[i] = v.re; }
__device__ __forceinline__ void store_out(zdouble* y, int64_t i, zdouble v) { y[i] = v; }

// y_g = d x_g + b_g U[iy][ix] (or B_g minus that).
template <typename S>
__global__ void __launch_bounds__(256) unpack_kernel(int G, int L, int Ng, int np, double dre,
                                                     double dim, const S* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     const zdouble* __restrict__ U,
                                                     const S* __restrict__ B, S* __restrict__ y) {
  const int ix = blockIdx.x * 256 + threadIdx.x, iy = blockIdx.y;
  if (ix >= G) return;
  const int64_t g = grid_index(ix, iy, Ng, np);
  zdouble v = zdouble(dre, dim) * zdouble(x[g]) + zdouble(b[g]) * U[int64_t(iy) * L + ix];
  if (B) v = zdouble(B[g]) - v;
  store_out(y, g, v);
}

// Wrapped kernel K0 on the L x L array: U[ky][kx] = K0(dx h, dy h), offsets dx = kx or
// kx - L (|dx| < G), zero at the origin and on offsets beyond the grid.
__global__ void __launch_bounds__(256) kernel_wrap_kernel(int G, int L, double h, k::KernelSpec ks,
                                                          zdouble* __restrict__ U) {
  const int kx = blockIdx.x * 256 + threadIdx.x, ky = blockIdx.y;
  if (kx >= L) return;
  const int dx = kx < G ? kx : (kx > L - G ? kx - L : INT32_MAX);
  const int dy = ky < G ? ky : (ky > L - G ? ky - L : INT32_MAX);
  zdouble v(0.0);
  if (dx != INT32_MAX && dy != INT32_MAX && (dx != 0 || dy != 0))
    v = k::kernel_entry<zdouble>(ks, dx * h, dy * h, 1.0, 0.0, 0.0, 1.0);
  U[int64_t(ky) * L + kx] = v;
}

__global__ void __launch_bounds__(256) scale_kernel(int64_t len, double s, zdouble* __restrict__ A) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < len) A[i] = A[i] * zdouble(s);
}

__global__ void __launch_bounds__(256) mult_kernel(int64_t len, const zdouble* __restrict__ K,
                                                   zdouble* __restrict__ A) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < len) A[i] = A[i] * K[i];
}

void fft_rows(hipStream_t st, bool inv, int L, int64_t nrows, zdouble* A, int64_t ld,
              const zdouble* tw) {
  if (nrows <= 0) return;
  const size_t shm = size_t(2) * L * sizeof(zdouble);
  const double ent = double(L) * nrows;
  HSSC_TIMED("fft_c128", st, 5.0 * ent * std::log2(double(L)), 2.0 * ent * sizeof(zdouble), {
    if (inv)
      hipLaunchKernelGGL(fft_rows_kernel<true>, dim3(nrows), dim3(FFT_T), shm, st, L, A, ld, tw);
    else
      hipLaunchKernelGGL(fft_rows_kernel<false>, dim3(nrows), dim3(FFT_T), shm, st, L, A, ld, tw);
  });
}

void transpose(hipStream_t st, int64_t R, int64_t C, int64_t rvalid, const zdouble* S, int64_t lds,
               zdouble* D, int64_t ldd) {
  hipLaunchKernelGGL(transpose_kernel, dim3((C + 31) / 32, (R + 31) / 32), dim3(256), 0, st, R, C,
                     rvalid, S, lds, D, ldd);
}

template <typename S>
void apply_one(const hssc_grid_op* op, const S* x, const S* B, S* y, hipStream_t st) {
  const int G = op->G, L = op->L;
  const int64_t LL = int64_t(L) * L;
  hipLaunchKernelGGL(pack_kernel<S>, dim3((L + 255) / 256, G), dim3(256), 0, st, G, L, op->Ng,
                     op->np, x, op->c, op->U);
  fft_rows(st, false, L, G, op->U, L, op->tw);
  transpose(st, L, L, G, op->U, L, op->V, L);  // rows >= G of U are the zero padding
  fft_rows(st, false, L, L, op->V, L, op->tw);
  hipLaunchKernelGGL(mult_kernel, dim3((LL + 255) / 256), dim3(256), 0, st, LL, op->khat, op->V);
  fft_rows(st, true, L, L, op->V, L, op->tw);
  transpose(st, L, G, L, op->V, L, op->U, L);  // only rows iy < G of the result are needed
  fft_rows(st, true, L, G, op->U, L, op->tw);
  hipLaunchKernelGGL(unpack_kernel<S>, dim3((G + 255) / 256, G), dim3(256), 0, st, G, L, op->Ng,
                     op->np, op->dre, op->dim, x, op->b, op->U, B, y);
}

int guarded_grid(const std::function<int()>& f) {
  try {
    return f();
  } catch (const GridError& e) {
    return e.status;
  } catch (const std::bad_alloc&) {
    return fail(HSSC_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

}  // namespace

void grid_op_apply(const hssc_grid_op* op, const void* X, const void* B, void* R, int64_t m,
                   hipStream_t st) {
  const int64_t N = op->N;
  for (int64_t j = 0; j < m; ++j) {
    if (op->cx)
      apply_one<zdouble>(op, static_cast<const zdouble*>(X) + j * N,
                         B ? static_cast<const zdouble*>(B) + j * N : nullptr,
                         static_cast<zdouble*>(R) + j * N, st);
    else
      apply_one<double>(op, static_cast<const double*>(X) + j * N,
                        B ? static_cast<const double*>(B) + j * N : nullptr,
                        static_cast<double*>(R) + j * N, st);
  }
}

int64_t grid_op_n(const hssc_grid_op* op) { return op->N; }
bool grid_op_complex(const hssc_grid_op* op) { return op->cx; }
int grid_op_device(const hssc_grid_op* op) { return op->device; }
uint64_t grid_op_uid(const hssc_grid_op* op) { return op->uid; }

}  // namespace hssc

using namespace hssc;

extern "C" {

int hssc_grid_op_create(const hssc_params* p, hssc_grid_op** out) {
  if (!p || !out) return fail(HSSC_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (int s = check_params(*p)) return s;
  if (p->Ng < 1 || p->np < 1) return fail(HSSC_ERR_INVALID_ARG, "bad grid");
  const int64_t G64 = int64_t(p->Ng) * p->np;
  int L = 2;
  while (L < 2 * G64 - 1) L <<= 1;
  if (L > 1024)  // two LDS rows of L complex doubles per workgroup: 32 KB at L = 1024
    return fail(HSSC_ERR_UNSUPPORTED, "grid operator: grids up to 512 x 512 points");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(HSSC_ERR_HIP, "no HIP device is available");
  return guarded_grid([&]() {
    // typically created while a tree build still runs on its streams: allocations take a
    // cached block or new memory rather than waiting for the device
    pool::NoSync ns;
    auto* op = new hssc_grid_op();
    static std::atomic<uint64_t> next_uid{1};
    op->uid = next_uid.fetch_add(1);
    try {
      op->p = *p;
      GCHK(hipGetDevice(&op->device));
      op->G = int(G64), op->L = L, op->Ng = p->Ng, op->np = p->np;
      op->N = G64 * G64;
      op->cx = p->flag_pot == HSSC_POT_SL_H_2D;
      const k::KernelSpec ks = kernel_spec(*p);
      if (op->cx) op->dre = ks.wre, op->dim = ks.wim;
      // per-point weights (row b, column c) of X_source
      std::vector<double> X(2 * op->N), bw(op->N), cw(op->N);
      source_points(p->Ng, p->np, X.data());
      {
        // the weights are independent per point: contiguous chunks on host threads (this
        // runs while the GPU still works on a tree build queued just before)
        const int64_t N = op->N;
        const int nw = int(std::max<int64_t>(
            1, std::min<int64_t>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())),
                                 N / 4096)));
        auto chunk = [&](int w) {
          for (int64_t g = N * w / nw; g < N * (w + 1) / nw; ++g) {
            bw[g] = point_weight(*p, X[g], X[N + g], true);
            cw[g] = point_weight(*p, X[g], X[N + g], false);
          }
        };
        std::vector<std::thread> th;
        for (int w = 1; w < nw; ++w) th.emplace_back(chunk, w);
        chunk(0);
        for (auto& x : th) x.join();
      }
      op->b = galloc<double>(op->N);
      op->c = galloc<double>(op->N);
      GCHK(hipMemcpy(op->b, bw.data(), op->N * sizeof(double), hipMemcpyHostToDevice));
      GCHK(hipMemcpy(op->c, cw.data(), op->N * sizeof(double), hipMemcpyHostToDevice));
      // twiddles exp(-2 pi i k / L), k < L / 2
      std::vector<zdouble> tw(L / 2);
      for (int kk = 0; kk < L / 2; ++kk) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * kk / L;
        tw[kk] = zdouble(double(std::cos(a)), double(std::sin(a)));
      }
      op->tw = galloc<zdouble>(L / 2);
      GCHK(hipMemcpy(op->tw, tw.data(), tw.size() * sizeof(zdouble), hipMemcpyHostToDevice));
      const int64_t LL = int64_t(L) * L;
      op->U = galloc<zdouble>(LL);
      op->V = galloc<zdouble>(LL);
      op->khat = galloc<zdouble>(LL);
      const hipStream_t st = nullptr;
      hipLaunchKernelGGL(kernel_wrap_kernel, dim3((L + 255) / 256, L), dim3(256), 0, st, op->G, L,
                         p->h, ks, op->U);
      fft_rows(st, false, L, L, op->U, L, op->tw);
      transpose(st, L, L, L, op->U, L, op->khat, L);
      fft_rows(st, false, L, L, op->khat, L, op->tw);
      hipLaunchKernelGGL(scale_kernel, dim3((LL + 255) / 256), dim3(256), 0, st, LL,
                         1.0 / double(LL), op->khat);
      GCHK(hipGetLastError());
      GCHK(hipStreamSynchronize(st));
    } catch (...) {
      delete op;
      throw;
    }
    *out = op;
    return HSSC_OK;
  });
}

int hssc_grid_op_apply_device(const hssc_grid_op* op, const void* dX, void* dY, int64_t m,
                              void* stream) {
  if (!op || !dX || !dY || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid grid-op arguments");
  return guarded_grid([&]() {
    grid_op_apply(op, dX, nullptr, dY, m, static_cast<hipStream_t>(stream));
    GCHK(hipGetLastError());
    return HSSC_OK;
  });
}

int hssc_grid_op_apply(const hssc_grid_op* op, const double* X, double* Y, int64_t m) {
  if (!op || !X || !Y || m < 1) return fail(HSSC_ERR_INVALID_ARG, "invalid grid-op arguments");
  const size_t bytes = size_t(op->N * m) * (op->cx ? sizeof(zdouble) : sizeof(double));
  return guarded_grid([&]() {
    char* d = galloc<char>(2 * bytes);
    GCHK(hipMemcpy(d, X, bytes, hipMemcpyHostToDevice));
    grid_op_apply(op, d, nullptr, d + bytes, m, nullptr);
    GCHK(hipGetLastError());
    GCHK(hipMemcpy(Y, d + bytes, bytes, hipMemcpyDeviceToHost));
    pool::release(d);
    return HSSC_OK;
  });
}

void hssc_grid_op_destroy(hssc_grid_op* op) { delete op; }

}  // extern "C"

// Host construction of the HSS2D tree, skeleton index sets and proxies.
//
// Follows HSS2D/HSS2D_build_tree_fsym.m (uniform tree, 'bdry' skeleton and proxy rules)
// and HSS2D/build_box_proxy.m.  Everything here is integer / comparison work on at most
// a few thousand points per box; the numerical work (interpolation operators) runs on the
// GPU afterwards.  The floating-point operation order matches the oracle so that index
// sets agree exactly.

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <numeric>
#include <random>
#include <thread>

#include "internal.h"

namespace hssc {

void source_points(int Ng, int np, double* X) {
  // Common/HSS_tree_parameters.m:144-160: np x np panels, Ng x Ng cell-centred nodes,
  // panel-major, y fastest inside a panel and across panels.
  const double hp = 2.0 / np;
  const double hh = 2.0 / (double(Ng) * np);
  const int64_t npp = int64_t(Ng) * Ng;
  const int64_t N = int64_t(np) * np * npp;
  for (int64_t g = 0; g < N; ++g) {
    const int64_t p = g / npp, loc = g % npp;
    const int64_t px = p / np, py = p % np;
    const int64_t jx = loc / Ng, iy = loc % Ng;
    X[g] = (-1.0 + double(px) * hp) + (double(jx) + 0.5) * hh;
    X[N + g] = (-1.0 + double(py) * hp) + (double(iy) + 0.5) * hh;
  }
}

namespace {
// one factor of the Helmholtz weight: the tanh plateau of Bump_function option 4
double bump_factor(double h, double x) {
  const double c = 4 * h, D = 10.0;
  return 0.5 * std::tanh(D * (-std::fabs(x) + 1 - c)) + 0.5;
}
}  // namespace

double point_weight(const hssc_params& p, double x, double y, bool row) {
  // Common/Kernel_Eval.m:111-159 (LOCAL_get_bc): b on the row points, c on the column points.
  //   sym = 0, kh > 0 (:115-128): b = Bump_function(., 4) with h = 0 (1 when proxy > 0), c = 1;
  //   TI (:155-158): 1;
  //   NTI (:129-154): b = 1 when proxy > 0, c = 1 when proxy < 0; otherwise
  //     Helmholtz (:133,:143): sqrt(Bump_function(., 4, h)), or sqrt of the Gaussian
  //     contrast option;
  //     Laplace (:135,:146-150): b = b_func(., 0.3, 0.6, 0.25); c the same when sym,
  //     b_func(., 0.2, -0.6, 0.25) otherwise.
  if (p.sym == 0 && p.kh > 0)
    return (row && p.proxy <= 0) ? bump_factor(0.0, x) * bump_factor(0.0, y) : 1.0;
  if (p.transinv) return 1.0;
  if ((row && p.proxy > 0) || (!row && p.proxy < 0)) return 1.0;
  if (p.kh > 0) {
    if (p.contrast == HSSC_CONTRAST_GAUSSIAN)
      return std::sqrt(p.contrast_amp *
                       std::exp(-(x * x + y * y) / (2.0 * p.contrast_width * p.contrast_width)));
    return std::sqrt(bump_factor(p.h, x) * bump_factor(p.h, y));
  }
  const bool second = !row && !p.sym;
  const double d1 = x - (second ? 0.2 : 0.3), d2 = y - (second ? -0.6 : 0.6);
  return 1.0 + 0.25 * std::exp(-d1 * d1 - d2 * d2);
}

namespace {

// point_weight of one tree build.  The Helmholtz weight sqrt(f(x) f(y)) is a product of two
// one-variable factors, and with dyadic coordinates every grid and proxy point lies on the
// half-h lattice: the factors are tabulated there once (the same expression, so the values
// are bit-identical) and looked up; other points and kernels evaluate point_weight.
class WeightTable {
 public:
  explicit WeightTable(const hssc_params& p) : p_(p) {
    if (p.transinv || !(p.kh > 0) || !(p.h > 0) || p.contrast != HSSC_CONTRAST_PLATEAU ||
        !p.sym)
      return;
    step_ = 0.5 * p.h;
    K_ = int64_t(std::ceil(4.0 / step_));
    if (K_ > (int64_t(1) << 22)) return;
    f_.resize(2 * K_ + 1);
    for (int64_t k = -K_; k <= K_; ++k) f_[k + K_] = bump_factor(p.h, double(k) * step_);
  }
  double operator()(double x, double y) const {
    if (f_.empty()) return point_weight(p_, x, y);
    const double kx = x / step_, ky = y / step_;
    if (kx != std::floor(kx) || ky != std::floor(ky) || std::fabs(kx) > double(K_) ||
        std::fabs(ky) > double(K_) || double(int64_t(kx)) * step_ != x ||
        double(int64_t(ky)) * step_ != y)
      return point_weight(p_, x, y);
    return std::sqrt(f_[int64_t(kx) + K_] * f_[int64_t(ky) + K_]);
  }

 private:
  const hssc_params& p_;
  double step_ = 0;
  int64_t K_ = 0;
  std::vector<double> f_;
};

void level_radii(int lev, double* colrad, double* rowrad) {
  // HSS2D/HSS2D_build_tree_fsym.m:322-326 (colrad = x half-width, rowrad = y).
  const int scl = lev / 2;
  *rowrad = std::ldexp(1.0, -scl);
  *colrad = *rowrad / double(1 << (lev % 2));
}

// Stable argsort by key (MATLAB sort is stable).
std::vector<int32_t> argsort_stable(const std::vector<double>& key) {
  std::vector<int32_t> idx(key.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  return idx;
}

// J = [skeleton in polar order ; residual in interface order]: LOCAL_orient_skeleton
// (:866-874, stable sort by radius, then by angle from th0) and LOCAL_orient_residual
// (:877-883, stable sort by the normal coordinate, then by minus the tangential one).
// LOCAL_Non_Preset_Skeleton_Info (:705-724) closes with the same two sorts.
void orient(const std::vector<int32_t>& sk, const std::vector<int32_t>& rs,
            const std::vector<double>& xs, const std::vector<double>& ys, double cx, double cy,
            int lev, std::vector<int32_t>* J) {
  const bool even = (lev % 2) == 0;
  const double th0 = even ? M_PI / 2 : 0.0;
  std::vector<double> th(sk.size()), rho(sk.size());
  for (size_t i = 0; i < sk.size(); ++i) {
    const double dx = xs[sk[i]] - cx, dy = ys[sk[i]] - cy;
    double t = std::atan2(dy, dx);
    if (t < th0) t = t + 2 * M_PI;
    th[i] = t;
    rho[i] = std::hypot(dx, dy);
  }
  std::vector<int32_t> J1 = argsort_stable(rho);
  std::vector<double> thJ1(sk.size());
  for (size_t i = 0; i < sk.size(); ++i) thJ1[i] = th[J1[i]];
  std::vector<int32_t> J2 = argsort_stable(thJ1);
  std::vector<double> nrm(rs.size()), tang(rs.size());
  for (size_t i = 0; i < rs.size(); ++i) {
    nrm[i] = even ? xs[rs[i]] : ys[rs[i]];
    tang[i] = even ? ys[rs[i]] : xs[rs[i]];
  }
  std::vector<int32_t> R1 = argsort_stable(nrm);
  std::vector<double> tR1(rs.size());
  for (size_t i = 0; i < rs.size(); ++i) tR1[i] = -tang[R1[i]];
  std::vector<int32_t> R2 = argsort_stable(tR1);

  J->clear();
  for (size_t i = 0; i < sk.size(); ++i) J->push_back(sk[J1[J2[i]]]);
  for (size_t i = 0; i < rs.size(); ++i) J->push_back(rs[R1[R2[i]]]);
}

// LOCAL_Preset_Skeleton_Info, 'bdry' branch (:804-864): the points within layers*h of the
// box boundary.  pad_to > count adds the residual points nearest the boundary (distance
// min(colrad - |x - cx|, rowrad - |y - cy|), ties by index) until there are pad_to: the
// levels above ID-selected skeletons have ragged counts, padded to the level's maximum.
void preset_skeleton(const std::vector<double>& xs, const std::vector<double>& ys, double cx,
                     double cy, int lev, const hssc_params& p, int64_t pad_to,
                     std::vector<int32_t>* J, int64_t* k) {
  double colrad, rowrad;
  level_radii(lev, &colrad, &rowrad);
  const double width = p.layers * p.h;
  const int64_t n = int64_t(xs.size());
  std::vector<int32_t> sk, rs;
  for (int64_t i = 0; i < n; ++i) {
    const bool bdry = (std::fabs(xs[i] - cx) > (colrad - width)) ||
                      (std::fabs(ys[i] - cy) > (rowrad - width));
    (bdry ? sk : rs).push_back(int32_t(i));
  }
  if (pad_to > int64_t(sk.size()) && !rs.empty()) {
    std::vector<double> d(rs.size());
    for (size_t i = 0; i < rs.size(); ++i)
      d[i] = std::min(colrad - std::fabs(xs[rs[i]] - cx), rowrad - std::fabs(ys[rs[i]] - cy));
    const std::vector<int32_t> o = argsort_stable(d);
    const size_t add = std::min<size_t>(size_t(pad_to) - sk.size(), rs.size());
    std::vector<char> moved(rs.size(), 0);
    for (size_t i = 0; i < add; ++i) sk.push_back(rs[o[i]]), moved[o[i]] = 1;
    std::vector<int32_t> keep;
    for (size_t i = 0; i < rs.size(); ++i)
      if (!moved[i]) keep.push_back(rs[i]);
    rs.swap(keep);
  }
  orient(sk, rs, xs, ys, cx, cy, lev, J);
  *k = int64_t(sk.size());
}

// Number of 'bdry' skeleton points of a box (the count preset_skeleton finds before padding).
int64_t bdry_count(const std::vector<double>& xs, const std::vector<double>& ys, double cx,
                   double cy, int lev, const hssc_params& p) {
  double colrad, rowrad;
  level_radii(lev, &colrad, &rowrad);
  const double width = p.layers * p.h;
  int64_t c = 0;
  for (size_t i = 0; i < xs.size(); ++i)
    c += (std::fabs(xs[i] - cx) > (colrad - width)) || (std::fabs(ys[i] - cy) > (rowrad - width));
  return c;
}

// MATLAB a:d:b for spans that are integer multiples of d up to rounding.
void colon(double a, double d, double b, std::vector<double>* out) {
  const int64_t n = int64_t(std::floor((b - a) / d + 1e-10)) + 1;
  out->clear();
  for (int64_t i = 0; i < n; ++i) out->push_back(a + d * double(i));
}

// HSS2D/build_box_proxy.m:52-85: four strips of exterior points at spacing h, `layers`
// deep, split by the corner trimming of :71-84 into X_proxy and X_extra.  The 'bdry' rule
// of LOCAL_Build_Proxy (HSS2D_build_tree_fsym.m:962-965) uses [X_proxy; X_extra]: the kept
// points of strips 1-4 first, then the trimmed points of strips 1-4.
void box_proxy(double cx, double cy, double colrad, double rowrad, int layers, double h,
               std::vector<double>* px, std::vector<double>* py) {
  const double width = (layers - 0.5) * h;
  const double r1 = layers * layers - 0.5, r2 = layers * layers + 0.5;
  const double ew = colrad + width - h / 2, eh = rowrad + width - h / 2;
  std::vector<double> xr, yr, ex, ey;
  px->clear();
  py->clear();
  // strip 1 (left): x fastest
  colon(cx - colrad - width, h, cx - colrad - h / 2, &xr);
  colon(cy - rowrad - width, h, cy + rowrad + width, &yr);
  for (double y : yr)
    for (double x : xr) {
      const double dx = x - cx, dy = y - cy;
      const bool rm = ((std::fabs(dx) > ew) &&
                       ((dy > rowrad + width - r1 * h) || (dy < -(rowrad + width - r2 * h)))) ||
                      (std::fabs(dy) > eh);
      if (!rm) px->push_back(x), py->push_back(y);
      else ex.push_back(x), ey.push_back(y);
    }
  // strip 2 (top): y fastest
  colon(cx - colrad + h / 2, h, cx + colrad - h / 2, &xr);
  colon(cy + rowrad + width, -h, cy + rowrad + h / 2, &yr);
  for (double x : xr)
    for (double y : yr) {
      const double dx = x - cx, dy = y - cy;
      const bool rm = (std::fabs(dy) > eh) &&
                      ((dx > colrad + width - r1 * h) || (dx < -(colrad + width - r2 * h)));
      if (!rm) px->push_back(x), py->push_back(y);
      else ex.push_back(x), ey.push_back(y);
    }
  // strip 3 (right): x fastest
  colon(cx + colrad + h / 2, h, cx + colrad + width, &xr);
  colon(cy + rowrad + width, -h, cy - rowrad - width, &yr);
  for (double y : yr)
    for (double x : xr) {
      const double dx = x - cx, dy = y - cy;
      const bool rm = ((std::fabs(dx) > ew) &&
                       ((dy > rowrad + width - r2 * h) || (dy < -(rowrad + width - r1 * h)))) ||
                      (std::fabs(dy) > eh);
      if (!rm) px->push_back(x), py->push_back(y);
      else ex.push_back(x), ey.push_back(y);
    }
  // strip 4 (bottom): y fastest
  colon(cx + colrad - h / 2, -h, cx - colrad + h / 2, &xr);
  colon(cy - rowrad - h / 2, -h, cy - rowrad - width, &yr);
  for (double x : xr)
    for (double y : yr) {
      const double dx = x - cx, dy = y - cy;
      const bool rm = (std::fabs(dy) > eh) &&
                      ((dx > colrad + width - r2 * h) || (dx < -(colrad + width - r1 * h)));
      if (!rm) px->push_back(x), py->push_back(y);
      else ex.push_back(x), ey.push_back(y);
    }
  px->insert(px->end(), ex.begin(), ex.end());
  py->insert(py->end(), ey.begin(), ey.end());
}

// MATLAB's rand after rng('default'): the Mersenne Twister MT19937 seeded with 5489, each
// double made from two 32-bit outputs a, b as ((a >> 5) 2^26 + (b >> 6)) / 2^53.
class DefaultRand {
 public:
  double operator()() {
    const uint32_t a = uint32_t(g_()) >> 5, b = uint32_t(g_()) >> 6;
    return (double(a) * 67108864.0 + double(b)) / 9007199254740992.0;
  }

 private:
  std::mt19937 g_{5489u};
};

// LOCAL_random_sampler (HSS2D/HSS2D_build_tree_fsym.m:979-1006): keeps the points within
// lay*h of the box and the others with probability 2^(-lambda (sz - Ld)), Ld the dyadic shell
// of their distance D (Chebyshev distance to the box) relative to the largest one, drawn
// with a fresh rng('default') stream.
void random_sampler(std::vector<double>* px, std::vector<double>* py, double cx, double cy,
                    double colrad, double rowrad, double lambda, int lay, double h) {
  const size_t n = px->size();
  if (n == 0) return;
  std::vector<double> D(n);
  double rad = 0;
  for (size_t i = 0; i < n; ++i) {
    D[i] = std::fabs(std::min(colrad - std::fabs((*px)[i] - cx), rowrad - std::fabs((*py)[i] - cy)));
    rad = std::max(rad, D[i]);
  }
  const double width = lay * h;
  const int sz = -int(std::floor(std::log2(width / rad))) + 1;
  std::vector<char> keep(n);
  for (size_t i = 0; i < n; ++i) keep[i] = D[i] < width;
  if (sz > 0) {
    DefaultRand rnd;
    for (size_t i = 0; i < n; ++i) {
      const int Ld = keep[i] ? sz
                             : std::min(-int(std::ceil(std::log2(D[i] / rad))), sz - 2) + 1;
      const double P = std::exp2(-lambda * double(sz - Ld));  // prob(Ld), :995
      keep[i] = rnd() < P;
    }
  }
  size_t o = 0;
  for (size_t i = 0; i < n; ++i)
    if (keep[i]) (*px)[o] = (*px)[i], (*py)[o] = (*py)[i], ++o;
  px->resize(o), py->resize(o);
}

// LOCAL_Build_Proxy (:886-977) for the 'bdry' and 'rand' rules.  'rand' thins the ring
// [X_proxy; X_extra] until it holds k0 = ceil(2/3 n) (even levels) or ceil(3/4 n) (odd)
// points, at most ten times, each time with lambda x 0.9 (:899-921; every pass thins the
// previous pass's points, as the reference does).
void build_proxy(double cx, double cy, double colrad, double rowrad, int lev, int64_t n_skel,
                 int lay_prox, const hssc_params& p, std::vector<double>* px,
                 std::vector<double>* py) {
  box_proxy(cx, cy, colrad, rowrad, lay_prox, p.h, px, py);
  if (p.proxy_rule != HSSC_PROXY_RAND) return;
  int64_t k0 = int64_t(std::ceil((lev % 2 == 0 ? 2.0 / 3.0 : 3.0 / 4.0) * double(n_skel)));
  k0 = std::min<int64_t>(k0, int64_t(px->size()));
  double lambda = p.lambda;
  int64_t k = 0;
  for (int c = 0; k0 > k && c < 10; ++c) {
    random_sampler(px, py, cx, cy, colrad, rowrad, lambda, p.layers, p.h);
    k = int64_t(px->size());
    if (k0 > k) lambda *= 0.9;
  }
}

int proxy_layers(int lev, const hssc_params& p) {
  // lay_prox of HSS2D/HSS2D_build_tree_fsym.m:339-347 (params.m = 1).
  double colrad, rowrad;
  level_radii(lev, &colrad, &rowrad);
  if (lev > 0) {
    if (p.proxy_layers > 0) return p.proxy_layers;  // caller override of the rule below
    if (p.flag_pot == HSSC_POT_SL_L_2D) return p.layers + 1;
    return 2 * int(std::lround(std::min(rowrad, colrad) / p.h));
  }
  return p.layers;
}

// Worker threads kept for one tree build: the per-level loops below are short (a few ms), so
// starting threads for each of them would cost about as much as the work.
class WorkPool {
 public:
  explicit WorkPool(int n) : n_(std::max(1, n)) {
    for (int w = 1; w < n_; ++w) th_.emplace_back([this, w]() { loop(w); });
  }
  ~WorkPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  int size() const { return n_; }
  // f(w) for w in [0, nthr), nthr <= size(), on the workers and the caller; returns when all
  // are done
  void run(int nthr, const std::function<void(int)>& f) {
    nthr = std::max(1, std::min(nthr, n_));
    {
      std::lock_guard<std::mutex> g(mu_);
      f_ = &f;
      nthr_ = nthr;
      pending_ = nthr - 1;
      ++gen_;
    }
    cv_.notify_all();
    f(0);
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return pending_ == 0; });
    f_ = nullptr;
  }

 private:
  void loop(int w) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* f;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        if (w >= nthr_) continue;
        f = f_;
      }
      (*f)(w);
      {
        std::lock_guard<std::mutex> g(mu_);
        --pending_;
      }
      done_.notify_one();
    }
  }
  int n_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* f_ = nullptr;
  int nthr_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace

int build_index_sets(const hssc_params& p, hssc_tree* t, const IdSelect& sel,
                     const LevelDone& done) {
  if (p.Ng < 1 || p.np < 1 || (p.np & (p.np - 1)) != 0)
    return fail(HSSC_ERR_INVALID_ARG, "np must be a power of two and Ng >= 1");
  if (p.np < 2) return fail(HSSC_ERR_INVALID_ARG, "np must be at least 2");
  if (p.sym != 1)
    return fail(HSSC_ERR_UNSUPPORTED, "only symmetric kernels (HSS2D_*_fsym) are supported");
  if (p.proxy != 0)
    return fail(HSSC_ERR_UNSUPPORTED, "HSS2D trees need params.proxy = 0 (b = c)");
  t->p = p;
  t->cx = p.flag_pot == HSSC_POT_SL_H_2D;
  const int depth = p.depth;
  t->depth = depth;
  const int64_t N = int64_t(p.np) * p.np * p.Ng * p.Ng;
  t->N = N;
  t->X.resize(2 * N);
  source_points(p.Ng, p.np, t->X.data());
  const double* Xx = t->X.data();
  const double* Xy = t->X.data() + N;
  // Power-of-two Ng and np make h, the grid, the box centres and the proxy rings dyadic
  // rationals with short mantissas: every coordinate and every difference is exact.
  const bool dyadic = (p.Ng & (p.Ng - 1)) == 0 && (p.np & (p.np - 1)) == 0;
  // per-point kernel weights b(x) = c(x), evaluated once per grid point
  WorkPool pool(int(std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u)));
  // contiguous chunks of [0, count) over the pool
  auto parallel_range = [&](int64_t count, const std::function<void(int64_t, int64_t)>& f) {
    const int nw = int(std::max<int64_t>(1, std::min<int64_t>(pool.size(), count / 256)));
    pool.run(nw, [&](int w) { f(count * w / nw, count * (w + 1) / nw); });
  };
  const WeightTable weight(p);
  std::vector<double> wgt(N);
  parallel_range(N, [&](int64_t g0, int64_t g1) {
    for (int64_t g = g0; g < g1; ++g) wgt[g] = weight(Xx[g], Xy[g]);
  });

  // Heap-numbered centres (bisect_level, :1289-1344).
  const int64_t nboxes = (int64_t(1) << (depth + 1)) - 1;
  t->cent.assign(2 * (nboxes + 1), 0.0);
  for (int lev = 0; lev < depth; ++lev) {
    const double s = 0.5 * std::ldexp(1.0, -(lev / 2));
    for (int64_t b = int64_t(1) << lev; b < (int64_t(1) << (lev + 1)); ++b) {
      const double cx = t->cent[2 * b], cy = t->cent[2 * b + 1];
      if (lev % 2 == 0) {
        t->cent[2 * (2 * b)] = cx + (-s), t->cent[2 * (2 * b) + 1] = cy + 0.0;
        t->cent[2 * (2 * b + 1)] = cx + s, t->cent[2 * (2 * b + 1) + 1] = cy + 0.0;
      } else {
        t->cent[2 * (2 * b)] = cx + 0.0, t->cent[2 * (2 * b) + 1] = cy + s;
        t->cent[2 * (2 * b + 1)] = cx + 0.0, t->cent[2 * (2 * b + 1) + 1] = cy + (-s);
      }
    }
  }
  t->Isrc.assign(nboxes + 1, {});
  t->Isk.assign(nboxes + 1, {});
  // Leaf slabs (Build_I_source_unif, :1203-1230).
  const double hp = 2.0 / p.np;
  const int64_t nmx = int64_t(p.Ng) * p.Ng;
  for (int64_t b = int64_t(1) << depth; b <= nboxes; ++b) {
    const int64_t px = int64_t(std::floor((t->cent[2 * b] + 1.0) / hp));
    const int64_t py = int64_t(std::floor((t->cent[2 * b + 1] + 1.0) / hp));
    const int64_t pn = py + px * p.np;
    auto& I = t->Isrc[b];
    I.resize(nmx);
    for (int64_t i = 0; i < nmx; ++i) I[i] = pn * nmx + i;
  }

  t->levels.assign(depth + 1, Level{});
  const bool TI = p.transinv != 0;
  bool ragged_below = false;  // skeleton sets of the level below differ between boxes
  // Upward skeleton pass (:321-622).
  for (int lev = depth; lev >= 0; --lev) {
    Level& L = t->levels[lev];
    L.lev = lev;
    L.nbox = int64_t(1) << lev;
    L.nkey = TI ? 1 : L.nbox;
    double colrad, rowrad;
    level_radii(lev, &colrad, &rowrad);
    const int lay_prox = proxy_layers(lev, p);
    const int64_t b0 = int64_t(1) << lev;
    // I_src of every box of the level (children's skeletons, :358-363)
    if (lev < depth)
      parallel_range(L.nbox, [&](int64_t i0, int64_t i1) {
        for (int64_t bi = i0; bi < i1; ++bi) {
          const int64_t b = b0 + bi;
          auto& I = t->Isrc[b];
          I = t->Isk[2 * b];
          I.insert(I.end(), t->Isk[2 * b + 1].begin(), t->Isk[2 * b + 1].end());
        }
      });
    L.n = int64_t(t->Isrc[b0].size());
    for (int64_t bi = 0; bi < L.nbox; ++bi)
      if (int64_t(t->Isrc[b0 + bi].size()) != L.n)
        return fail(HSSC_ERR_UNSUPPORTED, "non-uniform box sizes on a level");
    // For kh > 0 the two bottom levels take their skeletons from an ID of K(X_proxy, X_src)
    // (:374-395 selects, LOCAL_Non_Preset_Skeleton_Info :686-729 computes); the proxy is
    // the level's 'bdry' ring (:408-420).  Every level of a batch has one rank: boxes keep
    // the first k_max columns of their ID order, k_max the level's largest adaptive rank,
    // and preset levels above pad their 'bdry' sets to the level's largest count.
    const bool id_level = p.kh > 0 && !p.preset_all && lev >= depth - 1;
    L.id_skel = id_level;
    L.kbox.assign(L.nkey, -1);
    std::vector<int32_t> id_order;
    int64_t pad_to = 0;
    const int64_t nthr_all = std::max<int64_t>(1, std::min<int64_t>(L.nkey, pool.size()));
    auto parallel_keys = [&](const std::function<bool(int64_t)>& f) {
      std::vector<char> ok(nthr_all, 1);
      pool.run(int(nthr_all), [&](int w) {
        for (int64_t bi = w; bi < L.nkey; bi += nthr_all)
          if (!f(bi)) ok[w] = 0;
      });
      for (char c : ok)
        if (!c) return false;
      return true;
    };
    if (id_level) {
      if (!sel) return fail(HSSC_ERR_UNSUPPORTED, "ID skeletons (kh > 0) need the GPU");
      const int64_t n = L.n;
      std::vector<double> qx0, qy0;
      build_proxy(t->cent[2 * b0], t->cent[2 * b0 + 1], colrad, rowrad, lev, n, lay_prox, p, &qx0,
                  &qy0);
      const int64_t M = int64_t(qx0.size());
      hvec<double> Sx(L.nkey * n), Sy(L.nkey * n), Sw(L.nkey * n);
      hvec<double> Qx(L.nkey * M), Qy(L.nkey * M), Qw(L.nkey * M);
      const bool same = parallel_keys([&](int64_t bi) {
        const int64_t b = b0 + bi;
        const auto& I = t->Isrc[b];
        for (int64_t i = 0; i < n; ++i)
          Sx[bi * n + i] = Xx[I[i]], Sy[bi * n + i] = Xy[I[i]], Sw[bi * n + i] = wgt[I[i]];
        std::vector<double> qx, qy;
        build_proxy(t->cent[2 * b], t->cent[2 * b + 1], colrad, rowrad, lev, n, lay_prox, p, &qx,
                    &qy);
        if (int64_t(qx.size()) != M) return false;
        for (int64_t r = 0; r < M; ++r)
          Qx[bi * M + r] = qx[r], Qy[bi * M + r] = qy[r], Qw[bi * M + r] = weight(qx[r], qy[r]);
        return true;
      });
      if (!same) return fail(HSSC_ERR_UNSUPPORTED, "non-uniform proxy sizes");
      IdJob job{lev, L.nkey, n, M, Sx.data(), Sy.data(), Sw.data(), Qx.data(), Qy.data(), Qw.data()};
      id_order.resize(L.nkey * n);
      std::vector<int32_t> rank(L.nkey);
      if (int s = sel(p, job, id_order.data(), rank.data())) return s;
      for (int64_t bi = 0; bi < L.nkey; ++bi) {
        L.kbox[bi] = rank[bi];
        pad_to = std::max<int64_t>(pad_to, rank[bi]);
      }
    } else if (ragged_below && !TI) {
      parallel_keys([&](int64_t bi) {
        const int64_t b = b0 + bi;
        const auto& I = t->Isrc[b];
        std::vector<double> xs(L.n), ys(L.n);
        for (int64_t i = 0; i < L.n; ++i) xs[i] = Xx[I[i]], ys[i] = Xy[I[i]];
        L.kbox[bi] = int32_t(bdry_count(xs, ys, t->cent[2 * b], t->cent[2 * b + 1], lev, p));
        return true;
      });
      for (int64_t bi = 0; bi < L.nkey; ++bi) pad_to = std::max<int64_t>(pad_to, L.kbox[bi]);
    }
    // params.skel_extra_pts (:380-391 selects the preset levels below the root; :545-564 NTI,
    // :582-597 TI): the preset skeleton of every key is extended by the skeleton of its
    // interpolation residual K(Xp, X_rs) - K(Xp, X_sk) T (LOCAL_ID_rand_update, :1069-1124).
    // The GPU job factors K(Xp, X_src) in the preset J order with the skeleton columns pinned
    // first, so its later pivots are the residual's; every key then keeps the level's largest
    // count, as at the ID levels.
    const bool extra_level = p.skel_extra_pts && !id_level && lev > 0;
    if (extra_level) {
      if (!sel) return fail(HSSC_ERR_UNSUPPORTED, "skel_extra_pts needs the GPU");
      const int64_t n = L.n;
      std::vector<std::vector<int32_t>> J0(L.nkey);
      std::vector<int64_t> k0(L.nkey);
      std::vector<double> qx0, qy0;
      build_proxy(t->cent[2 * b0], t->cent[2 * b0 + 1], colrad, rowrad, lev, n, lay_prox, p, &qx0,
                  &qy0);
      const int64_t M = int64_t(qx0.size());
      hvec<double> Sx(L.nkey * n), Sy(L.nkey * n), Sw(L.nkey * n);
      hvec<double> Qx(L.nkey * M), Qy(L.nkey * M), Qw(L.nkey * M);
      const bool same = parallel_keys([&](int64_t bi) {
        const int64_t b = b0 + bi;
        const auto& I = t->Isrc[b];
        const double cx = t->cent[2 * b], cy = t->cent[2 * b + 1];
        std::vector<double> xs(n), ys(n);
        for (int64_t i = 0; i < n; ++i) xs[i] = Xx[I[i]], ys[i] = Xy[I[i]];
        preset_skeleton(xs, ys, cx, cy, lev, p, pad_to, &J0[bi], &k0[bi]);
        for (int64_t a = 0; a < n; ++a) {
          const int32_t j = J0[bi][a];
          Sx[bi * n + a] = xs[j], Sy[bi * n + a] = ys[j], Sw[bi * n + a] = wgt[I[j]];
        }
        std::vector<double> qx, qy;
        build_proxy(cx, cy, colrad, rowrad, lev, n, lay_prox, p, &qx, &qy);
        if (int64_t(qx.size()) != M) return false;
        for (int64_t r = 0; r < M; ++r)
          Qx[bi * M + r] = qx[r], Qy[bi * M + r] = qy[r], Qw[bi * M + r] = weight(qx[r], qy[r]);
        return true;
      });
      if (!same || std::any_of(k0.begin(), k0.end(), [&](int64_t v) { return v != k0[0]; }))
        return fail(HSSC_ERR_UNSUPPORTED, "non-uniform skeleton or proxy sizes");
      if (k0[0] < n) {
        IdJob job{lev, L.nkey, n, M, Sx.data(), Sy.data(), Sw.data(), Qx.data(), Qy.data(),
                  Qw.data()};
        job.pin = k0[0];
        std::vector<int32_t> ord(L.nkey * n), rank(L.nkey);
        if (int s = sel(p, job, ord.data(), rank.data())) return s;
        id_order.resize(L.nkey * n);
        pad_to = 0;
        for (int64_t bi = 0; bi < L.nkey; ++bi) {
          for (int64_t a = 0; a < n; ++a) id_order[bi * n + a] = J0[bi][ord[bi * n + a]];
          L.kbox[bi] = rank[bi];
          pad_to = std::max<int64_t>(pad_to, rank[bi]);
        }
      } else {
        id_order.resize(L.nkey * n);
        for (int64_t bi = 0; bi < L.nkey; ++bi)
          std::copy(J0[bi].begin(), J0[bi].end(), id_order.begin() + bi * n);
        pad_to = k0[0];
      }
    }
    const bool ordered = id_level || extra_level;  // skeleton = the head of id_order
    // skeleton, J order and proxy of one box (key slot = bi for NTI, 0 for TI)
    auto process = [&](int64_t bi) -> bool {
      const int64_t b = b0 + bi;
      const auto& I = t->Isrc[b];
      const int64_t n = L.n;
      std::vector<double> xs(n), ys(n);
      for (int64_t i = 0; i < n; ++i) xs[i] = Xx[I[i]], ys[i] = Xy[I[i]];
      const double cx = t->cent[2 * b], cy = t->cent[2 * b + 1];
      std::vector<int32_t> J;
      int64_t k;
      if (ordered) {
        const int32_t* o = id_order.data() + (TI ? 0 : bi) * n;
        const std::vector<int32_t> sk(o, o + pad_to), rs(o + pad_to, o + n);
        orient(sk, rs, xs, ys, cx, cy, lev, &J);
        k = pad_to;
      } else {
        preset_skeleton(xs, ys, cx, cy, lev, p, pad_to, &J, &k);
      }
      std::vector<double> qx, qy;
      build_proxy(cx, cy, colrad, rowrad, lev, n, lay_prox, p, &qx, &qy);
      if (bi == 0) {
        L.k = k;
        L.M = int64_t(qx.size());
        L.J.resize(L.nkey * n);
        L.Jinv.resize(L.nkey * n);
        L.Px.resize(L.nkey * n), L.Py.resize(L.nkey * n), L.Pw.resize(L.nkey * n);
        L.PNx.resize(L.nkey * n), L.PNy.resize(L.nkey * n), L.PNw.resize(L.nkey * n);
        L.Xpx.resize(L.nkey * L.M), L.Xpy.resize(L.nkey * L.M), L.Xpw.resize(L.nkey * L.M);
      }
      if (k != L.k || int64_t(qx.size()) != L.M) return false;
      const int64_t key = TI ? 0 : bi;
      for (int64_t a = 0; a < n; ++a) {
        L.J[key * n + a] = J[a];
        L.Jinv[key * n + J[a]] = int32_t(a);
        L.Px[key * n + a] = xs[J[a]];
        L.Py[key * n + a] = ys[J[a]];
        L.Pw[key * n + a] = wgt[I[J[a]]];
        L.PNx[key * n + a] = xs[a];
        L.PNy[key * n + a] = ys[a];
        L.PNw[key * n + a] = wgt[I[a]];
      }
      for (int64_t r = 0; r < L.M; ++r) {
        L.Xpx[key * L.M + r] = qx[r];
        L.Xpy[key * L.M + r] = qy[r];
        L.Xpw[key * L.M + r] = weight(qx[r], qy[r]);
      }
      return true;
    };
    if (!process(0)) return fail(HSSC_ERR_UNSUPPORTED, "non-uniform skeleton or proxy sizes");
    if (!TI && L.nbox > 1 && dyadic && !ordered && !ragged_below) {
      // All coordinates are exact dyadic rationals, so every box of the level is an exact
      // translate of box 0: the same J (same comparisons on bit-identical relative
      // coordinates) and the same proxy ring shifted by the centre difference.
      const int64_t n = L.n, M = L.M;
      const int64_t c0 = b0;
      auto copy_box = [&](int64_t bi) {
        const int64_t b = b0 + bi, key = bi;
        const auto& I = t->Isrc[b];
        const double dx = t->cent[2 * b] - t->cent[2 * c0], dy = t->cent[2 * b + 1] - t->cent[2 * c0 + 1];
        for (int64_t a = 0; a < n; ++a) {
          const int32_t j = L.J[a];
          L.J[key * n + a] = j;
          L.Jinv[key * n + j] = int32_t(a);
          L.Px[key * n + a] = Xx[I[j]];
          L.Py[key * n + a] = Xy[I[j]];
          L.Pw[key * n + a] = wgt[I[j]];
          L.PNx[key * n + a] = Xx[I[a]];
          L.PNy[key * n + a] = Xy[I[a]];
          L.PNw[key * n + a] = wgt[I[a]];
        }
        for (int64_t r = 0; r < M; ++r) {
          const double qx = L.Xpx[r] + dx, qy = L.Xpy[r] + dy;
          L.Xpx[key * M + r] = qx;
          L.Xpy[key * M + r] = qy;
          L.Xpw[key * M + r] = weight(qx, qy);
        }
      };
      const int64_t nthr = std::max<int64_t>(1, std::min<int64_t>(L.nbox - 1, pool.size()));
      pool.run(int(nthr), [&](int w) {
        for (int64_t bi = 1 + w; bi < L.nbox; bi += nthr) copy_box(bi);
      });
    } else if (!TI && L.nbox > 1) {
      // boxes are independent and write disjoint key slots: split them over threads
      const int64_t nthr = std::max<int64_t>(1, std::min<int64_t>(L.nbox - 1, pool.size()));
      std::vector<char> ok(nthr, 1);
      pool.run(int(nthr), [&](int w) {
        for (int64_t bi = 1 + w; bi < L.nbox; bi += nthr)
          if (!process(bi)) ok[w] = 0;
      });
      for (char c : ok)
        if (!c) return fail(HSSC_ERR_UNSUPPORTED, "non-uniform skeleton or proxy sizes");
    }
    parallel_range(L.nbox, [&](int64_t i0, int64_t i1) {
      for (int64_t bi = i0; bi < i1; ++bi) {
        const int64_t b = b0 + bi;
        const int64_t key = TI ? 0 : bi;
        auto& S = t->Isk[b];
        S.resize(L.k);
        for (int64_t a = 0; a < L.k; ++a) S[a] = t->Isrc[b][L.J[key * L.n + a]];
      }
    });
    for (auto& kb : L.kbox)
      if (kb < 0) kb = int32_t(L.k);
    if (ordered && !TI) ragged_below = true;
    L.q = L.n - L.k;
    if (done)
      if (int s = done(lev)) return s;
  }
  // Leaf gather map: global index of each J-ordered leaf point.
  const Level& LL = t->levels[depth];
  t->leafG.resize(LL.nbox * LL.n);
  for (int64_t bi = 0; bi < LL.nbox; ++bi) {
    const int64_t key = TI ? 0 : bi;
    const auto& I = t->Isrc[(int64_t(1) << depth) + bi];
    for (int64_t a = 0; a < LL.n; ++a) t->leafG[bi * LL.n + a] = int32_t(I[LL.J[key * LL.n + a]]);
  }
  return HSSC_OK;
}

}  // namespace hssc

// CPU oracle: an independent C++17 restatement of the reference's CPU training path.
//
// TEST INFRASTRUCTURE ONLY - never linked or loaded by the product. It is the checker
// that GPU results are compared against, and the CPU baseline bench.py times.
//
// Every function cites the reference file:line whose behaviour it restates. The
// restatement is pinned three ways (see tests/test_oracle_*.py):
//   * the reference's own hand-computed unit tests, ported as fixtures;
//   * the reference sources themselves, compiled unchanged into oracle/_ref/libref.so,
//     compared bit-for-bit on seeded inputs;
//   * closed forms / brute force (numpy bincount in row order, exhaustive split search).
//
// Leaf-wise growth (src/core/tree_leafwise.cpp) is outside the hot path and is rejected.

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "oracle_capi.h"

namespace orc {

struct DataError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct TrainError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

constexpr double kInf = std::numeric_limits<double>::infinity();

// Feature-parallel mode: the per-feature loops (binning, histograms, missing statistics, the
// split search) may run on several threads. Each feature's work is the same sequence of
// operations as on one thread, and results are combined in feature order, so every result is
// bit-identical to one thread (the reference itself has no threads). orc_set_threads.
class FeaturePool {
 public:
  ~FeaturePool() { resize(0); }
  unsigned size() const { return unsigned(workers_.size()) + 1; }
  void resize(unsigned threads) {  // threads including the caller
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
      ++epoch_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
    workers_.clear();
    stop_ = false;
    // A new worker starts from the current epoch: it waits for the next run, not for one past.
    for (unsigned k = 1; k < threads; ++k) workers_.emplace_back([this, e = epoch_] { work(e); });
  }
  // Calls fn(i) for i in [0, n) across the pool; the first exception (lowest i) is rethrown.
  template <class Fn>
  void run(uint32_t n, Fn&& fn) {
    std::vector<std::exception_ptr> errors(n);
    std::function<void(uint32_t)> task = [&](uint32_t i) {
      try {
        fn(i);
      } catch (...) {
        errors[i] = std::current_exception();
      }
    };
    {
      std::lock_guard<std::mutex> g(m_);
      task_ = &task;
      n_ = n;
      next_ = 0;
      active_ = unsigned(workers_.size());
      ++epoch_;
    }
    cv_.notify_all();
    drain();
    {
      std::unique_lock<std::mutex> g(m_);
      done_.wait(g, [this] { return active_ == 0; });
      task_ = nullptr;
    }
    for (auto& e : errors)
      if (e) std::rethrow_exception(e);
  }

 private:
  void drain() {
    for (uint32_t i; (i = next_.fetch_add(1)) < n_;) (*task_)(i);
  }
  void work(uint64_t seen) {
    for (;;) {
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || epoch_ != seen; });
        if (stop_) return;
        seen = epoch_;
      }
      drain();
      std::lock_guard<std::mutex> g(m_);
      if (--active_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  bool stop_ = false;
  uint64_t epoch_ = 0;
  unsigned active_ = 0;
  std::atomic<uint32_t> next_{0};
  uint32_t n_ = 0;
  std::function<void(uint32_t)>* task_ = nullptr;
};

static FeaturePool g_pool;
static std::mutex g_pool_user;  // one parallel loop at a time (the checker is called from one thread)

template <class Fn>
static void for_features(uint32_t features, Fn&& fn) {
  if (g_pool.size() <= 1 || features < 2) {
    for (uint32_t f = 0; f < features; ++f) fn(f);
    return;
  }
  std::lock_guard<std::mutex> g(g_pool_user);
  g_pool.run(features, fn);
}

// ---------------------------------------------------------------- binned dataset ----
// include/mpsboost/binned_dataset.hpp:40-46,81-126; storage is kept as uint16 here for
// both widths, the reported width follows max_bins (src/core/binned_dataset.cpp:132).
struct Dataset {
  uint64_t rows = 0;
  uint32_t features = 0;
  uint32_t max_bins = 0;
  std::vector<float> boundaries;
  std::vector<uint64_t> boundary_offset;
  std::vector<uint32_t> boundary_count;
  std::vector<uint32_t> bin_count;
  std::vector<uint64_t> missing_count;
  std::vector<uint16_t> bins;    // feature-major [f][row]
  std::vector<uint8_t> missing;  // feature-major [f][row]
  std::vector<uint8_t> any_missing;  // per feature: some row of this data is missing

  uint32_t bin(uint64_t row, uint32_t f) const {  // src/core/binned_dataset.cpp:255-266
    if (row >= rows || f >= features) throw DataError("Binned index is out of bounds");
    return bins[static_cast<size_t>(f) * rows + row];
  }
  bool is_missing(uint64_t row, uint32_t f) const {  // src/core/binned_dataset.cpp:268-275
    if (row >= rows || f >= features) throw DataError("missing mask index out of range");
    return missing[static_cast<size_t>(f) * rows + row] != 0;
  }
};

struct View {
  const uint8_t* data;
  uint64_t rows;
  uint32_t features;
  uint64_t row_stride;
  uint64_t col_stride;
  int scalar;  // 1 f32, 2 f64
};

static uint64_t mul_checked(uint64_t a, uint64_t b, const char* what) {
  if (a != 0 && b > std::numeric_limits<uint64_t>::max() / a)
    throw DataError(std::string(what) + ": multiplication overflow");
  return a * b;
}
static uint64_t add_checked(uint64_t a, uint64_t b, const char* what) {
  if (b > std::numeric_limits<uint64_t>::max() - a)
    throw DataError(std::string(what) + ": addition overflow");
  return a + b;
}

// src/core/binned_dataset.cpp:19-48
static void validate_view(const View& v, uint32_t max_bins) {
  if (v.data == nullptr) throw DataError("Input matrix data must not be null");
  if (v.rows == 0 || v.features == 0)
    throw DataError("Input matrix must contain at least one row and one feature");
  if (max_bins < 2 || max_bins > 65536) throw DataError("max_bins must be in [2, 65536]");
  uint64_t item = v.scalar == 1 ? 4 : (v.scalar == 2 ? 8 : 0);
  if (item == 0) throw DataError("Input scalar type is not supported");
  if (v.row_stride < item || v.col_stride < item)
    throw DataError("Input matrix stride is smaller than the element size");
  uint64_t last_row = mul_checked(v.rows - 1, v.row_stride, "input matrix row stride");
  uint64_t last_col =
      mul_checked(uint64_t(v.features - 1), v.col_stride, "input matrix column stride");
  add_checked(add_checked(last_row, last_col, "input matrix maximum offset"), item,
              "input matrix final offset");
  mul_checked(v.rows, v.features, "input matrix element count");
}

// src/core/binned_dataset.cpp:50-81: NaN passes through, Inf and out-of-float-range fail.
static float read_value(const View& v, uint64_t row, uint32_t f) {
  const uint8_t* p = v.data + row * v.row_stride + uint64_t(f) * v.col_stride;
  if (v.scalar == 1) {
    float x;
    std::memcpy(&x, p, 4);
    if (std::isnan(x)) return x;
    if (!std::isfinite(x)) throw DataError("Input matrix contains Inf");
    return x;
  }
  double x;
  std::memcpy(&x, p, 8);
  if (std::isnan(x)) return std::numeric_limits<float>::quiet_NaN();
  if (!std::isfinite(x)) throw DataError("Input matrix contains Inf");
  const double fmax = double(std::numeric_limits<float>::max());
  if (x > fmax || x < -fmax) throw DataError("float64 input value exceeds the finite float32 range");
  return static_cast<float>(x);
}

// Ceil-rounded integer quantile rank: part*total/divisor rounded up, minus one, computed
// without overflow (src/core/binned_dataset.cpp:83-93).
static uint64_t quantile_rank(uint64_t part, uint64_t total, uint64_t divisor) {
  const uint64_t q = total / divisor, r = total % divisor;
  const uint64_t frac_num = part * r;
  return part * q + frac_num / divisor + (frac_num % divisor ? 1 : 0) - 1;
}

// src/core/binned_dataset.cpp:95-118: sort the non-missing values, take the value at each
// quantile rank, keep it only if below the maximum and strictly above the previous one.
static std::vector<float> make_boundaries(std::vector<float> values, uint32_t max_bins) {
  values.erase(std::remove_if(values.begin(), values.end(), [](float x) { return std::isnan(x); }),
               values.end());
  std::vector<float> out;
  if (values.empty()) return out;
  std::sort(values.begin(), values.end());
  const uint64_t desired = std::min<uint64_t>(max_bins, values.size());
  const float top = values.back();
  for (uint64_t part = 1; part < desired; ++part) {
    const float c = values[quantile_rank(part, values.size(), desired)];
    if (c < top && (out.empty() || c > out.back())) out.push_back(c);
  }
  return out;
}

// src/core/binned_dataset_internal.cpp:38-76
static void validate_schema(const Dataset& d) {
  if (d.features == 0 || d.max_bins < 2 || d.max_bins > 65536)
    throw DataError("Binned schema feature count or max_bins is invalid");
  uint64_t expect = 0;
  for (uint32_t f = 0; f < d.features; ++f) {
    if (d.boundary_offset[f] != expect || d.bin_count[f] == 0 ||
        d.bin_count[f] != d.boundary_count[f] + 1 || d.bin_count[f] > d.max_bins)
      throw DataError("Binned schema boundary range or bin count is invalid");
    const uint64_t end = d.boundary_offset[f] + d.boundary_count[f];
    if (end > d.boundaries.size()) throw DataError("Binned schema boundary range is out of bounds");
    for (uint32_t i = 0; i < d.boundary_count[f]; ++i) {
      const float x = d.boundaries[d.boundary_offset[f] + i];
      if (!std::isfinite(x) || (i && x <= d.boundaries[d.boundary_offset[f] + i - 1]))
        throw DataError("Binned schema feature boundaries must be finite and strictly increasing");
    }
    expect = end;
  }
  if (expect != d.boundaries.size())
    throw DataError("Binned schema contains boundaries not referenced by a feature");
}

// QuantizeDense, src/core/binned_dataset.cpp:126-188. Missing values go to bin 0 and set
// the mask; other values take the lower_bound position among the boundaries.
static Dataset quantize(const View& v, uint32_t max_bins) {
  validate_view(v, max_bins);
  Dataset d;
  d.rows = v.rows;
  d.features = v.features;
  d.max_bins = max_bins;
  d.bins.assign(size_t(v.rows) * v.features, 0);
  d.missing.assign(size_t(v.rows) * v.features, 0);
  // Feature by feature (in parallel in feature-parallel mode); the first bad value in
  // feature-major read order is the one reported.
  std::vector<std::vector<float>> bounds(v.features);
  std::vector<uint64_t> nmiss(v.features, 0);
  for_features(v.features, [&](uint32_t f) {
    std::vector<float> col(v.rows);
    for (uint64_t r = 0; r < v.rows; ++r) {
      col[r] = read_value(v, r, f);
      nmiss[f] += std::isnan(col[r]) ? 1 : 0;
    }
    const std::vector<float> b = make_boundaries(col, max_bins);
    for (uint64_t r = 0; r < v.rows; ++r) {
      const bool miss = std::isnan(col[r]);
      const size_t at = size_t(f) * v.rows + r;
      d.missing[at] = miss ? 1 : 0;
      d.bins[at] = miss ? 0 : uint16_t(std::lower_bound(b.begin(), b.end(), col[r]) - b.begin());
    }
    bounds[f] = b;
  });
  for (uint32_t f = 0; f < v.features; ++f) {
    const std::vector<float>& b = bounds[f];
    d.boundary_offset.push_back(d.boundaries.size());
    d.boundary_count.push_back(uint32_t(b.size()));
    d.bin_count.push_back(uint32_t(b.size() + 1));
    d.missing_count.push_back(nmiss[f]);
    d.any_missing.push_back(nmiss[f] ? 1 : 0);
    d.boundaries.insert(d.boundaries.end(), b.begin(), b.end());
  }
  validate_schema(d);
  return d;
}

// TransformDense, src/core/binned_dataset.cpp:190-238: frozen boundaries, no refit.
static Dataset transform(const View& v, const Dataset& schema) {
  validate_view(v, schema.max_bins);
  validate_schema(schema);
  if (v.features != schema.features)
    throw DataError("Prediction input feature count does not match the binned schema");
  Dataset d = schema;
  d.rows = v.rows;
  d.bins.assign(size_t(v.rows) * v.features, 0);
  d.missing.assign(size_t(v.rows) * v.features, 0);
  d.any_missing.assign(v.features, 0);
  for (uint32_t f = 0; f < v.features; ++f) {
    const float* lo = schema.boundaries.data() + schema.boundary_offset[f];
    const float* hi = lo + schema.boundary_count[f];
    for (uint64_t r = 0; r < v.rows; ++r) {
      const float x = read_value(v, r, f);
      const bool miss = std::isnan(x);
      const size_t at = size_t(f) * v.rows + r;
      d.missing[at] = miss ? 1 : 0;
      d.any_missing[f] |= miss ? 1 : 0;
      d.bins[at] = miss ? 0 : uint16_t(std::lower_bound(lo, hi, x) - lo);
    }
  }
  return d;
}

// ------------------------------------------------------------------- objectives ----
// src/core/objective.cpp. Messages follow the reference so error tests read alike.
struct GH {
  double g = 0.0, h = 0.0;
};

static void need_finite(double x, const char* what) {
  if (!std::isfinite(x)) throw TrainError(std::string(what) + " must be finite");
}
static void need_same_nonempty(size_t a, size_t b) {  // objective.cpp:38-46
  if (a == 0) throw TrainError("Labels must not be empty");
  if (a != b) throw TrainError("Label and prediction lengths do not match");
}
static double denominator(double h, double lambda) {  // objective.cpp:22-36
  need_finite(h, "Hessian sum");
  need_finite(lambda, "reg_lambda");
  if (h < 0.0) throw TrainError("Hessian sum cannot be negative");
  if (lambda < 0.0) throw TrainError("reg_lambda cannot be negative");
  const double d = h + lambda;
  if (!std::isfinite(d) || d <= 0.0)
    throw TrainError("Hessian sum plus reg_lambda must be finite and positive");
  return d;
}
static double soft_threshold(double g, double alpha) {  // objective.cpp:75-86
  need_finite(g, "Gradient sum");
  need_finite(alpha, "reg_alpha");
  if (alpha < 0.0) throw TrainError("reg_alpha cannot be negative");
  const double m = std::fabs(g) - alpha;
  return m <= 0.0 ? 0.0 : std::copysign(m, g);
}
static double clip_leaf(double w, double max_delta) {  // objective.cpp:88-97
  need_finite(max_delta, "max_delta_step");
  if (max_delta < 0.0) throw TrainError("max_delta_step cannot be negative");
  if (max_delta == 0.0) return w;
  return std::max(-max_delta, std::min(max_delta, w));
}
static double node_score(double g, double h, double lambda, double alpha) {  // :289-301
  const double s = soft_threshold(g, alpha);
  const double d = denominator(h, lambda);
  const double r = s * s / d;
  if (!std::isfinite(r)) throw TrainError("Node score overflowed");
  return r;
}
static double leaf_weight(double g, double h, double lambda, double alpha, double max_delta) {
  // objective.cpp:303-318
  const double s = soft_threshold(g, alpha);
  const double w = clip_leaf(-s / denominator(h, lambda), max_delta);
  if (!std::isfinite(w)) throw TrainError("Leaf value overflowed");
  return w;
}
static double split_gain(double gl, double hl, double gr, double hr, double lambda, double alpha,
                         double gamma) {  // objective.cpp:320-356
  need_finite(gamma, "gamma");
  if (gamma < 0.0) throw TrainError("gamma cannot be negative");
  need_finite(hl, "left child Hessian sum");
  need_finite(hr, "right child Hessian sum");
  if (hl <= 0.0 || hr <= 0.0) throw TrainError("Split child Hessian sums must be strictly positive");
  const double gp = gl + gr, hp = hl + hr;
  if (!std::isfinite(gp) || !std::isfinite(hp)) throw TrainError("Parent node statistics overflowed");
  const double sl = node_score(gl, hl, lambda, alpha);
  const double sr = node_score(gr, hr, lambda, alpha);
  const double sp = node_score(gp, hp, lambda, alpha);
  const double gain = 0.5 * (sl + sr - sp) - gamma;
  if (!std::isfinite(gain)) throw TrainError("Split gain overflowed");
  return gain;
}

static std::vector<GH> squared_error(const double* y, const double* p, size_t n) {  // :101-118
  need_same_nonempty(n, n);
  std::vector<GH> out(n);
  for (size_t i = 0; i < n; ++i) {
    need_finite(y[i], "label");
    need_finite(p[i], "prediction");
    const double g = p[i] - y[i];
    if (!std::isfinite(g)) throw TrainError("Squared-error gradient overflowed");
    out[i] = {g, 1.0};
  }
  return out;
}
static double logistic(double z) {  // :120-128
  need_finite(z, "logit");
  if (z >= 0.0) return 1.0 / (1.0 + std::exp(-z));
  const double e = std::exp(z);
  return e / (1.0 + e);
}
static std::vector<GH> binary_logistic(const double* y, const double* z, size_t n) {  // :154-172
  need_same_nonempty(n, n);
  std::vector<GH> out(n);
  for (size_t i = 0; i < n; ++i) {
    need_finite(y[i], "label");
    if (!(y[i] == 0.0 || y[i] == 1.0)) throw TrainError("Binary logistic labels must be 0 or 1");
    const double p = logistic(z[i]);
    const double g = p - y[i], h = p * (1.0 - p);
    if (!std::isfinite(g) || !std::isfinite(h) || h < 0.0)
      throw TrainError("Binary logistic gradient/Hessian overflowed");
    out[i] = {g, h};
  }
  return out;
}
static std::vector<GH> quantile_grad(const double* y, const double* p, size_t n, double a) {
  // :174-192
  need_same_nonempty(n, n);
  need_finite(a, "quantile alpha");
  if (a <= 0.0 || a >= 1.0) throw TrainError("quantile alpha must be in (0, 1)");
  std::vector<GH> out(n);
  for (size_t i = 0; i < n; ++i) {
    need_finite(y[i], "label");
    need_finite(p[i], "prediction");
    out[i] = {p[i] < y[i] ? -a : 1.0 - a, 1.0};
  }
  return out;
}
static double safe_exp(double x, const char* what) {  // :63-73
  need_finite(x, what);
  if (x > 709.0) throw TrainError(std::string(what) + " exponential overflowed");
  const double r = std::exp(x);
  if (!std::isfinite(r)) throw TrainError(std::string(what) + " exponential overflowed");
  return r;
}
static void need_nonneg_label(double y, const char* obj) {
  need_finite(y, "label");
  if (y < 0.0) throw TrainError(std::string(obj) + " labels must be non-negative");
}
static std::vector<GH> poisson_grad(const double* y, const double* m, size_t n) {  // :194-211
  need_same_nonempty(n, n);
  std::vector<GH> out(n);
  for (size_t i = 0; i < n; ++i) {
    need_nonneg_label(y[i], "poisson");
    const double mu = safe_exp(m[i], "poisson log mean");
    const double g = mu - y[i], h = mu;
    if (!std::isfinite(g) || !std::isfinite(h) || h <= 0.0)
      throw TrainError("poisson gradient/Hessian overflowed");
    out[i] = {g, h};
  }
  return out;
}
static std::vector<GH> tweedie_grad(const double* y, const double* m, size_t n, double rho) {
  // :213-240
  need_same_nonempty(n, n);
  need_finite(rho, "tweedie variance power");
  if (rho <= 1.0 || rho >= 2.0) throw TrainError("tweedie variance power must be in (1, 2)");
  std::vector<GH> out(n);
  for (size_t i = 0; i < n; ++i) {
    need_nonneg_label(y[i], "tweedie");
    const double a = safe_exp((2.0 - rho) * m[i], "tweedie first term");
    const double b = safe_exp((1.0 - rho) * m[i], "tweedie second term");
    const double g = a - y[i] * b;
    const double h = (2.0 - rho) * a - (1.0 - rho) * y[i] * b;
    if (!std::isfinite(g) || !std::isfinite(h) || h <= 0.0)
      throw TrainError("tweedie gradient/Hessian overflowed");
    out[i] = {g, h};
  }
  return out;
}
static std::vector<double> softmax_row(const double* m, size_t k) {  // :130-152
  if (k < 2) throw TrainError("softmax requires at least two class margins");
  double top = -kInf;
  for (size_t j = 0; j < k; ++j) {
    need_finite(m[j], "softmax margin");
    top = std::max(top, m[j]);
  }
  std::vector<double> p(k);
  double z = 0.0;
  for (size_t j = 0; j < k; ++j) {
    p[j] = std::exp(m[j] - top);
    z += p[j];
  }
  if (!std::isfinite(z) || z <= 0.0) throw TrainError("softmax normalizer must be finite and positive");
  for (double& x : p) x /= z;
  return p;
}
static void check_class_label(double y, uint32_t k, const char* what) {
  need_finite(y, what);
  if (std::floor(y) != y || y < 0.0 || y >= double(k))
    throw TrainError("softmax labels must be encoded class ids");
}
static std::vector<GH> softmax_grad(const double* y, const double* margins, size_t n, uint32_t k,
                                    uint32_t target) {  // :242-287
  if (k < 2) throw TrainError("softmax class_count must be at least two");
  if (target >= k) throw TrainError("softmax target_class out of range");
  if (n == 0) throw TrainError("labels must be non-empty");
  std::vector<GH> out(n);
  for (size_t r = 0; r < n; ++r) {
    check_class_label(y[r], k, "softmax label");
    const std::vector<double> p = softmax_row(margins + r * k, k);
    const double ind = uint32_t(y[r]) == target ? 1.0 : 0.0;
    const double g = p[target] - ind, h = p[target] * (1.0 - p[target]);
    if (!std::isfinite(g) || !std::isfinite(h) || h < 0.0)
      throw TrainError("softmax gradient/Hessian overflowed");
    out[r] = {g, h};
  }
  return out;
}

// -------------------------------------------------------------- tree parameters ----
struct Params {
  uint32_t max_depth = 6, max_leaves = 0, max_active_leaves = 0;
  uint64_t min_samples_leaf = 1;
  double min_child_weight = 0.0, lambda = 1.0, alpha = 0.0, max_delta = 0.0, gamma = 0.0,
         min_gain = 0.0;
  bool random_threshold = false;
  uint32_t growth = 0, seed = 0;
  std::vector<int8_t> monotonic;
  std::vector<std::vector<uint32_t>> groups;
};

static Params to_params(const orc_tree_params& p) {
  Params q;
  q.max_depth = p.max_depth;
  q.max_leaves = p.max_leaves;
  q.max_active_leaves = p.max_active_leaves;
  q.min_samples_leaf = p.min_samples_leaf;
  q.min_child_weight = p.min_child_weight;
  q.lambda = p.reg_lambda;
  q.alpha = p.reg_alpha;
  q.max_delta = p.max_delta_step;
  q.gamma = p.gamma;
  q.min_gain = p.min_gain_to_split;
  if (p.split_strategy > 1) throw ArgError("unknown split strategy");
  q.random_threshold = p.split_strategy == 1;
  q.growth = p.growth_strategy;
  q.seed = p.random_seed;
  if (p.n_monotonic) q.monotonic.assign(p.monotonic, p.monotonic + p.n_monotonic);
  size_t at = 0;
  for (uint32_t i = 0; i < p.n_groups; ++i) {
    q.groups.emplace_back(p.group_features + at, p.group_features + at + p.group_sizes[i]);
    at += p.group_sizes[i];
  }
  return q;
}

static void validate_params(const Params& p) {  // src/core/tree_structure.cpp:16-51
  if (p.min_samples_leaf == 0) throw TrainError("min_samples_leaf must be at least 1");
  if (p.max_leaves == 1) throw TrainError("max_leaves must be zero or at least 2");
  if (p.max_active_leaves == 1) throw TrainError("max_active_leaves must be zero or at least 2");
  if (p.max_active_leaves && p.max_leaves && p.max_active_leaves > p.max_leaves)
    throw TrainError("max_active_leaves must not exceed max_leaves");
  auto nonneg = [](double x, const char* msg) {
    if (!std::isfinite(x) || x < 0.0) throw TrainError(msg);
  };
  nonneg(p.min_child_weight, "min_child_weight must be finite and non-negative");
  nonneg(p.lambda, "reg_lambda must be finite and non-negative");
  nonneg(p.alpha, "reg_alpha must be finite and non-negative");
  nonneg(p.max_delta, "max_delta_step must be finite and non-negative");
  nonneg(p.gamma, "gamma must be finite and non-negative");
  nonneg(p.min_gain, "min_gain_to_split must be finite and non-negative");
}

static uint32_t max_leaves_of(const Params& p) {  // src/core/tree_structure.cpp:54-62
  if (p.max_leaves) return p.max_leaves;
  if (p.max_depth >= 31) return std::numeric_limits<uint32_t>::max();
  return 1u << p.max_depth;
}

// ------------------------------------------------------------------- histograms ----
struct Stats {
  uint64_t count = 0;
  double g = 0.0, h = 0.0;
};
struct Cell {
  uint64_t count = 0;
  double g = 0.0, h = 0.0;
};
using FeatureHist = std::vector<Cell>;
using NodeHist = std::vector<FeatureHist>;

// CpuReferenceBackend::BuildHistograms, src/cpu/reference_backend.cpp:22-71: rows outer,
// features inner, FP64 sums in row order so every bin's order is the row order. The row and
// gradient checks run first, in row order, as the reference meets them on its first feature;
// the accumulation then runs feature by feature (each feature's rows in row order).
static NodeHist build_hist(const Dataset& d, const std::vector<uint64_t>& rows,
                           const std::vector<GH>& gh) {
  if (rows.empty()) throw TrainError("Node row set must not be empty");
  if (d.rows != gh.size()) throw TrainError("Gradient count does not match binned dataset row count");
  NodeHist hist(d.features);
  for (uint32_t f = 0; f < d.features; ++f) hist[f].resize(d.bin_count[f]);
  for (const uint64_t r : rows) {
    if (r >= d.rows) throw TrainError("Histogram row index is out of bounds");
    const GH& x = gh[r];
    if (!std::isfinite(x.g) || !std::isfinite(x.h) || x.h < 0.0)
      throw TrainError("Gradient/Hessian must be finite and Hessian non-negative");
  }
  for_features(d.features, [&](uint32_t f) {
    FeatureHist& fh = hist[f];
    for (const uint64_t r : rows) {
      const uint32_t b = d.bin(r, f);
      if (b >= fh.size()) throw TrainError("Binned value exceeds the feature histogram range");
      Cell& c = fh[b];
      ++c.count;
      c.g += gh[r].g;
      c.h += gh[r].h;
      if (!std::isfinite(c.g) || !std::isfinite(c.h))
        throw TrainError("Histogram FP64 accumulation overflowed");
    }
  });
  return hist;
}

// SubtractHistograms, src/core/tree_histogram.cpp:16-66. The Hessian difference may dip
// below zero by rounding; it is clamped to zero within 1e-5*max(1,|p|,|c|), else an error.
static double subtract_hessian(double p, double c) {
  if (!std::isfinite(p) || !std::isfinite(c) || p < 0.0 || c < 0.0)
    throw TrainError("Histogram subtraction Hessian values must be finite and non-negative");
  const double d = p - c;
  const double tol = 1e-5 * std::max({1.0, std::abs(p), std::abs(c)});
  if (d < -tol) throw TrainError("Histogram subtraction child Hessian exceeded parent");
  return d < 0.0 ? 0.0 : d;
}
static NodeHist subtract_hist(const NodeHist& parent, const NodeHist& child) {
  if (parent.size() != child.size()) throw TrainError("Histogram subtraction feature counts do not match");
  NodeHist out(parent.size());
  for (size_t f = 0; f < parent.size(); ++f) {
    if (parent[f].size() != child[f].size())
      throw TrainError("Histogram subtraction bin counts do not match");
    out[f].resize(parent[f].size());
    for (size_t b = 0; b < parent[f].size(); ++b) {
      const Cell& p = parent[f][b];
      const Cell& c = child[f][b];
      if (c.count > p.count) throw TrainError("Histogram subtraction child count exceeds parent count");
      const double g = p.g - c.g;
      if (!std::isfinite(p.g) || !std::isfinite(c.g) || !std::isfinite(g))
        throw TrainError("Histogram subtraction gradient values must be finite");
      out[f][b] = {p.count - c.count, g, subtract_hessian(p.h, c.h)};
    }
  }
  return out;
}

// SumRows, src/core/tree_split.cpp:119-140
static Stats sum_rows(const std::vector<uint64_t>& rows, const std::vector<GH>& gh) {
  Stats s;
  for (const uint64_t r : rows) {
    if (r >= gh.size()) throw TrainError("Node-statistics row index is out of bounds");
    const GH& x = gh[r];
    if (!std::isfinite(x.g) || !std::isfinite(x.h) || x.h < 0.0)
      throw TrainError("Gradient/Hessian must be finite and Hessian non-negative");
    ++s.count;
    s.g += x.g;
    s.h += x.h;
    if (!std::isfinite(s.g) || !std::isfinite(s.h)) throw TrainError("Node FP64 statistics overflowed");
  }
  return s;
}

// ---------------------------------------------------------------- split search ----
struct Split {
  bool valid = false;
  uint32_t feature = 0, threshold = 0;
  double gain = 0.0;
  bool default_left = true;
  Stats left, right;
  double l_lo = -kInf, l_hi = kInf, r_lo = -kInf, r_hi = kInf;
};

// StableRandomThreshold, src/core/tree_split.cpp:14-29 (splitmix-style integer hash).
static uint32_t random_threshold(uint32_t seed, uint32_t feature, uint32_t node, uint32_t depth,
                                 uint32_t n) {
  if (n == 0) throw TrainError("random threshold candidate count must be positive");
  uint64_t v = seed;
  v ^= uint64_t(feature) + 0x9E3779B97F4A7C15ULL;
  v *= 0xBF58476D1CE4E5B9ULL;
  v ^= (uint64_t(node) << 32U) | depth;
  v *= 0x94D049BB133111EBULL;
  v ^= v >> 31U;
  return uint32_t(v % n);
}

// Candidates arrive in (feature, threshold, missing-left-first) order; a later one only
// replaces the incumbent on a strictly larger gain, so exact ties keep the earliest
// (src/core/tree_split.cpp:31-45).
static bool better(const Split& c, const Split& inc) {
  if (!inc.valid) return true;
  if (c.gain != inc.gain) return c.gain > inc.gain;
  if (c.feature != inc.feature) return c.feature < inc.feature;
  return c.threshold < inc.threshold;
}

static Stats add(const Stats& a, const Stats& b) { return {a.count + b.count, a.g + b.g, a.h + b.h}; }
static Stats sub(const Stats& a, const Stats& b) { return {a.count - b.count, a.g - b.g, a.h - b.h}; }

struct Bounds {
  bool valid = true;
  double l_lo, l_hi, r_lo, r_hi;
};
// MonotonicBoundsForSplit, src/core/tree_monotonic.cpp:12-58
static Bounds monotone_bounds(const Stats& l, const Stats& r, double lo, double hi, uint32_t f,
                              const Params& p) {
  Bounds b{true, lo, hi, lo, hi};
  const int c = p.monotonic.empty() ? 0 : p.monotonic[f];
  if (c == 0) return b;
  const double lv = std::clamp(leaf_weight(l.g, l.h, p.lambda, p.alpha, p.max_delta), lo, hi);
  const double rv = std::clamp(leaf_weight(r.g, r.h, p.lambda, p.alpha, p.max_delta), lo, hi);
  if (c > 0) {
    if (lv > rv) return Bounds{false, lo, hi, lo, hi};
    b.l_hi = std::min(b.l_hi, rv);
    b.r_lo = std::max(b.r_lo, lv);
    return b;
  }
  if (lv < rv) return Bounds{false, lo, hi, lo, hi};
  b.l_lo = std::max(b.l_lo, rv);
  b.r_hi = std::min(b.r_hi, lv);
  return b;
}

// InteractionAllowsFeature, src/core/tree_interactions.cpp:32-45
static bool interaction_ok(const std::vector<uint32_t>& path, uint32_t f, const Params& p) {
  if (p.groups.empty()) return true;
  for (const auto& g : p.groups) {
    if (std::find(g.begin(), g.end(), f) == g.end()) continue;
    bool all = true;
    for (uint32_t x : path) all = all && std::find(g.begin(), g.end(), x) != g.end();
    if (all) return true;
  }
  return false;
}

// ConsiderMissingDirection, src/core/tree_split.cpp:65-115
static void consider(const Stats& left, const Stats& right, const Stats& miss, double lo,
                     double hi, bool dl, uint32_t f, uint32_t t, const Params& p, Split* best) {
  const Stats cl = dl ? add(left, miss) : left;
  const Stats cr = dl ? right : add(right, miss);
  if (cl.count < p.min_samples_leaf || cr.count < p.min_samples_leaf || cl.h < p.min_child_weight ||
      cr.h < p.min_child_weight || cl.h <= 0.0 || cr.h <= 0.0)
    return;
  const double gain = split_gain(cl.g, cl.h, cr.g, cr.h, p.lambda, p.alpha, p.gamma);
  if (gain <= 0.0 || gain < p.min_gain) return;
  const Bounds b = monotone_bounds(cl, cr, lo, hi, f, p);
  if (!b.valid) return;
  Split c;
  c.valid = true;
  c.feature = f;
  c.threshold = t;
  c.gain = gain;
  c.default_left = dl;
  c.left = cl;
  c.right = cr;
  c.l_lo = b.l_lo;
  c.l_hi = b.l_hi;
  c.r_lo = b.r_lo;
  c.r_hi = b.r_hi;
  if (better(c, *best)) *best = c;
}

// FindBestSplit, src/core/tree_split.cpp:142-234. Prefix over bins 0..B-2 in FP64, bin 0
// with the missing rows taken out, right = (parent - missing) - left.
static Split find_best_split(const NodeHist& hist, const Dataset& d, const Stats& parent, double lo,
                             double hi, const std::vector<uint32_t>& path, uint32_t node,
                             uint32_t depth, const Params& p, const std::vector<Stats>& missing) {
  if (hist.size() != d.features) throw TrainError("Histogram feature count does not match the dataset");
  if (missing.size() != d.features) throw TrainError("missing statistics feature count mismatch");
  // Each feature's best candidate in the reference's order, then the features in order: the
  // same best as one pass, since `better` orders candidates by (gain, feature, threshold).
  std::vector<Split> per_feature(d.features);
  for_features(d.features, [&](uint32_t f) {
    if (!interaction_ok(path, f, p)) return;
    const FeatureHist& bins = hist[f];
    if (bins.size() != d.bin_count[f]) throw TrainError("Histogram bin count does not match feature metadata");
    uint64_t total = 0;
    for (const Cell& c : bins) {
      if (c.count > std::numeric_limits<uint64_t>::max() - total)
        throw TrainError("Histogram sample count overflowed");
      total += c.count;
      if (!std::isfinite(c.g) || !std::isfinite(c.h) || c.h < 0.0)
        throw TrainError("Histogram G/H must be finite and Hessian non-negative");
    }
    if (total != parent.count) throw TrainError("Histogram sample count does not match node statistics");
    if (bins.size() < 2) return;
    const Stats& m = missing[f];
    if (m.count > parent.count || m.h > parent.h)
      throw TrainError("missing statistics exceed parent statistics");
    const Stats nonmiss = sub(parent, m);
    Stats left;
    Split& best = per_feature[f];
    for (uint32_t t = 0; t + 1 < bins.size(); ++t) {
      const Cell& c = bins[t];
      left.count += t == 0 ? c.count - m.count : c.count;
      left.g += t == 0 ? c.g - m.g : c.g;
      left.h += t == 0 ? c.h - m.h : c.h;
      if (!std::isfinite(left.g) || !std::isfinite(left.h))
        throw TrainError("Histogram prefix accumulation overflowed");
      const Stats right = sub(nonmiss, left);
      if (left.count > nonmiss.count) throw TrainError("non-missing histogram prefix exceeded parent");
      if (left.count == 0 && right.count == 0) continue;
      if (p.random_threshold &&
          t != random_threshold(p.seed, f, node, depth, uint32_t(bins.size() - 1)))
        continue;
      consider(left, right, m, lo, hi, true, f, t, p, &best);
      consider(left, right, m, lo, hi, false, f, t, p, &best);
    }
  });
  Split best;
  for (const Split& c : per_feature)
    if (c.valid && better(c, best)) best = c;
  return best;
}

struct Active {
  uint32_t node = 0, depth = 0;
  std::vector<uint64_t> rows;
  Stats stats;
  NodeHist cached;
  double lo = -kInf, hi = kInf;
  std::vector<uint32_t> path;
};

struct Prepared {
  bool valid = false;
  Split split;
  std::vector<uint64_t> left, right;
};

// PrepareSplitRows, src/core/tree_split.cpp:236-291: per-feature missing statistics in
// row order, the best split, then a stable partition of the node's rows.
static Prepared prepare(const Dataset& d, const Active& a, const NodeHist& hist,
                        const std::vector<GH>& gh, const Params& p) {
  if (hist.size() != d.features) throw TrainError("Histogram feature count does not match the dataset");
  std::vector<Stats> missing(d.features);
  for_features(d.features, [&](uint32_t f) {
    if (f < d.any_missing.size() && !d.any_missing[f]) return;  // no row of the feature is missing
    for (const uint64_t r : a.rows) {
      if (!d.is_missing(r, f)) continue;
      if (r >= gh.size()) throw TrainError("missing statistics row index out of range");
      ++missing[f].count;
      missing[f].g += gh[r].g;
      missing[f].h += gh[r].h;
    }
  });
  const Split s = find_best_split(hist, d, a.stats, a.lo, a.hi, a.path, a.node, a.depth, p, missing);
  Prepared out;
  if (!s.valid) return out;
  out.valid = true;
  out.split = s;
  for (const uint64_t r : a.rows) {
    const bool go_left = d.is_missing(r, s.feature) ? s.default_left : d.bin(r, s.feature) <= s.threshold;
    (go_left ? out.left : out.right).push_back(r);
  }
  if (out.left.size() != s.left.count || out.right.size() != s.right.count)
    throw TrainError("Sample partition counts do not match histogram statistics");
  return out;
}

// ------------------------------------------------------------------------ trees ----
struct Tree {
  uint32_t features = 0;
  std::vector<orc_node> nodes;
};

static orc_node make_leaf(const Stats& s, const Params& p, double lo, double hi) {
  // MakeBoundedLeaf, src/core/tree_structure.cpp:72-87
  if (lo > hi) throw TrainError("monotonic leaf bounds are inconsistent");
  orc_node n{};
  n.left_child = n.right_child = 0xFFFFFFFFu;
  n.leaf_value = std::clamp(leaf_weight(s.g, s.h, p.lambda, p.alpha, p.max_delta), lo, hi);
  n.default_left = 1;
  n.flags = 1;
  return n;
}

static uint32_t append(Tree* t, const orc_node& n) {  // src/core/tree_structure.cpp:95-102
  if (t->nodes.size() >= 0xFFFFFFFFu) throw TrainError("Tree node count exceeds the uint32 index range");
  t->nodes.push_back(n);
  return uint32_t(t->nodes.size() - 1);
}

static std::vector<uint32_t> extend_path(const std::vector<uint32_t>& path, uint32_t f) {
  std::vector<uint32_t> r = path;  // src/core/tree_interactions.cpp:47-55
  if (std::find(r.begin(), r.end(), f) == r.end()) r.push_back(f);
  return r;
}

// Histograms for the nodes of a layer without a cached one (tree_histogram.cpp:68-110),
// and for the pending smaller children (:112-133); per node, in layer order.
static std::vector<NodeHist> layer_hists(const Dataset& d, const std::vector<Active>& layer,
                                         const std::vector<GH>& gh) {
  std::vector<NodeHist> out;
  out.reserve(layer.size());
  for (const Active& a : layer) out.push_back(a.cached.empty() ? build_hist(d, a.rows, gh) : a.cached);
  return out;
}

// TrainLevelWiseRegressionTree, src/core/tree.cpp:49-147
static Tree train_tree(const Dataset& d, const std::vector<GH>& gh, const Params& p) {
  // TrainSingleRegressionTree validation, src/core/tree.cpp:151-177
  validate_params(p);
  if (!p.monotonic.empty() && p.monotonic.size() != d.features)
    throw TrainError("monotonic_constraints length must match feature count");
  for (int8_t c : p.monotonic)
    if (c < -1 || c > 1) throw TrainError("monotonic_constraints values must be -1, 0, or 1");
  for (const auto& g : p.groups) {  // src/core/tree_interactions.cpp:57-70
    if (g.empty()) throw TrainError("interaction constraint groups must be non-empty");
    for (uint32_t f : g)
      if (f >= d.features) throw TrainError("interaction constraint feature index out of range");
  }
  if (d.rows != gh.size()) throw TrainError("Gradient count does not match binned dataset row count");
  if (p.growth != 0) throw ArgError("the oracle restates level-wise growth only");

  Active root;
  root.rows.resize(d.rows);
  std::iota(root.rows.begin(), root.rows.end(), uint64_t{0});
  root.stats = sum_rows(root.rows, gh);
  Tree tree;
  tree.features = d.features;
  append(&tree, make_leaf(root.stats, p, -kInf, kInf));

  std::vector<Active> layer;
  layer.push_back(std::move(root));
  uint32_t leaves = 1;
  const uint32_t cap = max_leaves_of(p);
  while (!layer.empty() && leaves < cap) {
    if (layer.front().depth >= p.max_depth) break;
    std::vector<Active> next;
    struct Pending {
      size_t index;
      NodeHist parent;
    };
    std::vector<Pending> pending;
    const std::vector<NodeHist> hists = layer_hists(d, layer, gh);
    for (size_t i = 0; i < layer.size(); ++i) {
      Active& a = layer[i];
      if (a.depth >= p.max_depth) continue;
      if (leaves >= cap) break;
      Prepared s = prepare(d, a, hists[i], gh, p);
      if (!s.valid) continue;
      // ApplySplit, src/core/tree_structure.cpp:118-148: children appended left, right.
      const uint32_t li = append(&tree, make_leaf(s.split.left, p, s.split.l_lo, s.split.l_hi));
      const uint32_t ri = append(&tree, make_leaf(s.split.right, p, s.split.r_lo, s.split.r_hi));
      orc_node& parent = tree.nodes[a.node];
      parent.feature_index = s.split.feature;
      parent.threshold_bin = s.split.threshold;
      parent.left_child = li;
      parent.right_child = ri;
      parent.leaf_value = 0.0;
      parent.gain = s.split.gain;
      parent.default_left = s.split.default_left ? 1 : 0;
      parent.flags = 0;
      ++leaves;
      const uint32_t cd = a.depth + 1;
      const std::vector<uint32_t> path = extend_path(a.path, s.split.feature);
      Active l, r;
      l.node = li;
      l.depth = cd;
      l.rows = std::move(s.left);
      l.stats = s.split.left;
      l.lo = s.split.l_lo;
      l.hi = s.split.l_hi;
      l.path = path;
      r.node = ri;
      r.depth = cd;
      r.rows = std::move(s.right);
      r.stats = s.split.right;
      r.lo = s.split.r_lo;
      r.hi = s.split.r_hi;
      r.path = path;
      next.push_back(std::move(l));
      next.push_back(std::move(r));
      if (cd < p.max_depth) {
        // The smaller child is built; ties build the left one (src/core/tree.cpp:109-117).
        const bool build_left = s.split.left.count <= s.split.right.count;
        pending.push_back({next.size() - (build_left ? 2u : 1u), hists[i]});
      }
    }
    for (const Pending& q : pending) {
      next[q.index].cached = build_hist(d, next[q.index].rows, gh);
      next[q.index ^ 1u].cached = subtract_hist(q.parent, next[q.index].cached);
    }
    layer = std::move(next);
  }
  return tree;
}

// RegressionTree::Predict, src/core/tree.cpp:179-214
static std::vector<double> predict_tree(const Tree& t, const Dataset& d) {
  if (t.nodes.empty()) throw TrainError("Tree model has no root node");
  if (d.features != t.features) throw TrainError("Prediction feature count does not match the tree model");
  std::vector<double> out(d.rows);
  for (uint64_t r = 0; r < d.rows; ++r) {
    uint32_t i = 0;
    size_t seen = 0;
    while (true) {
      if (i >= t.nodes.size() || ++seen > t.nodes.size())
        throw TrainError("Tree node index is out of bounds or the structure contains a cycle");
      const orc_node& n = t.nodes[i];
      if (n.flags & 1) {
        if (!std::isfinite(n.leaf_value)) throw TrainError("Tree leaf value is not finite");
        out[r] = n.leaf_value;
        break;
      }
      if (n.feature_index >= t.features) throw TrainError("Tree branch feature index is out of bounds");
      const bool go_left = d.is_missing(r, n.feature_index) ? n.default_left != 0
                                                            : d.bin(r, n.feature_index) <= n.threshold_bin;
      i = go_left ? n.left_child : n.right_child;
    }
  }
  return out;
}

// ----------------------------------------------------------------------- models ----
struct Model {
  bool multiclass = false;
  uint32_t classes = 0;
  double lr = 0.1;
  std::vector<double> base;  // one value, or one per class
  Dataset schema;            // boundaries only; bins dropped
  std::vector<Tree> trees;
};

struct TrainCfg {
  uint32_t n_estimators, max_bins, objective;
  double lr, alpha, rho;
  Params tree;
};

static void validate_train(const TrainCfg& c) {  // src/core/trainer.cpp:22-42
  if (c.n_estimators == 0) throw TrainError("n_estimators must be at least 1");
  if (!std::isfinite(c.lr) || c.lr <= 0.0 || c.lr > 1.0) throw TrainError("learning_rate must be in (0, 1]");
  if (c.max_bins < 2 || c.max_bins > 65536) throw TrainError("max_bins must be in [2, 65536]");
  if (!std::isfinite(c.alpha) || c.alpha <= 0.0 || c.alpha >= 1.0)
    throw TrainError("quantile alpha must be in (0, 1)");
  if (!std::isfinite(c.rho) || c.rho <= 1.0 || c.rho >= 2.0)
    throw TrainError("tweedie variance power must be in (1, 2)");
}

static double total_weight(const double* y, const double* w, size_t n) {  // trainer.cpp:44-66
  if (n == 0) throw TrainError("labels must be non-empty");
  double t = 0.0;
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(w[i]) || w[i] < 0.0) throw TrainError("sample weights must be finite non-negative values");
    t += w[i];
    if (!std::isfinite(t)) throw TrainError("sample weight sum overflowed");
  }
  (void)y;
  if (t <= 0.0) throw TrainError("sample weights must contain positive total weight");
  return t;
}

// InitialBaseScore, src/core/trainer.cpp:68-176
static double base_score(const double* y, const double* w, size_t n, const TrainCfg& c) {
  const double tw = total_weight(y, w, n);
  switch (c.objective) {
    case 0: {  // weighted mean, :68-86
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(y[i])) throw TrainError("Labels must be finite");
        s += y[i] * w[i];
        if (!std::isfinite(s)) throw TrainError("Label accumulation overflowed");
      }
      return s / tw;
    }
    case 1: {  // logit of the weighted positive rate, :88-110
      double pos = 0.0;
      for (size_t i = 0; i < n; ++i) {
        if (y[i] == 0.0) continue;
        if (y[i] == 1.0) {
          pos += w[i];
          continue;
        }
        throw TrainError("binary-logistic labels must be 0 or 1");
      }
      double p = pos / tw;
      p = std::min(1.0 - 1e-15, std::max(1e-15, p));
      return std::log(p / (1.0 - p));
    }
    case 2: {  // weighted quantile, :112-138
      std::vector<std::pair<double, double>> v;
      for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(y[i])) throw TrainError("label must be finite");
        v.push_back({y[i], w[i]});
      }
      std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      const double thr = c.alpha * tw;
      double cum = 0.0;
      for (const auto& it : v) {
        cum += it.second;
        if (cum >= thr) return it.first;
      }
      return v.back().first;
    }
    case 3:
    case 4: {  // log of the weighted mean, :140-157
      const char* name = c.objective == 3 ? "poisson" : "tweedie";
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(y[i]) || y[i] < 0.0) throw TrainError(std::string(name) + " labels must be non-negative");
        s += y[i] * w[i];
        if (!std::isfinite(s)) throw TrainError(std::string(name) + " label sum overflowed");
      }
      return std::log(std::max(1e-15, s / tw));
    }
  }
  throw TrainError("unknown training objective");
}

static std::vector<GH> objective_grad(const double* y, const std::vector<double>& pred,
                                      const TrainCfg& c) {  // trainer.cpp:178-198
  const size_t n = pred.size();
  switch (c.objective) {
    case 0: return squared_error(y, pred.data(), n);
    case 1: return binary_logistic(y, pred.data(), n);
    case 2: return quantile_grad(y, pred.data(), n, c.alpha);
    case 3: return poisson_grad(y, pred.data(), n);
    case 4: return tweedie_grad(y, pred.data(), n, c.rho);
  }
  throw TrainError("unknown training objective");
}

static void apply_weights(std::vector<GH>* gh, const double* w) {  // trainer.cpp:200-216
  for (size_t i = 0; i < gh->size(); ++i) {
    (*gh)[i].g *= w[i];
    (*gh)[i].h *= w[i];
    if (!std::isfinite((*gh)[i].g) || !std::isfinite((*gh)[i].h) || (*gh)[i].h < 0.0)
      throw TrainError("weighted gradient/Hessian overflowed");
  }
}

static Dataset schema_of(const Dataset& d) {
  Dataset s = d;
  s.bins.clear();
  s.missing.clear();
  s.bins.shrink_to_fit();
  s.missing.shrink_to_fit();
  return s;
}

// TrainRegressionModel, src/core/trainer.cpp:226-266
static Model train_regressor(const Dataset& d, const double* y, const double* w, const TrainCfg& c) {
  validate_train(c);
  if (d.max_bins != c.max_bins) throw TrainError("Training data, labels, or max_bins contract does not match");
  total_weight(y, w, d.rows);
  Model m;
  m.schema = schema_of(d);
  m.lr = c.lr;
  m.base = {base_score(y, w, d.rows, c)};
  std::vector<double> pred(d.rows, m.base[0]);
  for (uint32_t round = 0; round < c.n_estimators; ++round) {
    std::vector<GH> gh = objective_grad(y, pred, c);
    apply_weights(&gh, w);
    Tree t = train_tree(d, gh, c.tree);
    const std::vector<double> u = predict_tree(t, d);
    for (size_t r = 0; r < pred.size(); ++r) {
      pred[r] += c.lr * u[r];
      if (!std::isfinite(pred[r])) throw TrainError("Boosting prediction update overflowed");
    }
    m.trees.push_back(std::move(t));
  }
  return m;
}

// TrainMulticlassSoftmaxModel, src/core/multiclass_trainer.cpp:104-149: per round, one
// tree per class in class order, each seeing the margins already updated by the earlier
// classes of the same round.
static Model train_multiclass(const Dataset& d, const double* y, const double* w, const TrainCfg& c,
                              uint32_t k) {
  validate_train(c);
  if (d.max_bins != c.max_bins) throw TrainError("training data, labels, or max_bins contract mismatch");
  const double tw = total_weight(y, w, d.rows);
  if (k < 2) throw TrainError("softmax class_count must be at least two");
  for (uint64_t r = 0; r < d.rows; ++r) {
    if (!std::isfinite(y[r])) throw TrainError("softmax labels must be finite");
    check_class_label(y[r], k, "softmax label");
  }
  Model m;
  m.multiclass = true;
  m.classes = k;
  m.lr = c.lr;
  m.schema = schema_of(d);
  std::vector<double> cw(k, 0.0);  // WeightedClassBaseScores, :39-67
  for (uint64_t r = 0; r < d.rows; ++r) {
    cw[uint32_t(y[r])] += w[r];
    if (!std::isfinite(cw[uint32_t(y[r])])) throw TrainError("softmax class prior overflowed");
  }
  for (double x : cw) {
    const double s = std::log(std::max(1e-15, x / tw));
    if (!std::isfinite(s)) throw TrainError("softmax base score is not finite");
    m.base.push_back(s);
  }
  std::vector<double> margins(d.rows * k);
  for (uint64_t r = 0; r < d.rows; ++r)
    for (uint32_t j = 0; j < k; ++j) margins[r * k + j] = m.base[j];
  for (uint32_t round = 0; round < c.n_estimators; ++round) {
    for (uint32_t cls = 0; cls < k; ++cls) {
      std::vector<GH> gh = softmax_grad(y, margins.data(), d.rows, k, cls);
      apply_weights(&gh, w);
      Tree t = train_tree(d, gh, c.tree);
      const std::vector<double> u = predict_tree(t, d);
      for (uint64_t r = 0; r < d.rows; ++r) {
        double& x = margins[r * k + cls];
        x += c.lr * u[r];
        if (!std::isfinite(x)) throw TrainError("softmax margin update overflowed");
      }
      m.trees.push_back(std::move(t));
    }
  }
  return m;
}

static bool same_schema(const Dataset& a, const Dataset& b) {
  return a.features == b.features && a.max_bins == b.max_bins && a.boundaries == b.boundaries;
}

// RegressionModel::Predict (trainer.cpp:268-282) / MulticlassModel::PredictMargins
// (multiclass_trainer.cpp:215-235).
static std::vector<double> predict_model(const Model& m, const Dataset& d) {
  if (!same_schema(m.schema, d))
    throw TrainError("Prediction data does not use the model's frozen binned schema");
  if (!m.multiclass) {
    std::vector<double> out(d.rows, m.base[0]);
    for (const Tree& t : m.trees) {
      const std::vector<double> u = predict_tree(t, d);
      for (size_t r = 0; r < out.size(); ++r) out[r] += m.lr * u[r];
    }
    return out;
  }
  const uint32_t k = m.classes;
  std::vector<double> out(d.rows * k);
  for (uint64_t r = 0; r < d.rows; ++r)
    for (uint32_t j = 0; j < k; ++j) out[r * k + j] = m.base[j];
  for (size_t i = 0; i < m.trees.size(); ++i) {
    const std::vector<double> u = predict_tree(m.trees[i], d);
    const uint32_t cls = uint32_t(i % k);
    for (uint64_t r = 0; r < d.rows; ++r) out[r * k + cls] += m.lr * u[r];
  }
  return out;
}

static TrainCfg to_cfg(const orc_train_params& p) {
  TrainCfg c;
  c.n_estimators = p.n_estimators;
  c.max_bins = p.max_bins;
  c.objective = p.objective;
  c.lr = p.learning_rate;
  c.alpha = p.objective_alpha;
  c.rho = p.tweedie_variance_power;
  c.tree = to_params(p.tree);
  if (c.objective > 4) throw ArgError("unknown training objective");
  return c;
}

static std::vector<GH> to_gh(const double* pairs, uint64_t n) {
  std::vector<GH> gh(n);
  for (uint64_t i = 0; i < n; ++i) gh[i] = {pairs[2 * i], pairs[2 * i + 1]};
  return gh;
}

}  // namespace orc

// ------------------------------------------------------------------------ C ABI ----
namespace {
template <class F>
int guarded(char* err, uint64_t errlen, F&& body) {
  auto put = [&](const char* msg) {
    if (err && errlen) {
      std::strncpy(err, msg, errlen - 1);
      err[errlen - 1] = 0;
    }
  };
  try {
    body();
    return 0;
  } catch (const orc::DataError& e) {
    put(e.what());
    return 1;
  } catch (const orc::TrainError& e) {
    put(e.what());
    return 2;
  } catch (const std::exception& e) {
    put(e.what());
    return 3;
  }
}
orc::View make_view(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                    int scalar) {
  return orc::View{static_cast<const uint8_t*>(data), rows, features, rs, cs, scalar};
}
}  // namespace

extern "C" {

int orc_quantize(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                 int scalar, uint32_t max_bins, void** out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    *out = new orc::Dataset(orc::quantize(make_view(data, rows, features, rs, cs, scalar), max_bins));
  });
}

int orc_transform(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                  int scalar, const void* schema, void** out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    *out = new orc::Dataset(orc::transform(make_view(data, rows, features, rs, cs, scalar),
                                           *static_cast<const orc::Dataset*>(schema)));
  });
}

void orc_dataset_free(void* d) { delete static_cast<orc::Dataset*>(d); }

void orc_set_threads(uint32_t threads) { orc::g_pool.resize(threads ? threads : 1); }

int orc_dataset_info(const void* p, uint64_t* rows, uint32_t* features, uint32_t* max_bins,
                     uint32_t* storage_bits, uint64_t* boundary_total) {
  const auto& d = *static_cast<const orc::Dataset*>(p);
  *rows = d.rows;
  *features = d.features;
  *max_bins = d.max_bins;
  *storage_bits = d.max_bins <= 256 ? 8 : 16;
  *boundary_total = d.boundaries.size();
  return 0;
}

int orc_dataset_meta(const void* p, uint32_t* bin_count, uint64_t* boundary_offset,
                     uint64_t* missing_count, float* boundaries) {
  const auto& d = *static_cast<const orc::Dataset*>(p);
  for (uint32_t f = 0; f < d.features; ++f) {
    bin_count[f] = d.bin_count[f];
    boundary_offset[f] = d.boundary_offset[f];
    missing_count[f] = d.missing_count[f];
  }
  std::copy(d.boundaries.begin(), d.boundaries.end(), boundaries);
  return 0;
}

int orc_dataset_bins(const void* p, uint16_t* bins, uint8_t* missing) {
  const auto& d = *static_cast<const orc::Dataset*>(p);
  std::copy(d.bins.begin(), d.bins.end(), bins);
  std::copy(d.missing.begin(), d.missing.end(), missing);
  return 0;
}

int orc_squared_error(const double* y, const double* p, uint64_t n, double* out, char* err,
                      uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto gh = orc::squared_error(y, p, n);
    for (uint64_t i = 0; i < n; ++i) {
      out[2 * i] = gh[i].g;
      out[2 * i + 1] = gh[i].h;
    }
  });
}

int orc_gradients(uint32_t objective, const double* y, const double* p, uint64_t n, double alpha,
                  uint32_t k, uint32_t target, double* out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    std::vector<orc::GH> gh;
    switch (objective) {
      case 0: gh = orc::squared_error(y, p, n); break;
      case 1: gh = orc::binary_logistic(y, p, n); break;
      case 2: gh = orc::quantile_grad(y, p, n, alpha); break;
      case 3: gh = orc::poisson_grad(y, p, n); break;
      case 4: gh = orc::tweedie_grad(y, p, n, alpha); break;
      case 5: gh = orc::softmax_grad(y, p, n, k, target); break;
      default: throw orc::ArgError("unknown objective");
    }
    for (uint64_t i = 0; i < n; ++i) {
      out[2 * i] = gh[i].g;
      out[2 * i + 1] = gh[i].h;
    }
  });
}

int orc_node_score(double g, double h, double l, double a, double* out, char* err, uint64_t n) {
  return guarded(err, n, [&] { *out = orc::node_score(g, h, l, a); });
}
int orc_leaf_weight(double g, double h, double l, double a, double m, double* out, char* err,
                    uint64_t n) {
  return guarded(err, n, [&] { *out = orc::leaf_weight(g, h, l, a, m); });
}
int orc_split_gain(double gl, double hl, double gr, double hr, double l, double a, double gamma,
                   double* out, char* err, uint64_t n) {
  return guarded(err, n, [&] { *out = orc::split_gain(gl, hl, gr, hr, l, a, gamma); });
}

int orc_histograms(const void* p, const double* pairs, const uint64_t* rows, uint64_t n_rows,
                   uint64_t* count, double* g, double* h, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const orc::Dataset*>(p);
    const auto hist = orc::build_hist(d, std::vector<uint64_t>(rows, rows + n_rows), orc::to_gh(pairs, d.rows));
    size_t at = 0;
    for (const auto& fh : hist)
      for (const auto& c : fh) {
        count[at] = c.count;
        g[at] = c.g;
        h[at] = c.h;
        ++at;
      }
  });
}

int orc_train_tree(const void* p, const double* pairs, const orc_tree_params* tp, orc_node* nodes,
                   uint32_t capacity, uint32_t* n_nodes, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const orc::Dataset*>(p);
    const orc::Tree t = orc::train_tree(d, orc::to_gh(pairs, d.rows), orc::to_params(*tp));
    *n_nodes = uint32_t(t.nodes.size());
    if (t.nodes.size() > capacity) throw orc::ArgError("node capacity too small");
    std::copy(t.nodes.begin(), t.nodes.end(), nodes);
  });
}

int orc_tree_predict(const void* p, const orc_node* nodes, uint32_t n_nodes, uint32_t tree_features,
                     double* out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const orc::Dataset*>(p);
    orc::Tree t;
    t.features = tree_features;
    t.nodes.assign(nodes, nodes + n_nodes);
    const auto u = orc::predict_tree(t, d);
    std::copy(u.begin(), u.end(), out);
  });
}

int orc_train_regressor(const void* p, const double* y, const double* w,
                        const orc_train_params* tp, void** model, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const orc::Dataset*>(p);
    *model = new orc::Model(orc::train_regressor(d, y, w, orc::to_cfg(*tp)));
  });
}

int orc_train_multiclass(const void* p, const double* y, const double* w, const orc_train_params* tp,
                         uint32_t k, void** model, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const orc::Dataset*>(p);
    *model = new orc::Model(orc::train_multiclass(d, y, w, orc::to_cfg(*tp), k));
  });
}

int orc_model_info(const void* p, uint32_t* trees, uint32_t* classes, double* lr, double* base) {
  const auto& m = *static_cast<const orc::Model*>(p);
  *trees = uint32_t(m.trees.size());
  *classes = m.multiclass ? m.classes : 1;
  *lr = m.lr;
  std::copy(m.base.begin(), m.base.end(), base);
  return 0;
}

int orc_model_tree(const void* p, uint32_t i, orc_node* nodes, uint32_t capacity, uint32_t* n) {
  const auto& m = *static_cast<const orc::Model*>(p);
  if (i >= m.trees.size()) return 3;
  *n = uint32_t(m.trees[i].nodes.size());
  if (*n > capacity) return 3;
  std::copy(m.trees[i].nodes.begin(), m.trees[i].nodes.end(), nodes);
  return 0;
}

int orc_model_predict(const void* p, const void* ds, double* out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto u = orc::predict_model(*static_cast<const orc::Model*>(p),
                                      *static_cast<const orc::Dataset*>(ds));
    std::copy(u.begin(), u.end(), out);
  });
}

void orc_model_free(void* p) { delete static_cast<orc::Model*>(p); }

}  // extern "C"

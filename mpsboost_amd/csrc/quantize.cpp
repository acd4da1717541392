// Host binning: QuantizeDense / TransformDense (src/core/binned_dataset.cpp:126-238).
//
// Boundaries are bit-identical to the reference's: the same float32 reads, the same
// ceil-rounded integer quantile ranks over the sorted non-missing values, the same
// "below the maximum and strictly increasing" filter, lower_bound bin assignment and NaN ->
// missing. What differs is the work split and the output layout: features are processed in
// groups of eight per worker thread (one pass over each 32-byte slice of the input rows),
// and codes are written row-major for the device, missing rows taking the code bin_count.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

#include "internal.hpp"

namespace mpsb {
namespace {

constexpr uint32_t kGroup = 8;

uint64_t mul_checked(uint64_t a, uint64_t b, const char* what) {
  if (a != 0 && b > std::numeric_limits<uint64_t>::max() / a)
    fail(MPSB_ERR_DATA, std::string(what) + ": multiplication overflow");
  return a * b;
}

uint64_t add_checked(uint64_t a, uint64_t b, const char* what) {
  if (b > std::numeric_limits<uint64_t>::max() - a)
    fail(MPSB_ERR_DATA, std::string(what) + ": addition overflow");
  return a + b;
}

}  // namespace

// ValidateDenseView, src/core/binned_dataset.cpp:19-48
void validate_dense_view(const View& v, uint32_t max_bins) {
  if (v.data == nullptr) fail(MPSB_ERR_DATA, "Input matrix data must not be null");
  if (v.rows == 0 || v.features == 0)
    fail(MPSB_ERR_DATA, "Input matrix must contain at least one row and one feature");
  if (max_bins < 2 || max_bins > 65536) fail(MPSB_ERR_DATA, "max_bins must be in [2, 65536]");
  const uint64_t item = v.scalar == 1 ? 4 : (v.scalar == 2 ? 8 : 0);
  if (item == 0) fail(MPSB_ERR_DATA, "Input scalar type is not supported");
  if (v.row_stride < item || v.col_stride < item)
    fail(MPSB_ERR_DATA, "Input matrix stride is smaller than the element size");
  const uint64_t last_row = mul_checked(v.rows - 1, v.row_stride, "input matrix row stride");
  const uint64_t last_col = mul_checked(uint64_t(v.features - 1), v.col_stride, "input matrix column stride");
  add_checked(add_checked(last_row, last_col, "input matrix maximum offset"), item, "input matrix final offset");
  mul_checked(v.rows, v.features, "input matrix element count");
  if (v.rows > 0xFFFFFFFFull) fail(MPSB_ERR_DATA, "row count exceeds the device's 32-bit row index");
}

namespace {

// A value as the reference reads it: float32 as is, float64 range-checked then rounded
// (src/core/binned_dataset.cpp:50-81). Inf and out-of-range values are recorded by their
// position in the reference's feature-major read order, so the error raised is the one the
// reference meets first; NaN passes through as missing.
struct BadValues {
  std::atomic<uint64_t> inf{~0ull}, range{~0ull};
  static void lower(std::atomic<uint64_t>& a, uint64_t at) {
    uint64_t cur = a.load();
    while (at < cur && !a.compare_exchange_weak(cur, at)) {
    }
  }
  void raise() const {
    const uint64_t i = inf.load(), r = range.load();
    if (i == ~0ull && r == ~0ull) return;
    if (i < r) fail(MPSB_ERR_DATA, "Input matrix contains Inf");
    fail(MPSB_ERR_DATA, "float64 input value exceeds the finite float32 range");
  }
};

struct Reader {
  const View& v;
  BadValues* bad;
  float operator()(uint64_t r, uint32_t f) const {
    const uint8_t* p = v.data + r * v.row_stride + uint64_t(f) * v.col_stride;
    if (v.scalar == 1) {
      float x;
      std::memcpy(&x, p, 4);
      if (std::isinf(x)) BadValues::lower(bad->inf, uint64_t(f) * v.rows + r);
      return x;
    }
    double x;
    std::memcpy(&x, p, 8);
    if (std::isnan(x)) return std::numeric_limits<float>::quiet_NaN();
    if (std::isinf(x)) {
      BadValues::lower(bad->inf, uint64_t(f) * v.rows + r);
      return 0.0f;
    }
    const double fmax = double(std::numeric_limits<float>::max());
    if (x > fmax || x < -fmax) {
      BadValues::lower(bad->range, uint64_t(f) * v.rows + r);
      return 0.0f;
    }
    return static_cast<float>(x);
  }
};

// Ceil-rounded integer rank of part/divisor of total, minus one, overflow-free
// (src/core/binned_dataset.cpp:83-93).
uint64_t quantile_rank(uint64_t part, uint64_t total, uint64_t divisor) {
  const uint64_t q = total / divisor, r = total % divisor, fn = part * r;
  return part * q + fn / divisor + (fn % divisor ? 1 : 0) - 1;
}

// src/core/binned_dataset.cpp:95-118. `values` holds the non-missing values (sorted here).
std::vector<float> boundaries_of(std::vector<float>& values, uint32_t max_bins) {
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

uint32_t worker_count(uint32_t requested, uint64_t work_items) {
  uint32_t n = requested ? requested : std::max(1u, std::thread::hardware_concurrency());
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(n, work_items)));
}

template <class F>
void parallel_for(uint32_t threads, uint64_t items, F&& body) {
  std::atomic<uint64_t> next{0};
  auto run = [&] {
    for (uint64_t i = next++; i < items; i = next++) body(i);
  };
  std::vector<std::thread> pool;
  for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(run);
  run();
  for (auto& t : pool) t.join();
}

// Row-major codes from the schema: lower_bound among the feature's boundaries, missing ->
// bin_count. Rows are split into blocks across workers.
void write_codes(const View& v, HostDataset& d, uint32_t threads) {
  d.codes.assign(size_t(d.rows) * d.row_stride * d.code_bytes, 0);
  BadValues bad;
  const Reader read{v, &bad};
  constexpr uint64_t kBlock = 1 << 14;
  const uint64_t blocks = (d.rows + kBlock - 1) / kBlock;
  parallel_for(worker_count(threads, blocks), blocks, [&](uint64_t b) {
    const uint64_t r0 = b * kBlock, r1 = std::min(d.rows, r0 + kBlock);
    for (uint64_t r = r0; r < r1; ++r) {
      for (uint32_t f = 0; f < d.features; ++f) {
        const float x = read(r, f);
        uint32_t code;
        if (std::isnan(x)) {
          code = d.bin_count[f];  // the feature's missing cell
        } else {
          const float* lo = d.boundaries.data() + d.boundary_offset[f];
          const float* hi = lo + (d.bin_count[f] - 1);
          code = uint32_t(std::lower_bound(lo, hi, x) - lo);
        }
        const size_t at = (size_t(r) * d.row_stride + f) * d.code_bytes;
        d.codes[at] = uint8_t(code);
        if (d.code_bytes == 2) d.codes[at + 1] = uint8_t(code >> 8);
      }
    }
  });
  bad.raise();
}

std::atomic<uint64_t> g_uid{1};

}  // namespace

// Device layout of a schema: code width, padded stride, histogram cells.
void HostDataset::finalize_layout() {
  if (miss_flag.size() != features) {
    miss_flag.assign(features, 0);
    for (uint32_t f = 0; f < features; ++f) miss_flag[f] = missing_count[f] != 0 ? 1 : 0;
  }
  uint32_t widest = 0;
  for (uint32_t f = 0; f < features; ++f) {
    const uint32_t n = bin_count[f] + (has_missing(f) ? 1u : 0u);
    widest = std::max(widest, n);
    if (has_missing(f) && bin_count[f] >= 65535)
      fail(MPSB_ERR_DATA, "a feature with missing values needs fewer than 65535 bins on the device");
  }
  code_bytes = widest <= 256 ? 1 : 2;
  const uint32_t per16 = 16 / code_bytes;
  row_stride = (features + per16 - 1) / per16 * per16;
  cell_offset.assign(features + 1, 0);
  for (uint32_t f = 0; f < features; ++f)
    cell_offset[f + 1] = cell_offset[f] + bin_count[f] + (has_missing(f) ? 1u : 0u);
  cells = cell_offset[features];
  uid = g_uid++;
}

// ValidateSchemaFields, src/core/binned_dataset_internal.cpp:38-76
void validate_schema(const HostDataset& d) {
  if (d.features == 0 || d.max_bins < 2 || d.max_bins > 65536)
    fail(MPSB_ERR_DATA, "Binned schema feature count or max_bins is invalid");
  uint64_t expect = 0;
  for (uint32_t f = 0; f < d.features; ++f) {
    if (d.boundary_offset[f] != expect || d.bin_count[f] == 0 || d.bin_count[f] > d.max_bins)
      fail(MPSB_ERR_DATA, "Binned schema boundary range or bin count is invalid");
    const uint64_t end = d.boundary_offset[f] + (d.bin_count[f] - 1);
    if (end > d.boundaries.size()) fail(MPSB_ERR_DATA, "Binned schema boundary range is out of bounds");
    for (uint64_t i = d.boundary_offset[f]; i < end; ++i) {
      if (!std::isfinite(d.boundaries[i]) || (i > d.boundary_offset[f] && d.boundaries[i] <= d.boundaries[i - 1]))
        fail(MPSB_ERR_DATA, "Binned schema feature boundaries must be finite and strictly increasing");
    }
    expect = end;
  }
  if (expect != d.boundaries.size())
    fail(MPSB_ERR_DATA, "Binned schema contains boundaries not referenced by a feature");
}

HostDataset quantize(const View& v, uint32_t max_bins, uint32_t threads) {
  validate_dense_view(v, max_bins);
  HostDataset d;
  d.rows = v.rows;
  d.features = v.features;
  d.max_bins = max_bins;
  const uint32_t groups = (v.features + kGroup - 1) / kGroup;
  std::vector<std::vector<float>> bounds(v.features);
  std::vector<uint64_t> missing(v.features, 0);
  BadValues bad;
  const Reader read{v, &bad};
  // One worker per group of eight features: gather their columns in one pass over the
  // rows, then sort and cut each column.
  parallel_for(worker_count(threads, groups), groups, [&](uint64_t gi) {
    const uint32_t f0 = uint32_t(gi) * kGroup, f1 = std::min(v.features, f0 + kGroup);
    std::vector<std::vector<float>> cols(f1 - f0);
    for (auto& c : cols) c.reserve(v.rows);
    for (uint64_t r = 0; r < v.rows; ++r) {
      for (uint32_t f = f0; f < f1; ++f) {
        const float x = read(r, f);
        if (std::isnan(x)) {
          ++missing[f];
        } else {
          cols[f - f0].push_back(x);
        }
      }
    }
    for (uint32_t f = f0; f < f1; ++f) {
      bounds[f] = boundaries_of(cols[f - f0], max_bins);
      std::vector<float>().swap(cols[f - f0]);
    }
  });
  bad.raise();
  for (uint32_t f = 0; f < v.features; ++f) {
    d.boundary_offset.push_back(d.boundaries.size());
    d.bin_count.push_back(uint32_t(bounds[f].size() + 1));
    d.missing_count.push_back(missing[f]);
    d.boundaries.insert(d.boundaries.end(), bounds[f].begin(), bounds[f].end());
  }
  validate_schema(d);
  d.finalize_layout();
  write_codes(v, d, threads);
  return d;
}

HostDataset transform(const View& v, const HostDataset& schema, uint32_t threads) {
  validate_dense_view(v, schema.max_bins);
  validate_schema(schema);
  if (v.features != schema.features)
    fail(MPSB_ERR_DATA, "Prediction input feature count does not match the binned schema");
  HostDataset d;
  d.rows = v.rows;
  d.features = schema.features;
  d.max_bins = schema.max_bins;
  d.boundaries = schema.boundaries;
  d.boundary_offset = schema.boundary_offset;
  d.bin_count = schema.bin_count;
  // The device layout must match the model's: the missing cell exists wherever the training
  // data had one. A NaN in a feature that had none at training time still routes by the
  // missing rule, so such features get the code too.
  d.missing_count = schema.missing_count;
  d.miss_flag = schema.miss_flag;
  // Bad values met here are raised by write_codes, which reads every value.
  BadValues unused;
  const Reader read{v, &unused};
  for (uint32_t f = 0; f < v.features; ++f) {
    if (d.miss_flag[f]) continue;
    for (uint64_t r = 0; r < v.rows; ++r) {
      if (std::isnan(read(r, f))) {
        d.miss_flag[f] = 1;
        break;
      }
    }
  }
  d.finalize_layout();
  write_codes(v, d, threads);
  return d;
}

}  // namespace mpsb

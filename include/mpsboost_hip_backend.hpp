// mpsboost_hip_backend.hpp - the reference's C++ backend interfaces implemented over the
// C ABI of libmpsboost_hip.so (mpsboost_hip.h). Header-only; compiled inside the
// reference's tree (it includes the reference's own include/mpsboost/backend.hpp), it
// is the drop-in replacement of MpsBackend (include/mpsboost/backend.hpp:137-195):
//
//   mpsboost::HipBackend backend;                       // device 0, FP64 row-order form
//   auto model = mpsboost::TrainRegressionModel(dataset, labels, weights, params,
//                                               backend, backend);
//
// The shared level-wise builder finds BuildLayerHistograms through dynamic_cast
// (src/core/tree_histogram.cpp:83-84,122-123) exactly as it finds MpsBackend's. For the
// whole fit on the device (histograms, scan, partition and leaf updates without host
// round trips per level) call mpsb_train_regressor directly instead.
//
// Error mapping: MPSB_ERR_DATA -> DataError, MPSB_ERR_TRAINING -> TrainingError, every
// other status -> BackendError("<stage>: <message>"), as the reference's backends raise.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "mpsboost/backend.hpp"
#include "mpsboost_hip.h"

namespace mpsboost {

class HipBackend final : public GradientComputer, public HistogramBuilder, public LayerHistogramBuilder {
 public:
  explicit HipBackend(int device = 0, mpsb_precision precision = MPSB_PRECISION_FP64) : precision_(precision) {
    Check(mpsb_session_create(device, &session_), "session");
  }
  ~HipBackend() override {
    if (dataset_) mpsb_dataset_free(dataset_);
    if (session_) mpsb_session_destroy(session_);
  }
  HipBackend(HipBackend&& o) noexcept { Swap(o); }
  HipBackend& operator=(HipBackend&& o) noexcept {
    Swap(o);
    return *this;
  }
  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;

  std::vector<GradientPair> ComputeSquaredError(const std::vector<double>& labels,
                                                 const std::vector<double>& predictions) const override {
    if (labels.size() != predictions.size()) throw TrainingError("Label and prediction lengths do not match");
    std::vector<GradientPair> out(labels.size());
    Check(mpsb_compute_squared_error(session_, labels.data(), predictions.data(), labels.size(),
                                     reinterpret_cast<double*>(out.data())),
          "gradients");
    return out;
  }

  NodeHistograms BuildHistograms(const BinnedDataset& dataset, const std::vector<std::uint64_t>& rows,
                                 const std::vector<GradientPair>& gradients) const override {
    const mpsb_dataset* ds = Upload(dataset);
    const std::size_t cells = Cells(dataset);
    std::vector<std::uint64_t> count(cells);
    std::vector<double> g(cells), h(cells);
    Check(mpsb_build_histograms(session_, ds, rows.data(), rows.size(), Pairs(gradients, dataset), precision_,
                                count.data(), g.data(), h.data()),
          "histogram");
    return Unflatten(dataset, count.data(), g.data(), h.data());
  }

  std::vector<NodeHistograms> BuildLayerHistograms(const BinnedDataset& dataset,
                                                   const std::vector<std::vector<std::uint64_t>>& node_rows,
                                                   const std::vector<GradientPair>& gradients) const override {
    const mpsb_dataset* ds = Upload(dataset);
    std::vector<std::uint64_t> offsets(1, 0), rows;
    for (const auto& r : node_rows) {
      rows.insert(rows.end(), r.begin(), r.end());
      offsets.push_back(rows.size());
    }
    const std::size_t cells = Cells(dataset);
    std::vector<std::uint64_t> count(cells * node_rows.size());
    std::vector<double> g(count.size()), h(count.size());
    Check(mpsb_build_layer_histograms(session_, ds, static_cast<std::uint32_t>(node_rows.size()), offsets.data(),
                                      rows.data(), Pairs(gradients, dataset), precision_, count.data(), g.data(),
                                      h.data()),
          "layer histogram");
    std::vector<NodeHistograms> out;
    out.reserve(node_rows.size());
    for (std::size_t i = 0; i < node_rows.size(); ++i)
      out.push_back(Unflatten(dataset, count.data() + i * cells, g.data() + i * cells, h.data() + i * cells));
    return out;
  }

  BackendTiming last_timing() const noexcept {
    mpsb_timing t{};
    BackendTiming out;
    if (mpsb_session_last_timing(session_, &t) != MPSB_OK) return out;
    out.gradient_seconds = t.gradient_seconds;
    out.histogram_encode_seconds = t.histogram_encode_seconds;
    out.histogram_command_seconds = t.histogram_command_seconds;
    out.hot_path_encode_seconds = t.hot_path_encode_seconds;
    out.hot_path_command_seconds = t.hot_path_command_seconds;
    out.pooled_buffer_reuse_count = t.pooled_buffer_reuse_count;
    out.pooled_buffer_allocation_count = t.pooled_buffer_allocation_count;
    return out;
  }

  mpsb_session* session() const noexcept { return session_; }

 private:
  static void Check(int status, const char* stage) {
    if (status == MPSB_OK) return;
    const std::string msg = mpsb_last_error();
    if (status == MPSB_ERR_DATA) throw DataError(msg);
    if (status == MPSB_ERR_TRAINING) throw TrainingError(msg);
    throw BackendError(std::string(stage) + ": " + msg);
  }

  static std::size_t Cells(const BinnedDataset& d) {
    std::size_t cells = 0;
    for (const auto& m : d.feature_metadata()) cells += m.bin_count;
    return cells;
  }

  static const double* Pairs(const std::vector<GradientPair>& gradients, const BinnedDataset& d) {
    static_assert(sizeof(GradientPair) == 2 * sizeof(double), "GradientPair must be two packed doubles");
    if (gradients.size() != d.rows()) throw TrainingError("Gradient count does not match binned dataset row count");
    return reinterpret_cast<const double*>(gradients.data());
  }

  // One device copy per dataset object: re-uploaded only when a different dataset (or a
  // moved one) arrives. The missing mask is read through IsMissing, the only public view.
  const mpsb_dataset* Upload(const BinnedDataset& d) const {
    if (dataset_ && cached_bins_ == d.bin_data() && cached_rows_ == d.rows() && cached_features_ == d.features())
      return dataset_;
    std::vector<std::uint32_t> bin_count;
    for (const auto& m : d.feature_metadata()) bin_count.push_back(m.bin_count);
    std::vector<std::uint8_t> missing(static_cast<std::size_t>(d.rows()) * d.features());
    for (std::uint32_t f = 0; f < d.features(); ++f) {
      if (d.feature_metadata()[f].missing_count == 0) continue;
      for (std::uint64_t r = 0; r < d.rows(); ++r)
        missing[static_cast<std::size_t>(f) * d.rows() + r] = d.IsMissing(r, f) ? 1 : 0;
    }
    mpsb_dataset* fresh = nullptr;
    Check(mpsb_dataset_from_bins(d.rows(), d.features(), d.max_bins(), bin_count.data(), d.bin_data(),
                                 d.storage() == BinStorage::kUInt8 ? 1 : 2, missing.data(), d.boundaries().data(),
                                 &fresh),
          "dataset upload");
    if (dataset_) mpsb_dataset_free(dataset_);
    dataset_ = fresh;
    cached_bins_ = d.bin_data();
    cached_rows_ = d.rows();
    cached_features_ = d.features();
    return dataset_;
  }

  static NodeHistograms Unflatten(const BinnedDataset& d, const std::uint64_t* count, const double* g,
                                  const double* h) {
    NodeHistograms out(d.features());
    std::size_t at = 0;
    for (std::uint32_t f = 0; f < d.features(); ++f) {
      out[f].resize(d.feature_metadata()[f].bin_count);
      for (auto& bin : out[f]) {
        bin.count = count[at];
        bin.gradient_sum = g[at];
        bin.hessian_sum = h[at];
        ++at;
      }
    }
    return out;
  }

  void Swap(HipBackend& o) noexcept {
    std::swap(session_, o.session_);
    std::swap(precision_, o.precision_);
    std::swap(dataset_, o.dataset_);
    std::swap(cached_bins_, o.cached_bins_);
    std::swap(cached_rows_, o.cached_rows_);
    std::swap(cached_features_, o.cached_features_);
  }

  mpsb_session* session_ = nullptr;
  mpsb_precision precision_ = MPSB_PRECISION_FP64;
  mutable mpsb_dataset* dataset_ = nullptr;
  mutable const void* cached_bins_ = nullptr;
  mutable std::uint64_t cached_rows_ = 0;
  mutable std::uint32_t cached_features_ = 0;
};

}  // namespace mpsboost

// Exposes the reference's own C++ CPU training path through the oracle C interface
// (oracle_capi.h), so tests can compare this repository's restatement (oracle.cpp)
// with the reference itself, bit for bit.
//
// TEST INFRASTRUCTURE ONLY. It is compiled by oracle/Makefile together with the
// reference sources where they lie (/root/reference/src/{core,cpu,io}); nothing is
// copied into this repository, and the result (oracle/_ref/libref.so) is git-ignored.
// The product never loads it.
//
// Built a second time with REF_BACKEND_HIP (oracle/_ref/libref_hip.so), the reference's
// own trainers run on include/mpsboost_hip_backend.hpp's HipBackend instead of
// CpuReferenceBackend: the drop-in boundary exercised by the reference's code itself.

#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpsboost/backend.hpp"
#include "mpsboost/binned_dataset.hpp"
#include "mpsboost/objective.hpp"
#include "mpsboost/trainer.hpp"
#include "mpsboost/tree.hpp"
#include "oracle_capi.h"
#ifdef REF_BACKEND_HIP
#include "mpsboost_hip_backend.hpp"
#endif

using namespace mpsboost;

#ifdef REF_BACKEND_HIP
using Backend = HipBackend;  // device 0, row-order FP64 form
#else
using Backend = const CpuReferenceBackend;
#endif

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
  } catch (const DataError& e) {
    put(e.what());
    return 1;
  } catch (const TrainingError& e) {
    put(e.what());
    return 2;
  } catch (const std::exception& e) {
    put(e.what());
    return 3;
  }
}

DenseMatrixView view_of(const void* data, uint64_t rows, uint32_t features, uint64_t rs,
                        uint64_t cs, int scalar) {
  DenseMatrixView v;
  v.data = data;
  v.rows = rows;
  v.features = features;
  v.row_stride_bytes = rs;
  v.column_stride_bytes = cs;
  v.scalar_type = scalar == 2 ? ScalarType::kFloat64 : ScalarType::kFloat32;
  v.source_contiguous = false;
  return v;
}

TreeTrainingParameters tree_params(const orc_tree_params& p) {
  TreeTrainingParameters t;
  t.max_depth = p.max_depth;
  t.max_leaves = p.max_leaves;
  t.max_active_leaves = p.max_active_leaves;
  t.min_samples_leaf = p.min_samples_leaf;
  t.min_child_weight = p.min_child_weight;
  t.reg_lambda = p.reg_lambda;
  t.reg_alpha = p.reg_alpha;
  t.max_delta_step = p.max_delta_step;
  t.gamma = p.gamma;
  t.min_gain_to_split = p.min_gain_to_split;
  if (p.split_strategy > 1) throw std::invalid_argument("unknown split strategy");
  t.split_strategy = p.split_strategy == 1 ? TreeTrainingParameters::SplitStrategy::kRandomThreshold
                                           : TreeTrainingParameters::SplitStrategy::kBestGain;
  t.growth_strategy = p.growth_strategy == 1 ? TreeTrainingParameters::GrowthStrategy::kLeafWise
                                             : TreeTrainingParameters::GrowthStrategy::kLevelWise;
  t.random_seed = p.random_seed;
  if (p.n_monotonic) t.monotonic_constraints.assign(p.monotonic, p.monotonic + p.n_monotonic);
  size_t at = 0;
  for (uint32_t i = 0; i < p.n_groups; ++i) {
    t.interaction_constraints.emplace_back(p.group_features + at,
                                           p.group_features + at + p.group_sizes[i]);
    at += p.group_sizes[i];
  }
  return t;
}

TrainingParameters train_params(const orc_train_params& p) {
  TrainingParameters t;
  t.n_estimators = p.n_estimators;
  t.learning_rate = p.learning_rate;
  t.max_bins = p.max_bins;
  if (p.objective > 4) throw std::invalid_argument("unknown training objective");
  t.objective = static_cast<TrainingParameters::Objective>(p.objective);
  t.tree = tree_params(p.tree);
  t.objective_alpha = p.objective_alpha;
  t.tweedie_variance_power = p.tweedie_variance_power;
  return t;
}

std::vector<GradientPair> pairs_of(const double* pairs, uint64_t n) {
  std::vector<GradientPair> gh(n);
  for (uint64_t i = 0; i < n; ++i) gh[i] = GradientPair{pairs[2 * i], pairs[2 * i + 1]};
  return gh;
}

void put_pairs(const std::vector<GradientPair>& gh, double* out) {
  for (size_t i = 0; i < gh.size(); ++i) {
    out[2 * i] = gh[i].gradient;
    out[2 * i + 1] = gh[i].hessian;
  }
}

orc_node node_of(const TreeNode& n) {
  orc_node o{};
  o.feature_index = n.feature_index;
  o.threshold_bin = n.threshold_bin;
  o.left_child = n.left_child;
  o.right_child = n.right_child;
  o.leaf_value = n.leaf_value;
  o.gain = n.gain;
  o.default_left = n.default_left ? 1 : 0;
  o.flags = n.flags;
  return o;
}

TreeNode tree_node_of(const orc_node& o) {
  TreeNode n;
  n.feature_index = o.feature_index;
  n.threshold_bin = o.threshold_bin;
  n.left_child = o.left_child;
  n.right_child = o.right_child;
  n.leaf_value = o.leaf_value;
  n.gain = o.gain;
  n.default_left = o.default_left != 0;
  n.flags = o.flags;
  return n;
}

struct ShimModel {
  bool multiclass = false;
  RegressionModel regression;
  MulticlassModel softmax;
};

}  // namespace

extern "C" {

int orc_quantize(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                 int scalar, uint32_t max_bins, void** out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    *out = new BinnedDataset(QuantizeDense(view_of(data, rows, features, rs, cs, scalar), max_bins));
  });
}

int orc_transform(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                  int scalar, const void* schema, void** out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    *out = new BinnedDataset(TransformDense(view_of(data, rows, features, rs, cs, scalar),
                                            static_cast<const BinnedDataset*>(schema)->schema()));
  });
}

void orc_dataset_free(void* d) { delete static_cast<BinnedDataset*>(d); }

int orc_dataset_info(const void* p, uint64_t* rows, uint32_t* features, uint32_t* max_bins,
                     uint32_t* storage_bits, uint64_t* boundary_total) {
  const auto& d = *static_cast<const BinnedDataset*>(p);
  *rows = d.rows();
  *features = d.features();
  *max_bins = d.max_bins();
  *storage_bits = d.storage() == BinStorage::kUInt8 ? 8 : 16;
  *boundary_total = d.boundaries().size();
  return 0;
}

int orc_dataset_meta(const void* p, uint32_t* bin_count, uint64_t* boundary_offset,
                     uint64_t* missing_count, float* boundaries) {
  const auto& d = *static_cast<const BinnedDataset*>(p);
  for (uint32_t f = 0; f < d.features(); ++f) {
    bin_count[f] = d.feature_metadata()[f].bin_count;
    boundary_offset[f] = d.feature_metadata()[f].boundary_offset;
    missing_count[f] = d.feature_metadata()[f].missing_count;
  }
  std::copy(d.boundaries().begin(), d.boundaries().end(), boundaries);
  return 0;
}

int orc_dataset_bins(const void* p, uint16_t* bins, uint8_t* missing) {
  const auto& d = *static_cast<const BinnedDataset*>(p);
  for (uint32_t f = 0; f < d.features(); ++f)
    for (uint64_t r = 0; r < d.rows(); ++r) {
      bins[f * d.rows() + r] = static_cast<uint16_t>(d.GetBin(r, f));
      missing[f * d.rows() + r] = d.IsMissing(r, f) ? 1 : 0;
    }
  return 0;
}

int orc_squared_error(const double* y, const double* p, uint64_t n, double* out, char* err,
                      uint64_t errlen) {
  return guarded(err, errlen, [&] {
    put_pairs(ComputeSquaredErrorGradients(std::vector<double>(y, y + n), std::vector<double>(p, p + n)),
              out);
  });
}

int orc_gradients(uint32_t objective, const double* y, const double* p, uint64_t n, double alpha,
                  uint32_t k, uint32_t target, double* out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const std::vector<double> labels(y, y + n);
    const std::vector<double> preds(p, p + (objective == 5 ? n * k : n));
    switch (objective) {
      case 0: put_pairs(ComputeSquaredErrorGradients(labels, preds), out); break;
      case 1: put_pairs(ComputeBinaryLogisticGradients(labels, preds), out); break;
      case 2: put_pairs(ComputeQuantileGradients(labels, preds, alpha), out); break;
      case 3: put_pairs(ComputePoissonGradients(labels, preds), out); break;
      case 4: put_pairs(ComputeTweedieGradients(labels, preds, alpha), out); break;
      case 5: put_pairs(ComputeMulticlassSoftmaxGradients(labels, preds, k, target), out); break;
      default: throw std::invalid_argument("unknown objective");
    }
  });
}

int orc_node_score(double g, double h, double l, double a, double* out, char* err, uint64_t n) {
  return guarded(err, n, [&] { *out = NodeScore(g, h, l, a); });
}
int orc_leaf_weight(double g, double h, double l, double a, double m, double* out, char* err,
                    uint64_t n) {
  return guarded(err, n, [&] { *out = LeafWeight(g, h, l, a, m); });
}
int orc_split_gain(double gl, double hl, double gr, double hr, double l, double a, double gamma,
                   double* out, char* err, uint64_t n) {
  return guarded(err, n, [&] { *out = SplitGain(gl, hl, gr, hr, l, a, gamma); });
}

int orc_histograms(const void* p, const double* pairs, const uint64_t* rows, uint64_t n_rows,
                   uint64_t* count, double* g, double* h, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const BinnedDataset*>(p);
    Backend backend;
    const NodeHistograms hist = backend.BuildHistograms(
        d, std::vector<uint64_t>(rows, rows + n_rows), pairs_of(pairs, d.rows()));
    size_t at = 0;
    for (const auto& fh : hist)
      for (const auto& c : fh) {
        count[at] = c.count;
        g[at] = c.gradient_sum;
        h[at] = c.hessian_sum;
        ++at;
      }
  });
}

int orc_train_tree(const void* p, const double* pairs, const orc_tree_params* tp, orc_node* nodes,
                   uint32_t capacity, uint32_t* n_nodes, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const BinnedDataset*>(p);
    Backend backend;
    const RegressionTree t =
        TrainSingleRegressionTree(d, pairs_of(pairs, d.rows()), tree_params(*tp), backend);
    *n_nodes = static_cast<uint32_t>(t.nodes().size());
    if (t.nodes().size() > capacity) throw std::invalid_argument("node capacity too small");
    for (size_t i = 0; i < t.nodes().size(); ++i) nodes[i] = node_of(t.nodes()[i]);
  });
}

int orc_tree_predict(const void* p, const orc_node* nodes, uint32_t n_nodes, uint32_t tree_features,
                     double* out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const BinnedDataset*>(p);
    std::vector<TreeNode> tn;
    for (uint32_t i = 0; i < n_nodes; ++i) tn.push_back(tree_node_of(nodes[i]));
    const RegressionTree t = RegressionTree::Restore(tree_features, std::move(tn));
    const std::vector<double> u = t.Predict(d);
    std::copy(u.begin(), u.end(), out);
  });
}

int orc_train_regressor(const void* p, const double* y, const double* w,
                        const orc_train_params* tp, void** model, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const BinnedDataset*>(p);
    Backend backend;
    auto* m = new ShimModel();
    try {
      m->regression = TrainRegressionModel(d, std::vector<double>(y, y + d.rows()),
                                           std::vector<double>(w, w + d.rows()), train_params(*tp),
                                           backend, backend);
    } catch (...) {
      delete m;
      throw;
    }
    *model = m;
  });
}

int orc_train_multiclass(const void* p, const double* y, const double* w, const orc_train_params* tp,
                         uint32_t k, void** model, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& d = *static_cast<const BinnedDataset*>(p);
    Backend backend;
    auto* m = new ShimModel();
    m->multiclass = true;
    try {
      m->softmax = TrainMulticlassSoftmaxModel(d, std::vector<double>(y, y + d.rows()),
                                               std::vector<double>(w, w + d.rows()),
                                               train_params(*tp), k, backend);
    } catch (...) {
      delete m;
      throw;
    }
    *model = m;
  });
}

int orc_model_info(const void* p, uint32_t* trees, uint32_t* classes, double* lr, double* base) {
  const auto& m = *static_cast<const ShimModel*>(p);
  if (m.multiclass) {
    *trees = m.softmax.tree_count();
    *classes = m.softmax.class_count();
    *lr = m.softmax.learning_rate();
    std::copy(m.softmax.base_scores().begin(), m.softmax.base_scores().end(), base);
  } else {
    *trees = m.regression.tree_count();
    *classes = 1;
    *lr = m.regression.learning_rate();
    base[0] = m.regression.base_score();
  }
  return 0;
}

int orc_model_tree(const void* p, uint32_t i, orc_node* nodes, uint32_t capacity, uint32_t* n) {
  const auto& m = *static_cast<const ShimModel*>(p);
  const auto& trees = m.multiclass ? m.softmax.trees() : m.regression.trees();
  if (i >= trees.size()) return 3;
  *n = static_cast<uint32_t>(trees[i].nodes().size());
  if (*n > capacity) return 3;
  for (uint32_t j = 0; j < *n; ++j) nodes[j] = node_of(trees[i].nodes()[j]);
  return 0;
}

int orc_model_predict(const void* p, const void* ds, double* out, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& m = *static_cast<const ShimModel*>(p);
    const auto& d = *static_cast<const BinnedDataset*>(ds);
    const std::vector<double> u = m.multiclass ? m.softmax.PredictMargins(d) : m.regression.Predict(d);
    std::copy(u.begin(), u.end(), out);
  });
}

void orc_model_free(void* p) { delete static_cast<ShimModel*>(p); }

int orc_model_serialize(const void* p, uint8_t* out, uint64_t capacity, uint64_t* size, char* err,
                        uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const auto& m = *static_cast<const ShimModel*>(p);
    const std::vector<uint8_t> bytes = m.multiclass ? SerializeModel(m.softmax) : SerializeModel(m.regression);
    *size = bytes.size();
    if (!out) return;
    if (capacity < bytes.size()) throw std::invalid_argument("model byte buffer is too small");
    std::memcpy(out, bytes.data(), bytes.size());
  });
}

int orc_model_deserialize(const uint8_t* data, uint64_t size, void** model, char* err, uint64_t errlen) {
  return guarded(err, errlen, [&] {
    const std::vector<uint8_t> bytes(data, data + size);
    auto* m = new ShimModel();
    try {
      m->multiclass = size >= 16 && data[12] == 1;  // model kind field of the header
      if (m->multiclass) {
        m->softmax = DeserializeMulticlassModel(bytes);
      } else {
        m->regression = DeserializeModel(bytes);
      }
    } catch (...) {
      delete m;
      throw;
    }
    *model = m;
  });
}

}  // extern "C"

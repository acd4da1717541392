/*
 * C interface shared by the two CPU checkers under oracle/:
 *   liboracle.so   - this repository's own restatement of the reference CPU trainer
 *                    (oracle.cpp), and
 *   _ref/libref.so - the reference's own C++ sources compiled unchanged, wrapped by
 *                    ref_shim.cpp.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (mpsboost_amd/, include/) links,
 * loads or calls these libraries; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do, and only to check or time against them.
 *
 * Both libraries export exactly the same symbols so one ctypes wrapper (oracle.py)
 * drives either, and tests can compare the restatement with the reference itself.
 * Every function returns 0 on success, or 1 (data error), 2 (training error),
 * 3 (invalid argument) with a message copied into err[0..errlen).
 */
#ifndef MPSBOOST_ORACLE_CAPI_H
#define MPSBOOST_ORACLE_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat tree node, field-for-field the reference TreeNode
 * (include/mpsboost/tree.hpp:31-42). */
typedef struct orc_node {
  uint32_t feature_index;
  uint32_t threshold_bin;
  uint32_t left_child;
  uint32_t right_child;
  double leaf_value;
  double gain;
  uint8_t default_left;
  uint8_t flags; /* bit 0 = leaf */
  uint8_t pad[6];
} orc_node;

/* TreeTrainingParameters (include/mpsboost/tree.hpp:47-72). Constraint arrays are
 * optional: n_monotonic == 0 means none; interaction groups are given flattened with
 * per-group sizes. */
typedef struct orc_tree_params {
  uint32_t max_depth;
  uint32_t max_leaves;
  uint32_t max_active_leaves;
  uint32_t split_strategy;  /* 0 best gain, 1 random threshold */
  uint64_t min_samples_leaf;
  double min_child_weight;
  double reg_lambda;
  double reg_alpha;
  double max_delta_step;
  double gamma;
  double min_gain_to_split;
  uint32_t growth_strategy; /* 0 level-wise (1 leaf-wise: rejected by the restatement) */
  uint32_t random_seed;
  const int8_t* monotonic;
  uint32_t n_monotonic;
  uint32_t n_groups;
  const uint32_t* group_sizes;
  const uint32_t* group_features;
} orc_tree_params;

/* TrainingParameters (include/mpsboost/trainer.hpp:20-36). objective: 0 squared error,
 * 1 binary logistic, 2 quantile, 3 poisson, 4 tweedie. */
typedef struct orc_train_params {
  uint32_t n_estimators;
  uint32_t max_bins;
  uint32_t objective;
  uint32_t pad;
  double learning_rate;
  double objective_alpha;
  double tweedie_variance_power;
  orc_tree_params tree;
} orc_train_params;

/* scalar_type: 1 float32, 2 float64 (ScalarType, include/mpsboost/binned_dataset.hpp:23). */
int orc_quantize(const void* data, uint64_t rows, uint32_t features, uint64_t row_stride_bytes,
                 uint64_t column_stride_bytes, int scalar_type, uint32_t max_bins, void** out,
                 char* err, uint64_t errlen);
int orc_transform(const void* data, uint64_t rows, uint32_t features, uint64_t row_stride_bytes,
                  uint64_t column_stride_bytes, int scalar_type, const void* schema_dataset,
                  void** out, char* err, uint64_t errlen);
void orc_dataset_free(void* dataset);
/* Threads for the per-feature loops (histograms, missing statistics); results are
 * bit-identical for any count. Default 1. */
void orc_set_threads(uint32_t threads);
int orc_dataset_info(const void* dataset, uint64_t* rows, uint32_t* features, uint32_t* max_bins,
                     uint32_t* storage_bits, uint64_t* boundary_total);
/* bin_count/missing_count/boundary_offset: [features]; boundaries: [boundary_total]. */
int orc_dataset_meta(const void* dataset, uint32_t* bin_count, uint64_t* boundary_offset,
                     uint64_t* missing_count, float* boundaries);
/* Feature-major [features][rows] bins and missing mask. */
int orc_dataset_bins(const void* dataset, uint16_t* bins, uint8_t* missing);

int orc_squared_error(const double* labels, const double* predictions, uint64_t n,
                      double* out_pairs, char* err, uint64_t errlen);
/* objective 0..4 as in orc_train_params plus 5 = softmax for one class; for softmax
 * predictions are row-major margins [n][class_count]. */
int orc_gradients(uint32_t objective, const double* labels, const double* predictions,
                  uint64_t n, double alpha, uint32_t class_count, uint32_t target_class,
                  double* out_pairs, char* err, uint64_t errlen);
int orc_node_score(double g, double h, double lambda, double alpha, double* out, char* err,
                   uint64_t errlen);
int orc_leaf_weight(double g, double h, double lambda, double alpha, double max_delta_step,
                    double* out, char* err, uint64_t errlen);
int orc_split_gain(double gl, double hl, double gr, double hr, double lambda, double alpha,
                   double gamma, double* out, char* err, uint64_t errlen);

/* CPU FP64 histogram in the reference view: cells laid out feature after feature with
 * bin_count cells each, missing rows counted in bin 0. */
int orc_histograms(const void* dataset, const double* pairs, const uint64_t* rows,
                   uint64_t n_rows, uint64_t* count, double* gradient, double* hessian,
                   char* err, uint64_t errlen);

int orc_train_tree(const void* dataset, const double* pairs, const orc_tree_params* params,
                   orc_node* nodes, uint32_t capacity, uint32_t* n_nodes, char* err,
                   uint64_t errlen);
int orc_tree_predict(const void* dataset, const orc_node* nodes, uint32_t n_nodes,
                     uint32_t tree_features, double* out, char* err, uint64_t errlen);

int orc_train_regressor(const void* dataset, const double* labels, const double* weights,
                        const orc_train_params* params, void** model, char* err,
                        uint64_t errlen);
int orc_train_multiclass(const void* dataset, const double* labels, const double* weights,
                         const orc_train_params* params, uint32_t class_count, void** model,
                         char* err, uint64_t errlen);
/* base_scores receives 1 value (regression) or class_count values (multiclass). */
int orc_model_info(const void* model, uint32_t* tree_count, uint32_t* class_count,
                   double* learning_rate, double* base_scores);
int orc_model_tree(const void* model, uint32_t tree, orc_node* nodes, uint32_t capacity,
                   uint32_t* n_nodes);
/* Regression: out[rows]; multiclass: row-major margins out[rows][class_count]. */
int orc_model_predict(const void* model, const void* dataset, double* out, char* err,
                      uint64_t errlen);
void orc_model_free(void* model);

/* Model files (compiled reference only, src/io/model_format.cpp): SerializeModel bytes of a
 * trained model (out == NULL queries *size), and DeserializeModel /
 * DeserializeMulticlassModel of bytes (kind read from the header), with the reference's
 * error text in err. */
int orc_model_serialize(const void* model, uint8_t* out, uint64_t capacity, uint64_t* size, char* err,
                        uint64_t errlen);
int orc_model_deserialize(const uint8_t* data, uint64_t size, void** model, char* err, uint64_t errlen);

#ifdef __cplusplus
}
#endif

#endif

/*
 * mpsboost_hip.h - C ABI of the MI355X (gfx950) histogram-boosting training backend.
 *
 * This is the drop-in boundary: plain C types, caller-owned buffers with explicit sizes,
 * opaque handles, integer status codes. No C++ or framework types cross it and no
 * exception escapes it. Each entry point names the reference interface it replaces
 * (paths are relative to the reference repository).
 *
 * Threading: one mpsb_session owns one HIP stream and its device memory; different
 * sessions may be used from different threads, one session is not re-entrant
 * (the reference's MpsBackend has the same rule: include/mpsboost/backend.hpp:137-149).
 *
 * Every function returning int returns MPSB_OK (0) or a status below; the message of the
 * last failure on the calling thread is mpsb_last_error().
 */
#ifndef MPSBOOST_HIP_H
#define MPSBOOST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes --- */
/* Distinct codes replace the reference's exception types (include/mpsboost/backend.hpp:
 * 22-25 BackendError, include/mpsboost/objective.hpp:17-20 TrainingError,
 * include/mpsboost/binned_dataset.hpp:18-21 DataError). */
enum mpsb_status {
  MPSB_OK = 0,
  MPSB_ERR_DATA = 1,               /* DataError: input data violates the binning contract */
  MPSB_ERR_TRAINING = 2,           /* TrainingError: math/contract failure during training */
  MPSB_ERR_INVALID_ARGUMENT = 3,   /* bad pointer, size or option */
  MPSB_ERR_DEVICE_UNAVAILABLE = 4, /* BackendError: no usable gfx950 device */
  MPSB_ERR_OUT_OF_MEMORY = 5,      /* BackendError: device allocation failed */
  MPSB_ERR_LAUNCH = 6,             /* BackendError: HIP launch or runtime failure */
  MPSB_ERR_INJECTED = 7,           /* BackendError: fault forced by mpsb_debug_inject_fault */
  MPSB_ERR_INTERNAL = 8,
  MPSB_ERR_TIMEOUT = 9             /* BackendError: a host wait for the device passed its deadline
                                    * (MPSB_WAIT_TIMEOUT_S seconds, default 120) */
};

/* Histogram arithmetic of a fit (the precision forms of this backend).
 *  FP64  - the reference CPU backend's arithmetic, operation for operation: FP64 gradients,
 *          every histogram cell summed in row order (bin 0 including the missing rows),
 *          missing statistics in row order, FP64 sibling subtraction with the Hessian clamp,
 *          root statistics in row order, sequential FP64 prefix scan. Trees, gains, leaves
 *          and predictions are bit-identical to CpuReferenceBackend + TrainRegressionModel.
 *          Each cell is a serial chain over the node's rows, so this form is for parity and
 *          moderate sizes, not throughput.
 *  FP32  - FP32 gradients (as the reference's device path computes them), FP32 per-block
 *          histograms, FP64 fixed-order cross-block reduction.
 *  INT16 / INT8 - per-round max-abs scaled, stochastically rounded gradients, exact
 *          integer histograms; FP64 split scan and leaf solve.
 *  FP64_FIXED - FP64 gradients summed exactly in 64-bit fixed point (power-of-two scale per
 *          round): order-independent, fully parallel, deterministic. Trees match FP64 except
 *          where two candidate gains agree to within the rounding of the row-order sums. */
enum mpsb_precision {
  MPSB_PRECISION_FP64 = 0,
  MPSB_PRECISION_FP32 = 1,
  MPSB_PRECISION_INT16 = 2,
  MPSB_PRECISION_INT8 = 3,
  MPSB_PRECISION_FP64_FIXED = 4
};

typedef struct mpsb_session mpsb_session;
typedef struct mpsb_dataset mpsb_dataset;
typedef struct mpsb_model mpsb_model;

/* Message of the last failure on this thread ("" if none). Valid until the next call. */
const char* mpsb_last_error(void);

/* --------------------------------------------------------------- device query ------- */
/* BackendInfo (include/mpsboost/backend.hpp:30-35). available == 0 is not an error. */
typedef struct mpsb_backend_info {
  int32_t available;
  uint32_t compute_units;
  uint64_t total_memory;
  char device_name[96];
  char arch[32];
} mpsb_backend_info;

/* QueryBackendInfo (include/mpsboost/backend.hpp:197-200). */
int mpsb_query_backend_info(mpsb_backend_info* out);

/* ------------------------------------------------------------------- sessions ------- */
/* MpsBackend::MpsBackend / ~MpsBackend (include/mpsboost/backend.hpp:144-149): one stream,
 * pooled device memory, loaded code objects (embedded in this library; there is no
 * metallib path). Fails with MPSB_ERR_DEVICE_UNAVAILABLE when no gfx950 device exists. */
int mpsb_session_create(int32_t device, mpsb_session** out);
void mpsb_session_destroy(mpsb_session* session);

/* BackendTiming (include/mpsboost/backend.hpp:95-103), extended with device-side times
 * measured by HIP events on the session stream. Times are seconds of the latest call. */
typedef struct mpsb_timing {
  double gradient_seconds;
  double histogram_encode_seconds;  /* host work before the histogram launches */
  double histogram_command_seconds; /* launch to completion, device clock */
  double hot_path_encode_seconds;
  double hot_path_command_seconds;
  uint64_t pooled_buffer_reuse_count;
  uint64_t pooled_buffer_allocation_count;
} mpsb_timing;

/* Histogram launch shape of a session. No counterpart in the reference (its Metal kernels use
 * fixed shapes); written by the autotuner (mpsboost_amd/autotune.py) or by a caller that
 * has measured its own shape class. Takes effect from the next fit. */
typedef struct mpsb_tuning {
  uint32_t lds_budget;        /* bytes of histogram LDS per block, 16384 .. device maximum */
  uint32_t target_blocks;     /* histogram blocks per level launch, 64 .. 65536 */
  uint32_t min_slab_rows;     /* smallest row slab per block, 64 .. 2^24 */
  uint32_t block;             /* threads per histogram block: 256, 512 or 1024 */
  uint32_t words_per_thread;  /* 32-bit code words loaded per thread: 1, 2 or 4 */
  uint32_t planner;           /* who plans each tree level: 0 automatic (the device, except for
                               * fits under 2^20 rows while other fits are live), 1 the host,
                               * 2 the device wherever the form and parameters allow it */
} mpsb_tuning;

int mpsb_session_get_tuning(const mpsb_session* session, mpsb_tuning* out);
int mpsb_session_set_tuning(mpsb_session* session, const mpsb_tuning* tuning);

/* MpsBackend::last_timing (include/mpsboost/backend.hpp:162). */
int mpsb_session_last_timing(const mpsb_session* session, mpsb_timing* out);

/* Test switch for failure handling: the countdown-th time stage `stage` is reached the call
 * fails with MPSB_ERR_INJECTED and a "<stage>: ..." message, before any device work of that
 * stage. stage: 0 off, 1 upload, 2 gradients, 3 histogram, 4 split scan, 5 partition,
 * 6 prediction, 7 allocation. No counterpart in the reference (its failure paths are only
 * reachable by real faults). */
int mpsb_debug_inject_fault(mpsb_session* session, int32_t stage, int32_t countdown);

/* --------------------------------------------------------------- binned datasets ---- */
/* QuantizeDense (include/mpsboost/binned_dataset.hpp:128-129, src/core/binned_dataset.cpp:
 * 126-188): identical boundaries, bins and missing mask to the reference; the stored layout
 * is row-major codes for the device. scalar_type 1 = float32, 2 = float64. threads = host
 * worker threads (0 = all cores). */
int mpsb_quantize_dense(const void* data, uint64_t rows, uint32_t features,
                        uint64_t row_stride_bytes, uint64_t column_stride_bytes,
                        int32_t scalar_type, uint32_t max_bins, uint32_t threads,
                        mpsb_dataset** out);

/* TransformDense (include/mpsboost/binned_dataset.hpp:131-135): frozen schema of `schema`. */
int mpsb_transform_dense(const void* data, uint64_t rows, uint32_t features,
                         uint64_t row_stride_bytes, uint64_t column_stride_bytes,
                         int32_t scalar_type, const mpsb_dataset* schema, uint32_t threads,
                         mpsb_dataset** out);

/* QuantizeDense / TransformDense with the work on the session's GPU (the same reference code
 * as the two entries above): the matrix is copied to the device in row chunks, each feature's
 * values are radix-sorted there, the quantile cut and the code assignment run there, and the
 * codes stay resident in the session for the next fit or prediction (a host copy is kept in
 * the dataset as well). Boundaries, bins and missing counts equal the host entries', except
 * that a zero boundary of a column holding both -0.0 and +0.0 may carry the other sign (the
 * reference's std::sort leaves that order open; the bins are the same). Errors as the host
 * entries: MPSB_ERR_DATA with the reference's messages. */
int mpsb_quantize_dense_device(mpsb_session* session, const void* data, uint64_t rows, uint32_t features,
                               uint64_t row_stride_bytes, uint64_t column_stride_bytes,
                               int32_t scalar_type, uint32_t max_bins, mpsb_dataset** out);
int mpsb_transform_dense_device(mpsb_session* session, const void* data, uint64_t rows, uint32_t features,
                                uint64_t row_stride_bytes, uint64_t column_stride_bytes,
                                int32_t scalar_type, const mpsb_dataset* schema, mpsb_dataset** out);

/* A dense input matrix copied once to device memory (bytes and strides as given), for
 * repeated sampled binning: the reference's forest materialises matrix[rows][:, features] on
 * the host and quantises it for every tree (src/mpsboost/estimators/forest_sampling.py:42-63).
 * Not tied to a session: any session on `device` may read it, concurrently. */
typedef struct mpsb_matrix mpsb_matrix;
int mpsb_matrix_upload(int32_t device, const void* data, uint64_t rows, uint32_t features,
                       uint64_t row_stride_bytes, uint64_t column_stride_bytes, int32_t scalar_type,
                       mpsb_matrix** out);
void mpsb_matrix_free(mpsb_matrix* matrix);
/* QuantizeDense (src/core/binned_dataset.cpp:126-188) of matrix[rows][:, features] on the
 * session's GPU, read in place through the two index lists: rows may repeat (bootstrap
 * samples), features are distinct column indices in output order. Same results and errors as
 * mpsb_quantize_dense_device on the materialised sub-matrix. */
int mpsb_quantize_matrix_sampled(mpsb_session* session, const mpsb_matrix* matrix, const uint64_t* rows,
                                 uint64_t n_rows, const uint32_t* features, uint32_t n_features,
                                 uint32_t max_bins, mpsb_dataset** out);

/* Adopt an already binned dataset: the reference's BinnedDataset exposes its feature-major
 * bins (bin_data(), include/mpsboost/binned_dataset.hpp:101-105), feature metadata and
 * missing mask; this copies them (bin_bytes 1 or 2) so no pointer is retained. */
int mpsb_dataset_from_bins(uint64_t rows, uint32_t features, uint32_t max_bins,
                           const uint32_t* bin_count, const void* bins_feature_major,
                           int32_t bin_bytes, const uint8_t* missing_feature_major,
                           const float* boundaries, mpsb_dataset** out);

void mpsb_dataset_free(mpsb_dataset* dataset);

typedef struct mpsb_dataset_info {
  uint64_t rows;
  uint32_t features;
  uint32_t max_bins;
  uint32_t storage_bits;     /* 8 or 16, the reference's storage width */
  uint32_t code_bytes;       /* 1 or 2, the device code width */
  uint32_t row_stride;       /* codes per row in the device layout */
  uint32_t cells;            /* histogram cells per node incl. one missing cell per feature with NaN */
  uint64_t boundary_total;
} mpsb_dataset_info;

int mpsb_dataset_info_get(const mpsb_dataset* dataset, mpsb_dataset_info* out);
/* The dataset's codes in the device layout (rows x row_stride codes of code_bytes, row-major,
 * missing = bin_count of the feature): *size receives the byte count; out may be NULL to query
 * it, else it must hold capacity >= *size bytes. BinnedDataset::bin_data()
 * (include/mpsboost/binned_dataset.hpp:101-105) is the reference's feature-major counterpart. */
int mpsb_dataset_codes(const mpsb_dataset* dataset, void* out, uint64_t capacity, uint64_t* size);
/* Per-feature metadata (FeatureBinMetadata, include/mpsboost/binned_dataset.hpp:40-46). */
int mpsb_dataset_meta(const mpsb_dataset* dataset, uint32_t* bin_count,
                      uint64_t* boundary_offset, uint64_t* missing_count, float* boundaries);
/* Reference (feature-major) view of bins and missing mask: [features][rows]. */
int mpsb_dataset_bins(const mpsb_dataset* dataset, uint16_t* bins, uint8_t* missing);

/* ------------------------------------------------- backend operations (test hooks) --- */
/* GradientComputer::ComputeSquaredError (include/mpsboost/backend.hpp:52-58): pairs
 * out[2i] = g, out[2i+1] = h, computed on the device in FP64. */
int mpsb_compute_squared_error(mpsb_session* session, const double* labels,
                               const double* predictions, uint64_t n, double* out_pairs);

/* HistogramBuilder::BuildHistograms (include/mpsboost/backend.hpp:64-75) for one node, in the
 * reference view: cells feature after feature, bin_count each, missing rows counted in bin 0.
 * pairs = [2*rows] (g, h) in FP64; precision as in mpsb_precision. Outputs are [cells]. */
int mpsb_build_histograms(mpsb_session* session, const mpsb_dataset* dataset,
                          const uint64_t* rows, uint64_t n_rows, const double* pairs,
                          int32_t precision, uint64_t* count, double* gradient, double* hessian);

/* LayerHistogramBuilder::BuildLayerHistograms (include/mpsboost/backend.hpp:124-135): all
 * nodes in one launch. Node i owns rows[node_offsets[i] .. node_offsets[i+1]); outputs are
 * [n_nodes][cells] in node order. */
int mpsb_build_layer_histograms(mpsb_session* session, const mpsb_dataset* dataset,
                                uint32_t n_nodes, const uint64_t* node_offsets,
                                const uint64_t* rows, const double* pairs, int32_t precision,
                                uint64_t* count, double* gradient, double* hessian);

/* Allocation-free histogram build. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). It computes what mpsb_build_histograms computes for one node, in the
 * fp32, int16, int8 and fp64_fixed forms (the int forms with the same rounding stream, so
 * their results are identical; fp32 within its rounding), from device-resident inputs into
 * device-resident outputs, and allocates nothing:
 *   codes  - the dataset's codes in the device layout (mpsb_dataset_codes: rows x row_stride
 *            codes of code_bytes each, row-major), on the current device;
 *   rows   - n_rows row ids (uint64), on the device; pairs - [2 * dataset rows] FP64 (g, h);
 *   count, gradient, hessian - [reference cells] outputs (cells feature after feature,
 *            bin_count each, bin 0 including the missing rows), on the device;
 *   stream - a hipStream_t of the current device (NULL: the null stream); the call returns
 *            after the stream has completed its work.
 * Limits: at most 1,048,576 dataset rows and node rows, 32,768 device histogram cells and
 * 64 feature tiles; beyond them MPSB_ERR_INVALID_ARGUMENT. Its scratch memory is one fixed
 * 42 MiB device array per GPU, held until the process ends; calls that overlap on one GPU race
 * on it and must not be made. Errors as mpsb_build_histograms (row ids past the dataset:
 * MPSB_ERR_TRAINING). */
int mpsb_build_histograms_noalloc(const mpsb_dataset* dataset, const void* codes, const uint64_t* rows,
                                  uint64_t n_rows, const double* pairs, int32_t precision, uint64_t* count,
                                  double* gradient, double* hessian, void* stream);

/* MpsBackend::BuildBaselineHistogramsForTest (include/mpsboost/backend.hpp:169-172): one
 * thread per cell scanning every node row, FP64; a slow, obviously-correct device kernel. */
int mpsb_build_baseline_histograms(mpsb_session* session, const mpsb_dataset* dataset,
                                   const uint64_t* rows, uint64_t n_rows, const double* pairs,
                                   uint64_t* count, double* gradient, double* hessian);

/* Split-scan options: TreeTrainingParameters fields used by the scan
 * (include/mpsboost/tree.hpp:57-66). */
typedef struct mpsb_split_params {
  uint64_t min_samples_leaf;
  double min_child_weight;
  double reg_lambda;
  double reg_alpha;
  double gamma;
  double min_gain_to_split;
} mpsb_split_params;

/* SplitScanCandidate (include/mpsboost/backend.hpp:108-119) plus the missing-value
 * direction, which the reference's device sketch lacks. */
typedef struct mpsb_split_candidate {
  uint32_t valid;
  uint32_t feature;
  uint32_t threshold_bin;
  uint32_t default_left;
  uint64_t left_count;
  uint64_t right_count;
  double left_gradient_sum;
  double left_hessian_sum;
  double right_gradient_sum;
  double right_hessian_sum;
  double gain;
} mpsb_split_candidate;

/* MpsBackend::ScanSplitsForTest (include/mpsboost/backend.hpp:173-180): the best candidate of
 * every feature (out[features]) for one node, with the reference's full FindBestSplit rules
 * (src/core/tree_split.cpp:142-234), FP64 on the device. Parent statistics: SumRows in row
 * order (src/core/tree_split.cpp:119-140) in the FP64 form, the node's histogram totals in
 * the others. The node's best overall is *best (may be NULL). */
int mpsb_scan_splits(mpsb_session* session, const mpsb_dataset* dataset, const uint64_t* rows,
                     uint64_t n_rows, const double* pairs, const mpsb_split_params* params,
                     int32_t precision, mpsb_split_candidate* out, mpsb_split_candidate* best);

/* MpsBackend::PartitionRowsForTest (include/mpsboost/backend.hpp:181-185): stable partition of
 * rows by "missing ? default_left : bin <= threshold_bin" (src/core/tree_split.cpp:273-289).
 * left/right must hold n_rows entries; the counts are returned. */
int mpsb_partition_rows(mpsb_session* session, const mpsb_dataset* dataset, const uint64_t* rows,
                        uint64_t n_rows, uint32_t feature, uint32_t threshold_bin,
                        int32_t default_left, uint64_t* left, uint64_t* n_left, uint64_t* right,
                        uint64_t* n_right);

/* MpsBackend::RunVectorAddForTest / RunVectorAdd (include/mpsboost/backend.hpp:167-168,205-207). */
int mpsb_vector_add(mpsb_session* session, const float* a, const float* b, uint64_t n, float* out);

/* -------------------------------------------------------------------- training ------ */
/* TreeNode (include/mpsboost/tree.hpp:31-42), 40 bytes. */
typedef struct mpsb_tree_node {
  uint32_t feature_index;
  uint32_t threshold_bin;
  uint32_t left_child;
  uint32_t right_child;
  double leaf_value;
  double gain;
  uint8_t default_left;
  uint8_t flags; /* bit 0: leaf */
  uint8_t pad[6];
} mpsb_tree_node;

/* TrainingParameters + TreeTrainingParameters (include/mpsboost/trainer.hpp:20-36,
 * include/mpsboost/tree.hpp:47-72). objective: 0 squared error (the device path), other
 * values are rejected by the device trainer. Constraint arrays may be NULL / 0. */
typedef struct mpsb_train_params {
  uint32_t n_estimators;
  uint32_t max_bins;
  uint32_t objective;
  uint32_t precision;      /* mpsb_precision */
  double learning_rate;
  double objective_alpha;
  double tweedie_variance_power;
  uint32_t max_depth;
  uint32_t max_leaves;
  uint32_t max_active_leaves;
  uint32_t split_strategy; /* 0 best gain, 1 random threshold */
  uint64_t min_samples_leaf;
  double min_child_weight;
  double reg_lambda;
  double reg_alpha;
  double max_delta_step;
  double gamma;
  double min_gain_to_split;
  uint32_t growth_strategy; /* 0 level-wise */
  uint32_t random_seed;
  const int8_t* monotonic;
  uint32_t n_monotonic;
  uint32_t n_groups;
  const uint32_t* group_sizes;
  const uint32_t* group_features;
  uint64_t rounding_seed;  /* stochastic-rounding stream of INT16/INT8 */
  uint32_t leaf_solve;     /* leaf values: 0 automatic (renewed in the FP32, INT16 and INT8 forms,
                            * from the histograms in the FP64 and FP64_FIXED forms), 1 from the
                            * histograms' sums, 2 renewed: once the tree is grown, each leaf is
                            * solved again from exact sums of the round's FP64 gradients over its
                            * rows (no counterpart in the reference, whose FP64 sums need none;
                            * Shi et al., Quantized Training of Gradient Boosting Decision Trees,
                            * NeurIPS 2022). Not for softmax in the FP32 form (no FP64 gradients
                            * are kept there) nor for trees of more than 65,535 nodes. */
  uint32_t near_ties;      /* INT16 / INT8: 0 auto (resolve in INT16, keep in INT8), 1 resolve,
                            * 2 keep. Resolving, the split scan bounds each gain's stochastic-
                            * rounding noise (six standard deviations); a node whose best
                            * candidate another could outrank within those bounds is scanned
                            * again over its contender features from exact fixed-point sums of
                            * the round's FP64 gradients (the FP64_FIXED form's arithmetic;
                            * device-planned trees of at most 1,024 features, at most 64 nodes a
                            * level). With leaf renewal, an INT16 fit then equals the FP64_FIXED
                            * fit. INT8's noise makes most nodes ambiguous, so resolving there
                            * costs more than the FP64_FIXED form itself. No counterpart in the
                            * reference. */
  uint32_t class_rounds;   /* softmax only: 0 the reference's class order (each class tree sees the
                            * margins the earlier class trees of its round updated,
                            * src/core/multiclass_trainer.cpp:132-147); 1 batched rounds, the
                            * Jacobi form: every class's gradients come from the round's starting
                            * margins (one softmax evaluation per row for all classes) and the
                            * round's K class trees grow together as one device-planned tree of K
                            * roots, so each level's histogram, reduction, split scan, planner and
                            * partition is one launch for all K trees. The stochastic-rounding
                            * forms share one scale over the round's K columns. Departs from the
                            * reference's results by design; FP64_FIXED, INT16 and INT8 forms,
                            * near ties kept, leaf cap not below 2^max_depth, K x 2^(max_depth-1)
                            * <= 2048, K x rows < 2^32. No counterpart in the reference. */
} mpsb_train_params;

/* TrainRegressionModel (include/mpsboost/trainer.hpp:133-139, src/core/trainer.cpp:226-266)
 * with the device-resident backend, as the reference's TrainRegressorMps wires it
 * (src/python/binding_model.cpp:34-45). weights may be NULL (all ones). */
int mpsb_train_regressor(mpsb_session* session, const mpsb_dataset* dataset,
                         const double* labels, const double* weights,
                         const mpsb_train_params* params, mpsb_model** out);

/* TrainMulticlassSoftmaxModel (include/mpsboost/trainer.hpp:141-147,
 * src/core/multiclass_trainer.cpp:104-149): per round one tree per class, class order,
 * each class seeing the margins already updated this round. labels are class ids. */
int mpsb_train_multiclass(mpsb_session* session, const mpsb_dataset* dataset,
                          const double* labels, const double* weights, uint32_t class_count,
                          const mpsb_train_params* params, mpsb_model** out);

typedef struct mpsb_model_info {
  uint32_t tree_count;
  uint32_t class_count;      /* 1 for regression */
  uint32_t features;
  uint32_t precision;
  double learning_rate;
  double base_scores[256];   /* base_score, or one per class */
  /* RegressionModel::objective / objective_alpha / tweedie_variance_power and
   * MulticlassModel::class_labels (src/python/binding_model.cpp:65-106). */
  uint32_t objective;        /* TrainingParameters::Objective order; 0 for softmax models */
  uint32_t max_bins;
  double objective_alpha;
  double tweedie_variance_power;
  double class_labels[256];  /* softmax models: one per class */
} mpsb_model_info;

int mpsb_model_info_get(const mpsb_model* model, mpsb_model_info* out);
/* RegressionModel::trees / RegressionTree::nodes (include/mpsboost/trainer.hpp:52,
 * include/mpsboost/tree.hpp:57). */
int mpsb_model_tree(const mpsb_model* model, uint32_t tree, mpsb_tree_node* nodes,
                    uint32_t capacity, uint32_t* n_nodes);
/* RegressionModel::Predict / MulticlassModel::PredictMargins (include/mpsboost/trainer.hpp:
 * 57-60,101), on the device. dataset must use the model's schema. out: [rows] or
 * row-major [rows][class_count]. */
int mpsb_model_predict(mpsb_session* session, const mpsb_model* model,
                       const mpsb_dataset* dataset, double* out);
void mpsb_model_free(mpsb_model* model);

/* The model's frozen QuantizationSchema as a dataset of zero rows (RegressionModel::schema,
 * include/mpsboost/trainer.hpp:38-85): pass it as `schema` to mpsb_transform_dense(_device)
 * to bin prediction inputs of a trained or loaded model. */
int mpsb_model_schema(const mpsb_model* model, mpsb_dataset** out);
/* MulticlassModel::SetClassLabels (src/python/binding_model.cpp:105-107): `count` finite,
 * strictly increasing labels, one per class. */
int mpsb_model_set_class_labels(mpsb_model* model, const double* labels, uint32_t count);

/* Work and time accounting of the fit that produced `model`, used for throughput and
 * roofline reporting. Byte counts are algorithmic (what the algorithm must move), not
 * measured traffic. */
typedef struct mpsb_fit_stats {
  uint64_t trees;
  uint64_t histogram_rows;        /* sum over built nodes of their row counts */
  uint64_t histogram_bytes;       /* histogram_rows * (features * code_bytes + gh bytes) */
  uint64_t histogram_launches;
  double histogram_seconds;       /* device time of the histogram kernels (HIP events) */
  double scan_seconds;
  double partition_seconds;
  double gradient_seconds;
  double reduce_seconds;
  double train_seconds;           /* host wall clock of the boosting loop */
  uint64_t levels;
  /* Stage times (the *_seconds above) are HIP-event times of a sample of the trees, every
   * 4th: timed_trees of them; the histogram bytes and launches of those trees: */
  uint64_t timed_trees;
  uint64_t histogram_timed_bytes;
  uint64_t histogram_timed_launches;
  /* Trees whose levels were planned on the device (k_plan); the others were planned on the
   * host, one level round trip each. */
  uint64_t device_planned_trees;
  double renewal_seconds;         /* leaf renewal (HIP events of the timed trees) */
  uint64_t renewed_trees;         /* trees whose leaves were renewed */
  uint64_t ambiguous_nodes;       /* near-tie resolution: nodes marked ambiguous */
  uint64_t resolved_nodes;        /* ... and scanned again exactly */
  uint64_t resolved_features;     /* contender features scanned again, over the resolved nodes */
  /* Host wall clock around the boosting loop: from the call to the first tree's launches
   * (labels uploaded, base score known, buffers set up); from the device's last kernel to the
   * returned model (tree and leaf read-back); and the base score's own serial pass over the
   * labels, which overlaps the set-up. */
  double setup_seconds;
  double finish_seconds;
  double base_score_seconds;
  /* Root histograms built without their count plane: every tree after a fit's first takes the
   * root's counts from it (a root's counts do not change from tree to tree). Unit-Hessian
   * int16 / int8 fits sum gradients alone in 4-byte cells of 64-feature tiles; the other int16 /
   * int8 fits [G|H] words, and unit-Hessian FP64_FIXED fits fixed-point G, in 8-byte cells. */
  uint64_t root_gradient_only_builds;
} mpsb_fit_stats;

int mpsb_model_fit_stats(const mpsb_model* model, mpsb_fit_stats* out);

/* ------------------------------------------------------------------ model files ----- */
/* SerializeModel (src/io/model_format.cpp:46-68, src/io/model_payload.cpp:143-185): the
 * reference's MPSBMOD 1.4 bytes (32-byte header, FNV-1a 64 checksum, 40-byte node records).
 * *size receives the byte count; out == NULL only queries it, else capacity must hold it. */
int mpsb_model_serialize(const mpsb_model* model, uint8_t* out, uint64_t capacity, uint64_t* size);
/* DeserializeModel / DeserializeMulticlassModel (src/io/model_format.cpp:70-123): either
 * model kind, minors 0-4, with the reference's validation and messages (MPSB_ERR_TRAINING,
 * schema faults MPSB_ERR_DATA). */
int mpsb_model_deserialize(const uint8_t* data, uint64_t size, mpsb_model** out);
/* The same with the loader's kind check: kind 0 = DeserializeModel ("model kind is
 * incompatible with regression loader"), 1 = DeserializeMulticlassModel, -1 = either. */
int mpsb_model_deserialize_kind(const uint8_t* data, uint64_t size, int32_t kind, mpsb_model** out);
/* SaveModelFile / LoadModelFile (src/io/model_file_io.cpp): the file is written to a
 * same-directory temporary, synchronised and renamed over `path`. */
int mpsb_model_save(const mpsb_model* model, const char* path);
int mpsb_model_load(const char* path, mpsb_model** out);
/* LoadModelFile (kind 0) / LoadMulticlassModelFile (kind 1), -1 = either. */
int mpsb_model_load_kind(const char* path, int32_t kind, mpsb_model** out);

#ifdef __cplusplus
}
#endif

#endif /* MPSBOOST_HIP_H */

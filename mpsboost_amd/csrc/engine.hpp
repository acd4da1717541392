// Session, fit state and host orchestration of the device hot path.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

namespace mpsb {

enum Buf : int {
  kBufCodes,
  kBufCodesCM,
  kBufCellOffset,
  kBufBinCount,
  kBufMissCell,
  kBufMissCode,
  kBufTiles,
  kBufWordOff,
  kBufCellLds,
  kBufGH,
  kBufGHB,
  kBufGH64,
  kBufLabels,
  kBufPred,
  kBufWeights,
  kBufPairs,
  kBufRowsA,
  kBufRowsB,
  kBufPartial,
  kBufHistA,
  kBufHistB,
  kBufDesc,       // per-level small descriptors (slabs, reduce tables, scan nodes, part nodes)
  kBufCand,
  kBufPerFeature,
  kBufChunkNode,
  kBufBlockChunk,
  kBufChunkLeft,
  kBufBits,
  kBufErr,
  kBufMaxabs,
  kBufMono,
  kBufMasks,
  kBufTmpA,
  kBufTmpB,
  kBufTmpC,
  kBufTrees,
  kBufSeqFeat,
  kBufNodeStats,
  kBufScales,
  kBufLeafAdd,
  kBufBinX,       // device binning: input chunk, keys, sorted keys, sort scratch, tables
  kBufBinKeys,
  kBufBinSorted,
  kBufBinTemp,
  kBufBinCand,
  kBufBinBounds,
  kBufBinMeta,
  kBufPlanPart,   // device-side level planning: partition nodes, layers, slabs, builds, scan
  kBufPlanLayerA,
  kBufPlanLayerB,
  kBufPlanSlabs,
  kBufPlanBuild,
  kBufPlanScan,
  kBufPlanCtl,
  kBufPlanTrees,
  kBufPlanLeafTable,
  kBufGradStash,  // softmax: FP64 (g, h) of the max pass for the write pass
  kBufRenew,      // leaf renewal: per-leaf fixed-point sums and counts, own max |g|, |h|
  kBufLeafInfo,   // device-planned trees: node and bounds per leaf ordinal
  kBufRenewNodes, // host-planned trees: renewed leaf values by tree and node
  kBufResolve,    // near-tie resolution: plan, masks, scan nodes, scales, candidates, counters
  kBufResolveHist,  // near-tie resolution: exact histograms of the resolved nodes
  kBufRootLayer,  // batched softmax rounds: the class roots (level 0 of the round's tree)
  kBufRootRows,   // batched softmax rounds: row of each level-0 position (position mod rows)
  kBufTilesG,     // root histograms without counts: tiles, cell and word offsets, the root's counts
  kBufCellLdsG,
  kBufWordOffG,
  kBufRootCount,
  kBufCount
};

enum FaultStage : int {
  kFaultNone = 0,
  kFaultUpload = 1,
  kFaultGradients = 2,
  kFaultHistogram = 3,
  kFaultScan = 4,
  kFaultPartition = 5,
  kFaultPredict = 6,
  kFaultAlloc = 7,
};

// Histogram launch shape. Defaults from sweeps on MI355X at 10M x 128 (block 256/512/1024 x
// LDS 32/64/96/160 KB x 1024/2048/4096 blocks x 1/2/4 code words per thread): 1024-thread
// blocks (16 waves) with a 64 KB tile (two blocks per CU) and 4 words per thread (with two
// rows in flight per thread) load fastest; ~1024 blocks per launch beat 2048 once the
// cross-block reduction is counted (fewer slab partials to write and re-read).
struct Tuning {
  uint32_t lds_budget = 65544;     // bytes of histogram LDS per block (32 features x 256 bins x 8 B + 1)
  uint32_t target_blocks = 1024;   // histogram blocks per level launch
  uint32_t min_slab_rows = 2048;
  uint32_t block = 1024;
  uint32_t wpt = 4;                // 32-bit code words loaded per thread (1, 2 or 4)
  uint32_t planner = 0;            // 0 automatic, 1 host, 2 device (mpsb_tuning::planner)
};

// Histogram launch shape of a dataset (engine.cpp): feature tiles of the row, each tile's LDS
// layout, per precision form.
struct TilePlan {
  std::vector<TileDesc> tiles;
  std::vector<uint32_t> word_off;  // [tile][64][4] LDS offset of each code (| kReplicated)
  std::vector<uint32_t> cell_lds;  // [cells] LDS offset of each output cell (| kReplicated)
  uint32_t lds_bytes = 0;
  uint32_t wpt = 1;
  bool interleaved = false;        // feature-interleaved tiles (plan_interleaved)
  uint32_t slots = 0;              // interleaved: feature slots per bin row (64, 32 or 16)
};
int hist_kind(int precision);
uint32_t bytes_per_cell(int kind, bool unit_h);
uint32_t gh_bytes(int precision, bool unit_h);
uint32_t slab_cap(int precision, bool unit_h);
uint32_t slab_cap_g32(int precision);
bool plan_g32_tiles(const HostDataset& d, uint32_t budget, TilePlan& plan);
TilePlan plan_tiles(const HostDataset& d, uint32_t bpc, uint32_t budget, uint32_t block, uint32_t want_wpt);
// mpsb_build_histograms_noalloc (noalloc.hip): one node's histograms from device-resident
// codes, rows and FP64 pairs into device outputs, with no device allocation.
void histograms_noalloc(const HostDataset& d, const void* codes, const uint64_t* rows, uint64_t n_rows,
                        const double* pairs, int precision, uint64_t* out_cnt, double* out_g, double* out_h,
                        hipStream_t stream);
// Raises the TrainingError of the earliest failing stage recorded in a device error word
// (objective: the regression objective, -1 for softmax); returns when flags is 0.
void raise_device_errors(uint32_t flags, uint32_t first, int objective);

struct Session {
  int device = 0;
  uint64_t id = 0;              // process-wide serial number (error messages)
  double wait_limit_s = 120.0;  // deadline of a host wait for a level signal
  hipStream_t stream = nullptr;
  std::array<void*, kBufCount> ptr{};
  std::array<size_t, kBufCount> cap{};
  uint64_t reuse = 0, allocs = 0;
  mpsb_timing timing{};
  int fault_stage = 0, fault_countdown = 0;
  uint64_t uploaded_uid = 0;
  uint32_t lds_max = 65536;
  uint32_t cus = 256;
  Tuning tune;
  void* pinned = nullptr;
  size_t pinned_cap = 0;
  std::vector<hipEvent_t> events;
  // Host-mapped coherent block the device writes each level's candidates into:
  // [0, 8) sequence number, [8, 12) error word, [64, ...) candidates.
  uint8_t* sig = nullptr;      // host address
  uint8_t* sig_dev = nullptr;  // the same block as the device addresses it
  size_t sig_cap = 0;
  uint64_t sig_seq = 0;

  ~Session();
  void* get(Buf b, size_t bytes);  // grow-only pooled device buffer
  void release(Buf b);             // free one pooled buffer (after the stream drains)
  void* host_staging(size_t bytes);
  void check_fault(int stage, const char* what);
  void sync(const char* stage);
  void upload(const HostDataset& d);
  void adopt_codes(const HostDataset& d);  // codes already written to kBufCodes on the device
  // The device-binned dataset whose codes only this session's code buffer holds; before the
  // buffer is overwritten (or the session ends) they are copied to that dataset's host side.
  std::weak_ptr<DeviceCodes> resident_codes;
  void evict_codes();
  hipEvent_t event(size_t i);
  void ensure_signal(size_t candidates);
  // Spin until the device has published `seq`; stream errors end the wait.
  void wait_signal(uint64_t seq, const char* stage);
};

struct Params {
  uint32_t n_estimators = 100;
  double lr = 0.1;
  uint32_t max_bins = 256;
  uint32_t objective = 0;
  int precision = MPSB_PRECISION_FP64;
  uint32_t max_depth = 6, max_leaves = 0, max_active_leaves = 0;
  uint64_t min_samples_leaf = 1;
  double min_child_weight = 0.0, lambda = 1.0, alpha = 0.0, max_delta = 0.0, gamma = 0.0, min_gain = 0.0;
  bool random_threshold = false;
  uint32_t growth = 0, seed = 0;
  std::vector<int8_t> monotonic;
  std::vector<std::vector<uint32_t>> groups;
  uint64_t rounding_seed = 0;
  double objective_alpha = 0.5, tweedie_power = 1.5;
  uint32_t leaf_solve = 0;  // mpsb_train_params::leaf_solve
  uint32_t near_ties = 0;   // mpsb_train_params::near_ties
  uint32_t class_rounds = 0;  // mpsb_train_params::class_rounds
};

Params params_from(const mpsb_train_params& p);
void validate_train_params(const Params& p, const HostDataset& d);

struct Model {
  uint32_t features = 0;
  uint32_t classes = 1;
  int precision = 0;
  double lr = 0.1;
  uint32_t objective = 0;          // TrainingParameters::Objective code (model files)
  double objective_alpha = 0.5;
  double tweedie_power = 1.5;
  std::vector<double> class_labels;  // softmax models: 0 .. classes-1 as trained here
  std::vector<double> base;
  HostDataset schema;  // boundaries / layout only, no codes
  std::vector<std::vector<mpsb_tree_node>> trees;
  mpsb_fit_stats stats{};
};

Model train_regressor(Session& s, const HostDataset& d, const double* labels, const double* weights,
                      const Params& p);
Model train_multiclass(Session& s, const HostDataset& d, const double* labels, const double* weights,
                       uint32_t classes, const Params& p);
void predict(Session& s, const Model& m, const HostDataset& d, double* out);

// QuantizeDense / TransformDense on the device (binning.hip): the same dataset as quantize() /
// transform(), left resident in the session.
HostDataset quantize_on_device(Session& s, const View& v, uint32_t max_bins);
HostDataset transform_on_device(Session& s, const View& v, const HostDataset& schema);

// A dense input matrix copied to device memory once (its original bytes and strides), shared
// read-only by every session on its device.
struct DeviceMatrix {
  int device = 0;
  uint8_t* data = nullptr;
  uint64_t rows = 0;
  uint32_t features = 0;
  uint64_t row_stride = 0, col_stride = 0;
  int scalar = 1;
  uint64_t bytes = 0;
  ~DeviceMatrix();
};
DeviceMatrix* upload_matrix(int device, const View& v);
// QuantizeDense of matrix[rows][:, features] (rows may repeat) from the resident matrix.
HostDataset quantize_sampled(Session& s, const DeviceMatrix& m, const uint64_t* rows, uint64_t n_rows,
                             const uint32_t* features, uint32_t n_features, uint32_t max_bins);

// Model files in the reference's MPSBMOD 1.4 format (model_io.cpp).
std::vector<uint8_t> serialize_model(const Model& m);
// expect_kind: -1 either kind, 0 the regression loader's check, 1 the multiclass loader's.
Model deserialize_model(const uint8_t* data, size_t size, int expect_kind = -1);
void save_model_file(const Model& m, const std::string& path);
Model load_model_file(const std::string& path, int expect_kind = -1);

// Test hooks (reference MpsBackend methods).
void hook_squared_error(Session& s, const double* y, const double* p, uint64_t n, double* out);
void hook_layer_histograms(Session& s, const HostDataset& d, uint32_t n_nodes, const uint64_t* offsets,
                           const uint64_t* rows, const double* pairs, int precision, uint64_t* cnt,
                           double* g, double* h);
void hook_baseline_histograms(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n,
                              const double* pairs, uint64_t* cnt, double* g, double* h);
void hook_scan(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n, const double* pairs,
               const mpsb_split_params& sp, int precision, mpsb_split_candidate* out,
               mpsb_split_candidate* best);
void hook_partition(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n, uint32_t feature,
                    uint32_t threshold, int default_left, uint64_t* left, uint64_t* n_left, uint64_t* right,
                    uint64_t* n_right);
void hook_vector_add(Session& s, const float* a, const float* b, uint64_t n, float* out);

// FP64 objective math on the host, same operations as src/core/objective.cpp:289-318.
double host_leaf_weight(double g, double h, double lambda, double alpha, double max_delta);

}  // namespace mpsb

// Internal types shared by the host units of libmpsboost_hip.so.
//
// Layout decisions (all sized for one MI355X: 288 GB HBM3E, 256 CUs in 8 XCDs, 160 KB LDS):
//  * bins live on the device as a ROW-MAJOR code matrix, row_stride codes per row (a
//    multiple of 16 bytes), so a node's rows are gathered as whole 128-byte lines;
//  * a feature with missing values gets one extra histogram cell (code = bin_count), so
//    missing statistics fall out of the histogram instead of a second pass over a mask;
//  * node histograms are three planes (count u32, G 8 bytes, H 8 bytes) over
//    cells = sum_f (bin_count_f + has_missing_f).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpsboost_hip.h"

namespace mpsb {

// Sticky error bits the kernels OR into the fit's error word; the host turns them into the
// reference's exception messages when the fit ends.
enum : uint32_t {
  kErrNonFinite = 1u,      // supplied gradient / Hessian not finite or Hessian negative
  kErrCount = 2u,          // histogram counts do not sum to the node count
  kErrSubtract = 4u,       // sibling subtraction: child exceeds parent
  kErrMissing = 8u,        // missing statistics exceed the parent
  kErrOverflow = 16u,      // prefix or score overflow
  kErrPrefix = 32u,        // non-missing prefix exceeded the parent
  kErrPredict = 64u,       // malformed tree during prediction
  kErrPrediction = 128u,   // a prediction / margin update left a non-finite value
  kErrObjective = 256u,    // the objective's gradient or Hessian overflowed
  kErrExpFirst = 512u,     // exp argument above 709 (Poisson log mean, Tweedie first term)
  kErrExpSecond = 1024u,   // Tweedie second term exp argument above 709
  kErrWeighted = 2048u,    // weighted gradient / Hessian overflowed
  // Raised by the device-side level planner (the host path throws them directly).
  kErrLeafGradient = 4096u,     // a leaf's gradient sum is not finite
  kErrLeafHessian = 8192u,      // a leaf's Hessian sum is not finite
  kErrLeafNegative = 16384u,    // a leaf's Hessian sum is negative
  kErrLeafDenominator = 32768u, // Hessian sum plus reg_lambda not finite and positive
  kErrLeafOverflow = 65536u,    // a leaf value overflowed
  kErrBounds = 131072u,         // monotone leaf bounds inconsistent
  kErrSplitCount = 262144u,     // child counts do not add up to the node count
};
// The error word is two 32-bit words: [0] every bit raised (sticky OR), [1] the first bit
// raised (0 while none), so the earliest failure is the one reported.

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(int status, const std::string& msg) { throw Error(status, msg); }

#define MPSB_HIP(stage, expr)                                                              \
  do {                                                                                     \
    hipError_t mpsb_e_ = (expr);                                                           \
    if (mpsb_e_ != hipSuccess) {                                                           \
      ::mpsb::fail(mpsb_e_ == hipErrorOutOfMemory ? MPSB_ERR_OUT_OF_MEMORY : MPSB_ERR_LAUNCH, \
                   std::string(stage) + ": " + hipGetErrorString(mpsb_e_));                \
    }                                                                                      \
  } while (0)

// ------------------------------------------------------------------ datasets -------
struct Session;

// Codes that a device binning left in a session's code buffer. They reach the host only when
// something reads them there (fetch), or when the session is about to overwrite that buffer
// or end while the dataset still exists (the session copies them out first). Binning and
// training on one session thus never move the code matrix over PCIe.
struct DeviceCodes {
  std::mutex m;
  Session* session = nullptr;  // whose code buffer holds them, until copied out
  size_t bytes = 0;
  bool copied = false;
  std::vector<uint8_t> host;
  const std::vector<uint8_t>& fetch();  // the codes on the host (copied once, under the lock)
  void copy_out_locked();               // with `m` held and `session` set
};

struct HostDataset {
  uint64_t rows = 0;
  uint32_t features = 0;
  uint32_t max_bins = 0;
  std::vector<float> boundaries;
  std::vector<uint64_t> boundary_offset;
  std::vector<uint32_t> bin_count;
  std::vector<uint64_t> missing_count;

  uint32_t code_bytes = 1;     // 1 when every feature fits bin_count + has_missing <= 256
  uint32_t row_stride = 0;     // codes per row; row_stride * code_bytes % 16 == 0
  std::vector<uint8_t> codes;  // rows * row_stride * code_bytes, row-major (empty with `device`)
  std::shared_ptr<DeviceCodes> device;  // device binning: the codes, copied to the host on demand
  // The codes on the host: `codes`, or the device binning's codes (copied on the first call).
  const std::vector<uint8_t>& host_codes() const { return device ? device->fetch() : codes; }
  std::vector<uint32_t> cell_offset;  // features + 1
  uint32_t cells = 0;
  uint64_t uid = 0;  // identity for the session's upload cache
  // Features whose device layout carries a missing cell: the ones with NaN in this data
  // (and, for transformed data, the ones that had NaN at training time).
  std::vector<uint8_t> miss_flag;

  bool has_missing(uint32_t f) const { return miss_flag[f] != 0; }
  // Code of (row, f) in `c`, the codes on the host (host_codes()).
  uint32_t code(const std::vector<uint8_t>& c, uint64_t row, uint32_t f) const {
    const size_t at = (size_t(row) * row_stride + f) * code_bytes;
    return code_bytes == 1 ? c[at] : uint32_t(c[at]) | (uint32_t(c[at + 1]) << 8);
  }
  void finalize_layout();  // code width, stride, cell offsets, uid
};

struct View {
  const uint8_t* data;
  uint64_t rows;
  uint32_t features;
  uint64_t row_stride;
  uint64_t col_stride;
  int scalar;
};

void validate_dense_view(const View& v, uint32_t max_bins);
HostDataset quantize(const View& v, uint32_t max_bins, uint32_t threads);
HostDataset transform(const View& v, const HostDataset& schema, uint32_t threads);
void validate_schema(const HostDataset& d);

// -------------------------------------------------------------- device structs ----
// Plain structs shared with the kernels (kernels.hip).
struct TileDesc {
  uint32_t word0;      // first 32-bit word of the tile inside a row
  uint32_t words;      // words per row in this tile
  uint32_t cell0;      // first histogram cell of the tile
  uint32_t ncells;     // output cells of the tile
  uint32_t lds_cells;  // LDS cells of the tile (the trash cell is lds_cells)
  uint32_t pad[3];
};

// LDS offsets in word_cell_off / cell_lds carry this flag for a feature of at most 16 cells
// given 16 replicas (one per row slot of a wavefront, 16 cells apart): lanes of one
// instruction then never update the same address, the replicas are summed at the flush.
constexpr uint32_t kReplicated = 0x80000000u;

struct SlabDesc {
  uint32_t begin;  // position range in the row buffer
  uint32_t end;
};

struct LevelHist {
  uint32_t* cnt;  // [slots][cells]
  uint64_t* g;    // FP64 bits (FP32 form) or int64 (integer forms)
  uint64_t* h;
};

// Parent statistics of a node: exact integers for the integer forms, FP64 otherwise.
struct NodeStats {
  uint64_t count;
  double g;
  double h;
  int64_t gq;
  int64_t hq;
};

struct ScanNode {
  uint32_t slot;         // histogram slot in the level
  uint32_t node_index;   // tree node index (random-threshold hash)
  uint32_t depth;
  uint32_t mask_word0;   // offset of the allowed-feature mask (interaction constraints)
  uint32_t parent_from_hist;  // 1: histogram totals of feature 0; 2: ScanLaunch::parent_stats[i]
  uint32_t pad;
  double lo, hi;         // monotone leaf bounds
  NodeStats parent;
};

// One node of a level in the device-side tree state: its tree node id, row segment and
// statistics, as the level planner hands it from one level to the next.
struct LayerNode {
  uint32_t node;
  uint32_t begin, count;
  uint32_t pad;
  double lo, hi;   // monotone leaf bounds
  double leaf;     // the node's leaf value (while it is a leaf)
  NodeStats stats;
};

struct Candidate {
  uint32_t valid;
  uint32_t feature;
  uint32_t threshold;
  uint32_t default_left;
  double gain;
  NodeStats left;
  NodeStats right;
  NodeStats parent;  // statistics the scan used for the node
  double l_lo, l_hi, r_lo, r_hi;
  // Stochastic-rounding forms: `bound` bounds the rounding noise of `gain`, `rival` is the
  // largest gain + bound among the node's (or feature's) other candidates; `ambiguous` marks a
  // node whose best candidate the rounding noise could have swapped (rival >= gain - bound).
  double bound, rival;
  uint32_t ambiguous, pad2;
};

static_assert(sizeof(Candidate) % 8 == 0, "candidates are copied as 8-byte words");

struct ScanParams {
  uint64_t min_samples_leaf;
  double min_child_weight, lambda, alpha, max_delta, gamma, min_gain;
  uint32_t random_threshold, seed;
  uint32_t has_mono, has_mask;
  double scale_g, scale_h;  // integer forms: value = q * scale
  uint32_t unit_h, integer;
  uint32_t ref_view;        // 1: bin 0 includes the missing rows (row-order FP64 form)
  uint32_t pad;
  const double* scales;     // device [mult_g, mult_h, value_g, value_h] of this round, or null
  // Stochastic-rounding forms: k of the noise bound k * quantum * sqrt(rows) of a sum (k = 3:
  // six standard deviations, a row's rounding error having variance at most quantum^2 / 4);
  // 0: no bounds, no ambiguity marks.
  double noise_k;
};

// kPartSplitLeaf: a split whose two children are leaves (the last level): every row gets its
// child's leaf value directly, with no scatter of rows into children.
enum PartAction : uint32_t { kPartLeaf = 0, kPartSplit = 1, kPartSplitLeaf = 2 };

struct PartNode {
  uint32_t begin, count;     // segment in the row buffer
  uint32_t action;
  uint32_t feature;
  uint32_t threshold;
  uint32_t default_left;
  uint32_t missing_code;     // code meaning missing for `feature` (0xFFFFFFFF if none)
  uint32_t left_count;
  uint32_t chunk0;           // first chunk of this segment
  uint32_t column_major;     // 1: read the split feature from the column-major copy
  uint32_t leaf_node;        // tree node id of the leaf (kPartLeaf) / the left child (kPartSplitLeaf)
  uint32_t leaf_node_right;  // kPartSplitLeaf: the right child's node id
  double leaf_value;         // added (times the learning rate) for leaves; the left child's
  double leaf_right;         // kPartSplitLeaf: the right child's leaf value
};

// Histogram accumulation kinds. The last three build a root that holds every row without its
// count plane: the root's per-cell counts are the same in every tree of a fit and come from the
// fit's first root histogram. kHistG32: unit-Hessian int16/int8 gradient sums in int32 cells;
// kHistQNC: int16/int8 [G:32|H:32] words; kHistF64NC: FP64_FIXED unit-Hessian G in int64 cells.
enum HistKind : int { kHistF32 = 0, kHistFix64 = 1, kHistQ = 2, kHistSeq = 3, kHistG32 = 4, kHistQNC = 5,
                      kHistF64NC = 6 };

// Root kinds without counts (see HistKind).
inline bool hist_kind_no_count(int kind) {
  const int k = kind & 0xFF;
  return k == kHistG32 || k == kHistQNC || k == kHistF64NC;
}

// ------------------------------------------------------------- kernel launchers ---
struct HistLaunch {
  const uint8_t* codes;
  uint32_t row_stride_bytes;
  const uint32_t* rows;  // nullptr: identity
  const void* gh;
  const SlabDesc* slabs;
  uint32_t n_slabs;
  const TileDesc* tiles;
  uint32_t n_tiles;
  const uint32_t* word_cell_off;  // [tile][64 words][4] LDS offset of each code (| kReplicated)
  const uint32_t* cell_lds;       // [cells] LDS offset of each output cell within its tile (| kReplicated)
  void* partial;                  // planes of n_slab_cap * cells
  uint64_t plane;                 // elements per partial plane
  uint32_t cells;
  uint32_t code_bytes;
  uint32_t lds_bytes;
  int kind;                       // HistKind
  int unit_h;
  int block;
  int wpt;                        // 32-bit code words loaded per thread: 1, 2 or 4
  uint32_t gh_by_pos;             // 1: gh[i] belongs to row position i, else to row index i
  uint32_t interleaved;           // feature-interleaved tiles (one-byte codes only): slots per bin row, 64, 32 or 16; 0: packed
  const uint32_t* dyn_slabs;      // non-null: n_slabs is a bound, the count is here
};
void launch_hist(const HistLaunch& a, hipStream_t s);

struct ReduceLaunch {
  const void* partial;
  uint64_t plane;
  uint32_t cells;
  const uint32_t* slab_first;   // [n_build + 1] slab ranges per built node
  const uint32_t* build_slot;   // [n_build] slot written
  const int32_t* sibling_slot;  // [n_build] slot of the sibling, -1 none
  const uint32_t* parent_slot;  // [n_build] slot in the parent level
  uint32_t n_build;
  const uint32_t* dyn_build;    // non-null: n_build is a bound, the count is here
  LevelHist out;
  LevelHist parent;
  int kind;
  int unit_h;
  uint32_t* err;
  const uint32_t* counts;       // kinds without counts (a root): the root's per-cell row counts
};
void launch_reduce(const ReduceLaunch& a, hipStream_t s);
void launch_copy_u32(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t s);

void launch_baseline_hist(const uint8_t* codes, uint32_t row_stride, uint32_t code_bytes,
                          const uint32_t* rows, uint32_t n_rows, const double* pairs,
                          const uint32_t* cell_feature, const uint32_t* cell_code, uint32_t cells,
                          uint64_t* cnt, double* g, double* h, hipStream_t s);

// Row-order FP64 form: one wavefront per (node, group of features), lane j walking the
// node's rows in order for its feature, so every cell's additions happen in row order.
struct SeqNode {
  uint32_t begin, count;  // segment in the row buffer
  uint32_t slot;          // histogram slot written
  uint32_t pad;
};
struct SeqLaunch {
  const uint8_t* codes;
  uint32_t row_stride;  // codes per row
  uint32_t code_bytes;
  const uint32_t* rows;  // nullptr: identity positions
  const double* gh;      // FP64 (g, h) pairs per row
  const SeqNode* nodes;
  uint32_t n_nodes;
  const uint32_t* feat_list;  // features handled, ascending
  uint32_t n_feat;
  uint32_t fg;                // features (lanes) per wavefront, <= 64
  uint32_t cpf;               // LDS cells per lane; 0: accumulate in the output planes
  uint32_t missing_only;      // 1: only the missing cells (siblings built by subtraction)
  const uint32_t* cell_offset;
  const uint32_t* bin_count;
  LevelHist out;
  uint32_t cells;
  uint32_t unit_h;
  uint32_t* err;
};
void launch_hist_seq(const SeqLaunch& a, hipStream_t s);
// sibling = parent - built, FP64 per cell with the reference's Hessian clamp.
void launch_sub_seq(const uint32_t* build_slot, const int32_t* sibling_slot, const uint32_t* parent_slot,
                    uint32_t n_build, LevelHist out, LevelHist parent, uint32_t cells, uint32_t unit_h,
                    uint32_t* err, hipStream_t s);
// SumRows in row order over positions [begin, begin + n) of `rows` (nullptr: identity).
void launch_sum_rows(const double* gh, const uint32_t* rows, uint32_t begin, uint32_t n, NodeStats* out,
                     hipStream_t s);

struct ScanLaunch {
  LevelHist hist;
  uint32_t cells;
  const uint32_t* cell_offset;  // [F+1]
  const uint32_t* bin_count;    // [F]
  const uint32_t* miss_cell;    // [F] cell of the missing rows, 0xFFFFFFFF if none
  uint32_t features;
  const ScanNode* nodes;
  uint32_t n_nodes;
  ScanParams p;
  const int8_t* mono;
  const uint32_t* masks;
  Candidate* out;
  Candidate* per_feature;  // [n_nodes][F]: required by the wave-parallel scan, optional else
  const NodeStats* parent_stats;  // [n_nodes], read for parent_from_hist == 2
  uint32_t* err;
  int sequential;          // 1: one thread per feature, the reference's sequential FP64 prefix
  const uint32_t* dyn_nodes;  // non-null: n_nodes is a bound, the count is here
  uint32_t* contend;       // noise_k > 0: [n_nodes][(F + 31) / 32] features that could hold the
                           // node's best split (their gain + bound reaches its gain - bound)
};
void launch_scan(const ScanLaunch& a, hipStream_t s);


// Copies n candidates and the error word to host-mapped memory, then publishes seq.
void launch_signal(const Candidate* cand, uint32_t n, const uint32_t* err, Candidate* host_cand,
                   uint32_t* host_flags, unsigned long long* host_seq, unsigned long long seq, hipStream_t s);

// Counts of one level written by the device-side planner and read by that level's kernels,
// which are launched with grids for the largest possible counts (blocks past a count exit).
struct LevelCtl {
  uint32_t n_part;    // partition nodes (the level's nodes)
  uint32_t n_chunks;  // partition chunks
  uint32_t n_build;   // histograms built for the next level
  uint32_t n_slabs;   // slabs of those builds
  uint32_t n_next;    // nodes of the next level (its split scan)
  // 1: the mark launch takes its chunks from the planner's per-XCD lists (block b: XCD b mod 8,
  // entry b / 8 of that XCD's list), each list the chunks whose rows lie in one eighth of the
  // row range, so the XCD's L2 gathers that eighth's leaf-id writes (the last level).
  uint32_t xcd_map;
  uint32_t xcd_off[8], xcd_len[8];
  uint32_t pad[2];
};

struct PartLaunch {
  const uint8_t* codes;
  const uint8_t* codes_cm;  // column-major copy [features][rows]
  uint64_t rows_total;
  uint32_t row_stride;  // codes per row
  uint32_t code_bytes;
  const uint32_t* rows_in;     // null: positions are the rows (the root)
  uint32_t* rows_out;
  const PartNode* nodes;
  uint32_t n_nodes;
  const uint32_t* chunk_node;  // [n_chunks]
  const uint32_t* block_chunk; // the planner's per-XCD chunk lists (LevelCtl::xcd_map), or null
  uint32_t n_chunks;
  uint32_t* chunk_left;        // [n_chunks]
  uint32_t* bits;              // one bit per row position
  double* pred;                // leaf updates; may be nullptr when no leaves
  double lr;
  uint32_t pred_stride;        // margins row stride (1 for regression)
  uint32_t pred_offset;        // class column
  uint16_t* leaf_node_out;     // non-null: leaves store their leaf id per row instead of updating pred
  uint32_t leaf_bytes;         // 1: leaf ids are one byte each (the array read as uint8_t), else 2
  const uint8_t* gh_in;        // position-ordered gradients moved with the rows (gh_bytes 0: none)
  uint8_t* gh_out;
  uint32_t gh_bytes;           // 0, 1, 2, 4, 8 or 16
  uint32_t no_scatter;         // 1: no kPartSplit node in the launch (rows are not moved)
  // Batched softmax rounds: the class trees' rows fill consecutive segments of class_rows
  // positions each (class k at [k * class_rows, (k + 1) * class_rows)), and a row's leaf id of
  // class k is stored at k * class_rows + row. 0: one tree.
  uint32_t class_rows;
  const LevelCtl* ctl;         // non-null: n_nodes / n_chunks are bounds, the counts are here
  uint32_t* err;
};
constexpr uint32_t kPartChunk = 4096;  // rows per partition chunk (256 threads x 16)


// Device-side tree state of one tree (level-wise growth): the planner of level d reads the
// level's candidates and layer and writes the partition of level d, the builds, slabs and
// scan nodes of level d + 1, the new tree nodes and their leaf values.
// A final leaf of a tree as leaf renewal needs it: its tree node and monotone bounds, by leaf id
// (device-planned trees: leaf ordinal; host-planned trees: tree node id).
struct LeafInfo {
  uint32_t node;
  uint32_t pad;
  double lo, hi;
};

struct PlanLaunch {
  uint32_t depth;              // of the level's nodes
  uint32_t max_depth;
  uint32_t cap;                // maximum leaves
  uint32_t features;
  uint64_t rows_total;
  double lambda, alpha, max_delta;
  const Candidate* cand;       // [n] best split of each node of the level
  const LayerNode* layer;      // [n] the level's nodes (level 0: null, the root)
  uint32_t layer_cap;          // entries of layer / cand
  const LevelCtl* ctl_in;      // count of the level's nodes (n_next of the previous level)
  LevelCtl* ctl;               // this level's counts
  LayerNode* next;             // the next level's nodes
  PartNode* part;              // [n] partition nodes
  uint32_t* chunk_node;        // [chunks]
  uint32_t* block_chunk;       // [chunks] the last level's per-XCD chunk lists
  SlabDesc* slabs;
  uint32_t* slab_first;        // [n_build + 1]
  uint32_t* build_slot;
  int32_t* sibling_slot;
  uint32_t* parent_slot;
  ScanNode* scan;              // the next level's scan nodes
  const uint32_t* bin_count;   // [features]
  const uint32_t* miss_cell;   // [features], 0xFFFFFFFF: no missing cell
  // The fit's tree slots: tree + slot * tree_nodes and tree_ctl + 4 * slot are this tree's, slot
  // read from *tree_slot, which the last level (depth + 1 == max_depth) advances, so every launch
  // of a tree has the same arguments from one tree to the next.
  mpsb_tree_node* tree;
  uint32_t* tree_ctl;          // [0] nodes, [1] leaves, [2] final leaves numbered so far, of each slot
  uint32_t* tree_slot;
  uint32_t tree_nodes;
  double* leaf_table;          // leaf value per leaf ordinal (deferred leaf update)
  LeafInfo* leaf_info;         // tree node and bounds per leaf ordinal (leaf renewal), or null
  uint64_t min_slab_rows, want_slabs, slab_cap;
  unsigned long long* stats;   // [0] rows built, [1] non-empty builds (accumulated)
  // Level 0's nodes: one root (roots 1: every row, root_layer unused) or, for a batched softmax
  // round, one root per class (root_layer[k]: tree node k, its segment of positions); the class
  // trees then share the tree slot as one tree of `roots` roots, numbered level by level.
  uint32_t roots;
  const LayerNode* root_layer;
  uint32_t* err;
};
void launch_plan(const PlanLaunch& a, hipStream_t s);
// phase 0: mark + scan + scatter; 1: mark only (chunk counts); 2: scan + scatter only.
void launch_partition(const PartLaunch& a, hipStream_t s, int phase = 0);

// Gradients: source 0 = a regression objective from (labels, pred, weights); 1 = FP64 pairs;
// 2 = softmax class column of class-major margins [K][rows]; 3 = every class column of a batched
// softmax round (pass 1 only: the FP64 pairs into the [K][rows] stash, the write pass then
// quantizes that stash as source 2 over K x rows).
// Regression objectives (the reference's TrainingParameters::Objective order): 0 squared error,
// 1 binary logistic, 2 quantile (obj_alpha), 3 Poisson, 4 Tweedie (obj_rho).
struct GradLaunch {
  int source;
  int objective;
  double obj_alpha, obj_rho;
  const double* labels;
  const double* pred;
  // Deferred leaf update of the previous tree: pred + lr * leaf_table[leaf_node[row]] first.
  const uint16_t* leaf_node;
  uint32_t leaf_bytes;        // 1: leaf ids are one byte each, else 2
  const double* leaf_table;
  double lr;
  double* pred_out;           // with leaf_node: the updated prediction is written here (pass 2)
  uint32_t pending_class;     // softmax: the margin column the deferred update belongs to
  uint32_t wide_margins;      // softmax: 64-bit addressed margin loads (required beyond 4 GiB of margins)
  const double* weights;
  const double* pairs;
  uint32_t classes, target;
  uint64_t n;
  int precision;
  void* gh;
  double* gh64;               // optional FP64 copy of (g,h) (test hooks)
  // Softmax class columns with maxabs: pass 1 keeps each row's FP64 (g, h) here [rows][2] and
  // pass 2 quantizes them instead of evaluating the softmax a second time.
  double* stash;
  // Regression rows (source 0): pass 2 also keeps each row's FP64 g (unit Hessian) or (g, h)
  // pair here, in row order, for the near-tie resolution and the leaf renewal of the round.
  double* gstash;
  unsigned long long* maxabs; // [2]: g, h as FP64 bit patterns (pass 1); pass 2 derives the scales
  double* scales;             // pass 2 writes [mult_g, mult_h, value_g, value_h] (fixed/integer forms)
  double scale_g, scale_h;    // multipliers when maxabs is null (q = value * scale)
  uint64_t seed;
  uint32_t unit_h;
  uint32_t* err;
};
void launch_grad_maxabs(const GradLaunch& a, hipStream_t s);
void launch_grad_write(const GradLaunch& a, hipStream_t s);

// Exact resolution of ambiguous splits (stochastic-rounding forms, device-planned trees): the
// nodes of a level whose best candidate the rounding noise could have swapped are scanned again,
// over the features that could hold their best split, from exact 64-bit fixed-point histograms
// of the round's FP64 gradients (the FP64_FIXED form's arithmetic); their candidates become the
// exact scan's choice, with the child statistics taken from the level's integer histograms (so
// the next level's sums stay in one arithmetic). Counts on the device; every launch is sized
// for the bound.
struct ResolveLaunch {
  Candidate* cand;                 // the level's candidates, rewritten for resolved nodes
  const uint32_t* contend;         // the level's contender masks [n][mask_words]
  const uint32_t* n_nodes_dyn;     // the level's node count (null: n_nodes)
  uint32_t n_nodes;                // at most 512
  const ScanNode* scan;            // the level's scan nodes
  const LayerNode* layer;          // the level's nodes (segments); null: the root, every row
  uint32_t rows_total;
  uint32_t features, mask_words;
  uint32_t r_max, s_max;           // resolved nodes (<= 64) and row slabs of a level at most
  NodeStats* res_parent;           // [r_max] exact totals (fixed point), zero between levels
  unsigned long long* res_stats;   // [0] ambiguous nodes, [1] resolved
  unsigned long long* res_stats_features;  // [0] contender features of the resolved nodes
  LevelHist res_hist;              // [r_max][cells] exact histograms, zero between levels
  Candidate* res_pf;               // [r_max][features] the exact scan's per-feature candidates
  GradLaunch g;                    // the round's gradients (no pending update; softmax: stash)
  const unsigned long long* maxabs;
  const uint32_t* rows;            // positions -> row ids (null: identity)
  const uint8_t* codes;            // row-major codes [rows][row_stride]
  uint32_t row_stride;
  const uint8_t* codes_cm;         // column-major codes [features][rows]
  uint32_t code_bytes;
  LevelHist hist;                  // the level's integer histograms
  uint32_t cells;
  const uint32_t* cell_offset;
  const uint32_t* bin_count;
  const uint32_t* miss_cell;
  const int8_t* mono;
  ScanParams p;                    // the level's scan parameters
  uint32_t* err;
};
// Contender features histogrammed in LDS at once: (G, H, count) cells, or (G, count) with a
// unit Hessian.
constexpr uint32_t kResolveFeatures = 8, kResolveFeaturesUH = 16;
// Two launches per level (exact histograms; then one block per resolved node: scan, replace,
// clear). Every block plans the level's resolved nodes and row slabs from the scan's flags.
void launch_resolve(const ResolveLaunch& a, hipStream_t s);

// Leaf renewal: once a tree's rows all carry their leaf id, every leaf's value is solved again
// from exact sums of the round's FP64 gradients over its rows (64-bit fixed point with a
// power-of-two scale from the round's max |g|, |h|: exact and order-independent), in place of
// the sums of the form's rounded histogram.
struct RenewLaunch {
  GradLaunch g;                      // the round's gradients (no pending update), softmax: g.stash
  const void* leaf_id;               // per row, leaf_bytes each
  uint32_t leaf_bytes;
  uint32_t n_ids;                    // ids in use are below this bound
  const unsigned long long* maxabs;  // max |g|, |h| of the round as FP64 bit patterns
  unsigned long long* own_maxabs;    // non-null: computed here first (the FP32 form has none)
  long long* sums;                   // [8][copy_stride][2] fixed-point G, H, cleared here
  uint32_t* counts;                  // [8][copy_stride], cleared here
  uint32_t copy_stride;              // ids per copy (>= n_ids); block b adds into copy b mod 8
  const LeafInfo* info;              // [n_ids]
  const uint32_t* tree_ctl;          // device-planned: per-slot counters ([2] leaves numbered)
  const uint32_t* tree_slot;         // device-planned: the next tree's slot (this tree = slot - 1)
  mpsb_tree_node* trees;             // device-planned tree slots, or null
  uint32_t tree_nodes;
  double* node_out;                  // host-planned: renewed value by tree node, or null
  double* leaf_table;                // [n_ids] deferred-update table, rewritten
  double lambda, alpha, max_delta;
  uint32_t* err;
};
constexpr uint32_t kRenewCopies = 8;  // copies of the per-leaf totals (RenewLaunch::sums, counts)
void launch_renew(const RenewLaunch& a, hipStream_t s);

void launch_iota_mod(uint32_t* rows, uint64_t n, uint32_t m, hipStream_t s);  // rows[i] = i mod m
void launch_transpose(const uint8_t* codes, uint32_t row_stride, uint32_t code_bytes, uint32_t F, uint64_t N,
                      uint8_t* cm, hipStream_t s);
void launch_iota(uint32_t* rows, uint32_t n, hipStream_t s);
void launch_vector_add(const float* a, const float* b, float* out, uint64_t n, hipStream_t s);
void launch_fill_f64(double* p, double v, uint64_t n, uint32_t stride, uint32_t offset,
                     hipStream_t s);
// End of a fit: applies the last tree's deferred update (leaf_node non-null) to column `offset`
// of pred and raises kErrPrediction for any non-finite value of that column.
void launch_final_update(double* pred, uint64_t n, uint32_t stride, uint32_t offset, const uint16_t* leaf_node,
                         uint32_t leaf_bytes, const double* leaf_table, double lr, uint32_t* err, hipStream_t s);

struct PredictLaunch {
  const uint8_t* codes;
  uint32_t row_stride;
  uint32_t code_bytes;
  uint64_t rows;
  const mpsb_tree_node* nodes;   // all trees concatenated
  const uint32_t* tree_first;    // [n_trees + 1]
  const uint32_t* miss_code;     // [F]
  uint32_t n_trees;
  uint32_t classes;
  uint32_t rows_per_block;       // rows staged in LDS per block (row pitch = row bytes + 4)
  double lr;
  double* out;                   // [rows][classes], pre-filled with base scores
  uint32_t* err;
};
void launch_predict(const PredictLaunch& a, hipStream_t s);

}  // namespace mpsb

// Host orchestration of the device-resident hot path.
//
// The C++ host keeps the reference's tree-growth control flow (src/core/tree.cpp:49-147): it
// decides, level by level, which nodes split, appends children in BFS order, picks the
// smaller child to histogram and the sibling to subtract, and computes leaf values in FP64.
// Everything proportional to rows runs on the device: gradients, histograms, subtraction,
// the split scan, the stable partition and the prediction update. One host synchronisation
// per tree level reads back the per-node split candidates (a few hundred bytes).

#include "engine.hpp"
#include "host_wait.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <future>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>

namespace mpsb {

constexpr double kInf = std::numeric_limits<double>::infinity();

// ==================================================================== session =========
Session::~Session() {
  if (device >= 0) (void)hipSetDevice(device);
  try {
    evict_codes();
  } catch (...) {
    // A failed copy leaves that dataset without host codes; its reads then fail loudly.
  }
  for (void* p : ptr)
    if (p) (void)hipFree(p);
  if (pinned) (void)hipHostFree(pinned);
  if (sig) (void)hipHostFree(sig);
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void Session::check_fault(int stage, const char* what) {
  if (fault_stage == stage && fault_countdown > 0 && --fault_countdown == 0) {
    fault_stage = 0;
    fail(MPSB_ERR_INJECTED, std::string(what) + ": injected fault");
  }
}

void* Session::get(Buf b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (cap[b] >= bytes) {
    ++reuse;
    return ptr[b];
  }
  check_fault(kFaultAlloc, "allocation");
  if (ptr[b]) {
    MPSB_HIP("stream synchronize", hipStreamSynchronize(stream));
    (void)hipFree(ptr[b]);
    ptr[b] = nullptr;
    cap[b] = 0;
  }
  // Grow with headroom so level-to-level size changes do not reallocate.
  const size_t want = bytes + bytes / 4;
  MPSB_HIP("device allocation", hipMalloc(&ptr[b], want));
  cap[b] = want;
  ++allocs;
  return ptr[b];
}

void Session::release(Buf b) {
  if (!ptr[b]) return;
  MPSB_HIP("stream synchronize", hipStreamSynchronize(stream));
  (void)hipFree(ptr[b]);
  ptr[b] = nullptr;
  cap[b] = 0;
}

void* Session::host_staging(size_t bytes) {
  if (pinned_cap >= bytes) return pinned;
  if (pinned) {
    MPSB_HIP("stream synchronize", hipStreamSynchronize(stream));
    (void)hipHostFree(pinned);
    pinned = nullptr;
  }
  const size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 20);
  MPSB_HIP("pinned allocation", hipHostMalloc(&pinned, want, hipHostMallocDefault));
  pinned_cap = want;
  return pinned;
}

void Session::sync(const char* stage) { MPSB_HIP(stage, hipStreamSynchronize(stream)); }

void Session::ensure_signal(size_t candidates) {
  const size_t bytes = 64 + candidates * sizeof(Candidate);
  if (sig_cap >= bytes) return;
  if (sig) {
    MPSB_HIP("stream synchronize", hipStreamSynchronize(stream));
    (void)hipHostFree(sig);
    sig = sig_dev = nullptr;
    sig_cap = 0;
  }
  const size_t want = std::max<size_t>(bytes * 2, 64 << 10);
  void* p = nullptr;
  MPSB_HIP("pinned allocation", hipHostMalloc(&p, want, hipHostMallocMapped | hipHostMallocCoherent));
  sig = static_cast<uint8_t*>(p);
  sig_cap = want;
  void* dp = nullptr;
  MPSB_HIP("pinned mapping", hipHostGetDevicePointer(&dp, p, 0));
  sig_dev = static_cast<uint8_t*>(dp);
  std::memset(sig, 0, 64);
  __atomic_store_n(reinterpret_cast<uint64_t*>(sig), sig_seq, __ATOMIC_RELEASE);
}

void Session::wait_signal(uint64_t seq, const char* stage) {
  const uint64_t* word = reinterpret_cast<const uint64_t*>(sig);
  hipError_t failed = hipSuccess;
  const WaitResult r = wait_for_sequence(word, seq, wait_limit_s, [&]() -> int {
    // A failed launch or a faulted kernel never publishes: surface the stream's error.
    const hipError_t q = hipStreamQuery(stream);
    if (q != hipSuccess && q != hipErrorNotReady) {
      failed = q;
      return 1;
    }
    // The stream drained without publishing: the signal kernel did not run.
    if (q == hipSuccess && __atomic_load_n(word, __ATOMIC_ACQUIRE) < seq) return 2;
    return 0;
  });
  if (r.published) return;
  if (r.stopped == 1) MPSB_HIP(stage, failed);
  char msg[256];
  if (r.stopped == 2)
    std::snprintf(msg, sizeof msg, "%s: session %llu: level signal %llu was not published (stream idle, last %llu)",
                  stage, (unsigned long long)id, (unsigned long long)seq, (unsigned long long)r.seen);
  else
    std::snprintf(msg, sizeof msg,
                  "%s: session %llu: level signal %llu not published after %.1f s (last %llu); "
                  "MPSB_WAIT_TIMEOUT_S sets the limit",
                  stage, (unsigned long long)id, (unsigned long long)seq, r.seconds, (unsigned long long)r.seen);
  fail(r.stopped == 2 ? MPSB_ERR_INTERNAL : MPSB_ERR_TIMEOUT, msg);
}

hipEvent_t Session::event(size_t i) {
  while (events.size() <= i) {
    // Timing only: no system-scope fence at record time (a fence writes back and invalidates
    // the L2s between the kernels it separates, which costs the stream several microseconds).
    hipEvent_t e;
    MPSB_HIP("event create", hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    events.push_back(e);
  }
  return events[i];
}

template <class T>
static T* dptr(Session& s, Buf b, size_t count) {
  return static_cast<T*>(s.get(b, count * sizeof(T)));
}

static void h2d(Session& s, void* dst, const void* src, size_t bytes, const char* stage) {
  if (bytes) MPSB_HIP(stage, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s.stream));
}
static void d2h(Session& s, void* dst, const void* src, size_t bytes, const char* stage) {
  if (bytes) MPSB_HIP(stage, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s.stream));
}

// Dataset upload: codes once per dataset (cached by identity), plus per-feature tables.
void Session::upload(const HostDataset& d) {
  if (uploaded_uid == d.uid) return;
  check_fault(kFaultUpload, "dataset upload");
  const std::vector<uint8_t>& host = d.host_codes();  // (another session's binning: copied out now)
  evict_codes();
  uploaded_uid = 0;  // until the copy below completes, the code buffer holds no dataset
  const size_t bytes = host.size();
  void* codes = get(kBufCodes, bytes);
  MPSB_HIP("dataset upload", hipMemcpyAsync(codes, host.data(), bytes, hipMemcpyHostToDevice, stream));
  adopt_codes(d);
}

// The resident device-binned dataset, if it still exists and its codes are only here, gets
// its host copy now: the code buffer is about to change.
void Session::evict_codes() {
  std::shared_ptr<DeviceCodes> dc = resident_codes.lock();
  resident_codes.reset();
  if (!dc) return;
  std::lock_guard<std::mutex> lock(dc->m);
  if (dc->session == this) {
    if (!dc->copied) dc->copy_out_locked();
    dc->session = nullptr;
  }
}

void DeviceCodes::copy_out_locked() {
  host.resize(bytes);
  Session& s = *session;
  MPSB_HIP("set device", hipSetDevice(s.device));
  MPSB_HIP("code read-back",
           hipMemcpyAsync(host.data(), s.ptr[kBufCodes], bytes, hipMemcpyDeviceToHost, s.stream));
  s.sync("code read-back");
  copied = true;
}

const std::vector<uint8_t>& DeviceCodes::fetch() {
  std::lock_guard<std::mutex> lock(m);
  if (!copied) {
    if (!session) fail(MPSB_ERR_INTERNAL, "device-binned codes are no longer on the device");
    copy_out_locked();
  }
  return host;
}

// The code buffer holds d's row-major codes: derive the column-major copy and the
// per-feature tables, then remember d as the resident dataset.
void Session::adopt_codes(const HostDataset& d) {
  void* codes = ptr[kBufCodes];
  void* cm = get(kBufCodesCM, size_t(d.rows) * d.features * d.code_bytes);
  launch_transpose(static_cast<const uint8_t*>(codes), d.row_stride, d.code_bytes, d.features, d.rows,
                   static_cast<uint8_t*>(cm), stream);
  std::vector<uint32_t> miss_cell(d.features), miss_code(d.features);
  for (uint32_t f = 0; f < d.features; ++f) {
    miss_cell[f] = d.has_missing(f) ? d.cell_offset[f] + d.bin_count[f] : 0xFFFFFFFFu;
    miss_code[f] = d.has_missing(f) ? d.bin_count[f] : 0xFFFFFFFFu;
  }
  h2d(*this, get(kBufCellOffset, (d.features + 1) * 4), d.cell_offset.data(), (d.features + 1) * 4, "metadata upload");
  h2d(*this, get(kBufBinCount, d.features * 4), d.bin_count.data(), d.features * 4, "metadata upload");
  h2d(*this, get(kBufMissCell, d.features * 4), miss_cell.data(), d.features * 4, "metadata upload");
  h2d(*this, get(kBufMissCode, d.features * 4), miss_code.data(), d.features * 4, "metadata upload");
  sync("dataset upload");
  uploaded_uid = d.uid;
}

// =============================================================== objective math ======
// src/core/objective.cpp:22-36,75-97,303-318 (host copies used for leaf values).
static double validated_denominator(double h, double lambda) {
  if (!std::isfinite(h)) fail(MPSB_ERR_TRAINING, "Hessian sum must be finite");
  if (!std::isfinite(lambda)) fail(MPSB_ERR_TRAINING, "reg_lambda must be finite");
  if (h < 0.0) fail(MPSB_ERR_TRAINING, "Hessian sum cannot be negative");
  if (lambda < 0.0) fail(MPSB_ERR_TRAINING, "reg_lambda cannot be negative");
  const double d = h + lambda;
  if (!std::isfinite(d) || d <= 0.0) fail(MPSB_ERR_TRAINING, "Hessian sum plus reg_lambda must be finite and positive");
  return d;
}

double host_leaf_weight(double g, double h, double lambda, double alpha, double max_delta) {
  if (!std::isfinite(g)) fail(MPSB_ERR_TRAINING, "Gradient sum must be finite");
  const double m = std::fabs(g) - alpha;
  const double s = m <= 0.0 ? 0.0 : std::copysign(m, g);
  double w = -s / validated_denominator(h, lambda);
  if (max_delta != 0.0) w = std::max(-max_delta, std::min(max_delta, w));
  if (!std::isfinite(w)) fail(MPSB_ERR_TRAINING, "Leaf value overflowed");
  return w;
}

// ================================================================ parameters ==========
Params params_from(const mpsb_train_params& p) {
  Params q;
  q.n_estimators = p.n_estimators;
  q.lr = p.learning_rate;
  q.max_bins = p.max_bins;
  q.objective = p.objective;
  q.precision = int(p.precision);
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
  if (p.split_strategy > 1) fail(MPSB_ERR_INVALID_ARGUMENT, "unknown split strategy");
  q.random_threshold = p.split_strategy == 1;
  q.growth = p.growth_strategy;
  q.seed = p.random_seed;
  if (p.n_monotonic) {
    if (!p.monotonic) fail(MPSB_ERR_INVALID_ARGUMENT, "monotonic constraints pointer is null");
    q.monotonic.assign(p.monotonic, p.monotonic + p.n_monotonic);
  }
  size_t at = 0;
  for (uint32_t i = 0; i < p.n_groups; ++i) {
    if (!p.group_sizes || !p.group_features) fail(MPSB_ERR_INVALID_ARGUMENT, "interaction constraint arrays are null");
    q.groups.emplace_back(p.group_features + at, p.group_features + at + p.group_sizes[i]);
    at += p.group_sizes[i];
  }
  q.rounding_seed = p.rounding_seed;
  q.objective_alpha = p.objective_alpha;
  q.tweedie_power = p.tweedie_variance_power;
  if (p.leaf_solve > 2) fail(MPSB_ERR_INVALID_ARGUMENT, "leaf_solve must be 0 (automatic), 1 (histograms) or 2 (renewed)");
  q.leaf_solve = p.leaf_solve;
  if (p.near_ties > 2) fail(MPSB_ERR_INVALID_ARGUMENT, "near_ties must be 0 (auto), 1 (resolve) or 2 (keep)");
  q.near_ties = p.near_ties;
  if (p.class_rounds > 1) fail(MPSB_ERR_INVALID_ARGUMENT, "class_rounds must be 0 (class order) or 1 (batched rounds)");
  q.class_rounds = p.class_rounds;
  return q;
}

// ValidateTrainingParameters (src/core/trainer.cpp:22-42), ValidateParameters
// (src/core/tree_structure.cpp:16-51) and the constraint checks of TrainSingleRegressionTree
// (src/core/tree.cpp:156-166), with the reference's messages.
void validate_train_params(const Params& p, const HostDataset& d) {
  if (p.n_estimators == 0) fail(MPSB_ERR_TRAINING, "n_estimators must be at least 1");
  if (!std::isfinite(p.lr) || p.lr <= 0.0 || p.lr > 1.0) fail(MPSB_ERR_TRAINING, "learning_rate must be in (0, 1]");
  if (p.max_bins < 2 || p.max_bins > 65536) fail(MPSB_ERR_TRAINING, "max_bins must be in [2, 65536]");
  if (!std::isfinite(p.objective_alpha) || p.objective_alpha <= 0.0 || p.objective_alpha >= 1.0)
    fail(MPSB_ERR_TRAINING, "quantile alpha must be in (0, 1)");
  if (!std::isfinite(p.tweedie_power) || p.tweedie_power <= 1.0 || p.tweedie_power >= 2.0)
    fail(MPSB_ERR_TRAINING, "tweedie variance power must be in (1, 2)");
  if (d.max_bins != p.max_bins) fail(MPSB_ERR_TRAINING, "Training data, labels, or max_bins contract does not match");
  if (p.min_samples_leaf == 0) fail(MPSB_ERR_TRAINING, "min_samples_leaf must be at least 1");
  if (p.max_leaves == 1) fail(MPSB_ERR_TRAINING, "max_leaves must be zero or at least 2");
  if (p.max_active_leaves == 1) fail(MPSB_ERR_TRAINING, "max_active_leaves must be zero or at least 2");
  if (p.max_active_leaves && p.max_leaves && p.max_active_leaves > p.max_leaves)
    fail(MPSB_ERR_TRAINING, "max_active_leaves must not exceed max_leaves");
  auto nonneg = [](double x, const char* msg) {
    if (!std::isfinite(x) || x < 0.0) fail(MPSB_ERR_TRAINING, msg);
  };
  nonneg(p.min_child_weight, "min_child_weight must be finite and non-negative");
  nonneg(p.lambda, "reg_lambda must be finite and non-negative");
  nonneg(p.alpha, "reg_alpha must be finite and non-negative");
  nonneg(p.max_delta, "max_delta_step must be finite and non-negative");
  nonneg(p.gamma, "gamma must be finite and non-negative");
  nonneg(p.min_gain, "min_gain_to_split must be finite and non-negative");
  if (!p.monotonic.empty() && p.monotonic.size() != d.features)
    fail(MPSB_ERR_TRAINING, "monotonic_constraints length must match feature count");
  for (int8_t c : p.monotonic)
    if (c < -1 || c > 1) fail(MPSB_ERR_TRAINING, "monotonic_constraints values must be -1, 0, or 1");
  for (const auto& g : p.groups) {
    if (g.empty()) fail(MPSB_ERR_TRAINING, "interaction constraint groups must be non-empty");
    for (uint32_t f : g)
      if (f >= d.features) fail(MPSB_ERR_TRAINING, "interaction constraint feature index out of range");
  }
  if (p.growth != 0) fail(MPSB_ERR_INVALID_ARGUMENT, "the device trainer grows trees level-wise only");
  if (p.precision < MPSB_PRECISION_FP64 || p.precision > MPSB_PRECISION_FP64_FIXED)
    fail(MPSB_ERR_INVALID_ARGUMENT, "unknown histogram precision");
}

static uint32_t max_leaves_of(const Params& p) {  // src/core/tree_structure.cpp:54-62
  if (p.max_leaves) return p.max_leaves;
  if (p.max_depth >= 31) return std::numeric_limits<uint32_t>::max();
  return 1u << p.max_depth;
}

// ==================================================================== fit state =======
int hist_kind(int precision) {
  switch (precision) {
    case MPSB_PRECISION_FP32: return kHistF32;
    case MPSB_PRECISION_FP64: return kHistSeq;
    case MPSB_PRECISION_FP64_FIXED: return kHistFix64;
    case MPSB_PRECISION_INT16: return kHistQ | (2 << 8);
    default: return kHistQ | (1 << 8);
  }
}

uint32_t bytes_per_cell(int kind, bool unit_h) {
  switch (kind & 0xFF) {
    case kHistF32: return unit_h ? 8 : 12;
    case kHistFix64:
    case kHistSeq: return unit_h ? 12 : 20;
    default: return unit_h ? 8 : 12;
  }
}

uint32_t gh_bytes(int precision, bool unit_h) {
  switch (precision) {
    case MPSB_PRECISION_FP32: return unit_h ? 4 : 8;
    case MPSB_PRECISION_FP64: return 16;
    case MPSB_PRECISION_FP64_FIXED: return unit_h ? 8 : 16;
    case MPSB_PRECISION_INT16: return unit_h ? 2 : 4;
    default: return unit_h ? 1 : 2;
  }
}

// Rows per slab cap: the packed integer accumulators must not overflow (see kernels.hip).
uint32_t slab_cap(int precision, bool unit_h) {
  if (precision == MPSB_PRECISION_INT16 && !unit_h) return 65535;
  if ((precision == MPSB_PRECISION_INT16 || precision == MPSB_PRECISION_INT8) && unit_h) return (1u << 24) - 1;
  return 0xFFFFFFFFu;
}


// Rows per slab of a gradient-only (kHistG32) histogram: a slab's int32 cell sums stay in range
// (32,767 x 65,535 and 127 x (2^24 - 1) are below 2^31).
uint32_t slab_cap_g32(int precision) { return precision == MPSB_PRECISION_INT16 ? 65535u : (1u << 24) - 1; }

// MPSB_HIST_INTERLEAVE=0 keeps the packed layout (for comparisons).
static bool interleave_enabled() {
  const char* v = std::getenv("MPSB_HIST_INTERLEAVE");
  return !(v && v[0] == '0');
}

// Feature-interleaved tiles for one-byte codes: a tile is 8 words of the row (32 features) and
// stores the cell (bin b, feature slot j) at LDS cell b * 32 + j, so every feature owns one
// 8-byte bank pair. The kernel rotates, lane by lane, which of its features a lane updates
// first: the 32 lanes of a half-wave then update 32 different features in each atomic
// instruction and never meet in a bank, whatever the bins. Forms whose cells are wider than
// 8 bytes (a [G|H] word and a count) do not fit 32 features of 256 bins in the budget; they
// take 16-feature tiles (4 words, cell b * 16 + j, one thread per row). Needs slots x (largest
// feature's cells) cells per tile within the budget; returns false (layout not usable)
// otherwise.
static bool plan_interleaved(const HostDataset& d, uint32_t bpc, uint32_t budget, uint32_t want_wpt,
                             TilePlan& plan) {
  if (d.code_bytes != 1) return false;
  const uint32_t max_cells = budget / bpc;  // including the trash cell
  uint32_t cmax = 1;
  for (uint32_t f = 0; f < d.features; ++f) cmax = std::max(cmax, d.cell_offset[f + 1] - d.cell_offset[f]);
  if (cmax > 256) return false;
  // 64 slots (a 128 KB tile of 8-byte cells: one block per CU, a row's 64-byte half per block)
  // only when the budget asks for that much; 32 within the default 64 KB; 16 for wider cells.
  uint32_t slots = 64;
  while (slots > 16 && slots * cmax + 1 > max_cells) slots /= 2;
  if (slots * cmax + 1 > max_cells) return false;
  const uint32_t tile_words = slots / 4;
  const uint32_t row_words = d.row_stride / 4;  // a multiple of 4
  const uint32_t words = std::min(row_words, ((d.features + 3) / 4 + 3) / 4 * 4);
  plan.wpt = slots == 32 ? want_wpt : 4u;  // 16 and 64 slots: four words per thread
  plan.interleaved = true;
  plan.slots = slots;
  plan.cell_lds.assign(d.cells, 0);
  for (uint32_t w = 0; w < words; w += tile_words) {
    TileDesc t{};
    t.word0 = w;
    t.words = std::min(tile_words, words - w);
    const uint32_t f0 = std::min(d.features, w * 4), f1 = std::min(d.features, (w + t.words) * 4);
    uint32_t c_tile = 1;
    for (uint32_t f = f0; f < f1; ++f) {
      const uint32_t c = d.cell_offset[f + 1] - d.cell_offset[f];
      c_tile = std::max(c_tile, c);
      for (uint32_t b = 0; b < c; ++b) plan.cell_lds[d.cell_offset[f] + b] = b * slots + (f - f0);
    }
    t.cell0 = d.cell_offset[f0];
    t.ncells = d.cell_offset[f1] - d.cell_offset[f0];
    t.lds_cells = slots * c_tile;
    plan.lds_bytes = std::max(plan.lds_bytes, (t.lds_cells + 1) * bpc);
    plan.tiles.push_back(t);
  }
  plan.word_off.assign(plan.tiles.size() * 64 * 4, 0);  // unused by interleaved tiles
  return true;
}

// Tiles of the gradient-only root histogram (kHistG32): 4-byte cells, so 64 one-byte features
// (16 words of the row, 64 B) fill a 64 KB tile at cell b * 64 + j, one LDS bank per feature,
// and a row is read in half as many tiles as with 8-byte cells. Four threads per row load four
// words each. False when the layout does not apply (wide codes, features of over 256 cells, a
// budget below 64 slots x the widest feature).
bool plan_g32_tiles(const HostDataset& d, uint32_t budget, TilePlan& plan) {
  plan = TilePlan{};
  if (d.code_bytes != 1) return false;
  uint32_t cmax = 1;
  for (uint32_t f = 0; f < d.features; ++f) cmax = std::max(cmax, d.cell_offset[f + 1] - d.cell_offset[f]);
  constexpr uint32_t slots = 64, tile_words = 16;
  if (cmax > 256 || (slots * cmax + 1) * 4 > budget) return false;
  const uint32_t row_words = d.row_stride / 4;  // a multiple of 4
  const uint32_t words = std::min(row_words, ((d.features + 3) / 4 + 3) / 4 * 4);
  plan.wpt = 4;
  plan.interleaved = true;
  plan.slots = slots;
  plan.cell_lds.assign(d.cells, 0);
  for (uint32_t w = 0; w < words; w += tile_words) {
    TileDesc t{};
    t.word0 = w;
    t.words = std::min(tile_words, words - w);
    const uint32_t f0 = std::min(d.features, w * 4), f1 = std::min(d.features, (w + t.words) * 4);
    uint32_t c_tile = 1;
    for (uint32_t f = f0; f < f1; ++f) {
      const uint32_t c = d.cell_offset[f + 1] - d.cell_offset[f];
      c_tile = std::max(c_tile, c);
      for (uint32_t b = 0; b < c; ++b) plan.cell_lds[d.cell_offset[f] + b] = b * slots + (f - f0);
    }
    t.cell0 = d.cell_offset[f0];
    t.ncells = d.cell_offset[f1] - d.cell_offset[f0];
    t.lds_cells = slots * c_tile;
    plan.lds_bytes = std::max(plan.lds_bytes, (t.lds_cells + 1) * 4);
    plan.tiles.push_back(t);
  }
  plan.word_off.assign(plan.tiles.size() * 64 * 4, 0);  // unused by interleaved tiles
  return true;
}

// Feature tiles: runs of whole 32-bit words of the row whose cells fit the LDS budget.
// Tiles are runs of 16-byte row granules (four 32-bit words) so a thread can load up to four
// words at once; a granule that alone overflows the budget falls back to single words.
TilePlan plan_tiles(const HostDataset& d, uint32_t bpc, uint32_t budget, uint32_t block, uint32_t want_wpt) {
  TilePlan plan;
  if (interleave_enabled() && plan_interleaved(d, bpc, budget, want_wpt, plan)) return plan;
  plan = TilePlan{};
  const uint32_t cpw = 4 / d.code_bytes;
  const uint32_t row_words = d.row_stride * d.code_bytes / 4;  // a multiple of 4
  const uint32_t words = std::min(row_words, ((d.features + cpw - 1) / cpw + 3) / 4 * 4);
  const uint32_t max_cells = budget / bpc;  // including the trash cell
  auto cells_of = [&](uint32_t w0, uint32_t w1) {
    const uint32_t f0 = std::min(d.features, w0 * cpw), f1 = std::min(d.features, w1 * cpw);
    return d.cell_offset[f1] - d.cell_offset[f0];
  };
  uint32_t gran = 4;
  for (uint32_t w = 0; w < words; w += 4)
    if (cells_of(w, w + 4) + 1 > max_cells) gran = 1;
  plan.wpt = gran == 4 ? want_wpt : 1;
  uint32_t w = 0;
  while (w < words) {
    uint32_t e = w + gran;
    if (cells_of(w, e) + 1 > max_cells)
      fail(MPSB_ERR_INVALID_ARGUMENT, "a feature's histogram does not fit the device LDS budget");
    while (e + gran <= words && e - w + gran <= 64 && (e - w + gran) / plan.wpt <= block &&
           cells_of(w, e + gran) + 1 <= max_cells)
      e += gran;
    TileDesc t{};
    t.word0 = w;
    t.words = e - w;
    t.cell0 = d.cell_offset[std::min(d.features, w * cpw)];
    t.ncells = cells_of(w, e);
    plan.tiles.push_back(t);
    w = e;
  }
  // Spare LDS of a tile goes to features of at most 16 cells: 16 replicas each (256 cells),
  // in feature order while the tile stays within the budget. The tile count never changes.
  plan.word_off.assign(plan.tiles.size() * 64 * 4, 0);
  plan.cell_lds.assign(d.cells, 0);
  for (size_t ti = 0; ti < plan.tiles.size(); ++ti) {
    TileDesc& t = plan.tiles[ti];
    const uint32_t f0 = std::min(d.features, t.word0 * cpw), f1 = std::min(d.features, (t.word0 + t.words) * cpw);
    uint32_t lds = t.ncells;
    std::vector<uint8_t> replicated(f1 - f0, 0);
    for (uint32_t f = f0; f < f1; ++f) {
      const uint32_t c = d.cell_offset[f + 1] - d.cell_offset[f];
      if (c <= 16 && lds + (256 - c) + 1 <= max_cells) {
        replicated[f - f0] = 1;
        lds += 256 - c;
      }
    }
    uint32_t at = 0;
    std::vector<uint32_t> lds_off(f1 - f0);
    for (uint32_t f = f0; f < f1; ++f) {
      const uint32_t c = d.cell_offset[f + 1] - d.cell_offset[f];
      const uint32_t flag = replicated[f - f0] ? kReplicated : 0u;
      lds_off[f - f0] = at | flag;
      for (uint32_t b = 0; b < c; ++b) plan.cell_lds[d.cell_offset[f] + b] = (at + b) | flag;
      at += replicated[f - f0] ? 256 : c;
    }
    t.lds_cells = at;
    plan.lds_bytes = std::max(plan.lds_bytes, (t.lds_cells + 1) * bpc);
    for (uint32_t k = 0; k < 64 * 4; ++k) plan.word_off[ti * 256 + k] = t.lds_cells;  // trash
    for (uint32_t wi = 0; wi < t.words; ++wi)
      for (uint32_t k = 0; k < cpw; ++k) {
        const uint32_t f = (t.word0 + wi) * cpw + k;
        if (f < d.features) plan.word_off[ti * 256 + wi * 4 + k] = lds_off[f - f0];
      }
  }
  return plan;
}

// Reads the error word of a launch sequence and raises the reference's TrainingError for the
// earliest failing stage. The reference's messages: src/core/trainer.cpp:200-266,
// src/core/objective.cpp:63-287, src/core/multiclass_trainer.cpp:129-146,
// src/core/tree_split.cpp:167-181, src/core/tree_histogram.cpp:16-66.
// objective: the regression objective's index, -1 for softmax class columns.
void raise_device_errors(uint32_t flags, uint32_t first, int objective) {
  if (!flags) return;
  // The earliest failing stage reports (as the host stops at it), its bits in the order below.
  constexpr uint32_t kGradBits = kErrPrediction | kErrExpFirst | kErrExpSecond | kErrObjective | kErrWeighted |
                                 kErrNonFinite;
  constexpr uint32_t kPlanBits = kErrBounds | kErrLeafGradient | kErrLeafHessian | kErrLeafNegative |
                                 kErrLeafDenominator | kErrLeafOverflow | kErrSplitCount;
  if (first) flags &= (first & kGradBits) ? kGradBits : (first & kPlanBits) ? kPlanBits : ~(kGradBits | kPlanBits);
  static const char* const kObjective[] = {"Squared-error gradient overflowed",
                                           "Binary logistic gradient/Hessian overflowed", "",
                                           "poisson gradient/Hessian overflowed",
                                           "tweedie gradient/Hessian overflowed"};
  const int obj = objective;
  if (flags & kErrPrediction)
    fail(MPSB_ERR_TRAINING, obj < 0 ? "softmax margin update overflowed" : "Boosting prediction update overflowed");
  if (flags & kErrExpFirst)
    fail(MPSB_ERR_TRAINING, obj == 3 ? "poisson log mean exponential overflowed" : "tweedie first term exponential overflowed");
  if (flags & kErrExpSecond) fail(MPSB_ERR_TRAINING, "tweedie second term exponential overflowed");
  if (flags & kErrObjective)
    fail(MPSB_ERR_TRAINING, obj < 0 ? "softmax gradient/Hessian overflowed" : kObjective[obj]);
  if (flags & kErrWeighted) fail(MPSB_ERR_TRAINING, "weighted gradient/Hessian overflowed");
  if (flags & kErrNonFinite) fail(MPSB_ERR_TRAINING, "Gradient/Hessian must be finite and Hessian non-negative");
  if (flags & kErrCount) fail(MPSB_ERR_TRAINING, "Histogram sample count does not match node statistics");
  if (flags & kErrSubtract) fail(MPSB_ERR_TRAINING, "Histogram subtraction child exceeded parent");
  if (flags & kErrMissing) fail(MPSB_ERR_TRAINING, "missing statistics exceed parent statistics");
  if (flags & kErrOverflow) fail(MPSB_ERR_TRAINING, "Split gain or node score overflowed");
  if (flags & kErrPrefix) fail(MPSB_ERR_TRAINING, "non-missing histogram prefix exceeded parent");
  // Level planner (the host's make_leaf / LeafWeight / count checks, same messages).
  if (flags & kErrBounds) fail(MPSB_ERR_TRAINING, "monotonic leaf bounds are inconsistent");
  if (flags & kErrLeafGradient) fail(MPSB_ERR_TRAINING, "Gradient sum must be finite");
  if (flags & kErrLeafHessian) fail(MPSB_ERR_TRAINING, "Hessian sum must be finite");
  if (flags & kErrLeafNegative) fail(MPSB_ERR_TRAINING, "Hessian sum cannot be negative");
  if (flags & kErrLeafDenominator)
    fail(MPSB_ERR_TRAINING, "Hessian sum plus reg_lambda must be finite and positive");
  if (flags & kErrLeafOverflow) fail(MPSB_ERR_TRAINING, "Leaf value overflowed");
  if (flags & kErrSplitCount) fail(MPSB_ERR_TRAINING, "Sample partition counts do not match histogram statistics");
  fail(MPSB_ERR_TRAINING, "prediction found a malformed tree");
}

struct Timer {
  Session& s;
  struct Mark {
    int stage;
    size_t a, b;
  };
  std::vector<Mark> marks;
  size_t next = 0;
  double* acc[8] = {};
  // Each event boundary costs the stream a few microseconds between kernels, so a fit times
  // its stages on a sample of its trees; `on` says whether the current tree is timed.
  bool on = true;
  static constexpr size_t kOff = ~size_t(0);
  explicit Timer(Session& ss) : s(ss) {}
  size_t begin() {
    if (!on) return kOff;
    const size_t i = next++;
    MPSB_HIP("event record", hipEventRecord(s.event(i), s.stream));
    return i;
  }
  void end(int stage, size_t a) {
    if (a == kOff) return;
    const size_t i = next++;
    MPSB_HIP("event record", hipEventRecord(s.event(i), s.stream));
    marks.push_back({stage, a, i});
  }
  // Called once the recorded work is known complete (a synchronisation or a level signal);
  // an event the runtime has not yet seen complete is waited for, never dropped.
  void resolve() {
    for (const Mark& m : marks) {
      float ms = 0.0f;
      hipError_t e = hipEventElapsedTime(&ms, s.event(m.a), s.event(m.b));
      if (e == hipErrorNotReady) {
        MPSB_HIP("event wait", hipEventSynchronize(s.event(m.b)));
        e = hipEventElapsedTime(&ms, s.event(m.a), s.event(m.b));
      }
      MPSB_HIP("event timing", e);
      if (acc[m.stage]) *acc[m.stage] += ms * 1e-3;
    }
    marks.clear();
    next = 0;
  }
};

constexpr uint32_t kTimeEvery = 16;  // stage timing on every 16th tree of a fit

enum TimerStage { kTGrad = 0, kTHist = 1, kTReduce = 2, kTScan = 3, kTPart = 4, kTRenew = 5 };

// Bump allocator for per-sync-epoch descriptors: tables are written into pinned staging and
// reach the device in one async copy per launch group (flush), not one per table. Reset
// after each stream synchronisation, when every earlier copy has completed.
struct DescArena {
  Session& s;
  uint8_t* host = nullptr;
  uint8_t* dev = nullptr;
  size_t cap = 0, used = 0, flushed = 0;
  explicit DescArena(Session& ss, size_t bytes) : s(ss) {
    host = static_cast<uint8_t*>(s.host_staging(bytes));
    dev = static_cast<uint8_t*>(s.get(kBufDesc, bytes));
    cap = bytes;
  }
  template <class T>
  T* put(const T* src, size_t n) {
    const size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
    if (used + bytes > cap) fail(MPSB_ERR_INTERNAL, "descriptor arena exhausted");
    std::memcpy(host + used, src, n * sizeof(T));
    T* d = reinterpret_cast<T*>(dev + used);
    used += bytes;
    return d;
  }
  void flush() {
    if (used > flushed)
      MPSB_HIP("descriptor upload",
               hipMemcpyAsync(dev + flushed, host + flushed, used - flushed, hipMemcpyHostToDevice, s.stream));
    flushed = used;
  }
  void reset() { used = flushed = 0; }
};

// Fits alive in this process (all sessions), for the planning choice of small fits.
static std::atomic<int> active_fits{0};
constexpr uint64_t kDevicePlanMinRowsShared = uint64_t(1) << 20;

struct Fit {
  // Counts this fit among the live ones for as long as it exists.
  struct Live {
    Live() { active_fits.fetch_add(1); }
    ~Live() { active_fits.fetch_sub(1); }
  } live;
  Session& s;
  const HostDataset& d;
  const Params& p;
  int precision;
  int kind;
  bool unit_h;
  bool softmax = false;  // gradients are softmax class columns (error messages)
  // Batched softmax rounds: the round's class trees grow as one device-planned tree of `roots`
  // roots, class k's rows at positions [k * rows, (k + 1) * rows). 1: one tree.
  uint32_t roots = 1;
  uint64_t positions() const { return uint64_t(d.rows) * roots; }
  TilePlan tiles;
  TileDesc* d_tiles = nullptr;
  uint32_t* d_word_off = nullptr;
  uint32_t* d_cell_lds = nullptr;
  // Root histograms without counts (kHistG32, kHistQNC, kHistF64NC): their kind and tiles, and
  // the root's per-cell counts taken from the fit's first root histogram.
  TilePlan tiles_g;
  TileDesc* d_tiles_g = nullptr;
  uint32_t* d_cell_lds_g = nullptr;
  uint32_t* d_word_off_g = nullptr;
  uint32_t* root_count = nullptr;
  bool root_free = false;  // the form builds its roots without counts after the first
  int root_kind = 0;  // HistKind of those roots
  uint32_t root_cell_bytes = 0;  // bytes per cell of their slab partials
  bool root_counts_ready = false;
  uint32_t* err = nullptr;
  uint32_t* host_err = nullptr;
  double scale_g = 1.0, scale_h = 1.0;  // value = q * scale (integer forms)
  LevelHist hist[2]{};
  uint32_t hist_slots[2] = {0, 0};
  uint64_t seed = 0;  // stochastic-rounding stream of the current round
  bool sequential_scan = false;  // the reference's serial FP64 prefix (row-order form)
  bool seq = false;              // row-order FP64 form
  uint32_t* d_seq_feat = nullptr;  // all features, then the features with a missing cell
  uint32_t n_miss_feat = 0;
  uint32_t seq_fg = 64, seq_cpf = 0;
  NodeStats* d_node_stats = nullptr;  // parent statistics computed on the device
  double* d_scales = nullptr;         // [mult_g, mult_h, value_g, value_h] of the round
  ScanLaunch last_scan{};             // the most recently planned split scan
  // Gradients the tiled histogram reads: position-ordered while a tree grows (they move with
  // the rows through each partition), row-indexed (the kBufGH buffer) otherwise.
  const void* gh_cur = nullptr;
  bool gh_pos = false;
  // Regression: leaves store lr * leaf per row here; the next gradient pass applies it.
  // Deferred leaf update (regression): the last tree's leaf node id per row and its table of
  // node leaf values (in the descriptor arena, which is only rewritten after the next level's
  // signal, i.e. after the gradient pass that reads the table has completed).
  uint16_t* leaf_node = nullptr;
  const double* leaf_table = nullptr;
  // Leaf renewal (mpsb_train_params::leaf_solve): after each tree, its leaves are solved again
  // from exact fixed-point sums of the round's FP64 gradients over their rows.
  bool renew = false;
  bool renew_softmax = false;      // gradients of softmax class columns (the max pass's stash)
  uint32_t renew_ids = 0;          // bound of leaf ids over the fit's trees
  long long* renew_sums = nullptr;
  uint32_t* renew_counts = nullptr;
  unsigned long long* renew_max = nullptr;  // own max |g|, |h| (forms without a scaled pass)
  LeafInfo* leaf_info = nullptr;   // host-planned tree: node id -> bounds (descriptor arena)
  uint32_t leaf_ids = 0;           // host-planned tree: its node count
  double* renew_nodes = nullptr;   // host-planned trees: [tree][renew_ids] renewed values
  std::vector<size_t> renewed_host;  // model tree index of each host-planned renewed tree
  bool last_tree_device = false;
  // Near-tie resolution (mpsb_train_params::near_ties): the stochastic-rounding forms' scans
  // bound each gain's rounding noise (noise_k) and mark ambiguous nodes; device-planned trees
  // scan those again exactly (ResolveLaunch).
  double noise_k = 0.0;
  bool resolve = false;
  // Resolved nodes and row slabs of a level at most; levels of more nodes are not resolved.
  static constexpr uint32_t kResolveNodes = 64, kResolveSlabs = 512, kResolveLevelNodes = 512;
  GradLaunch round_g{};            // the round's gradient launch (resolution and renewal read it)
  struct Resolve {
    NodeStats* parent = nullptr;
    unsigned long long* stats = nullptr;
    Candidate* per_feature = nullptr;
    uint32_t* contend = nullptr;   // the level scans' contender masks
    LevelHist hist{};
  } rs;
  Timer timer;
  DescArena arena;
  mpsb_fit_stats stats{};
  int8_t* d_mono = nullptr;

  Fit(Session& ss, const HostDataset& dd, const Params& pp, int prec, bool uh)
      : s(ss), d(dd), p(pp), precision(prec), kind(hist_kind(prec)), unit_h(uh), seed(pp.rounding_seed),
        timer(ss), arena(ss, size_t(32) << 20) {
    s.upload(d);
    d_node_stats = dptr<NodeStats>(s, kBufNodeStats, 1);
    d_scales = dptr<double>(s, kBufScales, 4);
    if (kind == kHistSeq) {
      seq = true;
      sequential_scan = true;
      std::vector<uint32_t> fl(d.features);
      uint32_t cpf = 1;
      for (uint32_t ft = 0; ft < d.features; ++ft) {
        fl[ft] = ft;
        cpf = std::max(cpf, d.cell_offset[ft + 1] - d.cell_offset[ft]);
      }
      for (uint32_t ft = 0; ft < d.features; ++ft)
        if (d.has_missing(ft)) fl.push_back(ft);
      n_miss_feat = uint32_t(fl.size()) - d.features;
      d_seq_feat = dptr<uint32_t>(s, kBufSeqFeat, fl.size());
      h2d(s, d_seq_feat, fl.data(), fl.size() * 4, "feature list upload");
      // Per-lane LDS cells up to a 64 KB block; wider features accumulate in the output.
      const uint32_t lane_bytes = cpf * (unit_h ? 12u : 20u);
      const uint32_t fit = std::min<uint32_t>(std::min(s.lds_max, 65536u) / lane_bytes, 64u);
      seq_fg = fit ? fit : 64u;
      seq_cpf = fit ? cpf : 0u;
    }
    if (!seq) {
      // Tuned tile budget first; features too wide for it get the whole LDS of a block.
      uint32_t budget = std::min(s.lds_max, s.tune.lds_budget);
      // 12-byte integer cells ([G|H] word and count, or fixed-point G and count) of fits of at
      // least 8,000,000 histogram positions per tree (a batched softmax round has K >= 2 blocks of rows)
      // take 32-slot tiles of 96 KB, one block per CU, instead of 16-slot tiles of 48 KB, unless
      // the tuning sets a budget. Measured on MI355X: batched softmax rounds at 4M x 64, K = 10,
      // 7.7 % faster and FP64_FIXED at 10M x 128 3.5 % faster; class-order softmax trees of 4M
      // rows ran 18 % slower with them.
      const uint64_t positions_hint = d.rows * (p.class_rounds == 1 ? 2ull : 1ull);
      if (s.tune.lds_budget == Tuning{}.lds_budget && bytes_per_cell(kind, unit_h) == 12 &&
          (kind & 0xFF) != kHistF32 && positions_hint >= 8000000ull)
        budget = std::min<uint32_t>(s.lds_max, (32 * 256 + 1) * 12);
      try {
        tiles = plan_tiles(d, bytes_per_cell(kind, unit_h), budget, s.tune.block, s.tune.wpt);
      } catch (const Error&) {
        tiles = plan_tiles(d, bytes_per_cell(kind, unit_h), s.lds_max, s.tune.block, s.tune.wpt);
      }
      d_tiles = dptr<TileDesc>(s, kBufTiles, tiles.tiles.size());
      d_word_off = dptr<uint32_t>(s, kBufWordOff, tiles.word_off.size());
      d_cell_lds = dptr<uint32_t>(s, kBufCellLds, tiles.cell_lds.size());
      h2d(s, d_tiles, tiles.tiles.data(), tiles.tiles.size() * sizeof(TileDesc), "tile upload");
      h2d(s, d_word_off, tiles.word_off.data(), tiles.word_off.size() * 4, "tile upload");
      h2d(s, d_cell_lds, tiles.cell_lds.data(), tiles.cell_lds.size() * 4, "tile upload");
    }
    // Roots without counts (a root's per-cell counts are the same in every tree): gradient
    // sums alone in 4-byte cells for the unit-Hessian integer forms, [G|H] words for the other
    // integer forms, fixed-point G for FP64_FIXED with a unit Hessian; 8-byte cells take the
    // 32-slot tiles of the default budget. MPSB_ROOT_GONLY=0 keeps every root packed.
    const char* genv = std::getenv("MPSB_ROOT_GONLY");
    const int k = kind & 0xFF;
    const uint32_t root_budget = std::min(s.lds_max, Tuning{}.lds_budget);
    if (!seq && !(genv && genv[0] == '0')) {
      if (k == kHistQ && unit_h && plan_g32_tiles(d, std::min(s.lds_max, s.tune.lds_budget), tiles_g)) {
        root_free = true;
        root_kind = kHistG32;
        root_cell_bytes = 4;
      } else if ((k == kHistQ && !unit_h) || (k == kHistFix64 && unit_h)) {
        try {
          tiles_g = plan_tiles(d, 8, root_budget, s.tune.block, s.tune.wpt);
          root_free = true;
          root_kind = k == kHistQ ? kHistQNC : kHistF64NC;
          root_cell_bytes = 8;
        } catch (const Error&) {
          root_free = false;  // a feature too wide for the budget: packed roots
        }
      }
    }
    if (root_free) {
      d_tiles_g = dptr<TileDesc>(s, kBufTilesG, tiles_g.tiles.size());
      d_cell_lds_g = dptr<uint32_t>(s, kBufCellLdsG, tiles_g.cell_lds.size());
      d_word_off_g = dptr<uint32_t>(s, kBufWordOffG, tiles_g.word_off.size());
      h2d(s, d_tiles_g, tiles_g.tiles.data(), tiles_g.tiles.size() * sizeof(TileDesc), "tile upload");
      h2d(s, d_cell_lds_g, tiles_g.cell_lds.data(), tiles_g.cell_lds.size() * 4, "tile upload");
      h2d(s, d_word_off_g, tiles_g.word_off.data(), tiles_g.word_off.size() * 4, "tile upload");
      root_count = dptr<uint32_t>(s, kBufRootCount, d.cells);
    }
    err = dptr<uint32_t>(s, kBufErr, 4);
    MPSB_HIP("error flag reset", hipMemsetAsync(err, 0, 8, s.stream));
    if (!p.monotonic.empty()) {
      d_mono = dptr<int8_t>(s, kBufMono, p.monotonic.size());
      h2d(s, d_mono, p.monotonic.data(), p.monotonic.size(), "constraint upload");
    }
    timer.acc[kTGrad] = &stats.gradient_seconds;
    timer.acc[kTHist] = &stats.histogram_seconds;
    timer.acc[kTReduce] = &stats.reduce_seconds;
    timer.acc[kTScan] = &stats.scan_seconds;
    timer.acc[kTPart] = &stats.partition_seconds;
    timer.acc[kTRenew] = &stats.renewal_seconds;
    s.sync("fit setup");
  }

  // Capacity of the output level only: growing it may drop its old contents (a finished
  // level), never those of the parent level being subtracted from.
  void ensure_hist_slots(int i, uint32_t slots) {
    if (slots <= hist_slots[i]) return;
    const uint32_t want = std::max(slots, hist_slots[i] * 2);
    const size_t n = size_t(want) * d.cells;
    uint8_t* base = static_cast<uint8_t*>(s.get(i ? kBufHistB : kBufHistA, n * 20));
    hist[i].g = reinterpret_cast<uint64_t*>(base);
    hist[i].h = reinterpret_cast<uint64_t*>(base + n * 8);
    hist[i].cnt = reinterpret_cast<uint32_t*>(base + n * 16);
    hist_slots[i] = want;
  }

  // The regression rows' FP64 gradient copy: g per row (unit Hessian) or (g, h) pairs.
  double* row_stash() { return dptr<double>(s, kBufGradStash, (unit_h ? 1u : 2u) * size_t(d.rows)); }

  // Read back the error word after a sync and raise the reference's TrainingError.
  // The reference's messages: src/core/trainer.cpp:200-266, src/core/objective.cpp:63-287,
  // src/core/multiclass_trainer.cpp:129-146, src/core/tree_split.cpp:167-181,
  // src/core/tree_histogram.cpp:16-66.
  void check_errors(uint32_t flags, uint32_t first = 0) {
    raise_device_errors(flags, first, softmax ? -1 : int(p.objective));
  }

  // ------------------------------------------------------------------ gradients ---
  void gradients(GradLaunch g) {
    s.check_fault(kFaultGradients, "gradients");
    g.objective = int(p.objective);
    g.obj_alpha = p.objective_alpha;
    g.obj_rho = p.tweedie_power;
    g.precision = precision;
    g.unit_h = unit_h ? 1u : 0u;
    g.err = err;
    g.gh = s.get(kBufGH, size_t(d.rows) * gh_bytes(precision, unit_h));
    g.seed = seed;
    // Regression rows: the write pass keeps the round's FP64 gradients in row order when the
    // resolution or the renewal will read them (one load per row instead of the inputs').
    g.gstash = g.source == 0 && (resolve || renew) ? row_stash() : nullptr;
    const size_t t0 = timer.begin();
    if (scaled()) {
      // Pass 1: max |g|, |h|; pass 2 derives this round's scales from them on the device
      // (no host round trip) and publishes the value scales for the split scan.
      auto* mx = dptr<unsigned long long>(s, kBufMaxabs, 2);
      MPSB_HIP("gradients", hipMemsetAsync(mx, 0, 16, s.stream));
      g.maxabs = mx;
      g.scales = d_scales;
      if (g.source == 2) g.stash = dptr<double>(s, kBufGradStash, 2 * size_t(d.rows));
      launch_grad_maxabs(g, s.stream);
    }
    launch_grad_write(g, s.stream);
    timer.end(kTGrad, t0);
    round_g = g;
  }

  // A batched softmax round's gradients: pass 1 evaluates the softmax once per row for all
  // classes (the previous round's class trees applied to their columns first) into the [K][rows]
  // stash and the round's max |g|, |h|; the write pass quantizes the K x rows stash as one
  // column, so every class shares the round's scales and class k, row i takes stochastic-
  // rounding index k * rows + i. Returns the launch the round's renewal reads.
  GradLaunch gradients_batched(GradLaunch g) {
    s.check_fault(kFaultGradients, "gradients");
    if (!scaled()) fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds need the fp64_fixed, int16 or int8 form");
    const uint64_t npos = positions();
    g.source = 3;
    g.precision = precision;
    g.unit_h = 0u;
    g.err = err;
    g.seed = seed;
    const size_t t0 = timer.begin();
    auto* mx = dptr<unsigned long long>(s, kBufMaxabs, 2);
    MPSB_HIP("gradients", hipMemsetAsync(mx, 0, 16, s.stream));
    g.maxabs = mx;
    g.scales = d_scales;
    g.stash = dptr<double>(s, kBufGradStash, 2 * size_t(npos));
    g.gh = s.get(kBufGH, size_t(npos) * gh_bytes(precision, unit_h));
    launch_grad_maxabs(g, s.stream);
    GradLaunch w = g;
    w.source = 2;
    w.n = npos;
    w.leaf_node = nullptr;
    w.leaf_table = nullptr;
    w.pred_out = nullptr;
    launch_grad_write(w, s.stream);
    timer.end(kTGrad, t0);
    round_g = w;
    return w;
  }

  // Near-tie resolution for this fit: INT16 unless kept, INT8 when asked for (its noise makes
  // most nodes ambiguous). The scans bound their gains' noise from here on (the root scan is
  // planned with the device plan).
  void resolve_prepare() {
    const bool on = p.near_ties == 1 || (p.near_ties == 0 && precision == MPSB_PRECISION_INT16);
    const bool sr = precision == MPSB_PRECISION_INT16 || precision == MPSB_PRECISION_INT8;
    noise_k = sr && on ? 3.0 : 0.0;
  }

  // Buffers of the resolution, from the pool before any tree (fits of at most 1024 features:
  // the scans' contender masks are built in LDS). The exact histograms and totals start at
  // zero; each level's launches leave them at zero.
  void resolve_setup() {
    resolve = noise_k > 0.0 && !seq && d.features <= 1024;
    if (!resolve) return;
    const uint32_t R = kResolveNodes, W = (d.features + 31) / 32;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
      const size_t at = off;
      off += (bytes + 255) & ~size_t(255);
      return at;
    };
    const size_t o_parent = carve(R * sizeof(NodeStats)), o_stats = carve(32),
                 o_pf = carve(size_t(R) * d.features * sizeof(Candidate)),
                 o_contend = carve(size_t(kResolveLevelNodes) * W * 4);
    uint8_t* b = static_cast<uint8_t*>(s.get(kBufResolve, off));
    rs.parent = reinterpret_cast<NodeStats*>(b + o_parent);
    rs.stats = reinterpret_cast<unsigned long long*>(b + o_stats);
    rs.per_feature = reinterpret_cast<Candidate*>(b + o_pf);
    rs.contend = reinterpret_cast<uint32_t*>(b + o_contend);
    MPSB_HIP("resolution reset", hipMemsetAsync(b, 0, o_pf, s.stream));  // totals and counts
    const size_t hn = size_t(R) * d.cells;
    uint8_t* hb = static_cast<uint8_t*>(s.get(kBufResolveHist, hn * 20));
    MPSB_HIP("resolution reset", hipMemsetAsync(hb, 0, hn * 20, s.stream));
    rs.hist.g = reinterpret_cast<uint64_t*>(hb);
    rs.hist.h = reinterpret_cast<uint64_t*>(hb + hn * 8);
    rs.hist.cnt = reinterpret_cast<uint32_t*>(hb + hn * 16);
    if (dp.on) dp.root_scan.contend = rs.contend;
    s.sync("resolution setup");
  }

  // Exact resolution of a level's ambiguous nodes (after the level's scan, before its planner).
  void resolve_level(const ScanLaunch& lvl, const LayerNode* layer, uint32_t n_bound, const uint32_t* n_dyn,
                     const uint32_t* rows) {
    if (!resolve || !lvl.contend) return;
    const size_t t0 = timer.begin();
    ResolveLaunch a{};
    a.cand = lvl.out;
    a.contend = rs.contend;
    a.n_nodes_dyn = n_dyn;
    a.n_nodes = n_bound;
    a.scan = lvl.nodes;
    a.layer = layer;
    a.rows_total = uint32_t(d.rows);
    a.features = d.features;
    a.mask_words = (d.features + 31) / 32;
    a.r_max = kResolveNodes;
    a.s_max = kResolveSlabs;
    a.res_parent = rs.parent;
    a.res_stats = rs.stats;
    a.res_stats_features = rs.stats + 2;
    a.res_hist = rs.hist;
    a.res_pf = rs.per_feature;
    a.g = round_g;
    a.g.leaf_node = nullptr;  // the gradient pass applied the previous tree already
    a.g.pred_out = nullptr;
    a.maxabs = static_cast<const unsigned long long*>(s.ptr[kBufMaxabs]);
    a.rows = rows;
    a.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
    a.row_stride = d.row_stride;
    a.codes_cm = static_cast<const uint8_t*>(s.ptr[kBufCodesCM]);
    a.code_bytes = d.code_bytes;
    a.hist = lvl.hist;
    a.cells = d.cells;
    a.cell_offset = lvl.cell_offset;
    a.bin_count = lvl.bin_count;
    a.miss_cell = lvl.miss_cell;
    a.mono = lvl.mono;
    a.p = lvl.p;
    a.err = err;
    launch_resolve(a, s.stream);
    timer.end(kTScan, t0);
  }

  // Leaf renewal for this fit: automatic in the FP32 and integer forms, explicit in the others;
  // it needs per-row leaf ids (the deferred update) and, for softmax class columns, the FP64
  // (g, h) that the scaled forms' max pass keeps. Buffers come from the pool before any tree.
  void renew_setup(uint64_t trees, bool softmax_columns) {
    const bool auto_form = precision == MPSB_PRECISION_FP32 || precision == MPSB_PRECISION_INT16 ||
                           precision == MPSB_PRECISION_INT8;
    renew = leaf_node && (p.leaf_solve == 2 || (p.leaf_solve == 0 && auto_form));
    if (renew && softmax_columns && !scaled()) {
      if (p.leaf_solve == 2)
        fail(MPSB_ERR_INVALID_ARGUMENT, "leaf renewal of softmax trees needs the fp64_fixed, int16 or int8 form");
      renew = false;
    }
    if (!renew) return;
    renew_softmax = softmax_columns;
    renew_ids = roots * uint32_t(std::min<uint64_t>(2 * uint64_t(max_leaves_of(p)), uint64_t(2) << p.max_depth) - 1);
    // Eight copies of the per-leaf totals, one per XCD of the summing blocks.
    const size_t ids8 = size_t(renew_ids) * kRenewCopies;
    uint8_t* b = static_cast<uint8_t*>(s.get(kBufRenew, ids8 * 20 + 64));
    renew_sums = reinterpret_cast<long long*>(b);
    renew_counts = reinterpret_cast<uint32_t*>(b + ids8 * 16);
    renew_max = reinterpret_cast<unsigned long long*>(b + ((ids8 * 20 + 7) & ~size_t(7)));
    if (dp.on) dp.leaf_info = dptr<LeafInfo>(s, kBufLeafInfo, dp.tree_nodes);
    renew_nodes = dptr<double>(s, kBufRenewNodes, size_t(trees) * renew_ids);
  }

  // Renews the leaves of the tree just grown (device- or host-planned) from the round's
  // gradients `base` (the launch the round's gradient pass was given).
  void renew_leaves(const GradLaunch& base, size_t model_index) {
    if (!renew) return;
    RenewLaunch r{};
    r.g = base;
    r.g.leaf_node = nullptr;  // the gradient pass applied the previous tree already
    r.g.leaf_table = nullptr;
    r.g.pred_out = nullptr;
    r.g.gh = nullptr;
    r.g.gh64 = nullptr;
    r.g.maxabs = nullptr;
    r.g.scales = nullptr;
    r.g.stash = renew_softmax ? static_cast<double*>(s.ptr[kBufGradStash]) : nullptr;
    r.g.gstash = base.source == 0 ? round_g.gstash : nullptr;
    r.g.objective = int(p.objective);
    r.g.obj_alpha = p.objective_alpha;
    r.g.obj_rho = p.tweedie_power;
    r.g.precision = precision;
    r.g.unit_h = unit_h ? 1u : 0u;
    r.g.err = err;
    r.leaf_id = leaf_node;
    r.leaf_bytes = leaf_id_bytes();
    r.maxabs = scaled() ? static_cast<const unsigned long long*>(s.ptr[kBufMaxabs]) : nullptr;
    r.own_maxabs = scaled() ? nullptr : renew_max;
    r.sums = renew_sums;
    r.counts = renew_counts;
    r.copy_stride = renew_ids;
    r.lambda = p.lambda;
    r.alpha = p.alpha;
    r.max_delta = p.max_delta;
    r.err = err;
    if (last_tree_device) {
      // Leaf ordinals of a device-planned tree are below its leaf count.
      r.n_ids = uint32_t(std::min<uint64_t>(dp.tree_nodes, uint64_t(roots) * std::min<uint64_t>(max_leaves_of(p),
                                                                                              uint64_t(1) << p.max_depth)));
      r.info = dp.leaf_info;
      r.tree_ctl = dp.tree_ctl;
      r.tree_slot = dp.slot;
      r.trees = dp.trees;
      r.tree_nodes = dp.tree_nodes;
      r.leaf_table = dp.leaf_table;
    } else {
      r.n_ids = leaf_ids;
      r.info = leaf_info;
      r.node_out = renew_nodes + renewed_host.size() * size_t(renew_ids);
      r.leaf_table = const_cast<double*>(leaf_table);
      renewed_host.push_back(model_index);
    }
    if (r.n_ids > renew_ids) fail(MPSB_ERR_INTERNAL, "leaf renewal: more leaf ids than planned");
    const size_t t0 = timer.begin();
    launch_renew(r, s.stream);
    timer.end(kTRenew, t0);
    ++stats.renewed_trees;
  }

  // Host-planned trees take their renewed leaf values (device-planned trees were renewed in
  // their device slots). After the fit's final synchronisation.
  void collect_renewals(std::vector<std::vector<mpsb_tree_node>>& trees) {
    if (renewed_host.empty()) return;
    std::vector<double> v(renewed_host.size() * size_t(renew_ids));
    d2h(s, v.data(), renew_nodes, v.size() * 8, "leaf renewal read-back");
    s.sync("leaf renewal read-back");
    for (size_t t = 0; t < renewed_host.size(); ++t)
      for (size_t i = 0; i < trees[renewed_host[t]].size(); ++i) {
        mpsb_tree_node& n = trees[renewed_host[t]][i];
        if (n.flags & 1) n.leaf_value = v[t * renew_ids + i];
      }
  }

  // End of a fit: everything completed, arena released, stage times resolved, error word read.
  void finish() {
    uint32_t flags[2] = {0, 0};
    d2h(s, flags, err, 8, "fit");
    s.sync("fit");
    arena.reset();
    timer.resolve();
    check_errors(flags[0], flags[1]);
  }

  // Fixed-point and integer forms: gradients carry a per-round scale.
  bool scaled() const { return precision != MPSB_PRECISION_FP32 && precision != MPSB_PRECISION_FP64; }

  // Host copy of the value scales (test hooks), after a synchronisation.
  void read_scales() {
    if (!scaled()) return;
    double sc[4];
    d2h(s, sc, d_scales, sizeof(sc), "scale readback");
    s.sync("scale readback");
    scale_g = sc[2];
    scale_h = sc[3];
  }

  // ---------------------------------------------------------------- histograms ---
  struct Build {
    uint32_t slot, begin, count;
    int32_t sibling;
    uint32_t parent_slot;
    uint32_t sib_begin = 0, sib_count = 0;  // sibling segment (row-order form)
  };

  // One level's histogram work, its descriptors already in the arena.
  struct LevelPlan {
    bool active = false, seq = false;
    int out = 0;
    uint64_t total = 0;
    HistLaunch h{};
    ReduceLaunch r{};
    SeqLaunch a{}, m{};
    const uint32_t* d_b = nullptr;
    const int32_t* d_s = nullptr;
    const uint32_t* d_p = nullptr;
    uint32_t n_sub = 0;
  };

  // Histograms of `builds` (rows = nullptr: identity row positions) into hist[out]; siblings
  // by subtraction from hist[out ^ 1].
  void build_level(const std::vector<Build>& builds, const uint32_t* rows, int out) {
    const LevelPlan L = plan_build(builds, rows, out);
    arena.flush();
    run_build(L);
  }

  void run_build(const LevelPlan& L) {
    if (!L.active) return;
    const size_t t0 = timer.begin();
    if (L.seq) {
      launch_hist_seq(L.a, s.stream);
      launch_sub_seq(L.d_b, L.d_s, L.d_p, L.n_sub, hist[L.out], hist[L.out ^ 1], d.cells, L.a.unit_h, err,
                     s.stream);
      launch_hist_seq(L.m, s.stream);
      timer.end(kTHist, t0);
    } else {
      launch_hist(L.h, s.stream);
      timer.end(kTHist, t0);
      const size_t t1 = timer.begin();
      launch_reduce(L.r, s.stream);
      timer.end(kTReduce, t1);
    }
    count_build(L);
  }

  // Histogram statistics of a host-planned build (device-planned levels count on the device).
  void count_build(const LevelPlan& L) {
    if (L.h.dyn_slabs) return;
    const uint64_t bytes = L.total * (uint64_t(d.features) * d.code_bytes + gh_bytes(precision, unit_h));
    ++stats.histogram_launches;
    stats.histogram_rows += L.total;
    stats.histogram_bytes += bytes;
    if (timer.on) {
      ++stats.histogram_timed_launches;
      stats.histogram_timed_bytes += bytes;
    }
  }

  // Rows per slab of a level of `total` built rows. A root without counts keeps the packed
  // form's slabs (the gradient-only root: half as many blocks over its half as many tiles, each
  // block summing twice the rows, and the partials and their reduction halve; measured faster
  // than as many blocks).
  uint64_t slab_rows_for(uint64_t total, bool count_free) const {
    const uint64_t want_slabs = std::max<uint64_t>(8, s.tune.target_blocks / tiles.tiles.size());
    const uint64_t slab_rows = std::max<uint64_t>(s.tune.min_slab_rows, (total + want_slabs - 1) / want_slabs);
    return std::min<uint64_t>(slab_rows, count_free && root_kind == kHistG32 ? slab_cap_g32(precision)
                                                                              : slab_cap(precision, unit_h));
  }

  // count_free: the root form without counts (root_kind) with its own tiles and slab cap; the counts
  // come from root_count, so only builds of every row without a sibling (roots) take it.
  LevelPlan plan_build(const std::vector<Build>& builds, const uint32_t* rows, int out, bool count_free = false) {
    LevelPlan L;
    if (builds.empty()) return L;
    s.check_fault(kFaultHistogram, "histogram");
    if (seq) return plan_build_seq(builds, rows, out);
    const TilePlan& tp = count_free ? tiles_g : tiles;
    const auto t_enc = std::chrono::steady_clock::now();
    uint64_t total = 0;
    uint32_t max_slot = 0;
    for (const Build& b : builds) {
      total += b.count;
      max_slot = std::max({max_slot, b.slot + 1, b.sibling >= 0 ? uint32_t(b.sibling) + 1 : 0u});
    }
    ensure_hist_slots(out, max_slot);
    const uint32_t nt = uint32_t(tp.tiles.size());
    const uint64_t slab_rows = slab_rows_for(total, count_free);
    std::vector<SlabDesc> slabs;
    std::vector<uint32_t> first, slot, parent;
    std::vector<int32_t> sib;
    for (const Build& b : builds) {
      first.push_back(uint32_t(slabs.size()));
      const uint64_t n = std::max<uint64_t>(1, (b.count + slab_rows - 1) / slab_rows);
      for (uint64_t i = 0; i < n; ++i) {
        const uint64_t lo = b.begin + i * slab_rows;
        slabs.push_back({uint32_t(lo), uint32_t(std::min<uint64_t>(b.begin + b.count, lo + slab_rows))});
      }
      slot.push_back(b.slot);
      sib.push_back(b.sibling);
      parent.push_back(b.parent_slot);
    }
    first.push_back(uint32_t(slabs.size()));
    const uint64_t plane = uint64_t(slabs.size()) * d.cells;
    void* partial = s.get(kBufPartial, plane * (count_free ? root_cell_bytes : 20));

    HistLaunch h{};
    h.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
    h.row_stride_bytes = d.row_stride * d.code_bytes;
    h.rows = rows;
    h.gh = gh_cur ? gh_cur : s.ptr[kBufGH];
    h.gh_by_pos = gh_pos ? 1u : 0u;
    h.slabs = arena.put(slabs.data(), slabs.size());
    h.n_slabs = uint32_t(slabs.size());
    h.tiles = count_free ? d_tiles_g : d_tiles;
    h.n_tiles = nt;
    h.word_cell_off = count_free ? d_word_off_g : d_word_off;
    h.cell_lds = count_free ? d_cell_lds_g : d_cell_lds;
    h.partial = partial;
    h.plane = plane;
    h.cells = d.cells;
    h.code_bytes = d.code_bytes;
    h.lds_bytes = tp.lds_bytes;
    h.kind = count_free ? (root_kind | (kind & ~0xFF)) : kind;
    h.unit_h = unit_h;
    h.block = count_free && root_kind == kHistG32 ? 1024 : int(s.tune.block);
    h.wpt = int(tp.wpt);
    h.interleaved = tp.interleaved ? tp.slots : 0u;

    ReduceLaunch r{};
    r.partial = partial;
    r.plane = plane;
    r.cells = d.cells;
    r.slab_first = arena.put(first.data(), first.size());
    r.build_slot = arena.put(slot.data(), slot.size());
    r.sibling_slot = arena.put(sib.data(), sib.size());
    r.parent_slot = arena.put(parent.data(), parent.size());
    r.n_build = uint32_t(builds.size());
    r.out = hist[out];
    r.parent = hist[out ^ 1];
    r.kind = h.kind;
    r.unit_h = unit_h;
    r.err = err;
    r.counts = count_free ? root_count : nullptr;
    s.timing.histogram_encode_seconds +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_enc).count();
    L.active = true;
    L.out = out;
    L.total = total;
    L.h = h;
    L.r = r;
    return L;
  }

  // Row-order FP64 form: built nodes by the serial row walk, siblings by FP64 subtraction
  // and a row-order pass over their rows for the missing cells.
  LevelPlan plan_build_seq(const std::vector<Build>& builds, const uint32_t* rows, int out) {
    uint32_t max_slot = 0;
    uint64_t total = 0;
    std::vector<SeqNode> full, miss;
    std::vector<uint32_t> bslot, pslot;
    std::vector<int32_t> sslot;
    for (const Build& b : builds) {
      max_slot = std::max({max_slot, b.slot + 1, b.sibling >= 0 ? uint32_t(b.sibling) + 1 : 0u});
      full.push_back({b.begin, b.count, b.slot, 0});
      bslot.push_back(b.slot);
      sslot.push_back(b.sibling);
      pslot.push_back(b.parent_slot);
      if (b.sibling >= 0) miss.push_back({b.sib_begin, b.sib_count, uint32_t(b.sibling), 0});
      total += b.count;
    }
    ensure_hist_slots(out, max_slot);
    SeqLaunch a{};
    a.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
    a.row_stride = d.row_stride;
    a.code_bytes = d.code_bytes;
    a.rows = rows;
    a.gh = static_cast<const double*>(s.ptr[kBufGH]);
    a.nodes = arena.put(full.data(), full.size());
    a.n_nodes = uint32_t(full.size());
    a.feat_list = d_seq_feat;
    a.n_feat = d.features;
    a.fg = seq_fg;
    a.cpf = seq_cpf;
    a.cell_offset = static_cast<const uint32_t*>(s.ptr[kBufCellOffset]);
    a.bin_count = static_cast<const uint32_t*>(s.ptr[kBufBinCount]);
    a.out = hist[out];
    a.cells = d.cells;
    a.unit_h = unit_h ? 1u : 0u;
    a.err = err;
    SeqLaunch m = a;
    m.missing_only = 1;
    m.feat_list = d_seq_feat + d.features;
    m.n_feat = n_miss_feat;
    m.nodes = miss.empty() ? nullptr : arena.put(miss.data(), miss.size());
    m.n_nodes = uint32_t(miss.size());
    LevelPlan L;
    L.active = L.seq = true;
    L.out = out;
    L.total = total;
    L.a = a;
    L.m = m;
    L.d_b = arena.put(bslot.data(), bslot.size());
    L.d_s = arena.put(sslot.data(), sslot.size());
    L.d_p = arena.put(pslot.data(), pslot.size());
    L.n_sub = uint32_t(bslot.size());
    return L;
  }

  // Parent statistics of one node by SumRows in row order (row-order form).
  void sum_rows(const uint32_t* rows, uint32_t begin, uint32_t n) {
    launch_sum_rows(static_cast<const double*>(s.ptr[kBufGH]), rows, begin, n, d_node_stats, s.stream);
  }

  // ---------------------------------------------------------------------- scan ---
  ScanParams scan_params() const {
    ScanParams q{};
    q.min_samples_leaf = p.min_samples_leaf;
    q.min_child_weight = p.min_child_weight;
    q.lambda = p.lambda;
    q.alpha = p.alpha;
    q.max_delta = p.max_delta;
    q.gamma = p.gamma;
    q.min_gain = p.min_gain;
    q.random_threshold = p.random_threshold ? 1u : 0u;
    q.seed = p.seed;
    q.has_mono = p.monotonic.empty() ? 0u : 1u;
    q.has_mask = p.groups.empty() ? 0u : 1u;
    q.scale_g = scale_g;
    q.scale_h = scale_h;
    q.scales = scaled() ? d_scales : nullptr;
    q.unit_h = unit_h ? 1u : 0u;
    q.integer = (precision == MPSB_PRECISION_FP32 || precision == MPSB_PRECISION_FP64) ? 0u : 1u;
    q.ref_view = precision == MPSB_PRECISION_FP64 ? 1u : 0u;
    q.noise_k = noise_k;
    return q;
  }

  // Best split of each node (one block per node) read back to the host; raises on device
  // error flags. per_feature (optional) receives [n][features] candidates.
  std::vector<Candidate> scan(const std::vector<ScanNode>& nodes, int in, const std::vector<uint32_t>& masks,
                              Candidate* per_feature_host = nullptr) {
    const ScanLaunch l = plan_scan(nodes, in, masks);
    arena.flush();
    run_scan(l);
    return collect_scan(l, per_feature_host);
  }

  void run_scan(const ScanLaunch& l) {
    const size_t t0 = timer.begin();
    launch_scan(l, s.stream);
    timer.end(kTScan, t0);
  }

  ScanLaunch plan_scan(const std::vector<ScanNode>& nodes, int in, const std::vector<uint32_t>& masks) {
    s.check_fault(kFaultScan, "split scan");
    ScanLaunch l{};
    l.hist = hist[in];
    l.cells = d.cells;
    l.cell_offset = static_cast<const uint32_t*>(s.ptr[kBufCellOffset]);
    l.bin_count = static_cast<const uint32_t*>(s.ptr[kBufBinCount]);
    l.miss_cell = static_cast<const uint32_t*>(s.ptr[kBufMissCell]);
    l.features = d.features;
    l.nodes = arena.put(nodes.data(), nodes.size());
    l.n_nodes = uint32_t(nodes.size());
    l.p = scan_params();
    l.mono = d_mono;
    l.masks = masks.empty() ? nullptr : arena.put(masks.data(), masks.size());
    l.out = dptr<Candidate>(s, kBufCand, nodes.size());
    l.per_feature = dptr<Candidate>(s, kBufPerFeature, nodes.size() * d.features);
    l.parent_stats = d_node_stats;
    l.err = err;
    l.sequential = sequential_scan ? 1 : 0;
    l.contend = resolve && nodes.size() <= kResolveLevelNodes ? rs.contend : nullptr;
    last_scan = l;
    return l;
  }

  // Candidates and the error word of the level. Training: the device copies them into
  // host-mapped memory and publishes a sequence number the host waits on (no stream sync,
  // no copy calls); stage times are resolved at the tree's end. Test hooks that also want
  // every feature's candidate read back through copies and a synchronisation.
  std::vector<Candidate> collect_scan(const ScanLaunch& l, Candidate* per_feature_host = nullptr) {
    std::vector<Candidate> out(l.n_nodes);
    uint32_t flags = 0;
    if (!per_feature_host) {
      s.ensure_signal(l.n_nodes);
      const uint64_t seq = ++s.sig_seq;
      launch_signal(l.out, l.n_nodes, err, reinterpret_cast<Candidate*>(s.sig_dev + 64),
                    reinterpret_cast<uint32_t*>(s.sig_dev + 8), reinterpret_cast<unsigned long long*>(s.sig_dev),
                    seq, s.stream);
      s.wait_signal(seq, "split scan");
      std::memcpy(out.data(), s.sig + 64, out.size() * sizeof(Candidate));
      uint32_t first = 0;
      std::memcpy(&flags, s.sig + 8, 4);
      std::memcpy(&first, s.sig + 12, 4);
      arena.reset();
      timer.resolve();
      ++stats.levels;
      check_errors(flags, first);
      return out;
    }
    d2h(s, out.data(), l.out, out.size() * sizeof(Candidate), "split scan");
    uint32_t words[2] = {0, 0};
    d2h(s, words, err, 8, "split scan");
    d2h(s, per_feature_host, l.per_feature, size_t(l.n_nodes) * d.features * sizeof(Candidate), "split scan");
    s.sync("split scan");
    arena.reset();
    timer.resolve();
    ++stats.levels;
    check_errors(words[0], words[1]);
    return out;
  }

  // --------------------------------------------------------------- partition ---
  void partition(const std::vector<PartNode>& nodes_in, const uint32_t* rows_in, uint32_t* rows_out,
                 double* pred, uint32_t pred_stride, uint32_t pred_offset, const void* gh_in = nullptr,
                 void* gh_out = nullptr) {
    const PartLaunch a = plan_partition(nodes_in, rows_in, rows_out, pred, pred_stride, pred_offset, gh_in, gh_out);
    arena.flush();
    run_partition(a);
  }

  void run_partition(const PartLaunch& a) {
    if (a.n_chunks == 0) return;
    const size_t t0 = timer.begin();
    launch_partition(a, s.stream);
    timer.end(kTPart, t0);
  }

  PartLaunch plan_partition(const std::vector<PartNode>& nodes_in, const uint32_t* rows_in, uint32_t* rows_out,
                            double* pred, uint32_t pred_stride, uint32_t pred_offset, const void* gh_in,
                            void* gh_out) {
    PartLaunch a{};
    if (nodes_in.empty()) return a;
    s.check_fault(kFaultPartition, "partition");
    std::vector<PartNode> nodes = nodes_in;
    std::vector<uint32_t> chunk_node;
    for (uint32_t i = 0; i < nodes.size(); ++i) {
      // Dense nodes (at least one row per 128-row line of a column) read the split feature
      // from the column-major copy; sparse ones from their own row lines.
      nodes[i].column_major = uint64_t(nodes[i].count) * 128 >= d.rows ? 1u : 0u;
      nodes[i].chunk0 = uint32_t(chunk_node.size());
      const uint32_t n = (nodes[i].count + kPartChunk - 1) / kPartChunk;
      for (uint32_t c = 0; c < n; ++c) chunk_node.push_back(i);
    }
    if (chunk_node.empty()) return a;
    a.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
    a.codes_cm = static_cast<const uint8_t*>(s.ptr[kBufCodesCM]);
    a.rows_total = d.rows;
    a.row_stride = d.row_stride;
    a.code_bytes = d.code_bytes;
    a.rows_in = rows_in;
    a.rows_out = rows_out;
    a.nodes = arena.put(nodes.data(), nodes.size());
    a.n_nodes = uint32_t(nodes.size());
    a.chunk_node = arena.put(chunk_node.data(), chunk_node.size());
    a.n_chunks = uint32_t(chunk_node.size());
    a.chunk_left = dptr<uint32_t>(s, kBufChunkLeft, chunk_node.size());
    a.bits = dptr<uint32_t>(s, kBufBits, chunk_node.size() * (kPartChunk / 32));
    a.pred = pred;
    a.lr = p.lr;
    a.pred_stride = pred_stride;
    a.pred_offset = pred_offset;
    a.leaf_node_out = leaf_node;  // deferred update: leaf node ids instead of pred += lr * leaf
    a.leaf_bytes = 2;
    a.gh_in = static_cast<const uint8_t*>(gh_in);
    a.gh_out = static_cast<uint8_t*>(gh_out);
    a.gh_bytes = gh_in ? gh_bytes(precision, unit_h) : 0u;
    a.no_scatter = 1u;
    for (const PartNode& nd : nodes)
      if (nd.action == kPartSplit) a.no_scatter = 0u;
    a.err = err;
    return a;
  }

  // ------------------------------------------------------ device-side level planning ---
  // The tiled forms without interaction constraints plan each level on the device (k_plan):
  // a tree's levels are launched back to back with grids for the largest counts the level
  // can have, and the host never waits inside a tree. Trees are read back when the fit ends.
  struct DevicePlan {
    bool on = false;
    uint32_t level_nodes = 0;  // most nodes of one level (2^(max_depth - 1))
    uint32_t tree_nodes = 0;   // most nodes of one tree
    uint32_t max_slabs = 0;
    uint32_t slab_base = 0;    // slabs of a level beyond one per build
    uint64_t base_chunks = 0;  // ceil(rows / chunk)
    uint64_t plane = 0;
    uint64_t slots = 0, used = 0;  // tree slots on the device
    PartNode* part = nullptr;
    LayerNode* layer[2] = {nullptr, nullptr};
    SlabDesc* slabs = nullptr;
    uint32_t* slab_first = nullptr;
    uint32_t* build_slot = nullptr;
    int32_t* sibling_slot = nullptr;
    uint32_t* parent_slot = nullptr;
    ScanNode* scan = nullptr;
    LevelCtl* ctl = nullptr;
    mpsb_tree_node* trees = nullptr;
    uint32_t* tree_ctl = nullptr;
    double* leaf_table = nullptr;
    LeafInfo* leaf_info = nullptr;        // leaf renewal: node and bounds per ordinal
    unsigned long long* stats = nullptr;  // [0] rows, [1] launches; [2], [3] of timed trees
    uint32_t* chunk_node = nullptr;
    uint32_t* block_chunk = nullptr;
    uint32_t* chunk_left = nullptr;
    uint32_t* bits = nullptr;
    uint32_t* root_rows = nullptr;        // batched rounds: level-0 position -> row
    LayerNode* root_layer = nullptr;      // batched rounds: the class roots
    Candidate* cand = nullptr;
    Candidate* per_feature = nullptr;
    void* partial = nullptr;
    LevelPlan root;            // the root level is the same for every tree
    LevelPlan root_free;       // ... without counts, once the root's counts are known
    ScanLaunch root_scan{};
    uint32_t* slot = nullptr;  // device counter of the tree slot the next tree fills
    uint32_t leaf_bytes = 2;   // width of the per-row leaf ordinals (1 while leaves <= 256)
    // A tree's launches, captured once and replayed for the untimed trees, with the pooled
    // buffers they were captured on.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool no_graph = false;  // a capture failed: launch directly
    std::array<void*, kBufCount> graph_ptrs{};
  };
  DevicePlan dp;

  ~Fit() { drop_tree_graph(); }

  // Width of the per-row leaf ids: device-planned trees store leaf ordinals (one byte while a
  // tree has at most 256 leaves), host-planned ones tree node ids.
  uint32_t leaf_id_bytes() const { return dp.on ? dp.leaf_bytes : 2u; }

  void drop_tree_graph() {
    if (!dp.exec && !dp.graph) return;
    (void)hipStreamSynchronize(s.stream);  // no replay still in flight
    if (dp.exec) (void)hipGraphExecDestroy(dp.exec);
    if (dp.graph) (void)hipGraphDestroy(dp.graph);
    dp.exec = nullptr;
    dp.graph = nullptr;
  }

  // Sets device planning up for a fit of `trees` trees when the form and parameters allow it
  // and its worst-case buffers stay modest; otherwise the host plans each level.
  void device_plan_setup(uint64_t trees) {
    const char* env = std::getenv("MPSB_HOST_PLAN");  // "1": plan on the host (comparison runs)
    if (seq || !p.groups.empty() || p.max_depth < 1 || p.max_depth > 12 || trees == 0 || (env && env[0] == '1') ||
        s.tune.planner == 1)
      return;
    // A small fit running beside other fits plans on the host: its level round trips overlap
    // the other sessions' kernels, where planner kernels and grids sized for the largest level
    // would take turns with them on the shared hardware queues. Either way gives the same model.
    if (s.tune.planner == 0 && active_fits.load() > 1 && d.rows < kDevicePlanMinRowsShared) return;
    DevicePlan q;
    const uint64_t npos = positions();
    q.level_nodes = roots << (p.max_depth - 1);
    q.tree_nodes =
        roots * uint32_t(std::min<uint64_t>(2 * uint64_t(max_leaves_of(p)), uint64_t(2) << p.max_depth) - 1);
    const uint64_t nt = tiles.tiles.size();
    const uint64_t want = std::max<uint64_t>(8, s.tune.target_blocks / nt);
    const uint64_t cap = slab_cap(precision, unit_h);
    // Slabs of one level: at most (built rows / slab rows) + one per build, where slab rows is
    // at least min(min_slab_rows, cap) and built rows / slab rows at most max(want, rows / cap).
    const uint64_t min_rows = std::min<uint64_t>(s.tune.min_slab_rows, cap);
    q.slab_base = uint32_t(std::min<uint64_t>(std::max<uint64_t>(want, (npos + cap - 1) / cap),
                                              (npos + min_rows - 1) / min_rows));
    q.max_slabs = q.slab_base + q.level_nodes + 1;
    q.base_chunks = (npos + kPartChunk - 1) / kPartChunk;
    q.plane = uint64_t(q.max_slabs) * d.cells;
    const uint64_t chunks = q.base_chunks + q.level_nodes;
    const uint64_t bytes = q.plane * 20 + 2 * uint64_t(q.level_nodes) * d.cells * 20 +
                           uint64_t(q.level_nodes) * d.features * sizeof(Candidate) +
                           trees * q.tree_nodes * sizeof(mpsb_tree_node);
    if (bytes > (uint64_t(8) << 30)) return;
    q.slots = trees;
    ensure_hist_slots(0, q.level_nodes);
    ensure_hist_slots(1, q.level_nodes);
    q.part = dptr<PartNode>(s, kBufPlanPart, q.level_nodes);
    q.layer[0] = dptr<LayerNode>(s, kBufPlanLayerA, q.level_nodes);
    q.layer[1] = dptr<LayerNode>(s, kBufPlanLayerB, q.level_nodes);
    q.slabs = dptr<SlabDesc>(s, kBufPlanSlabs, q.max_slabs);
    uint32_t* b = dptr<uint32_t>(s, kBufPlanBuild, 4 * size_t(q.level_nodes) + 4);
    q.slab_first = b;
    q.build_slot = b + q.level_nodes + 1;
    q.sibling_slot = reinterpret_cast<int32_t*>(q.build_slot + q.level_nodes);
    q.parent_slot = q.build_slot + 2 * q.level_nodes;
    q.scan = dptr<ScanNode>(s, kBufPlanScan, q.level_nodes);
    q.ctl = dptr<LevelCtl>(s, kBufPlanCtl, p.max_depth + 1);
    // Tree slots, their [nodes, leaves, final leaves, -] counters, the leaf table and the statistics words
    // share one allocation.
    const size_t tree_bytes = size_t(trees) * q.tree_nodes * sizeof(mpsb_tree_node);
    uint8_t* t = static_cast<uint8_t*>(s.get(kBufPlanTrees, tree_bytes + trees * 16 + 64));
    q.trees = reinterpret_cast<mpsb_tree_node*>(t);
    q.tree_ctl = reinterpret_cast<uint32_t*>(t + tree_bytes);
    q.stats = reinterpret_cast<unsigned long long*>(t + tree_bytes + trees * 16);
    q.slot = reinterpret_cast<uint32_t*>(q.stats + 4);
    MPSB_HIP("planner reset", hipMemsetAsync(q.stats, 0, 64, s.stream));
    q.leaf_table = dptr<double>(s, kBufPlanLeafTable, q.tree_nodes);
    q.chunk_node = dptr<uint32_t>(s, kBufChunkNode, chunks);
    q.block_chunk = dptr<uint32_t>(s, kBufBlockChunk, chunks);
    q.chunk_left = dptr<uint32_t>(s, kBufChunkLeft, chunks);
    q.bits = dptr<uint32_t>(s, kBufBits, chunks * (kPartChunk / 32));
    q.cand = dptr<Candidate>(s, kBufCand, q.level_nodes);
    q.per_feature = dptr<Candidate>(s, kBufPerFeature, size_t(q.level_nodes) * d.features);
    // A root plan without counts beside the packed one (every root, one per class in a batched
    // round, holds all rows).
    uint64_t root_partial = 0;
    if (root_free) {
      const uint64_t sr = slab_rows_for(uint64_t(roots) * d.rows, true);
      root_partial = uint64_t(roots) * std::max<uint64_t>(1, (d.rows + sr - 1) / sr) * d.cells * root_cell_bytes;
    }
    // One size for every plan below, so no later request moves the buffer under an earlier plan.
    s.get(kBufPartial, std::max<uint64_t>(q.plane * 20, root_partial));
    q.partial = s.get(kBufPartial, q.plane * 20);
    s.get(kBufGH, size_t(npos) * gh_bytes(precision, unit_h));  // the gradient pass reuses it
    // The root: every row from the identity order, its scan from the histogram totals. Its
    // descriptors stay in the arena for the whole fit (no level signal resets it). A batched
    // round has one root per class over its own block of positions, each position's row from a
    // fixed table (position mod rows), and the class roots as level 0's layer.
    gh_cur = s.ptr[kBufGH];
    gh_pos = true;
    std::vector<Build> root_builds;
    std::vector<ScanNode> root_scans;
    std::vector<LayerNode> root_layer;
    for (uint32_t k = 0; k < roots; ++k) {
      root_builds.push_back({k, uint32_t(k * d.rows), uint32_t(d.rows), -1, 0});
      ScanNode sn{};
      sn.slot = k;
      sn.node_index = k;
      sn.parent_from_hist = 1;
      sn.lo = -kInf;
      sn.hi = kInf;
      root_scans.push_back(sn);
      LayerNode ln{};
      ln.node = k;
      ln.begin = uint32_t(k * d.rows);
      ln.count = uint32_t(d.rows);
      ln.lo = -kInf;
      ln.hi = kInf;
      root_layer.push_back(ln);
    }
    if (roots > 1) {
      q.root_rows = dptr<uint32_t>(s, kBufRootRows, npos);
      launch_iota_mod(q.root_rows, npos, uint32_t(d.rows), s.stream);
      q.root_layer = dptr<LayerNode>(s, kBufRootLayer, roots);
      h2d(s, q.root_layer, root_layer.data(), root_layer.size() * sizeof(LayerNode), "root upload");
    }
    q.root = plan_build(root_builds, q.root_rows, 0);
    if (root_free) q.root_free = plan_build(root_builds, q.root_rows, 0, true);
    q.root_scan = plan_scan(root_scans, 0, {});
    q.root_scan.out = q.cand;
    q.root_scan.per_feature = q.per_feature;
    arena.flush();
    gh_cur = nullptr;
    gh_pos = false;
    q.leaf_bytes = uint64_t(roots) * max_leaves_of(p) <= 256 ? 1u : 2u;
    q.on = true;
    dp = q;
    leaf_table = dp.leaf_table;
  }

  // Stage times of trees whose events have all completed, without waiting.
  void resolve_if_done() {
    if (timer.next == 0) return;
    if (hipEventQuery(s.event(timer.next - 1)) == hipSuccess) timer.resolve();
  }

  // One tree with device planning; its nodes land in tree slot dp.used (the device's slot
  // counter advances with it). With deferred leaf updates every launch of a tree has the same
  // arguments from one tree to the next, so with MPSB_TREE_GRAPH=1 the untimed trees replay one
  // captured graph of them; the timed ones launch directly, as their stage events need. Direct
  // launches are the default.
  void grow_tree_device(double* pred, uint32_t stride, uint32_t offset) {
    if (dp.used >= dp.slots) fail(MPSB_ERR_INTERNAL, "device tree slots exhausted");
    resolve_if_done();
    // The stages in launch order, before any launch (fault injection).
    s.check_fault(kFaultHistogram, "histogram");
    s.check_fault(kFaultScan, "split scan");
    for (uint32_t dep = 0; dep < p.max_depth; ++dep) {
      s.check_fault(kFaultPartition, "partition");
      if (dep + 1 >= p.max_depth) break;
      s.check_fault(kFaultHistogram, "histogram");
      s.check_fault(kFaultScan, "split scan");
    }
    ++dp.used;
    stats.levels += p.max_depth;
    // A graph is captured only while this is the only live fit, so the capture does not run
    // beside other fits' allocations and synchronisations (replays of an existing graph are
    // plain launches and continue either way).
    const char* env = std::getenv("MPSB_TREE_GRAPH");  // "1": replay captured trees (opt-in)
    if (timer.on || !leaf_node || dp.no_graph || !(env && env[0] == '1') || (!dp.exec && active_fits.load() > 1)) {
      launch_tree_device(pred, stride, offset);
      return;
    }
    // Buffers the launches take from the pool exist before a capture (no allocation in it).
    dptr<uint32_t>(s, kBufRowsA, positions());
    dptr<uint32_t>(s, kBufRowsB, positions());
    s.get(kBufGHB, size_t(positions()) * gh_bytes(precision, unit_h));
    if (dp.exec && dp.graph_ptrs != s.ptr) drop_tree_graph();  // a pooled buffer moved
    if (!dp.exec) {
      const mpsb_fit_stats before = stats;
      hipGraph_t g = nullptr;
      hipError_t r = hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed);
      if (r == hipSuccess) {
        try {
          launch_tree_device(pred, stride, offset);
        } catch (...) {
          (void)hipStreamEndCapture(s.stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        r = hipStreamEndCapture(s.stream, &g);
      }
      hipGraphExec_t e = nullptr;
      if (r == hipSuccess) r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
      if (r != hipSuccess) {
        // No graph came of it: this tree and the rest of the fit launch directly.
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        stats = before;
        dp.no_graph = true;
        launch_tree_device(pred, stride, offset);
        return;
      }
      dp.graph = g;
      dp.exec = e;
      dp.graph_ptrs = s.ptr;
    } else {
      count_build(dp.root);  // the capture counted the captured tree's root
      if (dp.root_free.active && root_counts_ready) ++stats.root_gradient_only_builds;
    }
    MPSB_HIP("tree replay", hipGraphLaunch(dp.exec, s.stream));
  }

  // The launches of one device-planned tree on the session stream.
  void launch_tree_device(double* pred, uint32_t stride, uint32_t offset) {
    const uint64_t n = positions();
    uint32_t* rows[2] = {dptr<uint32_t>(s, kBufRowsA, n), dptr<uint32_t>(s, kBufRowsB, n)};
    void* gh[2] = {s.ptr[kBufGH], s.get(kBufGHB, size_t(n) * gh_bytes(precision, unit_h))};
    // Root histogram and scan (descriptors planned once per fit). The first tree's packed root
    // histogram gives the root's per-cell counts; later roots, in forms that take them, are
    // built without counts.
    if (dp.root_free.active && root_counts_ready) {
      run_build(dp.root_free);
      ++stats.root_gradient_only_builds;
    } else {
      run_build(dp.root);
      if (dp.root_free.active) {
        launch_copy_u32(root_count, dp.root.r.out.cnt, d.cells, s.stream);  // the root's slot 0
        root_counts_ready = true;
      }
    }
    run_scan(dp.root_scan);
    resolve_level(dp.root_scan, nullptr, 1, nullptr, nullptr);
    int rc = 0, hc = 0;
    const uint32_t gb = gh_bytes(precision, unit_h);
    for (uint32_t dep = 0; dep < p.max_depth; ++dep) {
      const uint32_t nodes = std::min(roots << dep, dp.level_nodes);
      PlanLaunch q{};
      q.depth = dep;
      q.max_depth = p.max_depth;
      // A batched round's leaf cap counts the leaves of all its class trees (the per-tree cap of
      // such a fit never binds: 2^max_depth at least).
      q.cap = uint32_t(std::min<uint64_t>(uint64_t(roots) * max_leaves_of(p), 0xFFFFFFFFull));
      q.features = d.features;
      q.rows_total = d.rows;
      q.lambda = p.lambda;
      q.alpha = p.alpha;
      q.max_delta = p.max_delta;
      q.cand = dp.cand;
      q.layer = dep ? dp.layer[dep & 1] : nullptr;
      q.layer_cap = dp.level_nodes;
      q.ctl_in = dep ? dp.ctl + dep - 1 : nullptr;
      q.ctl = dp.ctl + dep;
      q.next = dp.layer[(dep + 1) & 1];
      q.part = dp.part;
      q.chunk_node = dp.chunk_node;
      q.block_chunk = dp.block_chunk;
      q.slabs = dp.slabs;
      q.slab_first = dp.slab_first;
      q.build_slot = dp.build_slot;
      q.sibling_slot = dp.sibling_slot;
      q.parent_slot = dp.parent_slot;
      q.scan = dp.scan;
      q.bin_count = static_cast<const uint32_t*>(s.ptr[kBufBinCount]);
      q.miss_cell = static_cast<const uint32_t*>(s.ptr[kBufMissCell]);
      q.tree = dp.trees;
      q.tree_ctl = dp.tree_ctl;
      q.tree_slot = dp.slot;
      q.tree_nodes = dp.tree_nodes;
      q.leaf_table = dp.leaf_table;
      q.leaf_info = renew ? dp.leaf_info : nullptr;
      q.min_slab_rows = s.tune.min_slab_rows;
      q.want_slabs = std::max<uint64_t>(8, s.tune.target_blocks / tiles.tiles.size());
      q.slab_cap = slab_cap(precision, unit_h);
      q.stats = dp.stats + (timer.on ? 2 : 0);
      q.roots = roots;
      q.root_layer = dp.root_layer;
      q.err = err;
      launch_plan(q, s.stream);

      // Partition of this level.
      PartLaunch a{};
      a.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
      a.codes_cm = static_cast<const uint8_t*>(s.ptr[kBufCodesCM]);
      a.rows_total = d.rows;
      a.row_stride = d.row_stride;
      a.code_bytes = d.code_bytes;
      a.rows_in = dep ? rows[rc] : dp.root_rows;  // a single root's positions are its rows
      a.rows_out = rows[rc ^ 1];
      a.nodes = dp.part;
      a.n_nodes = nodes;
      a.chunk_node = dp.chunk_node;
      a.block_chunk = dp.block_chunk;
      a.n_chunks = uint32_t(dp.base_chunks + nodes);
      a.chunk_left = dp.chunk_left;
      a.bits = dp.bits;
      a.pred = leaf_node ? nullptr : pred;  // deferred: leaf ids only, the same for every class tree
      a.lr = p.lr;
      a.pred_stride = stride;
      a.pred_offset = offset;
      a.leaf_node_out = leaf_node;
      a.leaf_bytes = dp.leaf_bytes;
      a.gh_in = static_cast<const uint8_t*>(gh[rc]);
      a.gh_out = static_cast<uint8_t*>(gh[rc ^ 1]);
      a.gh_bytes = gb;
      a.no_scatter = dep + 1 >= p.max_depth ? 1u : 0u;  // the last level only assigns leaves
      a.class_rows = roots > 1 ? uint32_t(d.rows) : 0u;
      a.ctl = dp.ctl + dep;
      a.err = err;
      run_partition(a);
      rc ^= 1;
      if (dep + 1 >= p.max_depth) break;

      // Histograms and split scan of the next level.
      LevelPlan L = dp.root;
      L.out = hc ^ 1;
      L.total = 0;
      L.h.rows = rows[rc];
      L.h.gh = gh[rc];
      L.h.gh_by_pos = 1;
      L.h.slabs = dp.slabs;
      L.h.n_slabs = dp.slab_base + nodes + 1;  // this level builds at most `nodes` children
      L.h.dyn_slabs = &dp.ctl[dep].n_slabs;
      L.h.partial = dp.partial;
      L.h.plane = dp.plane;
      L.r.partial = dp.partial;
      L.r.plane = dp.plane;
      L.r.slab_first = dp.slab_first;
      L.r.build_slot = dp.build_slot;
      L.r.sibling_slot = dp.sibling_slot;
      L.r.parent_slot = dp.parent_slot;
      L.r.n_build = nodes;
      L.r.dyn_build = &dp.ctl[dep].n_build;
      L.r.out = hist[hc ^ 1];
      L.r.parent = hist[hc];
      run_build(L);
      ScanLaunch sc = dp.root_scan;
      sc.hist = hist[hc ^ 1];
      sc.nodes = dp.scan;
      sc.n_nodes = std::min(2 * nodes, dp.level_nodes);
      sc.dyn_nodes = &dp.ctl[dep].n_next;
      sc.masks = nullptr;
      sc.parent_stats = nullptr;
      sc.contend = resolve && sc.n_nodes <= kResolveLevelNodes ? rs.contend : nullptr;
      run_scan(sc);
      resolve_level(sc, dp.layer[(dep + 1) & 1], sc.n_nodes, sc.dyn_nodes, rows[rc]);
      hc ^= 1;
    }
  }

  // Copies the device trees into the model's tree list (placeholders at `where`).
  // The resolution's counts over the fit (either planner).
  void collect_resolution() {
    if (!resolve) return;
    unsigned long long rv[3] = {0, 0, 0};
    d2h(s, rv, rs.stats, sizeof(rv), "resolution read-back");
    s.sync("resolution read-back");
    stats.ambiguous_nodes += rv[0];
    stats.resolved_nodes += rv[1];
    stats.resolved_features += rv[2];
  }

  void collect_device_trees(std::vector<std::vector<mpsb_tree_node>>& trees, const std::vector<size_t>& where) {
    if (!dp.on || dp.used == 0) return;
    std::vector<uint32_t> ctl(4 * dp.used);
    std::vector<mpsb_tree_node> nodes(size_t(dp.used) * dp.tree_nodes);
    d2h(s, ctl.data(), dp.tree_ctl, ctl.size() * 4, "tree read-back");
    d2h(s, nodes.data(), dp.trees, nodes.size() * sizeof(mpsb_tree_node), "tree read-back");
    unsigned long long st[4];
    d2h(s, st, dp.stats, sizeof(st), "tree read-back");
    s.sync("tree read-back");
    for (size_t i = 0; i < where.size(); ++i) {
      const uint32_t count = std::min(ctl[4 * i], dp.tree_nodes);
      const mpsb_tree_node* b = nodes.data() + i * dp.tree_nodes;
      if (roots == 1) {
        trees[where[i]].assign(b, b + count);
        continue;
      }
      // A batched round's tree: roots 0..K-1 are the class trees' roots and every child comes
      // after its parent, so each class's nodes, in this order, are its own tree level by level
      // in node order (the numbering a tree grown alone gets); child ids are renumbered.
      std::vector<uint32_t> cls(count, 0u), id(count, 0u), next(roots, 0u);
      for (uint32_t j = 0; j < count; ++j) {
        if (j < roots) cls[j] = j;
        id[j] = next[cls[j]]++;
        if (b[j].flags & 1) continue;
        if (b[j].left_child >= count || b[j].right_child >= count || b[j].left_child <= j || b[j].right_child <= j)
          fail(MPSB_ERR_INTERNAL, "batched round tree read-back: child out of order");
        cls[b[j].left_child] = cls[b[j].right_child] = cls[j];
      }
      for (uint32_t k = 0; k < roots; ++k) trees[where[i] + k].clear();
      for (uint32_t j = 0; j < count; ++j) {
        mpsb_tree_node nd = b[j];
        if (!(nd.flags & 1)) {
          nd.left_child = id[nd.left_child];
          nd.right_child = id[nd.right_child];
        }
        trees[where[i] + cls[j]].push_back(nd);
      }
    }
    // Histogram statistics of the planned levels (the root builds were counted as launched).
    const uint64_t per_row = uint64_t(d.features) * d.code_bytes + gh_bytes(precision, unit_h);
    stats.histogram_rows += st[0] + st[2];
    stats.histogram_launches += st[1] + st[3];
    stats.histogram_bytes += (st[0] + st[2]) * per_row;
    stats.histogram_timed_launches += st[3];
    stats.histogram_timed_bytes += st[2] * per_row;
  }
};

// ================================================================== tree growth ======
struct Active {
  uint32_t node = 0, depth = 0;
  uint32_t begin = 0, count = 0;
  NodeStats stats{};
  double lo = -kInf, hi = kInf;
  std::vector<uint32_t> path;
};

static mpsb_tree_node make_leaf(const NodeStats& s, const Params& p, double lo, double hi) {
  // MakeBoundedLeaf, src/core/tree_structure.cpp:72-87
  if (lo > hi) fail(MPSB_ERR_TRAINING, "monotonic leaf bounds are inconsistent");
  mpsb_tree_node n{};
  n.left_child = n.right_child = 0xFFFFFFFFu;
  n.leaf_value = std::clamp(host_leaf_weight(s.g, s.h, p.lambda, p.alpha, p.max_delta), lo, hi);
  n.default_left = 1;
  n.flags = 1;
  return n;
}

static bool interaction_ok(const std::vector<uint32_t>& path, uint32_t f, const Params& p) {
  // InteractionAllowsFeature, src/core/tree_interactions.cpp:32-45
  if (p.groups.empty()) return true;
  for (const auto& g : p.groups) {
    if (std::find(g.begin(), g.end(), f) == g.end()) continue;
    bool all = true;
    for (uint32_t x : path) all = all && std::find(g.begin(), g.end(), x) != g.end();
    if (all) return true;
  }
  return false;
}

// One tree, level-wise (TrainLevelWiseRegressionTree, src/core/tree.cpp:49-147). Rows of
// every leaf get pred[row * stride + offset] += lr * leaf on the device.
static std::vector<mpsb_tree_node> grow_tree(Fit& f, double* pred, uint32_t stride, uint32_t offset) {
  Session& s = f.s;
  const Params& p = f.p;
  const uint32_t n = uint32_t(f.d.rows);
  uint32_t* rows[2] = {dptr<uint32_t>(s, kBufRowsA, n), dptr<uint32_t>(s, kBufRowsB, n)};
  int rc = 0, hc = 0;
  launch_iota(rows[rc], n, s.stream);
  // Position-ordered gradients for the tiled forms: at the root positions are rows.
  void* gh[2] = {nullptr, nullptr};
  if (!f.seq) {
    gh[0] = s.ptr[kBufGH];
    gh[1] = s.get(kBufGHB, size_t(n) * gh_bytes(f.precision, f.unit_h));
    f.gh_cur = gh[0];
    f.gh_pos = true;
  }

  const uint32_t mask_words = (f.d.features + 31) / 32;
  auto masks_for = [&](const std::vector<Active>& layer) {
    std::vector<uint32_t> m;
    if (p.groups.empty()) return m;
    m.assign(layer.size() * mask_words, 0);
    for (size_t i = 0; i < layer.size(); ++i)
      for (uint32_t ft = 0; ft < f.d.features; ++ft)
        if (interaction_ok(layer[i].path, ft, p)) m[i * mask_words + ft / 32] |= 1u << (ft % 32);
    return m;
  };
  auto scan_nodes = [&](const std::vector<Active>& layer, bool root) {
    std::vector<ScanNode> sn(layer.size());
    for (size_t i = 0; i < layer.size(); ++i) {
      sn[i].slot = uint32_t(i);
      sn[i].node_index = layer[i].node;
      sn[i].depth = layer[i].depth;
      sn[i].mask_word0 = uint32_t(i * mask_words);
      sn[i].parent_from_hist = root ? (f.seq ? 2u : 1u) : 0u;
      sn[i].lo = layer[i].lo;
      sn[i].hi = layer[i].hi;
      sn[i].parent = layer[i].stats;
    }
    return sn;
  };

  std::vector<mpsb_tree_node> nodes(1);
  std::vector<LeafInfo> bounds(1, LeafInfo{0u, 0u, -kInf, kInf});  // per node, for leaf renewal
  std::vector<Active> layer(1);
  layer[0].count = n;
  {
    // The fit's first root is packed and gives the root's counts; later roots, when the form
    // takes them, are built without counts.
    const bool count_free = f.root_free && f.root_counts_ready;
    const Fit::LevelPlan root = f.plan_build({{0, 0, n, -1, 0}}, nullptr, hc, count_free);
    const ScanLaunch sc = f.plan_scan(scan_nodes(layer, true), hc, masks_for(layer));
    f.arena.flush();
    f.run_build(root);
    if (count_free) {
      ++f.stats.root_gradient_only_builds;
    } else if (f.root_free) {
      launch_copy_u32(f.root_count, root.r.out.cnt, f.d.cells, s.stream);  // the root's slot 0
      f.root_counts_ready = true;
    }
    if (f.seq) f.sum_rows(nullptr, 0, n);  // InitializeTree's SumRows (src/core/tree.cpp:37-47)
    f.run_scan(sc);
    f.resolve_level(sc, nullptr, 1, nullptr, nullptr);
  }
  std::vector<Candidate> cand = f.collect_scan(f.last_scan);
  layer[0].stats = cand[0].parent;
  nodes[0] = make_leaf(layer[0].stats, p, -kInf, kInf);

  uint32_t leaves = 1;
  const uint32_t cap = max_leaves_of(p);
  while (!layer.empty()) {
    const bool final = layer[0].depth >= p.max_depth || leaves >= cap;
    std::vector<PartNode> part;
    std::vector<Active> next;
    std::vector<Fit::Build> pending;
    for (size_t i = 0; i < layer.size(); ++i) {
      const Active& a = layer[i];
      PartNode pn{};
      pn.begin = a.begin;
      pn.count = a.count;
      pn.action = kPartLeaf;
      pn.leaf_node = a.node;
      pn.leaf_value = nodes[a.node].leaf_value;
      if (final || leaves >= cap || !cand[i].valid) {
        part.push_back(pn);
        continue;
      }
      const Candidate& c = cand[i];
      // ApplySplit, src/core/tree_structure.cpp:118-148: left then right appended.
      const uint32_t li = uint32_t(nodes.size());
      nodes.push_back(make_leaf(c.left, p, c.l_lo, c.l_hi));
      bounds.push_back(LeafInfo{li, 0u, c.l_lo, c.l_hi});
      const uint32_t ri = uint32_t(nodes.size());
      nodes.push_back(make_leaf(c.right, p, c.r_lo, c.r_hi));
      bounds.push_back(LeafInfo{ri, 0u, c.r_lo, c.r_hi});
      mpsb_tree_node& parent = nodes[a.node];
      parent.feature_index = c.feature;
      parent.threshold_bin = c.threshold;
      parent.left_child = li;
      parent.right_child = ri;
      parent.leaf_value = 0.0;
      parent.gain = c.gain;
      parent.default_left = c.default_left ? 1 : 0;
      parent.flags = 0;
      ++leaves;
      if (c.left.count + c.right.count != a.count)
        fail(MPSB_ERR_TRAINING, "Sample partition counts do not match histogram statistics");
      pn.action = kPartSplit;
      pn.feature = c.feature;
      pn.threshold = c.threshold;
      pn.default_left = c.default_left;
      pn.missing_code = f.d.has_missing(c.feature) ? f.d.bin_count[c.feature] : 0xFFFFFFFFu;
      pn.left_count = uint32_t(c.left.count);
      part.push_back(pn);
      std::vector<uint32_t> path = a.path;
      if (std::find(path.begin(), path.end(), c.feature) == path.end()) path.push_back(c.feature);
      Active l, r;
      l.node = li;
      r.node = ri;
      l.depth = r.depth = a.depth + 1;
      l.begin = a.begin;
      l.count = uint32_t(c.left.count);
      r.begin = a.begin + l.count;
      r.count = uint32_t(c.right.count);
      l.stats = c.left;
      r.stats = c.right;
      l.lo = c.l_lo;
      l.hi = c.l_hi;
      r.lo = c.r_lo;
      r.hi = c.r_hi;
      l.path = r.path = path;
      next.push_back(std::move(l));
      next.push_back(std::move(r));
      if (a.depth + 1 < p.max_depth) {
        // The smaller child is built, the left one on ties (src/core/tree.cpp:109-117).
        const bool left_first = c.left.count <= c.right.count;
        const uint32_t built = uint32_t(next.size() - (left_first ? 2 : 1));
        pending.push_back({built, next[built].begin, next[built].count, int32_t(built ^ 1u), uint32_t(i),
                           next[built ^ 1u].begin, next[built ^ 1u].count});
      }
    }
    // When every child of this level is a leaf (depth limit or the leaf cap), the rows take
    // their leaves' values in this partition: no scatter and no separate leaf pass.
    if (!final && (next.empty() || next[0].depth >= p.max_depth || leaves >= cap)) {
      size_t k = 0;
      for (PartNode& pn : part) {
        if (pn.action != kPartSplit) continue;
        pn.action = kPartSplitLeaf;
        pn.leaf_node = next[k].node;
        pn.leaf_node_right = next[k + 1].node;
        pn.leaf_value = nodes[next[k].node].leaf_value;
        pn.leaf_right = nodes[next[k + 1].node].leaf_value;
        k += 2;
      }
      next.clear();
      pending.clear();
    }
    // The partition, the next level's histograms and its split scan are planned from the
    // candidates alone, so their descriptors reach the device in one copy and a level costs
    // one synchronisation: the read-back of its candidates.
    f.gh_cur = gh[rc ^ 1];  // gradients in the order the partition writes
    const PartLaunch pa = f.plan_partition(part, rows[rc], rows[rc ^ 1], pred, stride, offset, gh[rc], gh[rc ^ 1]);
    const bool more = !(final || next.empty());
    Fit::LevelPlan lp;
    bool do_scan = false;
    const LayerNode* segments = nullptr;  // the next level's row segments (near-tie resolution)
    if (more) {
      if (!pending.empty()) lp = f.plan_build(pending, rows[rc ^ 1], hc ^ 1);
      if (next[0].depth < p.max_depth && leaves < cap) {
        f.plan_scan(scan_nodes(next, false), pending.empty() ? hc : hc ^ 1, masks_for(next));
        do_scan = true;
        if (f.last_scan.contend) {
          std::vector<LayerNode> seg(next.size());
          for (size_t i = 0; i < next.size(); ++i) {
            seg[i].node = next[i].node;
            seg[i].begin = next[i].begin;
            seg[i].count = next[i].count;
          }
          segments = f.arena.put(seg.data(), seg.size());
        }
      }
    }
    if (!more && f.leaf_node) {  // this tree's last partition: publish its leaf values
      std::vector<double> table(nodes.size());
      for (size_t i = 0; i < nodes.size(); ++i) table[i] = nodes[i].leaf_value;
      f.leaf_table = f.arena.put(table.data(), table.size());
      if (f.renew) {
        f.leaf_info = f.arena.put(bounds.data(), bounds.size());
        f.leaf_ids = uint32_t(nodes.size());
      }
    }
    f.arena.flush();
    f.run_partition(pa);
    rc ^= 1;
    if (!more) break;
    f.run_build(lp);
    if (!pending.empty()) hc ^= 1;
    layer = std::move(next);
    if (do_scan) {
      f.run_scan(f.last_scan);
      f.resolve_level(f.last_scan, segments, uint32_t(layer.size()), nullptr, rows[rc]);
      cand = f.collect_scan(f.last_scan);
    }
  }
  // No synchronisation here: the next tree's first level signal (or the fit's final sync)
  // orders everything, and releases the descriptor arena and the stage timers.
  f.gh_cur = nullptr;
  f.gh_pos = false;
  return nodes;
}

// =================================================================== boosting =========
static double weighted_total(const double* w, uint64_t n) {  // src/core/trainer.cpp:44-66
  // Unit weights add up to n exactly (n < 2^53), the loop's result without the loop.
  if (!w) {
    if (n == 0) fail(MPSB_ERR_TRAINING, "sample weights must contain positive total weight");
    return double(n);
  }
  double t = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    const double x = w ? w[i] : 1.0;
    if (!std::isfinite(x) || x < 0.0) fail(MPSB_ERR_TRAINING, "sample weights must be finite non-negative values");
    t += x;
    if (!std::isfinite(t)) fail(MPSB_ERR_TRAINING, "sample weight sum overflowed");
  }
  if (t <= 0.0) fail(MPSB_ERR_TRAINING, "sample weights must contain positive total weight");
  return t;
}

// The model keeps the schema only: boundaries and layout, never the code matrix.
static HostDataset schema_copy(const HostDataset& d) {
  HostDataset s;
  s.rows = 0;
  s.features = d.features;
  s.max_bins = d.max_bins;
  s.boundaries = d.boundaries;
  s.boundary_offset = d.boundary_offset;
  s.bin_count = d.bin_count;
  s.missing_count = d.missing_count;
  s.code_bytes = d.code_bytes;
  s.row_stride = d.row_stride;
  s.cell_offset = d.cell_offset;
  s.cells = d.cells;
  s.uid = d.uid;
  s.miss_flag = d.miss_flag;
  return s;
}

// InitialBaseScore, src/core/trainer.cpp:68-176, with the reference's checks and messages.
static double initial_base_score(const double* y, const double* w, uint64_t n, const Params& p) {
  auto wt = [w](uint64_t i) { return w ? w[i] : 1.0; };
  switch (p.objective) {
    case 0: {  // WeightedMeanLabel :68-86
      const double tw = weighted_total(w, n);
      double sum = 0.0;
      for (uint64_t i = 0; i < n; ++i) {
        if (!std::isfinite(y[i])) fail(MPSB_ERR_TRAINING, "Labels must be finite");
        sum += y[i] * wt(i);
        if (!std::isfinite(sum)) fail(MPSB_ERR_TRAINING, "Label accumulation overflowed");
      }
      return sum / tw;
    }
    case 1: {  // WeightedBinaryLogitBaseScore :88-110
      const double tw = weighted_total(w, n);
      double pos = 0.0;
      if (!w) {
        // Unit weights: the positive total is the count of ones, exactly (n < 2^53), so the
        // loop needs no serial floating-point chain and no branch per label.
        uint64_t ones = 0, bad = 0;
        for (uint64_t i = 0; i < n; ++i) {
          const double v = y[i];
          ones += v == 1.0 ? 1u : 0u;
          bad += (v != 0.0 && v != 1.0) ? 1u : 0u;
        }
        if (bad) fail(MPSB_ERR_TRAINING, "binary-logistic labels must be 0 or 1");
        pos = double(ones);
      }
      for (uint64_t i = 0; w && i < n; ++i) {
        if (y[i] == 0.0) continue;
        if (y[i] == 1.0) {
          pos += wt(i);
          continue;
        }
        fail(MPSB_ERR_TRAINING, "binary-logistic labels must be 0 or 1");
      }
      const double eps = 1e-15;
      const double q = std::min(1.0 - eps, std::max(eps, pos / tw));
      return std::log(q / (1.0 - q));
    }
    case 2: {  // WeightedQuantileBaseScore :112-138
      const double tw = weighted_total(w, n);
      std::vector<std::pair<double, double>> v;
      v.reserve(n);
      for (uint64_t i = 0; i < n; ++i) {
        if (!std::isfinite(y[i])) fail(MPSB_ERR_TRAINING, "label must be finite");
        v.push_back({y[i], wt(i)});
      }
      std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      const double threshold = p.objective_alpha * tw;
      double cum = 0.0;
      for (const auto& it : v) {
        cum += it.second;
        if (cum >= threshold) return it.first;
      }
      return v.back().first;
    }
    case 3:
    case 4: {  // WeightedLogMeanBaseScore :140-157
      const std::string name = p.objective == 3 ? "poisson" : "tweedie";
      const double tw = weighted_total(w, n);
      double sum = 0.0;
      for (uint64_t i = 0; i < n; ++i) {
        if (!std::isfinite(y[i]) || y[i] < 0.0) fail(MPSB_ERR_TRAINING, name + " labels must be non-negative");
        sum += y[i] * wt(i);
        if (!std::isfinite(sum)) fail(MPSB_ERR_TRAINING, name + " label sum overflowed");
      }
      return std::log(std::max(1e-15, sum / tw));
    }
  }
  fail(MPSB_ERR_TRAINING, "unknown training objective");
}

// TrainRegressionModel, src/core/trainer.cpp:226-266: every objective's gradients on the device.
Model train_regressor(Session& s, const HostDataset& d, const double* y, const double* w, const Params& p) {
  validate_train_params(p, d);
  if (d.rows == 0) fail(MPSB_ERR_TRAINING, "Labels must not be empty");
  // The base score's serial pass over the labels runs on a host thread while the device is
  // set up and the labels are uploaded; its errors surface before any tree is grown.
  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  double base_seconds = 0.0;
  std::future<double> base_score = std::async(std::launch::async, [&] {
    const auto t0 = clock::now();
    const double b = initial_base_score(y, w, d.rows, p);
    base_seconds = std::chrono::duration<double>(clock::now() - t0).count();
    return b;
  });
  struct Join {  // never leave the thread reading y / w behind an early exit
    std::future<double>& f;
    ~Join() {
      if (f.valid()) f.wait();
    }
  } join{base_score};
  Model m;
  m.features = d.features;
  m.precision = p.precision;
  m.lr = p.lr;
  m.objective = p.objective;
  m.objective_alpha = p.objective_alpha;
  m.tweedie_power = p.tweedie_power;
  m.schema = schema_copy(d);
  // Unweighted squared error and quantile gradients have h = 1 exactly: the histograms then
  // carry g only and the Hessian sums are the counts.
  Fit f(s, d, p, p.precision, w == nullptr && (p.objective == 0 || p.objective == 2));
  double* labels = dptr<double>(s, kBufLabels, d.rows);
  double* pred = dptr<double>(s, kBufPred, d.rows);
  double* weights = w ? dptr<double>(s, kBufWeights, d.rows) : nullptr;
  h2d(s, labels, y, d.rows * 8, "label upload");
  if (w) h2d(s, weights, w, d.rows * 8, "weight upload");
  m.base = {base_score.get()};
  launch_fill_f64(pred, m.base[0], d.rows, 1, 0, s.stream);
  // Node ids must fit two bytes for the deferred update; deeper trees update pred in place.
  const uint64_t max_nodes = 2 * std::min<uint64_t>(max_leaves_of(p), uint64_t(1) << std::min(p.max_depth, 31u)) - 1;
  if (max_nodes <= 65535) f.leaf_node = dptr<uint16_t>(s, kBufLeafAdd, d.rows);
  f.resolve_prepare();
  f.device_plan_setup(p.n_estimators);
  f.resolve_setup();
  f.renew_setup(p.n_estimators, false);
  std::vector<size_t> planned;  // model trees grown with device planning
  const auto t_loop = clock::now();
  for (uint32_t round = 0; round < p.n_estimators; ++round) {
    GradLaunch g{};
    g.source = 0;
    g.labels = labels;
    g.pred = pred;
    if (round > 0 && f.leaf_node) {  // every row received exactly one leaf in the previous round
      g.leaf_node = f.leaf_node;
      g.leaf_bytes = f.leaf_id_bytes();
      g.leaf_table = f.leaf_table;
      g.lr = p.lr;
      g.pred_out = pred;
    }
    g.weights = weights;
    g.n = d.rows;
    f.timer.on = round % kTimeEvery == 0;
    f.stats.timed_trees += f.timer.on ? 1 : 0;
    f.gradients(g);
    f.last_tree_device = f.dp.on;
    if (f.dp.on) {
      f.grow_tree_device(pred, 1, 0);
      planned.push_back(m.trees.size());
      m.trees.emplace_back();
    } else {
      m.trees.push_back(grow_tree(f, pred, 1, 0));
    }
    f.renew_leaves(g, m.trees.size() - 1);
    ++f.stats.trees;
    f.seed += 0x9E3779B97F4A7C15ull;  // a fresh stochastic-rounding stream per round
  }
  if (p.n_estimators > 0)
    launch_final_update(pred, d.rows, 1, 0, f.leaf_node, f.leaf_id_bytes(), f.leaf_table, p.lr, f.err, s.stream);
  f.finish();
  const auto t_drained = clock::now();
  f.collect_device_trees(m.trees, planned);
  f.collect_resolution();
  f.collect_renewals(m.trees);
  f.stats.device_planned_trees = planned.size();
  const auto t_end = clock::now();
  f.stats.train_seconds = std::chrono::duration<double>(t_end - t_start).count();
  f.stats.setup_seconds = std::chrono::duration<double>(t_loop - t_start).count();
  f.stats.finish_seconds = std::chrono::duration<double>(t_end - t_drained).count();
  f.stats.base_score_seconds = base_seconds;
  m.stats = f.stats;
  s.timing.histogram_command_seconds = f.stats.histogram_seconds;
  s.timing.hot_path_command_seconds = f.stats.scan_seconds + f.stats.partition_seconds;
  s.timing.gradient_seconds = f.stats.gradient_seconds;
  return m;
}

// Batched softmax rounds (mpsb_train_params::class_rounds = 1; no counterpart in the
// reference): per round, every class's gradients from the round's starting margins, then the K
// class trees grown together as one device-planned tree of K roots (one launch per level and
// kernel for all of them), each class's rows in its own block of positions; the read-back splits
// the round's tree into its class trees, stored in the reference's round-major class order.
static Model train_multiclass_batched(Session& s, const HostDataset& d, Fit& f, Model& m, const double* labels,
                                      double* margins, const double* weights, uint32_t k, const Params& p,
                                      std::chrono::steady_clock::time_point t_start) {
  if (p.precision != MPSB_PRECISION_FP64_FIXED && p.precision != MPSB_PRECISION_INT16 &&
      p.precision != MPSB_PRECISION_INT8)
    fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds need the fp64_fixed, int16 or int8 form");
  if (p.near_ties == 1) fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds keep near ties (near_ties must not be resolve)");
  if (!p.groups.empty()) fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds do not take interaction constraints");
  if (p.max_depth < 1 || p.max_depth > 12 || (uint64_t(k) << (p.max_depth - 1)) > 2048)
    fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds need class_count x 2^(max_depth - 1) <= 2048");
  if (max_leaves_of(p) < (1u << p.max_depth))
    fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds need a leaf cap of at least 2^max_depth");
  if (uint64_t(k) * d.rows >= (uint64_t(1) << 32))
    fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds need class_count x rows below 2^32");
  f.roots = k;
  f.leaf_node = dptr<uint16_t>(s, kBufLeafAdd, f.positions());  // leaf ids [K][rows]
  f.noise_k = 0.0;  // near ties kept
  f.device_plan_setup(p.n_estimators);
  if (!f.dp.on) fail(MPSB_ERR_INVALID_ARGUMENT, "batched class rounds need the device-side level planner");
  f.resolve_setup();
  f.renew_setup(p.n_estimators, true);
  std::vector<size_t> planned;  // first model tree of each round
  for (uint32_t round = 0; round < p.n_estimators; ++round) {
    GradLaunch g{};
    g.labels = labels;
    g.pred = margins;
    g.weights = weights;
    g.classes = k;
    g.n = d.rows;
    if (round > 0) {  // every class column receives its previous tree's leaf first
      g.leaf_node = f.leaf_node;
      g.leaf_bytes = f.leaf_id_bytes();
      g.leaf_table = f.leaf_table;
      g.lr = p.lr;
      g.pred_out = margins;
    }
    f.timer.on = round % kTimeEvery == 0;
    f.stats.timed_trees += f.timer.on ? k : 0;
    const GradLaunch w = f.gradients_batched(g);
    f.last_tree_device = true;
    f.grow_tree_device(margins, 1, 0);
    planned.push_back(m.trees.size());
    for (uint32_t c = 0; c < k; ++c) m.trees.emplace_back();
    f.renew_leaves(w, m.trees.size() - 1);
    f.stats.trees += k;
    f.seed += 0x9E3779B97F4A7C15ull;
  }
  if (p.n_estimators > 0) {
    const uint32_t lb = f.leaf_id_bytes();
    for (uint32_t c = 0; c < k; ++c) {
      const auto* ids = reinterpret_cast<const uint16_t*>(reinterpret_cast<const uint8_t*>(f.leaf_node) +
                                                          size_t(c) * d.rows * lb);
      launch_final_update(margins + size_t(c) * d.rows, d.rows, 1, 0, ids, lb, f.leaf_table, p.lr, f.err, s.stream);
    }
  }
  f.finish();
  f.collect_device_trees(m.trees, planned);
  f.collect_renewals(m.trees);
  f.stats.device_planned_trees = planned.size() * k;
  f.stats.train_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  m.stats = f.stats;
  return m;
}

// TrainMulticlassSoftmaxModel, src/core/multiclass_trainer.cpp:104-149
Model train_multiclass(Session& s, const HostDataset& d, const double* y, const double* w, uint32_t k,
                       const Params& p) {
  validate_train_params(p, d);
  if (k < 2) fail(MPSB_ERR_TRAINING, "softmax class_count must be at least two");
  if (k > 256) fail(MPSB_ERR_INVALID_ARGUMENT, "at most 256 classes are supported");
  const double tw = weighted_total(w, d.rows);
  std::vector<double> cw(k, 0.0);
  for (uint64_t i = 0; i < d.rows; ++i) {
    if (!std::isfinite(y[i])) fail(MPSB_ERR_TRAINING, "softmax labels must be finite");
    if (std::floor(y[i]) != y[i] || y[i] < 0.0 || y[i] >= double(k))
      fail(MPSB_ERR_TRAINING, "softmax labels must be encoded class ids");
  }
  for (uint64_t i = 0; i < d.rows; ++i) {
    cw[uint32_t(y[i])] += w ? w[i] : 1.0;
    if (!std::isfinite(cw[uint32_t(y[i])])) fail(MPSB_ERR_TRAINING, "softmax class prior overflowed");
  }
  const auto t_start = std::chrono::steady_clock::now();
  Model m;
  m.features = d.features;
  m.classes = k;
  m.precision = p.precision;
  m.lr = p.lr;
  for (double c : cw) {  // WeightedClassBaseScores, src/core/multiclass_trainer.cpp:39-67
    const double sc = std::log(std::max(1e-15, c / tw));
    if (!std::isfinite(sc)) fail(MPSB_ERR_TRAINING, "softmax base score is not finite");
    m.base.push_back(sc);
  }
  for (uint32_t c = 0; c < k; ++c) m.class_labels.push_back(double(c));  // multiclass_trainer.cpp:124-127
  m.schema = schema_copy(d);
  Fit f(s, d, p, p.precision, false);
  f.softmax = true;
  double* labels = dptr<double>(s, kBufLabels, d.rows);
  double* margins = dptr<double>(s, kBufPred, d.rows * k);
  double* weights = w ? dptr<double>(s, kBufWeights, d.rows) : nullptr;
  h2d(s, labels, y, d.rows * 8, "label upload");
  if (w) h2d(s, weights, w, d.rows * 8, "weight upload");
  // Margins class-major, [k][rows]: each class column is contiguous.
  for (uint32_t c = 0; c < k; ++c) launch_fill_f64(margins + size_t(c) * d.rows, m.base[c], d.rows, 1, 0, s.stream);
  // Deferred margin update, as for regression: a tree's last partition writes each row's leaf
  // node id and the next class tree's softmax pass adds lr * leaf to the tree's column first.
  const uint64_t max_nodes = 2 * std::min<uint64_t>(max_leaves_of(p), uint64_t(1) << std::min(p.max_depth, 31u)) - 1;
  if (p.class_rounds == 1) return train_multiclass_batched(s, d, f, m, labels, margins, weights, k, p, t_start);
  if (max_nodes <= 65535) f.leaf_node = dptr<uint16_t>(s, kBufLeafAdd, d.rows);
  // Margins beyond 4 GiB take 64-bit addressed loads (MPSB_WIDE_MARGINS=1 forces them).
  const char* wide_env = std::getenv("MPSB_WIDE_MARGINS");
  const bool wide = uint64_t(d.rows) * k > (uint64_t(1) << 29) || (wide_env && wide_env[0] == '1');
  f.resolve_prepare();
  f.device_plan_setup(uint64_t(p.n_estimators) * k);
  f.resolve_setup();
  f.renew_setup(uint64_t(p.n_estimators) * k, true);
  std::vector<size_t> planned;
  for (uint32_t round = 0; round < p.n_estimators; ++round) {
    for (uint32_t c = 0; c < k; ++c) {
      GradLaunch g{};
      g.source = 2;
      g.labels = labels;
      g.pred = margins;
      g.weights = weights;
      g.classes = k;
      g.target = c;
      g.n = d.rows;
      g.wide_margins = wide ? 1u : 0u;
      if ((round > 0 || c > 0) && f.leaf_node) {  // the previous class tree's update
        g.leaf_node = f.leaf_node;
        g.leaf_bytes = f.leaf_id_bytes();
        g.leaf_table = f.leaf_table;
        g.lr = p.lr;
        g.pred_out = margins;
        g.pending_class = (c + k - 1) % k;
      }
      f.timer.on = f.stats.trees % kTimeEvery == 0;
      f.stats.timed_trees += f.timer.on ? 1 : 0;
      f.gradients(g);
      f.last_tree_device = f.dp.on;
      if (f.dp.on) {
        f.grow_tree_device(margins + size_t(c) * d.rows, 1, 0);
        planned.push_back(m.trees.size());
        m.trees.emplace_back();
      } else {
        m.trees.push_back(grow_tree(f, margins + size_t(c) * d.rows, 1, 0));
      }
      f.renew_leaves(g, m.trees.size() - 1);
      ++f.stats.trees;
      f.seed += 0x9E3779B97F4A7C15ull;
    }
  }
  if (p.n_estimators > 0)
    launch_final_update(margins + size_t(k - 1) * d.rows, d.rows, 1, 0, f.leaf_node, f.leaf_id_bytes(), f.leaf_table,
                        p.lr, f.err, s.stream);
  f.finish();
  f.collect_device_trees(m.trees, planned);
  f.collect_resolution();
  f.collect_renewals(m.trees);
  f.stats.device_planned_trees = planned.size();
  f.stats.train_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  m.stats = f.stats;
  return m;
}

// RegressionModel::Predict / MulticlassModel::PredictMargins on the device.
void predict(Session& s, const Model& m, const HostDataset& d, double* out) {
  if (d.features != m.features || d.max_bins != m.schema.max_bins || d.boundaries != m.schema.boundaries)
    fail(MPSB_ERR_TRAINING, "Prediction data does not use the model's frozen binned schema");
  s.check_fault(kFaultPredict, "prediction");
  s.upload(d);
  std::vector<mpsb_tree_node> all;
  std::vector<uint32_t> first{0};
  for (const auto& t : m.trees) {
    all.insert(all.end(), t.begin(), t.end());
    first.push_back(uint32_t(all.size()));
  }
  auto* nodes = dptr<mpsb_tree_node>(s, kBufTrees, all.size() + 1);
  auto* tf = dptr<uint32_t>(s, kBufTmpA, first.size());
  h2d(s, nodes, all.data(), all.size() * sizeof(mpsb_tree_node), "tree upload");
  h2d(s, tf, first.data(), first.size() * 4, "tree upload");
  auto* o = dptr<double>(s, kBufTmpB, d.rows * m.classes);
  for (uint32_t c = 0; c < m.classes; ++c) launch_fill_f64(o, m.base[c], d.rows, m.classes, c, s.stream);
  auto* err = dptr<uint32_t>(s, kBufErr, 4);
  MPSB_HIP("prediction", hipMemsetAsync(err, 0, 4, s.stream));
  PredictLaunch a{};
  a.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
  a.row_stride = d.row_stride;
  a.code_bytes = d.code_bytes;
  a.rows = d.rows;
  a.nodes = nodes;
  a.tree_first = tf;
  a.miss_code = static_cast<const uint32_t*>(s.ptr[kBufMissCode]);
  a.n_trees = uint32_t(m.trees.size());
  a.classes = m.classes;
  a.lr = m.lr;
  a.out = o;
  a.err = err;
  // Rows staged per block: up to 256, within 64 KB of LDS at the padded row pitch.
  const uint32_t pitch = d.row_stride * d.code_bytes + 4;
  if (pitch > 65536) fail(MPSB_ERR_INVALID_ARGUMENT, "a row of codes does not fit 64 KB of LDS for prediction");
  a.rows_per_block = std::min<uint32_t>(256, 65536 / pitch);
  launch_predict(a, s.stream);
  uint32_t flags = 0;
  d2h(s, out, o, d.rows * m.classes * 8, "prediction");
  d2h(s, &flags, err, 4, "prediction");
  s.sync("prediction");
  if (flags) fail(MPSB_ERR_TRAINING, "Tree node index is out of bounds or the structure contains a cycle");
}

// ================================================================== test hooks =======
void hook_squared_error(Session& s, const double* y, const double* pr, uint64_t n, double* out) {
  if (n == 0) fail(MPSB_ERR_TRAINING, "Labels must not be empty");
  s.check_fault(kFaultGradients, "gradients");
  auto* dy = dptr<double>(s, kBufTmpA, n);
  auto* dp = dptr<double>(s, kBufTmpB, n);
  auto* o = dptr<double>(s, kBufGH64, 2 * n);
  auto* scratch = s.get(kBufGH, n * 16);
  auto* err = dptr<uint32_t>(s, kBufErr, 4);
  MPSB_HIP("gradients", hipMemsetAsync(err, 0, 4, s.stream));
  h2d(s, dy, y, n * 8, "gradients");
  h2d(s, dp, pr, n * 8, "gradients");
  GradLaunch g{};
  g.source = 0;
  g.labels = dy;
  g.pred = dp;
  g.n = n;
  g.precision = MPSB_PRECISION_FP64;
  g.gh = scratch;
  g.gh64 = o;
  g.unit_h = 1;
  g.scale_g = 1.0;
  g.err = err;
  hipEvent_t a = s.event(0), b = s.event(1);
  MPSB_HIP("gradients", hipEventRecord(a, s.stream));
  launch_grad_write(g, s.stream);
  MPSB_HIP("gradients", hipEventRecord(b, s.stream));
  uint32_t flags = 0;
  d2h(s, out, o, n * 16, "gradients");
  d2h(s, &flags, err, 4, "gradients");
  s.sync("gradients");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  s.timing.gradient_seconds = ms * 1e-3;
  if (flags) fail(MPSB_ERR_TRAINING, "label and prediction must be finite");
}

// Uploads rows (as device uint32) and FP64 pairs, prepares g/h in `precision` for a fit-like
// context; returns the device row pointer.
static uint32_t* upload_rows(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n) {
  std::vector<uint32_t> r(n);
  for (uint64_t i = 0; i < n; ++i) {
    if (rows[i] >= d.rows) fail(MPSB_ERR_TRAINING, "Histogram row index is out of bounds");
    r[i] = uint32_t(rows[i]);
  }
  auto* dr = dptr<uint32_t>(s, kBufRowsA, std::max<uint64_t>(n, 1));
  h2d(s, dr, r.data(), n * 4, "row upload");
  s.sync("row upload");
  return dr;
}

static bool all_unit_h(const double* pairs, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i)
    if (pairs[2 * i + 1] != 1.0) return false;
  return true;
}

static void pairs_to_device(Fit& f, const double* pairs, uint64_t n) {
  auto* dp = dptr<double>(f.s, kBufPairs, 2 * n);
  h2d(f.s, dp, pairs, n * 16, "gradient upload");
  GradLaunch g{};
  g.source = 1;
  g.pairs = dp;
  g.n = n;
  f.gradients(g);
}

// Reference view of node histograms: bin 0 includes the missing rows.
static void to_reference_view(Fit& f, uint32_t n_nodes, uint64_t* cnt, double* g, double* h) {
  const HostDataset& d = f.d;
  const size_t cells = size_t(n_nodes) * d.cells;
  std::vector<uint32_t> c(cells);
  std::vector<uint64_t> gv(cells), hv(cells);
  d2h(f.s, c.data(), f.hist[0].cnt, cells * 4, "histogram readback");
  d2h(f.s, gv.data(), f.hist[0].g, cells * 8, "histogram readback");
  d2h(f.s, hv.data(), f.hist[0].h, cells * 8, "histogram readback");
  uint32_t flags = 0;
  d2h(f.s, &flags, f.err, 4, "histogram readback");
  f.s.sync("histogram");
  f.arena.reset();
  f.timer.resolve();
  f.check_errors(flags);
  f.read_scales();
  const bool integer = f.kind != kHistF32 && f.kind != kHistSeq;
  const bool merge = f.kind != kHistSeq;  // the row-order form already keeps missing rows in bin 0
  auto G = [&](size_t i) {
    if (integer) return double(int64_t(gv[i])) * f.scale_g;
    double x;
    std::memcpy(&x, &gv[i], 8);
    return x;
  };
  auto Hh = [&](size_t i) {
    if (f.unit_h) return double(c[i]);
    if (integer) return double(int64_t(hv[i])) * f.scale_h;
    double x;
    std::memcpy(&x, &hv[i], 8);
    return x;
  };
  size_t o = 0;
  for (uint32_t nd = 0; nd < n_nodes; ++nd) {
    const size_t base = size_t(nd) * d.cells;
    for (uint32_t ft = 0; ft < d.features; ++ft) {
      for (uint32_t b = 0; b < d.bin_count[ft]; ++b, ++o) {
        const size_t at = base + d.cell_offset[ft] + b;
        cnt[o] = c[at];
        g[o] = G(at);
        h[o] = Hh(at);
        if (merge && b == 0 && d.has_missing(ft)) {
          const size_t m = base + d.cell_offset[ft] + d.bin_count[ft];
          cnt[o] += c[m];
          g[o] += G(m);
          h[o] += Hh(m);
        }
      }
    }
  }
}

void hook_layer_histograms(Session& s, const HostDataset& d, uint32_t n_nodes, const uint64_t* offsets,
                           const uint64_t* rows, const double* pairs, int precision, uint64_t* cnt,
                           double* g, double* h) {
  if (precision < 0 || precision > MPSB_PRECISION_FP64_FIXED) fail(MPSB_ERR_INVALID_ARGUMENT, "unknown histogram precision");
  if (n_nodes == 0) return;
  for (uint32_t i = 0; i < n_nodes; ++i) {
    if (offsets[i + 1] < offsets[i]) fail(MPSB_ERR_INVALID_ARGUMENT, "node offsets must be non-decreasing");
    if (offsets[i + 1] == offsets[i]) fail(MPSB_ERR_TRAINING, "Node row set must not be empty");
  }
  Params p;
  p.max_bins = d.max_bins;
  Fit f(s, d, p, precision, all_unit_h(pairs, d.rows));
  uint32_t* dr = upload_rows(s, d, rows, offsets[n_nodes]);
  pairs_to_device(f, pairs, d.rows);
  std::vector<Fit::Build> builds;
  for (uint32_t i = 0; i < n_nodes; ++i)
    builds.push_back({i, uint32_t(offsets[i]), uint32_t(offsets[i + 1] - offsets[i]), -1, 0});
  f.build_level(builds, dr, 0);
  to_reference_view(f, n_nodes, cnt, g, h);
  s.timing.histogram_command_seconds = f.stats.histogram_seconds + f.stats.reduce_seconds;
}

void hook_baseline_histograms(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n,
                              const double* pairs, uint64_t* cnt, double* g, double* h) {
  if (n == 0) fail(MPSB_ERR_TRAINING, "Node row set must not be empty");
  s.check_fault(kFaultHistogram, "histogram");
  s.upload(d);
  uint32_t* dr = upload_rows(s, d, rows, n);
  auto* dp = dptr<double>(s, kBufPairs, 2 * d.rows);
  h2d(s, dp, pairs, d.rows * 16, "gradient upload");
  // Reference-view cells: (feature, bin), bin 0 also takes the missing code.
  std::vector<uint32_t> cf, cc;
  for (uint32_t ft = 0; ft < d.features; ++ft)
    for (uint32_t b = 0; b < d.bin_count[ft]; ++b) {
      cf.push_back(ft);
      const uint32_t miss = (b == 0 && d.has_missing(ft)) ? d.bin_count[ft] : 0xFFFFu;
      cc.push_back(b | (miss << 16));
    }
  const uint32_t cells = uint32_t(cf.size());
  auto* dcf = dptr<uint32_t>(s, kBufTmpA, cells);
  auto* dcc = dptr<uint32_t>(s, kBufTmpB, cells);
  h2d(s, dcf, cf.data(), cells * 4, "cell table upload");
  h2d(s, dcc, cc.data(), cells * 4, "cell table upload");
  uint8_t* out = static_cast<uint8_t*>(s.get(kBufTmpC, size_t(cells) * 24));
  launch_baseline_hist(static_cast<const uint8_t*>(s.ptr[kBufCodes]), d.row_stride, d.code_bytes, dr, uint32_t(n),
                       dp, dcf, dcc, cells, reinterpret_cast<uint64_t*>(out),
                       reinterpret_cast<double*>(out + size_t(cells) * 8),
                       reinterpret_cast<double*>(out + size_t(cells) * 16), s.stream);
  d2h(s, cnt, out, size_t(cells) * 8, "histogram readback");
  d2h(s, g, out + size_t(cells) * 8, size_t(cells) * 8, "histogram readback");
  d2h(s, h, out + size_t(cells) * 16, size_t(cells) * 8, "histogram readback");
  s.sync("baseline histogram");
}

static mpsb_split_candidate to_public(const Candidate& c) {
  mpsb_split_candidate o{};
  o.valid = c.valid;
  o.feature = c.feature;
  o.threshold_bin = c.threshold;
  o.default_left = c.default_left;
  o.left_count = c.left.count;
  o.right_count = c.right.count;
  o.left_gradient_sum = c.left.g;
  o.left_hessian_sum = c.left.h;
  o.right_gradient_sum = c.right.g;
  o.right_hessian_sum = c.right.h;
  o.gain = c.gain;
  return o;
}

void hook_scan(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n, const double* pairs,
               const mpsb_split_params& sp, int precision, mpsb_split_candidate* out, mpsb_split_candidate* best) {
  if (n == 0) fail(MPSB_ERR_TRAINING, "Node row set must not be empty");
  if (precision < 0 || precision > MPSB_PRECISION_FP64_FIXED) fail(MPSB_ERR_INVALID_ARGUMENT, "unknown histogram precision");
  Params p;
  p.max_bins = d.max_bins;
  p.min_samples_leaf = sp.min_samples_leaf;
  p.min_child_weight = sp.min_child_weight;
  p.lambda = sp.reg_lambda;
  p.alpha = sp.reg_alpha;
  p.gamma = sp.gamma;
  p.min_gain = sp.min_gain_to_split;
  if (p.min_samples_leaf == 0) fail(MPSB_ERR_TRAINING, "min_samples_leaf must be at least 1");
  Fit f(s, d, p, precision, all_unit_h(pairs, d.rows));
  uint32_t* dr = upload_rows(s, d, rows, n);
  pairs_to_device(f, pairs, d.rows);
  f.build_level({{0, 0, uint32_t(n), -1, 0}}, dr, 0);
  if (f.seq) f.sum_rows(dr, 0, uint32_t(n));
  std::vector<ScanNode> sn(1);
  sn[0].parent_from_hist = f.seq ? 2 : 1;
  sn[0].lo = -kInf;
  sn[0].hi = kInf;
  std::vector<Candidate> per(d.features);
  const std::vector<Candidate> c = f.scan(sn, 0, {}, per.data());
  for (uint32_t ft = 0; ft < d.features; ++ft) out[ft] = to_public(per[ft]);
  if (best) *best = to_public(c[0]);
}

void hook_partition(Session& s, const HostDataset& d, const uint64_t* rows, uint64_t n, uint32_t feature,
                    uint32_t threshold, int default_left, uint64_t* left, uint64_t* n_left, uint64_t* right,
                    uint64_t* n_right) {
  if (feature >= d.features) fail(MPSB_ERR_INVALID_ARGUMENT, "partition feature is out of range");
  s.check_fault(kFaultPartition, "partition");
  s.upload(d);
  *n_left = *n_right = 0;
  if (n == 0) return;
  uint32_t* din = upload_rows(s, d, rows, n);
  uint32_t* dout = dptr<uint32_t>(s, kBufRowsB, n);
  auto* err = dptr<uint32_t>(s, kBufErr, 4);
  MPSB_HIP("partition", hipMemsetAsync(err, 0, 4, s.stream));
  PartNode pn{};
  pn.begin = 0;
  pn.count = uint32_t(n);
  pn.action = kPartSplit;
  pn.feature = feature;
  pn.threshold = threshold;
  pn.default_left = default_left ? 1u : 0u;
  pn.missing_code = d.has_missing(feature) ? d.bin_count[feature] : 0xFFFFFFFFu;
  const uint32_t chunks = uint32_t((n + kPartChunk - 1) / kPartChunk);
  std::vector<uint32_t> chunk_node(chunks, 0);
  auto* dn = dptr<PartNode>(s, kBufTmpA, 1);
  auto* dcn = dptr<uint32_t>(s, kBufTmpB, chunks);
  h2d(s, dn, &pn, sizeof(pn), "partition");
  h2d(s, dcn, chunk_node.data(), chunks * 4, "partition");
  pn.column_major = uint64_t(n) * 128 >= d.rows ? 1u : 0u;
  h2d(s, dn, &pn, sizeof(pn), "partition");
  PartLaunch a{};
  a.codes = static_cast<const uint8_t*>(s.ptr[kBufCodes]);
  a.codes_cm = static_cast<const uint8_t*>(s.ptr[kBufCodesCM]);
  a.rows_total = d.rows;
  a.row_stride = d.row_stride;
  a.code_bytes = d.code_bytes;
  a.rows_in = din;
  a.rows_out = dout;
  a.nodes = dn;
  a.n_nodes = 1;
  a.chunk_node = dcn;
  a.n_chunks = chunks;
  a.chunk_left = dptr<uint32_t>(s, kBufChunkLeft, chunks);
  a.bits = dptr<uint32_t>(s, kBufBits, size_t(chunks) * (kPartChunk / 32));
  a.err = err;
  // Count pass first: the left count fixes where the right side starts.
  launch_partition(a, s.stream, 1);
  std::vector<uint32_t> counts(chunks);
  d2h(s, counts.data(), a.chunk_left, chunks * 4, "partition");
  s.sync("partition");
  uint64_t lefts = 0;
  for (uint32_t c : counts) lefts += c;
  pn.left_count = uint32_t(lefts);
  h2d(s, dn, &pn, sizeof(pn), "partition");
  launch_partition(a, s.stream, 2);
  std::vector<uint32_t> o(n);
  uint32_t flags = 0;
  d2h(s, o.data(), dout, n * 4, "partition");
  d2h(s, &flags, err, 4, "partition");
  s.sync("partition");
  if (flags) fail(MPSB_ERR_TRAINING, "Sample partition counts do not match histogram statistics");
  *n_left = lefts;
  *n_right = n - lefts;
  for (uint64_t i = 0; i < lefts; ++i) left[i] = o[i];
  for (uint64_t i = lefts; i < n; ++i) right[i - lefts] = o[i];
  ++s.reuse;
}

void hook_vector_add(Session& s, const float* a, const float* b, uint64_t n, float* out) {
  if (n == 0) return;
  auto* da = dptr<float>(s, kBufTmpA, n);
  auto* db = dptr<float>(s, kBufTmpB, n);
  auto* dc = dptr<float>(s, kBufTmpC, n);
  h2d(s, da, a, n * 4, "vector add");
  h2d(s, db, b, n * 4, "vector add");
  launch_vector_add(da, db, dc, n, s.stream);
  d2h(s, out, dc, n * 4, "vector add");
  s.sync("vector add");
}

}  // namespace mpsb

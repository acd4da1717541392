// extern "C" entry points of libmpsboost_hip.so (declared in include/mpsboost_hip.h).
// Every entry converts C++ exceptions into a status code plus a thread-local message, so no
// exception crosses the C boundary.

#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "engine.hpp"

struct mpsb_session {
  mpsb::Session s;
};
struct mpsb_dataset {
  mpsb::HostDataset d;
};
struct mpsb_model {
  mpsb::Model m;
};

struct mpsb_matrix {
  std::unique_ptr<mpsb::DeviceMatrix> m;
};

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F&& body) {
  try {
    g_error.clear();
    body();
    return MPSB_OK;
  } catch (const mpsb::Error& e) {
    g_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return MPSB_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_error = e.what();
    return MPSB_ERR_INTERNAL;
  }
}

void need(const void* p, const char* what) {
  if (!p) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, std::string(what) + " must not be null");
}

void bind(mpsb_session* s) {
  need(s, "session");
  MPSB_HIP("set device", hipSetDevice(s->s.device));
}

mpsb::View view_of(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs, int32_t scalar) {
  return mpsb::View{static_cast<const uint8_t*>(data), rows, features, rs, cs, scalar};
}

uint32_t env_u32(const char* name, uint32_t fallback) {
  const char* v = std::getenv(name);
  if (!v || !*v) return fallback;
  const long x = std::strtol(v, nullptr, 10);
  return x > 0 ? uint32_t(x) : fallback;
}

bool is_gfx950(const hipDeviceProp_t& p) { return std::strncmp(p.gcnArchName, "gfx950", 6) == 0; }

}  // namespace

extern "C" {

const char* mpsb_last_error(void) { return g_error.c_str(); }

int mpsb_query_backend_info(mpsb_backend_info* out) {
  return guarded([&] {
    need(out, "out");
    std::memset(out, 0, sizeof(*out));
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return;  // no device: available = 0
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return;
    out->available = is_gfx950(p) ? 1 : 0;
    out->compute_units = uint32_t(p.multiProcessorCount);
    out->total_memory = p.totalGlobalMem;
    std::strncpy(out->device_name, p.name, sizeof(out->device_name) - 1);
    std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
  });
}

int mpsb_session_create(int32_t device, mpsb_session** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      mpsb::fail(MPSB_ERR_DEVICE_UNAVAILABLE,
                 "no HIP device is visible; this backend needs an AMD Instinct MI355X (gfx950) with the ROCm driver loaded");
    if (device < 0 || device >= n) mpsb::fail(MPSB_ERR_DEVICE_UNAVAILABLE, "device index out of range");
    hipDeviceProp_t p;
    MPSB_HIP("device query", hipGetDeviceProperties(&p, device));
    if (!is_gfx950(p))
      mpsb::fail(MPSB_ERR_DEVICE_UNAVAILABLE, std::string("device is ") + p.gcnArchName + ", this library is built for gfx950");
    MPSB_HIP("set device", hipSetDevice(device));
    auto* s = new mpsb_session();
    static std::atomic<uint64_t> serial{0};
    s->s.device = device;
    s->s.id = ++serial;
    if (const char* w = std::getenv("MPSB_WAIT_TIMEOUT_S")) {
      const double v = std::atof(w);
      if (v > 0.0) s->s.wait_limit_s = v;
    }
    s->s.cus = uint32_t(p.multiProcessorCount);
    s->s.lds_max = uint32_t(p.sharedMemPerBlock);
    s->s.tune.lds_budget = env_u32("MPSB_LDS_BUDGET", s->s.tune.lds_budget);
    s->s.tune.target_blocks = env_u32("MPSB_TARGET_BLOCKS", s->s.tune.target_blocks);
    s->s.tune.min_slab_rows = env_u32("MPSB_MIN_SLAB_ROWS", s->s.tune.min_slab_rows);
    const uint32_t block = env_u32("MPSB_HIST_BLOCK", s->s.tune.block);
    s->s.tune.block = (block == 256 || block == 512 || block == 1024) ? block : 1024;
    const uint32_t wpt = env_u32("MPSB_HIST_WPT", s->s.tune.wpt);
    if (wpt == 1 || wpt == 2 || wpt == 4) s->s.tune.wpt = wpt;
    if (hipStreamCreateWithFlags(&s->s.stream, hipStreamNonBlocking) != hipSuccess) {
      delete s;
      mpsb::fail(MPSB_ERR_LAUNCH, "stream creation failed");
    }
    *out = s;
  });
}

void mpsb_session_destroy(mpsb_session* s) {
  if (s) {
    (void)hipSetDevice(s->s.device);
    delete s;
  }
}

int mpsb_session_last_timing(const mpsb_session* s, mpsb_timing* out) {
  return guarded([&] {
    need(s, "session");
    need(out, "out");
    *out = s->s.timing;
    out->pooled_buffer_reuse_count = s->s.reuse;
    out->pooled_buffer_allocation_count = s->s.allocs;
  });
}

int mpsb_session_get_tuning(const mpsb_session* s, mpsb_tuning* out) {
  return guarded([&] {
    need(s, "session");
    need(out, "out");
    const mpsb::Tuning& t = s->s.tune;
    *out = mpsb_tuning{t.lds_budget, t.target_blocks, t.min_slab_rows, t.block, t.wpt, t.planner};
  });
}

int mpsb_session_set_tuning(mpsb_session* s, const mpsb_tuning* t) {
  return guarded([&] {
    need(s, "session");
    need(t, "tuning");
    auto bad = [](const char* what) { mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, std::string("tuning: ") + what); };
    if (t->lds_budget < 16384 || t->lds_budget > s->s.lds_max) bad("lds_budget outside 16384 .. device LDS");
    if (t->target_blocks < 64 || t->target_blocks > 65536) bad("target_blocks outside 64 .. 65536");
    if (t->min_slab_rows < 64 || t->min_slab_rows > (1u << 24)) bad("min_slab_rows outside 64 .. 2^24");
    if (t->block != 256 && t->block != 512 && t->block != 1024) bad("block must be 256, 512 or 1024");
    if (t->words_per_thread != 1 && t->words_per_thread != 2 && t->words_per_thread != 4)
      bad("words_per_thread must be 1, 2 or 4");
    if (t->planner > 2) bad("planner must be 0 (automatic), 1 (host) or 2 (device)");
    s->s.tune = mpsb::Tuning{t->lds_budget, t->target_blocks, t->min_slab_rows, t->block, t->words_per_thread,
                             t->planner};
  });
}

int mpsb_debug_inject_fault(mpsb_session* s, int32_t stage, int32_t countdown) {
  return guarded([&] {
    need(s, "session");
    if (stage < 0 || stage > mpsb::kFaultAlloc) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "unknown fault stage");
    s->s.fault_stage = stage;
    s->s.fault_countdown = countdown;
  });
}

int mpsb_quantize_dense(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                        int32_t scalar, uint32_t max_bins, uint32_t threads, mpsb_dataset** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    auto* d = new mpsb_dataset();
    try {
      d->d = mpsb::quantize(view_of(data, rows, features, rs, cs, scalar), max_bins, threads);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

int mpsb_transform_dense(const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                         int32_t scalar, const mpsb_dataset* schema, uint32_t threads, mpsb_dataset** out) {
  return guarded([&] {
    need(out, "out");
    need(schema, "schema");
    *out = nullptr;
    auto* d = new mpsb_dataset();
    try {
      d->d = mpsb::transform(view_of(data, rows, features, rs, cs, scalar), schema->d, threads);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

int mpsb_quantize_dense_device(mpsb_session* s, const void* data, uint64_t rows, uint32_t features, uint64_t rs,
                               uint64_t cs, int32_t scalar, uint32_t max_bins, mpsb_dataset** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    bind(s);
    auto* d = new mpsb_dataset();
    try {
      d->d = mpsb::quantize_on_device(s->s, view_of(data, rows, features, rs, cs, scalar), max_bins);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

int mpsb_transform_dense_device(mpsb_session* s, const void* data, uint64_t rows, uint32_t features, uint64_t rs,
                                uint64_t cs, int32_t scalar, const mpsb_dataset* schema, mpsb_dataset** out) {
  return guarded([&] {
    need(out, "out");
    need(schema, "schema");
    *out = nullptr;
    bind(s);
    auto* d = new mpsb_dataset();
    try {
      d->d = mpsb::transform_on_device(s->s, view_of(data, rows, features, rs, cs, scalar), schema->d);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

int mpsb_matrix_upload(int32_t device, const void* data, uint64_t rows, uint32_t features, uint64_t rs, uint64_t cs,
                       int32_t scalar, mpsb_matrix** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    auto* m = new mpsb_matrix();
    try {
      m->m.reset(mpsb::upload_matrix(device, view_of(data, rows, features, rs, cs, scalar)));
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

void mpsb_matrix_free(mpsb_matrix* m) { delete m; }

int mpsb_quantize_matrix_sampled(mpsb_session* s, const mpsb_matrix* m, const uint64_t* rows, uint64_t n_rows,
                                 const uint32_t* features, uint32_t n_features, uint32_t max_bins,
                                 mpsb_dataset** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    need(m, "matrix");
    need(rows, "rows");
    need(features, "features");
    bind(s);
    if (m->m->device != s->s.device) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "matrix lives on another device");
    auto* d = new mpsb_dataset();
    try {
      d->d = mpsb::quantize_sampled(s->s, *m->m, rows, n_rows, features, n_features, max_bins);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

int mpsb_dataset_from_bins(uint64_t rows, uint32_t features, uint32_t max_bins, const uint32_t* bin_count,
                           const void* bins, int32_t bin_bytes, const uint8_t* missing, const float* boundaries,
                           mpsb_dataset** out) {
  return guarded([&] {
    need(out, "out");
    need(bin_count, "bin_count");
    need(bins, "bins");
    *out = nullptr;
    if (rows == 0 || features == 0) mpsb::fail(MPSB_ERR_DATA, "Input matrix must contain at least one row and one feature");
    if (bin_bytes != 1 && bin_bytes != 2) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "bin_bytes must be 1 or 2");
    if (rows > 0xFFFFFFFFull) mpsb::fail(MPSB_ERR_DATA, "row count exceeds the device's 32-bit row index");
    auto* d = new mpsb_dataset();
    try {
      mpsb::HostDataset& h = d->d;
      h.rows = rows;
      h.features = features;
      h.max_bins = max_bins;
      uint64_t nb = 0;
      for (uint32_t f = 0; f < features; ++f) {
        if (bin_count[f] == 0 || bin_count[f] > max_bins) mpsb::fail(MPSB_ERR_DATA, "bin count out of range");
        h.bin_count.push_back(bin_count[f]);
        h.boundary_offset.push_back(nb);
        nb += bin_count[f] - 1;
      }
      if (boundaries) {
        h.boundaries.assign(boundaries, boundaries + nb);
      } else {
        // Ordinal placeholders keep the schema valid when only bins are known.
        for (uint32_t f = 0; f < features; ++f)
          for (uint32_t b = 0; b + 1 < bin_count[f]; ++b) h.boundaries.push_back(float(b));
      }
      h.missing_count.assign(features, 0);
      if (missing)
        for (uint32_t f = 0; f < features; ++f)
          for (uint64_t r = 0; r < rows; ++r) h.missing_count[f] += missing[f * rows + r] ? 1 : 0;
      mpsb::validate_schema(h);
      h.finalize_layout();
      h.codes.assign(size_t(rows) * h.row_stride * h.code_bytes, 0);
      for (uint32_t f = 0; f < features; ++f)
        for (uint64_t r = 0; r < rows; ++r) {
          const size_t src = size_t(f) * rows + r;
          uint32_t code = bin_bytes == 1 ? static_cast<const uint8_t*>(bins)[src] : static_cast<const uint16_t*>(bins)[src];
          if (code >= bin_count[f]) mpsb::fail(MPSB_ERR_DATA, "Binned value exceeds the feature histogram range");
          if (missing && missing[src]) code = bin_count[f];
          const size_t at = (size_t(r) * h.row_stride + f) * h.code_bytes;
          h.codes[at] = uint8_t(code);
          if (h.code_bytes == 2) h.codes[at + 1] = uint8_t(code >> 8);
        }
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

void mpsb_dataset_free(mpsb_dataset* d) { delete d; }

int mpsb_dataset_info_get(const mpsb_dataset* d, mpsb_dataset_info* out) {
  return guarded([&] {
    need(d, "dataset");
    need(out, "out");
    out->rows = d->d.rows;
    out->features = d->d.features;
    out->max_bins = d->d.max_bins;
    out->storage_bits = d->d.max_bins <= 256 ? 8 : 16;
    out->code_bytes = d->d.code_bytes;
    out->row_stride = d->d.row_stride;
    out->cells = d->d.cells;
    out->boundary_total = d->d.boundaries.size();
  });
}

int mpsb_dataset_meta(const mpsb_dataset* d, uint32_t* bin_count, uint64_t* boundary_offset, uint64_t* missing_count,
                      float* boundaries) {
  return guarded([&] {
    need(d, "dataset");
    const auto& h = d->d;
    for (uint32_t f = 0; f < h.features; ++f) {
      if (bin_count) bin_count[f] = h.bin_count[f];
      if (boundary_offset) boundary_offset[f] = h.boundary_offset[f];
      if (missing_count) missing_count[f] = h.missing_count[f];
    }
    if (boundaries) std::memcpy(boundaries, h.boundaries.data(), h.boundaries.size() * sizeof(float));
  });
}

int mpsb_dataset_codes(const mpsb_dataset* d, void* out, uint64_t capacity, uint64_t* size) {
  return guarded([&] {
    need(d, "dataset");
    need(size, "size");
    const auto& c = d->d.host_codes();
    *size = c.size();
    if (!out) return;
    if (capacity < c.size()) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "codes buffer is too small");
    std::memcpy(out, c.data(), c.size());
  });
}

int mpsb_dataset_bins(const mpsb_dataset* d, uint16_t* bins, uint8_t* missing) {
  return guarded([&] {
    need(d, "dataset");
    const auto& h = d->d;
    const auto& c = h.host_codes();
    for (uint32_t f = 0; f < h.features; ++f)
      for (uint64_t r = 0; r < h.rows; ++r) {
        const uint32_t code = h.code(c, r, f);
        const bool miss = h.has_missing(f) && code == h.bin_count[f];
        if (bins) bins[size_t(f) * h.rows + r] = uint16_t(miss ? 0 : code);
        if (missing) missing[size_t(f) * h.rows + r] = miss ? 1 : 0;
      }
  });
}

int mpsb_compute_squared_error(mpsb_session* s, const double* y, const double* p, uint64_t n, double* out) {
  return guarded([&] {
    bind(s);
    if (n) {
      need(y, "labels");
      need(p, "predictions");
      need(out, "out");
    }
    mpsb::hook_squared_error(s->s, y, p, n, out);
  });
}

int mpsb_build_histograms(mpsb_session* s, const mpsb_dataset* d, const uint64_t* rows, uint64_t n_rows,
                          const double* pairs, int32_t precision, uint64_t* count, double* g, double* h) {
  return guarded([&] {
    bind(s);
    need(d, "dataset");
    need(pairs, "pairs");
    if (n_rows == 0) mpsb::fail(MPSB_ERR_TRAINING, "Node row set must not be empty");
    need(rows, "rows");
    const uint64_t offsets[2] = {0, n_rows};
    mpsb::hook_layer_histograms(s->s, d->d, 1, offsets, rows, pairs, precision, count, g, h);
  });
}

int mpsb_build_layer_histograms(mpsb_session* s, const mpsb_dataset* d, uint32_t n_nodes, const uint64_t* offsets,
                                const uint64_t* rows, const double* pairs, int32_t precision, uint64_t* count,
                                double* g, double* h) {
  return guarded([&] {
    bind(s);
    need(d, "dataset");
    need(offsets, "node_offsets");
    need(pairs, "pairs");
    if (n_nodes && offsets[n_nodes]) need(rows, "rows");
    mpsb::hook_layer_histograms(s->s, d->d, n_nodes, offsets, rows, pairs, precision, count, g, h);
  });
}

int mpsb_build_histograms_noalloc(const mpsb_dataset* d, const void* codes, const uint64_t* rows, uint64_t n_rows,
                                  const double* pairs, int32_t precision, uint64_t* count, double* g, double* h,
                                  void* stream) {
  return guarded([&] {
    need(d, "dataset");
    need(codes, "codes");
    need(pairs, "pairs");
    need(count, "count");
    need(g, "gradient");
    need(h, "hessian");
    if (n_rows) need(rows, "rows");
    mpsb::histograms_noalloc(d->d, codes, rows, n_rows, pairs, precision, count, g, h,
                             static_cast<hipStream_t>(stream));
  });
}

int mpsb_build_baseline_histograms(mpsb_session* s, const mpsb_dataset* d, const uint64_t* rows, uint64_t n_rows,
                                   const double* pairs, uint64_t* count, double* g, double* h) {
  return guarded([&] {
    bind(s);
    need(d, "dataset");
    need(pairs, "pairs");
    mpsb::hook_baseline_histograms(s->s, d->d, rows, n_rows, pairs, count, g, h);
  });
}

int mpsb_scan_splits(mpsb_session* s, const mpsb_dataset* d, const uint64_t* rows, uint64_t n_rows,
                     const double* pairs, const mpsb_split_params* params, int32_t precision,
                     mpsb_split_candidate* out, mpsb_split_candidate* best) {
  return guarded([&] {
    bind(s);
    need(d, "dataset");
    need(pairs, "pairs");
    need(params, "params");
    need(out, "out");
    mpsb::hook_scan(s->s, d->d, rows, n_rows, pairs, *params, precision, out, best);
  });
}

int mpsb_partition_rows(mpsb_session* s, const mpsb_dataset* d, const uint64_t* rows, uint64_t n_rows,
                        uint32_t feature, uint32_t threshold, int32_t default_left, uint64_t* left, uint64_t* n_left,
                        uint64_t* right, uint64_t* n_right) {
  return guarded([&] {
    bind(s);
    need(d, "dataset");
    need(n_left, "n_left");
    need(n_right, "n_right");
    if (n_rows) {
      need(rows, "rows");
      need(left, "left");
      need(right, "right");
    }
    mpsb::hook_partition(s->s, d->d, rows, n_rows, feature, threshold, default_left, left, n_left, right, n_right);
  });
}

int mpsb_vector_add(mpsb_session* s, const float* a, const float* b, uint64_t n, float* out) {
  return guarded([&] {
    bind(s);
    if (n) {
      need(a, "a");
      need(b, "b");
      need(out, "out");
    }
    mpsb::hook_vector_add(s->s, a, b, n, out);
  });
}

int mpsb_train_regressor(mpsb_session* s, const mpsb_dataset* d, const double* labels, const double* weights,
                         const mpsb_train_params* params, mpsb_model** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    need(d, "dataset");
    need(labels, "labels");
    need(params, "params");
    const mpsb::Params p = mpsb::params_from(*params);
    mpsb::validate_train_params(p, d->d);
    bind(s);
    auto* m = new mpsb_model();
    try {
      m->m = mpsb::train_regressor(s->s, d->d, labels, weights, p);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int mpsb_train_multiclass(mpsb_session* s, const mpsb_dataset* d, const double* labels, const double* weights,
                          uint32_t class_count, const mpsb_train_params* params, mpsb_model** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    need(d, "dataset");
    need(labels, "labels");
    need(params, "params");
    const mpsb::Params p = mpsb::params_from(*params);
    mpsb::validate_train_params(p, d->d);
    bind(s);
    auto* m = new mpsb_model();
    try {
      m->m = mpsb::train_multiclass(s->s, d->d, labels, weights, class_count, p);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int mpsb_model_info_get(const mpsb_model* m, mpsb_model_info* out) {
  return guarded([&] {
    need(m, "model");
    need(out, "out");
    std::memset(out, 0, sizeof(*out));
    out->tree_count = uint32_t(m->m.trees.size());
    out->class_count = m->m.classes;
    out->features = m->m.features;
    out->precision = uint32_t(m->m.precision);
    out->learning_rate = m->m.lr;
    for (size_t i = 0; i < m->m.base.size() && i < 256; ++i) out->base_scores[i] = m->m.base[i];
    out->objective = m->m.classes > 1 ? 0u : m->m.objective;
    out->max_bins = m->m.schema.max_bins;
    out->objective_alpha = m->m.objective_alpha;
    out->tweedie_variance_power = m->m.tweedie_power;
    for (size_t i = 0; i < m->m.class_labels.size() && i < 256; ++i) out->class_labels[i] = m->m.class_labels[i];
  });
}

int mpsb_model_schema(const mpsb_model* m, mpsb_dataset** out) {
  return guarded([&] {
    need(m, "model");
    need(out, "out");
    *out = nullptr;
    auto* d = new mpsb_dataset();
    d->d = m->m.schema;
    d->d.rows = 0;
    d->d.codes.clear();
    *out = d;
  });
}

int mpsb_model_set_class_labels(mpsb_model* m, const double* labels, uint32_t count) {
  return guarded([&] {
    need(m, "model");
    if (m->m.classes < 2) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "class labels belong to softmax models");
    // MulticlassModel::SetClassLabels, src/core/multiclass_trainer.cpp:198-213
    if (count != m->m.classes) mpsb::fail(MPSB_ERR_TRAINING, "softmax class label count does not match class_count");
    need(labels, "labels");
    for (uint32_t i = 0; i < count; ++i) {
      if (!std::isfinite(labels[i])) mpsb::fail(MPSB_ERR_TRAINING, "softmax class labels must be finite");
      if (i > 0 && !(labels[i - 1] < labels[i]))
        mpsb::fail(MPSB_ERR_TRAINING, "softmax class labels must be strictly increasing");
    }
    m->m.class_labels.assign(labels, labels + count);
  });
}

int mpsb_model_tree(const mpsb_model* m, uint32_t tree, mpsb_tree_node* nodes, uint32_t capacity, uint32_t* n_nodes) {
  return guarded([&] {
    need(m, "model");
    need(n_nodes, "n_nodes");
    if (tree >= m->m.trees.size()) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "tree index out of range");
    const auto& t = m->m.trees[tree];
    *n_nodes = uint32_t(t.size());
    if (t.size() > capacity) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "node capacity too small");
    need(nodes, "nodes");
    std::memcpy(nodes, t.data(), t.size() * sizeof(mpsb_tree_node));
  });
}

int mpsb_model_predict(mpsb_session* s, const mpsb_model* m, const mpsb_dataset* d, double* out) {
  return guarded([&] {
    bind(s);
    need(m, "model");
    need(d, "dataset");
    need(out, "out");
    mpsb::predict(s->s, m->m, d->d, out);
  });
}

void mpsb_model_free(mpsb_model* m) { delete m; }

int mpsb_model_fit_stats(const mpsb_model* m, mpsb_fit_stats* out) {
  return guarded([&] {
    need(m, "model");
    need(out, "out");
    *out = m->m.stats;
  });
}

int mpsb_model_serialize(const mpsb_model* m, uint8_t* out, uint64_t capacity, uint64_t* size) {
  return guarded([&] {
    need(m, "model");
    need(size, "size");
    const std::vector<uint8_t> bytes = mpsb::serialize_model(m->m);
    *size = bytes.size();
    if (out == nullptr) return;  // size query
    if (capacity < bytes.size()) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "model byte buffer is too small");
    std::memcpy(out, bytes.data(), bytes.size());
  });
}

int mpsb_model_deserialize(const uint8_t* data, uint64_t size, mpsb_model** out) {
  return mpsb_model_deserialize_kind(data, size, -1, out);
}

int mpsb_model_deserialize_kind(const uint8_t* data, uint64_t size, int32_t kind, mpsb_model** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    need(data, "data");
    if (kind < -1 || kind > 1) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "model kind must be -1, 0 or 1");
    auto* m = new mpsb_model();
    try {
      m->m = mpsb::deserialize_model(data, size, kind);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int mpsb_model_save(const mpsb_model* m, const char* path) {
  return guarded([&] {
    need(m, "model");
    need(path, "path");
    mpsb::save_model_file(m->m, path);
  });
}

int mpsb_model_load(const char* path, mpsb_model** out) { return mpsb_model_load_kind(path, -1, out); }

int mpsb_model_load_kind(const char* path, int32_t kind, mpsb_model** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    need(path, "path");
    if (kind < -1 || kind > 1) mpsb::fail(MPSB_ERR_INVALID_ARGUMENT, "model kind must be -1, 0 or 1");
    auto* m = new mpsb_model();
    try {
      m->m = mpsb::load_model_file(path, kind);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

}  // extern "C"

// Model files: the reference's native MPSBMOD format (version 1.4), so models trained here
// load in the reference and the other way round.
//
// Container (src/io/model_format.cpp:46-123): 8-byte magic "MPSBMOD\0", major u16 = 1,
// minor u16 = 4, model kind u32 (0 regression, 1 softmax multiclass), payload size u64,
// FNV-1a 64 of the payload; then the payload. All integers little-endian, floats as their
// IEEE bits. Payload (src/io/model_payload.cpp): schema (per-feature boundary offset/count,
// bin count, missing count; float boundaries), objective metadata, and trees of 40-byte
// node records. Older minors are read as the reference reads them (no objective before 1,
// no missing counts or default_left before 2, no quantile/Tweedie fields before 4).
// Validation follows the reference's loaders and their messages: schema fields
// (src/core/binned_dataset_internal.cpp:50-88), tree structure
// (src/core/tree_structure.cpp:150-187), model metadata (src/core/trainer.cpp:284-320,
// src/core/multiclass_trainer.cpp:151-197).
#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <string>
#include <system_error>
#include <vector>

#include "engine.hpp"

namespace mpsb {
namespace {

constexpr uint8_t kMagic[8] = {'M', 'P', 'S', 'B', 'M', 'O', 'D', 0};
constexpr uint16_t kMajor = 1, kMinor = 4;
constexpr uint32_t kKindRegression = 0, kKindMulticlass = 1;
constexpr size_t kHeader = 32;

[[noreturn]] void bad_model(const std::string& m) { fail(MPSB_ERR_TRAINING, m); }

uint64_t fnv1a(const uint8_t* p, size_t n) {  // src/io/model_format.cpp:16-28
  uint64_t h = 14695981039346656037ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
  return h;
}

struct Writer {
  std::vector<uint8_t> b;
  template <class T>
  void u(T v) {
    for (size_t i = 0; i < sizeof(T); ++i) b.push_back(uint8_t((uint64_t(v) >> (8 * i)) & 0xFF));
  }
  void f64(double v) {
    uint64_t x;
    std::memcpy(&x, &v, 8);
    u<uint64_t>(x);
  }
  void f32(float v) {
    uint32_t x;
    std::memcpy(&x, &v, 4);
    u<uint32_t>(x);
  }
};

struct Reader {
  const uint8_t* p;
  size_t n, at = 0;
  template <class T>
  T u(const char* field) {
    if (sizeof(T) > n - at) bad_model(std::string("Model data is truncated: ") + field);
    uint64_t v = 0;
    for (size_t i = 0; i < sizeof(T); ++i) v |= uint64_t(p[at++]) << (8 * i);
    return T(v);
  }
  double f64(const char* field) {
    const uint64_t x = u<uint64_t>(field);
    double v;
    std::memcpy(&v, &x, 8);
    if (!std::isfinite(v)) bad_model(std::string("Model floating-point field is not finite: ") + field);
    return v;
  }
  float f32(const char* field) {
    const uint32_t x = u<uint32_t>(field);
    float v;
    std::memcpy(&v, &x, 4);
    if (!std::isfinite(v)) bad_model(std::string("Model floating-point field is not finite: ") + field);
    return v;
  }
};

void write_schema_fields(Writer& w, const HostDataset& s) {
  w.u<uint64_t>(s.boundaries.size());
  for (uint32_t f = 0; f < s.features; ++f) {
    w.u<uint64_t>(s.boundary_offset[f]);
    w.u<uint32_t>(s.bin_count[f] - 1);
    w.u<uint32_t>(s.bin_count[f]);
    w.u<uint64_t>(s.missing_count[f]);
  }
  for (const float b : s.boundaries) w.f32(b);
}

void write_trees(Writer& w, const Model& m) {
  w.u<uint32_t>(uint32_t(m.trees.size()));
  for (const auto& t : m.trees) {
    w.u<uint32_t>(uint32_t(t.size()));
    for (const mpsb_tree_node& nd : t) {
      w.u<uint32_t>(nd.feature_index);
      w.u<uint32_t>(nd.threshold_bin);
      w.u<uint32_t>(nd.left_child);
      w.u<uint32_t>(nd.right_child);
      w.f64(nd.leaf_value);
      w.f64(nd.gain);
      w.u<uint8_t>(nd.default_left ? 1 : 0);
      w.u<uint8_t>(nd.flags);
      for (int i = 0; i < 6; ++i) w.u<uint8_t>(0);
    }
  }
}

// ValidateSchemaFields, src/core/binned_dataset_internal.cpp:50-88 (DataError).
void validate_schema_fields(const HostDataset& s, uint64_t boundary_total) {
  if (s.features == 0 || s.max_bins < 2 || s.max_bins > 65536)
    fail(MPSB_ERR_DATA, "Binned schema feature count or max_bins is invalid");
  uint64_t expected = 0;
  for (uint32_t f = 0; f < s.features; ++f) {
    const uint64_t count = s.bin_count[f] == 0 ? 0 : s.bin_count[f] - 1;
    if (s.boundary_offset[f] != expected || s.bin_count[f] == 0 || s.bin_count[f] > s.max_bins)
      fail(MPSB_ERR_DATA, "Binned schema boundary range or bin count is invalid");
    const uint64_t end = s.boundary_offset[f] + count;
    if (end > boundary_total) fail(MPSB_ERR_DATA, "Binned schema boundary range is out of bounds");
    for (uint64_t i = 0; i < count; ++i) {
      const float v = s.boundaries[s.boundary_offset[f] + i];
      if (!std::isfinite(v) || (i != 0 && v <= s.boundaries[s.boundary_offset[f] + i - 1]))
        fail(MPSB_ERR_DATA, "Binned schema feature boundaries must be finite and strictly increasing");
    }
    expected = end;
  }
  if (expected != boundary_total) fail(MPSB_ERR_DATA, "Binned schema contains boundaries not referenced by a feature");
}

void read_schema_fields(Reader& r, HostDataset& s, uint16_t minor) {
  const uint64_t total = r.u<uint64_t>("boundary_count");
  if (total > r.n / sizeof(float)) bad_model("Model boundary_count exceeds the payload range");
  s.boundary_offset.resize(s.features);
  s.bin_count.resize(s.features);
  s.missing_count.assign(s.features, 0);
  for (uint32_t f = 0; f < s.features; ++f) {
    s.boundary_offset[f] = r.u<uint64_t>("boundary_offset");
    const uint32_t bcount = r.u<uint32_t>("feature boundary_count");
    s.bin_count[f] = r.u<uint32_t>("feature bin_count");
    if (minor >= 2) s.missing_count[f] = r.u<uint64_t>("feature missing_count");
    if (s.bin_count[f] != bcount + 1u) fail(MPSB_ERR_DATA, "Binned schema boundary range or bin count is invalid");
  }
  s.boundaries.resize(total);
  for (uint64_t i = 0; i < total; ++i) s.boundaries[i] = r.f32("boundary");
  validate_schema_fields(s, total);
  s.rows = 0;
  s.finalize_layout();
}

// ValidateTreeStructure, src/core/tree_structure.cpp:150-187.
void validate_tree(uint32_t features, const std::vector<mpsb_tree_node>& nodes) {
  if (features == 0 || nodes.empty() || nodes.size() >= 0xFFFFFFFFull)
    bad_model("Model tree feature count or node count is invalid");
  std::vector<uint8_t> state(nodes.size(), 0);
  std::vector<std::pair<uint32_t, int>> stack{{0u, 0}};  // node, visited children
  // Iterative depth-first walk in the reference's order (left subtree, then right).
  if (0 >= nodes.size()) bad_model("Model tree child index is out of bounds");
  state[0] = 1;
  while (!stack.empty()) {
    auto& [i, k] = stack.back();
    const mpsb_tree_node& nd = nodes[i];
    if (k == 0) {
      if (nd.flags & 1u) {
        if (nd.flags != 1u || !std::isfinite(nd.leaf_value)) bad_model("Model tree leaf value is not finite");
        state[i] = 2;
        stack.pop_back();
        continue;
      }
      if (nd.flags != 0 || nd.feature_index >= features || !std::isfinite(nd.gain) || nd.gain <= 0.0 ||
          nd.left_child == nd.right_child)
        bad_model("Model tree branch fields are invalid");
    }
    if (k == 2) {
      state[i] = 2;
      stack.pop_back();
      continue;
    }
    const uint32_t child = k == 0 ? nd.left_child : nd.right_child;
    ++k;
    if (child >= nodes.size()) bad_model("Model tree child index is out of bounds");
    if (state[child] == 1) bad_model("Model tree structure contains a cycle");
    if (state[child] == 2) continue;
    state[child] = 1;
    stack.push_back({child, 0});
  }
  for (uint8_t s : state)
    if (s == 0) bad_model("Model tree contains unreachable nodes");
}

void read_trees(Reader& r, Model& m, uint16_t minor) {
  const uint32_t count = r.u<uint32_t>("tree_count");
  if (count == 0 || count > r.n / 41) bad_model("Model tree count is invalid");
  m.trees.reserve(count);
  for (uint32_t t = 0; t < count; ++t) {
    const uint32_t nn = r.u<uint32_t>("node_count");
    if (nn == 0 || nn > r.n / 41) bad_model("Model node count is invalid");
    std::vector<mpsb_tree_node> nodes(nn);
    for (auto& nd : nodes) {
      nd = mpsb_tree_node{};
      nd.feature_index = r.u<uint32_t>("feature_index");
      nd.threshold_bin = r.u<uint32_t>("threshold_bin");
      nd.left_child = r.u<uint32_t>("left_child");
      nd.right_child = r.u<uint32_t>("right_child");
      nd.leaf_value = r.f64("leaf_value");
      nd.gain = r.f64("gain");
      nd.default_left = 1;
      if (minor >= 2) {
        const uint8_t dl = r.u<uint8_t>("default_left");
        if (dl > 1) bad_model("model default_left must be 0 or 1");
        nd.default_left = dl;
      }
      nd.flags = r.u<uint8_t>("flags");
      for (int i = 0; i < (minor >= 2 ? 6 : 7); ++i)
        if (r.u<uint8_t>("node reserved") != 0) bad_model("Model node reserved fields must be zero");
    }
    validate_tree(m.features, nodes);
    m.trees.push_back(std::move(nodes));
  }
}

// Thresholds inside the bin range and leaves keeping default_left (src/core/trainer.cpp:
// 303-318, src/core/multiclass_trainer.cpp:173-187).
void validate_against_schema(const Model& m, const char* threshold_msg, const char* leaf_msg) {
  for (const auto& t : m.trees)
    for (const mpsb_tree_node& nd : t) {
      if (!(nd.flags & 1u) && nd.threshold_bin >= m.schema.bin_count[nd.feature_index] - 1) bad_model(threshold_msg);
      if ((nd.flags & 1u) && !nd.default_left) bad_model(leaf_msg);
    }
}

}  // namespace

std::vector<uint8_t> serialize_model(const Model& m) {
  Writer w;
  if (m.classes <= 1) {  // BuildPayload(RegressionModel), src/io/model_payload.cpp:143-171
    w.u<uint32_t>(m.features);
    w.u<uint32_t>(m.schema.max_bins);
    w.f64(m.base.at(0));
    w.f64(m.lr);
    w.u<uint32_t>(m.objective);
    w.f64(m.objective_alpha);
    w.f64(m.tweedie_power);
    write_schema_fields(w, m.schema);
  } else {  // BuildPayload(MulticlassModel), :173-185
    w.u<uint32_t>(m.classes);
    w.f64(m.lr);
    for (const double b : m.base) w.f64(b);
    for (uint32_t c = 0; c < m.classes; ++c) w.f64(c < m.class_labels.size() ? m.class_labels[c] : double(c));
    w.u<uint32_t>(m.features);
    w.u<uint32_t>(m.schema.max_bins);
    write_schema_fields(w, m.schema);
  }
  write_trees(w, m);
  Writer out;
  out.b.assign(kMagic, kMagic + 8);
  out.u<uint16_t>(kMajor);
  out.u<uint16_t>(kMinor);
  out.u<uint32_t>(m.classes <= 1 ? kKindRegression : kKindMulticlass);
  out.u<uint64_t>(w.b.size());
  out.u<uint64_t>(fnv1a(w.b.data(), w.b.size()));
  out.b.insert(out.b.end(), w.b.begin(), w.b.end());
  return out.b;
}

Model deserialize_model(const uint8_t* data, size_t size, int expect_kind) {
  if (size < kHeader || std::memcmp(data, kMagic, 8) != 0) bad_model("Model magic does not match or header is truncated");
  Reader h{data + 8, size - 8};
  const uint16_t major = h.u<uint16_t>("major");
  const uint16_t minor = h.u<uint16_t>("minor");
  const uint32_t kind = h.u<uint32_t>("model kind");
  const uint64_t psize = h.u<uint64_t>("payload size");
  const uint64_t sum = h.u<uint64_t>("checksum");
  if (major != kMajor) bad_model("Model major version is unsupported");
  if (expect_kind == int(kKindRegression) && kind != kKindRegression)
    bad_model("model kind is incompatible with regression loader");
  if (expect_kind == int(kKindMulticlass) && kind != kKindMulticlass)
    bad_model("model kind is incompatible with multiclass loader");
  if (kind != kKindRegression && kind != kKindMulticlass) bad_model("model kind is unsupported");
  if (psize != size - kHeader) bad_model("Model payload length does not match");
  const uint8_t* payload = data + kHeader;
  if (fnv1a(payload, psize) != sum) bad_model("Model integrity check failed");
  Reader r{payload, size_t(psize)};
  Model m;
  if (kind == kKindRegression) {  // ParsePayload, src/io/model_payload.cpp:187-249
    m.features = r.u<uint32_t>("features");
    m.schema.features = m.features;
    m.schema.max_bins = r.u<uint32_t>("max_bins");
    m.base = {r.f64("base_score")};
    m.lr = r.f64("learning_rate");
    m.objective = 0;
    if (minor >= 1) {
      m.objective = r.u<uint32_t>("objective");
      if (m.objective > 4) bad_model("model objective is unsupported");
    }
    if (minor >= 4) {
      m.objective_alpha = r.f64("objective_alpha");
      m.tweedie_power = r.f64("tweedie_variance_power");
    }
    read_schema_fields(r, m.schema, minor);
    read_trees(r, m, minor);
    if (r.at != r.n) bad_model("Model contains unrecognized trailing bytes");
    // RegressionModel::Restore, src/core/trainer.cpp:284-320
    if (!(m.lr > 0.0) || m.lr > 1.0 || m.trees.empty())
      bad_model("Model base score, learning rate, or tree count is invalid");
    if (!(m.objective_alpha > 0.0) || m.objective_alpha >= 1.0) bad_model("model quantile alpha is invalid");
    if (!(m.tweedie_power > 1.0) || m.tweedie_power >= 2.0) bad_model("model tweedie variance power is invalid");
    validate_against_schema(m, "Model tree split threshold exceeds the feature bin range",
                            "model leaf nodes must keep default_left true");
    m.classes = 1;
  } else {  // ParseMulticlassPayload, src/io/model_payload.cpp:251-290
    if (minor < 3) bad_model("multiclass model format is unsupported");
    m.classes = r.u<uint32_t>("class_count");
    m.lr = r.f64("learning_rate");
    if (m.classes > r.n / 16) bad_model("softmax model metadata is invalid");
    for (uint32_t c = 0; c < m.classes; ++c) m.base.push_back(r.f64("softmax base_score"));
    for (uint32_t c = 0; c < m.classes; ++c) m.class_labels.push_back(r.f64("softmax class label"));
    m.features = r.u<uint32_t>("features");
    m.schema.features = m.features;
    m.schema.max_bins = r.u<uint32_t>("max_bins");
    read_schema_fields(r, m.schema, minor);
    read_trees(r, m, minor);
    if (r.at != r.n) bad_model("Model contains unrecognized trailing bytes");
    // MulticlassModel::Restore, src/core/multiclass_trainer.cpp:151-197
    if (m.classes < 2 || m.trees.empty() || m.trees.size() % m.classes != 0 || !(m.lr > 0.0) || m.lr > 1.0)
      bad_model("softmax model metadata is invalid");
    for (uint32_t c = 1; c < m.classes; ++c)
      if (!(m.class_labels[c - 1] < m.class_labels[c]))
        bad_model("softmax model class labels must be strictly increasing");
    validate_against_schema(m, "softmax model split threshold is outside schema",
                            "softmax model leaf nodes must keep default_left true");
  }
  return m;
}

// SaveModelFile, src/io/model_file_io.cpp:22-70: a same-directory temporary file, written,
// synchronised and renamed over the target, so a reader never sees a partial model.
void save_model_file(const Model& m, const std::string& path) {
  if (path.empty()) bad_model("Model save path must not be empty");
  const std::vector<uint8_t> bytes = serialize_model(m);
  const std::filesystem::path target(path);
  const std::filesystem::path dir = target.parent_path().empty() ? std::filesystem::path(".") : target.parent_path();
  if (!std::filesystem::is_directory(dir)) bad_model("Model save directory does not exist");
  std::string pattern = (dir / ("." + target.filename().string() + ".mpsboost-XXXXXX")).string();
  std::vector<char> name(pattern.begin(), pattern.end());
  name.push_back('\0');
  const int fd = ::mkstemp(name.data());
  if (fd < 0) bad_model(std::string("Failed to create temporary model file: ") + std::strerror(errno));
  const std::string tmp(name.data());
  size_t done = 0;
  while (done < bytes.size()) {
    const ssize_t k = ::write(fd, bytes.data() + done, bytes.size() - done);
    if (k <= 0) {
      const std::string why = std::strerror(errno);
      ::close(fd);
      ::unlink(tmp.c_str());
      bad_model("Failed to write temporary model file: " + why);
    }
    done += size_t(k);
  }
  if (::fsync(fd) != 0 || ::close(fd) != 0) {
    const std::string why = std::strerror(errno);
    ::unlink(tmp.c_str());
    bad_model("Failed to synchronize temporary model file: " + why);
  }
  if (::rename(tmp.c_str(), target.c_str()) != 0) {
    const std::string why = std::strerror(errno);
    ::unlink(tmp.c_str());
    bad_model("Failed to atomically replace model file: " + why);
  }
}

// LoadModelFile, src/io/model_file_io.cpp:72-110.
Model load_model_file(const std::string& path, int expect_kind) {
  if (path.empty()) bad_model("Model load path must not be empty");
  std::error_code ec;
  const uintmax_t size = std::filesystem::file_size(path, ec);
  if (ec) bad_model("Unable to read model file size");
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) bad_model(std::string("Failed to open model file: ") + std::strerror(errno));
  std::vector<uint8_t> bytes(size);
  size_t at = 0;
  while (at < bytes.size()) {
    const ssize_t k = ::read(fd, bytes.data() + at, bytes.size() - at);
    if (k <= 0) {
      ::close(fd);
      bad_model(k == 0 ? "Model file ended before the expected read completed"
                       : std::string("Failed to read model file: ") + std::strerror(errno));
    }
    at += size_t(k);
  }
  ::close(fd);
  return deserialize_model(bytes.data(), bytes.size(), expect_kind);
}

}  // namespace mpsb

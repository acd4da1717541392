// Device binning: QuantizeDense / TransformDense (src/core/binned_dataset.cpp:50-238) with the
// reads, the sort, the quantile cut and the code assignment on the GPU.
//
//  1. The matrix goes to the device in row chunks. k_bin_gather reads each chunk through a
//     64 x 64 LDS tile (coalesced along the input's rows) with the reference's reads (float32
//     as is; float64 range-checked, then rounded to nearest; Inf -> error; NaN -> missing) and
//     writes order-preserving 32-bit keys feature-major, NaN -> 0xFFFFFFFF so missing values
//     sort last. The same pass counts missing values per feature and finds the first Inf /
//     out-of-range value in the reference's feature-major read order.
//  2. Each feature's keys are radix-sorted (rocprim); k_bin_candidates picks the values at the
//     ceil-rounded quantile ranks (QuantileRank, :83-93) and the maximum.
//  3. The host applies BuildBoundaries' filter (:95-118) to the <= max_bins - 1 candidates of
//     each feature: below the maximum and strictly increasing.
//  4. k_bin_codes writes the row-major device codes straight into the session's code buffer:
//     lower_bound among the feature's boundaries (staged in LDS), missing -> the missing cell.
//
// Bins, missing mask and boundaries equal the reference's. The one value the reference leaves
// to its sort is the sign of a zero boundary when a column holds both -0.0 and +0.0 (std::sort
// treats them as equal, this radix sort orders -0.0 first); the bins are the same either way.

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace mpsb {
namespace {

__device__ inline uint32_t order_key(float x) {
  if (isnan(x)) return 0xFFFFFFFFu;
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float key_value(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// ReadFiniteOrMissingFloat, src/core/binned_dataset.cpp:50-81. *bad: 1 Inf, 2 float64 outside
// the finite float32 range.
template <int kScalar, bool kAligned>
__device__ inline float read_value(const uint8_t* p, uint32_t* bad) {
  if (kScalar == 1) {
    float v;
    if (kAligned) {
      v = *reinterpret_cast<const float*>(p);
    } else {
      __builtin_memcpy(&v, p, 4);
    }
    if (isinf(v)) *bad = 1;
    return v;
  }
  double v;
  if (kAligned) {
    v = *reinterpret_cast<const double*>(p);
  } else {
    __builtin_memcpy(&v, p, 8);
  }
  if (isnan(v)) return __uint_as_float(0x7FC00000u);
  if (isinf(v)) {
    *bad = 1;
    return 0.0f;
  }
  constexpr double kFloatMax = 3.4028234663852886e38;
  if (v > kFloatMax || v < -kFloatMax) {
    *bad = 2;
    return 0.0f;
  }
  return static_cast<float>(v);
}

struct GatherLaunch {
  const uint8_t* x;  // device copy of the chunk: row r of the chunk at x + r * rs
  uint64_t row0, rows, rows_total;
  uint32_t features;
  uint64_t rs, cs;
  uint32_t* keys;                    // [features][rows_total]
  unsigned long long* missing;       // [features]
  unsigned long long* first_bad;     // [2]: feature-major index of the first Inf / range error
  const uint64_t* ridx;              // sampled binning: source row of output row r (or null)
  const uint32_t* fidx;              // sampled binning: source column of output feature f (or null)
};

constexpr uint32_t kTile = 64;

template <int kScalar, bool kAligned>
__global__ __launch_bounds__(256) void k_bin_gather(GatherLaunch a) {
  __shared__ uint32_t tile[kTile][kTile + 1];
  __shared__ uint32_t miss[kTile];
  const uint32_t ft_count = (a.features + kTile - 1) / kTile;
  const uint32_t f0 = (blockIdx.x % ft_count) * kTile;
  const uint64_t rt_count = (a.rows + kTile - 1) / kTile;
  const uint32_t per_ft = gridDim.x / ft_count;
  const uint32_t lane = threadIdx.x & (kTile - 1), quad = threadIdx.x / kTile;
  if (threadIdx.x < kTile) miss[threadIdx.x] = 0;
  __syncthreads();
  uint32_t my_miss = 0;
  unsigned long long bad_inf = ~0ull, bad_range = ~0ull;
  for (uint64_t rt = blockIdx.x / ft_count; rt < rt_count; rt += per_ft) {
    const uint64_t r0 = rt * kTile;
    // Load: (row r0 + quad + 4i, feature f0 + lane), consecutive lanes along the input row.
#pragma unroll 4
    for (uint32_t i = 0; i < kTile / 4; ++i) {
      const uint32_t rl = quad + 4 * i;
      const uint64_t r = r0 + rl;
      const uint32_t f = f0 + lane;
      uint32_t key = 0xFFFFFFFFu;
      if (r < a.rows && f < a.features) {
        uint32_t bad = 0;
        const uint64_t sr = a.ridx ? a.ridx[r] : r;
        const uint64_t sf = a.fidx ? a.fidx[f] : f;
        const float v = read_value<kScalar, kAligned>(a.x + sr * a.rs + sf * a.cs, &bad);
        if (bad) {
          const unsigned long long at = uint64_t(f) * a.rows_total + a.row0 + r;
          if (bad == 1) {
            bad_inf = at < bad_inf ? at : bad_inf;
          } else {
            bad_range = at < bad_range ? at : bad_range;
          }
        } else if (isnan(v)) {
          ++my_miss;
        }
        key = order_key(v);
      }
      tile[rl][lane] = key;
    }
    __syncthreads();
    // Store: (feature f0 + quad + 4i, row r0 + lane), consecutive lanes along the key column.
#pragma unroll 4
    for (uint32_t i = 0; i < kTile / 4; ++i) {
      const uint32_t fl = quad + 4 * i;
      const uint64_t r = r0 + lane;
      const uint32_t f = f0 + fl;
      if (r < a.rows && f < a.features) a.keys[uint64_t(f) * a.rows_total + a.row0 + r] = tile[lane][fl];
    }
    __syncthreads();
  }
  if (my_miss) atomicAdd(&miss[lane], my_miss);
  if (bad_inf != ~0ull) atomicMin(&a.first_bad[0], bad_inf);
  if (bad_range != ~0ull) atomicMin(&a.first_bad[1], bad_range);
  __syncthreads();
  if (threadIdx.x < kTile && miss[threadIdx.x])
    atomicAdd(&a.missing[f0 + threadIdx.x], static_cast<unsigned long long>(miss[threadIdx.x]));
}

// QuantileRank, src/core/binned_dataset.cpp:83-93: ceil(part * total / divisor) - 1 without
// overflow.
__device__ inline uint64_t quantile_rank(uint64_t part, uint64_t total, uint64_t divisor) {
  const uint64_t q = total / divisor, r = total % divisor, fn = part * r;
  return part * q + fn / divisor + (fn % divisor ? 1 : 0) - 1;
}

// cand[part - 1] = sorted value at the rank of `part` (1 <= part < desired); cand[max_bins - 1]
// = the maximum non-missing value.
__global__ __launch_bounds__(256) void k_bin_candidates(const uint32_t* sorted, uint64_t valid, uint32_t max_bins,
                                                        float* cand) {
  const uint64_t desired = valid < max_bins ? valid : max_bins;
  for (uint64_t part = 1 + blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; part < desired;
       part += uint64_t(gridDim.x) * blockDim.x)
    cand[part - 1] = key_value(sorted[quantile_rank(part, valid, desired)]);
  if (blockIdx.x == 0 && threadIdx.x == 0) cand[max_bins - 1] = key_value(sorted[valid - 1]);
}

struct CodesLaunch {
  const uint32_t* keys;       // [features][rows]
  uint64_t rows;
  uint32_t features;
  const float* bounds;        // every feature's boundaries, concatenated
  const uint32_t* bound_off;  // [groups * width + 1]: boundary offsets, padded features empty
  const uint32_t* miss_code;  // [features]
  uint8_t* codes;             // [rows][row_stride] codes of code_bytes each
  uint32_t row_stride, code_bytes;
  uint32_t width;             // features per group: 16 / code_bytes (one 16-byte store per row)
  uint32_t groups;
  uint32_t rows_per_block;
  uint32_t lds_floats;        // 0: search the boundaries in global memory
};

constexpr uint32_t kCodesLds = 12288;  // floats of boundaries staged per block (48 KB)

// lower_bound over n strictly increasing boundaries: the number of boundaries below x.
__device__ inline uint32_t lower_bound_of(const float* b, uint32_t n, float x) {
  uint32_t lo = 0;
  while (n > 0) {
    const uint32_t half = n >> 1;
    if (b[lo + half] < x) {
      lo += half + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_bin_codes(CodesLaunch a) {
  __shared__ float lds[kCodesLds];
  const uint32_t g = blockIdx.x % a.groups;
  const uint64_t r_begin = uint64_t(blockIdx.x / a.groups) * a.rows_per_block;
  const uint32_t fb = g * a.width;
  const uint32_t o0 = a.bound_off[fb];
  const float* bounds = a.bounds + o0;
  if (a.lds_floats) {
    const uint32_t n = a.bound_off[fb + a.width] - o0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = bounds[i];
    __syncthreads();
    bounds = lds;
  }
  const uint64_t r_end = min(a.rows, r_begin + a.rows_per_block);
  for (uint64_t r = r_begin + threadIdx.x; r < r_end; r += blockDim.x) {
    uint32_t packed[4] = {0, 0, 0, 0};
    for (uint32_t j = 0; j < a.width; ++j) {
      const uint32_t f = fb + j;
      if (f >= a.features) break;
      const float x = key_value(a.keys[uint64_t(f) * a.rows + r]);
      const uint32_t off = a.bound_off[f] - o0;
      const uint32_t code =
          isnan(x) ? a.miss_code[f] : lower_bound_of(bounds + off, a.bound_off[f + 1] - a.bound_off[f], x);
      if (a.code_bytes == 1) {
        packed[j >> 2] |= code << (8 * (j & 3));
      } else {
        packed[j >> 1] |= code << (16 * (j & 1));
      }
    }
    *reinterpret_cast<uint4*>(a.codes + (r * a.row_stride + fb) * a.code_bytes) =
        make_uint4(packed[0], packed[1], packed[2], packed[3]);
  }
}

template <class T>
T* buf(Session& s, Buf b, size_t count) {
  return static_cast<T*>(s.get(b, count * sizeof(T)));
}

void raise_bad(uint64_t inf_at, uint64_t range_at) {
  if (inf_at == ~0ull && range_at == ~0ull) return;
  if (inf_at < range_at) fail(MPSB_ERR_DATA, "Input matrix contains Inf");
  fail(MPSB_ERR_DATA, "float64 input value exceeds the finite float32 range");
}

void launch_gather(Session& s, const GatherLaunch& a, int scalar, bool aligned) {
  const uint32_t ft_count = (a.features + kTile - 1) / kTile;
  const uint64_t rt = (a.rows + kTile - 1) / kTile;
  const uint32_t per_ft = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(rt, 4096 / ft_count)));
  const dim3 grid(ft_count * per_ft), block(256);
  if (scalar == 1) {
    if (aligned) {
      hipLaunchKernelGGL((k_bin_gather<1, true>), grid, block, 0, s.stream, a);
    } else {
      hipLaunchKernelGGL((k_bin_gather<1, false>), grid, block, 0, s.stream, a);
    }
  } else if (aligned) {
    hipLaunchKernelGGL((k_bin_gather<2, true>), grid, block, 0, s.stream, a);
  } else {
    hipLaunchKernelGGL((k_bin_gather<2, false>), grid, block, 0, s.stream, a);
  }
  MPSB_HIP("device binning", hipGetLastError());
}

unsigned long long* clear_meta(Session& s, uint32_t features) {
  auto* meta = buf<unsigned long long>(s, kBufBinMeta, features + 2);
  MPSB_HIP("device binning", hipMemsetAsync(meta, 0, features * 8, s.stream));
  MPSB_HIP("device binning", hipMemsetAsync(meta + features, 0xFF, 16, s.stream));
  return meta;
}

// Missing counts per feature; raises the first bad value in feature-major read order.
std::vector<uint64_t> read_meta(Session& s, const unsigned long long* meta, uint32_t features) {
  std::vector<unsigned long long> host(features + 2);
  MPSB_HIP("device binning", hipMemcpyAsync(host.data(), meta, host.size() * 8, hipMemcpyDeviceToHost, s.stream));
  s.sync("device binning");
  raise_bad(host[features], host[features + 1]);
  return std::vector<uint64_t>(host.begin(), host.begin() + features);
}

// Steps 1: chunked upload and gather into feature-major keys. Returns the missing counts.
std::vector<uint64_t> gather_keys(Session& s, const View& v, uint32_t* keys) {
  const uint64_t item = v.scalar == 1 ? 4 : 8;
  const uint64_t row_span = uint64_t(v.features - 1) * v.col_stride + item;
  // Rows own disjoint byte ranges: upload in chunks of about 256 MB; otherwise in one piece.
  const bool by_rows = row_span <= v.row_stride;
  const uint64_t chunk_rows =
      by_rows ? std::max<uint64_t>(1, std::min<uint64_t>(v.rows, (uint64_t(256) << 20) / v.row_stride)) : v.rows;
  const uint64_t chunk_bytes = (chunk_rows - 1) * v.row_stride + row_span;
  auto* x = buf<uint8_t>(s, kBufBinX, chunk_bytes);
  auto* meta = clear_meta(s, v.features);
  const bool aligned = v.row_stride % item == 0 && v.col_stride % item == 0;
  const uint32_t ft_count = (v.features + kTile - 1) / kTile;
  for (uint64_t r0 = 0; r0 < v.rows; r0 += chunk_rows) {
    const uint64_t n = std::min(chunk_rows, v.rows - r0);
    const uint64_t bytes = (n - 1) * v.row_stride + row_span;
    MPSB_HIP("device binning upload",
             hipMemcpyAsync(x, v.data + r0 * v.row_stride, bytes, hipMemcpyHostToDevice, s.stream));
    GatherLaunch a{x, r0, n, v.rows, v.features, v.row_stride, v.col_stride, keys, meta, meta + v.features,
                   nullptr, nullptr};
    launch_gather(s, a, v.scalar, aligned);
  }
  (void)ft_count;
  return read_meta(s, meta, v.features);
}

// Step 4: row-major codes into the session's code buffer, a host copy, then the rest of the
// upload (column-major copy, per-feature tables).
void write_codes(Session& s, HostDataset& d, const uint32_t* keys) {
  const uint32_t width = 16 / d.code_bytes;
  const uint32_t groups = d.row_stride / width;
  std::vector<uint32_t> off(size_t(groups) * width + 1, 0), miss_code(d.features);
  for (uint32_t f = 0; f < groups * width; ++f) {
    const uint64_t n = f < d.features ? d.bin_count[f] - 1 : 0;
    const uint64_t next = uint64_t(off[f]) + n;
    if (next > 0xFFFFFFFFull) fail(MPSB_ERR_DATA, "device binning: too many boundaries");
    off[f + 1] = uint32_t(next);
  }
  uint32_t widest = 0;
  for (uint32_t g = 0; g < groups; ++g) widest = std::max(widest, off[(g + 1) * width] - off[g * width]);
  for (uint32_t f = 0; f < d.features; ++f) miss_code[f] = d.has_missing(f) ? d.bin_count[f] : 0u;
  auto* bounds = buf<float>(s, kBufBinBounds, d.boundaries.size() + 1);
  auto* tables = buf<uint32_t>(s, kBufBinCand, off.size() + d.features);
  if (!d.boundaries.empty())
    MPSB_HIP("device binning", hipMemcpyAsync(bounds, d.boundaries.data(), d.boundaries.size() * 4,
                                              hipMemcpyHostToDevice, s.stream));
  MPSB_HIP("device binning", hipMemcpyAsync(tables, off.data(), off.size() * 4, hipMemcpyHostToDevice, s.stream));
  MPSB_HIP("device binning", hipMemcpyAsync(tables + off.size(), miss_code.data(), d.features * 4,
                                            hipMemcpyHostToDevice, s.stream));
  const size_t code_bytes = size_t(d.rows) * d.row_stride * d.code_bytes;
  s.evict_codes();
  s.uploaded_uid = 0;  // the code buffer no longer holds the previously uploaded dataset
  auto* codes = buf<uint8_t>(s, kBufCodes, code_bytes);
  CodesLaunch a{};
  a.keys = keys;
  a.rows = d.rows;
  a.features = d.features;
  a.bounds = bounds;
  a.bound_off = tables;
  a.miss_code = tables + off.size();
  a.codes = codes;
  a.row_stride = d.row_stride;
  a.code_bytes = d.code_bytes;
  a.width = width;
  a.groups = groups;
  a.rows_per_block = 4096;
  a.lds_floats = widest <= kCodesLds ? widest : 0;
  const uint64_t row_blocks = (d.rows + a.rows_per_block - 1) / a.rows_per_block;
  hipLaunchKernelGGL(k_bin_codes, dim3(uint32_t(row_blocks * groups)), dim3(256), 0, s.stream, a);
  MPSB_HIP("device binning", hipGetLastError());
  // The codes stay in the session's code buffer; the host gets them only on demand.
  d.codes.clear();
  d.device = std::make_shared<DeviceCodes>();
  d.device->session = &s;
  d.device->bytes = code_bytes;
  s.adopt_codes(d);
  s.resident_codes = d.device;
}

void release_binning(Session& s) {
  for (Buf b : {kBufBinX, kBufBinKeys, kBufBinSorted, kBufBinTemp, kBufBinCand, kBufBinBounds}) s.release(b);
}

// Steps 2-4 on feature-major keys [F][N] with their missing counts.
HostDataset quantize_keys(Session& s, uint32_t* keys, const std::vector<uint64_t>& missing, uint64_t N, uint32_t F,
                          uint32_t max_bins) {
  size_t temp_bytes = 0;
  MPSB_HIP("device binning sort", rocprim::radix_sort_keys(nullptr, temp_bytes, keys, keys, N, 0, 32, s.stream));
  void* temp = s.get(kBufBinTemp, temp_bytes);
  auto* sorted = buf<uint32_t>(s, kBufBinSorted, N);
  auto* cand = buf<float>(s, kBufBinCand, size_t(F) * max_bins);
  for (uint32_t f = 0; f < F; ++f) {
    const uint64_t valid = N - missing[f];
    if (valid < 2 || max_bins < 2) continue;  // at most one non-missing value: no boundary
    MPSB_HIP("device binning sort",
             rocprim::radix_sort_keys(temp, temp_bytes, keys + f * N, sorted, N, 0, 32, s.stream));
    const uint64_t parts = std::min<uint64_t>(max_bins, valid);
    const uint32_t blocks = uint32_t(std::min<uint64_t>(64, (parts + 255) / 256));
    hipLaunchKernelGGL(k_bin_candidates, dim3(blocks), dim3(256), 0, s.stream, sorted, valid, max_bins,
                       cand + size_t(f) * max_bins);
    MPSB_HIP("device binning", hipGetLastError());
  }
  std::vector<float> host(size_t(F) * max_bins);
  MPSB_HIP("device binning", hipMemcpyAsync(host.data(), cand, host.size() * 4, hipMemcpyDeviceToHost, s.stream));
  s.sync("device binning");
  HostDataset d;
  d.rows = N;
  d.features = F;
  d.max_bins = max_bins;
  for (uint32_t f = 0; f < F; ++f) {  // BuildBoundaries' filter, src/core/binned_dataset.cpp:106-116
    const uint64_t valid = N - missing[f];
    const uint64_t desired = std::min<uint64_t>(max_bins, valid);
    const float* c = host.data() + size_t(f) * max_bins;
    const float top = c[max_bins - 1];
    const size_t first = d.boundaries.size();
    for (uint64_t part = 1; valid >= 2 && part < desired; ++part) {
      const float x = c[part - 1];
      if (x < top && (d.boundaries.size() == first || x > d.boundaries.back())) d.boundaries.push_back(x);
    }
    d.boundary_offset.push_back(first);
    d.bin_count.push_back(uint32_t(d.boundaries.size() - first + 1));
    d.missing_count.push_back(missing[f]);
  }
  validate_schema(d);
  d.finalize_layout();
  write_codes(s, d, keys);
  return d;
}

}  // namespace

HostDataset quantize_on_device(Session& s, const View& v, uint32_t max_bins) {
  validate_dense_view(v, max_bins);
  s.check_fault(kFaultUpload, "device binning");
  try {
    auto* keys = buf<uint32_t>(s, kBufBinKeys, v.rows * v.features);
    const std::vector<uint64_t> missing = gather_keys(s, v, keys);
    HostDataset d = quantize_keys(s, keys, missing, v.rows, v.features, max_bins);
    release_binning(s);
    return d;
  } catch (...) {
    release_binning(s);
    throw;
  }
}

DeviceMatrix::~DeviceMatrix() {
  if (data) {
    (void)hipSetDevice(device);
    (void)hipFree(data);
  }
}

DeviceMatrix* upload_matrix(int device, const View& v) {
  validate_dense_view(v, 256);
  const uint64_t item = v.scalar == 1 ? 4 : 8;
  auto m = std::make_unique<DeviceMatrix>();
  m->device = device;
  m->rows = v.rows;
  m->features = v.features;
  m->row_stride = v.row_stride;
  m->col_stride = v.col_stride;
  m->scalar = v.scalar;
  m->bytes = (v.rows - 1) * v.row_stride + uint64_t(v.features - 1) * v.col_stride + item;
  MPSB_HIP("set device", hipSetDevice(device));
  MPSB_HIP("matrix upload", hipMalloc(reinterpret_cast<void**>(&m->data), m->bytes));
  MPSB_HIP("matrix upload", hipMemcpy(m->data, v.data, m->bytes, hipMemcpyHostToDevice));
  return m.release();
}

// QuantizeDense of matrix[rows][:, features] (src/mpsboost/estimators/forest_sampling.py:
// 42-63 materialises that sub-matrix on the host, then fits on it): the sub-matrix is read
// straight from the resident matrix through the two index lists. The keys and scratch stay
// pooled in the session for the next tree.
HostDataset quantize_sampled(Session& s, const DeviceMatrix& m, const uint64_t* rows, uint64_t n_rows,
                             const uint32_t* features, uint32_t n_features, uint32_t max_bins) {
  if (n_rows == 0 || n_features == 0)
    fail(MPSB_ERR_DATA, "Input matrix must contain at least one row and one feature");
  if (max_bins < 2 || max_bins > 65536) fail(MPSB_ERR_DATA, "max_bins must be in [2, 65536]");
  if (n_rows > 0xFFFFFFFFull) fail(MPSB_ERR_DATA, "row count exceeds the device's 32-bit row index");
  uint64_t top = 0;  // the largest index, one pass with no early exit (vectorised)
  for (uint64_t i = 0; i < n_rows; ++i) top = rows[i] > top ? rows[i] : top;
  if (top >= m.rows) fail(MPSB_ERR_INVALID_ARGUMENT, "sampled row index out of range");
  for (uint32_t j = 0; j < n_features; ++j)
    if (features[j] >= m.features) fail(MPSB_ERR_INVALID_ARGUMENT, "sampled feature index out of range");
  s.check_fault(kFaultUpload, "device binning");
  auto* ridx = buf<uint64_t>(s, kBufBinX, n_rows + (n_features + 1) / 2 + 1);
  auto* fidx = reinterpret_cast<uint32_t*>(ridx + n_rows);
  MPSB_HIP("device binning", hipMemcpyAsync(ridx, rows, n_rows * 8, hipMemcpyHostToDevice, s.stream));
  MPSB_HIP("device binning", hipMemcpyAsync(fidx, features, n_features * 4, hipMemcpyHostToDevice, s.stream));
  auto* keys = buf<uint32_t>(s, kBufBinKeys, n_rows * n_features);
  auto* meta = clear_meta(s, n_features);
  const uint64_t item = m.scalar == 1 ? 4 : 8;
  const bool aligned = m.row_stride % item == 0 && m.col_stride % item == 0;
  GatherLaunch a{m.data, 0, n_rows, n_rows, n_features, m.row_stride, m.col_stride, keys, meta, meta + n_features,
                 ridx, fidx};
  launch_gather(s, a, m.scalar, aligned);
  const std::vector<uint64_t> missing = read_meta(s, meta, n_features);
  return quantize_keys(s, keys, missing, n_rows, n_features, max_bins);
}

HostDataset transform_on_device(Session& s, const View& v, const HostDataset& schema) {
  validate_dense_view(v, schema.max_bins);
  validate_schema(schema);
  if (v.features != schema.features)
    fail(MPSB_ERR_DATA, "Prediction input feature count does not match the binned schema");
  s.check_fault(kFaultUpload, "device binning");
  try {
    auto* keys = buf<uint32_t>(s, kBufBinKeys, v.rows * v.features);
    const std::vector<uint64_t> missing = gather_keys(s, v, keys);
    HostDataset d;
    d.rows = v.rows;
    d.features = schema.features;
    d.max_bins = schema.max_bins;
    d.boundaries = schema.boundaries;
    d.boundary_offset = schema.boundary_offset;
    d.bin_count = schema.bin_count;
    d.missing_count = schema.missing_count;
    // The training layout's missing cells, plus one wherever this data has NaN.
    d.miss_flag = schema.miss_flag;
    for (uint32_t f = 0; f < d.features; ++f)
      if (missing[f]) d.miss_flag[f] = 1;
    d.finalize_layout();
    write_codes(s, d, keys);
    release_binning(s);
    return d;
  } catch (...) {
    release_binning(s);
    throw;
  }
}

}  // namespace mpsb

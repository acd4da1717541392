// Allocation-free histogram build: mpsb_build_histograms on device-resident inputs and outputs,
// for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// Everything the build needs beyond the caller's inputs and outputs (the row ids as 32-bit
// positions, the gradients in the form's arithmetic, the round's scales, the tile tables, the
// per-slab partial histograms, the reduced node histogram and the error words) lives in one
// fixed device array, `scratch`, of kScratchBytes per GPU. Each GPU that runs this entry point
// holds its own copy of the array until the process ends. Its device address is looked up once
// per device (hipGetSymbolAddress, at the first call on that device) and only that address is
// handed to kernels and copies: in host code the array's name is a host-side stand-in.
//
// Every call on one GPU uses that GPU's one array, and the array keeps the last call's contents:
// the words that must start at zero are cleared by a kernel at the start of each call. Calls
// that overlap on one GPU (two threads, or two streams) race on the array and must not be made.

#include <algorithm>
#include <mutex>

#include "engine.hpp"

namespace mpsb {
namespace {

// Limits of one call.
constexpr uint64_t kNoAllocMaxRows = uint64_t(1) << 20;  // dataset rows and node rows
constexpr uint32_t kNoAllocMaxCells = 32768;             // device histogram cells per node
constexpr uint32_t kNoAllocMaxSlabs = 32;                // row slabs (partial histograms)
constexpr uint32_t kNoAllocMaxTiles = 64;                // feature tiles of a row

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Layout of the scratch array (byte offsets).
constexpr size_t kOffWords = 0;                                             // u32[64]
constexpr size_t kOffMaxabs = 256;                                          // u64[2]
constexpr size_t kOffScales = 512;                                          // f64[4]
constexpr size_t kOffTiles = 1024;                                          // TileDesc[64]
constexpr size_t kOffWordOff = align256(kOffTiles + kNoAllocMaxTiles * sizeof(TileDesc));
constexpr size_t kOffCellLds = align256(kOffWordOff + size_t(kNoAllocMaxTiles) * 256 * 4);
constexpr size_t kOffSlabs = align256(kOffCellLds + size_t(kNoAllocMaxCells) * 4);
constexpr size_t kOffReduce = align256(kOffSlabs + kNoAllocMaxSlabs * sizeof(SlabDesc));  // 5 x u32
constexpr size_t kOffViewMap = align256(kOffReduce + 64);                   // u32[2][cells]
constexpr size_t kOffHist = align256(kOffViewMap + size_t(kNoAllocMaxCells) * 8);  // g, h u64, cnt u32
constexpr size_t kOffRows = align256(kOffHist + size_t(kNoAllocMaxCells) * 20);    // u32[rows]
constexpr size_t kOffGH = align256(kOffRows + kNoAllocMaxRows * 4);                // <= 16 B per row
constexpr size_t kOffPartial = align256(kOffGH + kNoAllocMaxRows * 16);            // 20 B per slab cell
constexpr size_t kScratchBytes = align256(kOffPartial + size_t(kNoAllocMaxSlabs) * kNoAllocMaxCells * 20);
static_assert(kScratchBytes <= (size_t(48) << 20), "the allocation-free scratch stays under 48 MiB");

// Word indexes in kOffWords.
constexpr uint32_t kWordRowRange = 2;  // a row id >= the dataset's rows
constexpr uint32_t kWordNonUnitH = 3;  // some Hessian != 1.0 (the unit-Hessian forms do not apply)

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_setup_lock;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU
uint32_t g_lds_max[kMaxDevices] = {};

// The device address of this GPU's scratch array, looked up on its first use.
unsigned char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) fail(MPSB_ERR_INVALID_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lock(g_setup_lock);
  if (!g_scratch[device]) {
    void* p = nullptr;
    MPSB_HIP("allocation-free scratch lookup", hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(MPSB_ERR_LAUNCH, "allocation-free scratch lookup returned no address");
    int lds = 0;
    MPSB_HIP("device query", hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    g_lds_max[device] = uint32_t(lds);
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch[device];
}

__global__ void k_na_clear(uint32_t* words, unsigned long long* maxabs) {
  const uint32_t t = threadIdx.x;
  if (t < 64) words[t] = 0u;
  if (t < 2) maxabs[t] = 0ull;
}

// Row ids to 32-bit positions (an id past the dataset is flagged and read as row 0, so no
// kernel reads outside the codes), and the unit-Hessian test over the dataset's pairs.
__global__ void k_na_prepare(const uint64_t* __restrict__ rows, uint64_t n_rows, uint64_t data_rows,
                             uint32_t* __restrict__ rows32, const double* __restrict__ pairs, uint32_t* words) {
  const uint64_t n = n_rows > data_rows ? n_rows : data_rows;
  bool out_of_range = false, non_unit = false;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    if (i < n_rows) {
      const uint64_t r = rows[i];
      out_of_range |= r >= data_rows;
      rows32[i] = r < data_rows ? uint32_t(r) : 0u;
    }
    if (i < data_rows) non_unit |= pairs[2 * i + 1] != 1.0;
  }
  if (out_of_range) atomicOr(&words[kWordRowRange], 1u);
  if (non_unit) atomicOr(&words[kWordNonUnitH], 1u);
}

// The node histogram in the reference's view (cells feature after feature, bin 0 including the
// missing rows), from the device layout's cells: map[o] = device cell of reference cell o,
// map[cells + o] = its feature's missing cell to add, or 0xFFFFFFFF. Values: exact integers
// times the round's value scale (integer and fixed-point forms), FP64 bits (FP32 form), counts
// as the Hessian for unit-Hessian data; the same operations as the host's view of
// mpsb_build_histograms.
__global__ void k_na_view(const uint32_t* __restrict__ map, uint32_t n_ref, const uint32_t* __restrict__ cnt,
                          const uint64_t* __restrict__ g, const uint64_t* __restrict__ h,
                          const double* __restrict__ scales, int integer, int unit_h, uint64_t* out_cnt,
                          double* out_g, double* out_h) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_ref) return;
  const double sg = integer ? scales[2] : 1.0, sh = integer ? scales[3] : 1.0;
  auto G = [&](uint32_t i) {
    return integer ? double(static_cast<long long>(g[i])) * sg : __longlong_as_double(static_cast<long long>(g[i]));
  };
  auto H = [&](uint32_t i) {
    if (unit_h) return double(cnt[i]);
    return integer ? double(static_cast<long long>(h[i])) * sh : __longlong_as_double(static_cast<long long>(h[i]));
  };
  const uint32_t at = map[o], miss = map[n_ref + o];
  uint64_t c = cnt[at];
  double gv = G(at), hv = H(at);
  if (miss != 0xFFFFFFFFu) {
    c += cnt[miss];
    gv += G(miss);
    hv += H(miss);
  }
  out_cnt[o] = c;
  out_g[o] = gv;
  out_h[o] = hv;
}

template <class T>
void to_scratch(unsigned char* base, size_t off, const std::vector<T>& v, hipStream_t st) {
  if (!v.empty())
    MPSB_HIP("allocation-free table upload",
             hipMemcpyAsync(base + off, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

}  // namespace

void histograms_noalloc(const HostDataset& d, const void* codes, const uint64_t* rows, uint64_t n_rows,
                        const double* pairs, int precision, uint64_t* out_cnt, double* out_g, double* out_h,
                        hipStream_t st) {
  if (precision != MPSB_PRECISION_FP32 && precision != MPSB_PRECISION_INT16 && precision != MPSB_PRECISION_INT8 &&
      precision != MPSB_PRECISION_FP64_FIXED)
    fail(MPSB_ERR_INVALID_ARGUMENT, "the allocation-free histogram build takes the fp32, int16, int8 and fp64_fixed forms");
  if (n_rows == 0) fail(MPSB_ERR_TRAINING, "Node row set must not be empty");
  if (d.rows > kNoAllocMaxRows || n_rows > kNoAllocMaxRows)
    fail(MPSB_ERR_INVALID_ARGUMENT, "the allocation-free histogram build takes at most 1,048,576 dataset and node rows");
  if (d.cells > kNoAllocMaxCells)
    fail(MPSB_ERR_INVALID_ARGUMENT, "the allocation-free histogram build takes at most 32,768 histogram cells per node");
  int device = 0;
  MPSB_HIP("device query", hipGetDevice(&device));
  unsigned char* base = scratch_on(device);
  auto* words = reinterpret_cast<uint32_t*>(base + kOffWords);
  auto* maxabs = reinterpret_cast<unsigned long long*>(base + kOffMaxabs);
  auto* scales = reinterpret_cast<double*>(base + kOffScales);
  auto* rows32 = reinterpret_cast<uint32_t*>(base + kOffRows);

  // 1. Clear the words, convert and check the rows, test for unit Hessians.
  hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, st, words, maxabs);
  const uint64_t n = std::max<uint64_t>(n_rows, d.rows);
  const uint32_t blocks = uint32_t(std::min<uint64_t>(1024, (n + 255) / 256));
  hipLaunchKernelGGL(k_na_prepare, dim3(blocks), dim3(256), 0, st, rows, n_rows, d.rows, rows32, pairs, words);
  MPSB_HIP("allocation-free histogram", hipGetLastError());
  uint32_t head[4] = {0, 0, 0, 0};
  MPSB_HIP("allocation-free histogram", hipMemcpyAsync(head, words, sizeof head, hipMemcpyDeviceToHost, st));
  MPSB_HIP("allocation-free histogram", hipStreamSynchronize(st));
  if (head[kWordRowRange]) fail(MPSB_ERR_TRAINING, "Histogram row index is out of bounds");
  const bool unit_h = head[kWordNonUnitH] == 0;

  // 2. Launch shape, as a session with the default tuning would plan it.
  const Tuning tune;
  const int kind = hist_kind(precision);
  const uint32_t bpc = bytes_per_cell(kind, unit_h);
  const uint32_t lds_max = g_lds_max[device];
  TilePlan tiles;
  try {
    tiles = plan_tiles(d, bpc, std::min(lds_max, tune.lds_budget), tune.block, tune.wpt);
  } catch (const Error&) {
    tiles = plan_tiles(d, bpc, lds_max, tune.block, tune.wpt);
  }
  if (tiles.tiles.size() > kNoAllocMaxTiles)
    fail(MPSB_ERR_INVALID_ARGUMENT, "the allocation-free histogram build takes at most 64 feature tiles per row");
  const uint64_t nt = tiles.tiles.size();
  const uint64_t want = std::max<uint64_t>(8, tune.target_blocks / nt);
  uint64_t slab_rows = std::max<uint64_t>(tune.min_slab_rows, (n_rows + want - 1) / want);
  slab_rows = std::max<uint64_t>(slab_rows, (n_rows + kNoAllocMaxSlabs - 1) / kNoAllocMaxSlabs);
  slab_rows = std::min<uint64_t>(slab_rows, slab_cap(precision, unit_h));
  std::vector<SlabDesc> slabs;
  for (uint64_t lo = 0; lo < n_rows; lo += slab_rows)
    slabs.push_back({uint32_t(lo), uint32_t(std::min(n_rows, lo + slab_rows))});
  if (slabs.size() > kNoAllocMaxSlabs) fail(MPSB_ERR_INTERNAL, "allocation-free slab plan exceeds its maximum");
  const std::vector<uint32_t> reduce_tab = {0u, uint32_t(slabs.size()), 0u /*slot*/, 0xFFFFFFFFu /*no sibling*/,
                                            0u /*parent slot*/};
  // Reference-view cells and the device cells they read.
  std::vector<uint32_t> view;
  uint32_t n_ref = 0;
  for (uint32_t f = 0; f < d.features; ++f) n_ref += d.bin_count[f];
  view.resize(size_t(2) * n_ref);
  for (uint32_t f = 0, o = 0; f < d.features; ++f)
    for (uint32_t b = 0; b < d.bin_count[f]; ++b, ++o) {
      view[o] = d.cell_offset[f] + b;
      view[n_ref + o] = (b == 0 && d.has_missing(f)) ? d.cell_offset[f] + d.bin_count[f] : 0xFFFFFFFFu;
    }
  // The tables below are copied from host vectors: if anything fails before the final
  // synchronisation, the stream drains before they go out of scope.
  struct DrainOnThrow {
    hipStream_t st;
    bool armed = true;
    ~DrainOnThrow() {
      if (armed) (void)hipStreamSynchronize(st);
    }
  } drain{st};
  to_scratch(base, kOffTiles, tiles.tiles, st);
  to_scratch(base, kOffWordOff, tiles.word_off, st);
  to_scratch(base, kOffCellLds, tiles.cell_lds, st);
  to_scratch(base, kOffSlabs, slabs, st);
  to_scratch(base, kOffReduce, reduce_tab, st);
  to_scratch(base, kOffViewMap, view, st);

  // 3. Gradients in the form's arithmetic (the round's scales derived on the device).
  GradLaunch gl{};
  gl.source = 1;
  gl.pairs = pairs;
  gl.n = d.rows;
  gl.precision = precision;
  gl.unit_h = unit_h ? 1u : 0u;
  gl.err = words;
  gl.gh = base + kOffGH;
  gl.seed = 0;  // the stochastic-rounding stream of mpsb_build_histograms
  if (precision != MPSB_PRECISION_FP32) {
    gl.maxabs = maxabs;
    gl.scales = scales;
    launch_grad_maxabs(gl, st);
  }
  launch_grad_write(gl, st);

  // 4. Histogram slabs, then their fixed-order reduction into the node histogram.
  const uint64_t plane = uint64_t(slabs.size()) * d.cells;
  HistLaunch h{};
  h.codes = static_cast<const uint8_t*>(codes);
  h.row_stride_bytes = d.row_stride * d.code_bytes;
  h.rows = rows32;
  h.gh = base + kOffGH;
  h.gh_by_pos = 0;
  h.slabs = reinterpret_cast<const SlabDesc*>(base + kOffSlabs);
  h.n_slabs = uint32_t(slabs.size());
  h.tiles = reinterpret_cast<const TileDesc*>(base + kOffTiles);
  h.n_tiles = uint32_t(nt);
  h.word_cell_off = reinterpret_cast<const uint32_t*>(base + kOffWordOff);
  h.cell_lds = reinterpret_cast<const uint32_t*>(base + kOffCellLds);
  h.partial = base + kOffPartial;
  h.plane = plane;
  h.cells = d.cells;
  h.code_bytes = d.code_bytes;
  h.lds_bytes = tiles.lds_bytes;
  h.kind = kind;
  h.unit_h = unit_h ? 1 : 0;
  h.block = int(tune.block);
  h.wpt = int(tiles.wpt);
  h.interleaved = tiles.interleaved ? tiles.slots : 0u;
  launch_hist(h, st);
  const auto* tab = reinterpret_cast<const uint32_t*>(base + kOffReduce);
  LevelHist node{};
  node.g = reinterpret_cast<uint64_t*>(base + kOffHist);
  node.h = reinterpret_cast<uint64_t*>(base + kOffHist + size_t(kNoAllocMaxCells) * 8);
  node.cnt = reinterpret_cast<uint32_t*>(base + kOffHist + size_t(kNoAllocMaxCells) * 16);
  ReduceLaunch r{};
  r.partial = base + kOffPartial;
  r.plane = plane;
  r.cells = d.cells;
  r.slab_first = tab;
  r.build_slot = tab + 2;
  r.sibling_slot = reinterpret_cast<const int32_t*>(tab + 3);
  r.parent_slot = tab + 4;
  r.n_build = 1;
  r.out = node;
  r.parent = node;  // not read: the node has no sibling
  r.kind = kind;
  r.unit_h = unit_h ? 1 : 0;
  r.err = words;
  launch_reduce(r, st);

  // 5. The reference view into the caller's outputs, then the error word.
  const bool integer = (kind & 0xFF) != kHistF32;
  hipLaunchKernelGGL(k_na_view, dim3((n_ref + 255) / 256), dim3(256), 0, st,
                     reinterpret_cast<const uint32_t*>(base + kOffViewMap), n_ref, node.cnt, node.g, node.h, scales,
                     integer ? 1 : 0, unit_h ? 1 : 0, out_cnt, out_g, out_h);
  MPSB_HIP("allocation-free histogram", hipGetLastError());
  uint32_t flags[2] = {0, 0};
  MPSB_HIP("allocation-free histogram", hipMemcpyAsync(flags, words, sizeof flags, hipMemcpyDeviceToHost, st));
  MPSB_HIP("allocation-free histogram", hipStreamSynchronize(st));
  drain.armed = false;
  raise_device_errors(flags[0], flags[1], 0);
}

}  // namespace mpsb

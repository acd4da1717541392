// gfx950 kernels of the histogram-boosting hot path.
//
// Nothing here is a dense contraction, so there is no MFMA: every kernel is bound by HBM
// bandwidth or by LDS atomics. Wavefronts are 64 lanes; blocks are 256 threads (4 waves).
// All FP64 decision arithmetic is compiled with -ffp-contract=off so each operation rounds
// exactly as the reference's C++ does (src/core/objective.cpp:289-356).

#include <algorithm>
#include <type_traits>

#include "internal.hpp"

namespace mpsb {

// err[0] collects every bit raised, err[1] keeps the first one.
__device__ inline void raise_err(uint32_t* err, uint32_t bit) {
  if (err) {
    atomicOr(err, bit);
    atomicCAS(err + 1, 0u, bit);
  }
}

// ===================================================================== histogram ====
// Block -> (slab, tile). The T tiles of one slab run back to back on the same XCD
// (work-group b goes to XCD b % 8), so the slab's rows fetched from HBM by the first tile
// are L2 hits for the other T-1 tiles.
__device__ inline bool hist_item(uint32_t n_slabs, uint32_t n_tiles, uint32_t* slab,
                                 uint32_t* tile) {
  const uint32_t x = blockIdx.x & 7u, q = blockIdx.x >> 3;
  *tile = q % n_tiles;
  *slab = (q / n_tiles) * 8u + x;
  return *slab < n_slabs;
}

template <int KIND, bool UNIT_H>
struct GHType;
template <>
struct GHType<kHistF32, true> {
  using T = float;
};
template <>
struct GHType<kHistF32, false> {
  using T = float2;
};
template <>
struct GHType<kHistFix64, true> {
  using T = long long;
};
template <>
struct GHType<kHistFix64, false> {
  using T = longlong2;
};

// Integer-form element: int16 pair or int8 pair (QBYTES 2 or 1), or g alone when unit_h.
template <int QBYTES, bool UNIT_H>
struct QType;
template <>
struct QType<2, true> {
  using T = short;
};
template <>
struct QType<2, false> {
  using T = short2;
};
template <>
struct QType<1, true> {
  using T = signed char;
};
template <>
struct QType<1, false> {
  using T = char2;
};

// LDS accumulators. Cells n1 = ncells + 1 (the last is a trash cell for padding codes).
// kHistF32:   float g[n1], float h[n1] (unless unit_h), uint32 c[n1]
// kHistFix64: int64 g[n1], int64 h[n1] (unless unit_h), uint32 c[n1]
// kHistQ unit_h: uint64 pk[n1] = G:40 (signed) | count:24, one 64-bit atomic per update
// kHistQ:     uint64 pk[n1] = G:32 (signed) | H:32, uint32 c[n1]
// kHistG32:   int32 g[n1] (unit_h only; a slab's sum fits: at most slab_cap_g32 rows)
// kHistQNC:   uint64 pk[n1] = G:32 (signed) | H:32, no count (not unit_h)
// kHistF64NC: int64 g[n1], no count (unit_h only)
template <int KIND, bool UNIT_H, int QBYTES>
struct Acc {
  static constexpr int kPlanes8 = KIND == kHistF32 || KIND == kHistG32 ? 0 : (KIND == kHistFix64 ? (UNIT_H ? 1 : 2) : 1);
  static constexpr int kPlanes4 = KIND == kHistG32 ? 1 : (KIND == kHistQNC || KIND == kHistF64NC) ? 0 :
                                  KIND == kHistF32 ? (UNIT_H ? 2 : 3) : (KIND == kHistFix64 ? 1 : (UNIT_H ? 0 : 1));
};

// FP32 add into an LDS cell by compare-and-swap (the reference's AtomicAddFloat,
// src/mps/kernels/histogram.metal:22-33). On gfx950 this runs about 12x faster than the native
// ds_add_f32 at the histogram's access pattern (2.6 against 0.2 T lane updates/s, one
// conflict-free interleaved tile); the sum is the same FP32 addition in arrival order.
__device__ inline void lds_add_f32(float* cell, float v) {
  uint32_t* a = reinterpret_cast<uint32_t*>(cell);
  uint32_t seen = *a, want;
  do {
    want = seen;
    seen = atomicCAS(a, want, __float_as_uint(__uint_as_float(want) + v));
  } while (seen != want);
}

// Rotate the 4*W code bytes of a row left by `rot` bytes (byte i of the result is byte
// (i + rot) mod 4W of the input): words first, then the byte remainder with v_alignbyte.
template <int W>
__device__ inline void rotate_codes(const uint32_t* v, uint32_t rot, uint32_t* out) {
  const uint32_t q = rot >> 2, sh = rot & 3u;
  uint32_t t[W];
#pragma unroll
  for (int i = 0; i < W; ++i) {
    uint32_t x = v[i];
#pragma unroll
    for (int m = 1; m < W; ++m) x = q == uint32_t(m) ? v[(i + m) % W] : x;
    t[i] = x;
  }
#pragma unroll
  for (int i = 0; i < W; ++i) out[i] = __builtin_amdgcn_alignbyte(t[(i + 1) % W], t[i], sh);
}

// INTER: feature-interleaved tile (one-byte codes): cell (bin b, feature slot j) lives at
// b * SLOTS + j, one bank pair per feature. A thread's features are visited in an order rotated
// by its row slot, so in every atomic instruction the lanes of a half-wave update different
// features: no bank conflicts and no same-address collisions within the half. SLOTS is 32
// (8-byte cells: a 64 KB tile) or 16 (12-byte cells, [G|H] words plus a count plane: a 48 KB
// tile; lanes 16 apart share a slot but not a bank, as consecutive bins alternate bank halves),
// or 64 for kHistG32 (4-byte cells: 64 features in a 64 KB tile, one bank per feature, four
// threads per row whose rotations put the 64 lanes of a wave on 64 different slots).
template <typename CodeT, int KIND, bool UNIT_H, int QBYTES, int WPT, bool INTER, int SLOTS>
__global__ __launch_bounds__(1024) void k_hist(const uint8_t* __restrict__ codes,
                                              uint32_t row_stride_bytes,
                                              const uint32_t* __restrict__ rows,
                                              const void* __restrict__ gh_any,
                                              const SlabDesc* __restrict__ slabs, uint32_t n_slabs,
                                              const TileDesc* __restrict__ tiles, uint32_t n_tiles,
                                              const uint32_t* __restrict__ word_cell_off,
                                              const uint32_t* __restrict__ cell_lds,
                                              void* __restrict__ partial, uint64_t plane,
                                              uint32_t cells, uint32_t gh_by_pos,
                                              const uint32_t* __restrict__ dyn_slabs) {
  extern __shared__ __align__(16) unsigned char lds[];
  uint32_t slab, tile;
  if (!hist_item(dyn_slabs ? *dyn_slabs : n_slabs, n_tiles, &slab, &tile)) return;
  const TileDesc td = tiles[tile];
  const SlabDesc sd = slabs[slab];
  const uint32_t n1 = td.lds_cells + 1;

  using Q = typename QType<(QBYTES == 0 ? 2 : QBYTES), UNIT_H>::T;
  // LDS planes
  unsigned long long* p8a = reinterpret_cast<unsigned long long*>(lds);
  unsigned long long* p8b = p8a + n1;
  constexpr int P8 = Acc<KIND, UNIT_H, QBYTES>::kPlanes8;
  constexpr int P4 = Acc<KIND, UNIT_H, QBYTES>::kPlanes4;
  uint32_t* p4 = reinterpret_cast<uint32_t*>(lds + size_t(P8) * n1 * 8);
  float* fg = reinterpret_cast<float*>(p4);            // kHistF32: g
  float* fh = fg + n1;                                 // kHistF32: h (unless unit_h)
  uint32_t* cnt = p4 + size_t(P4 > 0 ? P4 - 1 : 0) * n1;  // count plane: the last 4-byte plane

  for (uint32_t i = threadIdx.x; i < n1 * P8; i += blockDim.x) p8a[i] = 0ull;
  for (uint32_t i = threadIdx.x; i < n1 * P4; i += blockDim.x) p4[i] = 0u;
  __syncthreads();

  // Each thread loads WPT consecutive 32-bit words (4*WPT bytes) of a row: with WPT = 4 a
  // 64-lane load instruction covers 32 rows of a 32-feature tile, which keeps enough
  // independent row gathers in flight to hide HBM latency.
  constexpr int CPW = 4 / int(sizeof(CodeT));  // codes per 32-bit word
  constexpr uint32_t CMASK = sizeof(CodeT) == 1 ? 0xFFu : 0xFFFFu;
  constexpr bool kInt = KIND == kHistQ || KIND == kHistG32 || KIND == kHistQNC;  // quantized gradients
  // The form of the gradients read: FP32 (a placeholder for the quantized kinds), fixed-point.
  constexpr int kGH = kInt ? kHistF32 : (KIND == kHistF64NC ? kHistFix64 : KIND);
  using GT = typename GHType<kGH, UNIT_H>::T;
  using VecT = typename std::conditional<WPT == 4, uint4, typename std::conditional<WPT == 2, uint2, uint32_t>::type>::type;
  const uint32_t tpr = td.words / WPT;  // threads per row; tile words are a multiple of 4
  const uint32_t rpp = blockDim.x / tpr;
  const uint32_t w = threadIdx.x % tpr, rs = threadIdx.x / tpr;
  const uint32_t rrep = (rs & 15u) << 4;  // this thread's replica of the replicated features
  constexpr uint32_t FPT = uint32_t(WPT * CPW);  // features per thread
  // Rotation: the row slot, plus half a thread's features for the upper half of the threads
  // of a row. Any 16 consecutive lanes then update 16 different feature slots modulo 16, and
  // any 32 consecutive lanes 32 different slots, so neither a 16-lane nor a 32-lane LDS pass
  // meets two lanes on one bank pair.
  const uint32_t rot = INTER ? ((rs + (FPT / 2u) * (w / (tpr > 1u ? tpr / 2u : 1u))) & (FPT - 1u)) : 0u;
  const uint32_t fbase = w * FPT;
  if (rs < rpp) {
    uint32_t off[WPT * CPW];
#pragma unroll
    for (int j = 0; j < WPT; ++j)
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if constexpr (INTER) {
          off[j * CPW + k] = fbase + ((uint32_t(j * CPW + k) + rot) & (FPT - 1u));  // feature slot
        } else {
          const uint32_t o = word_cell_off[(size_t(tile) * 64 + w * WPT + j) * 4 + k];
          off[j * CPW + k] = (o & ~kReplicated) + ((o & kReplicated) ? rrep : 0u);
        }
      }
    const VecT* codesv = reinterpret_cast<const VecT*>(codes);
    const size_t stride_v = row_stride_bytes / (4 * WPT);
    const size_t col_v = td.word0 / WPT + w;
    constexpr int U = 2;  // rows in flight per thread (4 needs 93 VGPRs: one block per CU)
    // Software pipeline: the code words and gradients of the next U rows are loaded while the
    // current rows' atomics run, and row ids two steps ahead, so no load waits on another
    // inside an iteration. Positions past the slab are clamped to its last row (in-bounds
    // loads whose results are never used); atomics only for positions inside the slab.
    const uint64_t step = uint64_t(rpp) * U;
    const uint64_t last = uint64_t(sd.end) - 1u;
    auto pos_at = [&](uint64_t p) -> uint32_t { return uint32_t(p < last ? p : last); };
    auto row_at = [&](uint32_t pp) -> uint32_t { return rows ? rows[pp] : pp; };
    uint64_t p = uint64_t(sd.begin) + rs;
    uint32_t rid[U] = {};
    VecT word[U] = {};
    GT gv[U] = {};
    Q qv[U] = {};
    auto load_rows = [&](uint64_t base, const uint32_t* r, VecT* wd, GT* g, Q* q) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t pp = pos_at(base + uint64_t(u) * rpp);
        wd[u] = codesv[size_t(r[u]) * stride_v + col_v];
        // Gradients travel with the rows through the partition, so the engine reads them
        // by position (contiguous); the test hooks pass them by row.
        const uint32_t gi = gh_by_pos ? pp : r[u];
        if constexpr (kInt) {
          q[u] = reinterpret_cast<const Q*>(gh_any)[gi];
        } else {
          g[u] = reinterpret_cast<const typename GHType<kGH, UNIT_H>::T*>(gh_any)[gi];
        }
      }
    };
    if (sd.begin < sd.end) {  // uniform over the block
#pragma unroll
      for (int u = 0; u < U; ++u) rid[u] = row_at(pos_at(p + uint64_t(u) * rpp));
      load_rows(p, rid, word, gv, qv);
#pragma unroll
      for (int u = 0; u < U; ++u) rid[u] = row_at(pos_at(p + step + uint64_t(u) * rpp));
    }
    for (; p < sd.end; p += step) {
      VecT word_n[U] = {};
      GT gv_n[U] = {};
      Q qv_n[U] = {};
      uint32_t rid_n[U];
      load_rows(p + step, rid, word_n, gv_n, qv_n);
#pragma unroll
      for (int u = 0; u < U; ++u) rid_n[u] = row_at(pos_at(p + 2 * step + uint64_t(u) * rpp));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (p + uint64_t(u) * rpp >= sd.end) continue;
        const uint32_t* wv = reinterpret_cast<const uint32_t*>(&word[u]);
        uint32_t rv[WPT];
        if constexpr (INTER) {
          rotate_codes<WPT>(wv, rot, rv);
          wv = rv;
        }
#pragma unroll
        for (int jk = 0; jk < WPT * CPW; ++jk) {
          const int j = jk / CPW, k = jk % CPW;
          const uint32_t code = (wv[j] >> (k * 8 * int(sizeof(CodeT)))) & CMASK;
          const uint32_t c = INTER ? code * uint32_t(SLOTS) + off[jk] : off[jk] + code;
          if constexpr (KIND == kHistF32) {
            if constexpr (UNIT_H) {
              lds_add_f32(&fg[c], gv[u]);
            } else {
              lds_add_f32(&fg[c], gv[u].x);
              lds_add_f32(&fh[c], gv[u].y);
            }
            atomicAdd(&cnt[c], 1u);
          } else if constexpr (KIND == kHistG32) {
            atomicAdd(&p4[c], static_cast<uint32_t>(static_cast<int32_t>(qv[u])));
          } else if constexpr (KIND == kHistQNC) {
            const long long gq = static_cast<long long>(qv[u].x);
            const unsigned long long hq = static_cast<unsigned long long>(static_cast<long long>(qv[u].y) & 0xFFFFFFFFll);
            atomicAdd(&p8a[c], (static_cast<unsigned long long>(gq) << 32) + hq);
          } else if constexpr (KIND == kHistF64NC) {
            atomicAdd(&p8a[c], static_cast<unsigned long long>(gv[u]));
          } else if constexpr (KIND == kHistFix64) {
            if constexpr (UNIT_H) {
              atomicAdd(&p8a[c], static_cast<unsigned long long>(gv[u]));
            } else {
              atomicAdd(&p8a[c], static_cast<unsigned long long>(gv[u].x));
              atomicAdd(&p8b[c], static_cast<unsigned long long>(gv[u].y));
            }
            atomicAdd(&cnt[c], 1u);
          } else {
            if constexpr (UNIT_H) {
              const long long gq = static_cast<long long>(qv[u]);
              atomicAdd(&p8a[c], (static_cast<unsigned long long>(gq) << 24) + 1ull);
            } else {
              const long long gq = static_cast<long long>(qv[u].x);
              const unsigned long long hq = static_cast<unsigned long long>(
                  static_cast<long long>(qv[u].y) & 0xFFFFFFFFll);
              atomicAdd(&p8a[c], (static_cast<unsigned long long>(gq) << 32) + hq);
              atomicAdd(&cnt[c], 1u);
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        word[u] = word_n[u];
        gv[u] = gv_n[u];
        qv[u] = qv_n[u];
        rid[u] = rid_n[u];
      }
    }
  }
  __syncthreads();

  // Flush the tile to this slab's partial slot (each slot written by exactly one block); a
  // replicated feature's 16 replicas are summed in replica order.
  const size_t base = size_t(slab) * cells + td.cell0;
  unsigned char* pbytes = reinterpret_cast<unsigned char*>(partial);
  for (uint32_t c = threadIdx.x; c < td.ncells; c += blockDim.x) {
    const uint32_t m = cell_lds[td.cell0 + c];
    const uint32_t l0 = m & ~kReplicated;
    const uint32_t nrep = (m & kReplicated) ? 16u : 1u;
    if constexpr (KIND == kHistF32) {
      float sg = 0.0f, sh = 0.0f;
      uint32_t sc = 0;
      for (uint32_t k = 0; k < nrep; ++k) {
        sg += fg[l0 + 16 * k];
        if constexpr (!UNIT_H) sh += fh[l0 + 16 * k];
        sc += cnt[l0 + 16 * k];
      }
      float* og = reinterpret_cast<float*>(pbytes);
      float* oh = og + plane;
      uint32_t* oc = reinterpret_cast<uint32_t*>(pbytes + 8 * plane);
      og[base + c] = sg;
      if constexpr (!UNIT_H) oh[base + c] = sh;
      oc[base + c] = sc;
    } else if constexpr (KIND == kHistFix64) {
      unsigned long long sg = 0, sh = 0;
      uint32_t sc = 0;
      for (uint32_t k = 0; k < nrep; ++k) {
        sg += p8a[l0 + 16 * k];
        if constexpr (!UNIT_H) sh += p8b[l0 + 16 * k];
        sc += cnt[l0 + 16 * k];
      }
      unsigned long long* og = reinterpret_cast<unsigned long long*>(pbytes);
      unsigned long long* oh = og + plane;
      uint32_t* oc = reinterpret_cast<uint32_t*>(pbytes + 16 * plane);
      og[base + c] = sg;
      if constexpr (!UNIT_H) oh[base + c] = sh;
      oc[base + c] = sc;
    } else if constexpr (KIND == kHistG32) {
      uint32_t sg = 0;
      for (uint32_t k = 0; k < nrep; ++k) sg += p4[l0 + 16 * k];
      reinterpret_cast<uint32_t*>(pbytes)[base + c] = sg;
    } else if constexpr (KIND == kHistQNC || KIND == kHistF64NC) {
      unsigned long long spk = 0;  // integer words add as integers, replicas too
      for (uint32_t k = 0; k < nrep; ++k) spk += p8a[l0 + 16 * k];
      reinterpret_cast<unsigned long long*>(pbytes)[base + c] = spk;
    } else {
      // Packed [G | count] words add as integers: a slab's count per cell stays below the
      // count field's range, so the replica sum is the packed sum.
      unsigned long long spk = 0;
      uint32_t sc = 0;
      for (uint32_t k = 0; k < nrep; ++k) {
        spk += p8a[l0 + 16 * k];
        if constexpr (!UNIT_H) sc += cnt[l0 + 16 * k];
      }
      unsigned long long* opk = reinterpret_cast<unsigned long long*>(pbytes);
      uint32_t* oc = reinterpret_cast<uint32_t*>(pbytes + 8 * plane);
      opk[base + c] = spk;
      if constexpr (!UNIT_H) oc[base + c] = sc;
    }
  }
}

template <typename CodeT, int KIND, bool UNIT_H, int QBYTES>
static void hist_dispatch(const HistLaunch& a, hipStream_t s) {
  const uint32_t groups = (a.n_slabs + 7) / 8;
  const uint32_t blocks = groups * 8 * a.n_tiles;
  if (blocks == 0) return;
#define MPSB_HIST(W, I, S)                                                                              \
  hipLaunchKernelGGL((k_hist<CodeT, KIND, UNIT_H, QBYTES, W, I, S>), dim3(blocks), dim3(a.block), a.lds_bytes, s, \
                     a.codes, a.row_stride_bytes, a.rows, a.gh, a.slabs, a.n_slabs, a.tiles, a.n_tiles, \
                     a.word_cell_off, a.cell_lds, a.partial, a.plane, a.cells, a.gh_by_pos, a.dyn_slabs)
  constexpr bool kCanInterleave = sizeof(CodeT) == 1;
  if constexpr (KIND == kHistG32) {
    // 64-slot tiles of 4-byte cells: four threads per row of a 16-word tile.
    if (!kCanInterleave || a.interleaved != 64 || a.wpt != 4 || !UNIT_H)
      fail(MPSB_ERR_INTERNAL, "gradient-only histograms need 64-slot tiles of one-byte codes");
    MPSB_HIST(4, kCanInterleave, 64);
    return;
  }
  if (kCanInterleave && a.interleaved == 64) {
    // 64-slot tiles of 8-byte cells (a 128 KB tile): four threads per row of a 16-word tile.
    if (a.wpt != 4) fail(MPSB_ERR_INTERNAL, "64-slot histogram tiles need four words per thread");
    MPSB_HIST(4, kCanInterleave, 64);
  } else if (kCanInterleave && a.interleaved == 16) {
    // 16-slot tiles are four words of the row: one thread per row (the planner sets wpt 4).
    if (a.wpt != 4) fail(MPSB_ERR_INTERNAL, "16-slot histogram tiles need four words per thread");
    MPSB_HIST(4, kCanInterleave, 16);
  } else if (kCanInterleave && a.interleaved) {
    if (a.wpt == 4) {
      MPSB_HIST(4, kCanInterleave, 32);
    } else if (a.wpt == 2) {
      MPSB_HIST(2, kCanInterleave, 32);
    } else {
      MPSB_HIST(1, kCanInterleave, 32);
    }
  } else if (a.wpt == 4) {
    MPSB_HIST(4, false, 32);
  } else if (a.wpt == 2) {
    MPSB_HIST(2, false, 32);
  } else {
    MPSB_HIST(1, false, 32);
  }
#undef MPSB_HIST
}

template <typename CodeT>
static void hist_by_kind(const HistLaunch& a, hipStream_t s) {
  // kind encodes the integer width too: kHistQ with code_bytes-independent QBYTES passed in
  // the upper bits of `kind` (kind & 0xFF = HistKind, kind >> 8 = bytes per quantized value).
  const int kind = a.kind & 0xFF, qb = a.kind >> 8;
  if (kind == kHistG32) {
    if (!a.unit_h) fail(MPSB_ERR_INTERNAL, "gradient-only histograms need a unit Hessian");
    qb == 2 ? hist_dispatch<CodeT, kHistG32, true, 2>(a, s) : hist_dispatch<CodeT, kHistG32, true, 1>(a, s);
  } else if (kind == kHistQNC) {
    if (a.unit_h) fail(MPSB_ERR_INTERNAL, "count-free [G|H] histograms need Hessians");
    qb == 2 ? hist_dispatch<CodeT, kHistQNC, false, 2>(a, s) : hist_dispatch<CodeT, kHistQNC, false, 1>(a, s);
  } else if (kind == kHistF64NC) {
    if (!a.unit_h) fail(MPSB_ERR_INTERNAL, "count-free fixed-point histograms need a unit Hessian");
    hist_dispatch<CodeT, kHistF64NC, true, 0>(a, s);
  } else if (kind == kHistF32) {
    a.unit_h ? hist_dispatch<CodeT, kHistF32, true, 0>(a, s) : hist_dispatch<CodeT, kHistF32, false, 0>(a, s);
  } else if (kind == kHistFix64) {
    a.unit_h ? hist_dispatch<CodeT, kHistFix64, true, 0>(a, s) : hist_dispatch<CodeT, kHistFix64, false, 0>(a, s);
  } else if (qb == 2) {
    a.unit_h ? hist_dispatch<CodeT, kHistQ, true, 2>(a, s) : hist_dispatch<CodeT, kHistQ, false, 2>(a, s);
  } else {
    a.unit_h ? hist_dispatch<CodeT, kHistQ, true, 1>(a, s) : hist_dispatch<CodeT, kHistQ, false, 1>(a, s);
  }
}

void launch_hist(const HistLaunch& a, hipStream_t s) {
  if (a.code_bytes == 1) {
    hist_by_kind<uint8_t>(a, s);
  } else {
    hist_by_kind<uint16_t>(a, s);
  }
}

// ======================================================================= reduce ======
// Node histogram = sum of its slabs' partials, in slab order (fixed order: the FP32 form is
// deterministic run to run); the sibling is parent - built in the same pass
// (SubtractHistograms, src/core/tree_histogram.cpp:16-66, including its Hessian clamp).
__device__ inline double subtract_hessian(double p, double c, uint32_t* err) {
  if (!isfinite(p) || !isfinite(c) || p < 0.0 || c < 0.0) {
    raise_err(err, kErrSubtract);
    return 0.0;
  }
  const double d = p - c;
  const double tol = 1e-5 * fmax(1.0, fmax(fabs(p), fabs(c)));
  if (d < -tol) raise_err(err, kErrSubtract);
  return d < 0.0 ? 0.0 : d;
}

// Block = 64 cells x 4 wavefronts; wavefront w sums its quarter of the node's slabs in slab
// order, then the four quarters are added in wavefront order: a fixed association, so the
// FP32 form's FP64 totals are identical run to run.
template <int KIND, bool UNIT_H>
__global__ __launch_bounds__(256) void k_reduce(const unsigned char* __restrict__ partial,
                                                uint64_t plane, uint32_t cells,
                                                const uint32_t* __restrict__ slab_first,
                                                const uint32_t* __restrict__ build_slot,
                                                const int32_t* __restrict__ sibling_slot,
                                                const uint32_t* __restrict__ parent_slot,
                                                uint32_t n_build, const uint32_t* __restrict__ dyn_build,
                                                LevelHist out, LevelHist parent, uint32_t* err) {
  __shared__ unsigned long long s_cnt[4][64];
  __shared__ double s_fg[4][64], s_fh[4][64];
  __shared__ long long s_ig[4][64], s_ih[4][64];
  // Blocks loop over builds (gridDim.y may be below the count: grids sized for the largest
  // level would be mostly idle).
  const uint32_t nb = dyn_build ? *dyn_build : n_build;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * 64 + lane;
  for (uint32_t b = blockIdx.y; b < nb; b += gridDim.y) {
    const uint32_t s0 = slab_first[b], s1 = slab_first[b + 1];
    const uint32_t per = (s1 - s0 + 3) / 4;
    const uint32_t q0 = s0 + wave * per, q1 = min(s1, q0 + per);
    uint64_t count = 0;
    double fg = 0.0, fh = 0.0;
    long long ig = 0, ih = 0;
    // kReduceU slabs per step, every load issued before the adds (the sums stay in slab order:
    // integer sums are exact, the FP32 form's FP64 additions keep their sequence).
    constexpr uint32_t kReduceU = 8;
    for (uint32_t s = q0; c < cells && s < q1; s += kReduceU) {
      uint64_t w0[kReduceU], w1[kReduceU];
      uint32_t w2[kReduceU];
  #pragma unroll
      for (uint32_t u = 0; u < kReduceU; ++u) {
        const bool in = s + u < q1;
        const size_t at = size_t(in ? s + u : s) * cells + c;
        w0[u] = w1[u] = 0;
        w2[u] = 0;
        if constexpr (KIND == kHistF32) {
          w0[u] = in ? reinterpret_cast<const uint32_t*>(partial)[at] : 0u;
          if constexpr (!UNIT_H) w1[u] = in ? reinterpret_cast<const uint32_t*>(partial)[plane + at] : 0u;
          w2[u] = in ? reinterpret_cast<const uint32_t*>(partial + 8 * plane)[at] : 0u;
        } else if constexpr (KIND == kHistFix64) {
          w0[u] = in ? reinterpret_cast<const unsigned long long*>(partial)[at] : 0ull;
          if constexpr (!UNIT_H) w1[u] = in ? reinterpret_cast<const unsigned long long*>(partial)[plane + at] : 0ull;
          w2[u] = in ? reinterpret_cast<const uint32_t*>(partial + 16 * plane)[at] : 0u;
        } else {
          w0[u] = in ? reinterpret_cast<const unsigned long long*>(partial)[at] : 0ull;
          if constexpr (!UNIT_H) w2[u] = in ? reinterpret_cast<const uint32_t*>(partial + 8 * plane)[at] : 0u;
        }
      }
  #pragma unroll
      for (uint32_t u = 0; u < kReduceU; ++u) {
        if (s + u >= q1) break;
        if constexpr (KIND == kHistF32) {
          fg += double(__uint_as_float(uint32_t(w0[u])));
          if constexpr (!UNIT_H) fh += double(__uint_as_float(uint32_t(w1[u])));
          count += w2[u];
        } else if constexpr (KIND == kHistFix64) {
          ig += static_cast<long long>(w0[u]);
          if constexpr (!UNIT_H) ih += static_cast<long long>(w1[u]);
          count += w2[u];
        } else if constexpr (UNIT_H) {
          count += w0[u] & 0xFFFFFFull;
          ig += static_cast<long long>(w0[u]) >> 24;
        } else {
          ig += static_cast<long long>(w0[u]) >> 32;
          ih += static_cast<long long>(w0[u] & 0xFFFFFFFFull);
          count += w2[u];
        }
      }
    }
    s_cnt[wave][lane] = count;
    s_fg[wave][lane] = fg;
    s_fh[wave][lane] = fh;
    s_ig[wave][lane] = ig;
    s_ih[wave][lane] = ih;
    __syncthreads();
    if (wave == 0 && c < cells) {
      count = 0;
      fg = fh = 0.0;
      ig = ih = 0;
      for (int w = 0; w < 4; ++w) {
        count += s_cnt[w][lane];
        fg += s_fg[w][lane];
        fh += s_fh[w][lane];
        ig += s_ig[w][lane];
        ih += s_ih[w][lane];
      }
      const size_t o = size_t(build_slot[b]) * cells + c;
      out.cnt[o] = uint32_t(count);
      if constexpr (KIND == kHistF32) {
        out.g[o] = __double_as_longlong(fg);
        if constexpr (!UNIT_H) out.h[o] = __double_as_longlong(fh);
      } else {
        out.g[o] = static_cast<uint64_t>(ig);
        if constexpr (!UNIT_H) out.h[o] = static_cast<uint64_t>(ih);
      }
      const int32_t sib = sibling_slot[b];
      if (sib >= 0) {  // the sibling: parent - built
        const size_t po = size_t(parent_slot[b]) * cells + c;
        const size_t so = size_t(sib) * cells + c;
        const uint32_t pc = parent.cnt[po];
        if (count > pc) raise_err(err, kErrSubtract);
        out.cnt[so] = pc - uint32_t(count);
        if constexpr (KIND == kHistF32) {
          const double pg = __longlong_as_double(parent.g[po]);
          const double g = pg - fg;
          if (!isfinite(pg) || !isfinite(g)) raise_err(err, kErrSubtract);
          out.g[so] = __double_as_longlong(g);
          if constexpr (!UNIT_H) out.h[so] = __double_as_longlong(subtract_hessian(__longlong_as_double(parent.h[po]), fh, err));
        } else {
          out.g[so] = static_cast<uint64_t>(static_cast<long long>(parent.g[po]) - ig);
          if constexpr (!UNIT_H) {
            const long long hd = static_cast<long long>(parent.h[po]) - ih;
            if (hd < 0) raise_err(err, kErrSubtract);
            out.h[so] = static_cast<uint64_t>(hd);
          }
        }
      }
    }
    __syncthreads();  // the sums of the next build reuse the LDS
  }
}

// Root histograms without counts (kHistG32, kHistQNC, kHistF64NC): the node's sums are its
// slabs' partials added as integers (exact), its counts the root's per-cell counts. Built nodes
// only: a root has no sibling. Block = 64 cells x 16 wavefronts, each summing a sixteenth of
// the slabs (a root has hundreds: the sums are latency-bound chains of loads, kept short).
constexpr uint32_t kReduceRootWaves = 16;
template <int KIND>
__global__ __launch_bounds__(64 * kReduceRootWaves) void k_reduce_root(const void* __restrict__ partial_any,
                                                    uint32_t cells, const uint32_t* __restrict__ slab_first,
                                                    const uint32_t* __restrict__ build_slot, uint32_t n_build,
                                                    const uint32_t* __restrict__ counts, LevelHist out) {
  constexpr uint32_t W = kReduceRootWaves;
  // G:32|H:32 words carry H in the low half: the pair sums apart.
  constexpr bool kPair = KIND == kHistQNC;
  using P = typename std::conditional<KIND == kHistG32, uint32_t, unsigned long long>::type;
  const P* __restrict__ partial = static_cast<const P*>(partial_any);
  __shared__ long long s_g[W][64];
  __shared__ long long s_h[kPair ? W : 1][64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * 64 + lane;
  for (uint32_t b = blockIdx.y; b < n_build; b += gridDim.y) {
    const uint32_t s0 = slab_first[b], s1 = slab_first[b + 1];
    const uint32_t per = (s1 - s0 + W - 1) / W;
    const uint32_t q0 = s0 + wave * per, q1 = min(s1, q0 + per);
    long long g = 0, h = 0;
    constexpr uint32_t kU = 8;  // slabs per step, every load issued before the adds
    for (uint32_t sl = q0; c < cells && sl < q1; sl += kU) {
      P v[kU];
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) v[u] = sl + u < q1 ? partial[size_t(sl + u) * cells + c] : P(0);
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) {
        if constexpr (KIND == kHistG32) {
          g += static_cast<int32_t>(v[u]);
        } else if constexpr (kPair) {
          g += static_cast<long long>(v[u]) >> 32;
          h += static_cast<long long>(v[u] & 0xFFFFFFFFull);
        } else {
          g += static_cast<long long>(v[u]);
        }
      }
    }
    s_g[wave][lane] = g;
    if constexpr (kPair) s_h[wave][lane] = h;
    __syncthreads();
    if (wave == 0 && c < cells) {
      long long tg = 0, th = 0;
      for (uint32_t k = 0; k < W; ++k) {
        tg += s_g[k][lane];
        if constexpr (kPair) th += s_h[k][lane];
      }
      const size_t o = size_t(build_slot[b]) * cells + c;
      out.cnt[o] = counts[c];
      out.g[o] = static_cast<uint64_t>(tg);
      if constexpr (kPair) out.h[o] = static_cast<uint64_t>(th);
    }
    __syncthreads();
  }
}

__global__ void k_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

void launch_copy_u32(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  const uint32_t blocks = uint32_t(std::min<uint64_t>((n + 255) / 256, 1024));
  hipLaunchKernelGGL(k_copy_u32, dim3(blocks), dim3(256), 0, s, dst, src, n);
}

void launch_reduce(const ReduceLaunch& a, hipStream_t s) {
  if (a.n_build == 0) return;
  if (hist_kind_no_count(a.kind)) {
    if (a.dyn_build || !a.counts) fail(MPSB_ERR_INTERNAL, "a root reduce without counts needs host-planned builds and counts");
    const dim3 grid((a.cells + 63) / 64, std::min(a.n_build, 64u)), block(64 * kReduceRootWaves);
#define MPSB_REDUCE_ROOT(K) \
  hipLaunchKernelGGL(k_reduce_root<K>, grid, block, 0, s, a.partial, a.cells, a.slab_first, a.build_slot, a.n_build, \
                     a.counts, a.out)
    switch (a.kind & 0xFF) {
      case kHistG32: MPSB_REDUCE_ROOT(kHistG32); break;
      case kHistQNC: MPSB_REDUCE_ROOT(kHistQNC); break;
      default: MPSB_REDUCE_ROOT(kHistF64NC);
    }
#undef MPSB_REDUCE_ROOT
    return;
  }
  // With the count on the device, n_build is a bound: at most 64 build rows of blocks, each
  // looping over builds.
  const dim3 grid((a.cells + 63) / 64, a.dyn_build ? std::min(a.n_build, 64u) : a.n_build);
  const unsigned char* p = reinterpret_cast<const unsigned char*>(a.partial);
  const int kind = a.kind & 0xFF;
#define MPSB_REDUCE(K, U)                                                                      \
  hipLaunchKernelGGL((k_reduce<K, U>), grid, dim3(256), 0, s, p, a.plane, a.cells, a.slab_first, \
                     a.build_slot, a.sibling_slot, a.parent_slot, a.n_build, a.dyn_build, a.out, a.parent, a.err)
  if (kind == kHistF32) {
    if (a.unit_h) MPSB_REDUCE(kHistF32, true); else MPSB_REDUCE(kHistF32, false);
  } else if (kind == kHistFix64) {
    if (a.unit_h) MPSB_REDUCE(kHistFix64, true); else MPSB_REDUCE(kHistFix64, false);
  } else {
    if (a.unit_h) MPSB_REDUCE(kHistQ, true); else MPSB_REDUCE(kHistQ, false);
  }
#undef MPSB_REDUCE
}

// ===================================================================== baseline ======
// One thread per reference-view cell scanning every node row in row order, FP64: the
// counterpart of the reference's histogram_baseline kernel, used as a second opinion.
__global__ void k_baseline(const uint8_t* __restrict__ codes, uint32_t row_stride, uint32_t code_bytes,
                           const uint32_t* __restrict__ rows, uint32_t n_rows,
                           const double* __restrict__ pairs, const uint32_t* __restrict__ cell_feature,
                           const uint32_t* __restrict__ cell_code, uint32_t cells,
                           unsigned long long* cnt, double* g, double* h) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const uint32_t f = cell_feature[c];
  const uint32_t want = cell_code[c] & 0xFFFFu;
  const uint32_t miss = cell_code[c] >> 16;  // 0xFFFF: no missing code merged into this cell
  unsigned long long n = 0;
  double sg = 0.0, sh = 0.0;
  for (uint32_t i = 0; i < n_rows; ++i) {
    const uint32_t r = rows[i];
    const size_t at = (size_t(r) * row_stride + f) * code_bytes;
    const uint32_t code = code_bytes == 1 ? codes[at] : (uint32_t(codes[at]) | (uint32_t(codes[at + 1]) << 8));
    if (code == want || code == miss) {
      ++n;
      sg += pairs[2 * size_t(r)];
      sh += pairs[2 * size_t(r) + 1];
    }
  }
  cnt[c] = n;
  g[c] = sg;
  h[c] = sh;
}

void launch_baseline_hist(const uint8_t* codes, uint32_t row_stride, uint32_t code_bytes,
                          const uint32_t* rows, uint32_t n_rows, const double* pairs,
                          const uint32_t* cell_feature, const uint32_t* cell_code, uint32_t cells,
                          uint64_t* cnt, double* g, double* h, hipStream_t s) {
  hipLaunchKernelGGL(k_baseline, dim3((cells + 127) / 128), dim3(128), 0, s, codes, row_stride,
                     code_bytes, rows, n_rows, pairs, cell_feature, cell_code, cells,
                     reinterpret_cast<unsigned long long*>(cnt), g, h);
}

// ============================================================ row-order FP64 form ======
// The reference's own arithmetic (CpuReferenceBackend::BuildHistograms,
// src/cpu/reference_backend.cpp:22-71): every cell is an FP64 sum over the node's rows in
// row order. Lane j of a wavefront owns feature feat_list[group * fg + j] and walks the
// rows one by one, so each of its cells sees its additions in row order; the rows' codes
// of all lanes come from one row line. A missing row adds to bin 0 (the reference's
// NaN -> bin 0) and to the feature's missing cell, which thereby holds PrepareSplitRows'
// row-order missing statistics (src/core/tree_split.cpp:244-259). missing_only recomputes
// just the missing cells of a sibling whose bins came from subtraction.
template <typename CodeT, bool GLOBAL>
__global__ __launch_bounds__(64) void k_hist_seq(SeqLaunch a) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr uint32_t U = 8;  // rows whose loads are issued ahead of the serial updates
  const uint32_t groups = (a.n_feat + a.fg - 1) / a.fg;
  const uint32_t node_i = blockIdx.x / groups, grp = blockIdx.x % groups;
  const uint32_t lane = threadIdx.x;
  const uint32_t k = grp * a.fg + lane;
  const bool active = lane < a.fg && k < a.n_feat;
  const SeqNode nd = a.nodes[node_i];
  const uint32_t f = active ? a.feat_list[k] : 0u;
  const uint32_t B = a.bin_count[f];
  const uint32_t off = a.cell_offset[f];
  const uint32_t ncell = a.cell_offset[f + 1] - off;
  const bool has_miss = ncell > B;  // the missing cell is local index B, code B
  const size_t obase = size_t(nd.slot) * a.cells + off;
  double* ag;
  double* ah;
  uint32_t* ac;
  if constexpr (GLOBAL) {
    ag = reinterpret_cast<double*>(a.out.g + obase);
    ah = reinterpret_cast<double*>(a.out.h + obase);
    ac = a.out.cnt + obase;
  } else {
    const uint32_t per = a.fg * a.cpf;
    ag = reinterpret_cast<double*>(lds) + size_t(lane) * a.cpf;
    ac = reinterpret_cast<uint32_t*>(lds + size_t(per) * (a.unit_h ? 8 : 16)) + size_t(lane) * a.cpf;
    ah = reinterpret_cast<double*>(lds + size_t(per) * 8) + size_t(lane) * a.cpf;
  }
  const uint32_t c0 = a.missing_only ? B : 0u;
  if (active) {
    for (uint32_t c = c0; c < ncell; ++c) {
      ag[c] = 0.0;
      if (!a.unit_h) ah[c] = 0.0;
      ac[c] = 0u;
    }
  }
  const CodeT* codes = reinterpret_cast<const CodeT*>(a.codes);
  for (uint32_t p0 = 0; p0 < nd.count; p0 += U) {
    const uint32_t m = min(U, nd.count - p0);
    uint32_t r[U], code[U];
    double g[U], h[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t pos = nd.begin + p0 + (u < m ? u : 0u);
      r[u] = a.rows ? a.rows[pos] : pos;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      g[u] = a.gh[2 * size_t(r[u])];
      h[u] = a.gh[2 * size_t(r[u]) + 1];
      code[u] = active ? uint32_t(codes[size_t(r[u]) * a.row_stride + f]) : 0u;
    }
    if (!active) continue;
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (u >= m) break;
      const uint32_t cd = code[u];
      if (cd >= ncell) {
        raise_err(a.err, kErrCount);
        continue;
      }
      const bool miss = has_miss && cd == B;
      if (!a.missing_only) {
        const uint32_t c = miss ? 0u : cd;
        ag[c] += g[u];
        if (!a.unit_h) ah[c] += h[u];
        ac[c] += 1u;
      }
      if (miss) {
        ag[B] += g[u];
        if (!a.unit_h) ah[B] += h[u];
        ac[B] += 1u;
      }
    }
  }
  if constexpr (!GLOBAL) {
    if (!active) return;
    for (uint32_t c = c0; c < ncell; ++c) {
      a.out.g[obase + c] = __double_as_longlong(ag[c]);
      if (!a.unit_h) a.out.h[obase + c] = __double_as_longlong(ah[c]);
      a.out.cnt[obase + c] = ac[c];
    }
  }
}

void launch_hist_seq(const SeqLaunch& a, hipStream_t s) {
  if (a.n_nodes == 0 || a.n_feat == 0) return;
  const uint32_t groups = (a.n_feat + a.fg - 1) / a.fg;
  const dim3 grid(a.n_nodes * groups);
  const size_t lds = a.cpf ? size_t(a.fg) * a.cpf * (a.unit_h ? 12 : 20) : 0;
  if (a.code_bytes == 1) {
    if (a.cpf) hipLaunchKernelGGL((k_hist_seq<uint8_t, false>), grid, dim3(64), lds, s, a);
    else hipLaunchKernelGGL((k_hist_seq<uint8_t, true>), grid, dim3(64), 0, s, a);
  } else {
    if (a.cpf) hipLaunchKernelGGL((k_hist_seq<uint16_t, false>), grid, dim3(64), lds, s, a);
    else hipLaunchKernelGGL((k_hist_seq<uint16_t, true>), grid, dim3(64), 0, s, a);
  }
}

// SubtractHistograms (src/core/tree_histogram.cpp:16-66) for the row-order form: every cell
// of the sibling is parent - built in FP64, the Hessian through subtract_hessian. Missing
// cells are overwritten afterwards by the row-order missing pass.
__global__ __launch_bounds__(256) void k_sub_seq(const uint32_t* __restrict__ build_slot,
                                                 const int32_t* __restrict__ sibling_slot,
                                                 const uint32_t* __restrict__ parent_slot, LevelHist out,
                                                 LevelHist parent, uint32_t cells, uint32_t unit_h,
                                                 uint32_t* err) {
  const uint32_t b = blockIdx.y;
  const int32_t sib = sibling_slot[b];
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (sib < 0 || c >= cells) return;
  const size_t bo = size_t(build_slot[b]) * cells + c;
  const size_t po = size_t(parent_slot[b]) * cells + c;
  const size_t so = size_t(sib) * cells + c;
  const uint32_t pc = parent.cnt[po], cc = out.cnt[bo];
  if (cc > pc) raise_err(err, kErrSubtract);
  out.cnt[so] = pc - cc;
  const double pg = __longlong_as_double(parent.g[po]), cg = __longlong_as_double(out.g[bo]);
  const double g = pg - cg;
  if (!isfinite(pg) || !isfinite(cg) || !isfinite(g)) raise_err(err, kErrSubtract);
  out.g[so] = __double_as_longlong(g);
  if (!unit_h)
    out.h[so] = __double_as_longlong(subtract_hessian(__longlong_as_double(parent.h[po]),
                                                      __longlong_as_double(out.h[bo]), err));
}

void launch_sub_seq(const uint32_t* build_slot, const int32_t* sibling_slot, const uint32_t* parent_slot,
                    uint32_t n_build, LevelHist out, LevelHist parent, uint32_t cells, uint32_t unit_h,
                    uint32_t* err, hipStream_t s) {
  if (n_build == 0) return;
  hipLaunchKernelGGL(k_sub_seq, dim3((cells + 255) / 256, n_build), dim3(256), 0, s, build_slot, sibling_slot,
                     parent_slot, out, parent, cells, unit_h, err);
}

// SumRows (src/core/tree_split.cpp:119-140): one lane adds the node's pairs in row order.
__global__ __launch_bounds__(64) void k_sum_rows(const double* __restrict__ gh, const uint32_t* __restrict__ rows,
                                                 uint32_t begin, uint32_t n, NodeStats* out) {
  if (threadIdx.x != 0) return;
  constexpr uint32_t U = 8;
  double g = 0.0, h = 0.0;
  for (uint32_t p0 = 0; p0 < n; p0 += U) {
    const uint32_t m = min(U, n - p0);
    double xg[U], xh[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t pos = begin + p0 + (u < m ? u : 0u);
      const uint32_t r = rows ? rows[pos] : pos;
      xg[u] = gh[2 * size_t(r)];
      xh[u] = gh[2 * size_t(r) + 1];
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (u >= m) break;
      g += xg[u];
      h += xh[u];
    }
  }
  out->count = n;
  out->g = g;
  out->h = h;
  out->gq = 0;
  out->hq = 0;
}

void launch_sum_rows(const double* gh, const uint32_t* rows, uint32_t begin, uint32_t n, NodeStats* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, s, gh, rows, begin, n, out);
}

// ========================================================================= scan ======
// FindBestSplit (src/core/tree_split.cpp:142-234) for every active node of a level: one
// block per node, one thread per feature running the reference's sequential FP64 prefix
// (so empty bins give bit-identical prefixes and exact ties, as on the CPU), then a block
// argmax under the reference's order: larger gain, then smaller feature, then smaller
// threshold, then missing-left first (src/core/tree_split.cpp:31-45,225-230).

__device__ inline double soft_threshold(double g, double alpha) {  // objective.cpp:75-86
  const double m = fabs(g) - alpha;
  return m <= 0.0 ? 0.0 : copysign(m, g);
}
__device__ inline double node_score(double g, double h, double lambda, double alpha, uint32_t* err) {
  const double s = soft_threshold(g, alpha);
  const double d = h + lambda;
  const double r = s * s / d;
  if (!isfinite(r) || !(d > 0.0)) raise_err(err, kErrOverflow);
  return r;
}
__device__ inline double leaf_weight(double g, double h, double lambda, double alpha, double max_delta) {
  const double s = soft_threshold(g, alpha);
  double w = -s / (h + lambda);
  if (max_delta != 0.0) w = fmax(-max_delta, fmin(max_delta, w));
  return w;
}
__device__ inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

// StableRandomThreshold, src/core/tree_split.cpp:14-29 (same integer hash).
__device__ inline uint32_t random_threshold(uint32_t seed, uint32_t f, uint32_t node, uint32_t depth,
                                            uint32_t n) {
  uint64_t v = seed;
  v ^= uint64_t(f) + 0x9E3779B97F4A7C15ull;
  v *= 0xBF58476D1CE4E5B9ull;
  v ^= (uint64_t(node) << 32) | depth;
  v *= 0x94D049BB133111EBull;
  v ^= v >> 31;
  return uint32_t(v % n);
}

// The integer forms' value scales are computed on the device with the gradients.
__device__ inline ScanParams with_scales(ScanParams p) {
  if (p.scales) {
    p.scale_g = p.scales[2];
    p.scale_h = p.scales[3];
  }
  return p;
}

struct Acc3 {
  uint32_t count;  // rows of a node: below 2^32 by the dataset contract
  double g, h;
  long long gq, hq;
};

__device__ inline Acc3 cell_value(const LevelHist& H, size_t at, const ScanParams& p) {
  Acc3 v;
  v.count = H.cnt[at];
  if (p.integer) {
    v.gq = static_cast<long long>(H.g[at]);
    v.hq = p.unit_h ? 0 : static_cast<long long>(H.h[at]);
    v.g = 0.0;
    v.h = 0.0;
  } else {
    v.g = __longlong_as_double(H.g[at]);
    v.h = p.unit_h ? 0.0 : __longlong_as_double(H.h[at]);
    v.gq = v.hq = 0;
  }
  return v;
}

// FP64 value of an accumulated statistic.
__device__ inline double gval(const Acc3& a, const ScanParams& p) {
  return p.integer ? double(a.gq) * p.scale_g : a.g;
}
__device__ inline double hval(const Acc3& a, const ScanParams& p) {
  if (p.unit_h) return double(a.count);
  return p.integer ? double(a.hq) * p.scale_h : a.h;
}
__device__ inline Acc3 add3(const Acc3& a, const Acc3& b, const ScanParams& p) {
  Acc3 r;
  r.count = a.count + b.count;
  if (p.integer) {
    r.gq = a.gq + b.gq;
    r.hq = a.hq + b.hq;
    r.g = r.h = 0.0;
  } else {
    r.g = a.g + b.g;
    r.h = a.h + b.h;
    r.gq = r.hq = 0;
  }
  return r;
}
__device__ inline Acc3 sub3(const Acc3& a, const Acc3& b, const ScanParams& p) {
  Acc3 r;
  r.count = a.count - b.count;
  if (p.integer) {
    r.gq = a.gq - b.gq;
    r.hq = a.hq - b.hq;
    r.g = r.h = 0.0;
  } else {
    r.g = a.g - b.g;
    r.h = a.h - b.h;
    r.gq = r.hq = 0;
  }
  return r;
}
__device__ inline NodeStats to_stats(const Acc3& a, const ScanParams& p) {
  NodeStats s;
  s.count = a.count;
  s.g = gval(a, p);
  s.h = hval(a, p);
  s.gq = a.gq;
  s.hq = a.hq;
  return s;
}

struct Best {
  bool valid;
  double gain;
  uint32_t f, t, dl;
  Acc3 l, r;
  double l_lo, l_hi, r_lo, r_hi;
  double bound;  // rounding-noise bound of `gain` (noise_k > 0)
  double rival;  // largest gain + bound of the other candidates seen
};

// Rounding-noise bound of a split gain 0.5 (S(G_L,H_L) + S(G_R,H_R) - S(G_P,H_P)),
// S = G^2 / (H + lambda), when each side's sums carry stochastic-rounding errors of at most
// e = k * quantum * sqrt(rows) (independent rows, per-row variance at most quantum^2 / 4):
// 0.5 |dS| <= (|G| e_G + e_G^2 / 2) / (H + lambda) + G^2 e_H / (2 (H + lambda)^2) per side
// (the L1 soft threshold only shrinks |G|). The sides are disjoint row sets, the parent their
// union.
// One side's term of that bound. A bound, not a value: single precision with hardware square
// root and reciprocal (about one ulp each); the sum is widened by 2^-10 relative to cover their
// rounding.
__device__ inline float noise_side(double g, double h, uint32_t n, const ScanParams& p) {
  const float kg = float(p.noise_k * p.scale_g), kh = p.unit_h ? 0.0f : float(p.noise_k * p.scale_h);
  const float rn = __builtin_amdgcn_sqrtf(float(n)), eg = kg * rn, eh = kh * rn;
  const float rd = __builtin_amdgcn_rcpf(float(h) + float(p.lambda)), ag = fabsf(float(g)), w = ag * rd;
  return (ag * eg + 0.5f * eg * eg) * rd + 0.5f * w * w * eh;
}
// The parent's side is the node's own, the same for every candidate: computed once per scan.
__device__ inline double gain_noise(double gl, double hl, uint32_t nl, double gr, double hr, uint32_t nr,
                                    float parent_side, const ScanParams& p) {
  const float b = noise_side(gl, hl, nl, p) + noise_side(gr, hr, nr, p) + parent_side;
  return double(b) * (1.0 + 0x1p-10);
}

// What a node's scan computes once for all its candidates: in the stochastic-rounding forms
// (whose gains carry a noise bound far above FP64 rounding) the parent's score from the node's
// own totals instead of from each candidate's left + right (the reference's order, which every
// exact form keeps), and the parent's side of the noise bound.
struct NodeTerms {
  bool hoisted;
  double sp;
  float noise_parent;
};

// a better than b (incumbent), src/core/tree_split.cpp:31-45 plus missing-left first.
__device__ inline bool better(bool av, double ag, uint32_t af, uint32_t at, uint32_t adl, bool bv,
                              double bg, uint32_t bf, uint32_t bt, uint32_t bdl) {
  if (!av) return false;
  if (!bv) return true;
  if (ag != bg) return ag > bg;
  if (af != bf) return af < bf;
  if (at != bt) return at < bt;
  return adl > bdl;
}

// ConsiderMissingDirection, src/core/tree_split.cpp:65-115
__device__ inline void consider(const Acc3& left, const Acc3& right, const Acc3& miss, bool dl,
                                uint32_t f, uint32_t t, const ScanNode& nd, const ScanParams& p,
                                const NodeTerms& nt, const int8_t* mono, Best* best, uint32_t* err) {
  const Acc3 cl = dl ? add3(left, miss, p) : left;
  const Acc3 cr = dl ? right : add3(right, miss, p);
  const double gl = gval(cl, p), hl = hval(cl, p), gr = gval(cr, p), hr = hval(cr, p);
  if (cl.count < p.min_samples_leaf || cr.count < p.min_samples_leaf || hl < p.min_child_weight ||
      hr < p.min_child_weight || hl <= 0.0 || hr <= 0.0)
    return;
  // SplitGain, src/core/objective.cpp:320-356: parent = left + right, both in FP64.
  const double sl = node_score(gl, hl, p.lambda, p.alpha, err);
  const double sr = node_score(gr, hr, p.lambda, p.alpha, err);
  double sp = nt.sp;
  if (!nt.hoisted) {
    const double gp = gl + gr, hp = hl + hr;
    if (!isfinite(gp) || !isfinite(hp)) raise_err(err, kErrOverflow);
    sp = node_score(gp, hp, p.lambda, p.alpha, err);
  }
  const double gain = 0.5 * (sl + sr - sp) - p.gamma;
  if (!isfinite(gain)) raise_err(err, kErrOverflow);
  const double bound = p.noise_k > 0.0 ? gain_noise(gl, hl, cl.count, gr, hr, cr.count, nt.noise_parent, p) : 0.0;
  if (!(gain > 0.0) || gain < p.min_gain) {
    // Rejected here, but its exact gain might pass: it rivals "no split".
    if (p.noise_k > 0.0 && isfinite(gain)) best->rival = fmax(best->rival, gain + bound);
    return;
  }
  // MonotonicBoundsForSplit, src/core/tree_monotonic.cpp:12-58
  double l_lo = nd.lo, l_hi = nd.hi, r_lo = nd.lo, r_hi = nd.hi;
  const int c = p.has_mono ? int(mono[f]) : 0;
  if (c != 0) {
    const double lv = clampd(leaf_weight(gl, hl, p.lambda, p.alpha, p.max_delta), nd.lo, nd.hi);
    const double rv = clampd(leaf_weight(gr, hr, p.lambda, p.alpha, p.max_delta), nd.lo, nd.hi);
    if (c > 0) {
      if (lv > rv) return;
      l_hi = fmin(l_hi, rv);
      r_lo = fmax(r_lo, lv);
    } else {
      if (lv < rv) return;
      l_lo = fmax(l_lo, rv);
      r_hi = fmin(r_hi, lv);
    }
  }
  if (better(true, gain, f, t, dl ? 1u : 0u, best->valid, best->gain, best->f, best->t, best->dl)) {
    if (best->valid) best->rival = fmax(best->rival, best->gain + best->bound);
    best->bound = bound;
    best->valid = true;
    best->gain = gain;
    best->f = f;
    best->t = t;
    best->dl = dl ? 1u : 0u;
    best->l = cl;
    best->r = cr;
    best->l_lo = l_lo;
    best->l_hi = l_hi;
    best->r_lo = r_lo;
    best->r_hi = r_hi;
  } else {
    best->rival = fmax(best->rival, gain + bound);
  }
}

__device__ inline void write_candidate(Candidate* o, const Best& b, const Acc3& parent,
                                       const ScanParams& p) {
  o->valid = b.valid ? 1u : 0u;
  o->feature = b.f;
  o->threshold = b.t;
  o->default_left = b.dl;
  o->gain = b.gain;
  o->left = to_stats(b.l, p);
  o->right = to_stats(b.r, p);
  o->parent = to_stats(parent, p);
  o->l_lo = b.l_lo;
  o->l_hi = b.l_hi;
  o->r_lo = b.r_lo;
  o->r_hi = b.r_hi;
  o->bound = b.bound;
  o->rival = b.rival;
  o->ambiguous = 0u;
  o->pad2 = 0u;
}

__global__ __launch_bounds__(256) void k_scan(LevelHist H, uint32_t cells,
                                              const uint32_t* __restrict__ cell_offset,
                                              const uint32_t* __restrict__ bin_count,
                                              const uint32_t* __restrict__ miss_cell, uint32_t F,
                                              const ScanNode* __restrict__ nodes, ScanParams pin,
                                              const int8_t* __restrict__ mono,
                                              const uint32_t* __restrict__ masks,
                                              Candidate* __restrict__ out,
                                              Candidate* __restrict__ per_feature,
                                              const NodeStats* __restrict__ parent_stats, uint32_t* err) {
  __shared__ double s_gain[256];
  __shared__ uint32_t s_key[256][4];  // valid, f, t, dl
  __shared__ uint32_t s_who[256];
  const ScanParams p = with_scales(pin);
  const ScanNode nd = nodes[blockIdx.x];
  const size_t base = size_t(nd.slot) * cells;

  // Parent statistics: given by the parent's split, or (root) the histogram totals of
  // feature 0 in cell order.
  Acc3 P;
  if (nd.parent_from_hist == 2) {
    const NodeStats& q = parent_stats[blockIdx.x];
    P = Acc3{uint32_t(q.count), q.g, q.h, q.gq, q.hq};
  } else if (nd.parent_from_hist) {
    P = Acc3{0, 0.0, 0.0, 0, 0};
    const uint32_t c1 = p.ref_view ? cell_offset[0] + bin_count[0] : cell_offset[1];
    for (uint32_t c = cell_offset[0]; c < c1; ++c) P = add3(P, cell_value(H, base + c, p), p);
  } else {
    P.count = uint32_t(nd.parent.count);
    P.g = nd.parent.g;
    P.h = nd.parent.h;
    P.gq = nd.parent.gq;
    P.hq = nd.parent.hq;
  }

  Best best;
  best.valid = false;
  best.gain = 0.0;
  best.f = best.t = best.dl = 0;
  best.bound = 0.0;
  best.rival = -INFINITY;
  for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
    Best fb;
    fb.valid = false;
    fb.gain = 0.0;
    fb.f = f;
    fb.t = fb.dl = 0;
    fb.bound = 0.0;
    fb.rival = -INFINITY;
    const bool allowed = !p.has_mask || ((masks[nd.mask_word0 + f / 32] >> (f % 32)) & 1u);
    const uint32_t B = bin_count[f], off = cell_offset[f], mc = miss_cell[f];
    if (allowed) {
      // Counts of every feature must add up to the node count (tree_split.cpp:167-181); in
      // the reference view the missing cell is a separate statistic, already inside bin 0.
      uint64_t total = 0;
      const uint32_t c1 = p.ref_view ? off + B : cell_offset[f + 1];
      for (uint32_t c = off; c < c1; ++c) total += H.cnt[base + c];
      if (total != P.count) raise_err(err, kErrCount);
    }
    if (allowed && B >= 2) {
      const Acc3 m = mc == 0xFFFFFFFFu ? Acc3{0, 0.0, 0.0, 0, 0} : cell_value(H, base + mc, p);
      if (m.count > P.count || hval(m, p) > hval(P, p)) raise_err(err, kErrMissing);
      const Acc3 nonmiss = sub3(P, m, p);
      Acc3 left{0, 0.0, 0.0, 0, 0};
      // The reference's order for the parent score (left + right per candidate).
      const NodeTerms nt{false, 0.0, p.noise_k > 0.0 ? noise_side(gval(P, p), hval(P, p), P.count, p) : 0.0f};
      const uint32_t rt = p.random_threshold ? random_threshold(p.seed, f, nd.node_index, nd.depth, B - 1) : 0u;
      for (uint32_t t = 0; t + 1 < B; ++t) {
        const Acc3 cv = cell_value(H, base + off + t, p);
        // Reference view: the prefix starts from bin 0 with the missing rows taken out.
        left = add3(left, (p.ref_view && t == 0) ? sub3(cv, m, p) : cv, p);
        if (!p.integer && (!isfinite(left.g) || !isfinite(left.h))) raise_err(err, kErrOverflow);
        const Acc3 right = sub3(nonmiss, left, p);
        if (left.count > nonmiss.count) raise_err(err, kErrPrefix);
        if (left.count == 0 && right.count == 0) continue;
        if (p.random_threshold && t != rt) continue;
        consider(left, right, m, true, f, t, nd, p, nt, mono, &fb, err);
        consider(left, right, m, false, f, t, nd, p, nt, mono, &fb, err);
      }
    }
    if (per_feature) write_candidate(&per_feature[size_t(blockIdx.x) * F + f], fb, P, p);
    if (better(fb.valid, fb.gain, fb.f, fb.t, fb.dl, best.valid, best.gain, best.f, best.t, best.dl)) best = fb;
  }

  s_gain[threadIdx.x] = best.gain;
  s_key[threadIdx.x][0] = best.valid ? 1u : 0u;
  s_key[threadIdx.x][1] = best.f;
  s_key[threadIdx.x][2] = best.t;
  s_key[threadIdx.x][3] = best.dl;
  s_who[threadIdx.x] = threadIdx.x;
  __syncthreads();
  for (uint32_t stride = blockDim.x / 2; stride > 0; stride /= 2) {
    if (threadIdx.x < stride) {
      const uint32_t a = threadIdx.x, b = threadIdx.x + stride;
      if (better(s_key[b][0], s_gain[b], s_key[b][1], s_key[b][2], s_key[b][3], s_key[a][0], s_gain[a],
                 s_key[a][1], s_key[a][2], s_key[a][3])) {
        s_gain[a] = s_gain[b];
        for (int k = 0; k < 4; ++k) s_key[a][k] = s_key[b][k];
        s_who[a] = s_who[b];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == s_who[0]) write_candidate(&out[blockIdx.x], best, P, p);
}

// ---- wave-parallel scan: one 64-lane wavefront per (node, feature) -----------------------
// Lane l owns thresholds [chunk + 4l, chunk + 4l + 4): a lane-local prefix of its four bins,
// an exclusive wave scan of lane totals, then every lane evaluates its thresholds in both
// missing directions. Integer forms are exact in any order; the FP32 form's prefixes are
// FP64 sums in a different association than the sequential reference scan (that form is not
// bit-exact anyway). The reference-exact form uses k_scan above.
__device__ inline Acc3 shfl_up3(const Acc3& v, int d) {
  Acc3 r;
  r.count = __shfl_up(v.count, d, 64);
  r.g = __shfl_up(v.g, d, 64);
  r.h = __shfl_up(v.h, d, 64);
  r.gq = __shfl_up(v.gq, d, 64);
  r.hq = __shfl_up(v.hq, d, 64);
  return r;
}

__device__ inline Acc3 wave_sum3(Acc3 v, const ScanParams& p) {
  for (int d = 1; d < 64; d <<= 1) {
    Acc3 o;
    o.count = __shfl_xor(v.count, d, 64);
    o.g = __shfl_xor(v.g, d, 64);
    o.h = __shfl_xor(v.h, d, 64);
    o.gq = __shfl_xor(v.gq, d, 64);
    o.hq = __shfl_xor(v.hq, d, 64);
    v = add3(v, o, p);
  }
  return v;
}

constexpr int kScanK = 4;

// INT / UH fix the form at compile time (integer sums; unit Hessian), so the accumulators
// carry only the fields that form uses (the register count sets how many waves run at once).
// One wavefront scans feature f of node nd and writes the feature's candidate to *o (lane 0):
// the wave-parallel prefix over the feature's bins, the missing direction, the wave argmax, and
// (with a noise bound) the feature's rival gain. parent_q: the node's totals when
// nd.parent_from_hist == 2.
template <bool INT, bool UH>
__device__ inline void scan_feature_wave(const LevelHist& H, uint32_t cells, const uint32_t* __restrict__ cell_offset,
                                         const uint32_t* __restrict__ bin_count,
                                         const uint32_t* __restrict__ miss_cell, uint32_t f, const ScanNode& nd,
                                         const NodeStats* parent_q, const ScanParams& p,
                                         const int8_t* __restrict__ mono, bool allowed, Candidate* o, uint32_t* err) {
  const uint32_t lane = threadIdx.x & 63u;
  const size_t base = size_t(nd.slot) * cells;
  const Acc3 zero{0, 0.0, 0.0, 0, 0};

  Acc3 P;
  if (nd.parent_from_hist == 2) {
    const NodeStats& q = *parent_q;
    P = Acc3{uint32_t(q.count), q.g, q.h, q.gq, q.hq};
  } else if (nd.parent_from_hist) {
    // Histogram totals of feature 0, reduced the same way by every wavefront.
    Acc3 part = zero;
    for (uint32_t c = cell_offset[0] + lane; c < cell_offset[1]; c += 64) part = add3(part, cell_value(H, base + c, p), p);
    P = wave_sum3(part, p);
  } else {
    P.count = uint32_t(nd.parent.count);
    P.g = nd.parent.g;
    P.h = nd.parent.h;
    P.gq = nd.parent.gq;
    P.hq = nd.parent.hq;
  }
  Best best;
  best.valid = false;
  best.gain = 0.0;
  best.f = f;
  best.t = best.dl = 0;
  best.bound = 0.0;
  best.rival = -INFINITY;
  const uint32_t B = bin_count[f], off = cell_offset[f], mc = miss_cell[f];
  if (allowed) {
    // Every cell of the feature is read once: the bins in the prefix loop (the last bin only
    // for the count check) and the missing cell beside them.
    const Acc3 m = mc == 0xFFFFFFFFu ? zero : cell_value(H, base + mc, p);
    if (B >= 2 && lane == 0 && (m.count > P.count || hval(m, p) > hval(P, p))) raise_err(err, kErrMissing);
    const Acc3 nonmiss = sub3(P, m, p);
    NodeTerms nt{false, 0.0, 0.0f};
    if (p.noise_k > 0.0) {
      const double gp = gval(P, p), hp = hval(P, p);
      if (!isfinite(gp) || !isfinite(hp)) raise_err(err, kErrOverflow);
      nt = NodeTerms{true, node_score(gp, hp, p.lambda, p.alpha, err), noise_side(gp, hp, P.count, p)};
    }
    const uint32_t rt = p.random_threshold && B >= 2 ? random_threshold(p.seed, f, nd.node_index, nd.depth, B - 1) : 0u;
    Acc3 carry = zero;
    for (uint32_t c0 = 0; c0 < B; c0 += 64 * kScanK) {
      Acc3 v[kScanK];
      Acc3 run = zero;
#pragma unroll
      for (int k = 0; k < kScanK; ++k) {
        const uint32_t t = c0 + lane * kScanK + k;
        v[k] = t < B ? cell_value(H, base + off + t, p) : zero;
        run = add3(run, v[k], p);
        v[k] = run;  // lane-local inclusive prefix
      }
      // Exclusive scan of lane totals.
      Acc3 incl = run;
      for (int d = 1; d < 64; d <<= 1) {
        const Acc3 o = shfl_up3(incl, d);
        if (lane >= uint32_t(d)) incl = add3(o, incl, p);
      }
      const Acc3 up = shfl_up3(incl, 1);
      const Acc3 excl = lane == 0 ? zero : up;
      const Acc3 chunk_total = [&] {
        Acc3 t;
        t.count = __shfl(incl.count, 63, 64);
        t.g = __shfl(incl.g, 63, 64);
        t.h = __shfl(incl.h, 63, 64);
        t.gq = __shfl(incl.gq, 63, 64);
        t.hq = __shfl(incl.hq, 63, 64);
        return t;
      }();
#pragma unroll
      for (int k = 0; k < kScanK; ++k) {
        const uint32_t t = c0 + lane * kScanK + k;
        if (t + 1 >= B) break;  // a split needs a bin on its right
        const Acc3 left = add3(add3(carry, excl, p), v[k], p);
        const Acc3 right = sub3(nonmiss, left, p);
        if (left.count > nonmiss.count) raise_err(err, kErrPrefix);
        if (left.count == 0 && right.count == 0) continue;
        if (p.random_threshold && t != rt) continue;
        consider(left, right, m, true, f, t, nd, p, nt, mono, &best, err);
        // Without a missing cell both directions are the same split with the same gain, and
        // the tie rule keeps missing-left: the second evaluation cannot win.
        if (mc != 0xFFFFFFFFu) consider(left, right, m, false, f, t, nd, p, nt, mono, &best, err);
      }
      carry = add3(carry, chunk_total, p);
    }
    // Every row of the node lands in one of the feature's cells.
    if (lane == 0 && uint64_t(carry.count) + m.count != P.count) raise_err(err, kErrCount);
  }
  // Wave argmax: larger gain, then smaller threshold, then missing-left first.
  double bg = best.gain;
  uint32_t bv = best.valid ? 1u : 0u, bt = best.t, bd = best.dl;
  for (int d = 1; d < 64; d <<= 1) {
    const double og = __shfl_xor(bg, d, 64);
    const uint32_t ov = __shfl_xor(bv, d, 64), ot = __shfl_xor(bt, d, 64), od = __shfl_xor(bd, d, 64);
    if (better(ov, og, f, ot, od, bv, bg, f, bt, bd)) {
      bg = og;
      bv = ov;
      bt = ot;
      bd = od;
    }
  }
  // The feature's rival: every other lane's best and every lane's other candidates.
  const bool winner = bv && best.valid && best.t == bt && best.dl == bd;
  double rival = winner ? best.rival : fmax(best.rival, best.valid ? best.gain + best.bound : -INFINITY);
  for (int d = 1; d < 64; d <<= 1) rival = fmax(rival, __shfl_xor(rival, d, 64));
  if (!bv) {
    if (lane == 0) {
      Best none;
      none.valid = false;
      none.gain = 0.0;
      none.f = f;
      none.t = none.dl = 0;
      none.l = none.r = zero;
      none.l_lo = none.l_hi = none.r_lo = none.r_hi = 0.0;
      none.bound = 0.0;
      none.rival = rival;
      write_candidate(o, none, P, p);
    }
  } else if (winner) {
    best.rival = rival;
    write_candidate(o, best, P, p);
  }
}

template <bool INT, bool UH>
__global__ __launch_bounds__(256) void k_scan_feature(LevelHist H, uint32_t cells,
                                                      const uint32_t* __restrict__ cell_offset,
                                                      const uint32_t* __restrict__ bin_count,
                                                      const uint32_t* __restrict__ miss_cell, uint32_t F,
                                                      const ScanNode* __restrict__ nodes, ScanParams pin,
                                                      const int8_t* __restrict__ mono,
                                                      const uint32_t* __restrict__ masks,
                                                      Candidate* __restrict__ per_feature,
                                                      const NodeStats* __restrict__ parent_stats, uint32_t* err,
                                                      const uint32_t* __restrict__ dyn_nodes) {
  ScanParams p = with_scales(pin);
  p.integer = INT ? 1u : 0u;
  p.unit_h = UH ? 1u : 0u;
  const uint32_t groups = (F + 3) / 4;
  const uint32_t node_i = blockIdx.x / groups;
  const uint32_t f = (blockIdx.x % groups) * 4 + (threadIdx.x >> 6);
  if (f >= F) return;  // whole wavefront leaves together
  if (dyn_nodes && node_i >= *dyn_nodes) return;
  const ScanNode nd = nodes[node_i];
  const bool allowed = !p.has_mask || ((masks[nd.mask_word0 + f / 32] >> (f % 32)) & 1u);
  scan_feature_wave<INT, UH>(H, cells, cell_offset, bin_count, miss_cell, f, nd,
                             parent_stats ? parent_stats + node_i : nullptr, p, mono, allowed,
                             &per_feature[size_t(node_i) * F + f], err);
}

// Per node: best of the per-feature candidates under the reference order.
__global__ __launch_bounds__(256) void k_scan_reduce(const Candidate* __restrict__ per_feature, uint32_t F,
                                                     Candidate* __restrict__ out,
                                                     const uint32_t* __restrict__ dyn_nodes, double accept,
                                                     int noise, uint32_t* __restrict__ contend) {
  __shared__ double s_gain[256];
  __shared__ uint32_t s_key[256][4];
  __shared__ uint32_t s_idx[256];
  __shared__ double s_rival[256];
  __shared__ uint32_t s_mask[32];  // F <= 1024 features
  // Blocks loop over nodes (with the count on the device the grid may be smaller).
  const uint32_t nn = dyn_nodes ? *dyn_nodes : gridDim.x;
  for (uint32_t node = blockIdx.x; node < nn; node += gridDim.x) {
    const Candidate* c = per_feature + size_t(node) * F;
    uint32_t bi = 0xFFFFFFFFu;
    double bg = 0.0;
    uint32_t bv = 0, bf = 0, bt = 0, bd = 0;
    for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
      const Candidate& x = c[f];
      if (better(x.valid, x.gain, x.feature, x.threshold, x.default_left, bv, bg, bf, bt, bd)) {
        bv = 1;
        bg = x.gain;
        bf = x.feature;
        bt = x.threshold;
        bd = x.default_left;
        bi = f;
      }
    }
    s_gain[threadIdx.x] = bg;
    s_key[threadIdx.x][0] = bv;
    s_key[threadIdx.x][1] = bf;
    s_key[threadIdx.x][2] = bt;
    s_key[threadIdx.x][3] = bd;
    s_idx[threadIdx.x] = bi;
    __syncthreads();
    for (uint32_t stride = blockDim.x / 2; stride > 0; stride /= 2) {
      if (threadIdx.x < stride) {
        const uint32_t x = threadIdx.x, y = threadIdx.x + stride;
        if (better(s_key[y][0], s_gain[y], s_key[y][1], s_key[y][2], s_key[y][3], s_key[x][0], s_gain[x],
                   s_key[x][1], s_key[x][2], s_key[x][3])) {
          s_gain[x] = s_gain[y];
          for (int k = 0; k < 4; ++k) s_key[x][k] = s_key[y][k];
          s_idx[x] = s_idx[y];
        }
      }
      __syncthreads();
    }
    if (noise) {
      // Ambiguous: some other candidate's gain + bound reaches the best's gain - bound, or
      // the best's gain - bound reaches the acceptance threshold of "no split". Contenders:
      // the features with a candidate that reaches it (the best's feature always).
      const uint32_t win = s_idx[0];
      const bool wvalid = win != 0xFFFFFFFFu && c[win].valid;
      const double low = wvalid ? c[win].gain - c[win].bound : accept;
      const uint32_t W = (F + 31) / 32;
      const bool track = contend && W <= 32;
      if (track)
        for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) s_mask[w] = 0u;
      __syncthreads();
      double r = -INFINITY;
      for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
        const Candidate& x = c[f];
        const double own = x.valid ? x.gain + x.bound : -INFINITY;
        r = fmax(r, f == win ? x.rival : fmax(x.rival, own));
        if (track && ((wvalid && f == win) || fmax(x.rival, own) >= low)) atomicOr(&s_mask[f / 32], 1u << (f % 32));
      }
      s_rival[threadIdx.x] = r;
      __syncthreads();
      if (track)
        for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) contend[size_t(node) * W + w] = s_mask[w];
      for (uint32_t stride = blockDim.x / 2; stride > 0; stride /= 2) {
        if (threadIdx.x < stride) s_rival[threadIdx.x] = fmax(s_rival[threadIdx.x], s_rival[threadIdx.x + stride]);
        __syncthreads();
      }
    }
    if (threadIdx.x == 0) {
      Candidate b = c[s_idx[0] == 0xFFFFFFFFu ? 0 : s_idx[0]];
      if (noise) {
        const bool valid = s_idx[0] != 0xFFFFFFFFu && b.valid;
        const double low = valid ? b.gain - b.bound : accept;
        b.rival = s_rival[0];
        b.ambiguous = (s_rival[0] >= low || (valid && low <= accept)) ? 1u : 0u;
      }
      out[node] = b;
    }
    __syncthreads();  // the next node reuses the LDS
  }
}

// End of a level: the level's candidates and error word go to host-mapped (coherent)
// memory, then the sequence number is published after a system-scope fence. The host waits
// on that number instead of a stream synchronisation plus two copies.
__global__ __launch_bounds__(256) void k_signal(const Candidate* __restrict__ cand, uint32_t n,
                                                const uint32_t* __restrict__ err, Candidate* host_cand,
                                                uint32_t* host_flags, unsigned long long* host_seq,
                                                unsigned long long seq) {
  const uint32_t words = n * uint32_t(sizeof(Candidate) / 8);
  const unsigned long long* src = reinterpret_cast<const unsigned long long*>(cand);
  unsigned long long* dst = reinterpret_cast<unsigned long long*>(host_cand);
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
  if (threadIdx.x == 0) {
    host_flags[0] = err ? err[0] : 0u;
    host_flags[1] = err ? err[1] : 0u;
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(host_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

void launch_signal(const Candidate* cand, uint32_t n, const uint32_t* err, Candidate* host_cand,
                   uint32_t* host_flags, unsigned long long* host_seq, unsigned long long seq, hipStream_t s) {
  hipLaunchKernelGGL(k_signal, dim3(1), dim3(256), 0, s, cand, n, err, host_cand, host_flags, host_seq, seq);
}

void launch_scan(const ScanLaunch& a, hipStream_t s) {
  if (a.n_nodes == 0) return;
  if (a.sequential) {
    hipLaunchKernelGGL(k_scan, dim3(a.n_nodes), dim3(256), 0, s, a.hist, a.cells, a.cell_offset,
                       a.bin_count, a.miss_cell, a.features, a.nodes, a.p, a.mono, a.masks, a.out,
                       a.per_feature, a.parent_stats, a.err);
    return;
  }
  const uint32_t groups = (a.features + 3) / 4;
#define MPSB_SCAN(I, U)                                                                                   \
  hipLaunchKernelGGL((k_scan_feature<I, U>), dim3(a.n_nodes * groups), dim3(256), 0, s, a.hist, a.cells,  \
                     a.cell_offset, a.bin_count, a.miss_cell, a.features, a.nodes, a.p, a.mono, a.masks,   \
                     a.per_feature, a.parent_stats, a.err, a.dyn_nodes)
  if (a.p.integer) {
    if (a.p.unit_h) {
      MPSB_SCAN(true, true);
    } else {
      MPSB_SCAN(true, false);
    }
  } else if (a.p.unit_h) {
    MPSB_SCAN(false, true);
  } else {
    MPSB_SCAN(false, false);
  }
#undef MPSB_SCAN
  // With the count on the device, n_nodes is a bound: at most 64 blocks, looping over nodes.
  const uint32_t rows = a.dyn_nodes ? std::min(a.n_nodes, 64u) : a.n_nodes;
  hipLaunchKernelGGL(k_scan_reduce, dim3(rows), dim3(256), 0, s, a.per_feature, a.features, a.out, a.dyn_nodes,
                     fmax(0.0, a.p.min_gain), a.p.noise_k > 0.0 ? 1 : 0, a.contend);
}

// ==================================================================== partition ======
// Stable partition of each split node's row segment (src/core/tree_split.cpp:273-289):
// pass 1 marks go-left bits and counts per 4096-row chunk (leaf segments instead add
// lr * leaf to the prediction of their rows, src/core/trainer.cpp:256-262), pass 2 scatters
// rows in order, each chunk adding up the counts of its node's earlier chunks (a level whose
// rows do not move only checks the counts per node).
__device__ inline uint32_t load_code(const uint8_t* codes, uint32_t row, uint32_t stride,
                                     uint32_t code_bytes, uint32_t f) {
  const size_t at = (size_t(row) * stride + f) * code_bytes;
  return code_bytes == 1 ? uint32_t(codes[at]) : uint32_t(*reinterpret_cast<const uint16_t*>(codes + at));
}

// Column-major copy of the codes ([feature][row]) for the partition of dense nodes: their
// rows, ascending, read a feature column almost contiguously instead of one 128-byte row line
// per row. Built once per uploaded dataset.
// Column-major copy of the codes (read by the partition of dense nodes): a block stages 256
// whole rows in LDS with coalesced 4-byte loads (row pitch padded by one word), then writes
// each feature's 256 codes as one contiguous run.
__global__ __launch_bounds__(256) void k_transpose(const uint8_t* __restrict__ codes, uint32_t row_stride,
                                                   uint32_t code_bytes, uint32_t F, uint64_t N,
                                                   uint8_t* __restrict__ cm, uint32_t rows_per_block) {
  extern __shared__ uint32_t stile[];
  const uint32_t row_words = row_stride * code_bytes / 4;
  const uint32_t pitch = row_words + 1;
  const uint64_t r0 = uint64_t(blockIdx.x) * rows_per_block;
  if (r0 >= N) return;
  const uint32_t nrows = uint32_t(min<uint64_t>(rows_per_block, N - r0));
  const uint32_t* src = reinterpret_cast<const uint32_t*>(codes) + r0 * row_words;
  for (uint32_t i = threadIdx.x; i < nrows * row_words; i += blockDim.x)
    stile[(i / row_words) * pitch + i % row_words] = src[i];
  __syncthreads();
  if (threadIdx.x >= nrows) return;
  const uint8_t* rb = reinterpret_cast<const uint8_t*>(stile + size_t(threadIdx.x) * pitch);
  for (uint32_t f = 0; f < F; ++f) {
    const size_t o = size_t(f) * N + r0 + threadIdx.x;
    if (code_bytes == 1) {
      cm[o] = rb[f];
    } else {
      reinterpret_cast<uint16_t*>(cm)[o] = reinterpret_cast<const uint16_t*>(rb)[f];
    }
  }
}
// Rows wider than 64 KB: one thread per output code (strided reads).
__global__ void k_transpose_wide(const uint8_t* __restrict__ codes, uint32_t row_stride, uint32_t code_bytes,
                                 uint32_t F, uint64_t N, uint8_t* __restrict__ cm) {
  const uint64_t total = uint64_t(F) * N;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < total; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t f = i / N, r = i % N;
    const size_t src = (size_t(r) * row_stride + f) * code_bytes;
    if (code_bytes == 1) {
      cm[i] = codes[src];
    } else {
      reinterpret_cast<uint16_t*>(cm)[i] = *reinterpret_cast<const uint16_t*>(codes + src);
    }
  }
}

void launch_transpose(const uint8_t* codes, uint32_t row_stride, uint32_t code_bytes, uint32_t F, uint64_t N,
                      uint8_t* cm, hipStream_t s) {
  const size_t pitch = size_t(row_stride * code_bytes / 4 + 1) * 4;
  if (pitch > 65536) {
    hipLaunchKernelGGL(k_transpose_wide, dim3(8192), dim3(256), 0, s, codes, row_stride, code_bytes, F, N, cm);
    return;
  }
  const uint32_t rows_per_block = uint32_t(std::min<size_t>(256, 65536 / pitch));
  hipLaunchKernelGGL(k_transpose, dim3(uint32_t((N + rows_per_block - 1) / rows_per_block)), dim3(256),
                     size_t(rows_per_block) * pitch, s, codes, row_stride, code_bytes, F, N, cm, rows_per_block);
}

// Each thread owns positions begin + j * 256 + tid, j < 16: all sixteen row ids (and then
// all sixteen dependent values) are loaded before any is used, so a wave keeps 16 gathers
// in flight instead of one.
constexpr uint32_t kPartPer = kPartChunk / 256;

// Rows of a leaf segment take the leaf: its node id per row (deferred update: the next
// round's gradient pass adds lr * leaf while it streams the rows anyway) or lr * leaf added to
// the prediction. Rows of a node are distinct, so the read-modify-writes are independent.
// A row's leaf id in the launch's width.
__device__ inline void store_leaf_id(const PartLaunch& a, uint32_t row, uint32_t id) {
  if (a.leaf_bytes == 1) {
    reinterpret_cast<uint8_t*>(a.leaf_node_out)[row] = static_cast<uint8_t>(id);
  } else {
    a.leaf_node_out[row] = static_cast<uint16_t>(id);
  }
}

// First leaf-id slot of the node's rows: its class's block in a batched round (a node's
// segment lies inside one class's positions), else 0.
__device__ inline uint32_t leaf_slot0(const PartLaunch& a, const PartNode& nd) {
  return a.class_rows ? (nd.begin / a.class_rows) * a.class_rows : 0u;
}

__device__ inline void part_leaf_rows(const PartLaunch& a, const PartNode& nd, const uint32_t (&r)[kPartPer]) {
  if (a.leaf_node_out) {
    const uint32_t s0 = leaf_slot0(a, nd);
#pragma unroll
    for (uint32_t j = 0; j < kPartPer; ++j)
      if (r[j] != 0xFFFFFFFFu) store_leaf_id(a, s0 + r[j], nd.leaf_node);
    return;
  }
  const double add = a.lr * nd.leaf_value;
  double* __restrict__ pred = a.pred;
  double v[kPartPer];
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j)
    v[j] = r[j] != 0xFFFFFFFFu ? pred[size_t(r[j]) * a.pred_stride + a.pred_offset] : 0.0;
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j)
    if (r[j] != 0xFFFFFFFFu) pred[size_t(r[j]) * a.pred_stride + a.pred_offset] = v[j] + add;
}

// A split whose children are both leaves: each row takes its child's leaf the same way.
__device__ inline void part_split_leaf_rows(const PartLaunch& a, const PartNode& nd, const uint32_t (&r)[kPartPer],
                                            const bool (&left)[kPartPer]) {
  if (a.leaf_node_out) {
    const uint32_t s0 = leaf_slot0(a, nd);
#pragma unroll
    for (uint32_t j = 0; j < kPartPer; ++j)
      if (r[j] != 0xFFFFFFFFu) store_leaf_id(a, s0 + r[j], left[j] ? nd.leaf_node : nd.leaf_node_right);
    return;
  }
  const double add_l = a.lr * nd.leaf_value, add_r = a.lr * nd.leaf_right;
  double* __restrict__ pred = a.pred;
  double v[kPartPer];
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j)
    v[j] = r[j] != 0xFFFFFFFFu ? pred[size_t(r[j]) * a.pred_stride + a.pred_offset] : 0.0;
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j)
    if (r[j] != 0xFFFFFFFFu) pred[size_t(r[j]) * a.pred_stride + a.pred_offset] = v[j] + (left[j] ? add_l : add_r);
}

// Go-left decision of each row of a split chunk from its split feature's code.
__device__ inline void part_decide(const PartLaunch& a, const PartNode& nd, const uint32_t (&r)[kPartPer],
                                   bool (&left)[kPartPer]) {
  uint32_t code[kPartPer];
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j) {
    code[j] = 0;
    if (r[j] == 0xFFFFFFFFu) continue;
    if (nd.column_major) {
      const size_t at = size_t(nd.feature) * a.rows_total + r[j];
      code[j] = a.code_bytes == 1 ? uint32_t(a.codes_cm[at]) : uint32_t(reinterpret_cast<const uint16_t*>(a.codes_cm)[at]);
    } else {
      code[j] = load_code(a.codes, r[j], a.row_stride, a.code_bytes, nd.feature);
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j)
    left[j] = r[j] != 0xFFFFFFFFu && (code[j] == nd.missing_code ? nd.default_left != 0 : code[j] <= nd.threshold);
}

__global__ __launch_bounds__(256) void k_part_mark(PartLaunch a) {
  uint32_t ch = blockIdx.x;
  if (a.ctl) {
    if (a.ctl->xcd_map) {  // block b runs on XCD b mod 8: entry b / 8 of that XCD's list
      const uint32_t x = blockIdx.x & 7u, k = blockIdx.x >> 3;
      if (k >= a.ctl->xcd_len[x]) return;
      ch = a.block_chunk[a.ctl->xcd_off[x] + k];
    }
    if (ch >= a.ctl->n_chunks) return;
  }
  const PartNode nd = a.nodes[a.chunk_node[ch]];
  const uint32_t local0 = (ch - nd.chunk0) * kPartChunk;
  const uint32_t begin = nd.begin + local0;
  const uint32_t end = nd.begin + min(nd.count, local0 + kPartChunk);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t* __restrict__ rows_in = a.rows_in;
  uint32_t r[kPartPer];
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j) {
    const uint32_t p = begin + j * 256 + threadIdx.x;
    r[j] = p < end ? (rows_in ? rows_in[p] : p) : 0xFFFFFFFFu;  // null rows_in: identity
  }
  if (nd.action == kPartLeaf) {
    part_leaf_rows(a, nd, r);
    if (threadIdx.x == 0) a.chunk_left[ch] = 0;
    return;
  }
  bool left[kPartPer];
  part_decide(a, nd, r, left);
  __shared__ uint32_t s_cnt[4];
  uint32_t mine = 0;
  if (nd.action == kPartSplitLeaf) {
    // Both children are leaves: each row takes its child's leaf value here, and the left
    // count still goes to the per-node check.
    part_split_leaf_rows(a, nd, r, left);
#pragma unroll
    for (uint32_t j = 0; j < kPartPer; ++j) {
      const unsigned long long m = __ballot(left[j]);
      if (lane == 0) mine += __popcll(m);
    }
    if (lane == 0) s_cnt[wave] = mine;
    __syncthreads();
    if (threadIdx.x == 0) a.chunk_left[ch] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    return;
  }
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(a.bits) + size_t(ch) * (kPartChunk / 64);
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j) {
    const unsigned long long m = __ballot(left[j]);
    if (lane == 0) {
      bits[j * 4 + wave] = m;
      mine += __popcll(m);
    }
  }
  if (lane == 0) s_cnt[wave] = mine;
  __syncthreads();
  if (threadIdx.x == 0) a.chunk_left[ch] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

// Exclusive prefix of chunk counts within each node (chunk_left becomes chunk bases): one
// block per node, 256 chunks per step.
__global__ __launch_bounds__(256) void k_part_scan(const PartNode* nodes, uint32_t n_nodes, uint32_t* chunk_left,
                                                   uint32_t n_chunks, const LevelCtl* ctl, uint32_t* err) {
  __shared__ uint32_t s[256];
  if (ctl) {
    if (blockIdx.x >= ctl->n_part) return;
    n_chunks = ctl->n_chunks;
  }
  const PartNode nd = nodes[blockIdx.x];
  if (nd.action == kPartLeaf) return;
  const uint32_t chunks = (nd.count + kPartChunk - 1) / kPartChunk;
  uint32_t run = 0;
  for (uint32_t c0 = 0; c0 < chunks; c0 += 256) {
    const uint32_t c = nd.chunk0 + c0 + threadIdx.x;
    const bool in = c0 + threadIdx.x < chunks && c < n_chunks;
    const uint32_t v = in ? chunk_left[c] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t d = 1; d < 256; d <<= 1) {  // Hillis-Steele inclusive scan
      const uint32_t add = threadIdx.x >= d ? s[threadIdx.x - d] : 0u;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (in) chunk_left[c] = run + s[threadIdx.x] - v;
    run += s[255];
    __syncthreads();
  }
  if (threadIdx.x == 0 && run != nd.left_count) raise_err(err, kErrCount);
}

template <typename G>
__device__ inline void scatter_rows(const PartLaunch& a, const PartNode& nd, uint32_t begin, uint32_t end,
                                    uint32_t local0, const unsigned long long* bits, const uint32_t* s_pre,
                                    uint32_t base_left) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint32_t* __restrict__ rows_in = a.rows_in;
  const G* __restrict__ gin = reinterpret_cast<const G*>(a.gh_in);
  uint32_t r[kPartPer];
  G g[kPartPer];
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j) {
    const uint32_t p = begin + j * 256 + threadIdx.x;
    r[j] = p < end ? (rows_in ? rows_in[p] : p) : 0xFFFFFFFFu;  // null rows_in: identity
    if (gin && p < end) g[j] = gin[p];
  }
#pragma unroll
  for (uint32_t j = 0; j < kPartPer; ++j) {
    if (r[j] == 0xFFFFFFFFu) continue;
    const uint32_t k = j * 4 + wave;
    const unsigned long long m = bits[k];
    const bool go_left = (m >> lane) & 1ull;
    const uint32_t lefts_before = base_left + s_pre[k] + __popcll(m & below);
    const uint32_t pos_in_node = local0 + j * 256 + threadIdx.x;
    // Destination inside the node's own segment; a left count that disagrees with the bits
    // (a contract violation) is reported instead of writing outside the segment.
    const uint32_t dst = go_left ? lefts_before : nd.left_count + (pos_in_node - lefts_before);
    if (dst < nd.count && (go_left ? lefts_before < nd.left_count : true)) {
      a.rows_out[nd.begin + dst] = r[j];
      if (gin) reinterpret_cast<G*>(a.gh_out)[nd.begin + dst] = g[j];
    } else {
      raise_err(a.err, kErrCount);
    }
  }
}

// G: the gradient element moved with each row (void: none), fixed at compile time so each
// variant keeps only its own registers.
template <typename G>
__global__ __launch_bounds__(256) void k_part_scatter(PartLaunch a) {
  const uint32_t ch = blockIdx.x;
  if (a.ctl && ch >= a.ctl->n_chunks) return;
  const PartNode nd = a.nodes[a.chunk_node[ch]];
  if (nd.action != kPartSplit) return;
  const uint32_t local0 = (ch - nd.chunk0) * kPartChunk;
  const uint32_t begin = nd.begin + local0;
  const uint32_t end = nd.begin + min(nd.count, local0 + kPartChunk);
  __shared__ uint32_t s_pre[kPartChunk / 64 + 1];
  const unsigned long long* bits =
      reinterpret_cast<const unsigned long long*>(a.bits) + size_t(ch) * (kPartChunk / 64);
  if (threadIdx.x < kPartChunk / 64) {  // wave 0: exclusive scan of the 64 word popcounts
    const uint32_t c = __popcll(bits[threadIdx.x]);
    uint32_t incl = c;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d, 64);
      if (threadIdx.x >= uint32_t(d)) incl += o;
    }
    s_pre[threadIdx.x] = incl - c;
    if (threadIdx.x == kPartChunk / 64 - 1) s_pre[kPartChunk / 64] = incl;
  }
  // Lefts in this node's earlier chunks: their counts from the mark pass, added up here.
  __shared__ uint32_t s_red[4];
  uint32_t earlier = 0;
  for (uint32_t c = nd.chunk0 + threadIdx.x; c < ch; c += 256) earlier += a.chunk_left[c];
  for (int d = 32; d >= 1; d >>= 1) earlier += __shfl_xor(earlier, d, 64);
  if ((threadIdx.x & 63u) == 0) s_red[threadIdx.x >> 6] = earlier;
  __syncthreads();
  const uint32_t base_left = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  // The node's last chunk checks the lefts against the split's count.
  if (threadIdx.x == 0 && local0 + kPartChunk >= nd.count && base_left + s_pre[kPartChunk / 64] != nd.left_count)
    raise_err(a.err, kErrCount);
  // The row's gradient moves with it (position-ordered gradients), as a whole element.
  if constexpr (std::is_void<G>::value) {
    PartLaunch b = a;
    b.gh_in = nullptr;
    scatter_rows<uint8_t>(b, nd, begin, end, local0, bits, s_pre, base_left);
  } else {
    scatter_rows<G>(a, nd, begin, end, local0, bits, s_pre, base_left);
  }
}

void launch_partition(const PartLaunch& a, hipStream_t s, int phase) {
  if (a.n_chunks == 0) return;
  // With per-XCD lists, XCD x's list has at most chunks / 8 + nodes entries.
  const uint32_t mark_grid = a.block_chunk && a.ctl ? a.n_chunks + 8 * a.n_nodes : a.n_chunks;
  if (phase != 2) hipLaunchKernelGGL(k_part_mark, dim3(mark_grid), dim3(256), 0, s, a);
  if (phase == 1) return;
  if (a.no_scatter) {  // no rows move: only the per-node count check
    hipLaunchKernelGGL(k_part_scan, dim3(a.n_nodes), dim3(256), 0, s, a.nodes, a.n_nodes, a.chunk_left, a.n_chunks,
                       a.ctl, a.err);
    return;
  }
  const dim3 grid(a.n_chunks), block(256);
  switch (a.gh_in ? a.gh_bytes : 0u) {
    case 1: hipLaunchKernelGGL(k_part_scatter<uint8_t>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_part_scatter<uint16_t>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_part_scatter<uint32_t>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_part_scatter<uint2>, grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL(k_part_scatter<uint4>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_part_scatter<void>, grid, block, 0, s, a);
  }
}

// ============================================================== level planner ======
// Level-wise growth planned on the device (TrainLevelWiseRegressionTree,
// src/core/tree.cpp:49-147): one block turns a level's best splits into the level's
// partition, the next level's histogram builds, slabs and split-scan nodes, and the new tree
// nodes with their leaf values, so a tree's levels run back to back with no host round trip.
// Node numbering, the leaf cap, the smaller-child choice and the slab split are those of the
// host planner; leaf values are MakeBoundedLeaf (src/core/tree_structure.cpp:72-87) with the
// same FP64 operations.
constexpr uint32_t kPlanThreads = 1024;
constexpr uint32_t kResolveMax = 64;  // resolved nodes of one level at most (ResolveLaunch::r_max <= this)
constexpr uint32_t kPlanMaxNodes = 2 * kPlanThreads;  // nodes of one level

// Exclusive prefix sum of v[0, n) in LDS (n <= 2 * blockDim.x), in place; returns the total.
__device__ uint32_t plan_prefix(uint32_t* v, uint32_t n, uint32_t* s_wave) {
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6, waves = blockDim.x / 64;
  const uint32_t a0 = 2 * t < n ? v[2 * t] : 0u, a1 = 2 * t + 1 < n ? v[2 * t + 1] : 0u;
  uint32_t x = a0 + a1;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= uint32_t(d)) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < waves; ++k) {
      const uint32_t c = s_wave[k];
      s_wave[k] = run;
      run += c;
    }
    s_wave[waves] = run;
  }
  __syncthreads();
  const uint32_t excl = s_wave[w] + x - (a0 + a1);
  if (2 * t < n) v[2 * t] = excl;
  if (2 * t + 1 < n) v[2 * t + 1] = excl + a0;
  const uint32_t total = s_wave[waves];
  __syncthreads();
  return total;
}

// MakeBoundedLeaf with LeafWeight's checks in the host's order; 0 or the error bit.
__device__ inline uint32_t bounded_leaf(double lambda, double alpha, double max_delta, double g, double h, double lo,
                                        double hi, double* out) {
  *out = 0.0;
  if (lo > hi) return kErrBounds;
  if (!isfinite(g)) return kErrLeafGradient;
  if (!isfinite(h)) return kErrLeafHessian;
  if (h < 0.0) return kErrLeafNegative;
  const double d = h + lambda;
  if (!isfinite(d) || d <= 0.0) return kErrLeafDenominator;
  const double m = fabs(g) - alpha;
  const double sg = m <= 0.0 ? 0.0 : copysign(m, g);
  double w = -sg / d;
  if (max_delta != 0.0) {
    const double t = w < max_delta ? w : max_delta;  // min(max_delta, w)
    w = -max_delta < t ? t : -max_delta;            // max(-max_delta, t)
  }
  if (!isfinite(w)) return kErrLeafOverflow;
  *out = w < lo ? lo : (hi < w ? hi : w);  // clamp(w, lo, hi)
  return 0u;
}

__device__ inline uint32_t plan_leaf(const PlanLaunch& a, const NodeStats& s, double lo, double hi,
                                     double* out) {
  return bounded_leaf(a.lambda, a.alpha, a.max_delta, s.g, s.h, lo, hi, out);
}

__device__ inline mpsb_tree_node plan_leaf_node(double value) {
  mpsb_tree_node n{};
  n.left_child = n.right_child = 0xFFFFFFFFu;
  n.leaf_value = value;
  n.default_left = 1;
  n.flags = 1;
  return n;
}

// What the planner needs of one node: its layer entry and its best split.
struct PlanIn {
  LayerNode L;
  uint32_t valid, feature, threshold, default_left;
  double gain;
  NodeStats left, right;
  double l_lo, l_hi, r_lo, r_hi;
};

__device__ inline PlanIn plan_in(const PlanLaunch& a, uint32_t i) {
  PlanIn x;
  const Candidate& c = a.cand[i];
  if (a.depth == 0 && a.roots > 1) {  // a batched round's class roots: their segments, scanned stats
    x.L = a.root_layer[i];
    x.L.stats = c.parent;
  } else if (a.depth == 0) {  // level 0's one node: every row, statistics from the root's own scan
    x.L = LayerNode{};
    x.L.count = uint32_t(a.rows_total);
    x.L.lo = -INFINITY;
    x.L.hi = INFINITY;
    x.L.stats = c.parent;
  } else {
    x.L = a.layer[i];
  }
  x.valid = c.valid;
  x.feature = c.feature;
  x.threshold = c.threshold;
  x.default_left = c.default_left;
  x.gain = c.gain;
  x.left = c.left;
  x.right = c.right;
  x.l_lo = c.l_lo;
  x.l_hi = c.l_hi;
  x.r_lo = c.r_lo;
  x.r_hi = c.r_hi;
  return x;
}

// A split's child counts add up to its node's count (each within it, so no sum wraps).
__device__ inline bool plan_counts_ok(const PlanIn& x) {
  return x.left.count <= x.L.count && x.right.count <= x.L.count && x.left.count + x.right.count == x.L.count;
}

constexpr uint32_t kPlanPer = kPlanMaxNodes / kPlanThreads;  // nodes per thread

// XCD of chunk c of node i's C partition chunks in the last level's lists: the eighth of the
// row range holding the chunk's middle (a node's rows are in ascending order, so chunk c covers
// about the fraction [c / C, (c + 1) / C) of them); a node of fewer than 8 chunks spreads them
// over the XCDs in turn.
__device__ inline uint32_t plan_xcd(uint32_t i, uint32_t c, uint32_t C) {
  return C >= 8 ? uint32_t((8ull * c + 4ull) / C) : (i + c) & 7u;
}

__global__ __launch_bounds__(kPlanThreads) void k_plan(PlanLaunch a) {
  __shared__ uint32_t s_wave[kPlanThreads / 64 + 1];
  __shared__ uint32_t s_rank[kPlanMaxNodes + 1];   // valid rank, then split rank per node
  __shared__ uint32_t s_base[kPlanMaxNodes + 1];   // chunk bases, then slab bases
  __shared__ uint32_t s_bcount[kPlanMaxNodes], s_bbegin[kPlanMaxNodes];  // per build: rows
  __shared__ uint32_t s_err_node;
  __shared__ unsigned long long s_rows;
  const uint32_t t = threadIdx.x;
  const uint32_t slot = *a.tree_slot;  // read before the first barrier; advanced at the end
  mpsb_tree_node* const tree = a.tree + size_t(slot) * a.tree_nodes;
  uint32_t* const tree_ctl = a.tree_ctl + 4 * size_t(slot);
  // Every input of this thread's nodes, loaded once, up to the buffers' capacity so the
  // loads do not wait for the node count (entries past it are never used).
  const uint32_t cap = a.depth == 0 ? a.roots : min(a.layer_cap, kPlanMaxNodes);
  PlanIn in[kPlanPer];
#pragma unroll
  for (uint32_t r = 0; r < kPlanPer; ++r) {
    const uint32_t i = t + r * blockDim.x;
    if (i < cap) in[r] = plan_in(a, i);
  }
  const uint32_t n = a.depth == 0 ? a.roots : min(a.ctl_in->n_next, cap);
  const uint32_t nodes0 = a.depth == 0 ? a.roots : tree_ctl[0];
  const uint32_t leaves0 = a.depth == 0 ? a.roots : tree_ctl[1];
  const uint32_t emitted0 = a.depth == 0 ? 0u : tree_ctl[2];  // final leaves numbered so far
  if (t == 0) {
    s_err_node = 0xFFFFFFFFu;
    s_rows = 0;
  }
  if (a.depth == 0) {  // the root leaf of each root (src/core/tree.cpp:37-47)
#pragma unroll
    for (uint32_t r = 0; r < kPlanPer; ++r) {
      const uint32_t i = t + r * blockDim.x;
      if (i >= a.roots) continue;
      double v;
      const uint32_t e = plan_leaf(a, in[r].L.stats, in[r].L.lo, in[r].L.hi, &v);
      if (e) raise_err(a.err, e);
      in[r].L.leaf = v;
      tree[i] = plan_leaf_node(v);
    }
  }
  const bool final = a.depth >= a.max_depth || leaves0 >= a.cap;

  // Nodes the host would split: valid, and within the leaf cap counted in node order. One
  // whose child counts do not add up to its own is an error (the host stops there) and stays
  // a leaf here; the first failing node in node order reports.
#pragma unroll
  for (uint32_t r = 0; r < kPlanPer; ++r) {
    const uint32_t i = t + r * blockDim.x;
    if (i < n) s_rank[i] = !final && in[r].valid ? 1u : 0u;
  }
  __syncthreads();
  plan_prefix(s_rank, n, s_wave);
  bool split[kPlanPer];
  uint32_t err[kPlanPer];
  double lv[kPlanPer], rv[kPlanPer];
#pragma unroll
  for (uint32_t r = 0; r < kPlanPer; ++r) {
    const uint32_t i = t + r * blockDim.x;
    split[r] = false;
    err[r] = 0;
    lv[r] = rv[r] = 0.0;
    if (i >= n) continue;
    const PlanIn& x = in[r];
    if (!final && x.valid && uint64_t(leaves0) + s_rank[i] < a.cap) {
      double l, rr;  // locals: the arrays' addresses would keep them in scratch
      err[r] = plan_leaf(a, x.left, x.l_lo, x.l_hi, &l);
      const uint32_t er = plan_leaf(a, x.right, x.r_lo, x.r_hi, &rr);
      lv[r] = l;
      rv[r] = rr;
      if (!err[r]) err[r] = er;
      if (!err[r] && !plan_counts_ok(x)) err[r] = kErrSplitCount;
      split[r] = plan_counts_ok(x);
    }
    if (err[r]) atomicMin(&s_err_node, i);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kPlanPer; ++r) {
    const uint32_t i = t + r * blockDim.x;
    if (i < n && i == s_err_node) raise_err(a.err, err[r]);
    if (i < n) s_rank[i] = split[r] ? 1u : 0u;
  }
  __syncthreads();
  const uint32_t n_split = plan_prefix(s_rank, n, s_wave);  // s_rank: rank among splits
  const bool children_leaves = a.depth + 1 >= a.max_depth || uint64_t(leaves0) + n_split >= a.cap;

  // Tree nodes (ApplySplit, src/core/tree_structure.cpp:118-148: left then right appended),
  // partition nodes, the next level and its builds.
#pragma unroll
  for (uint32_t r = 0; r < kPlanPer; ++r) {
    const uint32_t i = t + r * blockDim.x;
    if (i >= n) continue;
    const PlanIn& x = in[r];
    // Leaf ordinals number the tree's final leaves in emission order (level by level, node
    // order within a level): a node that stays a leaf takes one, a split into two leaves two.
    // The rows store the ordinal and the leaf table is indexed by it.
    const uint32_t ord = emitted0 + (i - s_rank[i]) + (children_leaves ? 2 * s_rank[i] : 0u);
    PartNode pn{};
    pn.begin = x.L.begin;
    pn.count = x.L.count;
    pn.action = kPartLeaf;
    pn.leaf_node = ord;
    pn.leaf_value = x.L.leaf;
    if (!split[r]) {
      a.leaf_table[ord] = x.L.leaf;
      if (a.leaf_info) a.leaf_info[ord] = LeafInfo{x.L.node, 0u, x.L.lo, x.L.hi};
    }
    pn.missing_code = 0xFFFFFFFFu;
    pn.column_major = uint64_t(x.L.count) * 128 >= a.rows_total ? 1u : 0u;
    if (split[r]) {
      const uint32_t li = nodes0 + 2 * s_rank[i], ri = li + 1;
      tree[li] = plan_leaf_node(lv[r]);
      tree[ri] = plan_leaf_node(rv[r]);
      mpsb_tree_node& parent = tree[x.L.node];
      parent.feature_index = x.feature;
      parent.threshold_bin = x.threshold;
      parent.left_child = li;
      parent.right_child = ri;
      parent.leaf_value = 0.0;
      parent.gain = x.gain;
      parent.default_left = x.default_left ? 1 : 0;
      parent.flags = 0;
      pn.action = children_leaves ? kPartSplitLeaf : kPartSplit;
      pn.feature = x.feature;
      pn.threshold = x.threshold;
      pn.default_left = x.default_left;
      pn.missing_code = a.miss_cell[x.feature] != 0xFFFFFFFFu ? a.bin_count[x.feature] : 0xFFFFFFFFu;
      pn.left_count = uint32_t(x.left.count);
      if (children_leaves) {
        pn.leaf_node = ord;
        pn.leaf_node_right = ord + 1;
        pn.leaf_value = lv[r];
        pn.leaf_right = rv[r];
        a.leaf_table[ord] = lv[r];
        a.leaf_table[ord + 1] = rv[r];
        if (a.leaf_info) {
          a.leaf_info[ord] = LeafInfo{li, 0u, x.l_lo, x.l_hi};
          a.leaf_info[ord + 1] = LeafInfo{ri, 0u, x.r_lo, x.r_hi};
        }
      } else {
        const uint32_t j = 2 * s_rank[i];
        LayerNode l{}, rr{};
        l.node = li;
        rr.node = ri;
        l.begin = x.L.begin;
        l.count = uint32_t(x.left.count);
        rr.begin = x.L.begin + l.count;
        rr.count = uint32_t(x.right.count);
        l.lo = x.l_lo;
        l.hi = x.l_hi;
        rr.lo = x.r_lo;
        rr.hi = x.r_hi;
        l.leaf = lv[r];
        rr.leaf = rv[r];
        l.stats = x.left;
        rr.stats = x.right;
        a.next[j] = l;
        a.next[j + 1] = rr;
        ScanNode sl{}, sr{};
        sl.slot = j;
        sr.slot = j + 1;
        sl.node_index = li;
        sr.node_index = ri;
        sl.depth = sr.depth = a.depth + 1;
        sl.lo = x.l_lo;
        sl.hi = x.l_hi;
        sr.lo = x.r_lo;
        sr.hi = x.r_hi;
        sl.parent = x.left;
        sr.parent = x.right;
        a.scan[j] = sl;
        a.scan[j + 1] = sr;
        // The smaller child is built, the left one on ties (src/core/tree.cpp:109-117).
        const uint32_t b = s_rank[i];
        const bool left_first = l.count <= rr.count;
        a.build_slot[b] = left_first ? j : j + 1;
        a.sibling_slot[b] = int32_t(left_first ? j + 1 : j);
        a.parent_slot[b] = i;
        s_bcount[b] = left_first ? l.count : rr.count;
        s_bbegin[b] = left_first ? l.begin : rr.begin;
      }
    }
    a.part[i] = pn;
    s_base[i] = (x.L.count + kPartChunk - 1) / kPartChunk;
  }
  __syncthreads();

  // Partition chunks: bases per node, then chunk -> node.
  const uint32_t n_chunks = plan_prefix(s_base, n, s_wave);
  if (t == 0) s_base[n] = n_chunks;
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kPlanPer; ++r) {
    const uint32_t i = t + r * blockDim.x;
    if (i < n) a.part[i].chunk0 = s_base[i];
  }
  // Chunk -> node: each wave takes whole nodes and its lanes write their chunk ids.
  const uint32_t lane = t & 63u, waves = blockDim.x / 64;
  for (uint32_t i = t / 64; i < n; i += waves)
    for (uint32_t k = s_base[i] + lane; k < s_base[i + 1]; k += 64) a.chunk_node[k] = i;
  // The last level's per-XCD chunk lists (LevelCtl::xcd_map): counts per XCD, their offsets,
  // then each chunk's place in its XCD's list (the order within a list is free: a chunk's
  // results do not depend on when it runs).
  __shared__ uint32_t s_xc[8], s_xo[8], s_xl[8];
  const bool xmap = children_leaves && a.block_chunk != nullptr;
  if (xmap) {
    if (t < 8) s_xc[t] = 0u;
    __syncthreads();
    for (uint32_t i = t / 64; i < n; i += waves) {
      const uint32_t C = s_base[i + 1] - s_base[i];
      for (uint32_t c = lane; c < C; c += 64) atomicAdd(&s_xc[plan_xcd(i, c, C)], 1u);
    }
    __syncthreads();
    if (t == 0) {
      uint32_t run = 0;
      for (uint32_t x = 0; x < 8; ++x) {
        s_xo[x] = run;
        s_xl[x] = s_xc[x];
        run += s_xc[x];
      }
    }
    __syncthreads();
    if (t < 8) s_xc[t] = 0u;
    __syncthreads();
    for (uint32_t i = t / 64; i < n; i += waves) {
      const uint32_t C = s_base[i + 1] - s_base[i];
      for (uint32_t c = lane; c < C; c += 64) {
        const uint32_t x = plan_xcd(i, c, C);
        a.block_chunk[s_xo[x] + atomicAdd(&s_xc[x], 1u)] = s_base[i] + c;
      }
    }
  }
  __syncthreads();

  // Builds of the next level and their slabs (the host's slab split of plan_build).
  const uint32_t n_build = children_leaves ? 0u : n_split;
  unsigned long long mine = 0;
  for (uint32_t b = t; b < n_build; b += blockDim.x) mine += s_bcount[b];
  for (int d = 32; d >= 1; d >>= 1) mine += __shfl_xor(mine, d, 64);
  if ((t & 63u) == 0 && mine) atomicAdd(&s_rows, mine);
  __syncthreads();
  const uint64_t total = s_rows;
  uint64_t slab_rows = max(a.min_slab_rows, (total + a.want_slabs - 1) / a.want_slabs);
  slab_rows = min(slab_rows, a.slab_cap);
  for (uint32_t b = t; b < n_build; b += blockDim.x)
    s_base[b] = uint32_t(max<uint64_t>(1, (uint64_t(s_bcount[b]) + slab_rows - 1) / slab_rows));
  __syncthreads();
  const uint32_t n_slabs = plan_prefix(s_base, n_build, s_wave);
  if (t == 0) s_base[n_build] = n_slabs;
  __syncthreads();
  for (uint32_t b = t; b <= n_build; b += blockDim.x) a.slab_first[b] = s_base[b];
  for (uint32_t b = t / 64; b < n_build; b += waves) {  // each wave takes whole builds
    const uint64_t end = uint64_t(s_bbegin[b]) + s_bcount[b];
    for (uint32_t k = s_base[b] + lane; k < s_base[b + 1]; k += 64) {
      const uint64_t lo = s_bbegin[b] + uint64_t(k - s_base[b]) * slab_rows;
      a.slabs[k] = SlabDesc{uint32_t(lo), uint32_t(min<uint64_t>(end, lo + slab_rows))};
    }
  }
  if (t == 0) {
    LevelCtl ctl{};
    ctl.n_part = n;
    ctl.n_chunks = n_chunks;
    ctl.n_build = n_build;
    ctl.n_slabs = n_slabs;
    ctl.n_next = children_leaves ? 0u : 2 * n_split;
    ctl.xcd_map = xmap ? 1u : 0u;
    for (uint32_t x = 0; x < 8; ++x) {
      ctl.xcd_off[x] = xmap ? s_xo[x] : 0u;
      ctl.xcd_len[x] = xmap ? s_xl[x] : 0u;
    }
    *a.ctl = ctl;
    tree_ctl[0] = nodes0 + 2 * n_split;
    tree_ctl[1] = leaves0 + n_split;
    tree_ctl[2] = emitted0 + (n - n_split) + (children_leaves ? 2 * n_split : 0u);
    if (n_build) {
      a.stats[0] += total;
      a.stats[1] += 1;
    }
    if (a.depth + 1 >= a.max_depth) *a.tree_slot = slot + 1;  // the tree's last level
  }
}

void launch_plan(const PlanLaunch& a, hipStream_t s) {
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, s, a);
}

// ==================================================================== gradients ======
// Regression objectives (src/core/objective.cpp:101-240), weights applied after
// (src/core/trainer.cpp:200-216). The FP32 form of squared error follows the reference device
// path: labels and predictions rounded to float, g subtracted in float
// (src/mps/kernels/gradients.metal:11-21), then weighted in FP64 and rounded to float for the
// histogram (src/mps/mps_context.mm:82-97). The reference computes every other objective on
// the host in FP64, so the FP32 form does too and only rounds the result.
// Returns 0 or the error bit naming the reference's exception.

// src/core/objective.cpp:120-128
__device__ inline double logistic(double z) {
  if (z >= 0.0) return 1.0 / (1.0 + exp(-z));
  const double e = exp(z);
  return e / (1.0 + e);
}

// OBJ >= 0 fixes the objective at compile time (each variant keeps only its own registers);
// -1 reads it from the launch.
template <int OBJ>
__device__ inline uint32_t objective_at(const GradLaunch& a, double y, double p, bool fp32_form, double* g,
                                        double* h) {
  switch (OBJ >= 0 ? OBJ : a.objective) {
    case 0:  // :101-118
      *g = fp32_form ? double(float(p) - float(y)) : p - y;
      *h = 1.0;
      return isfinite(*g) ? 0u : kErrObjective;
    case 1: {  // :154-172 (labels 0 / 1 were checked with the base score)
      const double q = logistic(p);
      *g = q - y;
      *h = q * (1.0 - q);
      return isfinite(*g) && isfinite(*h) && *h >= 0.0 ? 0u : kErrObjective;
    }
    case 2:  // :174-192
      *g = p < y ? -a.obj_alpha : 1.0 - a.obj_alpha;
      *h = 1.0;
      return 0u;
    case 3: {  // :194-211, exp guarded as SafeExp :63-73
      if (p > 709.0) return kErrExpFirst;
      const double mu = exp(p);
      if (!isfinite(mu)) return kErrExpFirst;
      *g = mu - y;
      *h = mu;
      return isfinite(*g) && isfinite(*h) && *h > 0.0 ? 0u : kErrObjective;
    }
    default: {  // Tweedie :213-240
      const double rho = a.obj_rho;
      const double x1 = (2.0 - rho) * p, x2 = (1.0 - rho) * p;
      if (x1 > 709.0) return kErrExpFirst;
      const double e1 = exp(x1);
      if (!isfinite(e1)) return kErrExpFirst;
      if (x2 > 709.0) return kErrExpSecond;
      const double e2 = exp(x2);
      if (!isfinite(e2)) return kErrExpSecond;
      *g = e1 - y * e2;
      *h = (2.0 - rho) * e1 - (1.0 - rho) * y * e2;
      return isfinite(*g) && isfinite(*h) && *h > 0.0 ? 0u : kErrObjective;
    }
  }
}

// One regression row's inputs: label, prediction (with the deferred leaf update of the
// previous round applied: the same single FP64 addition pred + lr * leaf the reference
// performs at the end of that round) and weight.
struct RowIn {
  double y, p, w;
};

// The inputs of rows i0 + u * stride, u < U, loaded together so a wave keeps every load in
// flight at once. Indexes past the end are clamped to the last row (loads stay in bounds and
// need no branch); the caller discards those rows.
template <uint32_t U>
__device__ inline void load_rows_at(const GradLaunch& a, const uint64_t (&idx)[U], RowIn (&v)[U]);

template <uint32_t U>
__device__ inline void load_rows(const GradLaunch& a, uint64_t i0, uint32_t stride, RowIn (&v)[U]) {
  uint64_t idx[U];
#pragma unroll
  for (uint32_t u = 0; u < U; ++u) idx[u] = min(i0 + uint64_t(u) * stride, a.n - 1);
  load_rows_at<U>(a, idx, v);
}

// The inputs of rows idx[u] (each below a.n), loaded together.
template <uint32_t U>
__device__ inline void load_rows_at(const GradLaunch& a, const uint64_t (&idx)[U], RowIn (&v)[U]) {
  uint32_t leaf[U];
  if (a.leaf_node) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u)
      leaf[u] = a.leaf_bytes == 1 ? reinterpret_cast<const uint8_t*>(a.leaf_node)[idx[u]] : a.leaf_node[idx[u]];
  }
#pragma unroll
  for (uint32_t u = 0; u < U; ++u) {
    v[u].y = a.labels[idx[u]];
    v[u].p = a.pred[idx[u]];
    v[u].w = a.weights ? a.weights[idx[u]] : 1.0;
  }
  if (a.leaf_node) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) v[u].p = v[u].p + a.lr * a.leaf_table[leaf[u]];
  }
}

// Weighted gradient and Hessian of one regression row.
template <int OBJ>
__device__ inline uint32_t grad_of(const GradLaunch& a, const RowIn& v, bool fp32_form, double* g, double* h) {
  if (!isfinite(v.p)) return kErrPrediction;
  const uint32_t e = objective_at<OBJ>(a, v.y, v.p, fp32_form, g, h);
  if (e) return e;
  *g *= v.w;
  *h *= v.w;
  return isfinite(*g) && isfinite(*h) && *h >= 0.0 ? 0u : kErrWeighted;
}

// Softmax class column (src/core/objective.cpp:130-152,242-287), weighted, of rows i0 + u * stride
// (u < R), from class-major margins [K][rows]. The labels, weights and leaf ids of the R rows are
// loaded first (indexes past the end are clamped to the last row, which the caller discards),
// then the K margins of each row are read twice, for the maximum and then for the exponentials
// (the second read hits the cache), so a row holds a few registers and more waves keep loads in
// flight. Any non-finite margin is the prediction error; the exponentials add up in class order.
// kBuffer (margins of at most 4 GiB): one buffer resource covers the margins and each load is
// the row's byte offset (one register for all classes) plus the class column's offset in a
// scalar register; otherwise each load takes a 64-bit address.
//
// With a.leaf_node, the previous tree's deferred update margin += lr * leaf is applied first to
// column a.pending_class (the reference's AddClassUpdate arithmetic); `store` writes the updated
// column back (the last pass that evaluates these rows).
template <uint32_t R, bool kBuffer>
__device__ inline void softmax_rows(const GradLaunch& a, uint64_t i0, uint32_t stride, bool store, double (&g)[R],
                                    double (&h)[R], uint32_t (&bad)[R]) {
  const uint32_t K = a.classes;
  const uint32_t pc = a.leaf_node ? a.pending_class : 0xFFFFFFFFu;
  uint32_t off[R], leaf[R];
  uint64_t row[R];
  double lab[R], w[R], upd[R], top[R], z[R], pk[R];
  bool finite[R];
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(a.pred), 0, -1, 0x00020000);
  const uint32_t col_bytes = uint32_t(a.n * 8u);
#pragma unroll
  for (uint32_t u = 0; u < R; ++u) {
    const uint64_t i = min(i0 + uint64_t(u) * stride, a.n - 1);
    row[u] = i;
    off[u] = uint32_t(i) * 8u;
    lab[u] = a.labels[i];
    w[u] = a.weights ? a.weights[i] : 1.0;
    if (a.leaf_node)
      leaf[u] = a.leaf_bytes == 1 ? reinterpret_cast<const uint8_t*>(a.leaf_node)[i] : a.leaf_node[i];
  }
  auto margin = [&](uint32_t u, uint32_t k) {
    if constexpr (kBuffer) {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off[u], k * col_bytes, 0));
    } else {
      return a.pred[uint64_t(k) * a.n + row[u]];
    }
  };
  if (a.leaf_node) {
#pragma unroll
    for (uint32_t u = 0; u < R; ++u) {
      upd[u] = margin(u, pc) + a.lr * a.leaf_table[leaf[u]];
      if (store && i0 + uint64_t(u) * stride < a.n) a.pred_out[uint64_t(pc) * a.n + row[u]] = upd[u];
    }
  }
#pragma unroll
  for (uint32_t u = 0; u < R; ++u) {
    top[u] = -INFINITY;
    finite[u] = true;
    z[u] = pk[u] = 0.0;
  }
  // Both class loops load class k+1's margins before they use class k's, so each wait covers
  // two columns of loads instead of one; the arithmetic runs in class order as before. The
  // loads are unconditional (the last step loads its own column again), the pending column's
  // update replaces its margin where the margin is used, and the two register sets alternate
  // (no copies), so the compiler waits only for the older column.
  auto column = [&](uint32_t k, double (&v)[R]) {
#pragma unroll
    for (uint32_t u = 0; u < R; ++u) v[u] = margin(u, k);
  };
  auto max_step = [&](uint32_t k, const double (&v)[R]) {
#pragma unroll
    for (uint32_t u = 0; u < R; ++u) {
      const double x = k == pc ? upd[u] : v[u];
      finite[u] = finite[u] && isfinite(x);
      top[u] = fmax(top[u], x);
    }
  };
  auto exp_step = [&](uint32_t k, const double (&v)[R]) {
#pragma unroll
    for (uint32_t u = 0; u < R; ++u) {
      const double e = exp((k == pc ? upd[u] : v[u]) - top[u]);
      z[u] += e;
      if (k == a.target) pk[u] = e;
    }
  };
  double va[R], vb[R];
  column(0, va);
  for (uint32_t k = 0; k < K; k += 2) {
    column(min(k + 1, K - 1), vb);
    max_step(k, va);
    if (k + 1 >= K) break;
    column(min(k + 2, K - 1), va);
    max_step(k + 1, vb);
  }
  column(0, va);
  for (uint32_t k = 0; k < K; k += 2) {
    column(min(k + 1, K - 1), vb);
    exp_step(k, va);
    if (k + 1 >= K) break;
    column(min(k + 2, K - 1), va);
    exp_step(k + 1, vb);
  }
#pragma unroll
  for (uint32_t u = 0; u < R; ++u) {
    g[u] = h[u] = 0.0;
    if (!finite[u]) {
      bad[u] = kErrPrediction;
      continue;
    }
    const double prob = pk[u] / z[u];
    const double gu = prob - (uint32_t(lab[u]) == a.target ? 1.0 : 0.0);
    const double hu = prob * (1.0 - prob);
    if (!isfinite(gu) || !isfinite(hu) || hu < 0.0) {
      bad[u] = kErrObjective;
      continue;
    }
    g[u] = gu * w[u];
    h[u] = hu * w[u];
    bad[u] = isfinite(g[u]) && isfinite(h[u]) && h[u] >= 0.0 ? 0u : kErrWeighted;
  }
}

__device__ inline uint32_t grad_at(const GradLaunch& a, uint64_t i, bool fp32_form, double* g, double* h) {
  if (a.source == 1) {
    *g = a.pairs[2 * i];
    *h = a.pairs[2 * i + 1];
    return isfinite(*g) && isfinite(*h) && *h >= 0.0 ? 0u : kErrNonFinite;
  }
  RowIn v[1];  // source 0 (softmax columns have their own kernel variant)
  load_rows<1>(a, i, 0, v);
  return grad_of<-1>(a, v[0], fp32_form, g, h);
}

// Rows per thread per step of the gradient passes: their loads are issued together.
constexpr uint32_t kGradRows = 4;
constexpr int kObjSoftmax = 16;  // kernel variant for softmax class columns

// Gradients and Hessians of rows i0 + u * blockDim.x (u < kGradRows); rows past the end get
// g = h = 0 and no error.
// `store`: softmax rows write their deferred margin update back (the last evaluating pass).
template <int OBJ>
__device__ inline void grad_rows(const GradLaunch& a, uint64_t i0, bool fp32_form, bool store,
                                 double (&g)[kGradRows], double (&h)[kGradRows], uint32_t (&bad)[kGradRows],
                                 RowIn (&v)[kGradRows]) {
  if constexpr (OBJ == kObjSoftmax) {
    if (a.wide_margins) {
      softmax_rows<kGradRows, false>(a, i0, blockDim.x, store, g, h, bad);
    } else {
      softmax_rows<kGradRows, true>(a, i0, blockDim.x, store, g, h, bad);
    }
  } else if (OBJ >= 0 || a.source == 0) {  // a fixed regression objective means regression rows
    load_rows<kGradRows>(a, i0, blockDim.x, v);
#pragma unroll
    for (uint32_t u = 0; u < kGradRows; ++u) bad[u] = grad_of<OBJ>(a, v[u], fp32_form, &g[u], &h[u]);
  } else if constexpr (OBJ < 0) {
#pragma unroll
    for (uint32_t u = 0; u < kGradRows; ++u) {
      const uint64_t i = i0 + uint64_t(u) * blockDim.x;
      bad[u] = i < a.n ? grad_at(a, i, fp32_form, &g[u], &h[u]) : 0u;
    }
  }
#pragma unroll
  for (uint32_t u = 0; u < kGradRows; ++u) {
    if (i0 + uint64_t(u) * blockDim.x >= a.n) {
      g[u] = h[u] = 0.0;
      bad[u] = 0u;
    }
  }
}

template <int OBJ>
__global__ __launch_bounds__(256) void k_grad_maxabs(GradLaunch a) {
  __shared__ double sg[4], sh[4];
  double mg = 0.0, mh = 0.0;
  const bool fp32_form = a.precision == MPSB_PRECISION_FP32;
  const uint64_t step = uint64_t(gridDim.x) * blockDim.x * kGradRows;
  for (uint64_t i0 = blockIdx.x * uint64_t(blockDim.x) * kGradRows + threadIdx.x; i0 < a.n; i0 += step) {
    double g[kGradRows], h[kGradRows];
    uint32_t bad[kGradRows];
    RowIn v[kGradRows];
    grad_rows<OBJ>(a, i0, fp32_form, a.stash != nullptr, g, h, bad, v);
#pragma unroll
    for (uint32_t u = 0; u < kGradRows; ++u) {
      if (bad[u]) {
        raise_err(a.err, bad[u]);
        g[u] = h[u] = 0.0;
      }
      if (OBJ == kObjSoftmax && a.stash) {
        const uint64_t i = i0 + uint64_t(u) * blockDim.x;
        if (i < a.n) reinterpret_cast<double2*>(a.stash)[i] = make_double2(g[u], h[u]);
      }
      if (OBJ != kObjSoftmax && a.gstash) {
        // Regression rows: this pass keeps the gradients for the write pass (which then reads
        // them instead of the inputs) and applies the deferred update, once.
        const uint64_t i = i0 + uint64_t(u) * blockDim.x;
        if (i < a.n) {
          if (a.unit_h) {
            a.gstash[i] = g[u];
          } else {
            reinterpret_cast<double2*>(a.gstash)[i] = make_double2(g[u], h[u]);
          }
          if (a.pred_out && a.source == 0) a.pred_out[i] = v[u].p;
        }
      }
      mg = fmax(mg, fabs(g[u]));
      mh = fmax(mh, fabs(h[u]));
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    mg = fmax(mg, __shfl_xor(mg, d, 64));
    mh = fmax(mh, __shfl_xor(mh, d, 64));
  }
  if ((threadIdx.x & 63u) == 0) {
    sg[threadIdx.x >> 6] = mg;
    sh[threadIdx.x >> 6] = mh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
      mg = fmax(mg, sg[w]);
      mh = fmax(mh, sh[w]);
    }
    // Non-negative doubles order like their bit patterns, so an integer max is exact. The
    // maxima only grow, so a block whose value is already reached skips its atomic.
    const unsigned long long bg = static_cast<unsigned long long>(__double_as_longlong(mg));
    const unsigned long long bh = static_cast<unsigned long long>(__double_as_longlong(mh));
    if (__hip_atomic_load(&a.maxabs[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bg) atomicMax(&a.maxabs[0], bg);
    if (__hip_atomic_load(&a.maxabs[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bh) atomicMax(&a.maxabs[1], bh);
  }
}

__device__ inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Unbiased stochastic rounding of v (|v| <= qmax) with a counter-based uniform keyed by
// (seed, row, channel): floor(v + u), u in [0, 1).
__device__ inline long long stochastic_round(double v, uint64_t seed, uint64_t i, uint32_t ch, long long qmax) {
  const uint64_t bits = mix64(seed ^ mix64(i * 2 + ch));
  const double u = double(bits >> 11) * (1.0 / 9007199254740992.0);
  long long q = static_cast<long long>(floor(v + u));
  return q > qmax ? qmax : (q < -qmax ? -qmax : q);
}

// Multipliers (q = value * mult) and value scales (value = q * scale) of the fixed-point and
// integer forms from this round's max |g|, |h|:
//  * FP64_FIXED: mult = 2^S with S = 62 - ceil(log2(rows * max|x|)), so |any sum| <= 2^62;
//  * INT16 / INT8: mult = qmax / max|x|, scale = max|x| / qmax.
__device__ inline void derive_scales(const GradLaunch& a, double* mg_out, double* mh_out, double* vg, double* vh) {
  const double m[2] = {__longlong_as_double(static_cast<long long>(a.maxabs[0])),
                       __longlong_as_double(static_cast<long long>(a.maxabs[1]))};
  double mult[2], val[2];
  for (int k = 0; k < 2; ++k) {
    if (a.precision == MPSB_PRECISION_FP64_FIXED) {
      int sh = 0;
      if (m[k] > 0.0 && isfinite(m[k])) {
        int e;
        frexp(m[k] * double(a.n), &e);
        sh = 62 - e;
      }
      mult[k] = ldexp(1.0, sh);
      val[k] = ldexp(1.0, -sh);
    } else {
      const double qmax = a.precision == MPSB_PRECISION_INT16 ? 32767.0 : 127.0;
      mult[k] = m[k] > 0.0 ? qmax / m[k] : 1.0;
      val[k] = m[k] > 0.0 ? m[k] / qmax : 1.0;
    }
  }
  *mg_out = mult[0];
  *mh_out = mult[1];
  *vg = val[0];
  *vh = a.unit_h ? 1.0 : val[1];
}

// Stores one row's gradient pair in the launch's precision.
__device__ inline void store_gh(const GradLaunch& a, uint64_t i, double g, double h) {
  switch (a.precision) {
    case MPSB_PRECISION_FP32:
      if (a.unit_h) {
        static_cast<float*>(a.gh)[i] = float(g);
      } else {
        static_cast<float2*>(a.gh)[i] = make_float2(float(g), float(h));
      }
      break;
    case MPSB_PRECISION_FP64:
      static_cast<double2*>(a.gh)[i] = make_double2(g, h);
      break;
    case MPSB_PRECISION_FP64_FIXED: {
      // Fixed point with a power-of-two scale: exact conversion, round to nearest even.
      const long long gq = static_cast<long long>(rint(g * a.scale_g));
      if (a.unit_h) {
        static_cast<long long*>(a.gh)[i] = gq;
      } else {
        static_cast<longlong2*>(a.gh)[i] = make_longlong2(gq, static_cast<long long>(rint(h * a.scale_h)));
      }
      break;
    }
    case MPSB_PRECISION_INT16: {
      const long long gq = stochastic_round(g * a.scale_g, a.seed, i, 0, 32767);
      if (a.unit_h) {
        static_cast<short*>(a.gh)[i] = short(gq);
      } else {
        const long long hq = stochastic_round(h * a.scale_h, a.seed, i, 1, 32767);
        static_cast<short2*>(a.gh)[i] = make_short2(short(gq), short(hq));
      }
      break;
    }
    default: {
      const long long gq = stochastic_round(g * a.scale_g, a.seed, i, 0, 127);
      if (a.unit_h) {
        static_cast<signed char*>(a.gh)[i] = static_cast<signed char>(gq);
      } else {
        const long long hq = stochastic_round(h * a.scale_h, a.seed, i, 1, 127);
        static_cast<char2*>(a.gh)[i] = make_char2(static_cast<signed char>(gq), static_cast<signed char>(hq));
      }
    }
  }
}

template <int OBJ>
__global__ __launch_bounds__(256) void k_grad_write(GradLaunch a) {
  const bool fp32_form = a.precision == MPSB_PRECISION_FP32;
  if (a.maxabs) {
    double vg, vh;
    derive_scales(a, &a.scale_g, &a.scale_h, &vg, &vh);
    if (a.scales && blockIdx.x == 0 && threadIdx.x == 0) {
      a.scales[0] = a.scale_g;
      a.scales[1] = a.scale_h;
      a.scales[2] = vg;
      a.scales[3] = vh;
    }
  }
  const uint64_t step = uint64_t(gridDim.x) * blockDim.x * kGradRows;
  for (uint64_t i0 = blockIdx.x * uint64_t(blockDim.x) * kGradRows + threadIdx.x; i0 < a.n; i0 += step) {
    double g[kGradRows], h[kGradRows];
    uint32_t bad[kGradRows];
    RowIn v[kGradRows];
    // Pass 1 kept these rows' gradients (failed rows hold zeros and have raised their error).
    const bool kept = a.maxabs && (OBJ == kObjSoftmax ? a.stash != nullptr : a.gstash != nullptr);
    if (kept) {
#pragma unroll
      for (uint32_t u = 0; u < kGradRows; ++u) {
        const uint64_t i = min(i0 + uint64_t(u) * blockDim.x, a.n - 1);
        if (OBJ == kObjSoftmax || !a.unit_h) {
          const double2 q = reinterpret_cast<const double2*>(OBJ == kObjSoftmax ? a.stash : a.gstash)[i];
          g[u] = q.x;
          h[u] = q.y;
        } else {
          g[u] = a.gstash[i];
          h[u] = 1.0;
        }
        bad[u] = 0u;
      }
    } else {
      grad_rows<OBJ>(a, i0, fp32_form, true, g, h, bad, v);
    }
#pragma unroll
    for (uint32_t u = 0; u < kGradRows; ++u) {
      const uint64_t i = i0 + uint64_t(u) * blockDim.x;
      if (i >= a.n) break;
      if (bad[u]) {
        raise_err(a.err, bad[u]);
        g[u] = h[u] = 0.0;
      }
      if (!kept && a.pred_out && a.source == 0) a.pred_out[i] = v[u].p;  // the deferred update
      if (a.gh64) {
        a.gh64[2 * i] = g[u];
        a.gh64[2 * i + 1] = h[u];
      }
      if (!kept && a.gstash) {
        if (a.unit_h) {
          a.gstash[i] = g[u];
        } else {
          reinterpret_cast<double2*>(a.gstash)[i] = make_double2(g[u], h[u]);
        }
      }
      store_gh(a, i, g[u], h[u]);
    }
  }
}

// Batched softmax rounds (the Jacobi form; no counterpart in the reference, whose class trees
// run in class order, src/core/multiclass_trainer.cpp:132-147): every class column's weighted
// (g, h) from the round's margins, one softmax evaluation per row for all classes. The previous
// round's class trees are applied first, each column's deferred update margin + lr * leaf
// (leaf ids [K][rows]) with the reference's AddClassUpdate arithmetic, and written back. The K
// margins of a row are read three times (maximum, exponential sum, probabilities: the later
// reads hit the cache) and written once, by the last pass. Pass 1 of the round: class k, row i
// goes to stash[k * rows + i], and max |g|, |h| over all of them to maxabs; the write pass then
// quantizes the K x rows stash like one column.
__device__ inline double class_margin(const GradLaunch& a, uint64_t at) {
  double m = a.pred[at];
  if (a.leaf_node) {
    const uint32_t id = a.leaf_bytes == 1 ? reinterpret_cast<const uint8_t*>(a.leaf_node)[at] : a.leaf_node[at];
    m = m + a.lr * a.leaf_table[id];
  }
  return m;
}

// Up to this many classes a row's K updated margins stay in registers: their loads (margins and
// leaf ids, then the leaf values) are issued together, once per row. Larger K reads them again
// in each of the three class loops. The values and their order are the same either way.
constexpr uint32_t kRegClasses = 16;

__global__ __launch_bounds__(256) void k_grad_softmax_all(GradLaunch a) {
  __shared__ double sg[4], sh[4];
  double mg = 0.0, mh = 0.0;
  const uint32_t K = a.classes;
  const uint64_t n = a.n;
  double2* stash = reinterpret_cast<double2*>(a.stash);
  // Class k's gradient pair of row i from its updated margin m, the row's largest margin and
  // the exponential sum.
  auto emit = [&](uint64_t at, uint32_t k, double m, uint32_t label, double w, bool finite, double top, double z) {
    if (a.leaf_node) a.pred_out[at] = m;  // the last read of this margin
    double g = 0.0, h = 0.0;
    uint32_t bad = 0u;
    if (!finite) {
      bad = kErrPrediction;
    } else {
      const double prob = exp(m - top) / z;
      const double gu = prob - (label == k ? 1.0 : 0.0);
      const double hu = prob * (1.0 - prob);
      if (!isfinite(gu) || !isfinite(hu) || hu < 0.0) {
        bad = kErrObjective;
      } else {
        g = gu * w;
        h = hu * w;
        if (!(isfinite(g) && isfinite(h) && h >= 0.0)) bad = kErrWeighted;
      }
    }
    if (bad) {
      raise_err(a.err, bad);
      g = h = 0.0;
    }
    stash[at] = make_double2(g, h);
    mg = fmax(mg, fabs(g));
    mh = fmax(mh, fabs(h));
  };
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t label = uint32_t(a.labels[i]);
    const double w = a.weights ? a.weights[i] : 1.0;
    double top = -INFINITY;
    bool finite = true;
    double z = 0.0;
    if (K <= kRegClasses) {
      double mk[kRegClasses];
      uint32_t idk[kRegClasses];
#pragma unroll
      for (uint32_t k = 0; k < kRegClasses; ++k) {
        if (k < K) {
          const uint64_t at = uint64_t(k) * n + i;
          mk[k] = a.pred[at];
          if (a.leaf_node)
            idk[k] = a.leaf_bytes == 1 ? reinterpret_cast<const uint8_t*>(a.leaf_node)[at] : a.leaf_node[at];
        }
      }
      if (a.leaf_node) {
#pragma unroll
        for (uint32_t k = 0; k < kRegClasses; ++k)
          if (k < K) mk[k] = mk[k] + a.lr * a.leaf_table[idk[k]];  // as class_margin
      }
#pragma unroll
      for (uint32_t k = 0; k < kRegClasses; ++k) {
        if (k < K) {
          finite = finite && isfinite(mk[k]);
          top = fmax(top, mk[k]);
        }
      }
#pragma unroll
      for (uint32_t k = 0; k < kRegClasses; ++k)
        if (k < K) z += exp(mk[k] - top);  // class order
#pragma unroll
      for (uint32_t k = 0; k < kRegClasses; ++k)
        if (k < K) emit(uint64_t(k) * n + i, k, mk[k], label, w, finite, top, z);
      continue;
    }
    for (uint32_t k = 0; k < K; ++k) {
      const double m = class_margin(a, uint64_t(k) * n + i);
      finite = finite && isfinite(m);
      top = fmax(top, m);
    }
    for (uint32_t k = 0; k < K; ++k) z += exp(class_margin(a, uint64_t(k) * n + i) - top);  // class order
    for (uint32_t k = 0; k < K; ++k) {
      const uint64_t at = uint64_t(k) * n + i;
      emit(at, k, class_margin(a, at), label, w, finite, top, z);
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    mg = fmax(mg, __shfl_xor(mg, d, 64));
    mh = fmax(mh, __shfl_xor(mh, d, 64));
  }
  if ((threadIdx.x & 63u) == 0) {
    sg[threadIdx.x >> 6] = mg;
    sh[threadIdx.x >> 6] = mh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
      mg = fmax(mg, sg[w]);
      mh = fmax(mh, sh[w]);
    }
    const unsigned long long bg = static_cast<unsigned long long>(__double_as_longlong(mg));
    const unsigned long long bh = static_cast<unsigned long long>(__double_as_longlong(mh));
    if (__hip_atomic_load(&a.maxabs[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bg) atomicMax(&a.maxabs[0], bg);
    if (__hip_atomic_load(&a.maxabs[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bh) atomicMax(&a.maxabs[1], bh);
  }
}

static uint32_t grid_for(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return uint32_t(b < 4096 ? (b ? b : 1) : 4096);
}

// Blocks of 256 threads, kGradRows rows per thread per step; the max pass ends each block in
// at most two device-scope atomics, so it keeps to fewer blocks.
static uint32_t grad_blocks(uint64_t n, uint32_t cap) {
  const uint64_t b = (n + 256 * kGradRows - 1) / (256 * kGradRows);
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(b, cap)));
}
void launch_grad_maxabs(const GradLaunch& a, hipStream_t s) {
  const dim3 grid(grad_blocks(a.n, 2048)), block(256);
  if (a.source == 3) {  // every class column of a batched softmax round
    hipLaunchKernelGGL(k_grad_softmax_all, dim3(grid_for(a.n)), block, 0, s, a);
    return;
  }
  switch (a.source == 0 ? a.objective : a.source == 2 ? kObjSoftmax : -1) {  // one variant per objective
    case kObjSoftmax: hipLaunchKernelGGL(k_grad_maxabs<kObjSoftmax>, grid, block, 0, s, a); break;
    case 0: hipLaunchKernelGGL(k_grad_maxabs<0>, grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_grad_maxabs<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_grad_maxabs<2>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_grad_maxabs<3>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_grad_maxabs<4>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_grad_maxabs<-1>, grid, block, 0, s, a);
  }
}
void launch_grad_write(const GradLaunch& a, hipStream_t s) {
  const dim3 grid(grad_blocks(a.n, 4096)), block(256);
  switch (a.source == 0 ? a.objective : a.source == 2 ? kObjSoftmax : -1) {
    case kObjSoftmax: hipLaunchKernelGGL(k_grad_write<kObjSoftmax>, grid, block, 0, s, a); break;
    case 0: hipLaunchKernelGGL(k_grad_write<0>, grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_grad_write<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_grad_write<2>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_grad_write<3>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_grad_write<4>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_grad_write<-1>, grid, block, 0, s, a);
  }
}

// ================================================================= leaf renewal ======
// Fixed-point shift of a channel: 2^shift * (rows * max|x|) stays below 2^62, so every sum of
// rows's rint(x * 2^shift) fits an int64 exactly (the FP64_FIXED form's scale).
__device__ inline int renew_shift(unsigned long long bits, uint64_t rows) {
  const double m = __longlong_as_double(static_cast<long long>(bits));
  if (!(m > 0.0) || !isfinite(m)) return 0;
  int e;
  frexp(m * double(rows), &e);
  return 62 - e;
}

constexpr uint32_t kRenewLds = 256;        // ids accumulated in LDS by the small variant (one-byte ids)
constexpr uint32_t kRenewBlocks = 2048;    // blocks of the per-leaf sums at most
constexpr uint32_t kRenewLdsLarge = 2048;  // by the large variant (40 KB of LDS); more ids: global atomics

__global__ void k_renew_clear(long long* sums, uint32_t* counts, uint32_t n_ids, unsigned long long* own_maxabs) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_ids; i += gridDim.x * blockDim.x) {
    sums[2 * i] = 0;
    sums[2 * i + 1] = 0;
    counts[i] = 0u;
  }
  if (own_maxabs && blockIdx.x == 0 && threadIdx.x < 2) own_maxabs[threadIdx.x] = 0ull;
}

// Per-leaf fixed-point sums of the round's gradients, rows in row order (coalesced). Ids below
// IDS accumulate in LDS per block first (a batched softmax round's tree has up to 2,048 leaves,
// which all rows would otherwise hit through a few thousand global addresses); the blocks'
// totals meet in int64 atomics, which are exact and independent of their order, in the copy of
// the block's XCD (block b runs on XCD b mod 8), so each address takes an eighth of the blocks.
template <int OBJ, uint32_t IDS>
__global__ __launch_bounds__(256) void k_renew_sum(RenewLaunch r) {
  __shared__ unsigned long long s_g[IDS], s_h[IDS];
  __shared__ uint32_t s_c[IDS];
  const GradLaunch& a = r.g;
  const bool local = r.leaf_bytes == 1 || r.n_ids <= IDS;  // every id below IDS (one-byte ids: below 256)
  const size_t copy = size_t(blockIdx.x % kRenewCopies) * r.copy_stride;
  long long* const sums = r.sums + 2 * copy;
  uint32_t* const counts = r.counts + copy;
  if (local) {
    for (uint32_t i = threadIdx.x; i < IDS; i += blockDim.x) {
      s_g[i] = s_h[i] = 0ull;
      s_c[i] = 0u;
    }
    __syncthreads();
  }
  const double mg = ldexp(1.0, renew_shift(r.maxabs[0], a.n));
  const double mh = ldexp(1.0, renew_shift(r.maxabs[1], a.n));
  const bool fp32_form = a.precision == MPSB_PRECISION_FP32;
  const uint64_t step = uint64_t(gridDim.x) * blockDim.x * kGradRows;
  for (uint64_t i0 = blockIdx.x * uint64_t(blockDim.x) * kGradRows + threadIdx.x; i0 < a.n; i0 += step) {
    double g[kGradRows], h[kGradRows];
    uint32_t bad[kGradRows];
    RowIn v[kGradRows];
    if (OBJ == kObjSoftmax) {
#pragma unroll
      for (uint32_t u = 0; u < kGradRows; ++u) {
        const double2 q = reinterpret_cast<const double2*>(a.stash)[min(i0 + uint64_t(u) * blockDim.x, a.n - 1)];
        g[u] = q.x;
        h[u] = q.y;
        bad[u] = 0u;
      }
    } else if (a.gstash) {  // the write pass's copy (its failed rows hold zeros)
#pragma unroll
      for (uint32_t u = 0; u < kGradRows; ++u) {
        const uint64_t i = min(i0 + uint64_t(u) * blockDim.x, a.n - 1);
        if (a.unit_h) {
          g[u] = a.gstash[i];
          h[u] = 1.0;
        } else {
          const double2 q = reinterpret_cast<const double2*>(a.gstash)[i];
          g[u] = q.x;
          h[u] = q.y;
        }
        bad[u] = 0u;
      }
    } else {
      grad_rows<OBJ>(a, i0, fp32_form, false, g, h, bad, v);
    }
#pragma unroll
    for (uint32_t u = 0; u < kGradRows; ++u) {
      const uint64_t i = i0 + uint64_t(u) * blockDim.x;
      if (i >= a.n) break;
      // A failed row raised its error in the gradient pass and counts as zero there too.
      const long long gq = bad[u] ? 0ll : static_cast<long long>(rint(g[u] * mg));
      const long long hq = bad[u] || a.unit_h ? 0ll : static_cast<long long>(rint(h[u] * mh));
      const uint32_t id = r.leaf_bytes == 1 ? reinterpret_cast<const uint8_t*>(r.leaf_id)[i]
                                            : reinterpret_cast<const uint16_t*>(r.leaf_id)[i];
      if (id >= r.n_ids) {
        raise_err(r.err, kErrSplitCount);  // a row without a leaf of this tree
        continue;
      }
      if (local) {
        atomicAdd(&s_g[id], static_cast<unsigned long long>(gq));
        if (!a.unit_h) atomicAdd(&s_h[id], static_cast<unsigned long long>(hq));
        atomicAdd(&s_c[id], 1u);
      } else {
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * id]), static_cast<unsigned long long>(gq));
        if (!a.unit_h)
          atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * id + 1]), static_cast<unsigned long long>(hq));
        atomicAdd(&counts[id], 1u);
      }
    }
  }
  if (local) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < min(r.n_ids, IDS); i += blockDim.x) {
      if (!s_c[i]) continue;
      atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * i]), s_g[i]);
      if (!a.unit_h) atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * i + 1]), s_h[i]);
      atomicAdd(&counts[i], s_c[i]);
    }
  }
}

// Every leaf in use: its value from the exact sums (MakeBoundedLeaf with its bounds), into the
// deferred-update table and the tree.
__global__ void k_renew_leaf(RenewLaunch r) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t used = r.n_ids;
  mpsb_tree_node* tree = r.trees;
  if (r.tree_slot) {
    const uint32_t slot = *r.tree_slot - 1u;  // the tree just grown (its last level advanced the slot)
    used = min(used, r.tree_ctl[4 * size_t(slot) + 2]);
    tree = r.trees + size_t(slot) * r.tree_nodes;
  }
  if (id >= used) return;
  long long sg = 0, sh = 0;  // the copies' int64 totals: exact, in copy order
  uint32_t cnt = 0;
  for (uint32_t k = 0; k < kRenewCopies; ++k) {
    const size_t at = size_t(k) * r.copy_stride + id;
    sg += r.sums[2 * at];
    sh += r.sums[2 * at + 1];
    cnt += r.counts[at];
  }
  if (cnt == 0u) return;
  const double g = ldexp(double(sg), -renew_shift(r.maxabs[0], r.g.n));
  const double h = r.g.unit_h ? double(cnt) : ldexp(double(sh), -renew_shift(r.maxabs[1], r.g.n));
  const LeafInfo& li = r.info[id];
  double w;
  const uint32_t e = bounded_leaf(r.lambda, r.alpha, r.max_delta, g, h, li.lo, li.hi, &w);
  if (e) {
    raise_err(r.err, e);
    return;
  }
  r.leaf_table[id] = w;
  if (tree) tree[li.node].leaf_value = w;
  if (r.node_out) r.node_out[li.node] = w;
}

void launch_renew(const RenewLaunch& r, hipStream_t s) {
  if (r.n_ids > r.copy_stride) fail(MPSB_ERR_INTERNAL, "leaf renewal: more leaf ids than the totals hold");
  const uint32_t n_clear = kRenewCopies * r.copy_stride;
  hipLaunchKernelGGL(k_renew_clear, dim3(std::max(1u, std::min(64u, (n_clear + 255) / 256))), dim3(256), 0, s,
                     r.sums, r.counts, n_clear, r.own_maxabs);
  RenewLaunch q = r;
  if (r.own_maxabs) {
    GradLaunch m = r.g;
    m.maxabs = r.own_maxabs;
    m.scales = nullptr;
    m.stash = nullptr;
    launch_grad_maxabs(m, s);
    q.maxabs = r.own_maxabs;
  }
  // Every block adds its per-leaf totals to its XCD's copy at its end, and atomics on one address
  // take turns: the copies keep those chains an eighth of the blocks long.
  const dim3 grid(grad_blocks(r.g.n, kRenewBlocks)), block(256);
  const int obj = r.g.source == 0 ? r.g.objective : r.g.source == 2 ? kObjSoftmax : -1;
  const bool large = r.leaf_bytes != 1 && r.n_ids > kRenewLds;  // two-byte ids
  if (large && obj == kObjSoftmax) {
    hipLaunchKernelGGL((k_renew_sum<kObjSoftmax, kRenewLdsLarge>), grid, block, 0, s, q);
  } else if (large && obj == 0) {
    hipLaunchKernelGGL((k_renew_sum<0, kRenewLdsLarge>), grid, block, 0, s, q);
  } else {
    switch (obj) {  // one variant per objective
      case kObjSoftmax: hipLaunchKernelGGL((k_renew_sum<kObjSoftmax, kRenewLds>), grid, block, 0, s, q); break;
      case 0: hipLaunchKernelGGL((k_renew_sum<0, kRenewLds>), grid, block, 0, s, q); break;
      case 1: hipLaunchKernelGGL((k_renew_sum<1, kRenewLds>), grid, block, 0, s, q); break;
      case 2: hipLaunchKernelGGL((k_renew_sum<2, kRenewLds>), grid, block, 0, s, q); break;
      case 3: hipLaunchKernelGGL((k_renew_sum<3, kRenewLds>), grid, block, 0, s, q); break;
      case 4: hipLaunchKernelGGL((k_renew_sum<4, kRenewLds>), grid, block, 0, s, q); break;
      default: hipLaunchKernelGGL((k_renew_sum<-1, kRenewLds>), grid, block, 0, s, q);
    }
  }
  hipLaunchKernelGGL(k_renew_leaf, dim3((r.n_ids + 255) / 256), dim3(256), 0, s, q);
}

// ========================================================== near-tie resolution ======
constexpr uint32_t kResolveCells = 260;  // cells of one contender feature in LDS (one-byte codes)
constexpr uint32_t kResolveSlabPairs = 4096;  // (row, feature) pairs of a slab at least
constexpr uint32_t kResolveSlabRows = 512;    // rows of a slab at least (amortise the slab's flush)

// The level's plan, computed by every block of both resolution kernels from the ambiguity
// flags of the level's scan (which these kernels never write, so every block plans the same):
// the ambiguous nodes in node order, at most r_max of them, and their row segments cut into
// slabs of about the same number of (row, contender feature) pairs, at most s_max + r_max
// slabs. Returns the number of resolved nodes.
struct ResolvePlan {
  uint32_t rank[512 + 1], wave[1024 / 64 + 1];
  uint32_t node[kResolveMax], begin[kResolveMax], count[kResolveMax], slab0[kResolveMax + 1];
  uint32_t rows[kResolveMax];  // rows of each slab of the node (its last slab: the rest)
  uint32_t per[kResolveMax];   // the node's contender features (1 for none): pairs per row
};

__device__ uint32_t resolve_plan(const ResolveLaunch& a, ResolvePlan& L) {
  const uint32_t t = threadIdx.x;
  const uint32_t bound = min(a.n_nodes, 512u);
  const uint32_t n = a.n_nodes_dyn ? min(*a.n_nodes_dyn, bound) : bound;
  for (uint32_t i = t; i < n; i += blockDim.x) L.rank[i] = a.cand[i].ambiguous ? 1u : 0u;
  __syncthreads();
  const uint32_t flagged = plan_prefix(L.rank, n, L.wave);
  const uint32_t nr = min(flagged, min(a.r_max, kResolveMax));
  if (nr == 0) return 0;  // the same for every block
  for (uint32_t i = t; i < n; i += blockDim.x) {
    const uint32_t r = L.rank[i];
    const bool amb = (i + 1 < n ? L.rank[i + 1] : flagged) != r;
    if (!amb || r >= nr) continue;
    L.node[r] = i;
    L.begin[r] = a.layer ? a.layer[i].begin : 0u;
    L.count[r] = a.layer ? a.layer[i].count : a.rows_total;
    uint32_t nf = 0;
    for (uint32_t w = 0; w < a.mask_words; ++w) nf += __popc(a.contend[size_t(i) * a.mask_words + w]);
    L.per[r] = max(nf, 1u);
  }
  __syncthreads();
  if (t == 0) {
    uint64_t total = 0;
    for (uint32_t r = 0; r < nr; ++r) total += uint64_t(L.count[r]) * L.per[r];
    const uint64_t pairs = max(uint64_t(kResolveSlabPairs), (total + a.s_max - 1) / max(1u, a.s_max));
    L.slab0[0] = 0;
    for (uint32_t r = 0; r < nr; ++r) {
      // At most 2^26 rows, so a slab's pairs in one feature group stay below 2^32.
      const uint64_t rows = min(max(uint64_t(kResolveSlabRows), (pairs + L.per[r] - 1) / L.per[r]), uint64_t(1) << 26);
      L.rows[r] = uint32_t(rows);
      L.slab0[r + 1] = L.slab0[r] + uint32_t((uint64_t(L.count[r]) + rows - 1) / rows);
    }
    if (blockIdx.x == 0 && a.res_stats) {
      atomicAdd(&a.res_stats[0], static_cast<unsigned long long>(flagged));
      atomicAdd(&a.res_stats[1], static_cast<unsigned long long>(nr));
    }
  }
  __syncthreads();
  return nr;
}

// Contender features of level node i in ascending order (thread 0); returns their count.
__device__ uint32_t resolve_contenders(const ResolveLaunch& a, uint32_t i, uint32_t* list, uint32_t* s_count) {
  if (threadIdx.x == 0) {
    uint32_t k = 0;
    for (uint32_t w = 0; w < a.mask_words; ++w)
      for (uint32_t m = a.contend[size_t(i) * a.mask_words + w]; m; m &= m - 1u) list[k++] = w * 32 + (__ffs(m) - 1u);
    *s_count = k;
  }
  __syncthreads();
  return *s_count;
}

constexpr uint32_t kResolveRowsPer = 8;  // rows per thread per step of the resolution's histograms

// FP64 gradient pairs of rows row[u] (in[u]; the others get zeros), their inputs loaded
// together: the max pass's stash (source 2), the write pass's row-order copy (regression rows
// with gstash), given pairs (source 1) or the regression inputs (source 0). A row the gradient
// pass rejected (its error is raised there) counts as zero here.
template <uint32_t U>
__device__ inline void resolve_grads(const GradLaunch& g, const uint32_t (&row)[U], const bool (&in)[U],
                                     double (&gv)[U], double (&hv)[U]) {
  uint32_t bad[U];
  if (g.source == 0 && g.gstash) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (g.unit_h) {
        gv[u] = g.gstash[row[u]];
        hv[u] = 1.0;
      } else {
        const double2 qv = reinterpret_cast<const double2*>(g.gstash)[row[u]];
        gv[u] = qv.x;
        hv[u] = qv.y;
      }
      bad[u] = 0u;
    }
  } else if (g.source == 2) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const double2 qv = reinterpret_cast<const double2*>(g.stash)[row[u]];
      gv[u] = qv.x;
      hv[u] = qv.y;
      bad[u] = 0u;
    }
  } else if (g.source == 1) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      gv[u] = g.pairs[2 * size_t(row[u])];
      hv[u] = g.pairs[2 * size_t(row[u]) + 1];
      bad[u] = isfinite(gv[u]) && isfinite(hv[u]) && hv[u] >= 0.0 ? 0u : kErrNonFinite;
    }
  } else {
    uint64_t idx[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) idx[u] = min<uint64_t>(row[u], g.n - 1);
    RowIn v[U];
    load_rows_at<U>(g, idx, v);
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) bad[u] = grad_of<-1>(g, v[u], false, &gv[u], &hv[u]);
  }
#pragma unroll
  for (uint32_t u = 0; u < U; ++u)
    if (!in[u] || bad[u]) gv[u] = hv[u] = 0.0;
}

// Near-tie resolution, first launch: each slab's block adds its rows' FP64 gradients, in the
// FP64_FIXED form's fixed point (multiplier 2^S, S = 62 - exponent(rows * max |g|), as its
// gradient pass writes them), to its node's exact histograms of the contender features (LDS
// cells for a group of features at a time, then integer atomics: exact and order-independent)
// and to the node's totals. The work items are (row, feature) pairs, a row's pairs on adjacent
// lanes: the lanes of one row load its row id and gradient inputs from the same addresses, and
// its codes from one 128-byte row line of the row-major codes when the node's rows are sparse
// and it has several contenders (one line per feature from the column-major copy otherwise).
template <bool UH>
__global__ __launch_bounds__(256) void k_resolve_hist(ResolveLaunch a) {
  constexpr uint32_t FG = UH ? kResolveFeaturesUH : kResolveFeatures;  // features per LDS group
  __shared__ unsigned long long s_g[FG][kResolveCells], s_h[UH ? 1 : FG][kResolveCells];
  __shared__ uint32_t s_c[FG][kResolveCells];
  __shared__ uint32_t s_nc[FG];
  __shared__ unsigned long long s_tg, s_th;
  __shared__ ResolvePlan L;
  __shared__ uint32_t s_flist[1024], s_nf;
  const uint32_t t = threadIdx.x;
  const uint32_t nr = resolve_plan(a, L);
  if (nr == 0) return;
  const uint32_t ns = L.slab0[nr];
  const GradLaunch& g = a.g;
  const double mg = ldexp(1.0, renew_shift(a.maxabs[0], g.n)), mh = ldexp(1.0, renew_shift(a.maxabs[1], g.n));
  uint32_t listed = 0xFFFFFFFFu;  // the node whose contender list is in s_flist
  for (uint32_t sl = blockIdx.x; sl < ns; sl += gridDim.x) {
    uint32_t r = 0;
    while (L.slab0[r + 1] <= sl) ++r;
    const uint64_t lo64 = uint64_t(L.begin[r]) + uint64_t(sl - L.slab0[r]) * L.rows[r];
    const uint32_t lo = uint32_t(lo64);
    const uint32_t hi = uint32_t(min<uint64_t>(uint64_t(L.begin[r]) + L.count[r], lo64 + L.rows[r]));
    if (listed != r) {
      resolve_contenders(a, L.node[r], s_flist, &s_nf);
      listed = r;
    }
    const uint32_t nf = s_nf;
    // The first feature group also sums the totals (a node with no contender: totals only).
    for (uint32_t g0 = 0; g0 < max(nf, 1u); g0 += FG) {
      const uint32_t ng = min(FG, nf - min(nf, g0));
      const uint32_t per = max(ng, 1u);  // pairs per row
      // Lines read per row: one row line, or one column segment per feature (one column line
      // per 64 rows at most when the node's rows are dense).
      const bool row_major =
          uint64_t(ng) * min<uint64_t>(64ull * L.count[r], a.rows_total) > 128ull * L.count[r];
      if (t < ng) s_nc[t] = a.cell_offset[s_flist[g0 + t] + 1] - a.cell_offset[s_flist[g0 + t]];
      for (uint32_t k = t; k < FG * kResolveCells; k += blockDim.x) {
        (&s_g[0][0])[k] = 0ull;
        if (!UH) (&s_h[0][0])[k] = 0ull;
        (&s_c[0][0])[k] = 0u;
      }
      if (t == 0) s_tg = s_th = 0ull;
      __syncthreads();
      unsigned long long tg = 0ull, th = 0ull;
      // kResolveRowsPer pairs per thread per step: their row ids, then their gradient inputs
      // and codes are loaded together (a wave keeps those loads in flight at once).
      const uint64_t pairs = uint64_t(hi - lo) * per;
      for (uint64_t j0 = t; j0 < pairs; j0 += uint64_t(blockDim.x) * kResolveRowsPer) {
        uint32_t row[kResolveRowsPer], k[kResolveRowsPer];
        bool in[kResolveRowsPer];
#pragma unroll
        for (uint32_t u = 0; u < kResolveRowsPer; ++u) {
          const uint64_t j = j0 + uint64_t(u) * blockDim.x;
          in[u] = j < pairs;
          const uint32_t jr = in[u] ? uint32_t(j) : 0u;  // below 2^26 rows * FG
          const uint32_t q = jr / per;
          k[u] = jr - q * per;
          row[u] = in[u] ? (a.rows ? a.rows[lo + q] : lo + q) : 0u;
        }
        double gv[kResolveRowsPer], hv[kResolveRowsPer];
        resolve_grads<kResolveRowsPer>(g, row, in, gv, hv);
        uint32_t code[kResolveRowsPer];
#pragma unroll
        for (uint32_t u = 0; u < kResolveRowsPer; ++u) {
          const uint32_t f = s_flist[g0 + min(k[u], max(ng, 1u) - 1u)];
          if (ng == 0) {
            code[u] = 0u;
          } else if (row_major) {
            code[u] = load_code(a.codes, row[u], a.row_stride, a.code_bytes, f);
          } else {
            const size_t at = size_t(f) * a.rows_total + row[u];
            code[u] = a.code_bytes == 1 ? uint32_t(a.codes_cm[at])
                                        : uint32_t(reinterpret_cast<const uint16_t*>(a.codes_cm)[at]);
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < kResolveRowsPer; ++u) {
          if (!in[u]) continue;
          const long long qg = static_cast<long long>(rint(gv[u] * mg));
          const long long qh = UH ? 0ll : static_cast<long long>(rint(hv[u] * mh));
          if (k[u] == 0) {  // each row once
            tg += static_cast<unsigned long long>(qg);
            th += static_cast<unsigned long long>(qh);
          }
          if (k[u] >= ng) continue;
          if (s_nc[k[u]] <= kResolveCells) {
            atomicAdd(&s_g[k[u]][code[u]], static_cast<unsigned long long>(qg));
            if (!UH) atomicAdd(&s_h[UH ? 0 : k[u]][code[u]], static_cast<unsigned long long>(qh));
            atomicAdd(&s_c[k[u]][code[u]], 1u);
          } else {
            const size_t cell = size_t(r) * a.cells + a.cell_offset[s_flist[g0 + k[u]]] + code[u];
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.res_hist.g[cell]), static_cast<unsigned long long>(qg));
            if (!UH)
              atomicAdd(reinterpret_cast<unsigned long long*>(&a.res_hist.h[cell]), static_cast<unsigned long long>(qh));
            atomicAdd(&a.res_hist.cnt[cell], 1u);
          }
        }
      }
      if (g0 == 0) {
        atomicAdd(&s_tg, tg);
        if (!UH) atomicAdd(&s_th, th);
      }
      __syncthreads();
      for (uint32_t k = 0; k < ng; ++k) {
        if (s_nc[k] > kResolveCells) continue;
        const size_t base = size_t(r) * a.cells + a.cell_offset[s_flist[g0 + k]];
        for (uint32_t c = t; c < s_nc[k]; c += blockDim.x) {
          if (!s_c[k][c]) continue;
          atomicAdd(reinterpret_cast<unsigned long long*>(&a.res_hist.g[base + c]), s_g[k][c]);
          if (!UH) atomicAdd(reinterpret_cast<unsigned long long*>(&a.res_hist.h[base + c]), s_h[UH ? 0 : k][c]);
          atomicAdd(&a.res_hist.cnt[base + c], s_c[k][c]);
        }
      }
      if (g0 == 0 && t == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.res_parent[r].gq), s_tg);
        if (!UH) atomicAdd(reinterpret_cast<unsigned long long*>(&a.res_parent[r].hq), s_th);
      }
      __syncthreads();  // the next group or slab reuses the LDS
    }
  }
}

// Near-tie resolution, second launch, one block per resolved node r (level node i): the exact
// scan of its contender features (one wavefront per feature), the best of them under the
// reference order, the node's candidate replaced (child statistics summed from the level's
// integer histogram for the chosen split, so the next level's sums stay in one arithmetic),
// and the node's exact cells and totals cleared for the next level. (Spreading a node's scans
// over several blocks would need a device-scope fence per block to hand their candidates to
// the last one; on this GPU such a fence writes back the XCD's L2, which costs more than the
// scans, above all beside other fits.)
template <bool UH>
__global__ __launch_bounds__(512) void k_resolve_node(ResolveLaunch a) {
  __shared__ ResolvePlan L;
  __shared__ uint32_t s_flist[1024], s_nf, s_best;
  __shared__ NodeStats s_par;
  const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u, waves = blockDim.x >> 6;
  a.res_stats = nullptr;  // counted by the first launch
  const uint32_t nr = resolve_plan(a, L);
  const uint32_t r = blockIdx.x;
  if (r >= nr) return;
  const uint32_t i = L.node[r];
  const uint32_t nf = resolve_contenders(a, i, s_flist, &s_nf);
  const GradLaunch& g = a.g;
  if (t == 0) s_par = NodeStats{L.count[r], 0.0, 0.0, a.res_parent[r].gq, a.res_parent[r].hq};
  __syncthreads();
  ScanParams q = a.p;
  q.scales = nullptr;
  q.scale_g = ldexp(1.0, -renew_shift(a.maxabs[0], g.n));
  q.scale_h = UH ? 1.0 : ldexp(1.0, -renew_shift(a.maxabs[1], g.n));
  q.integer = 1u;
  q.unit_h = UH ? 1u : 0u;
  q.noise_k = 0.0;
  q.ref_view = 0u;
  ScanNode nd = a.scan[i];
  nd.slot = r;
  nd.parent_from_hist = 2;
  Candidate* pf = a.res_pf + size_t(r) * a.features;
  for (uint32_t k = wave; k < nf; k += waves)
    scan_feature_wave<true, UH>(a.res_hist, a.cells, a.cell_offset, a.bin_count, a.miss_cell, s_flist[k], nd, &s_par,
                                q, a.mono, true, &pf[s_flist[k]], a.err);
  __syncthreads();
  if (wave == 0) {
    // The best contender under the reference order: each lane's share, then the wavefront's
    // (the candidates' features differ, so the order is total and any reduction order agrees).
    uint32_t bi = 0xFFFFFFFFu, bv = 0, bf = 0, bt = 0, bd = 0;
    double bg = 0.0;
    for (uint32_t k = lane; k < nf; k += 64) {
      const Candidate& x = pf[s_flist[k]];
      if (better(x.valid, x.gain, x.feature, x.threshold, x.default_left, bv, bg, bf, bt, bd)) {
        bv = 1;
        bg = x.gain;
        bf = x.feature;
        bt = x.threshold;
        bd = x.default_left;
        bi = s_flist[k];
      }
    }
    for (int d = 1; d < 64; d <<= 1) {
      const double og = __shfl_xor(bg, d, 64);
      const uint32_t ov = __shfl_xor(bv, d, 64), of = __shfl_xor(bf, d, 64), ot = __shfl_xor(bt, d, 64),
                     od = __shfl_xor(bd, d, 64), oi = __shfl_xor(bi, d, 64);
      if (better(ov, og, of, ot, od, bv, bg, bf, bt, bd)) {
        bv = ov;
        bg = og;
        bf = of;
        bt = ot;
        bd = od;
        bi = oi;
      }
    }
    if (lane == 0) {
      atomicAdd(&a.res_stats_features[0], static_cast<unsigned long long>(nf));
      s_best = bi;
    }
  }
  __syncthreads();
  if (wave == 0) {
    Candidate& c = a.cand[i];
    if (s_best == 0xFFFFFFFFu) {
      if (lane == 0) c.valid = 0u;
    } else {
      const Candidate e = pf[s_best];
      const ScanParams p = with_scales(a.p);
      const size_t base = size_t(a.scan[i].slot) * a.cells;
      const uint32_t f = e.feature, off = a.cell_offset[f], mc = a.miss_cell[f];
      Acc3 acc{0, 0.0, 0.0, 0, 0};
      for (uint32_t b = lane; b <= e.threshold; b += 64) acc = add3(acc, cell_value(a.hist, base + off + b, p), p);
      Acc3 left = wave_sum3(acc, p);
      if (e.default_left && mc != 0xFFFFFFFFu) left = add3(left, cell_value(a.hist, base + mc, p), p);
      const Acc3 P{uint32_t(c.parent.count), c.parent.g, c.parent.h, c.parent.gq, c.parent.hq};
      const Acc3 right = sub3(P, left, p);
      if (lane == 0) {
        c.valid = 1u;
        c.feature = f;
        c.threshold = e.threshold;
        c.default_left = e.default_left;
        c.gain = e.gain;
        c.left = to_stats(left, p);
        c.right = to_stats(right, p);
        c.l_lo = e.l_lo;
        c.l_hi = e.l_hi;
        c.r_lo = e.r_lo;
        c.r_hi = e.r_hi;
      }
    }
  }
  // Clear what this node used: its contender cells and totals.
  const size_t base = size_t(r) * a.cells;
  for (uint32_t k = 0; k < nf; ++k) {
    const uint32_t f = s_flist[k];
    for (uint32_t c = a.cell_offset[f] + t; c < a.cell_offset[f + 1]; c += blockDim.x) {
      a.res_hist.g[base + c] = 0ull;
      a.res_hist.h[base + c] = 0ull;
      a.res_hist.cnt[base + c] = 0u;
    }
  }
  if (t == 0) {
    a.res_parent[r].gq = 0;
    a.res_parent[r].hq = 0;
  }
}

void launch_resolve(const ResolveLaunch& a, hipStream_t s) {
  if (a.n_nodes == 0 || a.n_nodes > 512 || a.features > 1024 || a.r_max > kResolveMax)
    fail(MPSB_ERR_INTERNAL, "near-tie resolution launched outside its bounds");
  // Grids for the level's bounds: at most min(r_max, nodes) resolved nodes, and slabs of at
  // least kResolveSlabRows rows each (one per node at most more).
  const uint32_t nodes = std::min(a.r_max, a.n_nodes);
  const uint64_t row_slabs = (uint64_t(a.g.n) + kResolveSlabRows - 1) / kResolveSlabRows;
  const dim3 grid(uint32_t(std::min<uint64_t>(std::min(1024u, a.s_max) , row_slabs) + nodes)), block(256);
  if (a.g.unit_h) {
    hipLaunchKernelGGL(k_resolve_hist<true>, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_resolve_node<true>, dim3(nodes), dim3(512), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_resolve_hist<false>, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_resolve_node<false>, dim3(nodes), dim3(512), 0, s, a);
  }
}

// ====================================================================== utility ======
__global__ void k_iota(uint32_t* rows, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) rows[i] = i;
}
void launch_iota(uint32_t* rows, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(256), 0, s, rows, n);
}
__global__ void k_iota_mod(uint32_t* rows, uint64_t n, uint32_t m) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    rows[i] = uint32_t(i % m);
}
void launch_iota_mod(uint32_t* rows, uint64_t n, uint32_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_iota_mod, dim3(grid_for(n)), dim3(256), 0, s, rows, n, m);
}

__global__ void k_vector_add(const float* a, const float* b, float* out, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = a[i] + b[i];
}
void launch_vector_add(const float* a, const float* b, float* out, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_vector_add, dim3(grid_for(n)), dim3(256), 0, s, a, b, out, n);
}

__global__ void k_fill(double* p, double v, uint64_t n, uint32_t stride, uint32_t offset) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i * stride + offset] = v;
}
void launch_fill_f64(double* p, double v, uint64_t n, uint32_t stride, uint32_t offset, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, p, v, n, stride, offset);
}

// The reference checks every prediction / margin after each tree's update
// (src/core/trainer.cpp:256-261, src/core/multiclass_trainer.cpp:92-98); the gradient pass of
// the next tree does it for all trees but the last, this pass for the last.
__global__ void k_final_update(double* pred, uint64_t n, uint32_t stride, uint32_t offset, const uint16_t* leaf_node,
                               uint32_t leaf_bytes, const double* leaf_table, double lr, uint32_t* err) {
  bool bad = false;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    double* p = pred + i * stride + offset;
    double v = *p;
    if (leaf_node) {
      const uint32_t id = leaf_bytes == 1 ? reinterpret_cast<const uint8_t*>(leaf_node)[i] : leaf_node[i];
      v = v + lr * leaf_table[id];
      *p = v;
    }
    bad = bad || !isfinite(v);
  }
  if (bad) raise_err(err, kErrPrediction);
}

void launch_final_update(double* pred, uint64_t n, uint32_t stride, uint32_t offset, const uint16_t* leaf_node,
                         uint32_t leaf_bytes, const double* leaf_table, double lr, uint32_t* err, hipStream_t s) {
  hipLaunchKernelGGL(k_final_update, dim3(grid_for(n)), dim3(256), 0, s, pred, n, stride, offset, leaf_node,
                     leaf_bytes, leaf_table, lr, err);
}

// ===================================================================== predict =======
// RegressionTree::Predict (src/core/tree.cpp:179-214) for every tree, accumulated per row in
// tree order like RegressionModel::Predict (src/core/trainer.cpp:268-282).
// A block stages its rows' codes in LDS once (coalesced 4-byte loads; the row pitch is padded
// by one word so lanes reading one feature of different rows hit different banks), then each
// thread walks every tree for its row from LDS. The margin of a row is accumulated in a
// register in tree order, the reference's order of additions.
__global__ __launch_bounds__(256) void k_predict(PredictLaunch a) {
  extern __shared__ uint32_t srow[];
  const uint32_t row_words = a.row_stride * a.code_bytes / 4;
  const uint32_t pitch = row_words + 1;
  const uint64_t r0 = uint64_t(blockIdx.x) * a.rows_per_block;
  if (r0 >= a.rows) return;
  const uint32_t nrows = uint32_t(min<uint64_t>(a.rows_per_block, a.rows - r0));
  const uint32_t* src = reinterpret_cast<const uint32_t*>(a.codes) + r0 * row_words;
  for (uint32_t i = threadIdx.x; i < nrows * row_words; i += blockDim.x)
    srow[(i / row_words) * pitch + i % row_words] = src[i];
  __syncthreads();
  for (uint32_t rr = threadIdx.x; rr < nrows; rr += blockDim.x) {
    const uint8_t* rb = reinterpret_cast<const uint8_t*>(srow + size_t(rr) * pitch);
    const uint64_t r = r0 + rr;
    for (uint32_t c = 0; c < a.classes; ++c) {
      double acc = a.out[r * a.classes + c];
      for (uint32_t t = c; t < a.n_trees; t += a.classes) {
        const uint32_t first = a.tree_first[t], n = a.tree_first[t + 1] - first;
        uint32_t i = 0, seen = 0;
        double leaf = 0.0;
        while (true) {
          if (i >= n || ++seen > n) {
            raise_err(a.err, kErrPredict);
            break;
          }
          const mpsb_tree_node nd = a.nodes[first + i];
          if (nd.flags & 1u) {
            leaf = nd.leaf_value;
            break;
          }
          const uint32_t f = nd.feature_index;
          const uint32_t code = a.code_bytes == 1 ? uint32_t(rb[f]) : uint32_t(reinterpret_cast<const uint16_t*>(rb)[f]);
          const bool left = code == a.miss_code[f] ? nd.default_left != 0 : code <= nd.threshold_bin;
          i = left ? nd.left_child : nd.right_child;
        }
        acc += a.lr * leaf;
      }
      a.out[r * a.classes + c] = acc;
    }
  }
}
void launch_predict(const PredictLaunch& a, hipStream_t s) {
  const uint32_t row_words = a.row_stride * a.code_bytes / 4;
  const size_t lds = size_t(a.rows_per_block) * (row_words + 1) * 4;
  const uint64_t blocks = (a.rows + a.rows_per_block - 1) / a.rows_per_block;
  hipLaunchKernelGGL(k_predict, dim3(uint32_t(blocks)), dim3(256), lds, s, a);
}

}  // namespace mpsb

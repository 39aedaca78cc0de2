// Stable LSD radix sort of particles by cell id (8-bit digits).
//
// Binning (locate_particles, mesh.tcc:1097-1153) gives each particle a cell;
// the scatter kernels need the particles grouped by cell with ascending
// particle ids inside each cell. Sorting the pairs (cell, id) with a stable
// sort over the id-ordered input gives exactly that. Each pass is
//   hist:    per 4096-key tile, a 256-bin digit histogram (LDS atomics),
//   scan:    exclusive scan of the digit-major histogram table,
//   scatter: per tile, stable ranks from wave64 ballots (8 one-bit ballots
//            form each lane's digit-match mask), written at the scanned
//            offsets.
#include <hip/hip_runtime.h>

#include "mpmx_kernels.h"

namespace mpmx {

static constexpr int kSortBlock = 256;
static constexpr int kRounds = 16;
static constexpr int kTile = kSortBlock * kRounds;
static constexpr int kWaves = kSortBlock / 64;
static constexpr int kScanTile = 1024;

static inline int64_t tiles_for(int64_t n) { return (n + kTile - 1) / kTile; }

int64_t radix_sort_temp_ints(int64_t n) {
  const int64_t nb = tiles_for(n) > 0 ? tiles_for(n) : 1;
  const int64_t nt = (256 * nb + kScanTile - 1) / kScanTile;
  return 4 * n + 256 * nb + nt + 64;
}

// keys: cell in [0, ncells) else max_key; values: particle id
__global__ __launch_bounds__(kSortBlock) void k_rs_init(const int* cell, int64_t n, int max_key,
                                                        int* keys, int* vals) {
  const int64_t i = blockIdx.x * (int64_t)kSortBlock + threadIdx.x;
  if (i >= n) return;
  const int c = cell[i];
  keys[i] = (c >= 0 && c < max_key) ? c : max_key;
  vals[i] = (int)i;
}

__global__ __launch_bounds__(kSortBlock) void k_rs_hist(const int* keys, int64_t n, int shift,
                                                        int* hist, int64_t nblocks) {
  __shared__ int bins[256];
  bins[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = blockIdx.x * (int64_t)kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int64_t i = base + r * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&bins[(keys[i] >> shift) & 255], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = bins[threadIdx.x];
}

// Exclusive scan of `total` ints in place: each 1024-element tile is scanned
// by one block (4 consecutive elements per thread, wave prefix sums by
// shuffles), tile totals are scanned by one block, then added back.

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// block-wide exclusive scan of one value per thread (256 threads); returns
// the exclusive prefix and writes the block total to *total
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_incl_scan(v);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    before += (w < wave) ? wsum[w] : 0;
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

__global__ __launch_bounds__(256) void k_rs_scan_tiles(int* a, int64_t total, int* tile_sums) {
  const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * 4;
  int v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = (base + j < total) ? a[base + j] : 0;
  const int s = (v[0] + v[1]) + (v[2] + v[3]);
  int tot;
  int run = block_excl_scan(s, &tot);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (base + j < total) a[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// exclusive scan of the tile sums (any count) by one block
__global__ __launch_bounds__(256) void k_rs_scan_sums(int* sums, int64_t n) {
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b = 0; b < n; b += 256) {
    const int64_t i = b + threadIdx.x;
    const int v = i < n ? sums[i] : 0;
    int tot;
    const int ex = block_excl_scan(v, &tot);
    if (i < n) sums[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_rs_scan_add(int* a, int64_t total, const int* sums) {
  const int add = sums[blockIdx.x];
  const int64_t base = blockIdx.x * (int64_t)kScanTile;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t i = base + j * 256 + threadIdx.x;
    if (i < total) a[i] += add;
  }
}

__global__ __launch_bounds__(kSortBlock) void k_rs_scatter(const int* kin, const int* vin,
                                                           int* kout, int* vout, int64_t n,
                                                           int shift, const int* hist,
                                                           int64_t nblocks) {
  __shared__ int base[256];
  __shared__ int wcount[kWaves][256];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  base[tid] = hist[tid * nblocks + blockIdx.x];
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int64_t tbase = blockIdx.x * (int64_t)kTile;
  for (int r = 0; r < kRounds; ++r) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const int64_t i = tbase + r * kSortBlock + tid;
    const bool valid = i < n;
    int key = 0, val = 0, digit = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      digit = (key >> shift) & 255;
    }
    unsigned long long mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int bit = (digit >> b) & 1;
      const unsigned long long bal = __ballot(bit);
      mask &= bit ? bal : ~bal;
    }
    const int rank = __popcll(mask & lt);
    if (valid && rank == 0) wcount[wave][digit] = __popcll(mask);
    __syncthreads();
    if (valid) {
      int off = base[digit] + rank;
      for (int w = 0; w < wave; ++w) off += wcount[w][digit];
      kout[off] = key;
      vout[off] = val;
    }
    __syncthreads();
    int tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += wcount[w][tid];
    base[tid] += tot;
    __syncthreads();
  }
}

// Keys of an earlier sort's permutation with the current cells: keys[i] =
// cell[perm[i]] (max_key for an invalid cell). Particles that changed cell
// since that sort sit out of key order and form short runs of their own.
__global__ __launch_bounds__(kSortBlock) void k_rs_refresh(const int* cell, const int* perm,
                                                           int64_t n, int max_key, int* keys) {
  const int64_t i = blockIdx.x * (int64_t)kSortBlock + threadIdx.x;
  if (i >= n) return;
  const int c = cell[perm[i]];
  keys[i] = (c >= 0 && c < max_key) ? c : max_key;
}

void refresh_sorted_keys(const int* cell, const int* perm, int64_t n, int max_key, int* keys,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rs_refresh, dim3((unsigned)((n + kSortBlock - 1) / kSortBlock)),
                     dim3(kSortBlock), 0, st, cell, perm, n, max_key, keys);
}

void radix_sort_by_cell(const int* cell, int64_t n, int max_key, int* keys_out, int* vals_out,
                        int* tmp, hipStream_t st) {
  if (n <= 0) return;
  const int64_t nb = tiles_for(n);
  int* k0 = tmp;
  int* v0 = tmp + n;
  int* k1 = tmp + 2 * n;
  int* v1 = tmp + 3 * n;
  int* hist = tmp + 4 * n;
  int* tile_sums = hist + 256 * nb;
  const int64_t ntiles = (256 * nb + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_rs_init, dim3((unsigned)((n + kSortBlock - 1) / kSortBlock)),
                     dim3(kSortBlock), 0, st, cell, n, max_key, k0, v0);
  int bits = 1;
  while (bits < 31 && (1ll << bits) <= (long long)max_key) ++bits;
  const int passes = (bits + 7) / 8;
  const int* sk = k0;
  const int* sv = v0;
  for (int pass = 0; pass < passes; ++pass) {
    int* dk;
    int* dv;
    if (pass == passes - 1) {
      dk = keys_out;
      dv = vals_out;
    } else if (pass % 2 == 0) {
      dk = k1;
      dv = v1;
    } else {
      dk = k0;
      dv = v0;
    }
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nb), dim3(kSortBlock), 0, st, sk, n, shift, hist,
                       nb);
    hipLaunchKernelGGL(k_rs_scan_tiles, dim3((unsigned)ntiles), dim3(256), 0, st, hist, 256 * nb,
                       tile_sums);
    hipLaunchKernelGGL(k_rs_scan_sums, dim3(1), dim3(256), 0, st, tile_sums, ntiles);
    hipLaunchKernelGGL(k_rs_scan_add, dim3((unsigned)ntiles), dim3(256), 0, st, hist, 256 * nb,
                       tile_sums);
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nb), dim3(kSortBlock), 0, st, sk, sv, dk, dv,
                       n, shift, hist, nb);
    sk = dk;
    sv = dv;
  }
}

}  // namespace mpmx

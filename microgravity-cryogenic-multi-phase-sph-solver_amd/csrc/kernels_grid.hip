// Spatial hash on gfx950: cell keys, stable LSD radix sort by cell, cell
// offset table, fused gather of the persistent arrays, neighbour statistics.
//
// Replaces SpatialHash::build / reorder_particles / neighbor_stats
// (spatial_hash.cu:48-171).  The reference CPU path sorts with a serial stable
// counting sort; any stable sort by the same key gives the same permutation,
// so this LSD radix sort (8-bit digits, stable per pass) is bit-exact with it.
#include "kernels.hpp"

namespace cryo {

namespace {

constexpr int kBlock = 256;

__device__ inline int gtid() { return static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); }

// ---------------------------------------------------- kick-drift + cell keys
// sph_solver.cu:20-34 (kick_drift) fused with the keys kernel of
// spatial_hash.cu:61-68: the key uses the drifted position.
__global__ __launch_bounds__(kBlock) void k_kick_drift_keys(PV pv, GridDev g, DevState* st, Count nc, int do_kick,
                                                           int continuity, int tvf, uint32_t* keys,
                                                           uint32_t* vals) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  V3 xi = ld3(pv.s.x[i]);
  if (do_kick && is_fluid(ph)) {
    const float dt = st->dt;
    const float half = 0.5f * dt;
    const V3 v = ld3(pv.s.v[i]) + ld3(pv.t.a[i]) * half;
    pv.s.v[i] = st3(v);
    const V3 vt = tvf ? v + ld3(pv.t.ab[i]) * half : v;
    pv.s.vt[i] = st3(vt);
    const V3 dx = vt * dt;
    if (pv.s.xlo) {  // precision form: the displacement into the double-float position
      V3 lo = ld3(pv.s.xlo[i]);
      df_add(xi.x, lo.x, dx.x);
      df_add(xi.y, lo.y, dx.y);
      df_add(xi.z, lo.z, dx.z);
      pv.s.xlo[i] = st3(lo);
    } else {
      xi += dx;
    }
    pv.s.x[i] = st3(xi);
    if (continuity) pv.s.rho[i] += pv.t.drho[i] * dt;
  }
  keys[i] = ph == kDead ? static_cast<uint32_t>(g.num_cells) : static_cast<uint32_t>(cell_index(g, xi));
  vals[i] = static_cast<uint32_t>(i);
}

// ------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;

__global__ __launch_bounds__(kRsThreads) void k_radix_hist(const uint32_t* keys, Count nc, int shift,
                                                          uint32_t* tile_hist, int num_tiles) {
  const int n = nc.get();
  __shared__ uint32_t h[256];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRsTile;
  for (int k = 0; k < kRsItems; ++k) {
    const int i = base + k * kRsThreads + tid;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  tile_hist[tid * num_tiles + blockIdx.x] = h[tid];
}

// In-place exclusive scan of `total` counters by one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_one_block(uint32_t* data, int total, uint32_t* total_out) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int chunk = (total + 1023) / 1024;
  const int lo = tid * chunk;
  const int hi = lo + chunk < total ? lo + chunk : total;
  uint32_t s = 0;
  for (int k = lo; k < hi; ++k) s += data[k];
  part[tid] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 chunk sums.
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = tid > 0 ? part[tid - 1] : 0u;
  for (int k = lo; k < hi; ++k) {
    const uint32_t v = data[k];
    data[k] = run;
    run += v;
  }
  if (tid == 1023 && total_out) *total_out = part[1023];
}

// Exclusive scan of one digit's counts over all tiles (row d of the
// digit-major histogram), in place; the row total goes to row_total[d].
__global__ __launch_bounds__(256) void k_radix_scan_rows(uint32_t* tile_hist, int num_tiles, uint32_t* row_total) {
  __shared__ uint32_t part[256];
  uint32_t* row = tile_hist + static_cast<long long>(blockIdx.x) * num_tiles;
  const int tid = threadIdx.x;
  const int chunk = (num_tiles + 255) / 256;
  const int lo = tid * chunk;
  const int hi = lo + chunk < num_tiles ? lo + chunk : num_tiles;
  uint32_t s = 0;
  for (int k = lo; k < hi; ++k) s += row[k];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = tid > 0 ? part[tid - 1] : 0u;
  for (int k = lo; k < hi; ++k) {
    const uint32_t v = row[k];
    row[k] = run;
    run += v;
  }
  if (tid == 255) row_total[blockIdx.x] = part[255];
}

// Stable scatter: ranks inside a tile follow element order (round by round,
// wave by wave, lane by lane), tiles follow tile order through tile_off.
__global__ __launch_bounds__(kRsThreads) void k_radix_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                             uint32_t* keys_out, uint32_t* vals_out, Count nc,
                                                             int shift, const uint32_t* tile_off, int num_tiles,
                                                             const uint32_t* row_total) {
  const int n = nc.get();
  __shared__ uint32_t running[256];
  __shared__ uint32_t digit_base[256];
  __shared__ uint32_t wave_cnt[kRsThreads / 64][256];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  // exclusive scan of the 256 digit totals: where each digit's run starts
  digit_base[tid] = row_total[tid];
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t v = tid >= off ? digit_base[tid - off] : 0u;
    __syncthreads();
    digit_base[tid] += v;
    __syncthreads();
  }
  const uint32_t incl = digit_base[tid];
  __syncthreads();
  digit_base[tid] = incl - row_total[tid];
  running[tid] = 0;
  for (int w = 0; w < kRsThreads / 64; ++w) wave_cnt[w][tid] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRsTile;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int k = 0; k < kRsItems; ++k) {
    const int i = base + k * kRsThreads + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    // Lanes of this wave holding the same digit.
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    if (!valid) peers = 0;
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt_mask));
    const bool leader = valid && (peers & lt_mask) == 0;
    if (leader) wave_cnt[wave][d] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wave_cnt[w][d];
      const uint32_t dest = digit_base[d] + tile_off[d * num_tiles + blockIdx.x] + pos;
      keys_out[dest] = key;
      vals_out[dest] = vals_in[i];
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kRsThreads / 64; ++w) {
      add += wave_cnt[w][tid];
      wave_cnt[w][tid] = 0;
    }
    running[tid] += add;
    __syncthreads();
  }
}

// off[c] = first sorted index with key >= c, for c in [0, num_cells]:
// cell c holds [off[c], off[c+1]); off[num_cells] = live particle count.
__global__ __launch_bounds__(kBlock) void k_cell_offsets(const uint32_t* keys, Count nc, int num_cells, int* off,
                                                        DevState* st) {
  const int n = nc.get();
  const int c = gtid();
  if (c > num_cells) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < static_cast<uint32_t>(c)) lo = mid + 1;
    else hi = mid;
  }
  off[c] = lo;
  if (c == num_cells) {
    st->n_alive = lo;
    st->n = lo;
    st->merges_sorted = st->merges;
  }
}

// new[i] = old[perm[i]] for the 11 persistent arrays in one pass
// (spatial_hash.cu:117-138 uses 11 separate kernels).
__global__ __launch_bounds__(kBlock) void k_gather(Persist dst, Persist src, const uint32_t* perm,
                                                  const DevState* st) {
  const int i = gtid();
  if (i >= st->n_alive) return;
  const uint32_t j = perm[i];
  dst.x[i] = src.x[j];
  dst.v[i] = src.v[j];
  dst.vt[i] = src.vt[j];
  dst.nrm[i] = src.nrm[j];
  dst.m[i] = src.m[j];
  dst.m0[i] = src.m0[j];
  dst.rho[i] = src.rho[j];
  dst.T[i] = src.T[j];
  dst.phase[i] = src.phase[j];
  dst.flags[i] = src.flags[j];
  dst.id[i] = src.id[j];
  if (dst.xlo) dst.xlo[i] = src.xlo[j];
}

// spatial_hash.cu:140-171 (count and id-sum within radius, self included).
__global__ __launch_bounds__(kBlock) void k_neighbor_stats(const float4* x, const uint32_t* id, GridDev g, int n,
                                                          float r2max, int* counts, unsigned long long* sums) {
  const int i = gtid();
  if (i >= n) return;
  const V3 xi = ld3(x[i]);
  int c = 0;
  unsigned long long s = 0;
  for_each_nbr(g, xi, [&](int j) {
    if (norm2(xi - ld3(x[j])) < r2max) {
      ++c;
      s += id[j];
    }
  });
  counts[i] = c;
  sums[i] = s;
}

// CSR fill: offsets from a host scan of the counts.
__global__ __launch_bounds__(kBlock) void k_neighbor_fill(const float4* x, GridDev g, int n, float r2max,
                                                         const long long* offsets, int* indices) {
  const int i = gtid();
  if (i >= n) return;
  const V3 xi = ld3(x[i]);
  long long k = offsets[i];
  for_each_nbr(g, xi, [&](int j) {
    if (norm2(xi - ld3(x[j])) < r2max) indices[k++] = j;
  });
}

inline int blocks_for(long long n, int b = kBlock) { return static_cast<int>((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------ host wrappers
void launch_kick_drift_keys(const PV& pv, const GridDev& g, DevState* st, Count n, bool do_kick, bool continuity,
                            bool tvf, uint32_t* keys, uint32_t* vals, hipStream_t s) {
  if (n.ub <= 0) return;
  k_kick_drift_keys<<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, st, n, do_kick ? 1 : 0, continuity ? 1 : 0,
                                                     tvf ? 1 : 0, keys, vals);
}

int radix_tiles(int n) { return (n + kRsTile - 1) / kRsTile; }

void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, Count n,
                       int key_bits, uint32_t* tile_hist, uint32_t* row_total, hipStream_t s, bool* result_in_alt) {
  *result_in_alt = false;
  if (n.ub <= 0) return;
  const int tiles = radix_tiles(n.ub);
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  for (int shift = 0; shift < key_bits; shift += 8) {
    k_radix_hist<<<tiles, kRsThreads, 0, s>>>(ki, n, shift, tile_hist, tiles);
    k_radix_scan_rows<<<256, 256, 0, s>>>(tile_hist, tiles, row_total);
    k_radix_scatter<<<tiles, kRsThreads, 0, s>>>(ki, vi, ko, vo, n, shift, tile_hist, tiles, row_total);
    uint32_t* t = ki;
    ki = ko;
    ko = t;
    t = vi;
    vi = vo;
    vo = t;
    *result_in_alt = !*result_in_alt;
  }
}

// Per-cell kinds of the particles each cell holds (kCellWall | kCellFluid).
__global__ __launch_bounds__(kBlock) void k_cell_flags(const int* off, const uint8_t* phase, int num_cells,
                                                      uint8_t* flags) {
  const int c = gtid();
  if (c >= num_cells) return;
  int f = 0;
  for (int j = off[c], e = off[c + 1]; j < e; ++j) {
    const uint8_t ph = phase[j];
    f |= ph == kWall ? kCellWall : (ph == kLiquid ? kCellLiquid : (ph == kVapor ? kCellVapor : 0));
  }
  flags[c] = static_cast<uint8_t>(f);
}

void launch_cell_flags(const int* off, const uint8_t* phase, int num_cells, uint8_t* flags, hipStream_t s) {
  if (num_cells <= 0) return;
  k_cell_flags<<<blocks_for(num_cells), kBlock, 0, s>>>(off, phase, num_cells, flags);
}

void launch_cell_offsets(const uint32_t* keys, Count n, int num_cells, int* off, DevState* st, hipStream_t s) {
  k_cell_offsets<<<blocks_for(static_cast<long long>(num_cells) + 1), kBlock, 0, s>>>(keys, n, num_cells, off, st);
}

void launch_gather(const Persist& dst, const Persist& src, const uint32_t* perm, const DevState* st, int n_ub,
                   hipStream_t s) {
  if (n_ub <= 0) return;
  k_gather<<<blocks_for(n_ub), kBlock, 0, s>>>(dst, src, perm, st);
}

void launch_neighbor_stats(const float4* x, const uint32_t* id, const GridDev& g, int n, float r2max, int* counts,
                           unsigned long long* sums, hipStream_t s) {
  if (n <= 0) return;
  k_neighbor_stats<<<blocks_for(n), kBlock, 0, s>>>(x, id, g, n, r2max, counts, sums);
}

void launch_neighbor_fill(const float4* x, const GridDev& g, int n, float r2max, const long long* offsets,
                          int* indices, hipStream_t s) {
  if (n <= 0) return;
  k_neighbor_fill<<<blocks_for(n), kBlock, 0, s>>>(x, g, n, r2max, offsets, indices);
}

void launch_scan_one_block(uint32_t* data, int total, uint32_t* total_out, hipStream_t s) {
  k_scan_one_block<<<1, 1024, 0, s>>>(data, total, total_out);
}

}  // namespace cryo

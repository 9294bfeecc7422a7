// kernels_graph.hip — scans, radix sort, CSR and coloring for gfx950.
//
//  exclusive_scan       device-wide int scan in one launch (decoupled
//                       look-back over 1024-int tiles)
//  radix_sort_pairs     LSD radix sort, 4-bit digits, stable ranks from
//                       64-lane ballots (used for Morton codes and colors);
//                       two launches per pass (histogram, scatter)
//  tcsr_*               body -> contact-group lists in color-position order,
//                       each body's groups ascending as a serial scatter
//                       leaves them (gpu_csr_count/starts/scatter,
//                       kernels_6dof.py:2627-2692)
//  count_conflicts      gpu_count_color_conflicts (kernels_6dof.py:3313-3336),
//                       counted per emitting pair instead of per row
//  adj_*                gpu_body_adj_count/scatter (kernels_6dof.py:3045-3096)
//  k_color_greedy       the serial effect of gpu_color_spec_assign
//                       (kernels_6dof.py:3233-3262, greedy in body order) or
//                       of JP_ROUNDS gpu_color_round passes (:3099-3148,
//                       greedy in priority order, round index
//                       r(i) = max_blockers r(j) + [j > i]) computed as a
//                       dataflow: bodies are taken in processing order from a
//                       ticket counter and each waits only on bodies earlier
//                       in that order, so every wait ends.
//  k_color_resolve      gpu_color_spec_resolve (kernels_6dof.py:3265-3296)
#include "avbd_internal.h"

namespace avbd {

static size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
constexpr int kScanPer = 16;                // ints per thread
constexpr int kScanTile = 256 * kScanPer;   // ints per tile
constexpr int kScanMaxBlocks = 256;         // workgroups of one scan launch

__device__ __forceinline__ int wave_incl_scan(int v) {
  int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

// Block-wide exclusive scan of one int per thread (256 threads); returns the
// exclusive prefix and writes the block total to *total.
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[4];
  int lane = __lane_id(), wid = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int pre = 0, tot = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wid) pre += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return pre + inc - v;
}

// Single-pass scan with decoupled look-back: tile b (4096 ints) publishes
// its total, then its inclusive prefix once known, in one 64-bit status word
// {stamp << 1 | kind, value} (kind 1: inclusive prefix, 0: tile total). To
// find its carry, wavefront 0 of tile b reads the 64 status words below it
// at once, waits until all of them carry this call's stamp, adds the totals
// down to the nearest inclusive prefix, and steps down another 64 tiles
// when there is none. Tile 0 publishes its prefix at once, so every look-back
// ends. A workgroup takes its tile index from a ticket counter, so a tile
// waits only on tiles whose workgroups have already started, whatever order
// the dispatcher starts workgroups in; a spin cap still turns a wait that
// does not end into a flagged error instead of a hang.
//
// Workspace: hdr[0] the stamp of the last call, hdr[1] the count of tiles
// done in this call, hdr[2] the tile ticket (the last tile to finish resets
// both and advances hdr[0]), then the status words. A fresh workspace must
// be zero. One launch replaces the
// three of a reduce-then-scan, so two kernel boundaries (~10 us each on this
// GPU) are saved per scan.
constexpr int kScanSpinCap = 1 << 20;

__device__ __forceinline__ unsigned long long ld_status(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long* p, unsigned stamp, int kind,
                                          int value) {
  __hip_atomic_store(p, ((unsigned long long)((stamp << 1) | (unsigned)kind) << 32) |
                            (unsigned long long)(unsigned)value,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(256)
k_scan_1p(const int* in, int* out, int n, unsigned* hdr, unsigned long long* status, int* total,
          int* err, const int* nd) {
  __shared__ unsigned s_stamp;
  __shared__ int s_carry, s_tile;
  // with a device count, only the first ceil(n / tile) tickets scan
  n = dev_count(n, nd);
  const int nblk = gridDim.x;
  const int nwork = (n + kScanTile - 1) / kScanTile;
  // hdr[0] changes only after every workgroup of this call is done, so all
  // read the same value; stamps run 1 .. 2^31 - 1 (a zero workspace holds 0)
  if (threadIdx.x == 0) {
    unsigned st = (__hip_atomic_load(&hdr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u) &
                  0x7fffffffu;
    s_stamp = st == 0u ? 1u : st;
  }
  __syncthreads();
  const unsigned stamp = s_stamp;
  // Workgroups take tiles from the ticket counter until the tickets run past
  // the last tile: tiles are numbered in the order they are taken, and a
  // tile's predecessors are all held by running workgroups.
  for (;;) {
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(&hdr[2], 1u);
    __syncthreads();
    const int b = s_tile;
    __syncthreads();
    if (b >= nwork) {
      if (b == 0 && threadIdx.x == 0 && total) *total = 0;  // nothing to scan
      break;
    }
    const int base = b * kScanTile + threadIdx.x * kScanPer;
    int v[kScanPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      v[k] = (base + k < n) ? in[base + k] : 0;
      sum += v[k];
    }
    int tot;
    int pre = block_excl_scan(sum, &tot);
    if (threadIdx.x == 0) st_status(&status[b], stamp, b == 0 ? 1 : 0, tot);
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      int carry = 0;
      for (int p = b - 1; p >= 0; p -= 64) {
        const int idx = p - lane;
        unsigned long long w = 0ull;
        bool ready = idx < 0;
        int spins = 0;
        // the loop exit is uniform over the wavefront
        while (__builtin_amdgcn_ballot_w64(!ready) != 0) {
          if (!ready) {
            w = ld_status(&status[idx]);
            ready = (unsigned)(w >> 33) == stamp;
          }
          if (++spins > kScanSpinCap && !ready) {
            if (err) atomicExch(err, 1);
            ready = true;
          }
        }
        const int val = idx >= 0 ? (int)(unsigned)w : 0;
        const unsigned long long pm =
            __builtin_amdgcn_ballot_w64(idx >= 0 && ((w >> 32) & 1ull) != 0ull);
        if (pm != 0ull) {
          const int l0 = __builtin_ctzll(pm);
          carry += wave_sum(lane <= l0 ? val : 0);
          break;
        }
        carry += wave_sum(val);
      }
      if (lane == 0) {
        s_carry = carry;
        if (b > 0) st_status(&status[b], stamp, 1, carry + tot);
        if (b == nwork - 1 && total) *total = carry + tot;
      }
    }
    __syncthreads();
    pre += s_carry;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      if (base + k < n) out[base + k] = pre;
      pre += v[k];
    }
  }
  if (threadIdx.x == 0) {
    const unsigned done = atomicAdd(&hdr[1], 1u);
    if (done == (unsigned)(nblk - 1)) {
      // every workgroup of this call has drawn its last ticket
      __hip_atomic_store(&hdr[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&hdr[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&hdr[0], stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

size_t scan_workspace_bytes(int n) {
  return a256(16 + ((size_t)n / kScanTile + 2) * sizeof(unsigned long long));
}

void exclusive_scan(hipStream_t st, const int* in, int* out, int n, int* total_dev, void* ws,
                    int* err, const int* nd) {
  if (n <= 0) {
    if (total_dev) launch_fill_int(st, total_dev, 1, 0);
    return;
  }
  const int nblk = std::min((n + kScanTile - 1) / kScanTile, kScanMaxBlocks);
  unsigned* hdr = (unsigned*)ws;
  unsigned long long* status = (unsigned long long*)((char*)ws + 16);
  hipLaunchKernelGGL(k_scan_1p, dim3(nblk), dim3(256), 0, st, in, out, n, hdr, status, total_dev,
                     err, nd);
}

// ---------------------------------------------------------------------------
// LSD radix sort (keys unsigned, values int), 4-bit digits
// ---------------------------------------------------------------------------
constexpr int kRadixTile = 1024;

__global__ void __launch_bounds__(256)
k_radix_hist(const unsigned* __restrict__ keys, int n, int shift, int nblk,
             int* __restrict__ hist, const int* __restrict__ nd) {
  __shared__ int cnt[16];
  n = dev_count(n, nd);
  if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kRadixTile;
  for (int r = 0; r < 4; ++r) {
    int idx = base + r * 256 + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 15u], 1);
  }
  __syncthreads();
  if (threadIdx.x < 16) hist[threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// Each tile finds its 16 output offsets from the digit histogram itself
// (hist[d * nblk + b], digit-major): offset(d, b) = keys of smaller digits
// in all tiles + keys of digit d in tiles before b. Reading the whole
// histogram per tile costs 16 x nblk loads, a few KB, and saves a scan
// launch per pass.
__global__ void __launch_bounds__(256)
k_radix_scatter(const unsigned* __restrict__ kin, const int* __restrict__ vin, int n, int shift,
                int nblk, const int* __restrict__ hist, unsigned* __restrict__ kout,
                int* __restrict__ vout, const int* __restrict__ nd) {
  n = dev_count(n, nd);
  __shared__ int running[16];
  __shared__ int wcnt[4][16];
  __shared__ int part[2][16];  // per digit: keys in tiles before this one, keys in all tiles
  int lane = __lane_id(), wid = threadIdx.x >> 6;
  {
    // wavefront w sums digits 4w .. 4w+3 over the tiles
    const int d0 = wid * 4;
    int before[4], all[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) before[j] = all[j] = 0;
    for (int bb = lane; bb < nblk; bb += 64) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = hist[(d0 + j) * nblk + bb];
        all[j] += c;
        before[j] += bb < (int)blockIdx.x ? c : 0;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int sb = wave_sum(before[j]), sa = wave_sum(all[j]);
      if (lane == 0) {
        part[0][d0 + j] = sb;
        part[1][d0 + j] = sa;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int smaller = 0;  // keys of digits below d, all tiles
      for (int d = 0; d < 16; ++d) {
        running[d] = smaller + part[0][d];
        smaller += part[1][d];
      }
    }
  }
  __syncthreads();
  unsigned long long lt = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
  int base = blockIdx.x * kRadixTile;
  for (int r = 0; r < 4; ++r) {
    int idx = base + r * 256 + threadIdx.x;
    bool valid = idx < n;
    unsigned key = valid ? kin[idx] : 0u;
    int val = valid ? vin[idx] : 0;
    int digit = valid ? (int)((key >> shift) & 15u) : 16;
    int rank = 0;
#pragma unroll
    for (int d = 0; d < 16; ++d) {
      unsigned long long m = __ballot(digit == d);
      if (digit == d) rank = __popcll(m & lt);
      if (lane == 0) wcnt[wid][d] = __popcll(m);
    }
    __syncthreads();
    if (valid) {
      int pre = 0;
      for (int w = 0; w < wid; ++w) pre += wcnt[w][digit];
      int pos = running[digit] + pre + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
      int d = threadIdx.x;
      running[d] += wcnt[0][d] + wcnt[1][d] + wcnt[2][d] + wcnt[3][d];
    }
    __syncthreads();
  }
}

size_t radix_workspace_bytes(int n) {
  int nblk = (n + kRadixTile - 1) / kRadixTile;
  size_t h = (size_t)16 * nblk;
  return a256(h * sizeof(int)) + 256;
}

static void radix_sort_bits(hipStream_t st, unsigned* keys, int* vals, unsigned* keys_tmp,
                            int* vals_tmp, int n, int n_passes, void* ws,
                            const int* nd = nullptr) {
  if (n <= 1) return;
  int nblk = (n + kRadixTile - 1) / kRadixTile;
  int* hist = (int*)ws;  // 16 x nblk, digit-major
  unsigned* ka = keys;
  int* va = vals;
  unsigned* kb = keys_tmp;
  int* vb = vals_tmp;
  for (int pass = 0; pass < n_passes; ++pass) {
    int shift = 4 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(256), 0, st, ka, n, shift, nblk, hist, nd);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(256), 0, st, ka, va, n, shift, nblk,
                       hist, kb, vb, nd);
    unsigned* tk = ka; ka = kb; kb = tk;
    int* tv = va; va = vb; vb = tv;
  }
  if (ka != keys) {
    // errors surface at the caller's hipGetLastError / synchronisation
    (void)hipMemcpyAsync(keys, ka, n * sizeof(unsigned), hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(vals, va, n * sizeof(int), hipMemcpyDeviceToDevice, st);
  }
}

void radix_sort_pairs(hipStream_t st, unsigned* keys, int* vals, unsigned* keys_tmp,
                      int* vals_tmp, int n, void* ws) {
  radix_sort_bits(st, keys, vals, keys_tmp, vals_tmp, n, 8, ws);
}

void radix_sort_small(hipStream_t st, unsigned* keys, int* vals, unsigned* keys_tmp,
                      int* vals_tmp, int n, int key_bits, void* ws, const int* nd) {
  radix_sort_bits(st, keys, vals, keys_tmp, vals_tmp, n, (key_bits + 3) / 4, ws, nd);
}

// ---------------------------------------------------------------------------
// body -> group CSR in color-position order
// ---------------------------------------------------------------------------
// Position k of the color bucketing (bodies sorted by color, ascending within
// a color) owns entries [cpos[k], cpos[k+1]) of `csr`: its static groups in
// row order, then its dynamic groups in ascending group index, then padding
// {-1, -1} up to a multiple of four. Entry = {group, other body} (the body
// itself for floor and pin groups). Dynamic groups come from the broadphase
// pairs, which are in (a asc, b asc) order with their groups numbered in
// pair order: body i's groups as the pair's b (pairs (x, i), x < i) all
// precede its groups as the pair's a (a contiguous pair range), so the
// ascending group order is "as-b pairs in pair order, then the as-a range".
// This is the order the reference's CSR scatter produces when its row loop
// runs serially (gpu_csr_count / gpu_csr_scatter, kernels_6dof.py:2627-2692).

// Per kept pair: as-a range boundaries, as-b pair and group counts.
__global__ void __launch_bounds__(256)
k_tcsr_pairs(int kept, const int* __restrict__ pa, const int* __restrict__ pb,
             const int* __restrict__ succ, int* __restrict__ asb_cnt,
             int* __restrict__ asb_groups, int* __restrict__ start_a, int* __restrict__ end_a,
             const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  kept = dev_count(kept, nd);
  if (p >= kept) return;
  const int a = pa[p], b = pb[p];
  if (p == 0 || pa[p - 1] != a) start_a[a] = p;
  if (p == kept - 1 || pa[p + 1] != a) end_a[a] = p + 1;
  const int n = succ[p];
  if (n > 0) {
    atomicAdd(&asb_cnt[b], 1);
    atomicAdd(&asb_groups[b], n);
  }
}

__global__ void __launch_bounds__(256)
k_tcsr_asb_fill(int kept, const int* __restrict__ pb, const int* __restrict__ succ,
                const int* __restrict__ asb_start, int* __restrict__ asb_cur,
                int* __restrict__ asb_list, const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(kept, nd) || succ[p] <= 0) return;
  const int b = pb[p];
  asb_list[asb_start[b] + atomicAdd(&asb_cur[b], 1)] = p;
}

__device__ __forceinline__ int as_a_groups(int i, const int* start_a, const int* end_a,
                                           const int* sscan, const int* succ) {
  const int s = start_a[i], e = end_a[i];
  return e > s ? (sscan[e - 1] + succ[e - 1]) - sscan[s] : 0;
}

__global__ void __launch_bounds__(256)
k_tcsr_count(int nb, const int* __restrict__ cvals, const int* __restrict__ st_count,
             const int* __restrict__ asb_groups, const int* __restrict__ start_a,
             const int* __restrict__ end_a, const int* __restrict__ sscan,
             const int* __restrict__ succ, int* __restrict__ padc, int* __restrict__ pos_of) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int i = cvals[k];
  pos_of[i] = k;
  const int cnt = st_count[i] + asb_groups[i] + as_a_groups(i, start_a, end_a, sscan, succ);
  padc[k] = (cnt + 3) & ~3;
}

constexpr int kAsbLds = 16;

__global__ void __launch_bounds__(256)
k_tcsr_fill(int nb, const int* __restrict__ cvals, const int* __restrict__ st_start,
            const int* __restrict__ st_idx, const int* __restrict__ asb_start,
            int* __restrict__ asb_list, const int* __restrict__ start_a,
            const int* __restrict__ end_a, const int* __restrict__ sscan,
            const int* __restrict__ succ, const int* __restrict__ pa, const int* __restrict__ pb,
            int g_static, const int* __restrict__ cpos, int2* __restrict__ csr) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int i = cvals[k];
  int o = cpos[k];
  const int end = cpos[k + 1];
  for (int q = st_start[i]; q < st_start[i + 1]; ++q) csr[o++] = make_int2(st_idx[q], i);
  // as-b pairs: a handful per body, taken in ascending pair index. Up to
  // kAsbLds of them are sorted in an LDS column of this lane (a sort in
  // global memory is a chain of dependent loads); longer lists in place.
  __shared__ int srt[kAsbLds][256];
  const int l0 = asb_start[i], l1 = asb_start[i + 1], nl = l1 - l0;
  const int t = threadIdx.x;
  if (nl <= kAsbLds) {
    for (int u = 0; u < nl; ++u) srt[u][t] = asb_list[l0 + u];
    for (int u = 1; u < nl; ++u) {
      const int v = srt[u][t];
      int w = u - 1;
      while (w >= 0 && srt[w][t] > v) { srt[w + 1][t] = srt[w][t]; --w; }
      srt[w + 1][t] = v;
    }
  } else {
    for (int u = l0 + 1; u < l1; ++u) {
      const int v = asb_list[u];
      int w = u - 1;
      while (w >= l0 && asb_list[w] > v) { asb_list[w + 1] = asb_list[w]; --w; }
      asb_list[w + 1] = v;
    }
  }
  for (int u = 0; u < nl; ++u) {
    const int p = nl <= kAsbLds ? srt[u][t] : asb_list[l0 + u];
    const int g0 = g_static + sscan[p], n = succ[p], other = pa[p];
    for (int c = 0; c < n; ++c) csr[o++] = make_int2(g0 + c, other);
  }
  for (int p = start_a[i]; p < end_a[i]; ++p) {
    const int g0 = g_static + sscan[p], n = succ[p], other = pb[p];
    for (int c = 0; c < n; ++c) csr[o++] = make_int2(g0 + c, other);
  }
  while (o < end) csr[o++] = make_int2(-1, -1);
}

void launch_tcsr_pairs(hipStream_t st, int kept, const int* pa, const int* pb, const int* succ,
                       int* asb_cnt, int* asb_groups, int* start_a, int* end_a, const int* nd) {
  if (kept > 0)
    hipLaunchKernelGGL(k_tcsr_pairs, dim3((kept + 255) / 256), dim3(256), 0, st, kept, pa, pb, succ,
                       asb_cnt, asb_groups, start_a, end_a, nd);
}

void launch_tcsr_asb_fill(hipStream_t st, int kept, const int* pb, const int* succ,
                          const int* asb_start, int* asb_cur, int* asb_list, const int* nd) {
  if (kept > 0)
    hipLaunchKernelGGL(k_tcsr_asb_fill, dim3((kept + 255) / 256), dim3(256), 0, st, kept, pb, succ,
                       asb_start, asb_cur, asb_list, nd);
}

void launch_tcsr_count(hipStream_t st, int nb, const int* cvals, const int* st_count,
                       const int* asb_groups, const int* start_a, const int* end_a,
                       const int* sscan, const int* succ, int* padc, int* pos_of) {
  if (nb > 0)
    hipLaunchKernelGGL(k_tcsr_count, dim3((nb + 255) / 256), dim3(256), 0, st, nb, cvals, st_count,
                       asb_groups, start_a, end_a, sscan, succ, padc, pos_of);
}

// pos_map[k] = CSR position of the body now at color position k.
__global__ void __launch_bounds__(256)
k_tcsr_posmap(int nb, const int* __restrict__ cvals, const int* __restrict__ pos_of,
              int* __restrict__ pos_map) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nb) pos_map[k] = pos_of[cvals[k]];
}

void launch_tcsr_posmap(hipStream_t st, int nb, const int* cvals, const int* pos_of, int* pos_map) {
  if (nb > 0)
    hipLaunchKernelGGL(k_tcsr_posmap, dim3((nb + 255) / 256), dim3(256), 0, st, nb, cvals, pos_of,
                       pos_map);
}

void launch_tcsr_fill(hipStream_t st, int nb, const int* cvals, const int* st_start,
                      const int* st_idx, const int* asb_start, int* asb_list, const int* start_a,
                      const int* end_a, const int* sscan, const int* succ, const int* pa,
                      const int* pb, int g_static, const int* cpos, int2* csr) {
  if (nb > 0)
    hipLaunchKernelGGL(k_tcsr_fill, dim3((nb + 255) / 256), dim3(256), 0, st, nb, cvals, st_start,
                       st_idx, asb_start, asb_list, start_a, end_a, sscan, succ, pa, pb, g_static,
                       cpos, csr);
}

// ---------------------------------------------------------------------------
// coloring
// ---------------------------------------------------------------------------
// Grid-stride over the pairs with at most kConflictBlocks workgroups and one
// atomic per workgroup: the count is a single address, and same-address
// atomics serialise.
constexpr int kConflictBlocks = 240;

__global__ void __launch_bounds__(256)
k_count_conflicts(int n_pairs, const int* __restrict__ pa, const int* __restrict__ pb,
                  const int* __restrict__ succ, const int* __restrict__ rpc,
                  const int* __restrict__ color, int* out, const int* __restrict__ scene,
                  int* __restrict__ scene_dirty, const int* __restrict__ nd) {
  __shared__ int wsum[4];
  n_pairs = dev_count(n_pairs, nd);
  int c = 0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    const int ns = succ[p];
    if (ns > 0) {
      const int a = pa[p], b = pb[p];
      if (color[a] == color[b]) {
        c += ns * rpc[p];
        if (scene_dirty) scene_dirty[scene[a]] = 1;  // batched: the scene recolors
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if (__lane_id() == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    if (t) atomicAdd(out, t);
  }
}

void launch_count_conflicts(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                            const int* succ, const int* rpc, const int* color, int* out,
                            const int* scene, int* scene_dirty, const int* nd) {
  if (n_pairs > 0)
    hipLaunchKernelGGL(k_count_conflicts,
                       dim3(std::min((n_pairs + 255) / 256, kConflictBlocks)), dim3(256), 0, st,
                       n_pairs, pa, pb, succ, rpc, color, out, scene, scene_dirty, nd);
}

// Start of a recolor: every body uncolored, or (batched scenes) only the
// bodies of scenes marked dirty; the others keep their colors.
__global__ void __launch_bounds__(256)
k_color_reset(int nb, int* __restrict__ color, const int* __restrict__ scene,
              const int* __restrict__ scene_dirty, const int* __restrict__ nd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_count(nb, nd)) return;
  if (scene_dirty == nullptr || scene_dirty[scene[i]]) color[i] = -1;
}

void launch_color_reset(hipStream_t st, int nb, int* color, const int* scene,
                        const int* scene_dirty, const int* nd) {
  if (nb > 0)
    hipLaunchKernelGGL(k_color_reset, dim3((nb + 255) / 256), dim3(256), 0, st, nb, color, scene,
                       scene_dirty, nd);
}

__global__ void __launch_bounds__(256)
k_adj_count(int n_dyn, int g_first, const GroupGeo* __restrict__ geo, int* counts) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_dyn) return;
  int4 hd = geo[g_first + e].hdr;
  if (hd.y < 0 || hd.z < 0 || hd.y == hd.z) return;
  atomicAdd(&counts[hd.y], 1);
  atomicAdd(&counts[hd.z], 1);
}

void launch_adj_count(hipStream_t st, int n_dyn, int g_first, const GroupGeo* geo, int* counts) {
  if (n_dyn > 0)
    hipLaunchKernelGGL(k_adj_count, dim3((n_dyn + 255) / 256), dim3(256), 0, st, n_dyn, g_first,
                       geo, counts);
}

__global__ void __launch_bounds__(256)
k_adj_fill(int n_dyn, int g_first, const GroupGeo* __restrict__ geo, const int* __restrict__ starts,
           int* cursor, int* nbr) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_dyn) return;
  int4 hd = geo[g_first + e].hdr;
  int a = hd.y, b = hd.z;
  if (a < 0 || b < 0 || a == b) return;
  nbr[starts[a] + atomicAdd(&cursor[a], 1)] = b;
  nbr[starts[b] + atomicAdd(&cursor[b], 1)] = a;
}

void launch_adj_fill(hipStream_t st, int n_dyn, int g_first, const GroupGeo* geo,
                     const int* starts, int* cursor, int* nbr) {
  if (n_dyn > 0)
    hipLaunchKernelGGL(k_adj_fill, dim3((n_dyn + 255) / 256), dim3(256), 0, st, n_dyn, g_first, geo,
                       starts, cursor, nbr);
}

__device__ __forceinline__ int lowest_free_color(unsigned used) {
  unsigned fr = ~used;
  return fr ? (__ffs((int)fr) - 1) : (kMaxColors - 1);
}

__device__ __forceinline__ int ld_relaxed(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Body adjacency of the coloring: one entry per broadphase pair that emitted
// at least one contact group (the same neighbour sets the dynamic groups
// describe, without the per-contact-point duplicates).
__global__ void __launch_bounds__(256)
k_adj_count_pairs(int n_pairs, const int* __restrict__ pa, const int* __restrict__ pb,
                  const int* __restrict__ succ, int* counts, const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(n_pairs, nd) || succ[p] <= 0) return;
  int a = pa[p], b = pb[p];
  if (a < 0 || b < 0 || a == b) return;
  atomicAdd(&counts[a], 1);
  atomicAdd(&counts[b], 1);
}

__global__ void __launch_bounds__(256)
k_adj_fill_pairs(int n_pairs, const int* __restrict__ pa, const int* __restrict__ pb,
                 const int* __restrict__ succ, const int* __restrict__ starts, int* cursor,
                 int* nbr, const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(n_pairs, nd) || succ[p] <= 0) return;
  int a = pa[p], b = pb[p];
  if (a < 0 || b < 0 || a == b) return;
  nbr[starts[a] + atomicAdd(&cursor[a], 1)] = b;
  nbr[starts[b] + atomicAdd(&cursor[b], 1)] = a;
}

void launch_adj_count_pairs(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                            const int* succ, int* counts, const int* nd) {
  if (n_pairs > 0)
    hipLaunchKernelGGL(k_adj_count_pairs, dim3((n_pairs + 255) / 256), dim3(256), 0, st, n_pairs,
                       pa, pb, succ, counts, nd);
}

void launch_adj_fill_pairs(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                           const int* succ, const int* starts, int* cursor, int* nbr,
                           const int* nd) {
  if (n_pairs > 0)
    hipLaunchKernelGGL(k_adj_fill_pairs, dim3((n_pairs + 255) / 256), dim3(256), 0, st, n_pairs,
                       pa, pb, succ, starts, cursor, nbr, nd);
}

// Greedy coloring in a fixed processing order, computed as a dataflow: the
// body at position k of `order` takes the lowest color not used by its
// neighbours that come earlier in the order, which is what a serial loop over
// the order produces (gpu_color_spec_assign executed serially,
// kernels_6dof.py:3203-3243; Jones-Plassmann round numbers,
// solver_6dof.py:770-806). One workgroup of W wavefronts takes 64 W
// consecutive positions per ticket; tickets are handed out in order, so every
// position a workgroup waits on belongs to a workgroup that already holds a
// ticket.
//
// pub[i]: -1 until body i is resolved. Jacobi: the chosen color. JP:
// (min(round, 24) << 8) | (color + 1), color = -1 when round >= 24.
// pos: position of each body in `order` (JP); null means the identity order.
// Each lane collects up to kPollBatch earlier neighbours in LDS and polls
// all of them at once with relaxed loads (the published value is the only
// data exchanged), so one dependency step across tickets costs one memory
// round trip. A neighbour of the same ticket (e.g. the previous body of a
// row) is read from the workgroup's copy of the published values in LDS
// instead, and after every poll the lanes of a wavefront that wait only on
// such neighbours resolve among themselves from LDS until none makes
// progress, so a chain inside one ticket costs LDS accesses.
constexpr int kPollBatch = 16;
constexpr int kSpinCap = 1 << 20;

template <int W>
__global__ void __launch_bounds__(64 * W)
k_color_greedy(int nb, const int* __restrict__ order, const int* __restrict__ pos,
               const int* __restrict__ nbr_start, const int* __restrict__ nbr,
               const int* __restrict__ color_in, int* pub, int* ticket, int mode_jp, int* err,
               int* clamped, const int* __restrict__ scene, const int* __restrict__ scene_dirty,
               const int* __restrict__ nd) {
  nb = dev_count(nb, nd);
  // kT = 64 W positions per ticket. nj_s: a neighbour j to poll, or ~slot
  // when j is this ticket's body of that slot; lpub: this ticket's
  // published values, by slot (accessed with workgroup-scope relaxed
  // atomics: a volatile LDS access would be emitted as a flat access, which
  // serialises the poll)
  constexpr int kT = 64 * W;
  __shared__ int nj_s[kPollBatch][kT];
  __shared__ int lpub[kT];
  __shared__ int s_base;
  auto lpub_ld = [&](int l) {
    return __hip_atomic_load(&lpub[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  auto lpub_st = [&](int l, int v) {
    __hip_atomic_store(&lpub[l], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const int lane = threadIdx.x;  // slot in the ticket
  for (;;) {
    if (lane == 0) s_base = atomicAdd(ticket, kT);
    lpub_st(lane, -1);
    // the ticket and the cleared lpub are seen by every wavefront before any
    // of them reads them
    __syncthreads();
    const int base = s_base;
    if (base >= nb) return;
    const int k = base + lane;
    const int i = k < nb ? order[k] : -1;
    // a body of a clean scene (batched) keeps its color and has no neighbour
    // outside its scene, so nobody waits on it
    bool done = !(i >= 0 && (mode_jp || color_in[i] == -1)) ||
                (i >= 0 && scene_dirty != nullptr && !scene_dirty[scene[i]]);
    int my_pos = 0, q = 0, q1 = 0;
    if (!done) {
      my_pos = pos ? pos[i] : i;
      q = nbr_start[i];
      q1 = nbr_start[i + 1];
    }
    // pend: batch entries not yet resolved; loc: entries read from lpub;
    // gt (JP): entries whose body index exceeds i (their round counts +1)
    unsigned used = 0u, pend = 0u, loc = 0u, gt = 0u;
    int rmax = 0, spins = 0;
    // fold a resolved neighbour's published value into used / rmax
    auto absorb = [&](int t, int vt) {
      pend &= ~(1u << t);
      if (!mode_jp) {
        if (vt < kMaxColors) used |= 1u << vt;
      } else {
        const int cj = (vt & 255) - 1;
        const int rj = (vt >> 8) + (int)(gt >> t & 1u);
        if (cj >= 0) used |= 1u << cj;
        rmax = rj > rmax ? rj : rmax;
      }
    };
    auto publish = [&]() {
      // all 32 colors taken (jacobi) or past the round limit (JP): the
      // result may leave same-color neighbours
      if (used == 0xffffffffu || (mode_jp && rmax >= kJpRounds)) atomicOr(clamped, 1);
      int val;
      if (!mode_jp) {
        val = lowest_free_color(used);
      } else {
        const int r = rmax < kJpRounds ? rmax : kJpRounds;
        const int c = rmax < kJpRounds ? lowest_free_color(used) : -1;
        val = (r << 8) | (c + 1);
      }
      st_relaxed(&pub[i], val);
      lpub_st(lane, val);
      done = true;
    };
    // Per-lane state machine: collect a batch of earlier neighbours, poll it
    // until resolved, repeat, publish. The loop exit is uniform over the
    // wavefront (a ballot), so a lane's publishing store stays inside the
    // loop and a lane waiting on an earlier lane of the same wavefront sees
    // it on the next pass; with a per-lane exit the compiler may sink the
    // store past the loop, where it runs only once every lane has left.
    while (__builtin_amdgcn_ballot_w64(!done) != 0) {
      if (!done) {
        if (pend == 0u && q < q1) {
          int m = 0;
          gt = 0u;
          loc = 0u;
          for (; q < q1 && m < kPollBatch; ++q) {
            const int j = nbr[q];
            int pj = j;  // position of j in the processing order
            if (!mode_jp) {
              const int cj = color_in[j];
              if (cj >= 0) {
                if (cj < kMaxColors) used |= 1u << cj;
                continue;
              }
              if (j >= i) continue;
            } else {
              pj = pos ? pos[j] : j;
              if (pj >= my_pos) continue;
              if (j > i) gt |= 1u << m;
            }
            if (pj >= base) {
              nj_s[m][lane] = ~(pj - base);
              loc |= 1u << m;
            } else {
              nj_s[m][lane] = j;
            }
            pend |= 1u << m;
            ++m;
          }
        }
        if (pend != 0u) {
          // global entries first, all loads in flight at once; the LDS
          // entries in a loop of their own so no load mixes the two
          int jj[kPollBatch], v[kPollBatch], w[kPollBatch];
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t) jj[t] = nj_s[t][lane];
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t)
            v[t] = ((pend & ~loc) >> t & 1u) ? ld_relaxed(&pub[jj[t]]) : -1;
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t)
            w[t] = ((pend & loc) >> t & 1u) ? lpub_ld(~jj[t]) : -1;
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t) {
            const int vt = v[t] > w[t] ? v[t] : w[t];  // at most one is read
            if (vt >= 0) absorb(t, vt);
          }
          if (pend != 0u && ++spins > kSpinCap) {
            // A dependency never resolved: publish color 0 so waiters
            // drain, and flag the failure for the host.
            atomicExch(err, 1);
            st_relaxed(&pub[i], mode_jp ? 1 : 0);
            lpub_st(lane, mode_jp ? 1 : 0);
            done = true;
          }
        }
        if (!done && pend == 0u && q >= q1) publish();
      }
      // Lanes waiting only on lanes of this ticket: resolve from LDS while
      // any of them progresses (each round resolves at least one entry, so
      // this ends after at most 64 x kPollBatch rounds).
      for (;;) {
        const bool local_only = !done && pend != 0u && (pend & ~loc) == 0u;
        if (__builtin_amdgcn_ballot_w64(local_only) == 0) break;
        bool moved = false;
        if (local_only) {
          int w[kPollBatch];
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t) w[t] = nj_s[t][lane];
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t) w[t] = (pend >> t & 1u) ? lpub_ld(~w[t]) : -1;
#pragma unroll
          for (int t = 0; t < kPollBatch; ++t) {
            if (w[t] >= 0) {
              absorb(t, w[t]);
              moved = true;
            }
          }
          if (pend == 0u && q >= q1) publish();
        }
        if (__builtin_amdgcn_ballot_w64(moved) == 0) break;
      }
      if (__builtin_amdgcn_ballot_w64(!done && pend != 0u) != 0) __builtin_amdgcn_s_sleep(1);
    }
    // every slot is resolved before the next ticket reuses nj_s, lpub and
    // s_base
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256)
k_color_merge(int nb, int* color, const int* __restrict__ pub, int mode_jp,
              const int* __restrict__ scene, const int* __restrict__ scene_dirty,
              const int* __restrict__ nd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_count(nb, nd)) return;
  if (scene_dirty != nullptr && !scene_dirty[scene[i]]) return;
  if (mode_jp) color[i] = (pub[i] & 255) - 1;
  else if (color[i] == -1) color[i] = pub[i];
}

void launch_color_greedy(hipStream_t st, int nb, const int* order, const int* pos,
                         const int* nbr_start, const int* nbr, int* color_in_out, int* pub,
                         int* err, int* ticket, int mode_jp, int* clamped, const int* scene,
                         const int* scene_dirty, const int* nd) {
  if (nb <= 0) return;
  launch_fill_int(st, pub, nb, -1, nd);
  launch_fill_int(st, ticket, 1, 0);
  // Workgroups of 16 wavefronts (tickets of 1024 positions, so more of a
  // chain of consecutive bodies stays inside one ticket), enough of them to
  // cover the GPU; each loops over tickets.
  constexpr int kW = 16;
  const int blocks = std::min((nb + 64 * kW - 1) / (64 * kW), 4096 / kW);
  hipLaunchKernelGGL(k_color_greedy<kW>, dim3(blocks), dim3(64 * kW), 0, st, nb, order, pos,
                     nbr_start, nbr, color_in_out, pub, ticket, mode_jp, err, clamped, scene,
                     scene_dirty, nd);
  hipLaunchKernelGGL(k_color_merge, dim3((nb + 255) / 256), dim3(256), 0, st, nb, color_in_out, pub,
                     mode_jp, scene, scene_dirty, nd);
}

__global__ void __launch_bounds__(256)
k_color_resolve(int nb, const int* __restrict__ nbr_start, const int* __restrict__ nbr,
                const float* __restrict__ prio, const int* __restrict__ color,
                int* __restrict__ color_next, int* uncolored, const int* __restrict__ nd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_count(nb, nd)) return;
  int ci = color[i];
  int out = ci;
  if (ci >= 0) {
    float mp = prio[i];
    for (int q = nbr_start[i]; q < nbr_start[i + 1]; ++q) {
      int j = nbr[q];
      if (color[j] == ci && (prio[j] > mp || (prio[j] == mp && j < i))) {
        out = -1;
        break;
      }
    }
  }
  color_next[i] = out;
  if (out == -1) atomicAdd(uncolored, 1);
}

void launch_color_resolve(hipStream_t st, int nb, const int* nbr_start, const int* nbr,
                          const float* prio, const int* color, int* color_next, int* uncolored,
                          const int* nd) {
  if (nb > 0)
    hipLaunchKernelGGL(k_color_resolve, dim3((nb + 255) / 256), dim3(256), 0, st, nb, nbr_start, nbr,
                       prio, color, color_next, uncolored, nd);
}

// Jacobi rounds after the first, for a substep whose decisions stay on the
// device: the host-driven recolor repeats greedy + resolve until no body is
// uncolored (at most MAX_COLORS rounds, solver_6dof.py:795-844), reading the
// uncolored count back each round. Round 1 leaves bodies uncolored only when
// some body saw all 32 colors among its neighbours, which no scene here
// reaches, so the later rounds run on one workgroup: the bodies still
// uncolored are gathered in index order and colored by one thread exactly as
// the serial greedy does (lowest color not used by a colored neighbour, in
// place), then resolved as k_color_resolve does.
__global__ void __launch_bounds__(1024)
k_color_jacobi_tail(int nb, const int* __restrict__ nbr_start, const int* __restrict__ nbr,
                    const float* __restrict__ prio, int* color, int* color_next, int* uncolored,
                    int* clamped, const int* __restrict__ gate) {
  __shared__ int s_list[1024];
  __shared__ int s_wcnt[16];
  __shared__ int s_unc;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if ((gate != nullptr && *gate == 0) || *uncolored == 0) return;
  for (int round = 1; round < kMaxColors; ++round) {
    // assign, in index order, in chunks of 1024 bodies
    for (int base = 0; base < nb; base += 1024) {
      const int i = base + tid;
      const bool u = i < nb && color[i] == -1;
      const unsigned long long m = __ballot(u);
      if (lane == 0) s_wcnt[wid] = __popcll(m);
      __syncthreads();
      int off = 0, tot = 0;
      for (int w = 0; w < 16; ++w) {
        if (w < wid) off += s_wcnt[w];
        tot += s_wcnt[w];
      }
      if (u) s_list[off + __popcll(m & ((1ull << lane) - 1ull))] = i;
      __syncthreads();
      if (tid == 0) {
        for (int k = 0; k < tot; ++k) {
          const int b = s_list[k];
          unsigned used = 0u;
          for (int q = nbr_start[b]; q < nbr_start[b + 1]; ++q) {
            const int cj = color[nbr[q]];
            if (cj >= 0 && cj < kMaxColors) used |= 1u << cj;
          }
          if (used == 0xffffffffu) atomicOr(clamped, 1);
          color[b] = lowest_free_color(used);
        }
      }
      __syncthreads();
    }
    // resolve into color_next, then back into color
    if (tid == 0) s_unc = 0;
    __syncthreads();
    for (int i = tid; i < nb; i += 1024) {
      const int ci = color[i];
      int out = ci;
      if (ci >= 0) {
        const float mp = prio[i];
        for (int q = nbr_start[i]; q < nbr_start[i + 1]; ++q) {
          const int j = nbr[q];
          if (color[j] == ci && (prio[j] > mp || (prio[j] == mp && j < i))) {
            out = -1;
            break;
          }
        }
      }
      color_next[i] = out;
      if (out == -1) atomicAdd(&s_unc, 1);
    }
    __syncthreads();
    for (int i = tid; i < nb; i += 1024) color[i] = color_next[i];
    __syncthreads();
    if (s_unc == 0) break;
  }
  if (tid == 0) *uncolored = s_unc;
}

void launch_color_jacobi_tail(hipStream_t st, int nb, const int* nbr_start, const int* nbr,
                              const float* prio, int* color, int* color_next, int* uncolored,
                              int* clamped, const int* gate) {
  if (nb > 0)
    hipLaunchKernelGGL(k_color_jacobi_tail, dim3(1), dim3(1024), 0, st, nb, nbr_start, nbr, prio,
                       color, color_next, uncolored, clamped, gate);
}

// Color starts (kMaxColors + 1), the achieved color count (the highest
// non-empty color + 1, at least 1) and the first tile of each color for
// tiles of `bpw` bodies, from the per-color body counts.
__global__ void k_color_meta(const int* __restrict__ counts, int bpw, int* cstart, int* n_active,
                             int* tile_start) {
  if (threadIdx.x != 0) return;
  int c0 = 0, t = 0, mx = 0;
  for (int c = 0; c < kMaxColors; ++c) {
    const int n = counts[c];
    cstart[c] = c0;
    tile_start[c] = t;
    c0 += n;
    t += (n + bpw - 1) / bpw;
    if (n > 0) mx = c;
  }
  cstart[kMaxColors] = c0;
  tile_start[kMaxColors] = t;
  *n_active = mx + 1;
}

void launch_color_meta(hipStream_t st, const int* counts, int bpw, int* cstart, int* n_active,
                       int* tile_start) {
  hipLaunchKernelGGL(k_color_meta, dim3(1), dim3(64), 0, st, counts, bpw, cstart, n_active,
                     tile_start);
}

// Uncolored bodies (-1) take color 0; per-color body counts go through a
// block histogram in LDS, then one global atomic per color per block.
__global__ void __launch_bounds__(256)
k_color_finish(int nb, int* color, int* counts, const int* __restrict__ nd) {
  __shared__ int h[kMaxColors];
  nb = dev_count(nb, nd);
  if (threadIdx.x < kMaxColors) h[threadIdx.x] = 0;
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    int c = color[i];
    if (c == -1) { c = 0; color[i] = 0; }
    if (c >= 0 && c < kMaxColors) atomicAdd(&h[c], 1);
  }
  __syncthreads();
  if (threadIdx.x < kMaxColors && h[threadIdx.x] != 0) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

void launch_color_finish(hipStream_t st, int nb, int* color, int* counts, const int* gate) {
  launch_fill_int(st, counts, kMaxColors, 0, gate);
  if (nb > 0)
    hipLaunchKernelGGL(k_color_finish, dim3((nb + 255) / 256), dim3(256), 0, st, nb, color, counts,
                       gate);
}

__global__ void __launch_bounds__(256)
k_color_keys(int nb, const int* __restrict__ color, unsigned* keys, int* vals,
             const int* __restrict__ nd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dev_count(nb, nd)) return;
  keys[i] = (unsigned)color[i];
  vals[i] = i;
}

void radix_sort_small(hipStream_t st, unsigned* keys, int* vals, unsigned* keys_tmp,
                      int* vals_tmp, int n, int key_bits, void* ws, const int* nd);

// color_bodies: bodies bucketed by color, ascending within a color
// (gpu_color_bodies_scatter, kernels_6dof.py:3182-3199, in serial order).
void color_bucket(hipStream_t st, int nb, const int* color, unsigned* keys, int* vals,
                  unsigned* keys_tmp, int* vals_tmp, void* ws, const int* gate) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_color_keys, dim3((nb + 255) / 256), dim3(256), 0, st, nb, color, keys, vals,
                     gate);
  radix_sort_small(st, keys, vals, keys_tmp, vals_tmp, nb, 8, ws, gate);
}

}  // namespace avbd

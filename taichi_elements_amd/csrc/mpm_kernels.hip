// gfx950 kernels of the MLS-MPM substep.
//
// One substep (mpm_solver.py:789-799) is:
//   k_keys            block key of every particle + per-block counts   (build_pid, mpm:344-361;
//                     first substep, or positions set by the host)
//   k_bscan*          scan of the dense block window: particle offsets,
//                     transfer work items (chunks), active grid blocks  (replaces the pointer
//                                                                        SNodes and deactivate_all)
//   k_onesweep        stable LSD radix sort of slots by block key (with k_keys), or
//   k_offsets,        the steady-state neighbour sort: the previous order
//   k_nscatter        stably re-sorted from g2p's direction codes
//   k_p2g             per chunk: constitutive update + scatter into an
//                     LDS tile of (4+2)^d nodes, tile written to HBM     (p2g, mpm:487-584)
//   k_grid_op         per active node: sum of the overlapping tiles,
//                     normalisation, gravity, post-process              (mpm:586-692)
//   k_g2p             per chunk: LDS tile of grid velocity, gather,
//                     advection, max |v|, next binning's hints          (g2p, mpm:694-735)
//   k_reorder         every few substeps: particle state moved into the sorted order
// use_g2p2g replaces p2g / grid op / g2p with k_g2p2g + k_grid_op_g2p2g
// (mpm_g2p2g.h); the differentiable backward runs the adjoints of
// mpm_adjoint.h.  Particles are read and written through Particles<D, Q>
// (mpm_particles.h), instantiated for the float and the packed layout.
#include "mpm_kernels.h"
#include "mpm_adjoint.h"
#include "mpm_g2p2g.h"

namespace mpm {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int4 add4(int4 a, int4 b) {
  return make_int4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ int4 wave_inclusive_scan4(int4 v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int4 t;
    t.x = __shfl_up(v.x, off);
    t.y = __shfl_up(v.y, off);
    t.z = __shfl_up(v.z, off);
    t.w = __shfl_up(v.w, off);
    if (lane >= off) v = add4(v, t);
  }
  return v;
}

// Exclusive scan of one int4 per thread over a kThreads workgroup.
__device__ __forceinline__ int4 block_exscan4(int4 v, int4* total, int4* wsum) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int4 inc = wave_inclusive_scan4(v);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int4 pre = make_int4(0, 0, 0, 0), tot = make_int4(0, 0, 0, 0);
#pragma unroll
  for (int w = 0; w < kThreads / 64; w++) {
    if (w < wave) pre = add4(pre, wsum[w]);
    tot = add4(tot, wsum[w]);
  }
  __syncthreads();
  *total = tot;
  return make_int4(pre.x + inc.x - v.x, pre.y + inc.y - v.y, pre.z + inc.z - v.z,
                   pre.w + inc.w - v.w);
}

// ---------------------------------------------------------------------------
// binning: block keys and per-block counts (build_pid, mpm_solver.py:357-361)
// ---------------------------------------------------------------------------
template <int D, bool Q>
__global__ __launch_bounds__(kThreads) void k_keys(const Params* __restrict__ scenes,
                                                   const float* __restrict__ pf,
                                                   const int* __restrict__ pi, int64_t cap, int n,
                                                   int total_blocks, int* __restrict__ keys,
                                                   int* __restrict__ count, int* __restrict__ err) {
  using L = Layout<D>;
  const int lane = threadIdx.x & 63;
  for (int i0 = blockIdx.x * kThreads; i0 < n; i0 += gridDim.x * kThreads) {
    const int i = i0 + threadIdx.x;
    bool valid = false;
    int key = total_blocks;
    if (i < n) {
      const Params& P = scenes[pi[(int64_t)L::SCENE * cap + i]];
      float xi[3];
      Particles<D, Q>::load_x(pf, cap, i, xi);
      int lin = 0;
      valid = true;
#pragma unroll
      for (int a = D - 1; a >= 0; a--) {
        int base;
        float fx;
        base_fx(xi[a], P.inv_dx, base, fx);
        const int lb = (base >> 2) - P.blo[a];
        // the block and its +1 neighbour must lie inside the scene's window
        if (lb < 0 || lb >= P.bnl[a] - 1) valid = false;
        lin = lin * P.bn[a] + lb + (a == D - 1 ? P.zoff : 0);
      }
      key = valid ? lin : total_blocks;
      keys[i] = key;
      if (!valid) atomicOr(err, 1);
    }
    // wave-aggregated counting (count[total_blocks] counts escaped particles):
    // sorted input gives 1-3 distinct keys per wave
    const bool in = i < n;
    unsigned long long pending = __ballot(in);
    while (pending) {
      const int leader = __ffsll(pending) - 1;
      const int lk = __shfl(key, leader);
      const unsigned long long same = __ballot(in && key == lk);
      if (lane == leader) atomicAdd(&count[lk], (int)__popcll(same));
      pending &= ~same;
    }
  }
}

// Per-block values of the dense window scan: (particles, chunks, active grid
// block, non-empty particle block).  A grid block is active when one of the
// particle blocks b - {0,1}^d holds particles (their stencils reach it).
template <int D>
__device__ __forceinline__ int4 block_values(int bid, const int* __restrict__ count,
                                             const Params& P, int chunk) {
  const int c = count[bid];
  int lb[3];
  decode_local<D>(bid, P, lb);
  bool act = false;
  // p2g reaches blocks b + {0, 1}^d of a particle block b; the fused g2p2g
  // transfer (new positions up to one cell away) b + {-1, 0, 1}^d
  const int lo = P.g2p2g ? -1 : 0;
#pragma unroll
  for (int dz = (D == 3 ? -1 : 0); dz < (D == 3 ? 2 : 1); dz++)
#pragma unroll
    for (int dy = -1; dy < 2; dy++)
#pragma unroll
      for (int dx = -1; dx < 2; dx++) {
        if (dx < lo || dy < lo || (D == 3 && dz < lo)) continue;
        int q[3] = {lb[0] - dx, lb[1] - dy, lb[2] - dz};
        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= P.bn[0] || q[1] >= P.bn[1] ||
            (D == 3 && q[2] >= P.bn[2]))
          continue;
        act |= count[encode_local<D>(q, P)] > 0;
      }
  return make_int4(c, (c + chunk - 1) / chunk, act ? 1 : 0, c > 0 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// single-pass scans with decoupled look-back (Merrill & Garland 2016)
//
// Status word: value in bits 0-31, flag in bits 32-33 (1 = tile aggregate,
// 2 = inclusive prefix), launch epoch in bits 34-63.  Words left by earlier
// launches carry an older epoch and read as "not ready", so the status
// arrays are never cleared.  Tiles are numbered in the order workgroups start
// (an atomic ticket), so a tile only waits on tiles that are already running;
// the last workgroup to finish resets the ticket and advances the epoch,
// which keeps the kernels free of host-side per-launch state.
// ---------------------------------------------------------------------------
constexpr unsigned long long kFlagAgg = 1ull << 32, kFlagPre = 2ull << 32;
constexpr int kSpinLimit = 1 << 18;

__device__ __forceinline__ unsigned long long lb_epoch_bits(unsigned epoch) {
  return (unsigned long long)(epoch & 0x3fffffffu) << 34;
}

__device__ __forceinline__ void lb_publish(unsigned long long* p, unsigned long long flag,
                                           unsigned epoch, unsigned value) {
  __hip_atomic_store(p, lb_epoch_bits(epoch) | flag | (unsigned long long)value, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of channel `chan` for tile `tile` (status[t * stride + chan]).
// A bounded wait: past kSpinLimit polls it flags err bit 1 and returns what it has.
__device__ __forceinline__ unsigned lb_lookback(const unsigned long long* status, int stride,
                                                int tile, int chan, unsigned epoch, int* err) {
  const unsigned long long ep = lb_epoch_bits(epoch);
  unsigned excl = 0;
  for (int t = tile - 1; t >= 0; t--) {
    unsigned long long w = 0;
    int spins = 0;
    for (;;) {
      w = __hip_atomic_load(&status[(int64_t)t * stride + chan], __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      if ((w & ~0x3ffffffffull) == ep && (w & (kFlagAgg | kFlagPre))) break;
      if (++spins > kSpinLimit) {
        atomicOr(err, 2);
        return excl;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    excl += (unsigned)w;
    if (w & kFlagPre) break;
  }
  return excl;
}

// The same prefix, looked back by a whole wave: lane j reads the status of
// tile - 1 - j (64 predecessors per round trip instead of one); the nearest
// inclusive prefix in the window ends the sum.  Every lane returns the sum.
__device__ __forceinline__ unsigned lb_lookback_wave(const unsigned long long* status, int stride,
                                                     int tile, int chan, unsigned epoch,
                                                     int* err) {
  const unsigned long long ep = lb_epoch_bits(epoch);
  const int lane = threadIdx.x & 63;
  unsigned excl = 0;
  for (int top = tile - 1; top >= 0; top -= 64) {
    const int t = top - lane;
    unsigned long long w = 0;
    if (t >= 0) {
      int spins = 0;
      for (;;) {
        w = __hip_atomic_load(&status[(int64_t)t * stride + chan], __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        if ((w & ~0x3ffffffffull) == ep && (w & (kFlagAgg | kFlagPre))) break;
        if (++spins > kSpinLimit) {
          atomicOr(err, 2);
          w = ep | kFlagPre;  // give up: count this tile as a zero prefix
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    const unsigned long long pre = __ballot(t >= 0 && (w & kFlagPre) != 0);
    const int stop = pre ? __ffsll((unsigned long long)pre) - 1 : 63;
    unsigned v = (t >= 0 && lane <= stop) ? (unsigned)w : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    excl += v;
    if (pre) break;
  }
  return excl;
}

struct LBTicket {
  int tile;
  unsigned epoch;
};

__device__ __forceinline__ LBTicket lb_acquire(LBState* st, LBTicket* sh) {
  if (threadIdx.x == 0) {
    sh->tile = (int)atomicAdd(&st->tile_ctr, 1ull);
    sh->epoch = __hip_atomic_load(&st->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return *sh;
}

__device__ __forceinline__ void lb_release(LBState* st, int ntiles, unsigned epoch) {
  if (threadIdx.x == 0) {
    const unsigned long long d = atomicAdd(&st->done_ctr, 1ull);
    if (d == (unsigned long long)(ntiles - 1)) {
      __hip_atomic_store(&st->tile_ctr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&st->done_ctr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&st->epoch, epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---------------------------------------------------------------------------
// scan of the dense block window (one launch): particle offsets, chunk list,
// active grid blocks, and the radix digit histograms of the block keys
// weighted by the particle counts (count[total_blocks] holds the particles
// outside the window, whose key is total_blocks).
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void k_bscan(
    const Params* __restrict__ scenes, const int* __restrict__ count, int* __restrict__ count_next, int total_blocks,
    int chunk, int ntiles, int passes, int* __restrict__ pstart, int* __restrict__ pchunk_begin,
    int* __restrict__ pchunk_cnt, int4* __restrict__ chunks, int* __restrict__ active,
    int* __restrict__ totals, int* __restrict__ ghist, int* __restrict__ ghist_next,
    unsigned long long* __restrict__ status, LBState* __restrict__ lbs,
    const int4* __restrict__ tile_prefix, int* __restrict__ err, int* __restrict__ step) {
  __shared__ LBTicket tk_sh;
  // one launch per substep: the substep counter of the max-velocity slots
  if (step && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(step, 1);
  __shared__ int4 wsum[kThreads / 64];
  __shared__ unsigned excl_sh[4];
  __shared__ int lh[kMaxRadixPasses][kRadixBins];
  const Params& P = scenes[0];  // the combined window extents
  // tile_prefix != nullptr: exclusive tile prefixes were computed by
  // k_bscan_reduce + k_bscan_top (large windows, where a look-back chain over
  // thousands of tiles is slower than two extra launches)
  LBTicket tk;
  if (tile_prefix) {
    tk.tile = blockIdx.x;
    tk.epoch = 0;
  } else {
    tk = lb_acquire(lbs, &tk_sh);
  }
  const int tile = tk.tile;
  for (int i = threadIdx.x; i < kMaxRadixPasses * kRadixBins; i += kThreads) (&lh[0][0])[i] = 0;

  const int first = tile * kScanTile + threadIdx.x * kScanItems;
  int4 vals[kScanItems];
  int4 acc = make_int4(0, 0, 0, 0);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int bid = first + k;
    vals[k] = bid < total_blocks ? block_values<D>(bid, count, P, chunk) : make_int4(0, 0, 0, 0);
    acc = add4(acc, vals[k]);
  }
  int4 tot;
  int4 pre = block_exscan4(acc, &tot, wsum);  // contains __syncthreads (lh zeroed)

  // digit histograms of the keys, weighted by particle counts
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int bid = first + k;
    if (bid < total_blocks && vals[k].x > 0)
      for (int ps = 0; ps < passes; ps++) atomicAdd(&lh[ps][(bid >> (8 * ps)) & 255], vals[k].x);
  }
  if (tile == 0 && threadIdx.x == 0) {
    const int inval = count[total_blocks];
    if (inval > 0)
      for (int ps = 0; ps < passes; ps++)
        atomicAdd(&lh[ps][(total_blocks >> (8 * ps)) & 255], inval);
  }

  static_assert(kThreads == 256, "one wave per scanned channel");
  if (tile_prefix) {
    if (threadIdx.x < 4) {
      const int4 tp = tile_prefix[tile];
      const int c = threadIdx.x;
      excl_sh[c] = (unsigned)(c == 0 ? tp.x : c == 1 ? tp.y : c == 2 ? tp.z : tp.w);
    }
  } else {
    // wave c looks back channel c
    const int c = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned agg = (unsigned)(c == 0 ? tot.x : c == 1 ? tot.y : c == 2 ? tot.z : tot.w);
    unsigned excl = 0;
    if (tile == 0) {
      if (lane == 0) lb_publish(&status[c], kFlagPre, tk.epoch, agg);
    } else {
      if (lane == 0) lb_publish(&status[(int64_t)tile * 4 + c], kFlagAgg, tk.epoch, agg);
      excl = lb_lookback_wave(status, 4, tile, c, tk.epoch, err);
      if (lane == 0) lb_publish(&status[(int64_t)tile * 4 + c], kFlagPre, tk.epoch, excl + agg);
    }
    if (lane == 0) excl_sh[c] = excl;
  }
  __syncthreads();
  pre = add4(pre, make_int4((int)excl_sh[0], (int)excl_sh[1], (int)excl_sh[2], (int)excl_sh[3]));
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int bid = first + k;
    if (bid >= total_blocks) break;
    const int4 v = vals[k];
    pstart[bid] = pre.x;
    pchunk_begin[bid] = pre.y;
    pchunk_cnt[bid] = v.y;
    for (int c = 0; c < v.y; c++) {
      const int b0 = pre.x + c * chunk;
      const int b1 = min(b0 + chunk, pre.x + v.x);
      chunks[pre.y + c] = make_int4(bid, b0, b1, 0);
    }
    if (v.z) active[pre.z] = bid;
    count_next[bid] = 0;
    pre = add4(pre, v);
  }
  if (tile == 0 && threadIdx.x == 0) count_next[total_blocks] = 0;
  if (tile == ntiles - 1 && threadIdx.x == 0) {
    totals[kTotValid] = (int)excl_sh[0] + tot.x;
    totals[kTotChunks] = (int)excl_sh[1] + tot.y;
    totals[kTotActive] = (int)excl_sh[2] + tot.z;
    totals[kTotPBlocks] = (int)excl_sh[3] + tot.w;
  }
  // flush the histograms; clear next substep's copy
  for (int i = threadIdx.x; i < passes * kRadixBins; i += kThreads) {
    const int h = (&lh[0][0])[i];
    if (h) atomicAdd(&ghist[i], h);
  }
  if (tile == 0)
    for (int i = threadIdx.x; i < kMaxRadixPasses * kRadixBins; i += kThreads) ghist_next[i] = 0;
  __syncthreads();
  if (!tile_prefix) lb_release(lbs, ntiles, tk.epoch);
}

// Two-level mode of the block-window scan: per-tile sums, then one workgroup
// turns them into exclusive tile prefixes.
template <int D>
__global__ __launch_bounds__(kThreads) void k_bscan_reduce(const Params* __restrict__ scenes,
                                                           const int* __restrict__ count,
                                                           int total_blocks, int chunk,
                                                           int4* __restrict__ partials) {
  __shared__ int4 wsum[kThreads / 64];
  const Params& P = scenes[0];  // the combined window extents
  int4 acc = make_int4(0, 0, 0, 0);
  const int first = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
  for (int k = 0; k < kScanItems; k++)
    if (first + k < total_blocks) acc = add4(acc, block_values<D>(first + k, count, P, chunk));
  int4 tot;
  block_exscan4(acc, &tot, wsum);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Each thread owns a contiguous run of the tile sums: one pass of
// independent loads, one workgroup scan of the run totals, then the run's
// exclusive prefixes (one dependent round instead of one per 256 tiles).
__global__ __launch_bounds__(kThreads) void k_bscan_top(int4* __restrict__ partials, int nparts) {
  __shared__ int4 wsum[kThreads / 64];
  const int per = (nparts + kThreads - 1) / kThreads;
  const int i0 = min(nparts, (int)threadIdx.x * per), i1 = min(nparts, i0 + per);
  int4 acc = make_int4(0, 0, 0, 0);
#pragma unroll 8
  for (int i = i0; i < i1; i++) acc = add4(acc, partials[i]);
  int4 tot;
  int4 run = block_exscan4(acc, &tot, wsum);
#pragma unroll 8
  for (int i = i0; i < i1; i++) {
    const int4 v = partials[i];
    partials[i] = run;
    run = add4(run, v);
  }
}

// ---------------------------------------------------------------------------
// one pass of the stable LSD radix sort (8-bit digit), single launch.  Each
// wave ranks its own contiguous quarter of the tile with ballots and
// wave-private LDS counters (no workgroup barrier per round); one prefix over
// the four waves gives tile-local stable ranks; decoupled look-back over
// tiles gives the per-digit tile offsets; the digit totals come from the
// histogram k_bscan built.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_onesweep(
    const int* __restrict__ keys_in, const int* __restrict__ vals_in, int* __restrict__ keys_out,
    int* __restrict__ vals_out, int n, int shift, int ntiles, const int* __restrict__ ghist,
    unsigned long long* __restrict__ status, LBState* __restrict__ lbs, int* __restrict__ err) {
  constexpr int WAVES = kThreads / 64;
  constexpr int PER_WAVE = kRadixTile / WAVES;
  constexpr int ROUNDS = PER_WAVE / 64;
  __shared__ LBTicket tk_sh;
  __shared__ int wcnt[WAVES][kRadixBins];
  __shared__ int gbase[kRadixBins];
  __shared__ int4 wsum[WAVES];
  const LBTicket tk = lb_acquire(lbs, &tk_sh);
  const int tile = tk.tile;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int i0 = tile * kRadixTile + wave * PER_WAVE + lane;

  int key[ROUNDS], val[ROUNDS], lr[ROUNDS];
#pragma unroll
  for (int r = 0; r < ROUNDS; r++) {
    const int i = i0 + r * 64;
    key[r] = i < n ? keys_in[i] : 0;
    val[r] = i < n ? (vals_in ? vals_in[i] : i) : 0;
  }
#pragma unroll
  for (int w = 0; w < WAVES; w++) wcnt[w][threadIdx.x] = 0;
  int4 gtot;
  // block_exscan4 synchronises, which also publishes the zeroed counters
  const int gex = block_exscan4(make_int4(ghist[threadIdx.x], 0, 0, 0), &gtot, wsum).x;
  int* my = wcnt[wave];
#pragma unroll
  for (int r = 0; r < ROUNDS; r++) {
    const bool valid = i0 + r * 64 < n;
    const int d = (key[r] >> shift) & (kRadixBins - 1);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1;
      const unsigned long long bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const int rank = (int)__popcll(peers & lt);
    // every lane reads its digit's counter before the group leader bumps it;
    // a wave's LDS accesses are served in program order
    const int base = my[d];
    __asm__ volatile("" ::: "memory");
    if (valid && rank == 0) my[d] = base + (int)__popcll(peers);
    __asm__ volatile("" ::: "memory");
    lr[r] = base + rank;
  }
  __syncthreads();
  int agg = 0;
#pragma unroll
  for (int w = 0; w < WAVES; w++) {
    const int c = wcnt[w][threadIdx.x];
    wcnt[w][threadIdx.x] = agg;
    agg += c;
  }
  {
    unsigned excl = 0;
    unsigned long long* st = &status[(int64_t)tile * kRadixBins + threadIdx.x];
    if (tile == 0) {
      lb_publish(st, kFlagPre, tk.epoch, (unsigned)agg);
    } else {
      lb_publish(st, kFlagAgg, tk.epoch, (unsigned)agg);
      excl = lb_lookback(status, kRadixBins, tile, threadIdx.x, tk.epoch, err);
      lb_publish(st, kFlagPre, tk.epoch, excl + (unsigned)agg);
    }
    gbase[threadIdx.x] = gex + (int)excl;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ROUNDS; r++) {
    if (i0 + r * 64 < n) {
      const int d = (key[r] >> shift) & (kRadixBins - 1);
      const int pos = gbase[d] + my[d] + lr[r];
      if (pos >= 0 && pos < n) {
        keys_out[pos] = key[r];
        vals_out[pos] = val[r];
      } else {
        atomicOr(err, 4);  // histogram and keys disagree (cannot happen)
      }
    }
  }
  __syncthreads();
  lb_release(lbs, ntiles, tk.epoch);
}

// ---------------------------------------------------------------------------
// P2G (mpm_solver.py:487-584)
//
// One workgroup per chunk (<= kChunkMax particles of one grid block), taken
// kBatch particles at a time through phases 1-3; one node tile per chunk.
//  phase 1  one particle per thread: constitutive update (F, Jp written back)
//           and the particle's affine momentum terms staged in LDS;
//  phase 2  a stable counting sort of the chunk's particles by cell;
//  phase 3  cell-major scatter: lane l of a wave owns cell l of the block
//           (3D: 64 cells = 64 lanes; 2D: 16 cells, four lane groups), so the
//           27 (9) stencil nodes a wave updates in one instruction are
//           distinct and each wave / lane group accumulates into its own LDS
//           tile with plain read-modify-write -- no float atomics;
//  phase 4  the private tiles are summed in a fixed order and the (4+2)^d
//           node tile (momentum, mass) is written to HBM for the grid pass.
// The accumulation order is fixed, so p2g is bitwise reproducible.
// ---------------------------------------------------------------------------
// MODE 0: the forward transfer.  MODE 1: the scatter half of the g2p adjoint
// -- node o of particle p receives w_o (a + c4 Cbar (o - fx)) with
// a = vbar' + dt xbar' and Cbar the adjoint of the new C (gf holds the
// adjoints of the new particle state, same layout as pf); the mass channel
// stays zero.  Both scatter "q + Ad o" per node through the same LDS path.
// At most 128 VGPRs (4 waves per SIMD): the latency of the particle loads
// and of the LDS scatter is hidden better than at 3 waves with no spills
// (measured: 85 -> 75 us per 1M-particle substep, a few spilled registers).
template <int D, int MODE, bool Q = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_p2g(
    const Params* __restrict__ scenes, const int* __restrict__ slab, float* __restrict__ pf,
    const int* __restrict__ pi, int64_t cap, const int* __restrict__ perm,
    const int4* __restrict__ chunks, const int* __restrict__ totals, float4* __restrict__ tiles,
    const float* __restrict__ gf, float* __restrict__ pf_out) {
  using L = Layout<D>;
  constexpr int TN = Geo<D>::kTileNodes;
  constexpr int CELLS = Geo<D>::kNodesPerBlock;  // cells of a block (64 / 16)
  constexpr int GROUPS = 64 / CELLS;             // lane groups per wave
  constexpr int WAVES = kThreads / 64;
  constexpr int SUBS = WAVES * GROUPS;           // private tiles
  __shared__ float4 stage[4][kBatch];  // [record part][particle]: lane-contiguous
  __shared__ float4 ptile[SUBS][TN];
  __shared__ int wcnt[WAVES][CELLS];
  __shared__ int cstart[CELLS], ccount[CELLS];
  __shared__ unsigned char order[kBatch];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int nchunks = totals[kTotChunks];
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int4 ck = chunks[c];
    int lbc[3], b[3];
    decode_local<D>(ck.x, scenes[0], lbc);
    const Params& P = scenes[scene_of<D>(lbc, slab)];
    if (P.skip) continue;  // workgroup-uniform: the whole chunk belongs to one scene
    to_local<D>(lbc, P, b);
#pragma unroll
    for (int a = 0; a < D; a++) b[a] = (b[a] + P.blo[a]) * kBlock;  // first node of the block
    for (int i = threadIdx.x; i < SUBS * TN; i += kThreads)
      (&ptile[0][0])[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the chunk's particles, kBatch at a time, all into the same private tiles
    for (int b0 = ck.y; b0 < ck.z; b0 += kBatch) {
      const int nc = min(kBatch, ck.z - b0);
      for (int i = threadIdx.x; i < WAVES * CELLS; i += kThreads) (&wcnt[0][0])[i] = 0;

      // ---- phase 1: constitutive update and staging ----
      const int t = threadIdx.x;
      const bool valid = t < nc;
      int cell = 0;
      if (MODE == 1 && valid) {
        const int p = perm[b0 + t];
        const int mat = pi[(int64_t)L::MAT * cap + p];
        float fx[3] = {0.0f, 0.0f, 0.0f};
        int l[3] = {0, 0, 0};
  #pragma unroll
        for (int a = 0; a < D; a++) {
          int base;
          base_fx(pf[(int64_t)(L::X + a) * cap + p], P.inv_dx, base, fx[a]);
          l[a] = min(max(base - b[a], 0), kBlock - 1);
        }
        cell = D == 3 ? l[0] + 4 * (l[1] + 4 * l[2]) : l[0] + 4 * l[1];
        // a stationary particle keeps x, v, C in g2p: no grid adjoint
        const float on = mat == kStationary ? 0.0f : 1.0f;
        float Ad[3][3] = {{0.0f}}, q[3] = {0.0f, 0.0f, 0.0f};
  #pragma unroll
        for (int a = 0; a < D; a++) {
          float s = on * (gf[(int64_t)(L::V + a) * cap + p] + P.dt * gf[(int64_t)(L::X + a) * cap + p]);
  #pragma unroll
          for (int e = 0; e < D; e++) {
            Ad[a][e] = on * P.c4 * gf[(int64_t)(L::C + a * D + e) * cap + p];
            s -= Ad[a][e] * fx[e];
          }
          q[a] = s;
        }
        if (D == 3) {
          stage[0][t] = make_float4(Ad[0][0], Ad[0][1], Ad[0][2], Ad[1][0]);
          stage[1][t] = make_float4(Ad[1][1], Ad[1][2], Ad[2][0], Ad[2][1]);
          stage[2][t] = make_float4(Ad[2][2], q[0], q[1], q[2]);
          stage[3][t] = make_float4(fx[0], fx[1], fx[2], 0.0f);
        } else {
          stage[0][t] = make_float4(Ad[0][0], Ad[0][1], Ad[1][0], Ad[1][1]);
          stage[1][t] = make_float4(q[0], q[1], fx[0], fx[1]);
          stage[2][t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
      } else if (MODE == 0 && valid) {
        using PS = Particles<D, Q>;
        const int p = perm[b0 + t];
        const int mat = pi[(int64_t)L::MAT * cap + p];
        float xp[3], vp[3], C[3][3], F[3][3], stress[3][3];
        PS::load_x(pf, cap, p, xp);
        PS::load_v(pf, cap, p, vp);
        PS::load_C(pf, cap, p, C);
        PS::load_F(pf, cap, p, F);
        float Jp = P.plastic ? PS::load_Jp(pf, cap, p) : 1.0f;
        float fx[3] = {0.0f, 0.0f, 0.0f};
        int l[3] = {0, 0, 0};
  #pragma unroll
        for (int a = 0; a < D; a++) {
          int base;
          base_fx(xp[a], P.inv_dx, base, fx[a]);
          // the binning pass put this particle in this block from the same x
          l[a] = min(max(base - b[a], 0), kBlock - 1);
        }
        cell = D == 3 ? l[0] + 4 * (l[1] + 4 * l[2]) : l[0] + 4 * l[1];
        float mu_0 = P.mu_0, lambda_0 = P.lambda_0;
        if (P.youngs) {
          const float E = P.youngs[p];
          mu_0 = E * P.kmu;
          lambda_0 = E * P.kla;
        }
        constitutive<D>(mat, P, mu_0, lambda_0, F, C, Jp, stress);
        PS::store_F(pf_out, cap, p, F);  // the stress already used F before rounding (mpm:567)
        if (P.plastic) PS::store_Jp(pf_out, cap, p, Jp);
        float mass = P.p_mass;  // mpm_solver.py:571-573
        if (mat == kWater) mass *= P.water_density;
        // contribution to node o: w_o * (mass v + A (o - fx) dx) = w_o * (q + Adx o)
        // with A = stress_scale tau + mass C (mpm_solver.py:569-583)
        float Ad[3][3] = {{0.0f}}, q[3] = {0.0f, 0.0f, 0.0f};
  #pragma unroll
        for (int a = 0; a < D; a++) {
  #pragma unroll
          for (int e = 0; e < D; e++) Ad[a][e] = (P.stress_scale * stress[a][e] + mass * C[a][e]) * P.dx;
          float s = mass * vp[a];
  #pragma unroll
          for (int e = 0; e < D; e++) s -= Ad[a][e] * fx[e];
          q[a] = s;
        }
        if (D == 3) {
          stage[0][t] = make_float4(Ad[0][0], Ad[0][1], Ad[0][2], Ad[1][0]);
          stage[1][t] = make_float4(Ad[1][1], Ad[1][2], Ad[2][0], Ad[2][1]);
          stage[2][t] = make_float4(Ad[2][2], q[0], q[1], q[2]);
          stage[3][t] = make_float4(fx[0], fx[1], fx[2], mass);
        } else {
          stage[0][t] = make_float4(Ad[0][0], Ad[0][1], Ad[1][0], Ad[1][1]);
          stage[1][t] = make_float4(q[0], q[1], fx[0], fx[1]);
          stage[2][t] = make_float4(mass, 0.0f, 0.0f, 0.0f);
        }
      }
      lds_barrier();

      // ---- phase 2: stable counting sort of the chunk by cell ----
      unsigned long long peers = __ballot(valid);
  #pragma unroll
      for (int bit = 0; bit < (D == 3 ? 6 : 4); bit++) {
        const bool on = (cell >> bit) & 1;
        const unsigned long long bal = __ballot(on);
        peers &= on ? bal : ~bal;
      }
      const int rank = (int)__popcll(peers & lt);
      if (valid && rank == 0) wcnt[wave][cell] = (int)__popcll(peers);
      lds_barrier();
      if (threadIdx.x < 64) {  // one wave: per-cell totals and exclusive scan
        int run = 0;
        if (lane < CELLS) {
  #pragma unroll
          for (int w = 0; w < WAVES; w++) {
            const int cnt = wcnt[w][lane];
            wcnt[w][lane] = run;
            run += cnt;
          }
        }
        int inc = run;
  #pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int y = __shfl_up(inc, off);
          if (lane >= off) inc += y;
        }
        if (lane < CELLS) {
          ccount[lane] = run;
          cstart[lane] = inc - run;
        }
      }
      lds_barrier();
      if (valid) order[cstart[cell] + wcnt[wave][cell] + rank] = (unsigned char)t;
      lds_barrier();

      // ---- phase 3: cell-major scatter into private tiles ----
      {
        const int my_cell = scatter_cell_of_lane<D>(lane);
        const int sub = wave * GROUPS + lane / CELLS;
        const int lc[3] = {my_cell & 3, (my_cell >> 2) & 3, my_cell >> 4};
        const int cbeg = cstart[my_cell], ccnt = ccount[my_cell];
        float4* tp = ptile[sub];
        cell_scatter<D, SUBS>(sub, ccnt, cbeg, order, stage, lc, tp);
      }
      lds_barrier();
    }  // batches of the chunk

    // ---- phase 4: sum the private tiles, write the node tile ----
    float4* out = tiles + (int64_t)c * TN;
    for (int i = threadIdx.x; i < TN; i += kThreads) {
      float4 acc = ptile[0][i];
#pragma unroll
      for (int sdx = 1; sdx < SUBS; sdx++) {
        const float4 v = ptile[sdx][i];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
      }
      out[i] = acc;
    }
    lds_barrier();
  }
}

// ---------------------------------------------------------------------------
// grid operator (mpm_solver.py:586-692)
// ---------------------------------------------------------------------------
// Bounding-box walls zero the outward velocity components near the domain
// faces (mpm_solver.py:600-616); bit a of the result is set when component a
// of v (before the walls) passes unchanged.
// MODE 0: the grid operator (raw != 0: store the p2g sums unnormalised).
// MODE 1: its adjoint -- the tiles hold the g2p-adjoint scatter vbar; with
// the forward's raw sums (p, m) in grid, write (pbar, mbar) to gbar:
// v = post(p / m + dt g) gives vbar' = post^T(vbar) (apply_post_adjoint:
// walls, planes and spheres), pbar = vbar' / m and
// mbar = -sum_a pbar_a p_a / m.
template <int D, int MODE>
__global__ __launch_bounds__(kThreads) void k_grid_op(const Params* __restrict__ scenes,
                                                      const int* __restrict__ slab,
                                                      const int* __restrict__ active,
                                                      const int* __restrict__ totals,
                                                      const int* __restrict__ pchunk_begin,
                                                      const int* __restrict__ pchunk_cnt,
                                                      const float4* __restrict__ tiles,
                                                      float4* __restrict__ grid, int raw,
                                                      float4* __restrict__ gbar,
                                                      float4* __restrict__ raw_out) {
  constexpr int NPB = Geo<D>::kNodesPerBlock, TN = Geo<D>::kTileNodes;
  // Lanes per node (node_lanes), each summing the tiles of some of the SRC
  // particle blocks b - d (d in {0,1}^D) whose tiles reach the node; the
  // partial sums combine in a fixed tree (tree_sum), so the result does not
  // depend on the lane count.  A sparse window gets a lane per source block:
  // this latency-bound kernel then has SRC times the loads in flight.
  constexpr int SRC = D == 3 ? 8 : 4;
  const int64_t nodes = (int64_t)totals[kTotActive] * NPB;
  const int lanes = node_lanes(nodes, SRC);
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < nodes * lanes;
       t += (int64_t)gridDim.x * kThreads) {
    // groups of `lanes` lanes are aligned and enter or leave the loop together
    const int64_t nt = t / lanes;
    const int sub = (int)(t % lanes);
    const int blk = (int)(nt / NPB), node = (int)(nt % NPB);
    const int bid = active[blk];
    int lb[3];
    decode_local<D>(bid, scenes[0], lb);
    const Params& P = scenes[scene_of<D>(lb, slab)];
    if (P.skip) continue;  // uniform over the group: one node, one scene
    const int l[3] = {node & 3, (node >> 2) & 3, node >> 4};
    const float4 sum = tree_sum<SRC>(lanes, sub, [&](int src) {
      float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const int d[3] = {src & 1, (src >> 1) & 1, src >> 2};
      int q[3], tl[3];
      bool ok = true;
#pragma unroll
      for (int a = 0; a < D; a++) {
        tl[a] = l[a] + kBlock * d[a];
        q[a] = lb[a] - d[a];
        ok &= (tl[a] < kTile) & (q[a] >= 0);
      }
      if (ok) {
        const int nb = encode_local<D>(q, P);
        const int c0 = pchunk_begin[nb], cn = pchunk_cnt[nb];
        const int tn = D == 3 ? tl[0] + kTile * (tl[1] + kTile * tl[2]) : tl[0] + kTile * tl[1];
        for (int k = c0; k < c0 + cn; k++) {
          const float4 tv = tiles[(int64_t)k * TN + tn];
          acc.x += tv.x;
          acc.y += tv.y;
          if (D == 3) acc.z += tv.z;
          acc.w += tv.w;
        }
      }
      return acc;
    });
    if (sub != 0) continue;
    float m = sum.w, p[3] = {sum.x, sum.y, sum.z};
    if (MODE == 0 && raw_out) raw_out[(int64_t)bid * NPB + node] = sum;
    if (MODE == 1) {
      const float4 f = grid[(int64_t)bid * NPB + node];
      float pb[3] = {0.0f, 0.0f, 0.0f}, mb = 0.0f;
      if (node_has_mass(f.w)) {
        const float inv = 1.0f / f.w;
        const float pf4[3] = {f.x, f.y, f.z};
        float v[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int a = 0; a < D; a++) v[a] = pf4[a] * inv + P.dt * P.gravity[a];
        int I[3], lbl[3];
        to_local<D>(lb, P, lbl);
#pragma unroll
        for (int a = 0; a < D; a++) I[a] = (lbl[a] + P.blo[a]) * kBlock + l[a];
        float vb[3] = {p[0], p[1], p[2]};
        apply_post_adjoint<D>(P, I, v, vb);
#pragma unroll
        for (int a = 0; a < D; a++) {
          pb[a] = vb[a] * inv;
          mb -= pb[a] * (pf4[a] * inv);
        }
      }
      gbar[(int64_t)bid * NPB + node] = make_float4(pb[0], pb[1], D == 3 ? pb[2] : 0.0f, mb);
      continue;
    }
    if (!raw) {
      if (node_has_mass(m)) {  // mpm_solver.py:591-593
        const float inv = 1.0f / m;
#pragma unroll
        for (int a = 0; a < D; a++) p[a] = inv * p[a];
#pragma unroll
        for (int a = 0; a < D; a++) p[a] += P.dt * P.gravity[a];
      }
      int I[3], lbl[3];
      to_local<D>(lb, P, lbl);
#pragma unroll
      for (int a = 0; a < D; a++) I[a] = (lbl[a] + P.blo[a]) * kBlock + l[a];
      apply_post<D>(P, I, p);
    }
    grid[(int64_t)bid * NPB + node] = make_float4(p[0], p[1], D == 3 ? p[2] : 0.0f, m);
  }
}

// ---------------------------------------------------------------------------
// G2P (mpm_solver.py:694-735)
// ---------------------------------------------------------------------------
// W: minimum waves per SIMD the register allocation must allow (1: no
// constraint, 5 waves at its register count; 6: at most 80 registers, more
// grid-tile gathers in flight where each block holds few particles).
template <int D, bool Q = false, int W = 1>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(W))) void k_g2p(const Params* __restrict__ scenes,
                                                  const int* __restrict__ slab,
                                                  float* __restrict__ pf,
                                                  const int* __restrict__ pi, int64_t cap,
                                                  const int* __restrict__ perm,
                                                  const int4* __restrict__ chunks,
                                                  const int* __restrict__ totals,
                                                  const float4* __restrict__ grid,
                                                  unsigned* __restrict__ maxv,
                                                  const int* __restrict__ step, int stride,
                                                  int* __restrict__ code, int* __restrict__ nkey,
                                                  int* __restrict__ dcount,
                                                  int* __restrict__ count_next,
                                                  int* __restrict__ err,
                                                  float* __restrict__ pf_out, int n) {
  using L = Layout<D>;
  constexpr int NPB = Geo<D>::kNodesPerBlock, TN = Geo<D>::kTileNodes;
  // grid velocity tile as three float arrays: the per-particle gathers hit
  // fewer LDS bank conflicts than with one float4 per node (29 vs 32 us per
  // 1M-particle substep)
  __shared__ float tile_x[TN], tile_y[TN], tile_z[TN];
  __shared__ float red[kThreads / 64];
  __shared__ int dwcnt[kThreads / 64][32];
  __shared__ int dbase[32];  // particles per direction in the chunk's earlier batches
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  float vmax = 0.0f;
  const int nchunks = totals[kTotChunks];
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int4 ck = chunks[c];
    int lb[3], lbl[3];
    decode_local<D>(ck.x, scenes[0], lb);
    const Params& P = scenes[scene_of<D>(lb, slab)];
    to_local<D>(lb, P, lbl);
    // a skipped scene keeps its particles (and so their blocks) unchanged; its
    // grid was not rebuilt and is not read
    const bool skip = P.skip != 0;
    if (threadIdx.x < 32) dbase[threadIdx.x] = 0;
    for (int i = skip ? TN : threadIdx.x; i < TN; i += kThreads) {
      const int tl[3] = {i % kTile, (i / kTile) % kTile, i / (kTile * kTile)};
      int q[3], ln[3];
#pragma unroll
      for (int a = 0; a < D; a++) {
        const int hi = tl[a] >= kBlock;
        q[a] = lb[a] + hi;
        ln[a] = tl[a] - kBlock * hi;
      }
      const int gb = encode_local<D>(q, P);
      const int node = D == 3 ? ln[0] + 4 * (ln[1] + 4 * ln[2]) : ln[0] + 4 * ln[1];
      const float4 g = grid[(int64_t)gb * NPB + node];
      tile_x[i] = g.x;
      tile_y[i] = g.y;
      tile_z[i] = g.z;
    }
    lds_barrier();
    int b[3];
#pragma unroll
    for (int a = 0; a < D; a++) b[a] = (lbl[a] + P.blo[a]) * kBlock;
    // one particle per thread, kBatch particles of the chunk at a time
    for (int b0 = ck.y; b0 < ck.z; b0 += kBatch) {
      for (int i = threadIdx.x; i < (kThreads / 64) * 32; i += kThreads) (&dwcnt[0][0])[i] = 0;
      lds_barrier();
      const int j = b0 + threadIdx.x;
      const bool live = j < ck.z;
      int dir = 13;  // block offset of the particle after advection, (dz+1)*9 + (dy+1)*3 + (dx+1)
      int key_new = ck.x;
      if (live) {
        using PS = Particles<D, Q>;
        const int p = perm[j];
        const int mat = pi[(int64_t)L::MAT * cap + p];
        float xp[3] = {0.0f, 0.0f, 0.0f}, fx[3], w[3][3];
        int l[3];
        bool inside = true;
        PS::load_x(pf, cap, p, xp);
  #pragma unroll
        for (int a = 0; a < D; a++) {
          int base;
          base_fx(xp[a], P.inv_dx, base, fx[a]);
          bspline(fx[a], w[a]);
          l[a] = base - b[a];
          inside &= (l[a] >= 0) & (l[a] < kBlock);
        }
        // v = sum_o w_o g_o and C = 4 inv_dx sum_o w_o g_o (o - fx)^T
        // (mpm_solver.py:714-721), evaluated through the tensor-product
        // weights w_o = wx[ox] wy[oy] wz[oz]: P_e = sum_o w_o g_o (o_e - fx_e)
        // with the factor (o_e - fx_e) folded into axis e's weights (wd), so
        // C = 4 inv_dx P needs no subtraction of two nearly equal sums.  One
        // pass over the 27 nodes, per-axis partial sums instead of a 3x3 outer
        // product per node.
        float nv[3] = {0.0f, 0.0f, 0.0f}, pe[3][3] = {{0.0f}}, nC[3][3];
        float wd[3][3];
  #pragma unroll
        for (int a = 0; a < 3; a++)
  #pragma unroll
          for (int k = 0; k < 3; k++) wd[a][k] = a < D ? ((float)k - fx[a]) * w[a][k] : 0.0f;
        if (inside && !skip) {
  #pragma unroll
          for (int oz = 0; oz < (D == 3 ? 3 : 1); oz++) {
            float b0[3] = {0.0f, 0.0f, 0.0f}, bx[3] = {0.0f, 0.0f, 0.0f}, by[3] = {0.0f, 0.0f, 0.0f};
  #pragma unroll
            for (int oy = 0; oy < 3; oy++) {
              float a0[3] = {0.0f, 0.0f, 0.0f}, ax[3] = {0.0f, 0.0f, 0.0f};
  #pragma unroll
              for (int ox = 0; ox < 3; ox++) {
                const int node = D == 3 ? (l[0] + ox) + kTile * ((l[1] + oy) + kTile * (l[2] + oz))
                                        : (l[0] + ox) + kTile * (l[1] + oy);
                const float gg[3] = {tile_x[node], tile_y[node], D == 3 ? tile_z[node] : 0.0f};
  #pragma unroll
                for (int c = 0; c < D; c++) {
                  const float g = gg[c];
                  a0[c] += w[0][ox] * g;
                  ax[c] += wd[0][ox] * g;
                }
              }
  #pragma unroll
              for (int c = 0; c < D; c++) {
                b0[c] += w[1][oy] * a0[c];
                bx[c] += w[1][oy] * ax[c];
                by[c] += wd[1][oy] * a0[c];
              }
            }
            const float wz = D == 3 ? w[2][oz] : 1.0f;
  #pragma unroll
            for (int c = 0; c < D; c++) {
              nv[c] += wz * b0[c];
              pe[c][0] += wz * bx[c];
              pe[c][1] += wz * by[c];
              if (D == 3) pe[c][2] += wd[2][oz] * b0[c];
            }
          }
        }
  #pragma unroll
        for (int c = 0; c < D; c++)
  #pragma unroll
          for (int e = 0; e < D; e++) nC[c][e] = P.c4 * pe[c][e];
        float xn[3] = {xp[0], xp[1], xp[2]};
        if (skip) {
        } else if (mat != kStationary) {  // mpm_solver.py:722-724
          PS::store_v(pf_out, cap, p, nv);  // x advances by the stored v
  #pragma unroll
          for (int a = 0; a < D; a++) {
            xn[a] = xp[a] + P.dt * nv[a];
            vmax = fmaxf(vmax, fabsf(nv[a]));
          }
          PS::store_x(pf_out, cap, p, xn);  // xn: the stored x (binned next)
          PS::store_C(pf_out, cap, p, nC);
        } else {
          float vs[3];
          PS::load_v(pf, cap, p, vs);
  #pragma unroll
          for (int a = 0; a < D; a++) {
            const float va = vs[a];
            vmax = fmaxf(vmax, fabsf(va));
            if (!Q && pf_out != pf) {  // out of place: a stationary particle keeps x, v, C
              pf_out[(int64_t)(L::V + a) * cap + p] = va;
              pf_out[(int64_t)(L::X + a) * cap + p] = xp[a];
  #pragma unroll
              for (int e = 0; e < D; e++)
                pf_out[(int64_t)(L::C + a * D + e) * cap + p] = pf[(int64_t)(L::C + a * D + e) * cap + p];
            }
          }
        }
        // block of the new position for the next substep's binning
        // (build_pid, mpm_solver.py:357-361); particles move less than one block
        // per substep under any sane CFL number
        int dd[3] = {0, 0, 0}, q[3] = {0, 0, 0};
        bool near = true;
  #pragma unroll
        for (int a = 0; a < D; a++) {
          int base;
          float f;
          base_fx(xn[a], P.inv_dx, base, f);
          q[a] = (base >> 2) - P.blo[a];  // scene-local block
          dd[a] = q[a] - lbl[a];
          near &= (dd[a] >= -1) & (dd[a] <= 1) & (q[a] >= 0) & (q[a] < P.bnl[a] - 1);
        }
        if (near) {
          dir = (dd[2] + 1) * 9 + (dd[1] + 1) * 3 + (dd[0] + 1);
          q[D - 1] += P.zoff;
          key_new = encode_local<D>(q, P);
        } else {
          // left the window (bit 0) or moved more than one block (bit 3): the
          // particle stays binned in its old block and the caller is told
          atomicOr(err, (q[0] < 0 || q[0] >= P.bnl[0] - 1 || q[1] < 0 || q[1] >= P.bnl[1] - 1 ||
                         (D == 3 && (q[2] < 0 || q[2] >= P.bnl[2] - 1)))
                            ? 1
                            : 8);
        }
      }
      // stable rank of the particle among the chunk's particles with the same
      // direction; per-chunk direction counts; block counts of the next substep
      unsigned long long peers = __ballot(live);
  #pragma unroll
      for (int bit = 0; bit < 5; bit++) {
        const bool on = (dir >> bit) & 1;
        const unsigned long long bal = __ballot(on);
        peers &= on ? bal : ~bal;
      }
      const int rank = (int)__popcll(peers & lt);
      if (live && rank == 0) dwcnt[wave][dir] = (int)__popcll(peers);
      lds_barrier();
      if (threadIdx.x < 27) {
        const int d = threadIdx.x;
        int run = dbase[d];
  #pragma unroll
        for (int w = 0; w < kThreads / 64; w++) {
          const int cnt = dwcnt[w][d];
          dwcnt[w][d] = run;
          run += cnt;
        }
        dbase[d] = run;
      }
      lds_barrier();
      if (live) {
        code[j] = dir | ((dwcnt[wave][dir] + rank) << 8);
        nkey[j] = key_new;
      }
      lds_barrier();
    }  // batches of the chunk
    // per-chunk direction counts; block counts of the next substep
    if (threadIdx.x < 27) {
      const int d = threadIdx.x;
      const int run = dbase[d];
      dcount[(int64_t)c * 27 + d] = run;
      if (run > 0) {
        int q[3] = {lb[0] + d % 3 - 1, lb[1] + (d / 3) % 3 - 1, lb[2] + d / 9 - 1};
        atomicAdd(&count_next[encode_local<D>(q, P)], run);
      }
    }
    lds_barrier();
  }
  // out of place: particles outside every chunk (escaped the window; sorted
  // to the end) keep their whole state
  if (!Q && pf_out != pf) {
    for (int j = totals[kTotValid] + blockIdx.x * kThreads + threadIdx.x; j < n;
         j += gridDim.x * kThreads) {
      const int p = perm[j];
      for (int c = 0; c < L::NF; c++) pf_out[(int64_t)c * cap + p] = pf[(int64_t)c * cap + p];
    }
  }
  // compute_max_velocity (mpm_solver.py:726-735): non-negative floats order
  // like their bit patterns.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmax;
  lds_barrier();
  if (threadIdx.x == 0) {
    float m = red[0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; w++) m = fmaxf(m, red[w]);
    // one partial maximum per workgroup, reduced by the reader: a single
    // device-wide atomic serialised thousands of workgroups
    maxv[(int64_t)(step ? *step : 0) * stride + blockIdx.x] = __float_as_uint(m);
  }
}

// ---------------------------------------------------------------------------
// neighbour-bounded counting sort (steady-state binning)
//
// The order of a substep is the previous order stably sorted by the new
// block key.  Particles move less than one block per substep, so the members
// of new block k come from the 3^d old blocks k - d; in the previous order
// they are grouped by old block (ascending key), then by old chunk, then by
// position.  k_g2p already gave every particle its direction d and its
// stable rank among its old chunk's particles with the same d; k_offsets
// walks each new block's 3^d old neighbours in ascending key order and hands
// every (old chunk, d) pair its first output position; k_nscatter moves the
// particle indices.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void k_offsets(
    const Params* __restrict__ scenes, const int4* __restrict__ chunks,
    const int* __restrict__ totals,
    const int* __restrict__ pstart, const int* __restrict__ count,
    const int* __restrict__ pchunk_begin_prev, const int* __restrict__ pchunk_cnt_prev,
    const int* __restrict__ dcount, int* __restrict__ doff, int* __restrict__ err) {
  // one wave per new chunk; the wave of a block's first chunk does the block:
  // lane t < 3^d takes old neighbour t in ascending key order, sums its
  // chunks' counts toward this block, and a wave scan orders the neighbours
  const Params& P = scenes[0];  // the combined window extents
  const int nch = totals[kTotChunks];
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (kThreads / 64);
  for (int c = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); c < nch; c += waves) {
    const int4 ck = chunks[c];
    const int bid = ck.x;
    if (ck.y != pstart[bid]) continue;  // not the first chunk of its block (wave-uniform)
    int lb[3];
    decode_local<D>(bid, P, lb);
    constexpr int NB = D == 3 ? 27 : 9;
    int c0 = 0, cn = 0, dir = 0, mine = 0;
    if (lane < NB) {
      // ascending old key <=> descending offset (dz, dy, dx) in lexicographic order
      const int dx = 1 - lane % 3, dy = 1 - (lane / 3) % 3, dz = D == 3 ? 1 - lane / 9 : 0;
      const int q[3] = {lb[0] - dx, lb[1] - dy, lb[2] - dz};
      if (q[0] >= 0 && q[0] < P.bn[0] && q[1] >= 0 && q[1] < P.bn[1] &&
          (D == 2 || (q[2] >= 0 && q[2] < P.bn[2]))) {
        const int ob = encode_local<D>(q, P);
        c0 = pchunk_begin_prev[ob];
        cn = pchunk_cnt_prev[ob];
        dir = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
        for (int oc = c0; oc < c0 + cn; oc++) mine += dcount[(int64_t)oc * 27 + dir];
      }
    }
    int inc = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    int run = ck.y + inc - mine;
    for (int oc = c0; oc < c0 + cn; oc++) {
      doff[(int64_t)oc * 27 + dir] = run;
      run += dcount[(int64_t)oc * 27 + dir];
    }
    const int total = __shfl(inc, 63);
    if (lane == 0 && total != count[bid]) atomicOr(err, 16);  // counts disagree (cannot happen)
  }
}

__global__ __launch_bounds__(kThreads) void k_nscatter(
    const int4* __restrict__ chunks_prev, const int* __restrict__ totals_prev,
    const int* __restrict__ code, const int* __restrict__ nkey, const int* __restrict__ doff,
    const int* __restrict__ perm_old, int* __restrict__ perm_new, int* __restrict__ keys_new,
    int n, int total_blocks, int* __restrict__ err) {
  const int nch = totals_prev[kTotChunks];
  for (int c = blockIdx.x; c < nch; c += gridDim.x) {
    const int4 ck = chunks_prev[c];
    for (int j = ck.y + threadIdx.x; j < ck.z; j += kThreads) {
      const int cd = code[j];
      const int pos = doff[(int64_t)c * 27 + (cd & 255)] + (cd >> 8);
      if (pos >= 0 && pos < n) {
        perm_new[pos] = perm_old ? perm_old[j] : j;
        keys_new[pos] = nkey[j];
      } else {
        atomicOr(err, 16);
      }
    }
  }
  // particles outside the window (frozen since an earlier full sort) keep
  // their places at the end
  for (int j = totals_prev[kTotValid] + blockIdx.x * kThreads + threadIdx.x; j < n;
       j += gridDim.x * kThreads) {
    perm_new[j] = perm_old ? perm_old[j] : j;
    keys_new[j] = total_blocks;
  }
}

// ---------------------------------------------------------------------------
// particle bookkeeping kernels
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void k_reorder(const float* __restrict__ pf,
                                                      float* __restrict__ pf2,
                                                      const int* __restrict__ pi,
                                                      int* __restrict__ pi2,
                                                      const int* __restrict__ perm, int n,
                                                      int64_t cap) {
  using L = Layout<D>;
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const int p = perm[j];
#pragma unroll
    for (int c = 0; c < L::NF; c++) pf2[(int64_t)c * cap + j] = pf[(int64_t)c * cap + p];
#pragma unroll
    for (int c = 0; c < L::NI; c++) pi2[(int64_t)c * cap + j] = pi[(int64_t)c * cap + p];
  }
}

// Packed layout: the word count depends on the mode, so it is a parameter.
__global__ __launch_bounds__(kThreads) void k_reorder_words(const float* __restrict__ pf,
                                                            float* __restrict__ pf2,
                                                            const int* __restrict__ pi,
                                                            int* __restrict__ pi2,
                                                            const int* __restrict__ perm, int n,
                                                            int64_t cap, int nw, int ni) {
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const int p = perm[j];
#pragma unroll 4
    for (int c = 0; c < nw; c++) pf2[(int64_t)c * cap + j] = pf[(int64_t)c * cap + p];
    for (int c = 0; c < ni; c++) pi2[(int64_t)c * cap + j] = pi[(int64_t)c * cap + p];
  }
}

// New particles: seed_particle (mpm_solver.py:823-838).
template <int D, bool Q, bool G>
__global__ __launch_bounds__(kThreads) void k_append(const float* __restrict__ x,
                                                     const float* __restrict__ v,
                                                     const int* __restrict__ mat, int k,
                                                     int n_old, int scene, int plastic,
                                                     float* __restrict__ pf,
                                                     int* __restrict__ pi, int64_t cap) {
  using L = Layout<D>;
  using PS = Particles<D, Q, G>;
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= k) return;
  const int s = n_old + t;
  const int m = mat[t];
  float xs[3] = {0.0f, 0.0f, 0.0f}, vs[3] = {0.0f, 0.0f, 0.0f}, C[3][3], F[3][3];
#pragma unroll
  for (int a = 0; a < D; a++) {
    xs[a] = x[(int64_t)t * D + a];
    vs[a] = v ? v[(int64_t)t * D + a] : 0.0f;
#pragma unroll
    for (int e = 0; e < D; e++) {
      C[a][e] = 0.0f;
      F[a][e] = a == e ? 1.0f : 0.0f;
    }
  }
  PS::store_x(pf, cap, s, xs);
  PS::store_v(pf, cap, s, vs);
  PS::store_C(pf, cap, s, C);
  PS::store_F(pf, cap, s, F);
  if (!Q || plastic) PS::store_Jp(pf, cap, s, m == kSand ? 0.0f : 1.0f);  // mpm_solver.py:831-835
  pi[(int64_t)L::MAT * cap + s] = m;
  pi[(int64_t)L::ORIG * cap + s] = s;
  pi[(int64_t)L::SCENE * cap + s] = scene;
}

template <int D, bool G>
__global__ __launch_bounds__(kThreads) void k_unpack(const float* __restrict__ packed,
                                                     float* __restrict__ flt, int n, int64_t cap,
                                                     int plastic) {
  using PQ = Particles<D, true, G>;
  using PF = Particles<D, false>;
  for (int p = blockIdx.x * kThreads + threadIdx.x; p < n; p += gridDim.x * kThreads) {
    float x[3], v[3], C[3][3], F[3][3];
    PQ::load_x(packed, cap, p, x);
    PQ::load_v(packed, cap, p, v);
    PQ::load_C(packed, cap, p, C);
    PQ::load_F(packed, cap, p, F);
    PF::store_x(flt, cap, p, x);
    PF::store_v(flt, cap, p, v);
    PF::store_C(flt, cap, p, C);
    PF::store_F(flt, cap, p, F);
    PF::store_Jp(flt, cap, p, plastic ? PQ::load_Jp(packed, cap, p) : 1.0f);
  }
}

template <int D, bool G>
__global__ __launch_bounds__(kThreads) void k_pack(const float* __restrict__ flt,
                                                   float* __restrict__ packed, int n, int64_t cap,
                                                   int plastic) {
  using PQ = Particles<D, true, G>;
  using PF = Particles<D, false>;
  for (int p = blockIdx.x * kThreads + threadIdx.x; p < n; p += gridDim.x * kThreads) {
    float x[3], v[3], C[3][3], F[3][3];
    PF::load_x(flt, cap, p, x);
    PF::load_v(flt, cap, p, v);
    PF::load_C(flt, cap, p, C);
    PF::load_F(flt, cap, p, F);
    PQ::store_x(packed, cap, p, x);
    PQ::store_v(packed, cap, p, v);
    PQ::store_C(packed, cap, p, C);
    PQ::store_F(packed, cap, p, F);
    if (plastic) PQ::store_Jp(packed, cap, p, PF::load_Jp(flt, cap, p));
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather(const T* __restrict__ src, int64_t cap,
                                                     int ncomp, const int* __restrict__ orig,
                                                     int n, T* __restrict__ dst) {
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const int o = orig[j];
    for (int c = 0; c < ncomp; c++) dst[(int64_t)o * ncomp + c] = src[(int64_t)c * cap + j];
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter(const T* __restrict__ src, int64_t cap,
                                                      int ncomp, const int* __restrict__ orig,
                                                      int n, T* __restrict__ dst) {
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const int o = orig[j];
    for (int c = 0; c < ncomp; c++) dst[(int64_t)c * cap + j] = src[(int64_t)o * ncomp + c];
  }
}

__global__ __launch_bounds__(kThreads) void k_copy_segments(CopySegments S) {
  for (int s = 0; s < S.count; s++)
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < S.n[s];
         i += (int64_t)gridDim.x * kThreads)
      S.dst[s][i] = S.src[s][i];
}

__global__ __launch_bounds__(kThreads) void k_fill_u32(unsigned* __restrict__ p, unsigned v,
                                                       int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
// Workgroups of the grid operators (lanes per node x source block): enough
// that a large window needs few grid-stride iterations per lane, while the
// idle workgroups of a sparse window stay cheap.
constexpr int kGridOpGroups = 16384;

static inline int clamp_groups(int64_t g, int cap) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

struct TimerScope {
  Timer* t;
  hipStream_t st;
  TimerScope(Timer* t_, const char* name, hipStream_t s) : t(t_), st(s) {
    if (t) t->begin(name, st);
  }
  ~TimerScope() {
    if (t) t->end(st);
  }
};

// binning of one substep: block counts, the dense-window scan (chunks,
// active blocks) and the sorted order
template <int D>
static void launch_binning(const Params* P, SubstepBuffers& B, hipStream_t st, Timer* tm) {
  const int n = B.n;
  const int G = B.persistent_groups;
  if (n == 0) return;
  const int nparts = div_up(B.total_blocks, kScanTile);
  const int passes = (B.key_bits + 7) / 8;
  int* ghist = B.ghist + (B.parity & 1) * kMaxRadixPasses * kRadixBins;
  int* ghist_next = B.ghist + ((B.parity & 1) ^ 1) * kMaxRadixPasses * kRadixBins;
  const int chunk_bound = std::min(B.total_blocks, n) + div_up(n, B.chunk);
  if (B.full_sort) {
    // first substep (or positions changed from the host): block counts from
    // scratch, then a stable radix sort of the slots
    TimerScope ts(tm, "k_keys", st);
    launch_fill_u32((unsigned*)B.count, 0u, B.total_blocks + 1, st);
    if (B.quant)
      k_keys<D, true><<<clamp_groups(div_up(n, kThreads), 8192), kThreads, 0, st>>>(
          P, B.pf, B.pi, B.cap, n, B.total_blocks, B.keys, B.count, B.err);
    else
      k_keys<D, false><<<clamp_groups(div_up(n, kThreads), 8192), kThreads, 0, st>>>(
          P, B.pf, B.pi, B.cap, n, B.total_blocks, B.keys, B.count, B.err);
  }
  {
    TimerScope ts(tm, "k_bscan", st);
    const bool two_level = nparts > 256;
    if (two_level) {
      k_bscan_reduce<D><<<nparts, kThreads, 0, st>>>(P, B.count, B.total_blocks, B.chunk,
                                                     B.tile_prefix);
      k_bscan_top<<<1, kThreads, 0, st>>>(B.tile_prefix, nparts);
    }
    k_bscan<D><<<nparts, kThreads, 0, st>>>(
        P, B.count, B.count_next, B.total_blocks, B.chunk, nparts, B.full_sort ? passes : 0,
        B.pstart, B.pchunk_begin, B.pchunk_cnt, B.chunks, B.active, B.totals, ghist, ghist_next,
        B.lb_status_scan, B.lbs, two_level ? B.tile_prefix : nullptr, B.err, B.step);
  }
  if (B.full_sort) {
    TimerScope ts(tm, "k_radix", st);
    const int ntiles = div_up(n, kRadixTile);
    const int* kin = B.keys;
    const int* vin = nullptr;
    int* kbuf[2] = {B.keys_alt, B.keys_alt + B.cap};
    int* vbuf[2] = {B.perm, B.perm_alt};
    for (int pass = 0; pass < passes; pass++) {
      k_onesweep<<<ntiles, kThreads, 0, st>>>(kin, vin, kbuf[pass & 1], vbuf[pass & 1], n,
                                               pass * 8, ntiles, ghist + pass * kRadixBins,
                                               B.lb_status_radix, B.lbs, B.err);
      kin = kbuf[pass & 1];
      vin = vbuf[pass & 1];
    }
    if (vin != B.perm) std::swap(B.perm, B.perm_alt);
    B.keys_sorted = (int*)kin;
  } else {
    TimerScope ts(tm, "k_nsort", st);
    k_offsets<D><<<clamp_groups(div_up(chunk_bound, kThreads / 64), G), kThreads, 0, st>>>(
        P, B.chunks, B.totals, B.pstart, B.count, B.pchunk_begin_prev, B.pchunk_cnt_prev,
        B.dcount, B.doff, B.err);
    // the previous order: the last substep's permutation, or the slots
    // themselves right after a physical reorder
    const int* perm_old = B.perm_identity ? nullptr : B.perm;
    k_nscatter<<<clamp_groups(chunk_bound, G), kThreads, 0, st>>>(
        B.chunks_prev, B.totals_prev, B.code, B.nkey, B.doff, perm_old, B.perm_alt, B.keys_alt,
        n, B.total_blocks, B.err);
    std::swap(B.perm, B.perm_alt);
    B.keys_sorted = B.keys_alt;
  }
}

// binning, p2g and the grid operator of one substep
template <int D>
void launch_front(const Params* P, const int* slab, SubstepBuffers& B, int raw, hipStream_t st,
                  Timer* tm) {
  launch_binning<D>(P, B, st, tm);
  launch_transfer<D>(P, slab, B, raw, st, tm);
}

// one substep of the use_g2p2g mode: binning, the fused transfer from
// B.grid_in, and the grid operator into B.grid
template <int D>
static void launch_g2p2g(const Params* P, const int* slab, SubstepBuffers& B, int s_index,
                         hipStream_t st, Timer* tm) {
  const int n = B.n;
  const int G = B.persistent_groups;
  if (n == 0) return;
  launch_binning<D>(P, B, st, tm);
  const int chunk_bound = std::min(B.total_blocks, n) + div_up(n, B.chunk);
  {
    TimerScope ts(tm, "k_g2p2g", st);
    auto kern = B.quant ? k_g2p2g<D, true> : k_g2p2g<D, false>;
    kern<<<clamp_groups(chunk_bound, G), kThreads, 0, st>>>(
        P, slab, B.pf, B.pi, B.cap, B.perm, B.chunks, B.totals, B.grid_in, B.tiles, B.n_final,
        B.maxv, B.step, B.persistent_groups, B.code, B.nkey, B.dcount, B.count_next, B.err);
  }
  {
    TimerScope ts(tm, "k_grid_op", st);
    const int64_t node_bound = (int64_t)B.total_blocks * Geo<D>::kNodesPerBlock;
    k_grid_op_g2p2g<D><<<clamp_groups(div_up(node_bound * (D == 3 ? 8 : 4), kThreads), kGridOpGroups), kThreads,
                         0, st>>>(
        P, slab, B.active, B.totals, B.pchunk_begin, B.pchunk_cnt, B.tiles, B.grid);
  }
  std::swap(B.count, B.count_next);
}

// p2g and the grid operator on an existing binning
template <int D>
void launch_transfer(const Params* P, const int* slab, SubstepBuffers& B, int raw, hipStream_t st,
                     Timer* tm) {
  const int n = B.n;
  const int G = B.persistent_groups;
  if (n == 0) return;
  const int chunk_bound = std::min(B.total_blocks, n) + div_up(n, B.chunk);
  {
    TimerScope ts(tm, "k_p2g", st);
    auto kern = B.quant ? k_p2g<D, 0, true> : k_p2g<D, 0, false>;
    kern<<<clamp_groups(chunk_bound, G), kThreads, 0, st>>>(
        P, slab, B.pf, B.pi, B.cap, B.perm, B.chunks, B.totals, B.tiles, nullptr, B.pf_out);
  }
  {
    TimerScope ts(tm, "k_grid_op", st);
    const int64_t node_bound = (int64_t)B.total_blocks * Geo<D>::kNodesPerBlock;
    k_grid_op<D, 0><<<clamp_groups(div_up(node_bound * (D == 3 ? 8 : 4), kThreads), kGridOpGroups), kThreads, 0, st>>>(
        P, slab, B.active, B.totals, B.pchunk_begin, B.pchunk_cnt, B.tiles, B.grid, raw, nullptr,
        B.raw_out);
  }
}

// g2p of one substep (after launch_front); the max-velocity slot is the
// device substep counter B.step (s_index is kept for the interface)
template <int D>
void launch_g2p(const Params* P, const int* slab, SubstepBuffers& B, int s_index, hipStream_t st,
                Timer* tm) {
  const int n = B.n;
  const int G = B.persistent_groups;
  if (n == 0) return;
  const int chunk_bound = std::min(B.total_blocks, n) + div_up(n, B.chunk);
  {
    TimerScope ts(tm, "k_g2p", st);
    auto kern = B.g2p_waves >= 6 ? (B.quant ? k_g2p<D, true, 6> : k_g2p<D, false, 6>)
                                 : (B.quant ? k_g2p<D, true, 1> : k_g2p<D, false, 1>);
    kern<<<clamp_groups(chunk_bound, G), kThreads, 0, st>>>(
        P, slab, B.pf, B.pi, B.cap, B.perm, B.chunks, B.totals, B.grid,
        B.maxv, B.step, B.persistent_groups, B.code, B.nkey, B.dcount, B.count_next, B.err,
        B.pf_out, n);
  }
  std::swap(B.count, B.count_next);
}

template <int D>
void launch_substep(const Params* P, const int* slab, SubstepBuffers& B, int s_index, int raw,
                    hipStream_t st, Timer* tm) {
  if (B.g2p2g) {
    launch_g2p2g<D>(P, slab, B, s_index, st, tm);
    return;
  }
  launch_front<D>(P, slab, B, raw, st, tm);
  launch_g2p<D>(P, slab, B, s_index, st, tm);
}

template <int D>
void launch_substep_half(const Params* P, const int* slab, SubstepBuffers& B, int s_index,
                         int half, hipStream_t st, Timer* tm) {
  if (B.g2p2g) {
    if (half == 0) launch_g2p2g<D>(P, slab, B, s_index, st, tm);
    return;
  }
  if (half == 0)
    launch_front<D>(P, slab, B, 0, st, tm);
  else
    launch_g2p<D>(P, slab, B, s_index, st, tm);
}

template <int D>
void launch_adjoint(const Params* P, const int* slab, SubstepBuffers& B, const float* pf_in,
                    const float* gout, float* gin, float* gE, float4* gbar, float* abuf,
                    float* nu_part, float* gnu, hipStream_t st, Timer* tm) {
  const int n = B.n;
  const int G = B.persistent_groups;
  const int chunk_bound = n > 0 ? std::min(B.total_blocks, n) + div_up(n, B.chunk) : 1;
  const int groups = clamp_groups(chunk_bound, G);
  const int pgroups = clamp_groups(div_up(n, kThreads), G);
  if (n > 0) {
    {
      TimerScope ts(tm, "k_adj_affine", st);
      k_adj_affine<D><<<pgroups, kThreads, 0, st>>>(P, slab, pf_in, B.pi, B.cap, B.perm,
                                                    B.totals, abuf);
    }
    {
      TimerScope ts(tm, "k_g2p_adj_scatter", st);
      k_p2g<D, 1><<<groups, kThreads, 0, st>>>(P, slab, const_cast<float*>(pf_in), B.pi, B.cap,
                                               B.perm, B.chunks, B.totals, B.tiles, gout,
                                               nullptr);
    }
    {
      TimerScope ts(tm, "k_grid_op_adj", st);
      const int64_t node_bound = (int64_t)B.total_blocks * Geo<D>::kNodesPerBlock;
      k_grid_op<D, 1><<<clamp_groups(div_up(node_bound * (D == 3 ? 8 : 4), kThreads), kGridOpGroups), kThreads, 0, st>>>(
          P, slab, B.active, B.totals, B.pchunk_begin, B.pchunk_cnt, B.tiles, B.grid, 0, gbar,
          nullptr);
    }
    {
      TimerScope ts(tm, "k_p2g_adj", st);
      k_p2g_adj<D><<<groups, kThreads, 0, st>>>(P, slab, pf_in, B.pi, B.cap, B.perm, B.chunks,
                                                B.totals, B.grid, gbar, gout, abuf, gin, gE, n);
    }
    {
      TimerScope ts(tm, "k_adj_constitutive", st);
      k_adj_constitutive<D><<<pgroups, kThreads, 0, st>>>(P, slab, pf_in, B.pi, B.cap, B.perm,
                                                          B.totals, gout, gin, gE, nu_part);
    }
  } else {
    launch_fill_u32((unsigned*)nu_part, 0u, pgroups, st);
  }
  k_sum_partials<<<1, kThreads, 0, st>>>(nu_part, pgroups, gnu);
}

template <int D>
void launch_reorder(SubstepBuffers& B, float* pf_alt, int* pi_alt, hipStream_t st, Timer* tm) {
  TimerScope ts(tm, "k_reorder", st);
  if (B.quant)
    k_reorder_words<<<clamp_groups(div_up(B.n, kThreads), 8192), kThreads, 0, st>>>(
        B.pf, pf_alt, B.pi, pi_alt, B.perm, B.n, B.cap, B.nw, Layout<D>::NI);
  else
    k_reorder<D><<<clamp_groups(div_up(B.n, kThreads), 8192), kThreads, 0, st>>>(
        B.pf, pf_alt, B.pi, pi_alt, B.perm, B.n, B.cap);
}

template <int D>
void launch_append(SubstepBuffers& B, const float* x, const float* v, const int* mat, int k,
                   int n_old, int scene, int plastic, hipStream_t st) {
  if (k <= 0) return;
  const int g = div_up(k, kThreads);
  if (!B.quant)
    k_append<D, false, false><<<g, kThreads, 0, st>>>(x, v, mat, k, n_old, scene, plastic, B.pf,
                                                       B.pi, B.cap);
  else if (B.g2p2g)
    k_append<D, true, true><<<g, kThreads, 0, st>>>(x, v, mat, k, n_old, scene, plastic, B.pf,
                                                     B.pi, B.cap);
  else
    k_append<D, true, false><<<g, kThreads, 0, st>>>(x, v, mat, k, n_old, scene, plastic, B.pf,
                                                      B.pi, B.cap);
}

template <int D>
void launch_unpack(const float* packed, float* flt, int n, int64_t cap, bool g2p2g, bool plastic,
                   hipStream_t st) {
  if (n <= 0) return;
  const int g = clamp_groups(div_up(n, kThreads), 8192);
  if (g2p2g)
    k_unpack<D, true><<<g, kThreads, 0, st>>>(packed, flt, n, cap, plastic);
  else
    k_unpack<D, false><<<g, kThreads, 0, st>>>(packed, flt, n, cap, plastic);
}

template <int D>
void launch_pack(const float* flt, float* packed, int n, int64_t cap, bool g2p2g, bool plastic,
                 hipStream_t st) {
  if (n <= 0) return;
  const int g = clamp_groups(div_up(n, kThreads), 8192);
  if (g2p2g)
    k_pack<D, true><<<g, kThreads, 0, st>>>(flt, packed, n, cap, plastic);
  else
    k_pack<D, false><<<g, kThreads, 0, st>>>(flt, packed, n, cap, plastic);
}

void launch_gather_f(const float* src, int64_t cap, int ncomp, const int* orig, int n,
                     float* dst, hipStream_t st) {
  if (n > 0)
    k_gather<float><<<clamp_groups(div_up(n, kThreads), 8192), kThreads, 0, st>>>(src, cap, ncomp,
                                                                                 orig, n, dst);
}
void launch_gather_i(const int* src, int64_t cap, int ncomp, const int* orig, int n, int* dst,
                     hipStream_t st) {
  if (n > 0)
    k_gather<int><<<clamp_groups(div_up(n, kThreads), 8192), kThreads, 0, st>>>(src, cap, ncomp,
                                                                             orig, n, dst);
}
void launch_scatter_f(const float* src, int64_t cap, int ncomp, const int* orig, int n,
                      float* dst, hipStream_t st) {
  if (n > 0)
    k_scatter<float><<<clamp_groups(div_up(n, kThreads), 8192), kThreads, 0, st>>>(src, cap, ncomp,
                                                                                  orig, n, dst);
}
void launch_scatter_i(const int* src, int64_t cap, int ncomp, const int* orig, int n, int* dst,
                      hipStream_t st) {
  if (n > 0)
    k_scatter<int><<<clamp_groups(div_up(n, kThreads), 8192), kThreads, 0, st>>>(src, cap, ncomp,
                                                                              orig, n, dst);
}
void launch_copy_segments(const CopySegments& S, hipStream_t st) {
  int64_t most = 0;
  for (int s = 0; s < S.count; s++) most = std::max(most, S.n[s]);
  if (most > 0)
    k_copy_segments<<<clamp_groups(div_up(most, kThreads), 2048), kThreads, 0, st>>>(S);
}

// The device SVD (svd3 / svd2, the replacement of ti.svd) on n matrices,
// for tests: A, U, V (n, dim, dim) row-major, sig (n, dim).
__global__ __launch_bounds__(kThreads) void k_svd_batch(int dim, const float* __restrict__ A,
                                                        float* __restrict__ U,
                                                        float* __restrict__ sig,
                                                        float* __restrict__ V, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    float a[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    float u[3][3], v[3][3], s3[3] = {0.0f, 0.0f, 0.0f};
    const int dd = dim * dim;
    for (int r = 0; r < dim; r++)
      for (int c = 0; c < dim; c++) a[r][c] = A[i * dd + r * dim + c];
    if (dim == 3)
      svd3(a, u, s3, v);
    else
      svd2(a, u, s3, v);
    for (int r = 0; r < dim; r++) {
      sig[i * dim + r] = s3[r];
      for (int c = 0; c < dim; c++) {
        U[i * dd + r * dim + c] = u[r][c];
        V[i * dd + r * dim + c] = v[r][c];
      }
    }
  }
}

void launch_svd_batch(int dim, const float* A, float* U, float* sig, float* V, int64_t n,
                      hipStream_t st) {
  if (n > 0)
    k_svd_batch<<<clamp_groups(div_up(n, kThreads), 4096), kThreads, 0, st>>>(dim, A, U, sig, V,
                                                                            n);
}

void launch_fill_u32(unsigned* p, unsigned v, int64_t n, hipStream_t st) {
  if (n > 0)
    k_fill_u32<<<clamp_groups(div_up(n, kThreads), 4096), kThreads, 0, st>>>(p, v, n);
}

// Per-workgroup bounding box of the particle positions: partial[g * 6 + 2a]
// = min, partial[g * 6 + 2a + 1] = max of x_a (workgroup g).
template <int D, bool Q>
__global__ __launch_bounds__(kThreads) void k_aabb(const float* __restrict__ pf, int64_t cap,
                                                   int n, float* __restrict__ partial) {
  __shared__ float red[kThreads / 64][6];
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  for (int p = blockIdx.x * kThreads + threadIdx.x; p < n; p += gridDim.x * kThreads) {
    float x[3] = {0.0f, 0.0f, 0.0f};
    Particles<D, Q>::load_x(pf, cap, p, x);
#pragma unroll
    for (int a = 0; a < D; a++) {
      lo[a] = fminf(lo[a], x[a]);
      hi[a] = fmaxf(hi[a], x[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off));
    }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      red[wave][2 * a] = lo[a];
      red[wave][2 * a + 1] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    float v = red[0][c];
#pragma unroll
    for (int w = 1; w < kThreads / 64; w++) v = (c & 1) ? fmaxf(v, red[w][c]) : fminf(v, red[w][c]);
    partial[(int64_t)blockIdx.x * 6 + c] = v;
  }
}

void launch_aabb(int D, bool quant, const float* pf, int64_t cap, int n, float* partial,
                 int groups, hipStream_t st) {
  if (D == 3)
    (quant ? k_aabb<3, true> : k_aabb<3, false>)<<<groups, kThreads, 0, st>>>(pf, cap, n, partial);
  else
    (quant ? k_aabb<2, true> : k_aabb<2, false>)<<<groups, kThreads, 0, st>>>(pf, cap, n, partial);
}

// Grid contents carried into a moved block window: node i of new block b
// takes node i of the same global block in the old window, or zero.
__global__ __launch_bounds__(kThreads) void k_remap_grid(const float4* __restrict__ old_grid,
                                                         float4* __restrict__ new_grid,
                                                         WindowMap m, int64_t nodes, int npb) {
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < nodes;
       t += (int64_t)gridDim.x * kThreads) {
    const int64_t b = t / npb;
    const int node = (int)(t % npb);
    int64_t r = b;
    int64_t ob = 0, stride = 1;
    bool in = true;
    for (int a = 0; a < 3; a++) {
      const int c = (int)(r % m.new_bn[a]);
      r /= m.new_bn[a];
      const int oc = c + m.new_blo[a] - m.old_blo[a];
      in &= (oc >= 0) & (oc < m.old_bn[a]);
      ob += (int64_t)oc * stride;
      stride *= m.old_bn[a];
    }
    new_grid[t] = in ? old_grid[ob * npb + node] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

void launch_remap_grid(const float4* old_grid, float4* new_grid, const WindowMap& m,
                       int64_t nodes, int npb, hipStream_t st) {
  if (nodes > 0)
    k_remap_grid<<<clamp_groups(div_up(nodes, kThreads), 8192), kThreads, 0, st>>>(
        old_grid, new_grid, m, nodes, npb);
}

// compute_max_grid_velocity (mpm_solver.py:737-746): the largest velocity
// component magnitude on the grid after the last grid operator.  Nodes of
// active blocks without mass hold zero velocity, as the reference's
// freshly activated nodes do.  Non-negative floats order like their bits.
template <int D>
__global__ __launch_bounds__(kThreads) void k_grid_vmax(const float4* __restrict__ grid,
                                                        const int* __restrict__ active,
                                                        const int* __restrict__ totals,
                                                        unsigned* __restrict__ out) {
  constexpr int NPB = Geo<D>::kNodesPerBlock;
  const int64_t nodes = (int64_t)totals[kTotActive] * NPB;
  float m = 0.0f;
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < nodes;
       t += (int64_t)gridDim.x * kThreads) {
    const float4 g = grid[(int64_t)active[t / NPB] * NPB + t % NPB];
    m = fmaxf(m, fmaxf(fabsf(g.x), fabsf(g.y)));
    if (D == 3) m = fmaxf(m, fabsf(g.z));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

void launch_grid_vmax(int D, const float4* grid, const int* active, const int* totals,
                      int64_t total_blocks, unsigned* out, hipStream_t st) {
  launch_fill_u32(out, 0u, 1, st);
  const int npb = D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
  const int groups = clamp_groups(div_up(total_blocks * npb, kThreads), 2048);
  if (D == 3)
    k_grid_vmax<3><<<groups, kThreads, 0, st>>>(grid, active, totals, out);
  else
    k_grid_vmax<2><<<groups, kThreads, 0, st>>>(grid, active, totals, out);
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_grid_gather(const float4* __restrict__ grid,
                                                         const int* __restrict__ active,
                                                         int64_t count, NodeWindow w,
                                                         int* __restrict__ coords,
                                                         float* __restrict__ v,
                                                         float* __restrict__ m) {
  constexpr int NPB = Geo<D>::kNodesPerBlock;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * kThreads) {
    const int bid = active[i / NPB], l = (int)(i % NPB);
    const float4 g = grid[(int64_t)bid * NPB + l];
    const int lb0 = bid % w.bn[0];
    const int lb1 = D == 3 ? (bid / w.bn[0]) % w.bn[1] : bid / w.bn[0] - w.zoff;
    coords[i * D + 0] = 4 * (w.blo[0] + lb0) + (l & 3);
    coords[i * D + 1] = 4 * (w.blo[1] + lb1) + ((l >> 2) & 3);
    v[i * D + 0] = g.x;
    v[i * D + 1] = g.y;
    if (D == 3) {
      const int lb2 = bid / (w.bn[0] * w.bn[1]) - w.zoff;
      coords[i * D + 2] = 4 * (w.blo[2] + lb2) + (l >> 4);
      v[i * D + 2] = g.z;
    }
    m[i] = g.w;
  }
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_grid_scatter(float4* __restrict__ grid,
                                                          const int* __restrict__ active,
                                                          int64_t count,
                                                          const float* __restrict__ v) {
  constexpr int NPB = Geo<D>::kNodesPerBlock;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * kThreads) {
    float4* g = &grid[(int64_t)active[i / NPB] * NPB + i % NPB];
    g->x = v[i * D + 0];
    g->y = v[i * D + 1];
    if (D == 3) g->z = v[i * D + 2];
  }
}

void launch_grid_gather(int D, const float4* grid, const int* active, int64_t count,
                        const NodeWindow& w, int* coords, float* v, float* m, hipStream_t st) {
  if (count <= 0) return;
  const int groups = clamp_groups(div_up(count, kThreads), 4096);
  if (D == 3)
    k_grid_gather<3><<<groups, kThreads, 0, st>>>(grid, active, count, w, coords, v, m);
  else
    k_grid_gather<2><<<groups, kThreads, 0, st>>>(grid, active, count, w, coords, v, m);
}

void launch_grid_scatter(int D, float4* grid, const int* active, int64_t count, const float* v,
                         hipStream_t st) {
  if (count <= 0) return;
  const int groups = clamp_groups(div_up(count, kThreads), 4096);
  if (D == 3)
    k_grid_scatter<3><<<groups, kThreads, 0, st>>>(grid, active, count, v);
  else
    k_grid_scatter<2><<<groups, kThreads, 0, st>>>(grid, active, count, v);
}

template void launch_transfer<2>(const Params*, const int*, SubstepBuffers&, int, hipStream_t,
                                 Timer*);
template void launch_transfer<3>(const Params*, const int*, SubstepBuffers&, int, hipStream_t,
                                 Timer*);
template void launch_front<2>(const Params*, const int*, SubstepBuffers&, int, hipStream_t,
                              Timer*);
template void launch_front<3>(const Params*, const int*, SubstepBuffers&, int, hipStream_t,
                              Timer*);
template void launch_adjoint<2>(const Params*, const int*, SubstepBuffers&, const float*,
                                const float*, float*, float*, float4*, float*, float*, float*,
                                hipStream_t, Timer*);
template void launch_adjoint<3>(const Params*, const int*, SubstepBuffers&, const float*,
                                const float*, float*, float*, float4*, float*, float*, float*,
                                hipStream_t, Timer*);
template void launch_substep<2>(const Params*, const int*, SubstepBuffers&, int, int, hipStream_t,
                                Timer*);
template void launch_substep<3>(const Params*, const int*, SubstepBuffers&, int, int, hipStream_t,
                                Timer*);
template void launch_substep_half<2>(const Params*, const int*, SubstepBuffers&, int, int,
                                     hipStream_t, Timer*);
template void launch_substep_half<3>(const Params*, const int*, SubstepBuffers&, int, int,
                                     hipStream_t, Timer*);
template void launch_reorder<2>(SubstepBuffers&, float*, int*, hipStream_t, Timer*);
template void launch_reorder<3>(SubstepBuffers&, float*, int*, hipStream_t, Timer*);
template void launch_append<2>(SubstepBuffers&, const float*, const float*, const int*, int, int,
                               int, int, hipStream_t);
template void launch_append<3>(SubstepBuffers&, const float*, const float*, const int*, int, int,
                               int, int, hipStream_t);
template void launch_unpack<2>(const float*, float*, int, int64_t, bool, bool, hipStream_t);
template void launch_unpack<3>(const float*, float*, int, int64_t, bool, bool, hipStream_t);
template void launch_pack<2>(const float*, float*, int, int64_t, bool, bool, hipStream_t);
template void launch_pack<3>(const float*, float*, int, int64_t, bool, bool, hipStream_t);

}  // namespace mpm

// Launch interface of the substep kernels (implemented in mpm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.h"
#include "mpm_particles.h"

namespace mpm {

constexpr int kThreads = 256;
constexpr int kBatch = kThreads;  // particles a transfer workgroup holds at a time
constexpr int kChunkMax = 1024;  // particles per transfer work item (max), kBatch at a time
constexpr int kScanItems = 2;                       // per thread in the block scan
constexpr int kScanTile = kThreads * kScanItems;    // 512 blocks per workgroup
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kThreads * kRadixItems;  // 2048 keys per workgroup
constexpr int kRadixBins = 256;
constexpr int kMaxRadixPasses = 3;                  // block keys up to 24 bits

// Ticket / epoch state shared by the single-pass (look-back) scans.
struct LBState {
  unsigned long long tile_ctr;
  unsigned long long done_ctr;
  unsigned epoch;
  unsigned pad;
};

// Totals written on the device by the block scan (read by persistent kernels).
enum Total : int { kTotValid = 0, kTotChunks = 1, kTotActive = 2, kTotPBlocks = 3, kTotCount = 4 };

struct SubstepBuffers {
  float* pf;  // particle float components
  float* pf_out;  // where p2g / g2p write the new state (== pf: in place)
  int* pi;    // particle int components
  int64_t cap;
  int n;
  int* keys;        // [n] block key per slot
  int* keys_alt;    // radix ping-pong
  int* perm;        // [n] sorted slot order (radix output values)
  int* perm_alt;
  int* count;       // [total_blocks] particles per block (zero on entry)
  int* count_next;  // [total_blocks] zeroed for the next substep
  int* pstart;      // [total_blocks] first sorted position of the block
  int* pchunk_begin;
  int* pchunk_cnt;
  int4* chunks;     // (block, begin, end, 0)
  int* active;      // active grid blocks
  int4* partials;   // block-scan partial sums
  int* totals;      // kTot*
  int* ghist;       // radix digit histograms, 2 x [kMaxRadixPasses][256] (substep parity)
  unsigned long long* lb_status_scan;   // look-back status of k_bscan  [4 * tiles]
  unsigned long long* lb_status_radix;  // look-back status of k_onesweep [256 * tiles]
  LBState* lbs;
  float4* tiles;    // p2g per-chunk node tiles (momentum xyz, mass)
  float4* grid;     // dense block-major grid [total_blocks][nodes] of (v, mass)
  float4* raw_out;  // if set, the grid operator also stores the raw p2g sums here
  int* err;         // error bits
  unsigned* maxv;   // per-substep, per-g2p-workgroup max |v| (float bits)
  int* step;        // device substep counter selecting the maxv slot (null: slot 0);
                    // the block scan increments it, so it starts at -1
  int total_blocks;
  int key_bits;
  int chunk;
  int persistent_groups;
  int g2p_waves;  // g2p's occupancy target: 0 = the compiler's choice, 6 = six waves per SIMD
  int parity;  // substep counter parity (selects the digit-histogram copy)
  // neighbour-bounded sort: the previous substep's chunks and block tables,
  // direction codes / new keys per previous sorted position, per-chunk
  // direction counts and output offsets
  int4* chunks_prev;
  int* pchunk_begin_prev;
  int* pchunk_cnt_prev;
  int* totals_prev;
  int* code;
  int* nkey;
  int* dcount;      // [chunk_cap][27]
  int* doff;        // [chunk_cap][27]
  int4* tile_prefix;  // two-level block scan
  // use_g2p2g mode: the fused transfer reads grid_in and writes grid;
  // particles with an insertion index >= n_final skip the G2P
  bool g2p2g;
  float4* grid_in;
  int n_final;
  bool quant;         // packed particle layout (PackedLayout<D>)
  int nw;             // particle float-buffer words per slot
  bool full_sort;     // radix-sort the slots (first substep / positions set by the host)
  bool perm_identity; // the slots were just reordered into the previous order
  int* keys_sorted;   // output: block keys in sorted order
};

class Timer;

// P: device table of the scenes' parameters; slab: scene of each combined
// block coordinate along the stacking axis.
template <int D>
void launch_substep(const Params* P, const int* slab, SubstepBuffers& B, int substep_index,
                    int raw_grid, hipStream_t st, Timer* timer);

// One half of launch_substep for hosts that run their own grid
// post-processes between the halves: half 0 = binning, p2g and the grid
// operator (use_g2p2g: the whole fused substep), half 1 = g2p (use_g2p2g:
// nothing).  B carries over from half 0 to half 1.
template <int D>
void launch_substep_half(const Params* P, const int* slab, SubstepBuffers& B, int substep_index,
                         int half, hipStream_t st, Timer* timer);

// p2g + grid operator on the binning already in B
template <int D>
void launch_transfer(const Params* P, const int* slab, SubstepBuffers& B, int raw_grid,
                     hipStream_t st, Timer* timer);

// binning + p2g + grid operator only (raw_grid: keep the unnormalised sums)
template <int D>
void launch_front(const Params* P, const int* slab, SubstepBuffers& B, int raw_grid,
                  hipStream_t st, Timer* timer);

// Adjoint of one substep after launch_front(raw_grid = 1) on the same old
// state pf_in: gout = adjoints of the new state, gin = adjoints of the old
// state (both laid out like pf), gE (may be null) = adjoint of the
// per-particle Young's modulus, gnu[0] = adjoint of nu.  gbar holds
// total_blocks * nodes float4, abuf D*D * cap floats, nu_part
// persistent_groups floats.
template <int D>
void launch_adjoint(const Params* P, const int* slab, SubstepBuffers& B, const float* pf_in,
                    const float* gout, float* gin, float* gE, float4* gbar, float* abuf,
                    float* nu_part, float* gnu, hipStream_t st, Timer* timer);

template <int D>
void launch_reorder(SubstepBuffers& B, float* pf_alt, int* pi_alt, hipStream_t st, Timer* timer);

template <int D>
void launch_append(SubstepBuffers& B, const float* x, const float* v, const int* mat, int k,
                   int n_old, int scene, int plastic, hipStream_t st);
// Packed layout <-> float layout (Layout<D>, same cap) for the first n slots;
// unpacking fills absent components (C in use_g2p2g mode: 0, Jp without
// plasticity: 1).
template <int D>
void launch_unpack(const float* packed, float* flt, int n, int64_t cap, bool g2p2g, bool plastic,
                   hipStream_t st);
template <int D>
void launch_pack(const float* flt, float* packed, int n, int64_t cap, bool g2p2g, bool plastic,
                 hipStream_t st);

void launch_gather_f(const float* src, int64_t cap, int ncomp, const int* orig, int n,
                     float* dst, hipStream_t st);
void launch_gather_i(const int* src, int64_t cap, int ncomp, const int* orig, int n, int* dst,
                     hipStream_t st);
void launch_scatter_f(const float* src, int64_t cap, int ncomp, const int* orig, int n,
                      float* dst, hipStream_t st);
void launch_scatter_i(const int* src, int64_t cap, int ncomp, const int* orig, int n, int* dst,
                      hipStream_t st);
void launch_fill_u32(unsigned* p, unsigned v, int64_t n, hipStream_t st);
// svd3 / svd2 on n row-major dim x dim matrices (tests of the device SVD).
void launch_svd_batch(int dim, const float* A, float* U, float* sig, float* V, int64_t n,
                      hipStream_t st);

// Bounding box of the particle positions: groups x 6 partial (min, max) per axis.
void launch_aabb(int D, bool quant, const float* pf, int64_t cap, int n, float* partial,
                 int groups, hipStream_t st);
// A block window moved or resized: origin and extent (blocks, x fastest).
struct WindowMap {
  int old_blo[3], old_bn[3], new_blo[3], new_bn[3];
};
void launch_remap_grid(const float4* old_grid, float4* new_grid, const WindowMap& m,
                       int64_t nodes, int npb, hipStream_t st);
// out[0] = max over the nodes of the active blocks (count at totals[kTotActive])
// of max_a |v_a|, as float bits; out is cleared first.
void launch_grid_vmax(int D, const float4* grid, const int* active, const int* totals,
                      int64_t total_blocks, unsigned* out, hipStream_t st);

// Nodes of the active blocks (count at totals[kTotActive], NPB nodes each,
// node i = block active[i / NPB], local node i % NPB) <-> dense arrays:
// gather writes the global node coordinates coords[count * D] (block window
// origin blo, extent bn in blocks, the scene's offset zoff on the last
// axis), velocities v[count * D] and masses m[count]; scatter writes v back
// into the grid's velocity components (masses untouched).
struct NodeWindow {
  int blo[3], bn[3], zoff;
};
void launch_grid_gather(int D, const float4* grid, const int* active, int64_t count,
                        const NodeWindow& w, int* coords, float* v, float* m, hipStream_t st);
void launch_grid_scatter(int D, float4* grid, const int* active, int64_t count, const float* v,
                         hipStream_t st);

// Up to 8 int array copies in one launch.
struct CopySegments {
  const int* src[8];
  int* dst[8];
  int64_t n[8];
  int count;
};
void launch_copy_segments(const CopySegments& S, hipStream_t st);

// Quantised particle output (mpm_io.hip).  Components 0..D-1 are x, D..2D-1 v.
struct QuantRanges {
  float lo[6];
  float inv[6];  // 1 / (hi - lo), rounded to float32
};
void launch_minmax(int D, const float* pf, int64_t cap, int n, float* partial, int groups,
                   hipStream_t st);
void launch_quantize(int D, const float* pf, const int* orig, int64_t cap, int n,
                     const QuantRanges& q, unsigned* out, hipStream_t st);

// Mesh voxelization and seeding (mpm_mesh.hip).
struct VoxGrid {
  int res[3];  // voxels per axis
  int padding;
  double dx, inv_dx;  // voxel size
};
struct SeedParams {
  int sample_density;  // < 0: voxel counts only
  int ss3;             // super_sample^3
  int p_limit;         // first-axis voxel index bound (grid_size / 2 - padding)
  unsigned seed;
  float pdx;           // voxel size, float32
  float shift[3];      // translation
};
// events: res0 * res1 * (res2 + 1) ints, zeroed by the caller
void launch_vox_events(const double* tri, int64_t ntri, const VoxGrid& g, int* events,
                       hipStream_t st);
// Per column: voxel counts into voxels (may be null), seeded particles per
// column into counts (may be null); with out_x set, the positions at the
// column offsets (at most out_cap particles).
void launch_vox_columns(const VoxGrid& g, const int* events, int* voxels, const SeedParams& sp,
                        int* counts, const int64_t* offsets, float* out_x, int64_t out_cap,
                        hipStream_t st);
void launch_seed_attrs(int64_t n, const float v[3], int mat, float* vout, int* mout,
                       hipStream_t st);

// Event-bracketed kernel timing (enabled on demand).
class Timer {
 public:
  virtual ~Timer() {}
  virtual void begin(const char* name, hipStream_t st) = 0;
  virtual void end(hipStream_t st) = 0;
};

}  // namespace mpm

// Host runtime and C ABI of the MLS-MPM substep engine (include/mpm_abi.h).
//
// A solver handle owns, on one device and one HIP stream:
//  - particle state as structure-of-arrays in "slots" (float components
//    x, v, C, F, Jp; int components material, insertion index), kept
//    approximately sorted by grid block by a periodic physical reorder;
//  - the binning / scan / tile / grid buffers of the substep;
//  - the grid post-process list (bounding box and colliders).
// Colors never take part in a substep and stay on the host in insertion
// order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_abi.h"
#include "mpm_kernels.h"

using namespace mpm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(MPM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

struct EventTimer : public Timer {
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<int, double>> acc;
  std::vector<std::string> order;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    // timing only (read after a stream synchronisation): no system-scope
    // fence, whose cache write-back would be timed with the kernel before it
    hipEvent_t e;
    hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    return e;
  }
  void begin(const char* name, hipStream_t st) override {
    Pending p{name, get(), get()};
    hipEventRecord(p.a, st);
    pending.push_back(p);
  }
  void end(hipStream_t st) override { hipEventRecord(pending.back().b, st); }
  // Called after a stream synchronisation.
  void collect() {
    for (auto& p : pending) {
      float ms = 0.0f;
      hipEventElapsedTime(&ms, p.a, p.b);
      if (!acc.count(p.name)) order.push_back(p.name);
      auto& e = acc[p.name];
      e.first += 1;
      e.second += ms;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~EventTimer() override {
    for (auto& p : pending) {
      hipEventDestroy(p.a);
      hipEventDestroy(p.b);
    }
    for (auto e : pool) hipEventDestroy(e);
  }
};

// One simulated scene: its constants, gravity, grid post-process list and
// its slab of the combined block window.
struct SceneHost {
  mpm_config cfg{};
  double gravity[3] = {0, 0, 0};
  std::vector<PostOp> post;
  int blo[3] = {0, 0, 0}, bnl[3] = {1, 1, 1};
  int zoff = 0;
};

}  // namespace

struct mpm_solver {
  int D = 3;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<SceneHost> scenes;
  int active_scene = 0;       // target of the configuration calls and add_particles
  Params* h_params = nullptr;  // pinned staging of the scene table
  Params* d_params = nullptr;  // device scene table
  int* d_slab = nullptr;       // scene of each combined block coordinate on the last axis

  // combined block window of all scenes (stacked along the last axis)
  int bn[3] = {1, 1, 1};
  int total_blocks = 0;
  int key_bits = 0;

  // particles
  int64_t n = 0, cap = 0;
  float* pf = nullptr;
  float* pf_alt = nullptr;
  int* pi = nullptr;
  int* pi_alt = nullptr;
  std::vector<int32_t> color;  // insertion order, host side

  // substep buffers sized by particle capacity
  int* keys = nullptr;
  int* keys_alt = nullptr;  // 2 * cap
  int* perm = nullptr;
  int* perm_alt = nullptr;
  unsigned long long* lb_status_radix = nullptr;
  float4* tiles = nullptr;
  int4* chunks = nullptr;  // 2 x chunk_cap (substep parity)
  int64_t chunk_cap = 0;
  int* code = nullptr;     // cap
  int* nkey = nullptr;     // cap
  int* dcount = nullptr;   // 27 x chunk_cap
  int* doff = nullptr;     // 27 x chunk_cap
  int* keys_sorted = nullptr;
  // buffers sized by the block window
  int* count = nullptr;
  int* count_next = nullptr;
  int* pstart = nullptr;
  int* pchunk_begin = nullptr;  // 2 x total_blocks (substep parity)
  int* pchunk_cnt = nullptr;    // 2 x total_blocks
  int4* tile_prefix = nullptr;
  int* active = nullptr;
  unsigned long long* lb_status_scan = nullptr;
  LBState* lbs = nullptr;
  int* ghist = nullptr;
  float4* grid = nullptr;
  int* totals = nullptr;  // 2 x 16 (substep parity)
  int last_parity = 0;
  bool reordered = false;
  int* err = nullptr;
  unsigned* maxv = nullptr;
  int maxv_cap = 0;
  int last_batch = 0;

  bool have_perm = false;  // perm of the previous substep matches the slots
  int since_reorder = 1 << 30;  // reorder at the first opportunity
  int64_t substep_counter = 0;
  // mpm_substep_begin ran and mpm_substep_end has not: the buffers of the
  // substep in flight (binning, grid) between the two halves
  bool half_pending = false;
  SubstepBuffers half_B{};
  int* d_step = nullptr;  // device substep counter (max-velocity slot)
  unsigned* d_gvmax = nullptr;  // compute_max_grid_velocity result
  float* d_aabb = nullptr;      // unbounded: per-workgroup particle boxes (kAabbGroups x 6)
  std::vector<float> h_aabb;
  // HIP graphs of two steady-state substeps, keyed by the buffer pointers
  // and sizes their kernels were captured with
  bool use_graphs = true;
  struct GraphEntry {
    std::vector<intptr_t> key;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  int reorder_interval = 16;
  int g2p_waves = 0;   // g2p occupancy target (mpm_set_g2p_waves)
  int chunk = kBatch;  // particles per transfer work item (autotuned up to kChunkMax)
  int persistent_groups = 2048;
  int debug = 0;
  bool timing = false;
  EventTimer timer;

  // use_g2p2g mode: the fused transfer writes the next grid while reading
  // this one (grid_alt), particles with an insertion index >= n_final are
  // new (no G2P on their first substep)
  int g2p2g = 0;
  int g2p2g_clamp = 1;
  double g2p2g_cfl = 0.9;
  float4* grid_alt = nullptr;
  int64_t n_final = 0;

  // differentiable mode (mpm_diff_*): the particle state lives in caller
  // arrays; the handle supplies the block window, workspace and grid
  int64_t diff_n = 0;
  int* diff_pi = nullptr;          // 3 x diff_n: material, insertion index, scene 0
  float* diff_scratch = nullptr;   // NF x diff_n: the state p2g rewrites in backward
  float4* diff_gbar = nullptr;     // grid adjoint
  float* diff_abuf = nullptr;      // D*D x diff_n: affine matrices of the backward
  float* diff_nu_part = nullptr;   // per-workgroup nu adjoints
  bool diff_params_valid = false;  // the device scene table holds these values:
  float diff_dt = 0.0f;
  const float* diff_youngs = nullptr;
  double diff_nu = 0.0;

  // mpm_snapshot: a device copy of the particle slots and, in use_g2p2g
  // mode, of the grid the next substep reads (with n_final and, for an
  // unbounded scene, the block window that grid belongs to)
  float* snap_pf = nullptr;
  int* snap_pi = nullptr;
  float4* snap_grid = nullptr;
  int64_t snap_n = -1, snap_cap = 0, snap_n_final = 0, snap_grid_nodes = 0;
  int snap_nf = 0;
  int snap_blo[3] = {0, 0, 0}, snap_bnl[3] = {1, 1, 1};

  // particle storage: quant = 1 packs x, v and F (PackedLayout); plastic = 0
  // is support_plasticity=False (no Jp in the packed layout)
  int quant = 0;
  int plastic = 1;
  // unbounded (mpm_solver.py:143-148, 604-611): the walls sit at
  // +-grid_half cells and the block window follows the particles
  int unbounded = 0;
  int grid_half = 2048;

  // chunk size the binning uses: the fused g2p2g transfer holds a whole
  // chunk in one batch
  int eff_chunk() const { return g2p2g ? std::min(chunk, kBatch) : chunk; }

  // 32-bit words per particle slot in the float buffers
  int nf() const {
    if (!quant) return D == 3 ? Layout<3>::NF : Layout<2>::NF;
    return D == 3 ? PackedLayout<3>::words(g2p2g != 0, plastic != 0)
                  : PackedLayout<2>::words(g2p2g != 0, plastic != 0);
  }
};

namespace {

template <typename T>
int dalloc(T** p, int64_t count) {
  *p = nullptr;
  if (count <= 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
  if (e != hipSuccess)
    return fail(MPM_ENOMEM, std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                                " bytes failed: " + hipGetErrorString(e));
  return MPM_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

#define TRY(expr)           \
  do {                      \
    int r_ = (expr);        \
    if (r_ != MPM_OK) return r_; \
  } while (0)

int no_half_pending(const mpm_solver* s) {
  if (s->half_pending)
    return fail(MPM_EINVAL, "a substep begun by mpm_substep_begin has not been ended "
                            "(mpm_substep_end)");
  return MPM_OK;
}

int check_device_errors(mpm_solver* s) {
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->timing) s->timer.collect();
  int e = 0;
  HIP_TRY(hipMemcpy(&e, s->err, sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    int z = 0;
    HIP_TRY(hipMemcpy(s->err, &z, sizeof(int), hipMemcpyHostToDevice));
    if (e & 1)
      return fail(MPM_EDOMAIN,
                  "particles left the simulated window (positions outside the grid around the "
                  "domain); results are no longer valid");
    if (e & 8)
      return fail(MPM_EDOMAIN,
                  "a particle moved more than one 4-cell grid block in one substep (CFL number "
                  "above 4); reduce the time step");
    return fail(MPM_EHIP, "internal consistency check failed in the binning kernels (flags " +
                              std::to_string(e) + ")");
  }
  return MPM_OK;
}

void fill_params(Params& P, const mpm_solver* s, const SceneHost& sc, float dt, int skip) {
  const mpm_config& c = sc.cfg;
  P.dx = (float)c.dx;
  P.inv_dx = (float)c.inv_dx;
  P.dt = dt;
  P.p_mass = (float)c.p_mass;
  P.mu_0 = (float)c.mu_0;
  P.lambda_0 = (float)c.lambda_0;
  P.alpha = (float)c.alpha;
  P.water_density = (float)c.water_density;
  // (-dt * p_vol * 4 * inv_dx**2) evaluated left to right in f32 (mpm:569)
  P.stress_scale = ((-dt * (float)c.p_vol) * 4.0f) * (float)(c.inv_dx * c.inv_dx);
  P.c4 = (float)(4 * c.inv_dx);
  P.sand_coef = (float)((s->D * c.lambda_0 + 2 * c.mu_0) / (2 * c.mu_0));
  for (int a = 0; a < 3; a++) {
    P.gravity[a] = a < s->D ? (float)sc.gravity[a] : 0.0f;
    P.res[a] = a < s->D ? c.res[a] : 1;
    P.blo[a] = sc.blo[a];
    P.bnl[a] = sc.bnl[a];
    P.bn[a] = s->bn[a];
  }
  P.padding = c.padding;
  P.zoff = sc.zoff;
  P.skip = skip;
  P.youngs = nullptr;
  P.kmu = P.kla = P.dkmu = P.dkla = P.dsand_coef = 0.0f;
  P.g2p2g = s->g2p2g;
  P.plastic = s->plastic;
  P.v_allowed = s->g2p2g && s->g2p2g_clamp && s->g2p2g_cfl > 0
                    ? (float)(c.dx * s->g2p2g_cfl) / dt
                    : 0.0f;
  P.n_post = (int)sc.post.size();
  for (size_t k = 0; k < sc.post.size(); k++) P.post[k] = sc.post[k];
}

int free_particle_buffers(mpm_solver* s) {
  dfree(s->pf);
  dfree(s->pf_alt);
  dfree(s->pi);
  dfree(s->pi_alt);
  dfree(s->keys);
  dfree(s->keys_alt);
  dfree(s->perm);
  dfree(s->perm_alt);
  dfree(s->lb_status_radix);
  dfree(s->tiles);
  dfree(s->chunks);
  dfree(s->code);
  dfree(s->nkey);
  dfree(s->dcount);
  dfree(s->doff);
  return MPM_OK;
}

int alloc_chunk_buffers(mpm_solver* s);


// Captured graphs hold buffer addresses: drop them when buffers move.
void drop_graphs(mpm_solver* s) {
  if (!s->graphs.empty()) hipStreamSynchronize(s->stream);
  for (auto& g : s->graphs) hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

// The particle layout is about to change: only without particles; the
// buffers are reallocated by the next grow().
int drop_particle_layout(mpm_solver* s) {
  if (s->n > 0) return fail(MPM_EINVAL, "the particle storage can only change before particles are added");
  if (s->cap > 0) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    drop_graphs(s);
    free_particle_buffers(s);
    s->cap = 0;
    s->have_perm = false;
  }
  return MPM_OK;
}

int grow(mpm_solver* s, int64_t need) {
  if (need <= s->cap) return MPM_OK;
  drop_graphs(s);
  if (need > (int64_t)INT32_MAX / 2) return fail(MPM_EINVAL, "too many particles (max 2^30)");
  int64_t ncap = std::max<int64_t>(need, std::max<int64_t>(1024, s->cap * 3 / 2));
  const int nf = s->nf();
  float *pf = nullptr, *pf_alt = nullptr;
  int *pi = nullptr, *pi_alt = nullptr;
  TRY(dalloc(&pf, (int64_t)nf * ncap));
  TRY(dalloc(&pf_alt, (int64_t)nf * ncap));
  const int ni = s->D == 3 ? Layout<3>::NI : Layout<2>::NI;
  TRY(dalloc(&pi, ni * ncap));
  TRY(dalloc(&pi_alt, ni * ncap));
  if (s->n > 0) {
    HIP_TRY(hipMemcpy2DAsync(pf, ncap * sizeof(float), s->pf, s->cap * sizeof(float),
                             s->n * sizeof(float), nf, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpy2DAsync(pi, ncap * sizeof(int), s->pi, s->cap * sizeof(int),
                             s->n * sizeof(int), ni, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  free_particle_buffers(s);
  s->pf = pf;
  s->pf_alt = pf_alt;
  s->pi = pi;
  s->pi_alt = pi_alt;
  s->cap = ncap;
  TRY(dalloc(&s->keys, ncap));
  TRY(dalloc(&s->keys_alt, 2 * ncap));
  TRY(dalloc(&s->perm, ncap));
  TRY(dalloc(&s->perm_alt, ncap));
  const int64_t ntiles = (ncap + kRadixTile - 1) / kRadixTile;
  TRY(dalloc(&s->lb_status_radix, kRadixBins * ntiles));
  HIP_TRY(hipMemsetAsync(s->lb_status_radix, 0, sizeof(unsigned long long) * kRadixBins * ntiles,
                         s->stream));
  TRY(dalloc(&s->code, ncap));
  TRY(dalloc(&s->nkey, ncap));
  TRY(alloc_chunk_buffers(s));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->have_perm = false;
  return MPM_OK;
}

// Buffers sized by the number of transfer work items (chunks), which
// depends on the particle capacity and the chunk size.
int alloc_chunk_buffers(mpm_solver* s) {
  drop_graphs(s);
  dfree(s->tiles);
  dfree(s->chunks);
  dfree(s->dcount);
  dfree(s->doff);
  const int chunk = s->eff_chunk();
  s->chunk_cap = std::min<int64_t>(s->total_blocks, s->cap) + (s->cap + chunk - 1) / chunk;
  int TN = s->D == 3 ? Geo<3>::kTileNodes : Geo<2>::kTileNodes;
  if (s->g2p2g) TN = s->D == 3 ? (kBlock + 4) * (kBlock + 4) * (kBlock + 4) : (kBlock + 4) * (kBlock + 4);
  TRY(dalloc(&s->tiles, s->chunk_cap * TN));
  TRY(dalloc(&s->chunks, 2 * s->chunk_cap));
  TRY(dalloc(&s->dcount, 27 * s->chunk_cap));
  TRY(dalloc(&s->doff, 27 * s->chunk_cap));
  return MPM_OK;
}

void fill_buffers(mpm_solver* s, SubstepBuffers& B) {
  B.pf = s->pf;
  B.pf_out = s->pf;
  B.pi = s->pi;
  B.cap = s->cap;
  B.n = (int)s->n;
  B.keys = s->keys;
  B.keys_alt = s->keys_alt;
  B.perm = s->perm;
  B.perm_alt = s->perm_alt;
  const int cur = (int)(s->substep_counter & 1), prev = cur ^ 1;
  const int64_t tb = s->total_blocks;
  B.count = s->count;
  B.count_next = s->count_next;
  B.pstart = s->pstart;
  B.pchunk_begin = s->pchunk_begin + cur * tb;
  B.pchunk_cnt = s->pchunk_cnt + cur * tb;
  B.chunks = s->chunks + cur * s->chunk_cap;
  B.pchunk_begin_prev = s->pchunk_begin + prev * tb;
  B.pchunk_cnt_prev = s->pchunk_cnt + prev * tb;
  B.chunks_prev = s->chunks + prev * s->chunk_cap;
  B.totals_prev = s->totals + prev * 16;
  B.code = s->code;
  B.nkey = s->nkey;
  B.dcount = s->dcount;
  B.doff = s->doff;
  B.tile_prefix = s->tile_prefix;
  B.full_sort = !s->have_perm;
  B.perm_identity = s->reordered;
  B.g2p2g = s->g2p2g != 0;
  B.quant = s->quant != 0;
  B.nw = s->nf();
  B.grid_in = s->g2p2g ? s->grid : nullptr;
  B.n_final = (int)s->n_final;
  B.keys_sorted = nullptr;
  B.active = s->active;
  B.lb_status_scan = s->lb_status_scan;
  B.lb_status_radix = s->lb_status_radix;
  B.lbs = s->lbs;
  B.ghist = s->ghist;
  B.totals = s->totals + cur * 16;
  B.tiles = s->tiles;
  B.grid = s->g2p2g ? s->grid_alt : s->grid;
  B.raw_out = nullptr;
  B.err = s->err;
  B.maxv = s->maxv;
  B.step = s->d_step;
  B.total_blocks = s->total_blocks;
  B.key_bits = s->key_bits;
  B.chunk = s->eff_chunk();
  B.persistent_groups = s->persistent_groups;
  B.g2p_waves = s->g2p_waves;
  B.parity = (int)(s->substep_counter & 1);
}

void take_buffers(mpm_solver* s, const SubstepBuffers& B) {
  if (B.g2p2g && B.n > 0) {  // the output grid becomes the next input
    std::swap(s->grid, s->grid_alt);
    s->n_final = s->n;
  }
  s->perm = B.perm;
  s->perm_alt = B.perm_alt;
  s->keys_sorted = B.keys_sorted;
  s->last_parity = B.parity;
  s->count = B.count;
  s->count_next = B.count_next;
}

// The physical reorder of the particle slots when one is due, then B for
// the next substep.
void begin_buffers(mpm_solver* s, SubstepBuffers& B, Timer* tm) {
  fill_buffers(s, B);
  if (s->have_perm && s->reorder_interval > 0 && s->since_reorder >= s->reorder_interval) {
    if (s->D == 3)
      launch_reorder<3>(B, s->pf_alt, s->pi_alt, s->stream, tm);
    else
      launch_reorder<2>(B, s->pf_alt, s->pi_alt, s->stream, tm);
    std::swap(s->pf, s->pf_alt);
    std::swap(s->pi, s->pi_alt);
    s->since_reorder = 0;
    s->reordered = true;
    fill_buffers(s, B);
  }
}

// Host bookkeeping after a substep's kernels were launched.
void end_buffers(mpm_solver* s, const SubstepBuffers& B) {
  take_buffers(s, B);
  s->substep_counter++;
  s->have_perm = true;
  s->reordered = false;
  s->since_reorder++;
}

// One substep launched kernel by kernel (also what a graph captures).
int eager_substep(mpm_solver* s, int k, Timer* tm) {
  SubstepBuffers B;
  begin_buffers(s, B, tm);
  if (s->D == 3)
    launch_substep<3>(s->d_params, s->d_slab, B, k, s->debug & 1, s->stream, tm);
  else
    launch_substep<2>(s->d_params, s->d_slab, B, k, s->debug & 1, s->stream, tm);
  end_buffers(s, B);
  return MPM_OK;
}

// The next two substeps can run as one HIP graph: binning by the
// neighbour sort (not the first substep, not right after positions were
// set), even parity (the pair returns every double-buffered pointer to its
// place), no physical reorder due, g2p2g past its first substep, no
// per-kernel timing.
bool steady_pair(const mpm_solver* s, int left, const Timer* tm) {
  if (!s->use_graphs || tm || left < 2 || s->n == 0) return false;
  if (!s->have_perm || s->reordered || (s->substep_counter & 1)) return false;
  if (s->reorder_interval > 0 && s->since_reorder + 1 >= s->reorder_interval) return false;
  if (s->g2p2g && s->n_final != s->n) return false;
  return true;
}

std::vector<intptr_t> graph_key(const mpm_solver* s) {
  return {(intptr_t)s->pf, (intptr_t)s->pi, (intptr_t)s->perm, (intptr_t)s->perm_alt,
          (intptr_t)s->count, (intptr_t)s->count_next, (intptr_t)s->grid,
          (intptr_t)s->grid_alt, (intptr_t)s->maxv, (intptr_t)s->tiles, (intptr_t)s->chunks,
          (intptr_t)s->keys_alt, (intptr_t)s->dcount, (intptr_t)s->doff, (intptr_t)s->n,
          (intptr_t)s->cap, (intptr_t)s->debug, (intptr_t)s->chunk,
          (intptr_t)s->persistent_groups, (intptr_t)s->g2p2g, (intptr_t)s->g2p_waves};
}

// Two substeps as one graph launch: captured on first use (the capture
// itself advances the host-side bookkeeping, as the eager path does), then
// replayed with the same bookkeeping applied by hand.
int run_graph_pair(mpm_solver* s) {
  const std::vector<intptr_t> key = graph_key(s);
  for (auto& g : s->graphs) {
    if (g.key != key) continue;
    HIP_TRY(hipGraphLaunch(g.exec, s->stream));
    // bookkeeping of two substeps: every swapped pointer is back in place
    s->last_parity = 1;
    s->keys_sorted = s->keys_alt;
    s->substep_counter += 2;
    s->since_reorder += 2;
    if (s->g2p2g) s->n_final = s->n;
    return MPM_OK;
  }
  if (s->graphs.size() >= 8) {
    for (auto& g : s->graphs) hipGraphExecDestroy(g.exec);
    s->graphs.clear();
  }
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int r = eager_substep(s, 0, nullptr);
  if (r == MPM_OK) r = eager_substep(s, 1, nullptr);
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  if (r != MPM_OK) {
    if (graph) hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) return fail(MPM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(MPM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  s->graphs.push_back({key, exec});
  HIP_TRY(hipGraphLaunch(exec, s->stream));
  return MPM_OK;
}

}  // namespace

namespace {

// Device array freed on scope exit (after the stream work that uses it has
// been waited for by the caller).
template <typename T>
struct DeviceArray {
  T* p = nullptr;
  DeviceArray() = default;
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray() { dfree(p); }
  int alloc(int64_t count) { return dalloc(&p, count); }
};

// Float-layout copy (Layout<D>, same cap) of every slot of a solver with
// packed particle storage, on the solver's stream.
int unpacked(mpm_solver* s, DeviceArray<float>& flt) {
  const int nf = s->D == 3 ? Layout<3>::NF : Layout<2>::NF;
  TRY(flt.alloc((int64_t)nf * s->cap));
  if (s->D == 3)
    launch_unpack<3>(s->pf, flt.p, (int)s->n, s->cap, s->g2p2g != 0, s->plastic != 0, s->stream);
  else
    launch_unpack<2>(s->pf, flt.p, (int)s->n, s->cap, s->g2p2g != 0, s->plastic != 0, s->stream);
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}

// Voxelizer.__init__ (engine/voxelizer.py:18-38): every axis of the solver
// resolution rounded up to a power of two, times super_sample; voxel size
// dx / super_sample.
int voxel_grid(const int32_t* res, double dx, int32_t super_sample, int32_t padding,
               VoxGrid* g) {
  if (super_sample < 1 || super_sample > 16) return fail(MPM_EINVAL, "super_sample must be 1..16");
  if (padding < 0) return fail(MPM_EINVAL, "negative padding");
  if (!(dx > 0.0) || !std::isfinite(dx)) return fail(MPM_EINVAL, "dx must be positive");
  for (int k = 0; k < 3; k++) {
    if (res[k] < 1 || res[k] > (1 << 16)) return fail(MPM_EINVAL, "resolution out of range");
    int r = 1;
    while (r < res[k]) r *= 2;
    g->res[k] = r * super_sample;
  }
  g->padding = padding;
  g->dx = dx / super_sample;
  g->inv_dx = 1.0 / g->dx;
  return MPM_OK;
}

// Column-fill events of the triangles on stream st (events: res0 * res1 *
// (res2 + 1) ints).  Waits for the stream before returning.
int voxel_events(const VoxGrid& g, int64_t ntri, const double* triangles, hipStream_t st,
                 DeviceArray<int>& events) {
  if (ntri < 0) return fail(MPM_EINVAL, "negative triangle count");
  if (ntri > 0 && !triangles) return fail(MPM_EINVAL, "null triangles");
  for (int64_t i = 0; i < ntri * 9; i++)
    if (!std::isfinite(triangles[i])) return fail(MPM_EINVAL, "triangle coordinates must be finite");
  const int64_t nev = (int64_t)g.res[0] * g.res[1] * (g.res[2] + 1);
  TRY(events.alloc(nev));
  HIP_TRY(hipMemsetAsync(events.p, 0, sizeof(int) * nev, st));
  DeviceArray<double> tri;
  if (ntri > 0) {
    TRY(tri.alloc(ntri * 9));
    HIP_TRY(hipMemcpyAsync(tri.p, triangles, sizeof(double) * ntri * 9, hipMemcpyHostToDevice, st));
    launch_vox_events(tri.p, ntri, g, events.p, st);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(st));
  return MPM_OK;
}

// ---------------------------------------------------------------------------
// Static workspace of mpm_diff_substep_noalloc: a file-scope device array that
// every GPU the library runs on holds until the process ends.  Host code uses
// the device address hipGetSymbolAddress returns for it (resolved once per
// device when a solver is created on that device), never the array's name.
// ---------------------------------------------------------------------------
constexpr int64_t kScratchBytes = 64ll << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Resolves this device's copy of the workspace; the device must be current.
int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(MPM_EINVAL, "device index beyond the workspace table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return MPM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MPM_EHIP, std::string("workspace address: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return MPM_OK;
}

// Carves 256-byte aligned arrays out of one region (base == nullptr: only
// counts the bytes).
struct Carve {
  char* base;
  int64_t off = 0;
  template <typename T>
  T* take(int64_t count) {
    off = (off + 255) & ~(int64_t)255;
    T* p = base ? (T*)(base + off) : nullptr;
    off += (int64_t)sizeof(T) * std::max<int64_t>(count, 1);
    return p;
  }
};

// The workspace of one differentiable forward substep of handle s, carved
// from c (binning, chunk list, scan and sort state, node tiles, grid,
// per-workgroup maxima); the particle state stays in the caller's arrays.
void carve_diff_workspace(const mpm_solver* s, Carve& c, SubstepBuffers& B) {
  const int64_t n = s->diff_n, tb = s->total_blocks;
  const int NPB = s->D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
  const int TN = s->D == 3 ? Geo<3>::kTileNodes : Geo<2>::kTileNodes;
  const int64_t chunk_cap = std::min<int64_t>(tb, n) + (n + s->eff_chunk() - 1) / s->eff_chunk();
  const int64_t nparts = (tb + kScanTile - 1) / kScanTile;
  const int64_t rtiles = (n + kRadixTile - 1) / kRadixTile;
  B.keys = c.take<int>(n);
  B.keys_alt = c.take<int>(2 * n);
  B.perm = c.take<int>(n);
  B.perm_alt = c.take<int>(n);
  B.count = c.take<int>(tb + 1);
  B.count_next = c.take<int>(tb + 1);
  B.pstart = c.take<int>(tb);
  B.pchunk_begin = B.pchunk_begin_prev = c.take<int>(tb);
  B.pchunk_cnt = B.pchunk_cnt_prev = c.take<int>(tb);
  B.chunks = B.chunks_prev = c.take<int4>(chunk_cap);
  B.active = c.take<int>(tb);
  B.tile_prefix = c.take<int4>(nparts);
  B.totals = B.totals_prev = c.take<int>(16);
  B.ghist = c.take<int>(2 * kMaxRadixPasses * kRadixBins);
  B.lb_status_scan = c.take<unsigned long long>(4 * nparts);
  B.lb_status_radix = c.take<unsigned long long>(kRadixBins * rtiles);
  B.lbs = c.take<LBState>(1);
  B.code = c.take<int>(n);
  B.nkey = c.take<int>(n);
  B.dcount = c.take<int>(27 * chunk_cap);
  B.doff = nullptr;  // steady-state binning only
  B.maxv = c.take<unsigned>(s->persistent_groups);
  B.tiles = c.take<float4>(chunk_cap * TN);
  B.grid = c.take<float4>(tb * NPB);
}

}  // namespace

extern "C" {

const char* mpm_last_error(void) { return g_last_error.c_str(); }

// Buffers sized by the combined block window (bn, total_blocks): block
// counts, scans, chunk tables, active list, grid(s), slab table; zeroed.
int alloc_window_buffers(mpm_solver* s) {
  const int D = s->D;
  const int64_t tb = s->total_blocks;
  const int NPB = D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
  const int64_t nparts = (tb + kScanTile - 1) / kScanTile;
  TRY(dalloc(&s->count, tb + 1));
  TRY(dalloc(&s->count_next, tb + 1));
  TRY(dalloc(&s->lb_status_scan, 4 * nparts));
  TRY(dalloc(&s->pstart, tb));
  TRY(dalloc(&s->pchunk_begin, 2 * tb));
  TRY(dalloc(&s->pchunk_cnt, 2 * tb));
  TRY(dalloc(&s->active, tb));
  TRY(dalloc(&s->tile_prefix, nparts));
  TRY(dalloc(&s->grid, tb * NPB));
  if (s->g2p2g) TRY(dalloc(&s->grid_alt, tb * NPB));
  TRY(dalloc(&s->d_slab, s->bn[D - 1]));
  std::vector<int> slab(s->bn[D - 1], 0);
  for (size_t k = 0; k < s->scenes.size(); k++)
    for (int z = 0; z < s->scenes[k].bnl[D - 1]; z++) slab[s->scenes[k].zoff + z] = (int)k;
  HIP_TRY(hipMemcpyAsync(s->d_slab, slab.data(), sizeof(int) * slab.size(), hipMemcpyHostToDevice,
                         s->stream));
  HIP_TRY(hipMemsetAsync(s->count, 0, sizeof(int) * (tb + 1), s->stream));
  HIP_TRY(hipMemsetAsync(s->count_next, 0, sizeof(int) * (tb + 1), s->stream));
  HIP_TRY(hipMemsetAsync(s->lb_status_scan, 0, sizeof(unsigned long long) * 4 * nparts,
                         s->stream));
  HIP_TRY(hipMemsetAsync(s->grid, 0, sizeof(float4) * tb * NPB, s->stream));
  if (s->grid_alt) HIP_TRY(hipMemsetAsync(s->grid_alt, 0, sizeof(float4) * tb * NPB, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

void free_window_buffers(mpm_solver* s) {
  dfree(s->count);
  dfree(s->count_next);
  dfree(s->lb_status_scan);
  dfree(s->pstart);
  dfree(s->pchunk_begin);
  dfree(s->pchunk_cnt);
  dfree(s->active);
  dfree(s->tile_prefix);
  dfree(s->grid);
  dfree(s->grid_alt);
  dfree(s->d_slab);
}

// Moves the block window of a one-scene solver to [blo, blo + bnl) (blocks).
// The grid a use_g2p2g substep reads next is carried over; the next
// substep bins from scratch.
int set_window(mpm_solver* s, const int blo[3], const int bnl[3]) {
  const int D = s->D;
  int64_t tb = 1;
  for (int a = 0; a < D; a++) tb *= bnl[a];
  if (tb >= (1 << (8 * kMaxRadixPasses)) - 1)
    return fail(MPM_EDOMAIN, "the particles of the unbounded scene spread over more than 2^24 "
                             "grid blocks (the dense block window's limit)");
  drop_graphs(s);
  HIP_TRY(hipStreamSynchronize(s->stream));
  SceneHost& sc = s->scenes[0];
  WindowMap m{};
  for (int a = 0; a < 3; a++) {
    m.old_blo[a] = sc.blo[a];
    m.old_bn[a] = s->bn[a];
  }
  float4* keep = nullptr;  // use_g2p2g: the input grid of the next substep
  if (s->g2p2g) {
    keep = s->grid;
    s->grid = nullptr;
  }
  free_window_buffers(s);
  for (int a = 0; a < D; a++) {
    sc.blo[a] = blo[a];
    sc.bnl[a] = bnl[a];
    s->bn[a] = bnl[a];
  }
  s->total_blocks = (int)tb;
  s->key_bits = 1;
  while ((1ll << s->key_bits) <= tb) s->key_bits++;
  int r = alloc_window_buffers(s);
  if (r == MPM_OK && keep) {
    for (int a = 0; a < 3; a++) {
      m.new_blo[a] = sc.blo[a];
      m.new_bn[a] = s->bn[a];
    }
    const int NPB = D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
    launch_remap_grid(keep, s->grid, m, tb * NPB, NPB, s->stream);
    if (hipStreamSynchronize(s->stream) != hipSuccess) r = fail(MPM_EHIP, "grid remap failed");
  }
  dfree(keep);
  if (r != MPM_OK) return r;
  if (s->cap > 0) TRY(alloc_chunk_buffers(s));
  s->have_perm = false;
  return MPM_OK;
}

// Unbounded scene: before the next `seg` substeps, make sure every particle
// stays in the window for all of them (a particle moves less than one block
// per substep): its block and the next one, `seg + 1` blocks of margin.
// A window that does not cover that is moved to the particles' box with a
// margin of twice that on each side, clamped to the walls' extent.
constexpr int kAabbGroups = 1024;

int follow_particles(mpm_solver* s, int seg) {
  if (!s->unbounded || s->n == 0) return MPM_OK;
  const int D = s->D;
  const int groups = (int)std::min<int64_t>((s->n + kThreads - 1) / kThreads, kAabbGroups);
  if (!s->d_aabb) TRY(dalloc(&s->d_aabb, (int64_t)kAabbGroups * 6));  // kept until destroy
  launch_aabb(D, s->quant != 0, s->pf, s->cap, (int)s->n, s->d_aabb, groups, s->stream);
  std::vector<float>& h = s->h_aabb;
  h.resize((size_t)groups * 6);
  hipError_t e = hipMemcpyAsync(h.data(), s->d_aabb, sizeof(float) * h.size(),
                                hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(MPM_EHIP, std::string("particle box: ") + hipGetErrorString(e));
  const SceneHost& sc = s->scenes[0];
  const float inv_dx = (float)sc.cfg.inv_dx;
  const int margin = seg + 1;
  const int wall_blocks = s->grid_half / kBlock + 2;  // beyond the walls nothing moves
  int need_lo[3], need_hi[3], blo[3], bnl[3];
  bool fits = true;
  for (int a = 0; a < D; a++) {
    float lo = h[2 * a], hi = h[2 * a + 1];
    for (int g = 1; g < groups; g++) {
      lo = std::min(lo, h[(size_t)g * 6 + 2 * a]);
      hi = std::max(hi, h[(size_t)g * 6 + 2 * a + 1]);
    }
    if (!std::isfinite(lo) || !std::isfinite(hi))
      return fail(MPM_EDOMAIN, "particle positions are not finite");
    const auto block_of = [&](float x) {
      const double c = std::floor((double)(x * inv_dx) - 0.5);  // base cell
      return (int)std::floor(c / kBlock);
    };
    need_lo[a] = std::max(block_of(lo) - margin, -wall_blocks);
    need_hi[a] = std::min(block_of(hi) + 1 + margin, wall_blocks);  // inclusive
    fits &= need_lo[a] >= sc.blo[a] && need_hi[a] < sc.blo[a] + sc.bnl[a];
    blo[a] = std::max(need_lo[a] - margin, -wall_blocks);
    bnl[a] = std::min(need_hi[a] + margin, wall_blocks) - blo[a] + 1;
  }
  for (int a = D; a < 3; a++) {
    blo[a] = 0;
    bnl[a] = 1;
  }
  if (fits) return MPM_OK;
  TRY(set_window(s, blo, bnl));
  // the scene table carries the window
  for (size_t k = 0; k < s->scenes.size(); k++) {
    const float dt = s->h_params[k].dt;
    const int skip = s->h_params[k].skip;
    fill_params(s->h_params[k], s, s->scenes[k], dt, skip);
  }
  HIP_TRY(hipMemcpyAsync(s->d_params, s->h_params, sizeof(Params) * s->scenes.size(),
                         hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

static int create_impl(const mpm_config* cfgs, int nscenes, mpm_solver** out) {
  if (!cfgs || !out || nscenes < 1) return fail(MPM_EINVAL, "null argument or no scene");
  *out = nullptr;
  const int D = cfgs[0].dim;
  if (D != 2 && D != 3) return fail(MPM_EINVAL, "MPM solver supports only 2D and 3D simulations.");
  for (int k = 0; k < nscenes; k++) {
    const mpm_config* cfg = &cfgs[k];
    if (cfg->dim != D) return fail(MPM_EINVAL, "all scenes of a batch must have the same dim");
    for (int a = 0; a < D; a++)
      if (cfg->res[a] <= 0 || cfg->res[a] > 65536) return fail(MPM_EINVAL, "bad resolution");
    if (!(cfg->dx > 0) || !(cfg->inv_dx > 0)) return fail(MPM_EINVAL, "bad dx");
  }
  const mpm_config* cfg = &cfgs[0];
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(MPM_EINVAL, "device " + std::to_string(cfg->device) + " not present (" +
                                std::to_string(ndev) + " visible)");
  mpm_solver* s = new mpm_solver();
  s->D = D;
  s->device = cfg->device;
  if (cfg->reorder_interval > 0) s->reorder_interval = cfg->reorder_interval;
  if (cfg->reorder_interval < 0) s->reorder_interval = 0;
  if (cfg->chunk > 0) s->chunk = std::min(cfg->chunk, kChunkMax);
  // Each scene's block window: 2 blocks below its domain, enough above it for
  // the wall padding plus a margin; a particle block must also have its +1
  // neighbour inside the window.  Scenes are stacked along the last axis.
  s->scenes.resize(nscenes);
  for (int k = 0; k < nscenes; k++) {
    SceneHost& sc = s->scenes[k];
    sc.cfg = cfgs[k];
    sc.gravity[1] = -9.8;  // mpm_solver.py:283, 296
    for (int a = 0; a < 3; a++) {
      if (a < D) {
        sc.blo[a] = -2;
        sc.bnl[a] = (cfgs[k].res[a] + 3) / 4 + 5;
      } else {
        sc.blo[a] = 0;
        sc.bnl[a] = 1;
      }
    }
    sc.zoff = s->bn[D - 1] == 1 && k == 0 ? 0 : s->bn[D - 1];
    for (int a = 0; a < D - 1; a++) s->bn[a] = std::max(k == 0 ? 1 : s->bn[a], sc.bnl[a]);
    s->bn[D - 1] = sc.zoff + sc.bnl[D - 1];
  }
  int64_t tb = (int64_t)s->bn[0] * s->bn[1] * s->bn[2];
  if (tb >= (1 << (8 * kMaxRadixPasses)) - 1) {
    delete s;
    return fail(MPM_EINVAL, "grid too large for the dense block window");
  }
  s->total_blocks = (int)tb;
  s->key_bits = 1;
  while ((1ll << s->key_bits) <= tb) s->key_bits++;  // key == total_blocks marks escapes
  int r = MPM_OK;
  if (hipSetDevice(s->device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess) {
    delete s;
    return fail(MPM_EHIP, "cannot create a HIP stream");
  }
  if ((r = resolve_scratch(s->device)) != MPM_OK) {
    mpm_destroy(s);
    return r;
  }
  if ((r = dalloc(&s->lbs, 1)) || (r = dalloc(&s->ghist, 2 * kMaxRadixPasses * kRadixBins)) ||
      (r = dalloc(&s->totals, 32)) || (r = dalloc(&s->err, 1)) ||
      (r = dalloc(&s->maxv, 16 * s->persistent_groups)) ||
      (r = dalloc(&s->d_params, nscenes)) || (r = dalloc(&s->d_step, 1)) ||
      (r = dalloc(&s->d_gvmax, 1)) ||
      (r = alloc_window_buffers(s))) {
    mpm_destroy(s);
    return r;
  }
  s->maxv_cap = 16 * s->persistent_groups;
  if (hipHostMalloc((void**)&s->h_params, sizeof(Params) * nscenes) != hipSuccess) {
    s->h_params = nullptr;
    mpm_destroy(s);
    return fail(MPM_ENOMEM, "pinned host allocation failed");
  }
  if (hipMemsetAsync(s->lbs, 0, sizeof(LBState), s->stream) != hipSuccess ||
      hipMemsetAsync(s->ghist, 0, sizeof(int) * 2 * kMaxRadixPasses * kRadixBins, s->stream) !=
          hipSuccess ||
      hipMemsetAsync(s->totals, 0, sizeof(int) * 32, s->stream) != hipSuccess ||
      hipMemsetAsync(s->err, 0, sizeof(int), s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    mpm_destroy(s);
    return fail(MPM_EHIP, "initialising solver buffers failed");
  }
  *out = s;
  return MPM_OK;
}

int mpm_create(const mpm_config* cfg, mpm_solver** out) { return create_impl(cfg, 1, out); }

int mpm_create_batch(const mpm_config* cfgs, int32_t nscenes, mpm_solver** out) {
  return create_impl(cfgs, nscenes, out);
}

int mpm_set_active_scene(mpm_solver* s, int32_t scene) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (scene < 0 || scene >= (int)s->scenes.size()) return fail(MPM_EINVAL, "no such scene");
  s->active_scene = scene;
  return MPM_OK;
}

int mpm_destroy(mpm_solver* s) {
  if (!s) return MPM_OK;
  hipSetDevice(s->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  free_particle_buffers(s);
  dfree(s->count);
  dfree(s->count_next);
  dfree(s->pstart);
  dfree(s->pchunk_begin);
  dfree(s->pchunk_cnt);
  dfree(s->tile_prefix);
  dfree(s->active);
  dfree(s->lb_status_scan);
  dfree(s->lbs);
  dfree(s->ghist);
  for (auto& g : s->graphs) hipGraphExecDestroy(g.exec);
  s->graphs.clear();
  dfree(s->d_step);
  dfree(s->d_gvmax);
  dfree(s->d_aabb);
  dfree(s->grid);
  dfree(s->grid_alt);
  dfree(s->totals);
  dfree(s->err);
  dfree(s->maxv);
  dfree(s->d_params);
  dfree(s->d_slab);
  dfree(s->snap_pf);
  dfree(s->snap_pi);
  dfree(s->snap_grid);
  dfree(s->diff_pi);
  dfree(s->diff_scratch);
  dfree(s->diff_gbar);
  dfree(s->diff_abuf);
  dfree(s->diff_nu_part);
  if (s->h_params) hipHostFree(s->h_params);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return MPM_OK;
}

int mpm_set_gravity(mpm_solver* s, const double* g) {
  if (!s || !g) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  for (int a = 0; a < s->D; a++) s->scenes[s->active_scene].gravity[a] = g[a];
  return MPM_OK;
}

int mpm_add_bounding_box(mpm_solver* s, int32_t unbounded) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  SceneHost& sc = s->scenes[s->active_scene];
  std::vector<PostOp>& post = sc.post;
  if (post.size() >= (size_t)kMaxPost) return fail(MPM_EINVAL, "too many grid post-processes");
  PostOp op{};
  op.kind = kPostBBox;
  for (int a = 0; a < 3; a++) {  // mpm_solver.py:604-616
    const int res = a < s->D ? sc.cfg.res[a] : 1;
    op.lo[a] = unbounded ? -s->grid_half + sc.cfg.padding : sc.cfg.padding;
    op.hi[a] = unbounded ? s->grid_half - sc.cfg.padding : res - sc.cfg.padding;
  }
  post.push_back(op);
  return MPM_OK;
}

int mpm_add_surface_collider(mpm_solver* s, const double* point, const double* normal,
                             int32_t surface, double friction) {
  if (!s || !point || !normal) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (surface < 0 || surface > 2) return fail(MPM_EINVAL, "unknown surface type");
  if (surface == MPM_STICKY && friction != 0)
    return fail(MPM_EINVAL, "friction must be 0 on sticky surfaces.");
  std::vector<PostOp>& post = s->scenes[s->active_scene].post;
  if (post.size() >= (size_t)kMaxPost) return fail(MPM_EINVAL, "too many grid post-processes");
  double n2 = 0;
  for (int a = 0; a < s->D; a++) n2 += normal[a] * normal[a];
  if (!(n2 > 0)) return fail(MPM_EINVAL, "collider normal must be non-zero");
  const double scale = 1.0 / std::sqrt(n2);  // mpm_solver.py:654-655
  PostOp op{};
  op.kind = kPostPlane;
  op.surface = surface;
  op.friction = (float)friction;
  for (int a = 0; a < s->D; a++) {
    op.point[a] = (float)point[a];
    op.normal[a] = (float)(scale * normal[a]);
  }
  post.push_back(op);
  return MPM_OK;
}

int mpm_add_sphere_collider(mpm_solver* s, const double* center, double radius, int32_t surface) {
  if (!s || !center) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (surface < 0 || surface > 2) return fail(MPM_EINVAL, "unknown surface type");
  std::vector<PostOp>& post = s->scenes[s->active_scene].post;
  if (post.size() >= (size_t)kMaxPost) return fail(MPM_EINVAL, "too many grid post-processes");
  PostOp op{};
  op.kind = kPostSphere;
  op.surface = surface;
  op.radius = (float)radius;
  for (int a = 0; a < s->D; a++) op.point[a] = (float)center[a];
  post.push_back(op);
  return MPM_OK;
}

int mpm_clear_grid_postprocess(mpm_solver* s) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  s->scenes[s->active_scene].post.clear();
  return MPM_OK;
}

int mpm_add_particles(mpm_solver* s, int64_t k, const float* x, const float* v,
                      const int32_t* material, const int32_t* color) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (k < 0) return fail(MPM_EINVAL, "negative particle count");
  if (k == 0) return MPM_OK;
  if (!x || !material) return fail(MPM_EINVAL, "x and material are required");
  for (int64_t i = 0; i < k; i++)
    if (material[i] < 0 || material[i] > MPM_STATIONARY)
      return fail(MPM_EINVAL, "unknown material id " + std::to_string(material[i]));
  HIP_TRY(hipSetDevice(s->device));
  TRY(grow(s, s->n + k));
  const int D = s->D;
  float* dx = nullptr;
  float* dv = nullptr;
  int* dm = nullptr;
  TRY(dalloc(&dx, k * D));
  if (v) TRY(dalloc(&dv, k * D));
  TRY(dalloc(&dm, k));
  HIP_TRY(hipMemcpyAsync(dx, x, sizeof(float) * k * D, hipMemcpyHostToDevice, s->stream));
  if (v) HIP_TRY(hipMemcpyAsync(dv, v, sizeof(float) * k * D, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(dm, material, sizeof(int) * k, hipMemcpyHostToDevice, s->stream));
  SubstepBuffers B;
  fill_buffers(s, B);
  if (D == 3)
    launch_append<3>(B, dx, dv, dm, (int)k, (int)s->n, s->active_scene, s->plastic, s->stream);
  else
    launch_append<2>(B, dx, dv, dm, (int)k, (int)s->n, s->active_scene, s->plastic, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  dfree(dx);
  dfree(dv);
  dfree(dm);
  for (int64_t i = 0; i < k; i++) s->color.push_back(color ? color[i] : 0xFFFFFF);
  s->n += k;
  s->have_perm = false;
  s->since_reorder = 1 << 30;  // the new particles sit at the end: reorder soon
  return MPM_OK;
}

int mpm_num_particles(mpm_solver* s, int64_t* n) {
  if (!s || !n) return fail(MPM_EINVAL, "null argument");
  *n = s->n;
  return MPM_OK;
}

// Everything before the first substep of a run of nsub: the max-velocity
// slots, the scene table upload and the device substep counter.
static int prepare_substeps(mpm_solver* s, int32_t nsub, const float* dts, bool batch) {
  for (size_t k = 0; k < s->scenes.size(); k++)
    if (batch && !(std::isfinite(dts[k]) && dts[k] >= 0.0f))
      return fail(MPM_EINVAL, "time steps of a batch must be finite and >= 0");
  HIP_TRY(hipSetDevice(s->device));
  if ((int64_t)nsub * s->persistent_groups > s->maxv_cap) {
    drop_graphs(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s->maxv);
    TRY(dalloc(&s->maxv, (int64_t)nsub * s->persistent_groups));
    s->maxv_cap = nsub * s->persistent_groups;
  }
  launch_fill_u32(s->maxv, 0u, (int64_t)std::max(nsub, 1) * s->persistent_groups, s->stream);
  s->last_batch = nsub;
  if (s->n == 0) return MPM_OK;
  // scene table: the staging buffer is reused only once the last upload is done
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (size_t k = 0; k < s->scenes.size(); k++)
    fill_params(s->h_params[k], s, s->scenes[k], dts[k], batch && dts[k] == 0.0f);
  HIP_TRY(hipMemcpyAsync(s->d_params, s->h_params, sizeof(Params) * s->scenes.size(),
                         hipMemcpyHostToDevice, s->stream));
  launch_fill_u32((unsigned*)s->d_step, 0xFFFFFFFFu, 1, s->stream);  // slot k for substep k
  return MPM_OK;
}

// dts[k] is scene k's time step; with batch set, a zero time step means the
// scene sits the substeps out.
static int substeps_impl(mpm_solver* s, int32_t nsub, const float* dts, bool batch) {
  if (nsub < 0) return fail(MPM_EINVAL, "negative substep count");
  TRY(no_half_pending(s));
  TRY(prepare_substeps(s, nsub, dts, batch));
  if (s->n == 0) return MPM_OK;
  Timer* tm = s->timing ? &s->timer : nullptr;
  // unbounded: the window is checked against the particles every kSegment substeps
  constexpr int kSegment = 8;
  int next_check = 0;
  for (int k = 0; k < nsub;) {
    if (s->unbounded && k == next_check) {
      next_check = std::min(k + kSegment, nsub);
      TRY(follow_particles(s, next_check - k));
    }
    const int left = s->unbounded ? next_check - k : nsub - k;
    if (steady_pair(s, left, tm)) {
      TRY(run_graph_pair(s));
      k += 2;
      continue;
    }
    TRY(eager_substep(s, k, tm));
    k++;
  }
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}


int mpm_substeps(mpm_solver* s, int32_t nsub, float dt, double t0) {
  (void)t0;  // the built-in post-processes do not depend on t
  if (!s) return fail(MPM_EINVAL, "null argument");
  std::vector<float> dts(s->scenes.size(), dt);
  return substeps_impl(s, nsub, dts.data(), false);
}

int mpm_substeps_batch(mpm_solver* s, int32_t nsub, const float* dts) {
  if (!s || !dts) return fail(MPM_EINVAL, "null argument");
  return substeps_impl(s, nsub, dts, true);
}

int mpm_substep_begin(mpm_solver* s, float dt) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  if (s->scenes.size() != 1) return fail(MPM_EINVAL, "mpm_substep_begin: single-scene solvers only");
  TRY(no_half_pending(s));
  TRY(prepare_substeps(s, 1, &dt, false));
  if (s->n > 0) {
    if (s->unbounded) TRY(follow_particles(s, 1));
    Timer* tm = s->timing ? &s->timer : nullptr;
    SubstepBuffers& B = s->half_B;
    begin_buffers(s, B, tm);
    if (s->D == 3)
      launch_substep_half<3>(s->d_params, s->d_slab, B, 0, 0, s->stream, tm);
    else
      launch_substep_half<2>(s->d_params, s->d_slab, B, 0, 0, s->stream, tm);
    HIP_TRY(hipGetLastError());
  }
  s->half_pending = true;
  return MPM_OK;
}

int mpm_grid_node_count(mpm_solver* s, int64_t* count) {
  if (!s || !count) return fail(MPM_EINVAL, "null argument");
  if (!s->half_pending) return fail(MPM_EINVAL, "no substep begun (mpm_substep_begin)");
  *count = 0;
  if (s->n == 0) return check_device_errors(s);
  int nact = 0;
  HIP_TRY(hipMemcpyAsync(&nact, s->half_B.totals + kTotActive, sizeof(int),
                         hipMemcpyDeviceToHost, s->stream));
  TRY(check_device_errors(s));
  *count = (int64_t)nact * (s->D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock);
  return MPM_OK;
}

static int half_nodes(mpm_solver* s, int64_t count) {
  if (!s->half_pending) return fail(MPM_EINVAL, "no substep begun (mpm_substep_begin)");
  int64_t have = 0;
  TRY(mpm_grid_node_count(s, &have));
  if (count != have)
    return fail(MPM_EINVAL, "node count " + std::to_string(count) + " differs from the " +
                                std::to_string(have) + " nodes of the active blocks");
  return MPM_OK;
}

int mpm_grid_gather(mpm_solver* s, int32_t* coords, float* v, float* m, int64_t count) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(half_nodes(s, count));
  if (count == 0) return MPM_OK;
  if (!coords || !v || !m) return fail(MPM_EINVAL, "null argument");
  const SceneHost& sc = s->scenes[0];
  NodeWindow w{};
  for (int a = 0; a < 3; a++) {
    w.blo[a] = sc.blo[a];
    w.bn[a] = s->bn[a];
  }
  w.zoff = sc.zoff;
  launch_grid_gather(s->D, s->half_B.grid, s->half_B.active, count, w, coords, v, m, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int mpm_grid_scatter(mpm_solver* s, const float* v, int64_t count) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(half_nodes(s, count));
  if (count == 0) return MPM_OK;
  if (!v) return fail(MPM_EINVAL, "null argument");
  launch_grid_scatter(s->D, s->half_B.grid, s->half_B.active, count, v, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int mpm_substep_end(mpm_solver* s) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  if (!s->half_pending) return fail(MPM_EINVAL, "no substep begun (mpm_substep_begin)");
  s->half_pending = false;
  if (s->n == 0) return MPM_OK;
  HIP_TRY(hipSetDevice(s->device));
  Timer* tm = s->timing ? &s->timer : nullptr;
  SubstepBuffers& B = s->half_B;
  if (s->D == 3)
    launch_substep_half<3>(s->d_params, s->d_slab, B, 0, 1, s->stream, tm);
  else
    launch_substep_half<2>(s->d_params, s->d_slab, B, 0, 1, s->stream, tm);
  end_buffers(s, B);
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}

int mpm_read_max_velocity(mpm_solver* s, int32_t nsub, float* out) {
  if (!s || (!out && nsub > 0)) return fail(MPM_EINVAL, "null argument");
  if (nsub > s->last_batch) return fail(MPM_EINVAL, "more entries than substeps run");
  int r = check_device_errors(s);
  if (nsub > 0) {
    const int G = s->persistent_groups;
    std::vector<float> part((size_t)nsub * G);
    HIP_TRY(hipMemcpy(part.data(), s->maxv, sizeof(float) * part.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < nsub; k++) {
      float m = 0.0f;
      for (int g = 0; g < G; g++) m = std::max(m, part[(size_t)k * G + g]);
      out[k] = m;
    }
  }
  return r;
}

int mpm_sync(mpm_solver* s) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  return check_device_errors(s);
}

static int field_info(mpm_solver* s, int32_t field, int* comp0, int* ncomp, bool* is_int) {
  const int D = s->D;
  *is_int = false;
  switch (field) {
    case MPM_FIELD_X: *comp0 = 0; *ncomp = D; break;
    case MPM_FIELD_V: *comp0 = D; *ncomp = D; break;
    case MPM_FIELD_C: *comp0 = 2 * D; *ncomp = D * D; break;
    case MPM_FIELD_F: *comp0 = 2 * D + D * D; *ncomp = D * D; break;
    case MPM_FIELD_JP: *comp0 = 2 * D + 2 * D * D; *ncomp = 1; break;
    case MPM_FIELD_MATERIAL: *comp0 = 0; *ncomp = 1; *is_int = true; break;
    default: return fail(MPM_EINVAL, "unknown field " + std::to_string(field));
  }
  return MPM_OK;
}

int mpm_get_state(mpm_solver* s, int32_t field, void* host) {
  if (!s || !host) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (field == MPM_FIELD_COLOR) {
    std::memcpy(host, s->color.data(), sizeof(int32_t) * s->n);
    return MPM_OK;
  }
  int c0, nc;
  bool is_int;
  TRY(field_info(s, field, &c0, &nc, &is_int));
  HIP_TRY(hipSetDevice(s->device));
  if (s->n == 0) return MPM_OK;
  DeviceArray<float> flt;  // packed storage: the float layout of every slot
  const float* src = s->pf;
  if (s->quant && !is_int) {
    TRY(unpacked(s, flt));
    src = flt.p;
  }
  void* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, sizeof(float) * s->n * nc));
  const int* orig = s->pi + s->cap;  // Layout ORIG == 1
  if (is_int)
    launch_gather_i(s->pi + (int64_t)c0 * s->cap, s->cap, nc, orig, (int)s->n, (int*)tmp,
                    s->stream);
  else
    launch_gather_f(src + (int64_t)c0 * s->cap, s->cap, nc, orig, (int)s->n, (float*)tmp,
                    s->stream);
  int r = check_device_errors(s);
  hipError_t e = hipMemcpy(host, tmp, sizeof(float) * s->n * nc, hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (e != hipSuccess) return fail(MPM_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return r;
}

int mpm_set_state(mpm_solver* s, int32_t field, const void* host) {
  if (!s || !host) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (field == MPM_FIELD_COLOR) {
    std::memcpy(s->color.data(), host, sizeof(int32_t) * s->n);
    return MPM_OK;
  }
  int c0, nc;
  bool is_int;
  TRY(field_info(s, field, &c0, &nc, &is_int));
  if (is_int) {
    const int32_t* m = (const int32_t*)host;
    for (int64_t i = 0; i < s->n; i++)
      if (m[i] < 0 || m[i] > MPM_STATIONARY) return fail(MPM_EINVAL, "unknown material id");
  }
  HIP_TRY(hipSetDevice(s->device));
  if (s->n == 0) return MPM_OK;
  void* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, sizeof(float) * s->n * nc));
  HIP_TRY(hipMemcpyAsync(tmp, host, sizeof(float) * s->n * nc, hipMemcpyHostToDevice, s->stream));
  if (field == MPM_FIELD_X) s->have_perm = false;  // block keys must be recomputed
  const int* orig = s->pi + s->cap;
  int r = MPM_OK;
  if (is_int) {
    launch_scatter_i((const int*)tmp, s->cap, nc, orig, (int)s->n,
                     s->pi + (int64_t)c0 * s->cap, s->stream);
  } else if (s->quant) {  // unpack, replace the field, pack (the others round-trip exactly)
    DeviceArray<float> flt;
    r = unpacked(s, flt);
    if (r == MPM_OK) {
      launch_scatter_f((const float*)tmp, s->cap, nc, orig, (int)s->n,
                       flt.p + (int64_t)c0 * s->cap, s->stream);
      if (s->D == 3)
        launch_pack<3>(flt.p, s->pf, (int)s->n, s->cap, s->g2p2g != 0, s->plastic != 0, s->stream);
      else
        launch_pack<2>(flt.p, s->pf, (int)s->n, s->cap, s->g2p2g != 0, s->plastic != 0, s->stream);
      hipStreamSynchronize(s->stream);
    }
  } else {
    launch_scatter_f((const float*)tmp, s->cap, nc, orig, (int)s->n,
                     s->pf + (int64_t)c0 * s->cap, s->stream);
  }
  if (r == MPM_OK) r = check_device_errors(s);
  hipStreamSynchronize(s->stream);
  hipFree(tmp);
  return r;
}

int mpm_particle_info(mpm_solver* s, float* x, float* v, int32_t* material, int32_t* color) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (x) TRY(mpm_get_state(s, MPM_FIELD_X, x));
  if (v) TRY(mpm_get_state(s, MPM_FIELD_V, v));
  if (material) TRY(mpm_get_state(s, MPM_FIELD_MATERIAL, material));
  if (color) TRY(mpm_get_state(s, MPM_FIELD_COLOR, color));
  return MPM_OK;
}

int mpm_read_grid(mpm_solver* s, const int32_t* lo, const int32_t* nn, float* m, float* v) {
  if (!s || !lo || !nn) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  TRY(check_device_errors(s));
  const int D = s->D;
  const int NPB = D == 3 ? 64 : 16;
  const int64_t tb = s->total_blocks;
  std::vector<float> g((size_t)tb * 4 * NPB);
  HIP_TRY(hipMemcpy(g.data(), s->grid, sizeof(float) * g.size(), hipMemcpyDeviceToHost));
  // only blocks active in the last substep hold this substep's values
  int nact = 0;
  std::vector<int> tot(16);
  HIP_TRY(hipMemcpy(tot.data(), s->totals + s->last_parity * 16, sizeof(int) * 16,
                    hipMemcpyDeviceToHost));
  nact = tot[kTotActive];
  std::vector<int> act(std::max(nact, 1));
  if (nact > 0)
    HIP_TRY(hipMemcpy(act.data(), s->active, sizeof(int) * nact, hipMemcpyDeviceToHost));
  std::vector<char> is_act(tb, 0);
  for (int i = 0; i < nact; i++) is_act[act[i]] = 1;
  const int n0 = nn[0], n1 = nn[1], n2 = D == 3 ? nn[2] : 1;
  for (int k = 0; k < n2; k++)
    for (int j = 0; j < n1; j++)
      for (int i = 0; i < n0; i++) {
        const int I[3] = {lo[0] + i, lo[1] + j, D == 3 ? lo[2] + k : 0};
        const int64_t o = ((int64_t)k * n1 + j) * n0 + i;
        float mv = 0, vv[3] = {0, 0, 0};
        int lb[3], l[3];
        bool ok = true;
        const SceneHost& sc = s->scenes[s->active_scene];
        for (int a = 0; a < D; a++) {
          int b = I[a] >> 2;
          l[a] = I[a] - 4 * b;
          lb[a] = b - sc.blo[a];
          if (lb[a] < 0 || lb[a] >= sc.bnl[a]) ok = false;
        }
        lb[D - 1] += sc.zoff;
        if (ok) {
          const int64_t bid = D == 3 ? ((int64_t)lb[2] * s->bn[1] + lb[1]) * s->bn[0] + lb[0]
                                     : (int64_t)lb[1] * s->bn[0] + lb[0];
          if (is_act[bid]) {
            const int node = D == 3 ? l[0] + 4 * (l[1] + 4 * l[2]) : l[0] + 4 * l[1];
            const float* gp = &g[((size_t)bid * NPB + node) * 4];
            for (int a = 0; a < D; a++) vv[a] = gp[a];
            mv = gp[3];
          }
        }
        if (m) m[o] = mv;
        if (v)
          for (int a = 0; a < D; a++) v[o * D + a] = vv[a];
      }
  return MPM_OK;
}

int mpm_read_binning(mpm_solver* s, int32_t* keys, int32_t* perm, int32_t* slot_to_insertion) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  TRY(check_device_errors(s));
  if (!s->have_perm) return fail(MPM_EINVAL, "no substep has run since the last change");
  if (s->n == 0) return MPM_OK;
  if (keys) HIP_TRY(hipMemcpy(keys, s->keys_sorted, sizeof(int) * s->n, hipMemcpyDeviceToHost));
  if (perm) HIP_TRY(hipMemcpy(perm, s->perm, sizeof(int) * s->n, hipMemcpyDeviceToHost));
  if (slot_to_insertion)
    HIP_TRY(hipMemcpy(slot_to_insertion, s->pi + s->cap, sizeof(int) * s->n,
                      hipMemcpyDeviceToHost));
  return MPM_OK;
}

int mpm_set_debug(mpm_solver* s, int32_t flags) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  s->debug = flags;
  return MPM_OK;
}

int mpm_set_timing(mpm_solver* s, int32_t enabled) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(check_device_errors(s));
  s->timing = enabled != 0;
  s->timer.acc.clear();
  s->timer.order.clear();
  return MPM_OK;
}

int mpm_kernel_times(mpm_solver* s, char* names, int32_t names_cap, int32_t* counts, double* ms,
                     int32_t cap, int32_t* n_kernels) {
  if (!s || !n_kernels) return fail(MPM_EINVAL, "null argument");
  TRY(check_device_errors(s));
  std::string all;
  int k = 0;
  for (auto& name : s->timer.order) {
    if (k < cap) {
      if (counts) counts[k] = s->timer.acc[name].first;
      if (ms) ms[k] = s->timer.acc[name].second;
    }
    all += name;
    all += "\n";
    k++;
  }
  *n_kernels = k;
  if (names && names_cap > 0) {
    std::strncpy(names, all.c_str(), names_cap - 1);
    names[names_cap - 1] = 0;
  }
  return MPM_OK;
}

int mpm_host_svd(int32_t dim, const float* A, float* U, float* sig, float* V) {
  if (!A || !U || !sig || !V || (dim != 2 && dim != 3)) return fail(MPM_EINVAL, "bad argument");
  float a[3][3] = {{0}}, u[3][3], v[3][3], s3[3] = {0, 0, 0};
  for (int i = 0; i < dim; i++)
    for (int j = 0; j < dim; j++) a[i][j] = A[i * dim + j];
  if (dim == 3)
    svd3(a, u, s3, v);
  else
    svd2(a, u, s3, v);
  for (int i = 0; i < dim; i++) {
    sig[i] = s3[i];
    for (int j = 0; j < dim; j++) {
      U[i * dim + j] = u[i][j];
      V[i * dim + j] = v[i][j];
    }
  }
  return MPM_OK;
}

int mpm_device_svd(int32_t device, int32_t dim, int64_t n, const float* A, float* U, float* sig,
                   float* V) {
  if ((dim != 2 && dim != 3) || n < 0) return fail(MPM_EINVAL, "bad argument");
  if (n == 0) return MPM_OK;
  if (!A || !U || !sig || !V) return fail(MPM_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(device));
  const int64_t dd = (int64_t)dim * dim;
  DeviceArray<float> a, u, sg, v;
  TRY(a.alloc(n * dd));
  TRY(u.alloc(n * dd));
  TRY(v.alloc(n * dd));
  TRY(sg.alloc(n * dim));
  HIP_TRY(hipMemcpy(a.p, A, sizeof(float) * n * dd, hipMemcpyHostToDevice));
  launch_svd_batch(dim, a.p, u.p, sg.p, v.p, n, nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(U, u.p, sizeof(float) * n * dd, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(V, v.p, sizeof(float) * n * dd, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(sig, sg.p, sizeof(float) * n * dim, hipMemcpyDeviceToHost));
  return MPM_OK;
}

int mpm_last_counts(mpm_solver* s, int64_t* out4) {
  if (!s || !out4) return fail(MPM_EINVAL, "null argument");
  TRY(check_device_errors(s));
  int tot[16];
  HIP_TRY(hipMemcpy(tot, s->totals + s->last_parity * 16, sizeof(int) * 16,
                    hipMemcpyDeviceToHost));
  out4[0] = tot[kTotValid];
  out4[1] = tot[kTotChunks];
  out4[2] = tot[kTotActive];
  out4[3] = tot[kTotPBlocks];
  return MPM_OK;
}

// ---------------------------------------------------------------------------
// differentiable mode
// ---------------------------------------------------------------------------
int mpm_diff_setup(mpm_solver* s, int64_t n, const int32_t* material) {
  if (!s || (n > 0 && !material)) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->n != 0) return fail(MPM_EINVAL, "a differentiable handle holds no particles of its own");
  if (s->scenes.size() != 1) return fail(MPM_EINVAL, "differentiable mode runs one scene");
  if (n <= 0 || n > (int64_t)INT32_MAX / 2) return fail(MPM_EINVAL, "bad particle count");
  for (int64_t i = 0; i < n; i++)
    if (material[i] < MPM_WATER || material[i] > MPM_SAND)
      return fail(MPM_EINVAL,
                  "differentiable mode supports water, elastic, snow and sand particles");
  if (s->g2p2g) return fail(MPM_EINVAL, "differentiable mode uses the p2g / g2p transfers");
  if (s->unbounded) return fail(MPM_EINVAL, "differentiable mode keeps its block window");
  if (s->quant || !s->plastic)
    return fail(MPM_EINVAL, "differentiable mode uses the float particle layout with plasticity");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  TRY(grow(s, n));
  dfree(s->diff_pi);
  dfree(s->diff_scratch);
  dfree(s->diff_gbar);
  dfree(s->diff_abuf);
  dfree(s->diff_nu_part);
  const int NPB = s->D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
  TRY(dalloc(&s->diff_pi, 3 * n));
  TRY(dalloc(&s->diff_scratch, (int64_t)s->nf() * n));
  TRY(dalloc(&s->diff_gbar, (int64_t)s->total_blocks * NPB));
  TRY(dalloc(&s->diff_nu_part, s->persistent_groups));
  TRY(dalloc(&s->diff_abuf, (int64_t)s->D * s->D * n));
  std::vector<int32_t> host(3 * n);
  for (int64_t i = 0; i < n; i++) {
    host[i] = material[i];
    host[n + i] = (int32_t)i;
    host[2 * n + i] = 0;
  }
  HIP_TRY(hipMemcpy(s->diff_pi, host.data(), sizeof(int32_t) * host.size(),
                    hipMemcpyHostToDevice));
  s->diff_n = n;
  s->diff_params_valid = false;
  return MPM_OK;
}

namespace {

// scene table for a differentiable call, uploaded on st only when it changed
int diff_params(mpm_solver* s, float dt, const float* youngs, double nu, hipStream_t st) {
  if (!(dt > 0.0f) || !std::isfinite(dt)) return fail(MPM_EINVAL, "dt must be positive");
  if (!youngs) return fail(MPM_EINVAL, "null Young's modulus array");
  if (!(nu > -1.0 && nu < 0.5)) return fail(MPM_EINVAL, "nu must lie in (-1, 0.5)");
  if (s->diff_params_valid && s->diff_dt == dt && s->diff_youngs == youngs && s->diff_nu == nu)
    return MPM_OK;
  HIP_TRY(hipStreamSynchronize(st));  // the pinned staging copy may still be in flight
  Params& P = s->h_params[0];
  fill_params(P, s, s->scenes[0], dt, 0);
  P.youngs = youngs;
  P.kmu = (float)(1.0 / (2.0 * (1.0 + nu)));
  P.kla = (float)(nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
  P.dkmu = (float)(-1.0 / (2.0 * (1.0 + nu) * (1.0 + nu)));
  const double g = (1.0 + nu) * (1.0 - 2.0 * nu);
  P.dkla = (float)((1.0 + 2.0 * nu * nu) / (g * g));
  // sand's return-map coefficient (d lambda_0 + 2 mu_0) / (2 mu_0)
  // (mpm_solver.py:336-337) with this nu: E cancels, nu does not
  {
    const double kmu = 1.0 / (2.0 * (1.0 + nu)), kla = nu / g;
    const double dkmu = -1.0 / (2.0 * (1.0 + nu) * (1.0 + nu)), dkla = (1.0 + 2.0 * nu * nu) / (g * g);
    P.sand_coef = (float)((s->D * kla + 2.0 * kmu) / (2.0 * kmu));
    P.dsand_coef = (float)(0.5 * s->D * (dkla * kmu - kla * dkmu) / (kmu * kmu));
  }
  HIP_TRY(hipMemcpyAsync(s->d_params, s->h_params, sizeof(Params), hipMemcpyHostToDevice, st));
  s->diff_params_valid = true;
  s->diff_dt = dt;
  s->diff_youngs = youngs;
  s->diff_nu = nu;
  return MPM_OK;
}

// Saved binning of a differentiable substep: perm, chunks, per-block chunk
// ranges, active blocks and totals, each 256-byte aligned in one blob.
// Per-substep record of the differentiable forward, reused by its backward:
// the binning (sorted permutation, chunks, per-block chunk ranges, active
// list, totals) and the raw p2g grid sums (momentum, mass) of the active
// nodes, so that the backward needs neither a sort nor a p2g.
struct BinLayout {
  int64_t perm, chunks, pbeg, pcnt, active, totals, grid, bytes;
};

BinLayout bin_layout(const mpm_solver* s) {
  auto al = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
  const int64_t n = s->diff_n, tb = s->total_blocks;
  BinLayout L;
  L.perm = 0;
  L.chunks = al(4 * n);
  L.pbeg = L.chunks + al(16 * s->chunk_cap);
  L.pcnt = L.pbeg + al(4 * tb);
  L.active = L.pcnt + al(4 * tb);
  L.totals = L.active + al(4 * tb);
  L.grid = L.totals + 256;
  L.bytes = L.grid + 16 * tb * (s->D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock);
  return L;
}

void diff_buffers(mpm_solver* s, SubstepBuffers& B, float* pf) {
  fill_buffers(s, B);
  B.step = nullptr;
  B.pf = pf;
  B.pf_out = pf;
  B.pi = s->diff_pi;
  B.cap = s->diff_n;
  B.n = (int)s->diff_n;
  B.full_sort = true;  // every call is a function of its input state alone
  B.perm_identity = false;
}

}  // namespace

int mpm_set_g2p2g(mpm_solver* s, int32_t enabled, double allowed_cfl, int32_t clamp) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->substep_counter != 0 || s->diff_n != 0)
    return fail(MPM_EINVAL, "choose the transfer mode before the first substep");
  if (!(allowed_cfl >= 0.0)) return fail(MPM_EINVAL, "g2p2g_allowed_cfl must be >= 0");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int want = enabled ? 1 : 0;
  if (want != s->g2p2g && s->quant) TRY(drop_particle_layout(s));  // C present or not
  s->g2p2g_cfl = allowed_cfl;
  s->g2p2g_clamp = clamp ? 1 : 0;
  if (want != s->g2p2g) {
    s->g2p2g = want;
    dfree(s->grid_alt);
    if (want) {
      const int NPB = s->D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
      const int64_t nodes = (int64_t)s->total_blocks * NPB;
      TRY(dalloc(&s->grid_alt, nodes));
      HIP_TRY(hipMemset(s->grid_alt, 0, sizeof(float4) * nodes));
      HIP_TRY(hipMemset(s->grid, 0, sizeof(float4) * nodes));
    }
    if (s->cap > 0) TRY(alloc_chunk_buffers(s));  // tile size depends on the mode
    s->n_final = 0;
  }
  return MPM_OK;
}

int mpm_set_storage(mpm_solver* s, int32_t quant, int32_t support_plasticity) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->diff_n != 0) return fail(MPM_EINVAL, "the differentiable mode keeps the float layout");
  const int q = quant ? 1 : 0, pl = support_plasticity ? 1 : 0;
  if (q == s->quant && pl == s->plastic) return MPM_OK;
  HIP_TRY(hipSetDevice(s->device));
  TRY(drop_particle_layout(s));
  s->quant = q;
  s->plastic = pl;
  return MPM_OK;
}

int mpm_particle_bytes(mpm_solver* s, int32_t* bytes) {
  if (!s || !bytes) return fail(MPM_EINVAL, "null argument");
  const int ni = s->D == 3 ? Layout<3>::NI : Layout<2>::NI;
  *bytes = 4 * (s->nf() + ni) + 4;  // device words + the host-side colour
  return MPM_OK;
}

int mpm_max_grid_velocity(mpm_solver* s, float* out) {
  if (!s || !out) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->scenes.size() != 1) return fail(MPM_EINVAL, "the grid of a batch holds many scenes");
  HIP_TRY(hipSetDevice(s->device));
  *out = 0.0f;
  if (s->n == 0) return MPM_OK;  // an empty grid, as before the first substep
  launch_grid_vmax(s->D, s->grid, s->active, s->totals + s->last_parity * 16, s->total_blocks,
                   s->d_gvmax, s->stream);
  unsigned bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, s->d_gvmax, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  std::memcpy(out, &bits, sizeof(float));
  return MPM_OK;
}

int mpm_set_unbounded(mpm_solver* s, int32_t enabled) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (enabled && s->scenes.size() != 1)
    return fail(MPM_EINVAL, "unbounded scenes are supported for one-scene solvers");
  if (enabled && s->diff_n != 0)
    return fail(MPM_EINVAL, "the differentiable mode keeps its block window");
  s->unbounded = enabled ? 1 : 0;
  // grid_size: 4096, doubled until it exceeds twice the largest resolution
  // (mpm_solver.py:74, 143-148); the walls sit at +-grid_size/2
  const mpm_config& c = s->scenes[0].cfg;
  int64_t grid_size = 4096;
  int rmax = 0;
  for (int a = 0; a < s->D; a++) rmax = std::max(rmax, c.res[a]);
  if (enabled)
    while (grid_size <= 2 * (int64_t)rmax) grid_size *= 2;
  s->grid_half = (int)(grid_size / 2);
  return MPM_OK;
}

int mpm_set_graphs(mpm_solver* s, int32_t enabled) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  s->use_graphs = enabled != 0;
  if (!s->use_graphs) drop_graphs(s);
  return MPM_OK;
}

int mpm_set_tuning(mpm_solver* s, int32_t chunk, int32_t groups) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (chunk != 0 && (chunk < 16 || chunk > kChunkMax))
    return fail(MPM_EINVAL, "chunk must lie in [16, " + std::to_string(kChunkMax) + "]");
  if (groups != 0 && (groups < 64 || groups > 16384))
    return fail(MPM_EINVAL, "groups must lie in [64, 16384]");
  if (s->diff_n != 0) return fail(MPM_EINVAL, "tune a differentiable handle before mpm_diff_setup");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (groups != 0 && groups != s->persistent_groups) {
    drop_graphs(s);
    dfree(s->maxv);
    s->persistent_groups = groups;
    TRY(dalloc(&s->maxv, 16 * (int64_t)groups));
    s->maxv_cap = 16 * groups;
    launch_fill_u32(s->maxv, 0u, s->maxv_cap, s->stream);
    s->last_batch = 0;
  }
  if (chunk != 0 && chunk != s->chunk) {
    s->chunk = chunk;
    if (s->cap > 0) TRY(alloc_chunk_buffers(s));
  }
  // the next substep bins from scratch (chunk boundaries changed)
  s->have_perm = false;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int mpm_snapshot(mpm_solver* s, int32_t op) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->diff_n != 0) return fail(MPM_EINVAL, "a differentiable handle keeps its state in caller arrays");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int ni = s->D == 3 ? Layout<3>::NI : Layout<2>::NI;
  const int NPB = s->D == 3 ? Geo<3>::kNodesPerBlock : Geo<2>::kNodesPerBlock;
  if (op == MPM_SNAPSHOT_RELEASE) {
    dfree(s->snap_pf);
    dfree(s->snap_pi);
    dfree(s->snap_grid);
    s->snap_n = -1;
    return MPM_OK;
  }
  if (op == MPM_SNAPSHOT_SAVE) {
    TRY(check_device_errors(s));
    const int64_t nodes = s->g2p2g ? (int64_t)s->total_blocks * NPB : 0;
    if (s->snap_cap != s->cap || s->snap_nf != s->nf() || !s->snap_pf) {
      dfree(s->snap_pf);
      dfree(s->snap_pi);
      s->snap_cap = 0;
      if (s->cap > 0) {
        TRY(dalloc(&s->snap_pf, (int64_t)s->nf() * s->cap));
        TRY(dalloc(&s->snap_pi, (int64_t)ni * s->cap));
      }
      s->snap_cap = s->cap;
      s->snap_nf = s->nf();
    }
    if (s->snap_grid_nodes != nodes) {
      dfree(s->snap_grid);
      s->snap_grid_nodes = 0;
      if (nodes > 0) TRY(dalloc(&s->snap_grid, nodes));
      s->snap_grid_nodes = nodes;
    }
    if (s->cap > 0) {
      HIP_TRY(hipMemcpyAsync(s->snap_pf, s->pf, sizeof(float) * s->nf() * s->cap,
                             hipMemcpyDeviceToDevice, s->stream));
      HIP_TRY(hipMemcpyAsync(s->snap_pi, s->pi, sizeof(int) * ni * s->cap,
                             hipMemcpyDeviceToDevice, s->stream));
    }
    if (nodes > 0)
      HIP_TRY(hipMemcpyAsync(s->snap_grid, s->grid, sizeof(float4) * nodes,
                             hipMemcpyDeviceToDevice, s->stream));
    s->snap_n = s->n;
    s->snap_n_final = s->n_final;
    for (int a = 0; a < 3; a++) {
      s->snap_blo[a] = s->scenes[0].blo[a];
      s->snap_bnl[a] = s->scenes[0].bnl[a];
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MPM_OK;
  }
  if (op != MPM_SNAPSHOT_RESTORE) return fail(MPM_EINVAL, "unknown snapshot operation");
  if (s->snap_n < 0) return fail(MPM_EINVAL, "no snapshot saved");
  if (s->snap_n != s->n || s->snap_cap != s->cap || s->snap_nf != s->nf())
    return fail(MPM_EINVAL, "particles were added or the storage changed since the snapshot");
  if (s->unbounded) {  // the g2p2g grid belongs to the window it was saved with
    bool moved = false;
    for (int a = 0; a < s->D; a++)
      moved |= s->scenes[0].blo[a] != s->snap_blo[a] || s->scenes[0].bnl[a] != s->snap_bnl[a];
    if (moved) TRY(set_window(s, s->snap_blo, s->snap_bnl));
  }
  if (s->g2p2g && s->snap_grid_nodes != (int64_t)s->total_blocks * NPB)
    return fail(MPM_EINVAL, "the grid window changed since the snapshot");
  if (s->cap > 0) {
    HIP_TRY(hipMemcpyAsync(s->pf, s->snap_pf, sizeof(float) * s->nf() * s->cap,
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->pi, s->snap_pi, sizeof(int) * ni * s->cap,
                           hipMemcpyDeviceToDevice, s->stream));
  }
  if (s->g2p2g && s->snap_grid_nodes > 0)
    HIP_TRY(hipMemcpyAsync(s->grid, s->snap_grid, sizeof(float4) * s->snap_grid_nodes,
                           hipMemcpyDeviceToDevice, s->stream));
  s->n_final = s->snap_n_final;
  s->have_perm = false;  // slots moved: bin from scratch
  s->reordered = false;
  // the saved slots need not be in block order (a snapshot taken right after
  // particles were added holds insertion order, which makes every transfer's
  // particle loads scattered): reorder at the first opportunity
  s->since_reorder = 1 << 30;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int mpm_set_g2p_waves(mpm_solver* s, int32_t waves) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (waves != 0 && waves != 6) return fail(MPM_EINVAL, "g2p waves per SIMD must be 0 or 6");
  s->g2p_waves = waves;
  return MPM_OK;
}

int mpm_get_g2p_waves(mpm_solver* s, int32_t* waves) {
  if (!s || !waves) return fail(MPM_EINVAL, "null argument");
  *waves = s->g2p_waves;
  return MPM_OK;
}

int mpm_get_tuning(mpm_solver* s, int32_t* chunk, int32_t* groups) {
  if (!s || !chunk || !groups) return fail(MPM_EINVAL, "null argument");
  *chunk = s->chunk;
  *groups = s->persistent_groups;
  return MPM_OK;
}

int mpm_diff_binning_bytes(mpm_solver* s, int64_t* bytes) {
  if (!s || !bytes) return fail(MPM_EINVAL, "null argument");
  if (s->diff_n == 0) return fail(MPM_EINVAL, "mpm_diff_setup first");
  *bytes = bin_layout(s).bytes;
  return MPM_OK;
}

int mpm_diff_substep(mpm_solver* s, float dt, const float* state_in, float* state_out,
                     const float* youngs, double nu, void* binning, void* stream) {
  if (!s || !state_in || !state_out) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->diff_n == 0) return fail(MPM_EINVAL, "mpm_diff_setup first");
  hipStream_t st = (hipStream_t)stream;  // NULL: the null stream
  HIP_TRY(hipSetDevice(s->device));
  TRY(diff_params(s, dt, youngs, nu, st));
  SubstepBuffers B;
  diff_buffers(s, B, const_cast<float*>(state_in));
  B.pf_out = state_out;  // p2g / g2p write the new state out of place
  if (binning) B.raw_out = (float4*)((char*)binning + bin_layout(s).grid);
  if (s->D == 3)
    launch_substep<3>(s->d_params, s->d_slab, B, 0, 0, st, nullptr);
  else
    launch_substep<2>(s->d_params, s->d_slab, B, 0, 0, st, nullptr);
  if (binning) {
    const BinLayout L = bin_layout(s);
    char* b = (char*)binning;
    const int64_t tb = s->total_blocks;
    CopySegments S{};
    const int* src[6] = {B.perm, (const int*)B.chunks, B.pchunk_begin, B.pchunk_cnt, B.active,
                         B.totals};
    const int64_t off[6] = {L.perm, L.chunks, L.pbeg, L.pcnt, L.active, L.totals};
    const int64_t cnt[6] = {s->diff_n, 4 * s->chunk_cap, tb, tb, tb, 16};
    for (int i = 0; i < 6; i++) {
      S.src[i] = src[i];
      S.dst[i] = (int*)(b + off[i]);
      S.n[i] = cnt[i];
    }
    S.count = 6;
    launch_copy_segments(S, st);
  }
  take_buffers(s, B);
  s->substep_counter++;
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}

int mpm_diff_substep_noalloc(mpm_solver* s, float dt, const float* state_in, float* state_out,
                             const float* youngs, double nu, void* stream) {
  if (!s || !state_in || !state_out) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->diff_n == 0) return fail(MPM_EINVAL, "mpm_diff_setup first");
  Carve need{nullptr};
  SubstepBuffers B;
  carve_diff_workspace(s, need, B);
  if (need.off > kScratchBytes)
    return fail(MPM_EINVAL, "mpm_diff_substep_noalloc: this scene needs " +
                                std::to_string(need.off) + " bytes of workspace, beyond the " +
                                std::to_string(kScratchBytes) +
                                " bytes of the static workspace (use mpm_diff_substep)");
  char* base = (char*)g_scratch_ptr[s->device];
  if (!base) return fail(MPM_EINVAL, "the workspace of this device is not resolved");
  hipStream_t st = (hipStream_t)stream;  // NULL: the null stream
  HIP_TRY(hipSetDevice(s->device));
  TRY(diff_params(s, dt, youngs, nu, st));
  diff_buffers(s, B, const_cast<float*>(state_in));
  B.pf_out = state_out;  // p2g / g2p write the new state out of place
  Carve c{base};
  carve_diff_workspace(s, c, B);
  B.raw_out = nullptr;
  B.step = nullptr;
  B.parity = 0;
  // what the kernels expect to start at zero (the workspace keeps the last
  // call's contents): the digit histograms and the look-back scan and sort
  // state (the block counts are cleared by the binning itself); fill kernels,
  // not memsets, so that a captured graph replays them correctly
  const int64_t tb = s->total_blocks;
  const int64_t nparts = (tb + kScanTile - 1) / kScanTile;
  const int64_t rtiles = (s->diff_n + kRadixTile - 1) / kRadixTile;
  launch_fill_u32((unsigned*)B.ghist, 0u, 2 * kMaxRadixPasses * kRadixBins, st);
  launch_fill_u32((unsigned*)B.lb_status_scan, 0u, 2 * 4 * nparts, st);
  launch_fill_u32((unsigned*)B.lb_status_radix, 0u, 2 * kRadixBins * rtiles, st);
  launch_fill_u32((unsigned*)B.lbs, 0u, sizeof(LBState) / 4, st);
  if (s->D == 3)
    launch_substep<3>(s->d_params, s->d_slab, B, 0, 0, st, nullptr);
  else
    launch_substep<2>(s->d_params, s->d_slab, B, 0, 0, st, nullptr);
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}

int mpm_diff_substep_backward(mpm_solver* s, float dt, const float* state_in,
                              const float* youngs, double nu, const void* binning,
                              const float* grad_out, float* grad_in, float* grad_youngs,
                              float* grad_nu, void* stream) {
  if (!s || !state_in || !grad_out || !grad_in || !grad_nu)
    return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (s->diff_n == 0) return fail(MPM_EINVAL, "mpm_diff_setup first");
  hipStream_t st = (hipStream_t)stream;  // NULL: the null stream
  HIP_TRY(hipSetDevice(s->device));
  TRY(diff_params(s, dt, youngs, nu, st));
  // the forward's binning and raw grid sums: from its record, or recomputed
  // from the old state (p2g then writes its F / Jp into scratch)
  SubstepBuffers B;
  diff_buffers(s, B, const_cast<float*>(state_in));
  B.pf_out = s->diff_scratch;
  if (binning) {
    const BinLayout L = bin_layout(s);
    char* b = (char*)binning;
    B.perm = (int*)(b + L.perm);
    B.chunks = (int4*)(b + L.chunks);
    B.pchunk_begin = (int*)(b + L.pbeg);
    B.pchunk_cnt = (int*)(b + L.pcnt);
    B.active = (int*)(b + L.active);
    B.totals = (int*)(b + L.totals);
    B.grid = (float4*)(b + L.grid);
  } else {
    if (s->D == 3)
      launch_front<3>(s->d_params, s->d_slab, B, 1, st, nullptr);
    else
      launch_front<2>(s->d_params, s->d_slab, B, 1, st, nullptr);
  }
  if (s->D == 3)
    launch_adjoint<3>(s->d_params, s->d_slab, B, state_in, grad_out, grad_in, grad_youngs,
                      s->diff_gbar, s->diff_abuf, s->diff_nu_part, grad_nu, st, nullptr);
  else
    launch_adjoint<2>(s->d_params, s->d_slab, B, state_in, grad_out, grad_in, grad_youngs,
                      s->diff_gbar, s->diff_abuf, s->diff_nu_part, grad_nu, st, nullptr);
  if (!binning) {  // the block scan ran: keep the substep-parity bookkeeping
    take_buffers(s, B);
    s->substep_counter++;
  }
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}

int mpm_diff_sync(mpm_solver* s, void* stream) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return check_device_errors(s);
}

int mpm_quantize_particles(mpm_solver* s, float* ranges, uint32_t* x_and_v, uint8_t* color) {
  if (!s || !ranges || !x_and_v || !color) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  // np.min over no particles raises ValueError (particle_io.py:42)
  if (s->n == 0) return fail(MPM_EINVAL, "no particles to write");
  TRY(check_device_errors(s));
  HIP_TRY(hipSetDevice(s->device));
  const int D = s->D, NC = 2 * D, n = (int)s->n;
  const int groups = (int)std::min<int64_t>((s->n + kThreads - 1) / kThreads, 1024);
  float* partial = nullptr;
  unsigned* out = nullptr;
  HIP_TRY(hipMalloc((void**)&partial, sizeof(float) * groups * NC * 2));
  if (hipMalloc((void**)&out, sizeof(unsigned) * s->n * D) != hipSuccess) {
    hipFree(partial);
    return fail(MPM_ENOMEM, "device allocation for the quantised particles failed");
  }
  std::vector<float> part((size_t)groups * NC * 2);
  DeviceArray<float> flt;  // packed storage: read x and v from the float layout
  const float* src = s->pf;
  if (s->quant) {
    const int r = unpacked(s, flt);
    if (r != MPM_OK) {
      hipFree(partial);
      hipFree(out);
      return r;
    }
    src = flt.p;
  }
  launch_minmax(D, src, s->cap, n, partial, groups, s->stream);
  hipError_t e = hipMemcpyAsync(part.data(), partial, sizeof(float) * part.size(),
                                hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  QuantRanges q{};
  if (e == hipSuccess) {
    for (int c = 0; c < NC; c++) {
      float lo = part[c * 2], hi = part[c * 2 + 1];
      for (int g = 1; g < groups; g++) {
        lo = std::min(lo, part[((size_t)g * NC + c) * 2]);
        hi = std::max(hi, part[((size_t)g * NC + c) * 2 + 1]);
      }
      // max(lo + 1e-5, hi) with the float32 lo promoted to a float64 sum and
      // the result stored back into the float32 array (particle_io.py:45-47)
      const double widened = (double)lo + 1e-5;
      if (widened > (double)hi) hi = (float)widened;
      const int kind = c < D ? 0 : 1, d = c < D ? c : c - D;
      ranges[(kind * D + d) * 2 + 0] = lo;
      ranges[(kind * D + d) * 2 + 1] = hi;
      q.lo[c] = lo;
      const float width = hi - lo;  // float32 subtraction
      q.inv[c] = (float)(1.0 / (double)width);  // correctly rounded 1 / width
    }
    launch_quantize(D, src, s->pi + s->cap /* Layout ORIG == 1 */, s->cap, n, q, out, s->stream);
    e = hipMemcpyAsync(x_and_v, out, sizeof(unsigned) * s->n * D, hipMemcpyDeviceToHost,
                       s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  }
  hipFree(partial);
  hipFree(out);
  if (e != hipSuccess) return fail(MPM_EHIP, std::string("quantising particles: ") + hipGetErrorString(e));
  // colours live on the host: int32 -> float32 -> uint32 as the reference's
  // float32 staging array does (particle_io.py:62-74), then R, G, B bytes
  for (int64_t i = 0; i < s->n; i++) {
    const float f = (float)s->color[i];
    const uint32_t u = (uint32_t)(int64_t)f;
    for (int c = 0; c < 3; c++) color[i * 3 + c] = (uint8_t)((u >> (8 * (2 - c))) & 255u);
  }
  return check_device_errors(s);
}


int mpm_voxelize(int32_t device, const int32_t* res, double dx, int32_t super_sample,
                 int32_t padding, int64_t ntri, const double* triangles, int32_t* voxel_res,
                 int32_t* voxels) {
  if (!res || !voxel_res) return fail(MPM_EINVAL, "null argument");
  VoxGrid g{};
  TRY(voxel_grid(res, dx, super_sample, padding, &g));
  for (int k = 0; k < 3; k++) voxel_res[k] = g.res[k];
  if (!voxels) return MPM_OK;
  HIP_TRY(hipSetDevice(device));
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const int64_t nvox = (int64_t)g.res[0] * g.res[1] * g.res[2];
  int rc = MPM_OK;
  {
    DeviceArray<int> events, vox;
    rc = voxel_events(g, ntri, triangles, st, events);
    if (rc == MPM_OK) rc = vox.alloc(nvox);
    if (rc == MPM_OK) {
      SeedParams sp{};
      sp.sample_density = -1;
      launch_vox_columns(g, events.p, vox.p, sp, nullptr, nullptr, nullptr, 0, st);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(voxels, vox.p, sizeof(int) * nvox, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) rc = fail(MPM_EHIP, std::string("voxelize: ") + hipGetErrorString(e));
    }
    hipStreamSynchronize(st);
  }
  hipStreamDestroy(st);
  return rc;
}

int mpm_add_mesh(mpm_solver* s, int64_t ntri, const double* triangles, int32_t material,
                 int32_t color, int32_t sample_density, const float* velocity,
                 const float* translation, int32_t super_sample, uint32_t seed, int64_t* added) {
  if (!s) return fail(MPM_EINVAL, "null argument");
  TRY(no_half_pending(s));
  if (added) *added = 0;
  if (s->D != 3) return fail(MPM_EINVAL, "add_mesh needs a 3D solver");
  if (material < 0 || material > MPM_STATIONARY)
    return fail(MPM_EINVAL, "unknown material id " + std::to_string(material));
  if (sample_density < 0 || sample_density > 4096)
    return fail(MPM_EINVAL, "sample_density must be 0..4096");
  TRY(check_device_errors(s));
  HIP_TRY(hipSetDevice(s->device));
  const mpm_config& cfg = s->scenes[s->active_scene].cfg;
  VoxGrid g{};
  TRY(voxel_grid(cfg.res, cfg.dx, super_sample, cfg.padding, &g));
  SeedParams sp{};
  sp.sample_density = sample_density;
  sp.ss3 = super_sample * super_sample * super_sample;
  sp.p_limit = 4096 / 2 - cfg.padding;  // bounded grid_size (mpm_solver.py:74, 1026-1029)
  sp.seed = seed;
  sp.pdx = (float)(cfg.dx / super_sample);
  for (int k = 0; k < 3; k++) sp.shift[k] = translation ? translation[k] : 0.0f;
  const int64_t ncol = (int64_t)g.res[0] * g.res[1];

  DeviceArray<int> events, counts;
  TRY(voxel_events(g, ntri, triangles, s->stream, events));
  TRY(counts.alloc(ncol));
  launch_vox_columns(g, events.p, nullptr, sp, counts.p, nullptr, nullptr, 0, s->stream);
  HIP_TRY(hipGetLastError());
  std::vector<int> hcount(ncol);
  HIP_TRY(hipMemcpyAsync(hcount.data(), counts.p, sizeof(int) * ncol, hipMemcpyDeviceToHost,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  std::vector<int64_t> offsets(ncol);
  int64_t total = 0;
  for (int64_t c = 0; c < ncol; c++) {
    offsets[c] = total;
    total += hcount[c];
  }
  if (total == 0) return MPM_OK;
  if (s->n + total > (int64_t)INT32_MAX)
    return fail(MPM_EINVAL, "mesh seeds " + std::to_string(total) + " particles: too many");
  TRY(grow(s, s->n + total));
  DeviceArray<int64_t> doff;
  DeviceArray<float> dx, dv;
  DeviceArray<int> dm;
  TRY(doff.alloc(ncol));
  TRY(dx.alloc(total * 3));
  TRY(dv.alloc(total * 3));
  TRY(dm.alloc(total));
  HIP_TRY(hipMemcpyAsync(doff.p, offsets.data(), sizeof(int64_t) * ncol, hipMemcpyHostToDevice,
                         s->stream));
  launch_vox_columns(g, events.p, nullptr, sp, nullptr, doff.p, dx.p, total, s->stream);
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  launch_seed_attrs(total, velocity ? velocity : zero, material, dv.p, dm.p, s->stream);
  SubstepBuffers B;
  fill_buffers(s, B);
  launch_append<3>(B, dx.p, dv.p, dm.p, (int)total, (int)s->n, s->active_scene, s->plastic,
                   s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->color.insert(s->color.end(), (size_t)total, color);
  s->n += total;
  s->have_perm = false;
  s->since_reorder = 1 << 30;  // the new particles sit at the end: reorder soon
  if (added) *added = total;
  return check_device_errors(s);
}

}  // extern "C"

// C-ABI layer of libuclbs_hip.so: context, level state resident in HBM,
// host<->device population transfer and step dispatch.  The exported entry
// points are declared in include/uclbs_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/uclbs_hip.h"
#include "kernels.h"
#include "lattice.h"
#include "philox.h"
#include "cluster.h"

// Staging area of uclbs_mlp_forward_noalloc: the features (12 floats per
// row) and the network outputs (1 float per row) of at most
// kNoAllocRows rows.  One copy per device, held until the process ends; every
// call on a device shares it.  Host code never uses this name as an address
// (it is a host-side stand-in): the device address comes from
// hipGetSymbolAddress, once per device, in uclbs_create.
constexpr int kNoAllocRows = 65536;
constexpr size_t kScratchFloats = (size_t)kNoAllocRows * 13;
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

thread_local std::string g_err;

// device address of `scratch` per device, resolved once per device
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

// with `device` current: the device's copy of `scratch`
hipError_t scratch_address(int device, float** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err(UCLBS_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                     __FILE__, __LINE__);                                                 \
  } while (0)

struct Level {
  bool defined = false;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  uclbs::Box box{};
  void* f[2] = {nullptr, nullptr};
  int cur = 0;  // index of the buffer holding the current post-collision state
  double* alpha = nullptr;
  unsigned* defer_count = nullptr;  // far-from-equilibrium queue: 2 x (kFarShards + 1) lengths
  uint2* defer_list = nullptr;
  unsigned defer_cap = 0;
  int far_parity = 0;  // which half of defer_count this step appends to
  // link tables (reference form) and the compact form the kernels use
  int32_t face_code[6] = {0, 0, 0, 0, 0, 0};
  int32_t* state = nullptr;   // [27][n]
  uint8_t* flags = nullptr;   // [n]
  uint32_t* blist = nullptr;  // boundary fluid nodes, ascending
  uint32_t* bbox_of = nullptr;  // a level of many boxes: the box of each boundary node
  int64_t nb = 0, nsolid = 0;
  bool has_links = false;
  // D3Q39: second population, multipliers, repair list, step counter
  void* g[2] = {nullptr, nullptr};
  double* lm = nullptr;
  unsigned* neg_count = nullptr;
  unsigned* neg_list = nullptr;
  unsigned* near_list = nullptr;  // D3Q39 near-face nodes (rebuilt when the faces change)
  bool table39 = false;           // a D3Q39 link table is resident
  uint8_t* cls39 = nullptr;       // shifted D3Q39 table: node classes (0 plain, 1 near a face,
                                  // 2 wall links, 3 solid)
  unsigned* wall_list = nullptr;  // nodes of class 2
  unsigned n_wall = 0;
  double* wall_f = nullptr;       // rebuilt f, g of the wall links, [n_wall][39] each
  double* wall_g = nullptr;
  int ghost_depth = 0;            // ghost layer of a refined level (1 on D3Q27, 4 on D3Q39)
  uclbs::Gas39 gas39{};           // D3Q39 constants of this level (dt, epsilon, beta scaling)
  unsigned n_near = 0;
  bool near_dirty = true;
  double* rep = nullptr;
  int64_t step = 0;
  // placement in the level's global index space; levels >= 1 carry a
  // one-node ghost shell filled from the coarser level
  int origin[3] = {0, 0, 0};
  uint32_t* ghosts = nullptr;
  unsigned nghost = 0;
  bool refined = false;
  bool stale = false;   // its coarser level was regridded: regrid or define it again
  bool relink = false;  // regridded from a level with a link table: build or set its links
  // a level of many boxes (uclbs_define_level_boxes)
  bool multibox = false;
  std::vector<int32_t> blo, bhi;  // valid region of each box, [nbox][3], level index space
  std::vector<uclbs::BoxDesc> hbox;
  uclbs::BoxDesc* d_boxes = nullptr;
  uclbs::BoxTile* d_tiles = nullptr;
  unsigned ntiles = 0;
  int tile_nodes = 256;
  uint3* d_copies = nullptr;  // same-level ghost copies {dst, src, population mask}
  uint4* d_alias = nullptr;   // {mirrored valid node, its box strides, 0} of every kFlagMirror node
  unsigned ncopies = 0;
  uclbs::C2FStencil* d_c2f = nullptr;  // ghosts interpolated from the coarser level
  unsigned nc2f = 0;
  uint3* d_inject = nullptr;  // {coarse node, fine node, all populations} of fine_to_coarse
  unsigned ninject = 0;
  int dom_lo[3] = {0, 0, 0}, dom_hi[3] = {0, 0, 0};  // level 0: the periodic domain
  // host lookup of the box holding a level position: buckets of edge bk_e over
  // the bounding box of the valid regions, each listing the boxes it meets
  int bk_lo[3] = {0, 0, 0}, bk_n[3] = {0, 0, 0}, bk_e = 8;
  std::vector<std::vector<int>> buckets;
};

// a microphone (mic_setup.f95:45-225): its level, the eight corners of the
// cell it falls in and their weights
struct Mic {
  int number = 0;     // 1-based, names the file microphoneNNNNN.txt
  int level = 0;
  int valid = 0;      // 0 invalid, 1 sampled, 2 valid but outside the level's boxes
  int mask = 0;       // recorded fields: 32 p, 16 T, 8 rho, 4 u, 2 v, 1 w
  int32_t pos[8][3] = {};   // corner positions in the level's index space
  int64_t node[8] = {};     // stored node of each corner (-1: not a fluid node)
  double w[8] = {};
  unsigned slot = 0;        // first of its eight entries in the level's probe list
  std::string path;
};

struct MicLevel {
  int lvl = 0;
  unsigned m = 0;              // probed nodes (eight per sampled microphone)
  uint32_t* d_nodes = nullptr;
  size_t off = 0;              // doubles from the start of a sample
};

struct Ctx {
  uclbs_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // boundary-list kernel, concurrent with the streaming kernel
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_fork = nullptr, ev_join = nullptr;
  std::vector<Level> levels;
  unsigned long long* d_stats = nullptr;
  // inlet / outlet / freestream equilibria (full FP64 populations)
  int n_in = 0, n_out = 0;
  double* d_feq_in = nullptr;
  double* d_feq_out = nullptr;
  double* d_feq_fs = nullptr;
  bool keep_alpha = false;
  int64_t steps_done = 0;
  // D3Q39 gas constants and the reference-state populations of the FP32 storage
  bool gas_set = false;
  int shifted39 = 0;
  double gas39_mach = 0.0;
  uclbs::Gas39 gas39{};
  uclbs::Ref39 ref39{};
  uclbs::Ref39 fs39{};        // freestream equilibria fs_fieq / fs_gieq (full values)
  double* d_ref39 = nullptr;  // f_ref[39] then g_ref[39]
  // D3Q39 inlets / outlets with stored states: types and equilibrium rows
  // (inlets, then outlets; f then g, 78 doubles a row)
  std::vector<int> in_types39, out_types39;
  double* d_bc39 = nullptr;
  int64_t err_seen = 0;       // D3Q39 error-path counts already reported
  // learned alpha: FP32 master parameters, Adam moments, bf16 inference copy
  float* mlp_p = nullptr;
  float* mlp_g = nullptr;
  float* mlp_m = nullptr;
  float* mlp_v = nullptr;
  short* mlp_wbf = nullptr;
  float mlp_ymean = 0.0f, mlp_ystd = 1.0f;
  bool mlp_ready = false;
  int mlp_t = 0;
  uclbs::MlpSamples samples;
  bool samples_normalised = false;
  uclbs::MlpWork work{};
  int work_batch = 0;
  float* work_buf = nullptr;
  float* batch_x = nullptr;
  unsigned* batch_idx = nullptr;
  double* d_scalar = nullptr;  // loss, mean, std
  // this device's staging area of uclbs_mlp_forward_noalloc
  float* scratch = nullptr;
  // captured multi-level advance (uclbs_advance with use_graph)
  hipGraphExec_t graph = nullptr;
  std::vector<int> graph_sig;
  int64_t structure = 0;  // bumped by every change of levels or links
  // physical placement of the level-0 domain (amrex_problo, amrex_probhi);
  // unset: problo 0 and unit spacing at level 0
  bool placed = false;
  double problo[3] = {0.0, 0.0, 0.0}, probhi[3] = {0.0, 0.0, 0.0};
  // Plot3D writer: flow variables of one level, the packed blocks on the
  // device and in pinned host memory, and the thread that writes the files
  double* out_fields = nullptr;
  size_t out_fields_cap = 0;  // doubles
  char* out_dev = nullptr;
  char* out_host = nullptr;
  size_t out_cap = 0;  // bytes of out_dev and of out_host
  hipEvent_t out_ev = nullptr;
  std::thread out_thread;
  int out_rc = 0;
  std::string out_msg;
  // microphones: set-up, probe lists per level, the device ring of samples
  std::vector<Mic> mics;
  std::vector<MicLevel> mic_levels;
  double* mic_ring = nullptr;
  size_t mic_stride = 0;  // doubles per sample
  int mic_cap = 0, mic_count = 0;
  int64_t mic_structure = -1;
  int64_t regrid_copied = 0;  // nodes the last uclbs_regrid_level took from the old level
};

void release_output(Ctx* c);  // joins the writer thread, frees the output buffers

int hip_rc(hipError_t e) {
  return e == hipSuccess ? UCLBS_OK : set_err(UCLBS_E_HIP, "HIP error: %s", hipGetErrorString(e));
}

// A change of what the captured kernels read (levels, link tables, faces,
// boundary tables) drops the captured advance; the next uclbs_advance with
// use_graph captures it again.
void invalidate_graph(Ctx* c) {
  c->structure++;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  c->graph_sig.clear();
}

size_t elem_size(const Ctx* c) {
  return c->cfg.precision == UCLBS_PREC_FP64 ? sizeof(double) : sizeof(float);
}

int q_of(const Ctx* c) { return c->cfg.lattice; }

int check_level(Ctx* c, int lvl, Level** out) {
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (lvl < 0 || lvl >= (int)c->levels.size() || !c->levels[lvl].defined)
    return set_err(UCLBS_E_STATE, "level %d is not defined", lvl);
  *out = &c->levels[lvl];
  return UCLBS_OK;
}

int set_device(Ctx* c) {
  HIP_TRY(hipSetDevice(c->device));
  return UCLBS_OK;
}

void free_level(Level& L) {
  for (auto& p : L.f)
    if (p) { (void)hipFree(p); p = nullptr; }
  if (L.alpha) { (void)hipFree(L.alpha); L.alpha = nullptr; }
  if (L.defer_count) { (void)hipFree(L.defer_count); L.defer_count = nullptr; }
  if (L.defer_list) { (void)hipFree(L.defer_list); L.defer_list = nullptr; }
  if (L.state) { (void)hipFree(L.state); L.state = nullptr; }
  if (L.flags) { (void)hipFree(L.flags); L.flags = nullptr; }
  if (L.blist) { (void)hipFree(L.blist); L.blist = nullptr; }
  if (L.bbox_of) { (void)hipFree(L.bbox_of); L.bbox_of = nullptr; }
  for (auto& p : L.g)
    if (p) { (void)hipFree(p); p = nullptr; }
  if (L.lm) { (void)hipFree(L.lm); L.lm = nullptr; }
  if (L.neg_count) { (void)hipFree(L.neg_count); L.neg_count = nullptr; }
  if (L.neg_list) { (void)hipFree(L.neg_list); L.neg_list = nullptr; }
  if (L.near_list) { (void)hipFree(L.near_list); L.near_list = nullptr; }
  L.n_near = 0;
  L.near_dirty = true;
  L.table39 = false;
  for (void* p : {(void*)L.cls39, (void*)L.wall_list, (void*)L.wall_f, (void*)L.wall_g})
    if (p) (void)hipFree(p);
  L.cls39 = nullptr;
  L.wall_list = nullptr;
  L.wall_f = L.wall_g = nullptr;
  L.n_wall = 0;
  if (L.rep) { (void)hipFree(L.rep); L.rep = nullptr; }
  if (L.ghosts) { (void)hipFree(L.ghosts); L.ghosts = nullptr; }
  L.nghost = 0;
  for (void* p : {(void*)L.d_boxes, (void*)L.d_tiles, (void*)L.d_copies, (void*)L.d_c2f,
                  (void*)L.d_inject, (void*)L.d_alias})
    if (p) (void)hipFree(p);
  L.d_alias = nullptr;
  L.d_boxes = nullptr;
  L.d_tiles = nullptr;
  L.d_copies = nullptr;
  L.d_c2f = nullptr;
  L.d_inject = nullptr;
  L.ntiles = L.ncopies = L.nc2f = L.ninject = 0;
  L.multibox = false;
  L.blo.clear();
  L.bhi.clear();
  L.hbox.clear();
  L.refined = false;
  L.step = 0;
  L.has_links = false;
  L.nb = L.nsolid = 0;
  L.defined = false;
}

uclbs::StepArgs step_args(const Ctx* c, int lvl = 0) {
  uclbs::StepArgs a{};
  // acoustic scaling (dx, dt halved per level): the lattice viscosity doubles
  // per level, as beta_1_calculator's 2^(lvl+1) mu does (amr_processes.f95:1763)
  a.beta = 1.0 / (6.0 * c->cfg.nu * std::ldexp(1.0, lvl) + 1.0);
  a.tol = c->cfg.alpha_tol;
  a.maxit = c->cfg.alpha_maxit;
  a.alpha_mode = c->cfg.alpha_mode;
  a.shifted = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  // FP64 storage: series exact to round-off.  FP32 storage: the truncation of
  // the series tiers adds at most tol / 2 to the Newton residual bound tol
  a.thr = uclbs::make_series_thr(a.shifted ? 0.5 * a.tol : 1e-16);
  return a;
}

uclbs::BcArgs bc_args(const Ctx* c, const Level* L) {
  uclbs::BcArgs bc{};
  bc.state = L->state;
  bc.flags = L->flags;
  bc.list = L->blist;
  bc.nlist = L->nb;
  bc.n_in = c->n_in;
  bc.n_out = c->n_out;
  bc.feq_in = c->d_feq_in;
  bc.feq_out = c->d_feq_out;
  bc.feq_fs = c->d_feq_fs;
  bc.boxes = L->multibox ? L->d_boxes : nullptr;
  bc.box_of = L->bbox_of;
  return bc;
}

// product-form D3Q27 equilibrium on the host (FP64), for the boundary tables
void feq27_host(double rho, const double u[3], double* feq) {
  double p[3][3];
  for (int d = 0; d < 3; ++d) {
    const double s = std::sqrt(1.0 + 3.0 * u[d] * u[d]);
    const double bb = 2.0 - s;
    p[d][1] = bb;
    p[d][2] = bb * (2.0 * u[d] + s) / (1.0 - u[d]);
    p[d][0] = bb * (1.0 - u[d]) / (2.0 * u[d] + s);
  }
  for (int i = 0; i < 27; ++i)
    feq[i] = rho * uclbs::W27[i] * p[0][uclbs::CX27[i] + 1] * p[1][uclbs::CY27[i] + 1] *
             p[2][uclbs::CZ27[i] + 1];
}

size_t level_bytes(const Ctx* c, const Level* L) {
  return (size_t)L->box.n * q_of(c) * elem_size(c);
}

// the other population buffer mirrors the current one after a state change, so
// nodes no kernel writes (solid, null) hold valid values in both
int mirror_buffers(Ctx* c, Level* L) {
  HIP_TRY(hipMemcpyAsync(L->f[1 - L->cur], L->f[L->cur], level_bytes(c, L),
                         hipMemcpyDeviceToDevice, c->stream));
  return UCLBS_OK;
}

// far-from-equilibrium queue: kFarShards shards of the streaming kernel's
// blocks (ceil(blocks / kFarShards) blocks of 256 nodes each) and one shard of
// nb entries for the boundary kernel
hipError_t alloc_far_queue(Level& L, int64_t nb) {
  const int64_t blocks = (L.box.n + 255) / 256;
  const int64_t cap = (blocks + uclbs::kFarShards - 1) / uclbs::kFarShards * 256;
  if (L.defer_list) { (void)hipFree(L.defer_list); L.defer_list = nullptr; }
  L.defer_cap = (unsigned)cap;
  return hipMalloc(&L.defer_list, (size_t)(cap * uclbs::kFarShards + nb + 1) * sizeof(uint2));
}

// flags + ascending boundary list from the link table
int finalize_links(Ctx* c, Level* L) {
  invalidate_graph(c);
  const int64_t n = L->box.n;
  if (!L->flags) HIP_TRY(hipMalloc(&L->flags, (size_t)n));
  HIP_TRY(uclbs::launch_flags27(L->state, L->box, L->face_code, L->flags, c->stream));
  std::vector<uint8_t> hf((size_t)n);
  HIP_TRY(hipMemcpyAsync(hf.data(), L->flags, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (L->refined) {
    // a refined level's ghost shell is filled from the coarser level, never stepped
    const int G = L->ghost_depth;
    const int nx = L->box.nx, ny = L->box.ny, nz = L->box.nz;
    for (int64_t k = 0; k < n; ++k) {
      const int x = (int)(k % nx), y = (int)((k / nx) % ny), z = (int)(k / ((int64_t)nx * ny));
      if (x < G || y < G || z < G || x >= nx - G || y >= ny - G || z >= nz - G)
        hf[(size_t)k] = uclbs::kFlagGhost;
    }
    HIP_TRY(hipMemcpy(L->flags, hf.data(), (size_t)n, hipMemcpyHostToDevice));
  }
  std::vector<uint32_t> list;
  int64_t nsolid = 0;
  for (int64_t k = 0; k < n; ++k) {
    if (hf[k] == uclbs::kFlagBoundary) list.push_back((uint32_t)k);
    nsolid += hf[k] == uclbs::kFlagSolid;
  }
  if (L->blist) { (void)hipFree(L->blist); L->blist = nullptr; }
  L->nb = (int64_t)list.size();
  L->nsolid = nsolid;
  HIP_TRY(alloc_far_queue(*L, L->nb));
  if (L->nb) {
    HIP_TRY(hipMalloc(&L->blist, list.size() * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(L->blist, list.data(), list.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
  }
  L->has_links = true;
  L->relink = false;
  return UCLBS_OK;
}

// ev: optional [2*nsteps] events recorded around each step's streaming kernel
// every face code must have its table before a kernel reads it
int check_face_tables(const Ctx* c, const Level* L) {
  for (int f = 0; f < 6; ++f) {
    const int32_t v = L->face_code[f];
    if (v <= -101 && v >= -199 && (-v - 100) > c->n_in)
      return set_err(UCLBS_E_STATE, "face %d uses inlet %d but %d inlet states are set", f,
                     -v - 100, c->n_in);
    if (v <= -201 && v >= -299 && (-v - 200) > c->n_out)
      return set_err(UCLBS_E_STATE, "face %d uses outlet %d but %d outlet states are set", f,
                     -v - 200, c->n_out);
    if (v == -1000 && !c->d_feq_fs)
      return set_err(UCLBS_E_STATE, "face %d is a freestream face but no freestream state is set", f);
  }
  return UCLBS_OK;
}

// nodes of a D3Q39 box with faces whose pulls may cross a face: stepped by
// their own launch; the list is rebuilt when the faces change
static int build_near39(Ctx* c, Level* L) {
  if (!L->near_dirty) return UCLBS_OK;
  if (L->table39) {
    // node classes from the table: solid, wall links (i >= 1 coded <= -1001),
    // near a face, plain; the near-face and wall launches take their lists
    const int64_t n = L->box.n;
    std::vector<int32_t> st((size_t)n * uclbs::Q39);
    HIP_TRY(hipMemcpy(st.data(), L->state, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<uint8_t> cls((size_t)n);
    std::vector<unsigned> near, wall;
    const uclbs::Box& bx = L->box;
    for (int64_t k = 0; k < n; ++k) {
      uint8_t cl = 0;
      if (st[(size_t)k] < 0) {
        cl = 3;
      } else {
        bool w = false;
        for (int i = 1; i < uclbs::Q39 && !w; ++i) w = st[(size_t)i * n + k] <= -1001;
        const int x = (int)(k % bx.nx), y = (int)((k / bx.nx) % bx.ny), z = (int)(k / ((int64_t)bx.nx * bx.ny));
        if (w) cl = 2;
        else if (uclbs::near_face39(x, y, z, L->face_code, bx.nx, bx.ny, bx.nz, c->shifted39 != 0))
          cl = 1;
      }
      cls[(size_t)k] = cl;
      if (cl == 1) near.push_back((unsigned)k);
      if (cl == 2) wall.push_back((unsigned)k);
    }
    for (void* p : {(void*)L->cls39, (void*)L->wall_list, (void*)L->wall_f, (void*)L->wall_g,
                    (void*)L->near_list})
      if (p) HIP_TRY(hipFree(p));
    L->cls39 = nullptr;
    L->wall_list = L->near_list = nullptr;
    L->wall_f = L->wall_g = nullptr;
    HIP_TRY(hipMalloc(&L->cls39, (size_t)n));
    HIP_TRY(hipMemcpy(L->cls39, cls.data(), (size_t)n, hipMemcpyHostToDevice));
    L->n_near = (unsigned)near.size();
    L->n_wall = (unsigned)wall.size();
    if (!near.empty()) {
      HIP_TRY(hipMalloc(&L->near_list, near.size() * sizeof(unsigned)));
      HIP_TRY(hipMemcpy(L->near_list, near.data(), near.size() * sizeof(unsigned), hipMemcpyHostToDevice));
    }
    if (!wall.empty()) {
      HIP_TRY(hipMalloc(&L->wall_list, wall.size() * sizeof(unsigned)));
      HIP_TRY(hipMemcpy(L->wall_list, wall.data(), wall.size() * sizeof(unsigned), hipMemcpyHostToDevice));
      HIP_TRY(hipMalloc(&L->wall_f, wall.size() * uclbs::Q39 * sizeof(double)));
      HIP_TRY(hipMalloc(&L->wall_g, wall.size() * uclbs::Q39 * sizeof(double)));
    }
    L->near_dirty = false;
    invalidate_graph(c);
    return UCLBS_OK;
  }
  std::vector<unsigned> near;
  const uclbs::Box& bx = L->box;
  for (int z = 0; z < bx.nz; ++z)
    for (int y = 0; y < bx.ny; ++y)
      for (int x = 0; x < bx.nx; ++x)
        if (uclbs::near_face39(x, y, z, L->face_code, bx.nx, bx.ny, bx.nz, c->shifted39 != 0))
          near.push_back((unsigned)(x + (int64_t)bx.nx * (y + (int64_t)bx.ny * z)));
  if (L->near_list) { HIP_TRY(hipFree(L->near_list)); L->near_list = nullptr; }
  L->n_near = (unsigned)near.size();
  if (!near.empty()) {
    HIP_TRY(hipMalloc(&L->near_list, near.size() * sizeof(unsigned)));
    HIP_TRY(hipMemcpy(L->near_list, near.data(), near.size() * sizeof(unsigned),
                      hipMemcpyHostToDevice));
  }
  L->near_dirty = false;
  return UCLBS_OK;
}

// the finest level of a hierarchy: level 0 and the refined levels nested under it
int finest_level(const Ctx* c) {
  int n = 1;
  while (n < (int)c->levels.size() && c->levels[n].defined && c->levels[n].refined) ++n;
  return n - 1;
}

// rows of the inlet / outlet table the faces of L read (inlet_bc.f95:31-53,
// outlet_bc.f95:33-58): type 3 the freestream pair itself (freestream_bc.f95,
// -2); an inlet its own row; an outlet of type 6 none (pure outflow, -1); an
// outlet of type 2 the row of the inlet of the same number (press_temp_bc.f95
// reads inl_fieq for outlets too); other outlets their own row
int resolve_bc39(Ctx* c, Level* L, uclbs::Faces39* fc) {
  const int n_in = (int)c->in_types39.size(), n_out = (int)c->out_types39.size();
  fc->bc = c->d_bc39;
  for (int f = 0; f < 6; ++f) {
    const int v = L->face_code[f];
    fc->tab[f] = -1;
    if (v <= -101 && v >= -199) {
      const int j = -v - 100;
      if (j > n_in) return set_err(UCLBS_E_STATE, "face %d: inlet %d has no state (uclbs_set_bc_states39)", f, j);
      fc->tab[f] = c->in_types39[j - 1] == 3 ? -2 : j - 1;
    } else if (v <= -201 && v >= -299) {
      const int j = -v - 200;
      if (j > n_out) return set_err(UCLBS_E_STATE, "face %d: outlet %d has no state (uclbs_set_bc_states39)", f, j);
      const int t = c->out_types39[j - 1];
      if (t == 6) {
        fc->tab[f] = -1;
      } else if (t == 3) {
        fc->tab[f] = -2;
      } else if (t == 2) {
        if (j > n_in)
          return set_err(UCLBS_E_STATE, "face %d: outlet %d of type 2 reads inlet %d's state, which is not set", f, j, j);
        fc->tab[f] = j - 1;
      } else {
        fc->tab[f] = n_in + j - 1;
      }
    }
  }
  return UCLBS_OK;
}

int launch_steps39(Ctx* c, Level* L, int nsteps, hipEvent_t* ev) {
  uclbs::Step39 a{};
  a.gas = L->gas39;
  a.ghost = L->refined ? L->ghost_depth : 0;
  a.ref = c->ref39;
  a.fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  a.shifted = c->shifted39;
  a.stats = c->d_stats;
  a.neg_count = L->neg_count;
  a.neg_list = L->neg_list;
  a.rep = L->rep;
  a.lm = L->lm;
  a.alpha = c->keep_alpha ? L->alpha : nullptr;
  a.has_faces = 0;
  for (int f = 0; f < 6; ++f) {
    a.faces.code[f] = L->face_code[f];
    a.has_faces |= !L->refined && L->face_code[f] != uclbs::kFacePeriodic;
  }
  for (int i = 0; i < uclbs::Q39; ++i) {
    a.faces.f[i] = c->fs39.f[i];
    a.faces.g[i] = c->fs39.g[i];
  }
  {
    int rc = resolve_bc39(c, L, &a.faces);
    if (rc) return rc;
  }
  if (a.has_faces || L->table39) {
    int rc = build_near39(c, L);
    if (rc) return rc;
  }
  a.near_list = L->near_list;
  a.n_near = L->n_near;
  if (L->table39) {
    a.has_faces = 1;
    a.walls.cls = L->cls39;
    a.walls.state = L->state;
    a.walls.list = L->wall_list;
    a.walls.n = L->n_wall;
    a.walls.wf = L->wall_f;
    a.walls.wg = L->wall_g;
    a.walls.mach = c->gas39_mach;
  }
  for (int s = 0; s < nsteps; ++s) {
    a.fin = L->f[L->cur];
    a.gin = L->g[L->cur];
    a.fout = L->f[1 - L->cur];
    a.gout = L->g[1 - L->cur];
    // every level's start-up ramps read the finest level's counter, the
    // reference's timestep(amr_max_lvl) (calculate_alpha.f95:185-187,
    // collision.f95:277-278, find_q_bar.f95:36, find_little_q_bar.f95:26)
    a.step = (long long)c->levels[finest_level(c)].step;
    a.walls.lvl_step = (long long)L->step;
    if (ev) HIP_TRY(hipEventRecord(ev[2 * s], c->stream));
    HIP_TRY(uclbs::launch_d3q39_step(L->box, a, c->stream));
    if (ev) HIP_TRY(hipEventRecord(ev[2 * s + 1], c->stream));
    L->cur = 1 - L->cur;
    L->step++;
    c->steps_done++;
  }
  return UCLBS_OK;
}

// A level of many boxes: same-level ghost copies, the batched streaming kernel
// over all boxes, the far-from-equilibrium kernel.  Ghosts that come from the
// coarser level are filled by uclbs_interlevel before the step.
int launch_steps_boxes(Ctx* c, Level* L, int nsteps, hipEvent_t* ev) {
  if (c->cfg.alpha_mode == UCLBS_ALPHA_LEARNED)
    return set_err(UCLBS_E_UNSUPPORTED, "learned alpha runs on single-box levels");
  uclbs::StepArgs a = step_args(c, (int)(L - c->levels.data()));
  a.defer_list = L->defer_list;
  a.defer_cap = L->defer_cap;
  const unsigned nn = (unsigned)L->box.n;
  for (int s = 0; s < nsteps; ++s) {
    a.defer_count = L->defer_count + L->far_parity * (uclbs::kFarShards + 1);
    a.defer_next = L->defer_count + (1 - L->far_parity) * (uclbs::kFarShards + 1);
    HIP_TRY(uclbs::launch_copy_pairs27(L->f[L->cur], nn, L->f[L->cur], nn, L->d_copies, L->ncopies,
                                       a.shifted, c->stream));
    if (L->state && L->nb) {
      // boundary-list nodes on the side stream, concurrent with the box kernel
      HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
      HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
      HIP_TRY(uclbs::launch_d3q27_boundary(L->box, L->f[L->cur], L->f[1 - L->cur], L->alpha,
                                           c->keep_alpha, c->d_stats, a, bc_args(c, L), c->side));
      HIP_TRY(hipEventRecord(c->ev_join, c->side));
    }
    if (ev) HIP_TRY(hipEventRecord(ev[2 * s], c->stream));
    HIP_TRY(uclbs::launch_d3q27_step_boxes(nn, L->f[L->cur], L->f[1 - L->cur], L->alpha,
                                           c->keep_alpha, c->d_stats, a, L->flags, L->d_alias,
                                           L->d_tiles, L->ntiles, L->d_boxes, L->tile_nodes,
                                           L->state != nullptr, c->stream));
    if (ev) HIP_TRY(hipEventRecord(ev[2 * s + 1], c->stream));
    if (L->state && L->nb) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (a.shifted) {
      HIP_TRY(uclbs::launch_d3q27_far(L->box, L->f[1 - L->cur], L->alpha, c->keep_alpha,
                                      c->d_stats, a, c->stream));
      L->far_parity ^= 1;
    }
    L->cur = 1 - L->cur;
    c->steps_done++;
  }
  return UCLBS_OK;
}

int launch_steps(Ctx* c, Level* L, int nsteps, hipEvent_t* ev = nullptr) {
  if (L->stale)
    return set_err(UCLBS_E_STATE, "level %d: its coarser level was regridded (regrid or define it "
                                  "again)", (int)(L - c->levels.data()));
  if (L->relink)
    return set_err(UCLBS_E_STATE, "level %d was regridded from a level with a link table: build "
                                  "or set its links", (int)(L - c->levels.data()));
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) return launch_steps39(c, L, nsteps, ev);
  if (L->multibox) return launch_steps_boxes(c, L, nsteps, ev);
  uclbs::StepArgs a = step_args(c, (int)(L - c->levels.data()));
  if (L->has_links) {
    int rc = check_face_tables(c, L);
    if (rc) return rc;
  }
  a.defer_list = L->defer_list;
  a.defer_cap = L->defer_cap;
  for (int f = 0; f < 6; ++f) a.face_code[f] = L->face_code[f];
  const bool learned = c->cfg.alpha_mode == UCLBS_ALPHA_LEARNED;
  if (learned && !c->mlp_ready)
    return set_err(UCLBS_E_STATE, "learned alpha: no trained network (uclbs_mlp_train or "
                                  "uclbs_mlp_set_params)");
  if (learned && c->cfg.precision != UCLBS_PREC_FP32_SHIFTED)
    return set_err(UCLBS_E_UNSUPPORTED, "learned alpha runs on FP32-shifted storage");
  for (int s = 0; s < nsteps; ++s) {
    if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
      return set_err(UCLBS_E_UNSUPPORTED, "lattice %d not supported by step", c->cfg.lattice);
    if (learned)
      HIP_TRY(uclbs::launch_alpha_guess27(L->box, static_cast<const float*>(L->f[L->cur]),
                                          L->has_links ? L->flags : nullptr, c->mlp_wbf, c->mlp_p,
                                          c->mlp_ymean, c->mlp_ystd, L->alpha, c->stream));
    if (!L->has_links && !(L->box.per[0] && L->box.per[1] && L->box.per[2]))
      return set_err(UCLBS_E_UNSUPPORTED,
                     "level without link tables must be periodic on all axes");
    // learned mode: the streaming kernel reads its start from alpha[]
    const int keep = c->keep_alpha || learned;
    a.defer_count = L->defer_count + L->far_parity * (uclbs::kFarShards + 1);
    a.defer_next = L->defer_count + (1 - L->far_parity) * (uclbs::kFarShards + 1);
    if (L->has_links && L->nb) {
      // boundary-list nodes on the side stream, concurrent with the streaming
      // kernel: neighbour velocities, then the boundary collide
      HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
      HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
      HIP_TRY(uclbs::launch_d3q27_boundary(L->box, L->f[L->cur], L->f[1 - L->cur], L->alpha,
                                           keep, c->d_stats, a, bc_args(c, L), c->side));
      HIP_TRY(hipEventRecord(c->ev_join, c->side));
    }
    if (ev) HIP_TRY(hipEventRecord(ev[2 * s], c->stream));
    if (L->has_links)
      HIP_TRY(uclbs::launch_d3q27_step_links(L->box, L->f[L->cur], L->f[1 - L->cur], L->alpha,
                                             keep, c->d_stats, a, bc_args(c, L),
                                             c->stream));
    else
      HIP_TRY(uclbs::launch_d3q27_step_periodic(L->box, L->f[L->cur], L->f[1 - L->cur], L->alpha,
                                                keep, c->d_stats, a, c->stream));
    if (ev) HIP_TRY(hipEventRecord(ev[2 * s + 1], c->stream));
    // the far kernel takes the queue of both kernels
    if (L->has_links && L->nb) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (a.shifted) {
      HIP_TRY(uclbs::launch_d3q27_far(L->box, L->f[1 - L->cur], L->alpha, c->keep_alpha,
                                      c->d_stats, a, c->stream));
      L->far_parity ^= 1;
    }
    L->cur = 1 - L->cur;
    c->steps_done++;
  }
  return UCLBS_OK;
}

}  // namespace

extern "C" {

const char* uclbs_last_error(void) { return g_err.c_str(); }

int uclbs_abi_version(void) { return UCLBS_ABI_VERSION; }

int uclbs_create(const uclbs_config* cfg, void** out) {
  if (!cfg || !out) return set_err(UCLBS_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->lattice != UCLBS_LATTICE_D3Q27 && cfg->lattice != UCLBS_LATTICE_D3Q39)
    return set_err(UCLBS_E_ARG, "lattice must be 27 or 39, got %d", cfg->lattice);
  if (cfg->precision != UCLBS_PREC_FP32_SHIFTED && cfg->precision != UCLBS_PREC_FP64)
    return set_err(UCLBS_E_ARG, "bad precision %d", cfg->precision);
  if (cfg->alpha_mode < 0 || cfg->alpha_mode > 2)
    return set_err(UCLBS_E_ARG, "bad alpha_mode %d", cfg->alpha_mode);
  if (!(cfg->nu > 0.0)) return set_err(UCLBS_E_ARG, "nu must be > 0");
  Ctx* c = new Ctx();
  c->cfg = *cfg;
  if (c->cfg.alpha_maxit <= 0) c->cfg.alpha_maxit = 30;
  if (c->cfg.alpha_maxit > 63) c->cfg.alpha_maxit = 63;
  if (!(c->cfg.alpha_tol > 0.0)) c->cfg.alpha_tol = 1e-10;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipMalloc(&c->d_stats, sizeof(unsigned long long) * uclbs::kStatSlots * uclbs::kStatFields);
  if (e == hipSuccess)
    e = hipMemset(c->d_stats, 0,
                  sizeof(unsigned long long) * uclbs::kStatSlots * uclbs::kStatFields);
  if (e == hipSuccess) e = scratch_address(c->device, &c->scratch);
  if (e != hipSuccess) {
    int rc = set_err(UCLBS_E_HIP, "context setup failed: %s", hipGetErrorString(e));
    uclbs_destroy(c);
    return rc;
  }
  *out = c;
  return UCLBS_OK;
}

int uclbs_destroy(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return UCLBS_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release_output(c);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  for (auto& L : c->levels) free_level(L);
  if (c->d_stats) (void)hipFree(c->d_stats);
  if (c->d_feq_in) (void)hipFree(c->d_feq_in);
  if (c->d_feq_out) (void)hipFree(c->d_feq_out);
  if (c->d_feq_fs) (void)hipFree(c->d_feq_fs);
  if (c->d_ref39) (void)hipFree(c->d_ref39);
  if (c->d_bc39) (void)hipFree(c->d_bc39);
  for (void* p : {(void*)c->mlp_p, (void*)c->mlp_g, (void*)c->mlp_m, (void*)c->mlp_v,
                  (void*)c->mlp_wbf, (void*)c->samples.X, (void*)c->samples.y,
                  (void*)c->samples.node_of, (void*)c->samples.m_of, (void*)c->samples.count,
                  (void*)c->samples.ballots, (void*)c->samples.blk, (void*)c->work_buf, (void*)c->batch_x, (void*)c->batch_idx, (void*)c->d_scalar})
    if (p) (void)hipFree(p);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->side) { (void)hipStreamSynchronize(c->side); (void)hipStreamDestroy(c->side); }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return UCLBS_OK;
}

int uclbs_define_level(void* ctx, int lvl, const int32_t lo[3], const int32_t hi[3],
                       const int32_t periodic[3]) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !lo || !hi || !periodic) return set_err(UCLBS_E_ARG, "null argument");
  if (lvl < 0 || lvl > 31) return set_err(UCLBS_E_ARG, "level %d out of range", lvl);
  int rc = set_device(c);
  if (rc) return rc;
  if ((int)c->levels.size() <= lvl) c->levels.resize(lvl + 1);
  Level& L = c->levels[lvl];
  free_level(L);
  invalidate_graph(c);
  // ghost layer of a refined level: one node on D3Q27, four on D3Q39 (links
  // reach three nodes, four along x on the shifted lattice)
  const int G = lvl > 0 ? (c->cfg.lattice == UCLBS_LATTICE_D3Q39 ? 4 : 1) : 0;
  if (lvl > 0) {
    // a refined level: nested in level lvl-1 with ratio 2, not periodic, and
    // the coarse nodes its ghost shell interpolates from inside the coarser
    // level's valid nodes (no wrap across a periodic coarse face)
    if (lvl - 1 >= (int)c->levels.size() || !c->levels[lvl - 1].defined)
      return set_err(UCLBS_E_STATE, "define level %d before level %d", lvl - 1, lvl);
    const Level& P = c->levels[lvl - 1];
    if (P.multibox)
      return set_err(UCLBS_E_UNSUPPORTED, "a single-box level does not nest in a level of many "
                                          "boxes (use uclbs_define_level_boxes)");
    for (int d = 0; d < 3; ++d) {
      if (periodic[d]) return set_err(UCLBS_E_ARG, "a refined level is not periodic");
      const int cv_lo = P.lo[d], cv_hi = P.hi[d];
      const int g_lo = (lo[d] - G) >> 1, g_hi = (hi[d] + G + 1) >> 1;
      if (g_lo < cv_lo || g_hi > cv_hi)
        return set_err(UCLBS_E_ARG, "level %d box [%d,%d] on axis %d needs coarse nodes "
                                    "[%d,%d] outside level %d's [%d,%d]",
                       lvl, lo[d], hi[d], d, g_lo, g_hi, lvl - 1, cv_lo, cv_hi);
    }
  }
  int n[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = hi[d] - lo[d] + 1 + 2 * G;
    if (n[d] < 3) return set_err(UCLBS_E_ARG, "box extent %d on axis %d is below 3", n[d], d);
    L.origin[d] = lo[d] - G;
    L.lo[d] = lo[d];
    L.hi[d] = hi[d];
    L.box.per[d] = periodic[d] != 0;
    // faces default to periodic or to a bounce-back wall (-10)
    L.face_code[d] = L.face_code[d + 3] = L.box.per[d] ? uclbs::kFacePeriodic : -10;
  }
  L.box.nx = n[0];
  L.box.ny = n[1];
  L.box.nz = n[2];
  L.box.n = (int64_t)n[0] * n[1] * n[2];
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) {
    if (!c->gas_set) return set_err(UCLBS_E_STATE, "D3Q39: call uclbs_set_gas before defining levels");
    for (int d = 0; d < 3; ++d)
      if (n[d] < 5) return set_err(UCLBS_E_ARG, "D3Q39 box extent %d on axis %d is below 5", n[d], d);
  }
  const size_t bytes = (size_t)L.box.n * q_of(c) * elem_size(c);
  for (auto& p : L.f) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      free_level(L);
      return set_err(UCLBS_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
  }
  HIP_TRY(hipMalloc(&L.alpha, (size_t)L.box.n * sizeof(double)));
  if (L.box.n > (1ll << 30)) return set_err(UCLBS_E_ARG, "box exceeds 2^30 nodes");
  HIP_TRY(hipMalloc(&L.defer_count, 2 * (uclbs::kFarShards + 1) * sizeof(unsigned)));
  HIP_TRY(hipMemset(L.defer_count, 0, 2 * (uclbs::kFarShards + 1) * sizeof(unsigned)));
  L.far_parity = 0;
  HIP_TRY(alloc_far_queue(L, 0));
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) {
    for (auto& p : L.g) HIP_TRY(hipMalloc(&p, bytes));
    HIP_TRY(hipMalloc(&L.lm, (size_t)L.box.n * uclbs::kM39 * sizeof(double)));
    HIP_TRY(hipMalloc(&L.neg_count, sizeof(unsigned)));
    HIP_TRY(hipMemset(L.neg_count, 0, sizeof(unsigned)));
    HIP_TRY(hipMalloc(&L.neg_list, (size_t)L.box.n * sizeof(unsigned)));
    HIP_TRY(hipMalloc(&L.rep, (size_t)L.box.n * uclbs::Q39 * sizeof(double)));
  }
  L.step = 0;
  L.cur = 0;
  L.defined = true;
  L.ghost_depth = G;
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) {
    // timing_calcs.f95:24-26 on level lvl: dt and epsilon halve per level
    L.gas39 = c->gas39;
    L.gas39.lvl = lvl;
    L.gas39.dt = c->gas39.dt / std::ldexp(1.0, lvl);
    L.gas39.eps = L.gas39.dt / 1.0;
  }
  if (lvl > 0) {
    // flags (ghost layer / interior) and the ghost list
    L.refined = true;
    std::vector<uint8_t> hf((size_t)L.box.n, uclbs::kFlagInterior);
    std::vector<uint32_t> gl;
    for (int cz = 0; cz < n[2]; ++cz)
      for (int cy = 0; cy < n[1]; ++cy)
        for (int cx = 0; cx < n[0]; ++cx)
          if (cx < G || cy < G || cz < G || cx >= n[0] - G || cy >= n[1] - G || cz >= n[2] - G) {
            const size_t k = (size_t)cx + (size_t)n[0] * ((size_t)cy + (size_t)n[1] * cz);
            hf[k] = uclbs::kFlagGhost;
            gl.push_back((uint32_t)k);
          }
    HIP_TRY(hipMalloc(&L.flags, (size_t)L.box.n));
    HIP_TRY(hipMemcpy(L.flags, hf.data(), hf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&L.ghosts, gl.size() * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(L.ghosts, gl.data(), gl.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    L.nghost = (unsigned)gl.size();
    L.has_links = true;  // the streaming kernel skips the ghost shell by its flag
    L.nb = 0;
    for (int f = 0; f < 6; ++f) L.face_code[f] = -10;
  }
  return UCLBS_OK;
}


}  // extern "C"

// ---- levels of many boxes: host tables -------------------------------------------

// buckets of the valid regions (host lookup of the box that holds a position)
static void build_buckets(Level& L) {
  const int nb = (int)L.hbox.size();
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = INT32_MAX;
    hi[d] = INT32_MIN;
    for (int b = 0; b < nb; ++b) {
      lo[d] = std::min(lo[d], L.blo[3 * b + d]);
      hi[d] = std::max(hi[d], L.bhi[3 * b + d]);
    }
  }
  int ext = 1;
  for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d] + 1);
  L.bk_e = std::max(8, (ext + 63) / 64);
  size_t total = 1;
  for (int d = 0; d < 3; ++d) {
    L.bk_lo[d] = lo[d];
    L.bk_n[d] = (hi[d] - lo[d]) / L.bk_e + 1;
    total *= (size_t)L.bk_n[d];
  }
  L.buckets.assign(total, {});
  for (int b = 0; b < nb; ++b) {
    int c0[3], c1[3];
    for (int d = 0; d < 3; ++d) {
      c0[d] = (L.blo[3 * b + d] - lo[d]) / L.bk_e;
      c1[d] = (L.bhi[3 * b + d] - lo[d]) / L.bk_e;
    }
    for (int z = c0[2]; z <= c1[2]; ++z)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int x = c0[0]; x <= c1[0]; ++x)
          L.buckets[(size_t)x + (size_t)L.bk_n[0] * ((size_t)y + (size_t)L.bk_n[1] * z)].push_back(b);
  }
}

// stored index of the valid node at level position G (-1 if no valid node of
// the level is there).  Periodic axes wrap: a single box over its extent,
// level 0 of many boxes over its domain.
static int64_t valid_index(const Level& L, const int G[3], int* box_out = nullptr) {
  int g[3] = {G[0], G[1], G[2]};
  if (!L.multibox) {
    int loc[3];
    const int n[3] = {L.box.nx, L.box.ny, L.box.nz};
    for (int d = 0; d < 3; ++d) {
      if (L.box.per[d]) g[d] = ((g[d] - L.origin[d]) % n[d] + n[d]) % n[d] + L.origin[d];
      if (g[d] < L.lo[d] || g[d] > L.hi[d]) return -1;
      loc[d] = g[d] - L.origin[d];
    }
    return (int64_t)loc[0] + (int64_t)n[0] * (loc[1] + (int64_t)n[1] * loc[2]);
  }
  if (L.box.per[0]) {
    for (int d = 0; d < 3; ++d) {
      const int n = L.dom_hi[d] - L.dom_lo[d] + 1;
      g[d] = ((g[d] - L.dom_lo[d]) % n + n) % n + L.dom_lo[d];
    }
  }
  int cell[3];
  for (int d = 0; d < 3; ++d) {
    const int r = g[d] - L.bk_lo[d];
    if (r < 0) return -1;
    cell[d] = r / L.bk_e;
    if (cell[d] >= L.bk_n[d]) return -1;
  }
  for (int b : L.buckets[(size_t)cell[0] + (size_t)L.bk_n[0] * ((size_t)cell[1] + (size_t)L.bk_n[1] * cell[2])]) {
    const int* lo = &L.blo[3 * b];
    const int* hi = &L.bhi[3 * b];
    if (g[0] < lo[0] || g[0] > hi[0] || g[1] < lo[1] || g[1] > hi[1] || g[2] < lo[2] || g[2] > hi[2])
      continue;
    const uclbs::BoxDesc& bx = L.hbox[b];
    if (box_out) *box_out = b;
    return (int64_t)bx.off + (g[0] - lo[0] + 1) +
           (int64_t)bx.nx * ((g[1] - lo[1] + 1) + (int64_t)bx.ny * (g[2] - lo[2] + 1));
  }
  return -1;
}

// A level of many boxes with a link table: flags of the valid nodes, the
// boundary list with each node's box, and the ghost nodes that show a
// boundary or solid node turned from mirrors (which repeat a plain update)
// into copies of that node, made before each step.
static int finalize_links_boxes(Ctx* c, Level* L) {
  invalidate_graph(c);
  const int64_t n = L->box.n;
  HIP_TRY(uclbs::launch_flags_boxes27(L->state, n, L->flags, c->stream));
  std::vector<uint8_t> hf((size_t)n);
  HIP_TRY(hipMemcpyAsync(hf.data(), L->flags, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<uint3> copies;
  std::vector<uint32_t> list, box_of;
  int64_t nsolid = 0;
  for (size_t b = 0; b < L->hbox.size(); ++b) {
    const uclbs::BoxDesc& bx = L->hbox[b];
    const unsigned nz = bx.n / (bx.nx * bx.ny);
    for (unsigned z = 0; z < nz; ++z)
      for (unsigned y = 0; y < bx.ny; ++y)
        for (unsigned x = 0; x < bx.nx; ++x) {
          const unsigned k = bx.off + x + bx.nx * (y + bx.ny * z);
          const bool shell = x == 0 || y == 0 || z == 0 || x + 1 == bx.nx || y + 1 == bx.ny || z + 1 == nz;
          if (!shell) {
            if (hf[k] == uclbs::kFlagBoundary) {
              list.push_back(k);
              box_of.push_back((uint32_t)b);
            }
            nsolid += hf[k] == uclbs::kFlagSolid;
            continue;
          }
          if (hf[k] != uclbs::kFlagMirror) continue;
          const int G[3] = {L->blo[3 * b] - 1 + (int)x, L->blo[3 * b + 1] - 1 + (int)y,
                            L->blo[3 * b + 2] - 1 + (int)z};
          const int64_t src = valid_index(*L, G, nullptr);
          if (src >= 0 && hf[(size_t)src] != uclbs::kFlagInterior) {
            hf[k] = uclbs::kFlagGhost;
            // a solid node's populations are read by no update (a fluid
            // node's link to it is a wall link, rebuilt by the boundary
            // kernel), so its ghosts are not refreshed each step; those of
            // boundary nodes are
            if (hf[(size_t)src] != uclbs::kFlagSolid)
              copies.push_back(make_uint3(k, (unsigned)src, uclbs::kAllPops27));
          }
        }
  }
  HIP_TRY(hipMemcpy(L->flags, hf.data(), (size_t)n, hipMemcpyHostToDevice));
  if (L->d_copies) { HIP_TRY(hipFree(L->d_copies)); L->d_copies = nullptr; }
  L->ncopies = (unsigned)copies.size();
  if (!copies.empty()) {
    HIP_TRY(hipMalloc(&L->d_copies, copies.size() * sizeof(uint3)));
    HIP_TRY(hipMemcpy(L->d_copies, copies.data(), copies.size() * sizeof(uint3), hipMemcpyHostToDevice));
  }
  if (L->blist) { HIP_TRY(hipFree(L->blist)); L->blist = nullptr; }
  if (L->bbox_of) { HIP_TRY(hipFree(L->bbox_of)); L->bbox_of = nullptr; }
  L->nb = (int64_t)list.size();
  L->nsolid = nsolid;
  if (L->nb) {
    HIP_TRY(hipMalloc(&L->blist, list.size() * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(L->blist, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&L->bbox_of, box_of.size() * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(L->bbox_of, box_of.data(), box_of.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  // the far queue's last shard takes the boundary nodes
  if (L->defer_list) { HIP_TRY(hipFree(L->defer_list)); L->defer_list = nullptr; }
  HIP_TRY(hipMalloc(&L->defer_list,
                    ((size_t)L->defer_cap * uclbs::kFarShards + (size_t)L->nb + 1) * sizeof(uint2)));
  L->has_links = true;
  L->relink = false;
  return UCLBS_OK;
}

// coarse stencil of fine position G: per axis the coincident coarse node
// (G even) or the two neighbours with weight 1/2 (G odd), tensor order z, y, x
// -- the order k27_coarse_to_fine sums in
static bool c2f_stencil(const Level& C, const int G[3], uclbs::C2FStencil& st) {
  int c0[3], odd[3];
  for (int d = 0; d < 3; ++d) {
    odd[d] = G[d] & 1;
    c0[d] = G[d] >> 1;
  }
  int q = 0;
  for (int cz = 0; cz < 2; ++cz)
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx) {
        const double wx = odd[0] ? 0.5 : (cx ? 0.0 : 1.0), wy = odd[1] ? 0.5 : (cy ? 0.0 : 1.0),
                     wz = odd[2] ? 0.5 : (cz ? 0.0 : 1.0);
        const double w = wx * wy * wz;
        st.src[q] = 0;
        st.w[q] = 0.0f;
        if (w != 0.0) {
          const int P[3] = {c0[0] + cx * odd[0], c0[1] + cy * odd[1], c0[2] + cz * odd[2]};
          const int64_t k = valid_index(C, P);
          if (k < 0) return false;
          st.src[q] = (unsigned)k;
          st.w[q] = (float)w;
        }
        ++q;
      }
  return true;
}

template <typename T>
static hipError_t upload_vec(T** d, const std::vector<T>& h) {
  if (*d) (void)hipFree(*d);
  *d = nullptr;
  if (h.empty()) return hipSuccess;
  hipError_t e = hipMalloc(d, h.size() * sizeof(T));
  if (e == hipSuccess) e = hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

static hipError_t build_tiles(Level& L, int tile_nodes) {
  std::vector<uclbs::BoxTile> tiles;
  for (unsigned b = 0; b < (unsigned)L.hbox.size(); ++b)
    for (unsigned s = 0; s < L.hbox[b].n; s += (unsigned)tile_nodes) tiles.push_back({b, s});
  L.tile_nodes = tile_nodes;
  L.ntiles = (unsigned)tiles.size();
  return upload_vec(&L.d_tiles, tiles);
}

extern "C" {

int uclbs_define_level_boxes(void* ctx, int lvl, int nbox, const int32_t* lo, const int32_t* hi) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !lo || !hi) return set_err(UCLBS_E_ARG, "null argument");
  if (lvl < 0 || lvl > 31) return set_err(UCLBS_E_ARG, "level %d out of range", lvl);
  if (nbox < 1) return set_err(UCLBS_E_ARG, "no boxes");
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "levels of many boxes are D3Q27 in this library");
  int rc = set_device(c);
  if (rc) return rc;
  if (lvl > 0 && (lvl - 1 >= (int)c->levels.size() || !c->levels[lvl - 1].defined))
    return set_err(UCLBS_E_STATE, "define level %d before level %d", lvl - 1, lvl);
  // validity and disjointness of the valid regions
  int64_t vol = 0, stored = 0;
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < 3; ++d)
      if (hi[3 * b + d] < lo[3 * b + d])
        return set_err(UCLBS_E_ARG, "box %d is empty on axis %d", b, d);
  for (int b = 0; b < nbox; ++b) {
    int64_t v = 1, sv = 1;
    for (int d = 0; d < 3; ++d) {
      v *= hi[3 * b + d] - lo[3 * b + d] + 1;
      sv *= hi[3 * b + d] - lo[3 * b + d] + 3;
    }
    vol += v;
    stored += sv;
    for (int o = 0; o < b; ++o) {
      bool meet = true;
      for (int d = 0; d < 3; ++d)
        meet = meet && lo[3 * b + d] <= hi[3 * o + d] && lo[3 * o + d] <= hi[3 * b + d];
      if (meet) return set_err(UCLBS_E_ARG, "boxes %d and %d overlap", o, b);
    }
  }
  if (stored > (1ll << 30)) return set_err(UCLBS_E_ARG, "level exceeds 2^30 stored nodes");
  if ((int)c->levels.size() <= lvl) c->levels.resize(lvl + 1);
  Level& L = c->levels[lvl];
  free_level(L);
  invalidate_graph(c);
  L.multibox = true;
  L.blo.assign(lo, lo + 3 * nbox);
  L.bhi.assign(hi, hi + 3 * nbox);
  unsigned off = 0;
  for (int b = 0; b < nbox; ++b) {
    uclbs::BoxDesc bx{};
    const unsigned e[3] = {(unsigned)(hi[3 * b] - lo[3 * b] + 3), (unsigned)(hi[3 * b + 1] - lo[3 * b + 1] + 3),
                           (unsigned)(hi[3 * b + 2] - lo[3 * b + 2] + 3)};
    bx.off = off;
    bx.nx = e[0];
    bx.ny = e[1];
    bx.n = e[0] * e[1] * e[2];
    off += bx.n;
    L.hbox.push_back(bx);
  }
  for (int d = 0; d < 3; ++d) {
    L.dom_lo[d] = INT32_MAX;
    L.dom_hi[d] = INT32_MIN;
    for (int b = 0; b < nbox; ++b) {
      L.dom_lo[d] = std::min(L.dom_lo[d], lo[3 * b + d]);
      L.dom_hi[d] = std::max(L.dom_hi[d], hi[3 * b + d]);
    }
  }
  if (lvl == 0) {
    // level 0: the boxes tile their bounding box, periodic on every axis
    int64_t dv = 1;
    for (int d = 0; d < 3; ++d) dv *= L.dom_hi[d] - L.dom_lo[d] + 1;
    if (dv != vol) {
      free_level(L);
      return set_err(UCLBS_E_ARG, "level 0 boxes cover %lld of the %lld nodes of their bounding "
                                  "box (they must tile it)", (long long)vol, (long long)dv);
    }
  }
  // a flat stored "box" for the element-wise kernels and the far queue
  L.box.nx = (int)off;
  L.box.ny = L.box.nz = 1;
  L.box.n = off;
  for (int d = 0; d < 3; ++d) {
    L.box.per[d] = lvl == 0;
    L.lo[d] = L.dom_lo[d];
    L.hi[d] = L.dom_hi[d];
    L.face_code[d] = L.face_code[d + 3] = -10;
  }
  build_buckets(L);
  // ghost sources
  std::vector<uint8_t> hf((size_t)off, uclbs::kFlagInterior);
  std::vector<uint4> alias((size_t)off, make_uint4(0u, 0u, 0u, 0u));
  std::vector<uint3> copies;  // none: mirrors are updated by the streaming kernel
  std::vector<uclbs::C2FStencil> c2f;
  const Level* C = lvl > 0 ? &c->levels[lvl - 1] : nullptr;
  for (int b = 0; b < nbox; ++b) {
    const uclbs::BoxDesc& bx = L.hbox[b];
    const unsigned nz = bx.n / (bx.nx * bx.ny);
    for (unsigned z = 0; z < nz; ++z)
      for (unsigned y = 0; y < bx.ny; ++y)
        for (unsigned x = 0; x < bx.nx; ++x) {
          if (x > 0 && y > 0 && z > 0 && x + 1 < bx.nx && y + 1 < bx.ny && z + 1 < nz) continue;
          const unsigned k = bx.off + x + bx.nx * (y + bx.ny * z);
          hf[k] = uclbs::kFlagGhost;
          // populations a valid node pulls from this ghost: c_d = +1 across a
          // low face, -1 across a high face
          const unsigned loc[3] = {x, y, z}, ext[3] = {bx.nx, bx.ny, nz};
          unsigned mask = 0;
          for (int i = 0; i < 27; ++i) {
            const int ci[3] = {uclbs::CX27[i], uclbs::CY27[i], uclbs::CZ27[i]};
            bool need = true;
            for (int d = 0; d < 3; ++d) {
              if (loc[d] == 0) need = need && ci[d] == 1;
              if (loc[d] + 1 == ext[d]) need = need && ci[d] == -1;
            }
            if (need) mask |= 1u << i;
          }
          const int G[3] = {lo[3 * b] - 1 + (int)x, lo[3 * b + 1] - 1 + (int)y, lo[3 * b + 2] - 1 + (int)z};
          int sbox = -1;
          const int64_t src = valid_index(L, G, &sbox);
          if (src >= 0) {
            hf[k] = uclbs::kFlagMirror;
            const uclbs::BoxDesc& sb = L.hbox[sbox];
            alias[k] = make_uint4((unsigned)src, sb.nx, sb.nx * sb.ny, 0u);
            continue;
          }
          uclbs::C2FStencil st{};
          st.dst = k;
          st.mask = mask;
          if (!C || !c2f_stencil(*C, G, st)) {
            free_level(L);
            return set_err(UCLBS_E_ARG, "ghost node (%d,%d,%d) of box %d of level %d is neither a "
                                        "valid node of the level nor interpolable from level %d",
                           G[0], G[1], G[2], b, lvl, lvl - 1);
          }
          c2f.push_back(st);
        }
  }
  // fine_to_coarse: coarse node C <- fine node 2C for every even valid fine node
  std::vector<uint3> inj;
  if (C) {
    for (int b = 0; b < nbox; ++b) {
      int c0[3], c1[3];
      for (int d = 0; d < 3; ++d) {
        c0[d] = (lo[3 * b + d] + 1) >> 1;
        c1[d] = hi[3 * b + d] >> 1;
      }
      for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
          for (int x = c0[0]; x <= c1[0]; ++x) {
            const int P[3] = {x, y, z}, Fp[3] = {2 * x, 2 * y, 2 * z};
            const int64_t kc = valid_index(*C, P), kf = valid_index(L, Fp);
            if (kc < 0 || kf < 0) {
              free_level(L);
              return set_err(UCLBS_E_ARG, "level %d box %d is not nested in level %d", lvl, b, lvl - 1);
            }
            inj.push_back(make_uint3((unsigned)kc, (unsigned)kf, uclbs::kAllPops27));
          }
    }
  }
  const size_t bytes = (size_t)off * q_of(c) * elem_size(c);
  for (auto& p : L.f) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      free_level(L);
      return set_err(UCLBS_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
  }
  HIP_TRY(hipMalloc(&L.alpha, (size_t)off * sizeof(double)));
  HIP_TRY(hipMalloc(&L.defer_count, 2 * (uclbs::kFarShards + 1) * sizeof(unsigned)));
  HIP_TRY(hipMemset(L.defer_count, 0, 2 * (uclbs::kFarShards + 1) * sizeof(unsigned)));
  L.far_parity = 0;
  // queue shards sized for tiles of 64 nodes (the smallest tile shape)
  {
    const int64_t cap = (((int64_t)off + 63) / 64 + nbox + uclbs::kFarShards - 1) / uclbs::kFarShards * 256;
    L.defer_cap = (unsigned)cap;
    HIP_TRY(hipMalloc(&L.defer_list, (size_t)(cap * uclbs::kFarShards + 1) * sizeof(uint2)));
  }
  HIP_TRY(hipMalloc(&L.flags, (size_t)off));
  HIP_TRY(hipMemcpy(L.flags, hf.data(), hf.size(), hipMemcpyHostToDevice));
  HIP_TRY(upload_vec(&L.d_boxes, L.hbox));
  HIP_TRY(build_tiles(L, 256));
  HIP_TRY(upload_vec(&L.d_copies, copies));
  L.ncopies = (unsigned)copies.size();
  HIP_TRY(upload_vec(&L.d_alias, alias));
  HIP_TRY(upload_vec(&L.d_c2f, c2f));
  L.nc2f = (unsigned)c2f.size();
  HIP_TRY(upload_vec(&L.d_inject, inj));
  L.ninject = (unsigned)inj.size();
  L.refined = lvl > 0;
  L.step = 0;
  L.cur = 0;
  L.defined = true;
  return UCLBS_OK;
}

int uclbs_upload_f(void* ctx, int lvl, const double* f, int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "D3Q39 state moves through uclbs_upload_state / "
                                        "uclbs_download_state");
  if (!f || n_nodes != L->box.n)
    return set_err(UCLBS_E_ARG, "upload: expected %lld nodes, got %lld", (long long)L->box.n,
                   (long long)n_nodes);
  if ((rc = set_device(c))) return rc;
  const size_t bytes = (size_t)L->box.n * q_of(c) * sizeof(double);
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes));
  hipError_t e = hipMemcpyAsync(tmp, f, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = uclbs::launch_d3q27_from_fp64(L->box, tmp, L->f[L->cur],
                                       c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->f[1 - L->cur], L->f[L->cur], level_bytes(c, L),
                       hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "upload failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_download_f(void* ctx, int lvl, double* f, int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "D3Q39 state moves through uclbs_upload_state / "
                                        "uclbs_download_state");
  if (!f || n_nodes != L->box.n)
    return set_err(UCLBS_E_ARG, "download: expected %lld nodes, got %lld", (long long)L->box.n,
                   (long long)n_nodes);
  if ((rc = set_device(c))) return rc;
  const size_t bytes = (size_t)L->box.n * q_of(c) * sizeof(double);
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes));
  hipError_t e = uclbs::launch_d3q27_to_fp64(L->box, L->f[L->cur], tmp,
                                             c->cfg.precision == UCLBS_PREC_FP32_SHIFTED,
                                             c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(f, tmp, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "download failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_download_macro(void* ctx, int lvl, double* rho, double* u, double* v, double* w) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "D3Q39 state moves through uclbs_upload_state / "
                                        "uclbs_download_state");
  if (!rho || !u || !v || !w) return set_err(UCLBS_E_ARG, "null output array");
  if ((rc = set_device(c))) return rc;
  const size_t bytes = (size_t)L->box.n * sizeof(double);
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, 4 * bytes));
  const int64_t n = L->box.n;
  hipError_t e = uclbs::launch_d3q27_macro(L->box, L->f[L->cur],
                                           c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, tmp,
                                           tmp + n, tmp + 2 * n, tmp + 3 * n, c->stream);
  double* outs[4] = {rho, u, v, w};
  for (int k = 0; k < 4 && e == hipSuccess; ++k)
    e = hipMemcpyAsync(outs[k], tmp + k * n, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "macro failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_set_keep_alpha(void* ctx, int keep) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  c->keep_alpha = keep != 0;
  return UCLBS_OK;
}

int uclbs_download_alpha(void* ctx, int lvl, double* alpha) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (!alpha) return set_err(UCLBS_E_ARG, "null output array");
  if (!c->keep_alpha) return set_err(UCLBS_E_STATE, "alpha is not kept (uclbs_set_keep_alpha)");
  if ((rc = set_device(c))) return rc;
  const size_t bytes = (size_t)L->box.n * sizeof(double);
  hipError_t e = hipMemcpyAsync(alpha, L->alpha, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "alpha download failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_init_tgv(void* ctx, int lvl, double u0, double noise_amp, uint64_t seed) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "init_tgv is a D3Q27 initialiser");
  if (L->multibox && lvl > 0)
    return set_err(UCLBS_E_UNSUPPORTED, "init_tgv fills level 0 (refined levels: uclbs_init_from_coarse)");
  if (!(fabs(u0) < 0.5) || !(noise_amp >= 0.0))
    return set_err(UCLBS_E_ARG, "init_tgv: |u0| must be < 0.5 and noise_amp >= 0");
  if ((rc = set_device(c))) return rc;
  if (L->multibox) {
    // every stored node of every box from its domain position (ghosts wrap)
    uclbs::Box dom{};
    dom.nx = L->dom_hi[0] - L->dom_lo[0] + 1;
    dom.ny = L->dom_hi[1] - L->dom_lo[1] + 1;
    dom.nz = L->dom_hi[2] - L->dom_lo[2] + 1;
    dom.n = (int64_t)dom.nx * dom.ny * dom.nz;
    for (size_t b = 0; b < L->hbox.size(); ++b) {
      const uclbs::BoxDesc& bx = L->hbox[b];
      const int ext[3] = {(int)bx.nx, (int)bx.ny, (int)(bx.n / (bx.nx * bx.ny))};
      const int org[3] = {L->blo[3 * b] - 1 - L->dom_lo[0], L->blo[3 * b + 1] - 1 - L->dom_lo[1],
                          L->blo[3 * b + 2] - 1 - L->dom_lo[2]};
      HIP_TRY(uclbs::launch_d3q27_init_tgv_region(dom, L->f[L->cur], L->box.n, bx.off, ext, org,
                                                  c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, u0,
                                                  noise_amp, seed, c->stream));
    }
    if ((rc = mirror_buffers(c, L))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return UCLBS_OK;
  }
  HIP_TRY(uclbs::launch_d3q27_init_tgv(L->box, L->f[L->cur],
                                       c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, u0,
                                       noise_amp, seed, c->stream));
  if ((rc = mirror_buffers(c, L))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UCLBS_OK;
}

int uclbs_init_uniform(void* ctx, int lvl, double rho, double u, double v, double w) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "init_uniform is a D3Q27 initialiser");
  if (!(rho > 0.0) || !(fabs(u) < 0.5 && fabs(v) < 0.5 && fabs(w) < 0.5))
    return set_err(UCLBS_E_ARG, "init_uniform: rho must be > 0 and |u| < 0.5");
  if ((rc = set_device(c))) return rc;
  HIP_TRY(uclbs::launch_d3q27_init_uniform(L->box, L->f[L->cur],
                                           c->cfg.precision == UCLBS_PREC_FP32_SHIFTED,
                                           L->has_links ? L->flags : nullptr, rho, u, v, w,
                                           c->stream));
  if ((rc = mirror_buffers(c, L))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UCLBS_OK;
}

int uclbs_set_faces(void* ctx, int lvl, const int32_t face_code[6]) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (L->multibox) return set_err(UCLBS_E_UNSUPPORTED, "no link tables on levels of many boxes");
  if (L->refined) return set_err(UCLBS_E_UNSUPPORTED, "level %d is a refined level (no link tables)", lvl);
  if (!face_code) return set_err(UCLBS_E_ARG, "null face_code");
  const bool q39 = c->cfg.lattice == UCLBS_LATTICE_D3Q39;
  for (int f = 0; f < 6; ++f) {
    const int32_t v = face_code[f];
    const bool per = v == uclbs::kFacePeriodic;
    const bool ok = q39 ? (per || v == -10 || v == -1000 || v == -6 || v == -11 ||
                           (v <= -101 && v >= -199) || (v <= -201 && v >= -299))
                        : (per || v == -10 || v == -1000 || (v <= -101 && v >= -199) ||
                           (v <= -201 && v >= -299));
    if (!ok)
      return set_err(UCLBS_E_ARG, "face %d: code %d is not a %s boundary code", f, v,
                     q39 ? "D3Q39" : "D3Q27");
    if (q39 && v == -11 && f % 3 == 0 && c->shifted39)
      return set_err(UCLBS_E_ARG, "face %d: no inviscid face normal to the shift axis", f);
    if (per != (bool)L->box.per[f % 3])
      return set_err(UCLBS_E_ARG, "face %d: periodic code must match the level's periodic axes", f);
  }
  for (int f = 0; f < 6; ++f) L->face_code[f] = face_code[f];
  L->near_dirty = true;
  invalidate_graph(c);
  return UCLBS_OK;
}

int uclbs_set_bc_states(void* ctx, int n_in, const double* in_states, int n_out,
                        const double* out_states, const double* fs_state) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (n_in < 0 || n_in > 99 || n_out < 0 || n_out > 99 || (n_in && !in_states) ||
      (n_out && !out_states))
    return set_err(UCLBS_E_ARG, "bad inlet/outlet counts");
  int rc = set_device(c);
  if (rc) return rc;
  // a captured advance holds the old tables' addresses
  invalidate_graph(c);
  auto table = [&](int m, const double* st, double** dptr) -> int {
    if (*dptr) { (void)hipFree(*dptr); *dptr = nullptr; }
    if (m == 0) return UCLBS_OK;
    std::vector<double> h((size_t)m * 27);
    for (int j = 0; j < m; ++j) {
      const double* s4 = st + 4 * j;
      if (!(s4[0] > 0.0) || !(fabs(s4[1]) < 0.5 && fabs(s4[2]) < 0.5 && fabs(s4[3]) < 0.5))
        return set_err(UCLBS_E_ARG, "boundary state %d: rho must be > 0 and |u| < 0.5", j);
      feq27_host(s4[0], s4 + 1, h.data() + 27 * j);
    }
    HIP_TRY(hipMalloc(dptr, h.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(*dptr, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return UCLBS_OK;
  };
  if ((rc = table(n_in, in_states, &c->d_feq_in))) return rc;
  if ((rc = table(n_out, out_states, &c->d_feq_out))) return rc;
  if ((rc = table(fs_state ? 1 : 0, fs_state, &c->d_feq_fs))) return rc;
  c->n_in = n_in;
  c->n_out = n_out;
  return UCLBS_OK;
}

int uclbs_build_links(void* ctx, int lvl, const double origin[3], double dx, int64_t ntri,
                      const double* tri, const double* nrm, int64_t* n_solid, int64_t* n_tie) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  const bool q39 = c->cfg.lattice == UCLBS_LATTICE_D3Q39;
  if (L->multibox && L->box.per[0])
    return set_err(UCLBS_E_UNSUPPORTED, "no link tables on a periodic level of many boxes");
  if (L->refined && q39)
    return set_err(UCLBS_E_UNSUPPORTED, "level %d is a refined D3Q39 level (no link tables)", lvl);
  if (!origin || !(dx > 0.0) || ntri < 0 || ntri > (1 << 26) || (ntri && (!tri || !nrm)))
    return set_err(UCLBS_E_ARG, "build_links: bad geometry arguments");
  if ((rc = set_device(c))) return rc;
  const int64_t n = L->box.n;
  uclbs::LinkGeom g{};
  g.box = L->box;
  for (int d = 0; d < 3; ++d) g.origin[d] = origin[d];
  g.dx = dx;
  for (int f = 0; f < 6; ++f) {
    // an outflow face (outlet type 6) gives its links the code of the node
    // outside, a fluid node of level 0 (state_directions_lvl.f95:131-132)
    // (the shifted lattice's outflow links take the downstream node's code)
    g.face_code[f] = (q39 && !c->shifted39 && L->face_code[f] == -6) ? 0 : L->face_code[f];
  }
  g.ntri = (int)ntri;
  std::vector<double> tbb((size_t)(ntri > 0 ? ntri : 1) * 6);
  double gbb[6] = {-1e300, 1e300, -1e300, 1e300, -1e300, 1e300};
  for (int64_t t = 0; t < ntri; ++t)
    for (int d = 0; d < 3; ++d) {
      const double a = tri[t * 9 + d], b = tri[t * 9 + 3 + d], e = tri[t * 9 + 6 + d];
      double mx = a > b ? a : b; mx = mx > e ? mx : e;
      double mn = a < b ? a : b; mn = mn < e ? mn : e;
      tbb[t * 6 + 2 * d] = mx;
      tbb[t * 6 + 2 * d + 1] = mn;
      if (mx > gbb[2 * d]) gbb[2 * d] = mx;
      if (mn < gbb[2 * d + 1]) gbb[2 * d + 1] = mn;
    }
  for (int i = 0; i < 6; ++i) g.gbb[i] = gbb[i];
  std::vector<uclbs::BvhNode> bvh;
  std::vector<int32_t> bvh_idx;
  if (uclbs::build_bvh(tbb.data(), (int)ntri, &bvh, &bvh_idx) > uclbs::kBvhMaxDepth)
    return set_err(UCLBS_E_ARG, "build_links: triangle hierarchy deeper than %d",
                   uclbs::kBvhMaxDepth);
  double *dtri = nullptr, *dnrm = nullptr, *dtbb = nullptr;
  uclbs::BvhNode* dbvh = nullptr;
  int32_t* dbidx = nullptr;
  unsigned long long* dcnt = nullptr;
  const size_t nt = (size_t)(ntri > 0 ? ntri : 1);
  const size_t nb = bvh.empty() ? 1 : bvh.size();
  hipError_t e = hipMalloc(&dtri, nt * 9 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dbvh, nb * sizeof(uclbs::BvhNode));
  if (e == hipSuccess) e = hipMalloc(&dbidx, nt * sizeof(int32_t));
  if (e == hipSuccess && !bvh.empty())
    e = hipMemcpy(dbvh, bvh.data(), bvh.size() * sizeof(uclbs::BvhNode), hipMemcpyHostToDevice);
  if (e == hipSuccess && ntri)
    e = hipMemcpy(dbidx, bvh_idx.data(), (size_t)ntri * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&dnrm, nt * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dtbb, nt * 6 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dcnt, 2 * sizeof(unsigned long long));
  if (e == hipSuccess && ntri) e = hipMemcpy(dtri, tri, (size_t)ntri * 9 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && ntri) e = hipMemcpy(dnrm, nrm, (size_t)ntri * 3 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dtbb, tbb.data(), nt * 6 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dcnt, 0, 2 * sizeof(unsigned long long));
  if (e == hipSuccess && !L->state) e = hipMalloc(&L->state, (size_t)n * q_of(c) * sizeof(int32_t));
  unsigned long long hc[2] = {0, 0};
  if (e == hipSuccess) {
    g.tri = dtri;
    g.nrm = dnrm;
    g.tbb = dtbb;
    g.bvh = dbvh;
    g.bvh_idx = dbidx;
    if (L->multibox) {
      // box by box: the box's stored nodes (ghost shell included) at level
      // position blo - 1 + (a, b, c), i.e. origin + (blo - 1 + (a, b, c)) dx;
      // built into a scratch table and placed at the box's offset
      size_t maxn = 0;
      for (const uclbs::BoxDesc& bx : L->hbox) maxn = std::max<size_t>(maxn, bx.n);
      // Every node's solid test casts along one direction, nearly along +x:
      // the 64 rays of a wavefront (consecutive nodes) walk the same part of
      // the triangle hierarchy, and the segment's bounding box -- all that the
      // candidate test looks at -- stays a thin tube.  Rays away from each
      // small box's centre fan out over all directions, their long diagonal
      // segments keep most of the geometry as candidates, and a wavefront
      // walks the union of their paths (C4's 793 boxes around the sphere: 51
      // ms per box).  The parity of a cast does not depend on its direction
      // (ties re-cast), so the table is the same.  The length spans the boxes
      // and the geometry together.
      {
        double lo3[3], hi3[3];
        for (int d = 0; d < 3; ++d) {
          lo3[d] = g.ntri > 0 ? std::min(g.gbb[2 * d + 1], origin[d]) : origin[d];
          hi3[d] = g.ntri > 0 ? std::max(g.gbb[2 * d], origin[d]) : origin[d];
        }
        for (size_t b = 0; b < L->hbox.size(); ++b)
          for (int d = 0; d < 3; ++d) {
            lo3[d] = std::min(lo3[d], origin[d] + (double)(L->blo[3 * b + d] - 1) * dx);
            hi3[d] = std::max(hi3[d], origin[d] + (double)(L->bhi[3 * b + d] + 1) * dx);
          }
        double diag = 0.0;
        for (int d = 0; d < 3; ++d) diag += (hi3[d] - lo3[d]) * (hi3[d] - lo3[d]);
        const double r[3] = {1.0, 7.31e-4, 3.77e-4};
        const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (int d = 0; d < 3; ++d) g.ray[d] = r[d] / rn;
        g.ray_len = 2.0 * std::sqrt(diag) + 4.0 * dx;
      }
      int32_t* tmp = nullptr;
      e = hipMalloc(&tmp, maxn * 27 * sizeof(int32_t));
      for (size_t b = 0; b < L->hbox.size() && e == hipSuccess; ++b) {
        const uclbs::BoxDesc& bx = L->hbox[b];
        uclbs::LinkGeom gb = g;
        gb.box.nx = (int)bx.nx;
        gb.box.ny = (int)bx.ny;
        gb.box.nz = (int)(bx.n / (bx.nx * bx.ny));
        gb.box.n = bx.n;
        for (int d = 0; d < 3; ++d) {
          gb.box.per[d] = 1;  // a valid node's links stay in its stored box
          gb.origin[d] = origin[d] + (double)(L->blo[3 * b + d] - 1) * dx;
        }
        for (int f = 0; f < 6; ++f) gb.face_code[f] = uclbs::kFacePeriodic;
        e = uclbs::launch_build_links27(gb, tmp, dcnt, c->stream);
        if (e == hipSuccess)
          e = hipMemcpy2DAsync(L->state + bx.off, (size_t)n * sizeof(int32_t), tmp,
                               (size_t)bx.n * sizeof(int32_t), (size_t)bx.n * sizeof(int32_t), 27,
                               hipMemcpyDeviceToDevice, c->stream);
      }
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (tmp) (void)hipFree(tmp);
    } else {
      e = !q39 ? uclbs::launch_build_links27(g, L->state, dcnt, c->stream)
          : c->shifted39 ? uclbs::launch_build_links39s(g, L->state, dcnt, c->stream)
                         : uclbs::launch_build_links39(g, L->state, dcnt, c->stream);
    }
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hc, dcnt, sizeof hc, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dtri);
  (void)hipFree(dnrm);
  (void)hipFree(dtbb);
  (void)hipFree(dbvh);
  (void)hipFree(dbidx);
  (void)hipFree(dcnt);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "build_links failed: %s", hipGetErrorString(e));
  if (q39) {
    L->table39 = true;
    L->near_dirty = true;
    L->nsolid = (int64_t)hc[0];
    invalidate_graph(c);
  } else if ((rc = L->multibox ? finalize_links_boxes(c, L) : finalize_links(c, L))) {
    return rc;
  }
  if (n_solid) *n_solid = (int64_t)hc[0];
  if (n_tie) *n_tie = (int64_t)hc[1];
  return UCLBS_OK;
}

int uclbs_set_links(void* ctx, int lvl, const int32_t* state, int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (L->multibox) return set_err(UCLBS_E_UNSUPPORTED, "no link tables on levels of many boxes");
  if (L->refined && c->cfg.lattice == UCLBS_LATTICE_D3Q39)
    return set_err(UCLBS_E_UNSUPPORTED, "level %d is a refined D3Q39 level (no link tables)", lvl);
  if (!state || n_nodes != L->box.n) return set_err(UCLBS_E_ARG, "set_links: wrong node count");
  if ((rc = set_device(c))) return rc;
  const size_t bytes = (size_t)n_nodes * q_of(c) * sizeof(int32_t);
  if (!L->state) HIP_TRY(hipMalloc(&L->state, bytes));
  HIP_TRY(hipMemcpy(L->state, state, bytes, hipMemcpyHostToDevice));
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) {
    L->table39 = true;
    L->near_dirty = true;
    invalidate_graph(c);
    return UCLBS_OK;
  }
  return finalize_links(c, L);
}

int uclbs_download_links(void* ctx, int lvl, int32_t* state, int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (!L->state) return set_err(UCLBS_E_STATE, "level %d has no link table", lvl);
  if (!state || n_nodes != L->box.n) return set_err(UCLBS_E_ARG, "download_links: wrong node count");
  if ((rc = set_device(c))) return rc;
  HIP_TRY(hipMemcpy(state, L->state, (size_t)n_nodes * q_of(c) * sizeof(int32_t),
                    hipMemcpyDeviceToHost));
  return UCLBS_OK;
}

int uclbs_link_counts(void* ctx, int lvl, int64_t* n_boundary, int64_t* n_solid) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (L->table39)
    return set_err(UCLBS_E_UNSUPPORTED, "D3Q39 tables: n_solid comes from uclbs_build_links");
  if (!L->has_links) return set_err(UCLBS_E_STATE, "level %d has no link table", lvl);
  if (n_boundary) *n_boundary = L->nb;
  if (n_solid) *n_solid = L->nsolid;
  return UCLBS_OK;
}

int uclbs_step_level_async(void* ctx, int lvl, int nsteps) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (nsteps < 0) return set_err(UCLBS_E_ARG, "nsteps < 0");
  if ((rc = set_device(c))) return rc;
  return launch_steps(c, L, nsteps);
}

int uclbs_step_level(void* ctx, int lvl, int nsteps) {
  int rc = uclbs_step_level_async(ctx, lvl, nsteps);
  if (rc) return rc;
  if ((rc = uclbs_sync(ctx))) return rc;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) {
    // the reference stops in error_out on these (calculate_alpha.f95:303-313,
    // quick_calcs.f95:149-153, 207-212)
    uclbs_stats st;
    if ((rc = uclbs_get_stats(ctx, &st))) return rc;
    const int64_t errs = st.n_nonpos + st.n_fallback;
    if (errs > c->err_seen) {
      c->err_seen = errs;
      return set_err(UCLBS_E_NUMERIC,
                     "D3Q39 step: %lld nodes without a real alpha root, %lld outside the "
                     "conductivity fit",
                     (long long)st.n_nonpos, (long long)st.n_fallback);
    }
  }
  return UCLBS_OK;
}

int uclbs_time_steps(void* ctx, int lvl, int nsteps, double* elapsed_ms) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (!elapsed_ms || nsteps <= 0) return set_err(UCLBS_E_ARG, "bad timing arguments");
  if ((rc = set_device(c))) return rc;
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  if ((rc = launch_steps(c, L, nsteps))) return rc;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *elapsed_ms = ms;
  return UCLBS_OK;
}

int uclbs_time_kernels(void* ctx, int lvl, int nsteps, double* step_ms, double* main_ms) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (!step_ms || !main_ms || nsteps <= 0 || nsteps > 100000)
    return set_err(UCLBS_E_ARG, "bad timing arguments");
  if ((rc = set_device(c))) return rc;
  std::vector<hipEvent_t> ev((size_t)2 * nsteps, nullptr);
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
  if (e == hipSuccess) rc = launch_steps(c, L, nsteps, ev.data());
  if (e == hipSuccess && rc == UCLBS_OK) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess && rc == UCLBS_OK) e = hipEventSynchronize(c->ev1);
  double tot = 0.0;
  if (e == hipSuccess && rc == UCLBS_OK) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    *step_ms = ms;
    for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
      e = hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]);
      tot += ms;
    }
    *main_ms = tot;
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc) return rc;
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "kernel timing failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_sync(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  int rc = set_device(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UCLBS_OK;
}

int uclbs_get_stats(void* ctx, uclbs_stats* out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !out) return set_err(UCLBS_E_ARG, "null argument");
  int rc = set_device(c);
  if (rc) return rc;
  std::vector<unsigned long long> h(uclbs::kStatSlots * uclbs::kStatFields);
  HIP_TRY(hipMemcpyAsync(h.data(), c->d_stats, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int64_t acc[uclbs::kStatFields] = {0};
  for (int s = 0; s < uclbs::kStatSlots; ++s)
    for (int k = 0; k < uclbs::kStatFields; ++k) acc[k] += (int64_t)h[s * uclbs::kStatFields + k];
  out->node_updates = acc[uclbs::ST_NODES];
  out->newton_iters = acc[uclbs::ST_ITERS];
  out->n_trivial = acc[uclbs::ST_TRIVIAL];
  out->n_bound = acc[uclbs::ST_BOUND];
  out->n_nonpos = acc[uclbs::ST_NONPOS];
  out->n_noconv = acc[uclbs::ST_NOCONV];
  out->n_fallback = acc[uclbs::ST_FALLBACK];
  out->n_logpath = acc[uclbs::ST_LOGPATH];
  out->n_ibb_links = acc[uclbs::ST_IBB];
  out->n_repair = acc[uclbs::ST_REPAIR];
  return UCLBS_OK;
}

int uclbs_reset_stats(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  int rc = set_device(c);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->d_stats, 0,
                         sizeof(unsigned long long) * uclbs::kStatSlots * uclbs::kStatFields,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->err_seen = 0;
  return UCLBS_OK;
}

// ---- D3Q39 ----------------------------------------------------------------

int uclbs_set_gas(void* ctx, const uclbs_gas* gas) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !gas) return set_err(UCLBS_E_ARG, "null argument");
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q39)
    return set_err(UCLBS_E_STATE, "uclbs_set_gas needs a D3Q39 context");
  if (!(gas->gama > 1.0) || !(gas->temperature > 0.0) || !(gas->pressure > 0.0) ||
      !(gas->mach >= 0.0) || !(gas->dx > 0.0) || gas->startup_timesteps <= 0)
    return set_err(UCLBS_E_ARG, "bad gas parameters");
  for (auto& L : c->levels)
    if (L.defined) return set_err(UCLBS_E_STATE, "uclbs_set_gas after levels were defined");
  uclbs::Gas39& g = c->gas39;
  g = uclbs::Gas39{};
  // basic_fluids_calcs.f95:17-27, 31-38, 86
  const double R = 287.052874;
  g.gama = gas->gama;
  g.R = R;
  g.rho_ref = gas->pressure / (R * gas->temperature);
  g.t_ref = gas->temperature;
  g.p_ref = gas->pressure;
  g.cv = 1.0 / (gas->gama - 1.0);
  g.cp = gas->gama / (gas->gama - 1.0);
  g.lead_g = 2.0 * g.cv - 3.0;
  const double temperature = 1.0 / gas->gama, rho_inf = 1.0;
  const double sos = std::sqrt(gas->gama * R * g.t_ref);
  g.molec_weight = (double)28.97f / (6.022e23 * 1000);
  // viscosity_and_timing.f95:29-68
  const double S = (double)110.4f, visc_ref = (double)1.716E-5f, temp_ref = (double)273.15f;
  const double viscosity =
      visc_ref * std::pow(g.t_ref / temp_ref, 1.5) * ((temp_ref + S) / (g.t_ref + S));
  g.mu_ref = g.rho_ref * sos * gas->mach * 1;
  g.kappa_ref = 1.0;  // kappa_bottom = 1 for the start-up conductivity
  double kappa_base = 0.0;
  if (!uclbs::therm_cond39(rho_inf, temperature, g, &kappa_base))
    return set_err(UCLBS_E_ARG, "freestream outside the conductivity fit");
  g.kappa_ref = rho_inf * sos * sos * sos / temperature;
  g.init_beta1 = 0.5001;
  g.init_beta2 = 0.5001;
  const double b1 = 1.0 / ((2.0 * viscosity) / (rho_inf * temperature) + 1.0);
  const double b2 = 1.0 / ((2.0 * kappa_base / g.kappa_ref) / (rho_inf * g.cp * temperature) + 1.0);
  g.startup = gas->startup_timesteps;
  g.beta1_scale = (b1 - g.init_beta1) / (double)g.startup;
  g.beta2_scale = (b2 - g.init_beta2) / (double)g.startup;
  // timing_calcs.f95:24-26 (level 0)
  g.lvl = 0;
  g.dt = gas->dx / sos;
  g.eps = g.dt / 1.0;
  g.turbulent = gas->turbulent != 0;
  g.maxit = gas->lm_maxit > 0 ? gas->lm_maxit : 50;
  c->shifted39 = gas->shifted < 0 ? (gas->mach >= 1.2 ? 1 : 0) : (gas->shifted != 0);
  c->gas39_mach = gas->mach;
  // freestream equilibria (freestream_basis_calcs.f90:35-71: the multiplier
  // solve of rho = 1, u = Mach along x, T = 1/gamma; fs_gieq = (2 cv - 3) T
  // fs_fieq); FP32 storage keeps deviations from them
  c->ref39 = uclbs::Ref39{};
  {
    const double u0 = gas->mach, T = temperature, rho = 1.0;
    const double qf = rho * u0 * u0 + 5 * rho * T;
    const double tg[uclbs::kM39] = {rho, rho * u0, 0, 0, rho * (T + u0 * u0), rho * T, rho * T,
                                    0, 0, 0, qf * u0, 0, 0};
    double lm[uclbs::kM39] = {-1.5 * std::log(2 * 3.141592653589793 * T) - u0 * u0 / (2 * T),
                              u0 / T, 0, 0, -1 / (2 * T), -1 / (2 * T), -1 / (2 * T), 0, 0, 0, 0, 0, 0};
    double lmp[uclbs::kM39];
    const int it = c->shifted39 ? uclbs::lm_newton39<true>(rho, tg, lm, lmp, 200)
                                : uclbs::lm_newton39<false>(rho, tg, lm, lmp, 200);
    if (it < 0) return set_err(UCLBS_E_NUMERIC, "freestream multipliers did not converge");
    for (int i = 0; i < uclbs::Q39; ++i) {
      const double e = c->shifted39 ? uclbs::expo39<true>(lm, rho, i) : uclbs::expo39<false>(lm, rho, i);
      c->fs39.f[i] = e;
      c->fs39.g[i] = g.lead_g * T * e;
    }
  }
  if (c->cfg.precision == UCLBS_PREC_FP32_SHIFTED) c->ref39 = c->fs39;
  int rc = set_device(c);
  if (rc) return rc;
  if (!c->d_ref39) HIP_TRY(hipMalloc(&c->d_ref39, 2 * uclbs::Q39 * sizeof(double)));
  HIP_TRY(hipMemcpy(c->d_ref39, &c->ref39, 2 * uclbs::Q39 * sizeof(double), hipMemcpyHostToDevice));
  c->gas_set = true;
  return UCLBS_OK;
}

// One inlet or outlet table row: the multiplier solve of the stored state from
// the reference's starting guesses (inlet_outlet_basis_calcs.f90: chi -2.5,
// pi_aa -0.3, the rest 0; shifted-lattice outlets zeta = u, lambda = u / 25).
static int bc39_row(Ctx* c, const double* st, bool outlet, double* row) {
  const double rho = st[0], u = st[1], v = st[2], w = st[3], T = st[4];
  if (!(rho > 0.0) || !(T > 0.0) || !std::isfinite(u + v + w))
    return set_err(UCLBS_E_ARG, "inlet/outlet state: rho and T must be positive");
  const double qf = rho * (u * u + v * v + w * w) + 5 * rho * T;
  const double tg[uclbs::kM39] = {rho, rho * u, rho * v, rho * w, rho * (T + u * u),
                                  rho * (T + v * v), rho * (T + w * w), rho * u * v, rho * u * w,
                                  rho * v * w, qf * u, qf * v, qf * w};
  double lm[uclbs::kM39] = {-2.5, 0, 0, 0, -0.3, -0.3, -0.3, 0, 0, 0, 0, 0, 0};
  if (outlet && c->shifted39) {
    lm[1] = u; lm[2] = v; lm[3] = w;
    lm[10] = u / 25; lm[11] = v / 25; lm[12] = w / 25;
  }
  double lmp[uclbs::kM39];
  const int it = c->shifted39 ? uclbs::lm_newton39<true>(rho, tg, lm, lmp, 200)
                              : uclbs::lm_newton39<false>(rho, tg, lm, lmp, 200);
  if (it < 0) return set_err(UCLBS_E_NUMERIC, "inlet/outlet multipliers did not converge");
  for (int i = 0; i < uclbs::Q39; ++i)
    row[i] = c->shifted39 ? uclbs::expo39<true>(lm, rho, i) : uclbs::expo39<false>(lm, rho, i);
  return UCLBS_OK;
}

int uclbs_set_bc_states39(void* ctx, int n_in, const double* in_states, const int32_t* in_types,
                          int n_out, const double* out_states, const int32_t* out_types) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q39)
    return set_err(UCLBS_E_UNSUPPORTED, "set_bc_states39 is for D3Q39 contexts (uclbs_set_bc_states for D3Q27)");
  if (!c->gas_set) return set_err(UCLBS_E_STATE, "set_bc_states39 needs uclbs_set_gas first");
  if (n_in < 0 || n_in > 99 || n_out < 0 || n_out > 99 || (n_in && (!in_states || !in_types)) ||
      (n_out && (!out_states || !out_types)))
    return set_err(UCLBS_E_ARG, "bad inlet/outlet counts");
  for (int j = 0; j < n_in; ++j) {
    const int t = in_types[j];
    if (t != 1 && t != 2 && t != 3 && t != 4 && t != 7)
      return set_err(UCLBS_E_ARG, "inlet %d: type %d is not 1, 2, 3, 4 or 7", j + 1, t);
  }
  for (int j = 0; j < n_out; ++j) {
    const int t = out_types[j];
    if (t != 1 && t != 2 && t != 3 && t != 4 && t != 6 && t != 7)
      return set_err(UCLBS_E_ARG, "outlet %d: type %d is not 1, 2, 3, 4, 6 or 7", j + 1, t);
  }
  const double lead = c->gas39.lead_g;
  std::vector<double> tab((size_t)std::max(1, n_in + n_out) * 2 * uclbs::Q39, 0.0);
  // inlet_outlet_basis_calcs.f90: each stored state's row, then g = (2 cv - 3)
  // T f over every row of its family with that state's T (the reference
  // assigns the whole g array in the loop, so the last such state's T holds
  // for all rows solved before it); a type-3 row is the freestream pair
  auto family = [&](int n, const double* st, const int32_t* ty, int base, bool outlet) -> int {
    for (int j = 0; j < n; ++j) {
      double* row = &tab[(size_t)(base + j) * 2 * uclbs::Q39];
      const int t = ty[j];
      if (t == 1 || t == 2 || t == 4 || t == 7) {
        int rc = bc39_row(c, st + 5 * j, outlet, row);
        if (rc) return rc;
        const double T = st[5 * j + 4];
        for (int k = 0; k < n; ++k) {
          double* r = &tab[(size_t)(base + k) * 2 * uclbs::Q39];
          for (int i = 0; i < uclbs::Q39; ++i) r[uclbs::Q39 + i] = r[i] * lead * T;
        }
      } else if (t == 3) {
        for (int i = 0; i < uclbs::Q39; ++i) {
          row[i] = c->fs39.f[i];
          row[uclbs::Q39 + i] = c->fs39.g[i];
        }
      }
    }
    return UCLBS_OK;
  };
  int rc = family(n_in, in_states, in_types, 0, false);
  if (!rc) rc = family(n_out, out_states, out_types, n_in, true);
  if (rc) return rc;
  if ((rc = set_device(c))) return rc;
  invalidate_graph(c);
  if (c->d_bc39) { (void)hipFree(c->d_bc39); c->d_bc39 = nullptr; }
  HIP_TRY(hipMalloc(&c->d_bc39, tab.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(c->d_bc39, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  c->in_types39.assign(in_types, in_types + n_in);
  c->out_types39.assign(out_types, out_types + n_out);
  return UCLBS_OK;
}

static int check39(Ctx* c, int lvl, Level** L) {
  int rc = check_level(c, lvl, L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q39)
    return set_err(UCLBS_E_STATE, "needs a D3Q39 context");
  return set_device(c);
}

int uclbs_init_fields(void* ctx, int lvl, const double* rho, const double* u, const double* v,
                      const double* w, const double* temp) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check39(c, lvl, &L);
  if (rc) return rc;
  if (!rho || !u || !v || !w || !temp) return set_err(UCLBS_E_ARG, "null field");
  const size_t n = (size_t)L->box.n;
  double* d = nullptr;
  unsigned long long* bad = nullptr;
  HIP_TRY(hipMalloc(&d, 5 * n * sizeof(double)));
  hipError_t e = hipMalloc(&bad, sizeof(unsigned long long));
  const double* src[5] = {rho, u, v, w, temp};
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(unsigned long long), c->stream);
  for (int k = 0; k < 5 && e == hipSuccess; ++k)
    e = hipMemcpyAsync(d + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = uclbs::launch_d3q39_init(L->box, d, L->f[L->cur], L->g[L->cur], L->lm,
                                 c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, c->shifted39,
                                 c->gas39, c->ref39, bad, c->stream);
  const size_t bytes = n * uclbs::Q39 * elem_size(c);
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->f[1 - L->cur], L->f[L->cur], bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->g[1 - L->cur], L->g[L->cur], bytes, hipMemcpyDeviceToDevice, c->stream);
  unsigned long long nbad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (bad) (void)hipFree(bad);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "init_fields failed: %s", hipGetErrorString(e));
  if (nbad) return set_err(UCLBS_E_NUMERIC, "multiplier Newton failed at %llu nodes", nbad);
  return UCLBS_OK;
}

int uclbs_upload_state(void* ctx, int lvl, const double* f, const double* g, const double* lm,
                       int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check39(c, lvl, &L);
  if (rc) return rc;
  if (!f || !g || !lm || n_nodes != L->box.n)
    return set_err(UCLBS_E_ARG, "upload_state: null array or %lld nodes (expected %lld)",
                   (long long)n_nodes, (long long)L->box.n);
  const size_t n = (size_t)L->box.n, pb = n * uclbs::Q39 * sizeof(double);
  const int fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, pb));
  hipError_t e = hipSuccess;
  const double* src[2] = {f, g};
  void* dst[2] = {L->f[L->cur], L->g[L->cur]};
  void* dst2[2] = {L->f[1 - L->cur], L->g[1 - L->cur]};
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipMemcpyAsync(tmp, src[k], pb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
      e = uclbs::launch_d3q39_from_fp64(L->box, tmp, dst[k], fp32, c->d_ref39 + k * uclbs::Q39,
                                        c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dst2[k], dst[k], n * uclbs::Q39 * elem_size(c), hipMemcpyDeviceToDevice,
                         c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->lm, lm, n * uclbs::kM39 * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "upload_state failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_download_state(void* ctx, int lvl, double* f, double* g, double* lm, int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check39(c, lvl, &L);
  if (rc) return rc;
  if (n_nodes != L->box.n)
    return set_err(UCLBS_E_ARG, "download_state: %lld nodes (expected %lld)", (long long)n_nodes,
                   (long long)L->box.n);
  const size_t n = (size_t)L->box.n, pb = n * uclbs::Q39 * sizeof(double);
  const int fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, pb));
  hipError_t e = hipSuccess;
  double* out[2] = {f, g};
  const void* src[2] = {L->f[L->cur], L->g[L->cur]};
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    if (!out[k]) continue;
    e = uclbs::launch_d3q39_to_fp64(L->box, src[k], tmp, fp32, c->d_ref39 + k * uclbs::Q39,
                                    c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out[k], tmp, pb, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess && lm)
    e = hipMemcpyAsync(lm, L->lm, n * uclbs::kM39 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "download_state failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_download_macro39(void* ctx, int lvl, double* rho, double* u, double* v, double* w,
                           double* temp, double* press, int64_t n_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check39(c, lvl, &L);
  if (rc) return rc;
  if (n_nodes != L->box.n)
    return set_err(UCLBS_E_ARG, "download_macro39: %lld nodes (expected %lld)",
                   (long long)n_nodes, (long long)L->box.n);
  const size_t n = (size_t)L->box.n, bytes = n * sizeof(double);
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, 6 * bytes));
  hipError_t e = uclbs::launch_d3q39_macro(L->box, L->f[L->cur], L->g[L->cur],
                                           c->cfg.precision == UCLBS_PREC_FP32_SHIFTED,
                                           c->shifted39, c->ref39, L->gas39.cv, tmp, c->stream);
  double* outs[6] = {rho, u, v, w, temp, press};
  for (int k = 0; k < 6 && e == hipSuccess; ++k)
    if (outs[k]) e = hipMemcpyAsync(outs[k], tmp + k * n, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "download_macro39 failed: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

static uclbs::LevelMap level_map(const Level& L) {
  uclbs::LevelMap m{};
  m.box = L.box;
  for (int d = 0; d < 3; ++d) {
    m.origin[d] = L.origin[d];
    m.per[d] = L.box.per[d];
  }
  return m;
}

// all_pops: a ghost takes all 27 interpolated populations, not only those its
// valid neighbours pull (for quantities computed on the ghost shell)
static int interlevel_fill(Ctx* c, int lvl, double theta, bool all_nodes, bool all_pops = false) {
  Level* F;
  int rc = check_level(c, lvl, &F);
  if (rc) return rc;
  if (F->stale)
    return set_err(UCLBS_E_STATE, "level %d: its coarser level was regridded (regrid or define it "
                                  "again)", lvl);
  if (!F->refined) return set_err(UCLBS_E_ARG, "level %d is not a refined level", lvl);
  if (!(theta >= 0.0 && theta <= 1.0)) return set_err(UCLBS_E_ARG, "theta outside [0, 1]");
  if ((rc = set_device(c))) return rc;
  const Level& C = c->levels[lvl - 1];
  const int sh = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39) {
    const uclbs::LevelMap cm = level_map(C), fm = level_map(*F);
    HIP_TRY(uclbs::launch_coarse_to_fine39(
        C.f[1 - C.cur], C.f[C.cur], C.g[1 - C.cur], C.g[C.cur], C.lm, cm, F->f[F->cur],
        F->g[F->cur], F->lm, fm, all_nodes ? nullptr : F->ghosts,
        all_nodes ? (unsigned)F->box.n : F->nghost, theta, sh, c->shifted39, c->ref39, C.gas39,
        c->d_stats, c->stream));
    if (all_nodes) {
      HIP_TRY(hipMemcpyAsync(F->f[1 - F->cur], F->f[F->cur], level_bytes(c, F),
                             hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(F->g[1 - F->cur], F->g[F->cur], level_bytes(c, F),
                             hipMemcpyDeviceToDevice, c->stream));
    }
    return UCLBS_OK;
  }
  if (F->multibox) {
    if (!all_nodes)
      return hip_rc(uclbs::launch_c2f_stencils27(C.f[1 - C.cur], C.f[C.cur], (unsigned)C.box.n,
                                                 F->f[F->cur], (unsigned)F->box.n, F->d_c2f,
                                                 F->nc2f, theta, sh, c->d_stats, c->stream,
                                                 all_pops));
    // start-up: every valid node (and every ghost the coarse level reaches)
    std::vector<uclbs::C2FStencil> all;
    for (size_t b = 0; b < F->hbox.size(); ++b) {
      const uclbs::BoxDesc& bx = F->hbox[b];
      const unsigned nz = bx.n / (bx.nx * bx.ny);
      for (unsigned z = 0; z < nz; ++z)
        for (unsigned y = 0; y < bx.ny; ++y)
          for (unsigned x = 0; x < bx.nx; ++x) {
            const bool ghost = x == 0 || y == 0 || z == 0 || x + 1 == bx.nx || y + 1 == bx.ny || z + 1 == nz;
            const int G[3] = {F->blo[3 * b] - 1 + (int)x, F->blo[3 * b + 1] - 1 + (int)y,
                              F->blo[3 * b + 2] - 1 + (int)z};
            uclbs::C2FStencil st{};
            st.dst = bx.off + x + bx.nx * (y + bx.ny * z);
            st.mask = uclbs::kAllPops27;
            if (c2f_stencil(C, G, st)) all.push_back(st);
            else if (!ghost) return set_err(UCLBS_E_STATE, "fine node outside the coarse level");
          }
    }
    uclbs::C2FStencil* d = nullptr;
    HIP_TRY(upload_vec(&d, all));
    hipError_t e = uclbs::launch_c2f_stencils27(C.f[1 - C.cur], C.f[C.cur], (unsigned)C.box.n,
                                                F->f[F->cur], (unsigned)F->box.n, d,
                                                (unsigned)all.size(), theta, sh, c->d_stats, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(F->f[1 - F->cur], F->f[F->cur], level_bytes(c, F), hipMemcpyDeviceToDevice,
                         c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return hip_rc(e);
  }
  const uclbs::LevelMap cm = level_map(C), fm = level_map(*F);
  if (all_nodes) {
    HIP_TRY(uclbs::launch_coarse_to_fine27(C.f[1 - C.cur], C.f[C.cur], cm, F->f[F->cur], fm, nullptr,
                                           (unsigned)F->box.n, theta, sh, c->d_stats, c->stream));
    HIP_TRY(hipMemcpyAsync(F->f[1 - F->cur], F->f[F->cur], level_bytes(c, F),
                           hipMemcpyDeviceToDevice, c->stream));
  } else {
    HIP_TRY(uclbs::launch_coarse_to_fine27(C.f[1 - C.cur], C.f[C.cur], cm, F->f[F->cur], fm,
                                           F->ghosts, F->nghost, theta, sh, c->d_stats, c->stream));
  }
  return UCLBS_OK;
}

int uclbs_interlevel(void* ctx, int lvl, double theta) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  return interlevel_fill(c, lvl, theta, false);
}

int uclbs_init_from_coarse(void* ctx, int lvl) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  int rc = interlevel_fill(c, lvl, 1.0, true);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UCLBS_OK;
}

int uclbs_fine_to_coarse(void* ctx, int lvl) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* F;
  int rc = check_level(c, lvl, &F);
  if (rc) return rc;
  if (!F->refined) return set_err(UCLBS_E_ARG, "level %d is not a refined level", lvl);
  if (F->stale)
    return set_err(UCLBS_E_STATE, "level %d: its coarser level was regridded (regrid or define it "
                                  "again)", lvl);
  if ((rc = set_device(c))) return rc;
  Level& C = c->levels[lvl - 1];
  if (F->multibox)
    return hip_rc(uclbs::launch_copy_pairs27(F->f[F->cur], (unsigned)F->box.n, C.f[C.cur],
                                             (unsigned)C.box.n, F->d_inject, F->ninject,
                                             c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, c->stream));
  uclbs::Range3 cover{};
  for (int d = 0; d < 3; ++d) {
    cover.lo[d] = (F->lo[d] + 1) >> 1;  // coarse nodes whose fine image 2C is valid
    cover.hi[d] = F->hi[d] >> 1;
  }
  if (c->cfg.lattice == UCLBS_LATTICE_D3Q39)
    return hip_rc(uclbs::launch_fine_to_coarse39(F->f[F->cur], F->g[F->cur], F->lm, level_map(*F),
                                                 C.f[C.cur], C.g[C.cur], C.lm, level_map(C), cover,
                                                 c->cfg.precision == UCLBS_PREC_FP32_SHIFTED,
                                                 c->stream));
  HIP_TRY(uclbs::launch_fine_to_coarse27(F->f[F->cur], level_map(*F), C.f[C.cur], level_map(C), cover,
                                         c->cfg.precision == UCLBS_PREC_FP32_SHIFTED, c->stream));
  return UCLBS_OK;
}

int uclbs_regrid_level(void* ctx, int lvl, int nbox, const int32_t* lo, const int32_t* hi) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* F;
  int rc = check_level(c, lvl, &F);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "regrid: D3Q27 levels (levels of many boxes are D3Q27)");
  if (!F->refined) return set_err(UCLBS_E_ARG, "regrid: level %d is not a refined level", lvl);
  if ((rc = set_device(c))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  // The old level steps aside; the new one is defined in its place (its
  // interpolation and injection tables against the coarser level as it is
  // now) and filled as AMReX's remake fills a regridded level
  // (amr_processes.f95:670-1080): interpolated from the coarser level, then
  // the old level's valid nodes copied wherever the new level has a node at
  // the same position.  A stale level (its coarser level regridded first)
  // regrids the same way: its own data is still good.
  Level old = c->levels[lvl];
  c->levels[lvl] = Level();
  rc = uclbs_define_level_boxes(ctx, lvl, nbox, lo, hi);
  if (!rc) rc = uclbs_init_from_coarse(ctx, lvl);
  if (rc) {
    free_level(c->levels[lvl]);
    c->levels[lvl] = old;
    return rc;
  }
  Level& N = c->levels[lvl];
  std::vector<uint3> pairs;
  for (size_t b = 0; b < N.hbox.size(); ++b) {
    const uclbs::BoxDesc& bx = N.hbox[b];
    const unsigned nz = bx.n / (bx.nx * bx.ny);
    for (unsigned z = 0; z < nz; ++z)
      for (unsigned y = 0; y < bx.ny; ++y)
        for (unsigned x = 0; x < bx.nx; ++x) {
          const int G[3] = {N.blo[3 * b] - 1 + (int)x, N.blo[3 * b + 1] - 1 + (int)y,
                            N.blo[3 * b + 2] - 1 + (int)z};
          const int64_t src = valid_index(old, G);
          if (src >= 0)
            pairs.push_back(make_uint3(bx.off + x + bx.nx * (y + bx.ny * z), (unsigned)src,
                                       uclbs::kAllPops27));
        }
  }
  // {dst (new level), src (old level), populations}, into both buffers
  uint3* dp = nullptr;
  hipError_t e = hipSuccess;
  if (!pairs.empty()) {
    e = hipMalloc(&dp, pairs.size() * sizeof(uint3));
    if (e == hipSuccess)
      e = hipMemcpy(dp, pairs.data(), pairs.size() * sizeof(uint3), hipMemcpyHostToDevice);
    const int sh = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
      e = uclbs::launch_copy_pairs27(old.f[old.cur], (unsigned)old.box.n,
                                     N.f[k == 0 ? N.cur : 1 - N.cur], (unsigned)N.box.n, dp,
                                     (unsigned)pairs.size(), sh, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (dp) (void)hipFree(dp);
  N.step = old.step;
  N.relink = old.state != nullptr;
  c->regrid_copied = (int64_t)pairs.size();
  free_level(old);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "regrid: %s", hipGetErrorString(e));
  // a finer level's interpolation and injection tables point into the old level
  for (int l = lvl + 1; l < (int)c->levels.size() && c->levels[l].defined; ++l) c->levels[l].stale = true;
  invalidate_graph(c);
  return UCLBS_OK;
}

int uclbs_regrid_copied(void* ctx, int64_t* n) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !n) return set_err(UCLBS_E_ARG, "null argument");
  *n = c->regrid_copied;
  return UCLBS_OK;
}

int uclbs_clear_level(void* ctx, int lvl) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (lvl == 0) return set_err(UCLBS_E_ARG, "clear_level: level 0 stays (destroy the context)");
  if ((rc = set_device(c))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int l = (int)c->levels.size() - 1; l >= lvl; --l) free_level(c->levels[l]);
  c->levels.resize(lvl);
  invalidate_graph(c);
  return UCLBS_OK;
}

int uclbs_tag_ggm(void* ctx, int lvl, int quantity, int scheme, double limit, double* ggm,
                  uint8_t* tags, int64_t* ntagged) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27)
    return set_err(UCLBS_E_UNSUPPORTED, "tag_ggm: D3Q27 levels");
  if (L->multibox && !L->refined)
    return set_err(UCLBS_E_UNSUPPORTED, "tag_ggm: a level 0 of many boxes (periodic) is not tagged");
  if (quantity != 1 && quantity != 2 && quantity != 3 && quantity != 4 && quantity != 7)
    return set_err(UCLBS_E_ARG, "tag_ggm: quantity %d (1 density, 2 u, 3 v, 4 w, 7 |u|)", quantity);
  if (scheme < 0 || scheme > 2)
    return set_err(UCLBS_E_ARG, "tag_ggm: scheme %d (0 central, 1 upwind, 2 downwind)", scheme);
  if ((rc = set_device(c))) return rc;
  // a refined level's ghost shell holds last step's values; the stencils
  // reach into it, so it is filled from the coarser level's current state
  // first (the level's next step fills it again before stepping)
  if (L->refined && (rc = interlevel_fill(c, lvl, 1.0, false, true))) return rc;
  const int64_t n = L->box.n;
  const int shifted = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  double* work = nullptr;     // one box: 6 n doubles; many: 4 n stored + 6 nd dense + n ggm
  uint8_t* dtag = nullptr;    // tags of the stored nodes
  uint8_t* dense8 = nullptr;  // many boxes: filled and dense tags
  int32_t* dstate = nullptr;  // many boxes with a link table: dense link codes
  int* dblo = nullptr;
  double* ggm_out = nullptr;
  hipError_t e = hipSuccess;
  if (!L->multibox) {
    int vlo[3], vhi[3];
    for (int d = 0; d < 3; ++d) {
      vlo[d] = L->lo[d] - L->origin[d];
      vhi[d] = L->hi[d] - L->origin[d];
    }
    e = hipMalloc(&work, (size_t)n * 6 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dtag, (size_t)n);
    if (e == hipSuccess)
      e = uclbs::launch_d3q27_ggm(L->box, L->f[L->cur], shifted, L->state, quantity, scheme,
                                  limit, vlo, vhi, work, work + 5 * n, dtag, c->stream);
    ggm_out = work + 5 * n;
  } else {
    // the bounding box of the stored nodes (ghost shells included)
    const size_t nbox = L->hbox.size();
    int dlo[3], dhi[3];
    unsigned maxn = 0;
    for (int d = 0; d < 3; ++d) {
      dlo[d] = L->blo[d] - 1;
      dhi[d] = L->bhi[d] + 1;
    }
    for (size_t b = 0; b < nbox; ++b) {
      maxn = std::max(maxn, L->hbox[b].n);
      for (int d = 0; d < 3; ++d) {
        dlo[d] = std::min(dlo[d], L->blo[3 * b + d] - 1);
        dhi[d] = std::max(dhi[d], L->bhi[3 * b + d] + 1);
      }
    }
    uclbs::Box db{};
    db.nx = dhi[0] - dlo[0] + 1;
    db.ny = dhi[1] - dlo[1] + 1;
    db.nz = dhi[2] - dlo[2] + 1;
    db.n = (int64_t)db.nx * db.ny * db.nz;
    const size_t nd = (size_t)db.n;
    e = hipMalloc(&work, ((size_t)n * 5 + nd * 6) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dtag, (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&dense8, 2 * nd);
    if (e == hipSuccess && L->state) e = hipMalloc(&dstate, 27 * nd * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&dblo, 3 * nbox * sizeof(int));
    if (e == hipSuccess)
      e = hipMemcpyAsync(dblo, L->blo.data(), 3 * nbox * sizeof(int), hipMemcpyHostToDevice,
                         c->stream);
    ggm_out = work + 4 * (size_t)n + 6 * nd;
    if (e == hipSuccess)
      e = uclbs::launch_d3q27_ggm_boxes(L->f[L->cur], n, shifted, L->state, L->d_boxes, dblo,
                                        (unsigned)nbox, maxn, dlo, db, quantity, scheme, limit,
                                        work, work + 4 * (size_t)n, dense8, dense8 + nd, dstate,
                                        ggm_out, dtag, c->stream);
  }
  if (e == hipSuccess && ggm)
    e = hipMemcpyAsync(ggm, ggm_out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream);
  std::vector<uint8_t> ht;
  if (e == hipSuccess && (tags || ntagged)) {
    ht.resize((size_t)n);
    e = hipMemcpyAsync(ht.data(), dtag, (size_t)n, hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (void* p : {(void*)work, (void*)dtag, (void*)dense8, (void*)dstate, (void*)dblo})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "tag_ggm: %s", hipGetErrorString(e));
  if (tags) std::memcpy(tags, ht.data(), (size_t)n);
  if (ntagged) {
    int64_t t = 0;
    for (uint8_t v : ht) t += v;
    *ntagged = t;
  }
  return UCLBS_OK;
}

int uclbs_cluster_tags(const uint8_t* tags, const int32_t lo[3], const int32_t hi[3],
                       double efficiency, int blocking, int buffer, int max_size, int max_boxes,
                       int32_t* box_lo, int32_t* box_hi, int* nbox) {
  if (!tags || !lo || !hi || !nbox) return set_err(UCLBS_E_ARG, "null argument");
  std::vector<uclbs::ClusterBox> out;
  const char* msg = uclbs::cluster_tags(tags, lo, hi, efficiency, blocking, buffer, max_size, out);
  if (msg) return set_err(UCLBS_E_ARG, "cluster_tags: %s", msg);
  *nbox = (int)out.size();
  if ((int)out.size() > max_boxes)
    return set_err(UCLBS_E_ARG, "cluster_tags: %d boxes, room for %d", (int)out.size(), max_boxes);
  if (!out.empty() && (!box_lo || !box_hi)) return set_err(UCLBS_E_ARG, "null box arrays");
  for (size_t b = 0; b < out.size(); ++b)
    for (int d = 0; d < 3; ++d) {
      box_lo[3 * b + d] = out[b].lo[d];
      box_hi[3 * b + d] = out[b].hi[d];
    }
  return UCLBS_OK;
}

}  // extern "C"

// One step of level lvl and, recursively, two substeps of each finer level:
// ghosts from the coarser level at its previous state (theta 0) and half way
// (theta 1/2), then the finer level injected back -- the phase order of
// loop.f95:104-139 under the level schedule of valid_levels.f95:21-59.
static int advance_rec(Ctx* c, int lvl) {
  int rc = launch_steps(c, &c->levels[lvl], 1);
  if (rc) return rc;
  const int f = lvl + 1;
  if (f < (int)c->levels.size() && c->levels[f].defined && c->levels[f].refined) {
    for (double theta : {0.0, 0.5}) {
      if ((rc = interlevel_fill(c, f, theta, false))) return rc;
      if ((rc = advance_rec(c, f))) return rc;
    }
    rc = uclbs_fine_to_coarse(c, f);
  }
  return rc;
}

// what a captured advance depends on: buffer and queue parities of every
// level, the kernel variants and the level structure
static std::vector<int> graph_signature(const Ctx* c) {
  std::vector<int> sig = {(int)c->structure, (int)c->keep_alpha, c->cfg.alpha_mode,
                          (int)c->levels.size()};
  // addresses of the boundary tables the captured kernels read
  for (const void* p : {(const void*)c->d_feq_in, (const void*)c->d_feq_out, (const void*)c->d_feq_fs,
                        (const void*)c->d_bc39}) {
    sig.push_back((int)(uintptr_t)p);
    sig.push_back((int)((uintptr_t)p >> 32));
  }
  for (const Level& L : c->levels) {
    sig.push_back(L.defined);
    sig.push_back(L.cur);
    sig.push_back(L.far_parity);
    sig.push_back(L.tile_nodes);
    sig.push_back((int)(uintptr_t)L.near_list);
    sig.push_back((int)((uintptr_t)L.near_list >> 32));
    sig.push_back((int)L.n_near);
    for (int f = 0; f < 6; ++f) sig.push_back(L.face_code[f]);
  }
  return sig;
}

static int levels_in_use(const Ctx* c) { return finest_level(c) + 1; }

extern "C" {

int uclbs_advance(void* ctx, int ncoarse, int use_graph) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (ncoarse < 0) return set_err(UCLBS_E_ARG, "ncoarse < 0");
  if (c->levels.empty() || !c->levels[0].defined) return set_err(UCLBS_E_STATE, "level 0 undefined");
  int rc = set_device(c);
  if (rc) return rc;
  const int nl = levels_in_use(c);
  const bool q39 = c->cfg.lattice == UCLBS_LATTICE_D3Q39;
  // the D3Q39 start-up ramps read the finest level's step counter, which a
  // replayed graph would repeat: graphs only once it is past the start-up
  const bool ramp = q39 && c->levels[nl - 1].step < c->levels[nl - 1].gas39.startup;
  const int pairs = (use_graph && !ramp) ? ncoarse / 2 : 0;
  if (pairs > 0) {
    // host-side tables are built before the capture
    if (q39)
      for (int l = 0; l < nl; ++l) {
        Level& L = c->levels[l];
        bool faces = false;
        for (int f = 0; f < 6; ++f) faces = faces || (!L.refined && L.face_code[f] != uclbs::kFacePeriodic);
        if ((faces || L.table39) && (rc = build_near39(c, &L))) return rc;
      }
    const std::vector<int> sig = graph_signature(c);
    if (!c->graph || c->graph_sig != sig) {
      if (c->graph) (void)hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
      // capture two coarse steps: every level's buffer and queue parities
      // return to where they were, so the graph replays as it stands
      const int64_t done0 = c->steps_done;
      std::vector<int64_t> steps0;
      std::vector<int> cur0, par0;
      for (const Level& L : c->levels) {
        steps0.push_back(L.step);
        cur0.push_back(L.cur);
        par0.push_back(L.far_parity);
      }
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = advance_rec(c, 0);
      if (!rc) rc = advance_rec(c, 0);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      bool parities_return = true;
      for (size_t l = 0; l < cur0.size(); ++l)
        parities_return = parities_return && c->levels[l].cur == cur0[l] &&
                          c->levels[l].far_parity == par0[l];
      // the captured work has not run yet: counters and buffer parities are
      // those of the start (a capture that failed part-way may have flipped
      // the parities of launches that were never recorded)
      c->steps_done = done0;
      for (size_t l = 0; l < steps0.size(); ++l) {
        c->levels[l].step = steps0[l];
        c->levels[l].cur = cur0[l];
        c->levels[l].far_parity = par0[l];
      }
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return set_err(UCLBS_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
      hipGraphExec_t x = nullptr;
      const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess)
        return set_err(UCLBS_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
      if (!parities_return) {
        (void)hipGraphExecDestroy(x);
        return set_err(UCLBS_E_STATE, "captured advance does not return to its start state");
      }
      c->graph = x;
      c->graph_sig = sig;
    }
    for (int k = 0; k < pairs; ++k) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    for (int l = 0; l < nl; ++l) {
      c->steps_done += (int64_t)pairs * (2ll << l);  // level l steps 2^l times per coarse step
      c->levels[l].step += (int64_t)pairs * (2ll << l);
    }
  }
  for (int k = 2 * pairs; k < ncoarse; ++k)
    if ((rc = advance_rec(c, 0))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UCLBS_OK;
}

int uclbs_autotune(void* ctx, int lvl, int* tile_nodes) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (!L->multibox) return set_err(UCLBS_E_UNSUPPORTED, "autotune picks the tile shape of a level of many boxes");
  if ((rc = set_device(c))) return rc;
  // keep the state: the trial steps run on a copy of it
  const size_t bytes = level_bytes(c, L);
  const size_t sbytes = sizeof(unsigned long long) * uclbs::kStatSlots * uclbs::kStatFields;
  void* save = nullptr;
  unsigned long long* ssave = nullptr;
  HIP_TRY(hipMalloc(&save, bytes));
  hipError_t e = hipMalloc(&ssave, sbytes);
  if (e == hipSuccess) e = hipMemcpyAsync(save, L->f[L->cur], bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ssave, c->d_stats, sbytes, hipMemcpyDeviceToDevice, c->stream);
  const int cur0 = L->cur, par0 = L->far_parity;
  const int64_t done0 = c->steps_done, step0 = L->step;
  const int shapes[3] = {64, 128, 256};
  int best = L->tile_nodes;
  float best_ms = 1e30f;
  for (int t = 0; t < 3 && e == hipSuccess && !rc; ++t) {
    e = build_tiles(*L, shapes[t]);
    // one untimed and three timed steps per shape (an even total, so the
    // buffer and queue parities end where they began)
    if (e == hipSuccess) rc = launch_steps(c, L, 1);
    if (!rc && e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (!rc && e == hipSuccess) rc = launch_steps(c, L, 3);
    if (!rc && e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(c->ev1);
    float ms = 0.0f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (!rc && e == hipSuccess && ms < best_ms) {
      best_ms = ms;
      best = shapes[t];
    }
  }
  L->cur = cur0;
  L->far_parity = par0;
  c->steps_done = done0;
  L->step = step0;
  if (e == hipSuccess) e = hipMemcpyAsync(L->f[L->cur], save, bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(c->d_stats, ssave, sbytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = build_tiles(*L, best);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(save);
  if (ssave) (void)hipFree(ssave);
  if (rc) return rc;
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "autotune failed: %s", hipGetErrorString(e));
  if (tile_nodes) *tile_nodes = best;
  return UCLBS_OK;
}

int uclbs_set_timestep(void* ctx, int lvl, int64_t step) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (step < 0) return set_err(UCLBS_E_ARG, "negative timestep");
  L->step = step;
  return UCLBS_OK;
}

// ---- learned alpha ------------------------------------------------------------

static int mlp_alloc(Ctx* c) {
  if (c->mlp_p) return UCLBS_OK;
  const size_t pb = uclbs::kMlpParams * sizeof(float);
  HIP_TRY(hipMalloc(&c->mlp_p, pb));
  HIP_TRY(hipMalloc(&c->mlp_g, pb));
  HIP_TRY(hipMalloc(&c->mlp_m, pb));
  HIP_TRY(hipMalloc(&c->mlp_v, pb));
  HIP_TRY(hipMalloc(&c->mlp_wbf, (64 * 16 + 2 * 64 * 64) * sizeof(short)));
  HIP_TRY(hipMalloc(&c->d_scalar, 4 * sizeof(double)));
  HIP_TRY(hipMemset(c->mlp_m, 0, pb));
  HIP_TRY(hipMemset(c->mlp_v, 0, pb));
  c->mlp_t = 0;
  return UCLBS_OK;
}

static int mlp_work(Ctx* c, int B) {
  if (c->work_batch >= B) return UCLBS_OK;
  if (c->work_buf) (void)hipFree(c->work_buf);
  if (c->batch_x) (void)hipFree(c->batch_x);
  if (c->batch_idx) (void)hipFree(c->batch_idx);
  c->work_buf = nullptr;
  c->batch_x = nullptr;
  c->batch_idx = nullptr;
  c->work_batch = 0;
  const size_t per = (size_t)B * 64;
  const size_t npart = (size_t)((B + 255) / 256) * (64 * 64 + 64);
  HIP_TRY(hipMalloc(&c->work_buf, (9 * per + 3 * (size_t)B + npart) * sizeof(float)));
  float* q = c->work_buf;
  uclbs::MlpWork& w = c->work;
  w.Z1 = q; q += per; w.A1 = q; q += per; w.Z2 = q; q += per; w.A2 = q; q += per;
  w.Z3 = q; q += per; w.A3 = q; q += per; w.D1 = q; q += per; w.D2 = q; q += per;
  w.D3 = q; q += per; w.Z4 = q; q += B; w.A4 = q; q += B; w.D4 = q; q += B;
  w.part = q;
  HIP_TRY(hipMalloc(&c->batch_x, (size_t)B * 12 * sizeof(float)));
  HIP_TRY(hipMalloc(&c->batch_idx, (size_t)B * sizeof(unsigned)));
  c->work_batch = B;
  return UCLBS_OK;
}

static int mlp_publish(Ctx* c) {
  HIP_TRY(uclbs::launch_mlp_pack(c->mlp_p, c->mlp_wbf, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->mlp_ready = true;
  return UCLBS_OK;
}

int uclbs_mlp_init(void* ctx, uint64_t seed) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  int rc = set_device(c);
  if (rc || (rc = mlp_alloc(c))) return rc;
  // uniform(-sqrt(6/fan_in), +sqrt(6/fan_in)) weights from Philox4x32-10, zero biases
  std::vector<float> p(uclbs::kMlpParams, 0.0f);
  const int fan[4] = {12, 64, 64, 64}, outs[4] = {64, 64, 64, 1};
  size_t k = 0;
  for (int l = 0; l < 4; ++l) {
    const double lim = std::sqrt(6.0 / fan[l]);
    for (int t = 0; t < fan[l] * outs[l]; ++t, ++k)
      p[k] = (float)((2.0 * uclbs::philox_uniform(k, 77u, seed) - 1.0) * lim);
    k += outs[l];
  }
  HIP_TRY(hipMemcpy(c->mlp_p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(c->mlp_m, 0, p.size() * sizeof(float)));
  HIP_TRY(hipMemset(c->mlp_v, 0, p.size() * sizeof(float)));
  c->mlp_t = 0;
  c->mlp_ymean = 0.0f;
  c->mlp_ystd = 1.0f;
  return mlp_publish(c);
}

int uclbs_mlp_set_params(void* ctx, const float* params, float y_mean, float y_std) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !params) return set_err(UCLBS_E_ARG, "null argument");
  if (!(y_std > 0.0f)) return set_err(UCLBS_E_ARG, "y_std must be > 0");
  int rc = set_device(c);
  if (rc || (rc = mlp_alloc(c))) return rc;
  HIP_TRY(hipMemcpy(c->mlp_p, params, uclbs::kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  c->mlp_ymean = y_mean;
  c->mlp_ystd = y_std;
  return mlp_publish(c);
}

int uclbs_mlp_get_params(void* ctx, float* params, float* y_mean, float* y_std) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !params || !y_mean || !y_std) return set_err(UCLBS_E_ARG, "null argument");
  if (!c->mlp_p) return set_err(UCLBS_E_STATE, "no network");
  int rc = set_device(c);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(params, c->mlp_p, uclbs::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  *y_mean = c->mlp_ymean;
  *y_std = c->mlp_ystd;
  return UCLBS_OK;
}

int uclbs_mlp_loss_grad(void* ctx, const float* X, const float* y, int B, float* grad,
                        double* loss) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !X || !y || !grad || !loss || B <= 0) return set_err(UCLBS_E_ARG, "bad argument");
  if (!c->mlp_p) return set_err(UCLBS_E_STATE, "no network (uclbs_mlp_init)");
  int rc = set_device(c);
  if (rc || (rc = mlp_work(c, B))) return rc;
  float *dX = nullptr, *dy = nullptr;
  HIP_TRY(hipMalloc(&dX, (size_t)B * 12 * sizeof(float)));
  hipError_t e = hipMalloc(&dy, (size_t)B * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dX, X, (size_t)B * 12 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dy, y, (size_t)B * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = uclbs::launch_mlp_loss_grad(c->mlp_p, dX, dy, nullptr, B, c->work, c->mlp_g, c->d_scalar,
                                    c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpy(grad, c->mlp_g, uclbs::kMlpParams * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(loss, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dX);
  if (dy) (void)hipFree(dy);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "mlp_loss_grad: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_mlp_forward(void* ctx, const float* X, int B, int use_mfma, float* yhat) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !X || !yhat || B <= 0) return set_err(UCLBS_E_ARG, "bad argument");
  if (!c->mlp_ready) return set_err(UCLBS_E_STATE, "no network");
  int rc = set_device(c);
  if (rc || (rc = mlp_work(c, B))) return rc;
  float *dX = nullptr, *dy = nullptr;
  HIP_TRY(hipMalloc(&dX, (size_t)B * 12 * sizeof(float)));
  hipError_t e = hipMalloc(&dy, (size_t)B * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dX, X, (size_t)B * 12 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = use_mfma ? uclbs::launch_mlp_infer(dX, B, c->mlp_wbf, c->mlp_p, dy, c->stream)
                 : uclbs::launch_mlp_forward_fp32(c->mlp_p, dX, B, c->work, dy, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(yhat, dy, (size_t)B * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dX);
  if (dy) (void)hipFree(dy);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "mlp_forward: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

int uclbs_mlp_forward_noalloc(void* ctx, const float* X, int B, float* yhat) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !X || !yhat || B <= 0) return set_err(UCLBS_E_ARG, "bad argument");
  if (B > kNoAllocRows)
    return set_err(UCLBS_E_ARG, "mlp_forward_noalloc: %d rows exceed the %d-row staging area", B,
                   kNoAllocRows);
  if (!c->mlp_ready) return set_err(UCLBS_E_STATE, "no network");
  if (!c->scratch) return set_err(UCLBS_E_STATE, "no staging area on this device");
  int rc = set_device(c);
  if (rc) return rc;
  // features in rows [0, B) of the first part, outputs in the second; rows past
  // B (left by an earlier, larger call) are never read
  float* dX = c->scratch;
  float* dy = c->scratch + (size_t)kNoAllocRows * 12;
  HIP_TRY(hipMemcpyAsync(dX, X, (size_t)B * 12 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(uclbs::launch_mlp_infer(dX, B, c->mlp_wbf, c->mlp_p, dy, c->stream));
  HIP_TRY(hipMemcpyAsync(yhat, dy, (size_t)B * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UCLBS_OK;
}

int uclbs_mlp_reserve(void* ctx, int64_t capacity) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || capacity <= 0 || capacity > INT32_MAX) return set_err(UCLBS_E_ARG, "bad capacity");
  int rc = set_device(c);
  if (rc) return rc;
  uclbs::MlpSamples& m = c->samples;
  for (void* p : {(void*)m.X, (void*)m.y, (void*)m.node_of, (void*)m.m_of, (void*)m.count,
                  (void*)m.ballots, (void*)m.blk})
    if (p) (void)hipFree(p);
  m = uclbs::MlpSamples{};
  HIP_TRY(hipMalloc(&m.X, (size_t)capacity * 12 * sizeof(float)));
  HIP_TRY(hipMalloc(&m.y, (size_t)capacity * sizeof(float)));
  HIP_TRY(hipMalloc(&m.node_of, (size_t)capacity * sizeof(unsigned)));
  HIP_TRY(hipMalloc(&m.m_of, (size_t)capacity * sizeof(float)));
  HIP_TRY(hipMalloc(&m.count, sizeof(unsigned long long)));
  HIP_TRY(hipMemset(m.count, 0, sizeof(unsigned long long)));
  m.cap = capacity;
  c->samples_normalised = false;
  return UCLBS_OK;
}

static int sample_count(Ctx* c, unsigned long long* n) {
  HIP_TRY(hipMemcpyAsync(n, c->samples.count, sizeof(*n), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (*n > (unsigned long long)c->samples.cap) *n = (unsigned long long)c->samples.cap;
  return UCLBS_OK;
}

int uclbs_mlp_collect(void* ctx, int lvl, int stride, int64_t* n_samples) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int rc = check_level(c, lvl, &L);
  if (rc) return rc;
  if (!n_samples || stride < 1) return set_err(UCLBS_E_ARG, "bad argument");
  if (!c->samples.cap) return set_err(UCLBS_E_STATE, "no sample buffer (uclbs_mlp_reserve)");
  if (c->cfg.lattice != UCLBS_LATTICE_D3Q27 || c->cfg.precision != UCLBS_PREC_FP32_SHIFTED ||
      c->cfg.alpha_mode != UCLBS_ALPHA_NEWTON)
    return set_err(UCLBS_E_UNSUPPORTED, "samples come from the plain D3Q27 FP32-shifted solver");
  if (c->samples_normalised) return set_err(UCLBS_E_STATE, "samples already used for training");
  if ((rc = set_device(c))) return rc;
  // ordering buffers for the level's launch (one entry per 256 nodes)
  uclbs::MlpSamples& sm = c->samples;
  const size_t nb = (size_t)((L->box.n + 255) / 256);
  if (sm.blk_n < nb) {
    if (sm.ballots) (void)hipFree(sm.ballots);
    if (sm.blk) (void)hipFree(sm.blk);
    sm.ballots = sm.blk = nullptr;
    sm.blk_n = 0;
    HIP_TRY(hipMalloc(&sm.ballots, nb * 4 * sizeof(unsigned long long)));
    HIP_TRY(hipMalloc(&sm.blk, nb * sizeof(unsigned long long)));
    sm.blk_n = nb;
  }
  unsigned long long first = 0, last = 0;
  if ((rc = sample_count(c, &first))) return rc;
  HIP_TRY(uclbs::launch_mlp_collect(L->box, static_cast<const float*>(L->f[L->cur]),
                                    L->has_links ? L->flags : nullptr, stride,
                                    (int)(c->steps_done % stride), c->samples, c->stream));
  // the plain step produces the targets
  const bool keep = c->keep_alpha;
  c->keep_alpha = true;
  rc = launch_steps(c, L, 1);
  c->keep_alpha = keep;
  if (rc) return rc;
  if ((rc = sample_count(c, &last))) return rc;
  HIP_TRY(uclbs::launch_mlp_targets(L->alpha, c->samples, first, last, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *n_samples = (int64_t)last;
  return UCLBS_OK;
}

int uclbs_mlp_download_samples(void* ctx, float* X, float* y, int64_t capacity, int64_t* n) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !n) return set_err(UCLBS_E_ARG, "null argument");
  if (!c->samples.cap) return set_err(UCLBS_E_STATE, "no sample buffer");
  int rc = set_device(c);
  if (rc) return rc;
  unsigned long long cnt = 0;
  if ((rc = sample_count(c, &cnt))) return rc;
  *n = (int64_t)cnt;
  const size_t k = (size_t)std::min<int64_t>((int64_t)cnt, capacity);
  if (X && k) HIP_TRY(hipMemcpy(X, c->samples.X, k * 12 * sizeof(float), hipMemcpyDeviceToHost));
  if (y && k) HIP_TRY(hipMemcpy(y, c->samples.y, k * sizeof(float), hipMemcpyDeviceToHost));
  return UCLBS_OK;
}

int uclbs_mlp_download_sample_nodes(void* ctx, uint32_t* node, int64_t capacity, int64_t* n) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !n) return set_err(UCLBS_E_ARG, "null argument");
  if (!c->samples.cap) return set_err(UCLBS_E_STATE, "no sample buffer");
  int rc = set_device(c);
  if (rc) return rc;
  unsigned long long cnt = 0;
  if ((rc = sample_count(c, &cnt))) return rc;
  *n = (int64_t)cnt;
  const size_t k = (size_t)std::min<int64_t>((int64_t)cnt, capacity);
  if (node && k)
    HIP_TRY(hipMemcpy(node, c->samples.node_of, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return UCLBS_OK;
}

int uclbs_mlp_train(void* ctx, int64_t steps, int batch, double lr, uint64_t seed,
                    double* loss_first, double* loss_last) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || steps <= 0 || batch <= 0 || batch > (1 << 20) || !(lr > 0.0))
    return set_err(UCLBS_E_ARG, "bad training arguments");
  if (!c->mlp_p) return set_err(UCLBS_E_STATE, "no network (uclbs_mlp_init)");
  int rc = set_device(c);
  if (rc || (rc = mlp_work(c, batch))) return rc;
  unsigned long long n = 0;
  if ((rc = sample_count(c, &n))) return rc;
  if (n < 2) return set_err(UCLBS_E_STATE, "no training samples (uclbs_mlp_collect)");
  if (!c->samples_normalised) {
    HIP_TRY(uclbs::launch_mlp_normalise(c->samples.y, n, c->d_scalar + 2, c->stream));
    double ms[2];
    HIP_TRY(hipMemcpyAsync(ms, c->d_scalar + 2, sizeof(ms), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!std::isfinite(ms[0]) || !std::isfinite(ms[1]))
      return set_err(UCLBS_E_NUMERIC, "mlp_train: non-finite training targets (mean %g, std %g)",
                     ms[0], ms[1]);
    c->mlp_ymean = (float)ms[0];
    c->mlp_ystd = (float)ms[1];
    c->samples_normalised = true;
  }
  double l0 = 0.0, l1 = 0.0;
  for (int64_t t = 0; t < steps; ++t) {
    HIP_TRY(uclbs::launch_mlp_gather(c->samples.X, n, seed, (uint64_t)t, batch, c->batch_x,
                                     c->batch_idx, c->stream));
    HIP_TRY(uclbs::launch_mlp_loss_grad(c->mlp_p, c->batch_x, c->samples.y, c->batch_idx, batch,
                                        c->work, c->mlp_g, c->d_scalar, c->stream));
    // learning rate: linear warm-up over the first 2% of the steps, then a
    // cosine decay to 5% of lr (keeps the late steps from undoing the fit)
    const double warm = std::max<double>(1.0, 0.02 * (double)steps);
    const double prog = (double)t / (double)steps;
    const double lrt = t < warm ? lr * (t + 1) / warm
                                : lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * prog)));
    HIP_TRY(uclbs::launch_mlp_adam(c->mlp_p, c->mlp_g, c->mlp_m, c->mlp_v, ++c->mlp_t, (float)lrt,
                                   c->stream));
    if (t == 0 || t == steps - 1) {
      double l;
      HIP_TRY(hipMemcpyAsync(&l, c->d_scalar, sizeof(l), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      (t == 0 ? l0 : l1) = l;
    }
  }
  if (loss_first) *loss_first = l0;
  if (loss_last) *loss_last = steps == 1 ? l0 : l1;
  if (!std::isfinite(l0) || !std::isfinite(steps == 1 ? l0 : l1))
    return set_err(UCLBS_E_NUMERIC, "mlp_train: non-finite loss (first %g, last %g)", l0, l1);
  return mlp_publish(c);
}


// ---- AMReX patch data on the device -------------------------------------------

namespace {

int fab_args(Ctx* c, int lvl, int field, const void* fab, int ghost, int64_t cstride, Level** L,
             int* ncomp) {
  int rc = check_level(c, lvl, L);
  if (rc) return rc;
  if ((*L)->refined || (*L)->multibox)
    return set_err(UCLBS_E_UNSUPPORTED, "FAB exchange is for one-box levels that are not refined");
  const bool q39 = c->cfg.lattice == UCLBS_LATTICE_D3Q39;
  if (field == UCLBS_FIELD_F) *ncomp = q_of(c);
  else if (q39 && field == UCLBS_FIELD_G) *ncomp = uclbs::Q39;
  else if (q39 && field == UCLBS_FIELD_LM) *ncomp = uclbs::kM39;
  else return set_err(UCLBS_E_ARG, "FAB field %d does not exist on this lattice", field);
  const uclbs::Box& b = (*L)->box;
  const int64_t cells = (int64_t)(b.nx + 2 * ghost) * (b.ny + 2 * ghost) * (b.nz + 2 * ghost);
  if (!fab || ghost < 0 || ghost > 8 || cstride < cells)
    return set_err(UCLBS_E_ARG, "FAB: ghost width %d / component stride %lld do not hold a %dx%dx%d box",
                   ghost, (long long)cstride, b.nx, b.ny, b.nz);
  if ((rc = set_device(c))) return rc;
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, fab) != hipSuccess ||
      (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)) {
    (void)hipGetLastError();
    return set_err(UCLBS_E_ARG, "FAB pointer is not device memory");
  }
  return UCLBS_OK;
}

}  // namespace

int uclbs_import_fab(void* ctx, int lvl, int field, const double* fab, int ghost,
                     int64_t comp_stride) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int ncomp = 0;
  int rc = fab_args(c, lvl, field, fab, ghost, comp_stride, &L, &ncomp);
  if (rc) return rc;
  const bool fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  hipError_t e = hipSuccess;
  if (field == UCLBS_FIELD_LM) {
    e = uclbs::launch_fab_gather(fab, ghost, comp_stride, L->box, ncomp, L->lm, c->stream);
  } else {
    double* tmp = nullptr;
    e = hipMalloc(&tmp, (size_t)L->box.n * ncomp * sizeof(double));
    if (e == hipSuccess)
      e = uclbs::launch_fab_gather(fab, ghost, comp_stride, L->box, ncomp, tmp, c->stream);
    void* cur = field == UCLBS_FIELD_F ? L->f[L->cur] : L->g[L->cur];
    void* other = field == UCLBS_FIELD_F ? L->f[1 - L->cur] : L->g[1 - L->cur];
    if (e == hipSuccess) {
      if (c->cfg.lattice == UCLBS_LATTICE_D3Q27)
        e = uclbs::launch_d3q27_from_fp64(L->box, tmp, cur, fp32, c->stream);
      else
        e = uclbs::launch_d3q39_from_fp64(L->box, tmp, cur, fp32,
                                          c->d_ref39 + (field == UCLBS_FIELD_G ? uclbs::Q39 : 0),
                                          c->stream);
    }
    // the other buffer mirrors the imported one (nodes no kernel writes)
    if (e == hipSuccess)
      e = hipMemcpyAsync(other, cur, level_bytes(c, L) / q_of(c) * ncomp,
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "import_fab: %s", hipGetErrorString(e));
  invalidate_graph(c);
  return UCLBS_OK;
}

int uclbs_export_fab(void* ctx, int lvl, int field, double* fab, int ghost, int64_t comp_stride) {
  Ctx* c = static_cast<Ctx*>(ctx);
  Level* L;
  int ncomp = 0;
  int rc = fab_args(c, lvl, field, fab, ghost, comp_stride, &L, &ncomp);
  if (rc) return rc;
  const bool fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  hipError_t e = hipSuccess;
  if (field == UCLBS_FIELD_LM) {
    e = uclbs::launch_fab_scatter(L->lm, L->box, ncomp, ghost, comp_stride, fab, c->stream);
  } else {
    double* tmp = nullptr;
    e = hipMalloc(&tmp, (size_t)L->box.n * ncomp * sizeof(double));
    const void* cur = field == UCLBS_FIELD_F ? L->f[L->cur] : L->g[L->cur];
    if (e == hipSuccess) {
      if (c->cfg.lattice == UCLBS_LATTICE_D3Q27)
        e = uclbs::launch_d3q27_to_fp64(L->box, cur, tmp, fp32, c->stream);
      else
        e = uclbs::launch_d3q39_to_fp64(L->box, cur, tmp, fp32,
                                        c->d_ref39 + (field == UCLBS_FIELD_G ? uclbs::Q39 : 0),
                                        c->stream);
    }
    if (e == hipSuccess)
      e = uclbs::launch_fab_scatter(tmp, L->box, ncomp, ghost, comp_stride, fab, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(UCLBS_E_HIP, "export_fab: %s", hipGetErrorString(e));
  return UCLBS_OK;
}

// ---- checkpoint / restart ------------------------------------------------------
// File: "UCLBSCKP", version, lattice, precision, level count; per defined
// level a header (level index, node count and box extents, refined / many-box
// flags, step counter, face codes, the D3Q39 gas constants of the level, a
// hash of the link table), then the current f (and g) buffer in storage form
// and the D3Q39 multipliers, then a checksum of that payload.  Restoring
// needs a context set up the same way (levels, faces, gas, links): every
// header field is checked against it before the level is overwritten.

namespace {

constexpr char kCkMagic[8] = {'U', 'C', 'L', 'B', 'S', 'C', 'K', 'P'};
constexpr uint32_t kCkVersion = 1;

struct CkFile {
  char magic[8];
  uint32_t version, lattice, precision, nlev;
};

struct CkLevel {
  int32_t lvl, flags;  // flags: 1 g and multipliers present, 2 link table present
  int64_t n;
  int32_t nx, ny, nz, refined, multibox, shifted;
  int64_t step;
  int32_t face_code[6];
  uint64_t table_hash;
  uint64_t bytes_f, bytes_g, bytes_lm;
  uclbs::Gas39 gas;
};

// FNV-1a over 64-bit words (a tail shorter than a word byte by byte)
uint64_t ck_hash(const void* data, size_t bytes, uint64_t h = 1469598103934665603ull) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  size_t k = 0;
  for (; k + 8 <= bytes; k += 8) {
    uint64_t w;
    std::memcpy(&w, p + k, 8);
    h = (h ^ w) * 1099511628211ull;
  }
  for (; k < bytes; ++k) h = (h ^ p[k]) * 1099511628211ull;
  return h;
}

int ck_table_hash(Ctx* c, Level* L, uint64_t* h) {
  *h = 0;
  if (!L->state) return UCLBS_OK;
  std::vector<int32_t> t((size_t)L->box.n * q_of(c));
  HIP_TRY(hipMemcpyAsync(t.data(), L->state, t.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *h = ck_hash(t.data(), t.size() * sizeof(int32_t));
  return UCLBS_OK;
}

int ck_header(Ctx* c, Level* L, int lvl, CkLevel* h) {
  std::memset(h, 0, sizeof *h);
  const bool q39 = c->cfg.lattice == UCLBS_LATTICE_D3Q39;
  h->lvl = lvl;
  h->flags = (q39 ? 1 : 0) | (L->state ? 2 : 0);
  h->n = L->box.n;
  h->nx = L->box.nx;
  h->ny = L->box.ny;
  h->nz = L->box.nz;
  h->refined = L->refined;
  h->multibox = L->multibox;
  h->shifted = q39 ? c->shifted39 : 0;
  h->step = L->step;
  for (int f = 0; f < 6; ++f) h->face_code[f] = L->face_code[f];
  h->bytes_f = level_bytes(c, L);
  h->bytes_g = q39 ? level_bytes(c, L) : 0;
  h->bytes_lm = q39 ? (uint64_t)L->box.n * uclbs::kM39 * sizeof(double) : 0;
  if (q39) h->gas = L->gas39;
  return ck_table_hash(c, L, &h->table_hash);
}

}  // namespace

int uclbs_checkpoint_write(void* ctx, const char* path) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !path) return set_err(UCLBS_E_ARG, "checkpoint_write: null argument");
  int rc = set_device(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return set_err(UCLBS_E_ARG, "checkpoint_write: cannot open %s", path);
  CkFile fh{};
  std::memcpy(fh.magic, kCkMagic, 8);
  fh.version = kCkVersion;
  fh.lattice = (uint32_t)c->cfg.lattice;
  fh.precision = (uint32_t)c->cfg.precision;
  for (const Level& L : c->levels) fh.nlev += L.defined ? 1 : 0;
  bool ok = std::fwrite(&fh, sizeof fh, 1, fp) == 1;
  std::vector<char> buf;
  for (int lvl = 0; ok && lvl < (int)c->levels.size(); ++lvl) {
    Level* L = &c->levels[lvl];
    if (!L->defined) continue;
    CkLevel h;
    if ((rc = ck_header(c, L, lvl, &h))) { std::fclose(fp); return rc; }
    ok = std::fwrite(&h, sizeof h, 1, fp) == 1;
    uint64_t sum = 1469598103934665603ull;
    const void* src[3] = {L->f[L->cur], h.bytes_g ? L->g[L->cur] : nullptr,
                          h.bytes_lm ? L->lm : nullptr};
    const uint64_t len[3] = {h.bytes_f, h.bytes_g, h.bytes_lm};
    for (int k = 0; ok && k < 3; ++k) {
      if (!len[k]) continue;
      buf.resize(len[k]);
      hipError_t e = hipMemcpyAsync(buf.data(), src[k], len[k], hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        std::fclose(fp);
        return set_err(UCLBS_E_HIP, "checkpoint_write: %s", hipGetErrorString(e));
      }
      sum = ck_hash(buf.data(), len[k], sum);
      ok = std::fwrite(buf.data(), 1, len[k], fp) == len[k];
    }
    ok = ok && std::fwrite(&sum, sizeof sum, 1, fp) == 1;
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) return set_err(UCLBS_E_ARG, "checkpoint_write: write to %s failed", path);
  return UCLBS_OK;
}

int uclbs_checkpoint_read(void* ctx, const char* path) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !path) return set_err(UCLBS_E_ARG, "checkpoint_read: null argument");
  int rc = set_device(c);
  if (rc) return rc;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return set_err(UCLBS_E_ARG, "checkpoint_read: cannot open %s", path);
  auto fail = [&](int code, const char* why) {
    std::fclose(fp);
    return set_err(code, "checkpoint_read %s: %s", path, why);
  };
  CkFile fh{};
  if (std::fread(&fh, sizeof fh, 1, fp) != 1 || std::memcmp(fh.magic, kCkMagic, 8) != 0)
    return fail(UCLBS_E_ARG, "not a checkpoint file");
  if (fh.version != kCkVersion) return fail(UCLBS_E_ARG, "unknown checkpoint version");
  if (fh.lattice != (uint32_t)c->cfg.lattice || fh.precision != (uint32_t)c->cfg.precision)
    return fail(UCLBS_E_STATE, "lattice or storage precision differs from the context");
  uint32_t ndef = 0;
  for (const Level& L : c->levels) ndef += L.defined ? 1 : 0;
  if (fh.nlev != ndef) return fail(UCLBS_E_STATE, "level count differs from the context");
  std::vector<char> buf[3];
  for (uint32_t r = 0; r < fh.nlev; ++r) {
    CkLevel h;
    if (std::fread(&h, sizeof h, 1, fp) != 1) return fail(UCLBS_E_ARG, "truncated file");
    Level* L;
    if (h.lvl < 0 || check_level(c, h.lvl, &L) != UCLBS_OK)
      return fail(UCLBS_E_STATE, "a saved level is not defined in the context");
    CkLevel mine;
    if ((rc = ck_header(c, L, h.lvl, &mine))) { std::fclose(fp); return rc; }
    mine.step = h.step;  // the one field restored, not compared
    if (std::memcmp(&h, &mine, sizeof h) != 0)
      return fail(UCLBS_E_STATE, "level set-up (extent, faces, gas, link table) differs from the context");
    uint64_t sum = 1469598103934665603ull, saved;
    const uint64_t len[3] = {h.bytes_f, h.bytes_g, h.bytes_lm};
    for (int k = 0; k < 3; ++k) {
      buf[k].resize(len[k]);
      if (len[k] && std::fread(buf[k].data(), 1, len[k], fp) != len[k])
        return fail(UCLBS_E_ARG, "truncated file");
      sum = ck_hash(buf[k].data(), len[k], sum);
    }
    if (std::fread(&saved, sizeof saved, 1, fp) != 1) return fail(UCLBS_E_ARG, "truncated file");
    if (saved != sum) return fail(UCLBS_E_ARG, "checksum mismatch (damaged file)");
    void* dst[3] = {L->f[L->cur], len[1] ? L->g[L->cur] : nullptr, len[2] ? L->lm : nullptr};
    void* dst2[2] = {L->f[1 - L->cur], len[1] ? L->g[1 - L->cur] : nullptr};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
      if (len[k]) e = hipMemcpyAsync(dst[k], buf[k].data(), len[k], hipMemcpyHostToDevice, c->stream);
    // the other buffer mirrors the restored one (nodes no kernel writes)
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
      if (len[k]) e = hipMemcpyAsync(dst2[k], dst[k], len[k], hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      std::fclose(fp);
      return set_err(UCLBS_E_HIP, "checkpoint_read: %s", hipGetErrorString(e));
    }
    L->step = h.step;
  }
  std::fclose(fp);
  invalidate_graph(c);
  return UCLBS_OK;
}

}  // extern "C"

// ---- output: Plot3D files and microphones -------------------------------------------

namespace {

// spacing and domain corner of level lvl: AMReX geometry with refinement ratio
// 2 over the level-0 domain (dx = (probhi - problo) / cells, the domain's
// first node at problo); unplaced contexts use problo 0 and spacing 1 at level 0
void level_geom(const Ctx* c, int lvl, double lo[3], double dx[3], int dlo[3]) {
  const Level& L0 = c->levels[0];
  for (int d = 0; d < 3; ++d) {
    const int n0 = L0.hi[d] - L0.lo[d] + 1;
    lo[d] = c->placed ? c->problo[d] : 0.0;
    const double hi = c->placed ? c->probhi[d] : (double)n0;
    dx[d] = (hi - lo[d]) / ((double)n0 * (double)(1 << lvl));
    dlo[d] = L0.lo[d] * (1 << lvl);
  }
}

// one Plot3D block: the valid region of one box of one level
struct P3Block {
  int lvl = 0;
  int dims[3] = {0, 0, 0};
  double start[3] = {0, 0, 0}, dx[3] = {0, 0, 0};
  size_t nv = 0;
  size_t dbl_off = 0;  // byte offsets of its fields and its IBLANK in the packed buffer
  size_t int_off = 0;
};

// Fortran unformatted sequential records (gfortran layout): 4-byte length
// markers before and after each subrecord; a record longer than maxsub bytes
// is split into subrecords whose leading marker is negative when more follow
// and whose trailing marker is negative when one precedes
struct RecWriter {
  std::FILE* fp = nullptr;
  int64_t maxsub = 2147483639;
  int64_t rem = 0, sub = 0, sublen = 0;
  bool first = true, ok = true;
  void marker(int64_t v) {
    const int32_t m = (int32_t)v;
    ok = ok && std::fwrite(&m, 4, 1, fp) == 1;
  }
  void open_sub() {
    sublen = std::min(rem, maxsub);
    sub = sublen;
    marker(rem > sublen ? -sublen : sublen);
  }
  void close_sub() {
    marker(first ? sublen : -sublen);
    first = false;
    rem -= sublen;
  }
  void begin(int64_t total) {
    rem = total;
    first = true;
    open_sub();
  }
  void put(const void* p, size_t n) {
    const char* q = static_cast<const char*>(p);
    while (n && ok) {
      if (sub == 0) {
        close_sub();
        open_sub();
      }
      const size_t k = (size_t)std::min<int64_t>((int64_t)n, sub);
      ok = ok && std::fwrite(q, 1, k, fp) == k;
      q += k;
      n -= k;
      sub -= (int64_t)k;
    }
  }
  void end() { close_sub(); }
};

struct P3Job {
  Ctx* c = nullptr;
  std::string grid, data;
  uclbs_plot3d_header h{};
  std::vector<P3Block> blocks;
};

// the files of plot3d_output.f95:76-305 from the packed host copy
void write_plot3d(P3Job job) {
  Ctx* c = job.c;
  auto fail = [&](const char* what, const std::string& path) {
    c->out_rc = UCLBS_E_ARG;
    c->out_msg = std::string("plot3d: ") + what + " " + path;
  };
  const hipError_t e = hipEventSynchronize(c->out_ev);
  if (e != hipSuccess) {
    c->out_rc = UCLBS_E_HIP;
    c->out_msg = std::string("plot3d: copy to the host failed: ") + hipGetErrorString(e);
    return;
  }
  const int32_t nb = (int32_t)job.blocks.size();
  std::vector<int32_t> dims;
  for (const P3Block& b : job.blocks) dims.insert(dims.end(), b.dims, b.dims + 3);
  const int64_t maxsub = job.h.max_subrecord > 0 ? job.h.max_subrecord : 2147483639;
  std::vector<char> iobuf(1 << 23);
  // grid: block count, block extents, then per block x, y, z (REAL) and IBLANK
  std::FILE* fp = std::fopen(job.grid.c_str(), "wb");
  if (!fp) return fail("cannot create", job.grid);
  std::setvbuf(fp, iobuf.data(), _IOFBF, iobuf.size());
  RecWriter r{fp, maxsub};
  r.begin(4);
  r.put(&nb, 4);
  r.end();
  r.begin((int64_t)dims.size() * 4);
  r.put(dims.data(), dims.size() * 4);
  r.end();
  std::vector<float> tmp;
  for (const P3Block& b : job.blocks) {
    r.begin((int64_t)b.nv * 16);
    for (int ax = 0; ax < 3; ++ax) {
      // start + dx (i - lo) in double, stored as REAL (plot3d_output.f95:270-272)
      tmp.resize(b.nv);
      size_t k = 0;
      for (int cz = 0; cz < b.dims[2]; ++cz)
        for (int cy = 0; cy < b.dims[1]; ++cy)
          for (int cx = 0; cx < b.dims[0]; ++cx) {
            const int i = ax == 0 ? cx : (ax == 1 ? cy : cz);
            tmp[k++] = (float)(b.start[ax] + b.dx[ax] * (double)i);
          }
      r.put(tmp.data(), b.nv * 4);
    }
    r.put(c->out_host + b.int_off, b.nv * 4);
    r.end();
  }
  const bool gok = r.ok && std::fclose(fp) == 0;
  if (!gok) return fail("write failed:", job.grid);
  // solution: block count, extents with the variable count 6 and a 0, then
  // per block the header record and rho, u, v, w, T, alpha (FP64)
  fp = std::fopen(job.data.c_str(), "wb");
  if (!fp) return fail("cannot create", job.data);
  std::setvbuf(fp, iobuf.data(), _IOFBF, iobuf.size());
  r = RecWriter{fp, maxsub};
  r.begin(4);
  r.put(&nb, 4);
  r.end();
  dims.push_back(6);
  dims.push_back(0);
  r.begin((int64_t)dims.size() * 4);
  r.put(dims.data(), dims.size() * 4);
  r.end();
  const uclbs_plot3d_header& h = job.h;
  const double hd1[7] = {h.mach, 0.0, h.reynolds, h.time, h.gamma, 0.0, h.t_ref};
  const double hd2[7] = {h.stag_temperature, h.cv, h.cp, h.gas_r, h.gas_r, h.mach, h.time};
  const int32_t zero = 0, ts = h.timestep;
  for (const P3Block& b : job.blocks) {
    r.begin(120);
    r.put(hd1, 56);
    r.put(&zero, 4);
    r.put(hd2, 56);
    r.put(&ts, 4);
    r.end();
    r.begin((int64_t)b.nv * 48);
    r.put(c->out_host + b.dbl_off, b.nv * 48);
    r.end();
  }
  const bool dok = r.ok && std::fclose(fp) == 0;
  if (!dok) return fail("write failed:", job.data);
}

int join_output(Ctx* c) {
  if (c->out_thread.joinable()) c->out_thread.join();
  const int rc = c->out_rc;
  c->out_rc = 0;
  if (rc) return set_err(rc, "%s", c->out_msg.c_str());
  return UCLBS_OK;
}

void free_mics(Ctx* c) {
  for (MicLevel& ml : c->mic_levels)
    if (ml.d_nodes) (void)hipFree(ml.d_nodes);
  c->mic_levels.clear();
  c->mics.clear();
  if (c->mic_ring) (void)hipFree(c->mic_ring);
  c->mic_ring = nullptr;
  c->mic_stride = 0;
  c->mic_cap = c->mic_count = 0;
  c->mic_structure = -1;
}

// Fortran F12.5 editing: 12 asterisks when the value does not fit
void f12_5(double v, std::string& out) {
  char buf[48];
  if (std::isnan(v)) std::snprintf(buf, sizeof buf, "%12s", "NaN");
  else if (std::isinf(v)) std::snprintf(buf, sizeof buf, "%12s", v > 0 ? "Infinity" : "-Infinity");
  else if (std::snprintf(buf, sizeof buf, "%12.5f", v) > 12)
    std::snprintf(buf, sizeof buf, "************");
  out += buf;
}

void release_output(Ctx* c) {
  if (c->out_thread.joinable()) c->out_thread.join();
  if (c->out_fields) (void)hipFree(c->out_fields);
  if (c->out_dev) (void)hipFree(c->out_dev);
  if (c->out_host) (void)hipHostFree(c->out_host);
  if (c->out_ev) (void)hipEventDestroy(c->out_ev);
  c->out_fields = nullptr;
  c->out_dev = c->out_host = nullptr;
  c->out_ev = nullptr;
  c->out_fields_cap = c->out_cap = 0;
  free_mics(c);
}

}  // namespace

extern "C" {

int uclbs_set_placement(void* ctx, const double problo[3], const double probhi[3]) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !problo || !probhi) return set_err(UCLBS_E_ARG, "null argument");
  for (int d = 0; d < 3; ++d)
    if (!(probhi[d] > problo[d]) || !std::isfinite(problo[d]) || !std::isfinite(probhi[d]))
      return set_err(UCLBS_E_ARG, "placement: probhi must exceed problo on axis %d", d);
  for (int d = 0; d < 3; ++d) {
    c->problo[d] = problo[d];
    c->probhi[d] = probhi[d];
  }
  c->placed = true;
  return UCLBS_OK;
}

int uclbs_output_wait(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  return join_output(c);
}

int uclbs_plot3d_write(void* ctx, const char* dir, const uclbs_plot3d_header* h) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !h) return set_err(UCLBS_E_ARG, "null argument");
  if (h->timestep < 0 || h->timestep > 999999999)
    return set_err(UCLBS_E_ARG, "plot3d: timestep %d does not fit the file names' nine digits",
                   h->timestep);
  if (h->max_level < 0) return set_err(UCLBS_E_ARG, "plot3d: max_level below 0");
  if (h->max_subrecord < 0 || h->max_subrecord > 2147483639 ||
      (h->max_subrecord > 0 && h->max_subrecord < 4))
    return set_err(UCLBS_E_ARG, "plot3d: max_subrecord outside [4, 2147483639] (0: default)");
  Level* L0;
  int rc = check_level(c, 0, &L0);
  if (rc) return rc;
  if ((rc = join_output(c))) return rc;  // the previous files are complete
  if ((rc = set_device(c))) return rc;
  const bool d39 = c->cfg.lattice == UCLBS_LATTICE_D3Q39;
  const int fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  int top = 0;
  while (top < h->max_level && top + 1 < (int)c->levels.size() && c->levels[top + 1].defined) ++top;
  // blocks in level order, a level's boxes in their order (plot3d_output.f95:101-131)
  P3Job job;
  job.c = c;
  job.h = *h;
  size_t dbl = 0, ints = 0, max_n = 0;
  for (int l = 0; l <= top; ++l) {
    const Level& L = c->levels[l];
    max_n = std::max(max_n, (size_t)L.box.n);
    double plo[3], dx[3];
    int dlo[3];
    level_geom(c, l, plo, dx, dlo);
    const int nbox = L.multibox ? (int)L.hbox.size() : 1;
    for (int b = 0; b < nbox; ++b) {
      P3Block k;
      k.lvl = l;
      k.nv = 1;
      for (int d = 0; d < 3; ++d) {
        const int lo = L.multibox ? L.blo[3 * b + d] : L.lo[d];
        const int hi = L.multibox ? L.bhi[3 * b + d] : L.hi[d];
        k.dims[d] = hi - lo + 1;
        k.nv *= (size_t)k.dims[d];
        k.dx[d] = dx[d];
        k.start[d] = plo[d] + ((double)lo - (double)dlo[d]) * dx[d];  // get_real_coords
      }
      k.dbl_off = dbl;
      dbl += k.nv * 48;
      job.blocks.push_back(k);
      ints += k.nv * 4;
    }
  }
  {
    size_t io = dbl;  // the IBLANK arrays follow every block's fields
    for (P3Block& k : job.blocks) {
      k.int_off = io;
      io += k.nv * 4;
    }
  }
  const size_t bytes = dbl + ints;
  if (max_n * 6 > c->out_fields_cap) {
    if (c->out_fields) (void)hipFree(c->out_fields);
    c->out_fields = nullptr;
    c->out_fields_cap = 0;
    HIP_TRY(hipMalloc(&c->out_fields, max_n * 6 * sizeof(double)));
    c->out_fields_cap = max_n * 6;
  }
  if (bytes > c->out_cap) {
    if (c->out_dev) (void)hipFree(c->out_dev);
    if (c->out_host) (void)hipHostFree(c->out_host);
    c->out_dev = c->out_host = nullptr;
    c->out_cap = 0;
    HIP_TRY(hipMalloc(&c->out_dev, bytes));
    HIP_TRY(hipHostMalloc((void**)&c->out_host, bytes, hipHostMallocDefault));
    c->out_cap = bytes;
  }
  if (!c->out_ev) HIP_TRY(hipEventCreateWithFlags(&c->out_ev, hipEventDisableTiming));
  size_t bi = 0;
  for (int l = 0; l <= top; ++l) {
    Level& L = c->levels[l];
    const int64_t n = L.box.n;
    double* of = c->out_fields;
    // rho, u, v, w, T, alpha of every stored node (calculate_flow_variables.f95:
    // 201-221; alpha 2 where the entropic estimate is not kept, its value where
    // the reference does not compute it)
    if (d39) {
      HIP_TRY(uclbs::launch_d3q39_macro(L.box, L.f[L.cur], L.g[L.cur], fp32, c->shifted39,
                                        c->ref39, L.gas39.cv, of, c->stream));
    } else {
      HIP_TRY(uclbs::launch_d3q27_macro(L.box, L.f[L.cur], fp32, of, of + n, of + 2 * n,
                                        of + 3 * n, c->stream));
      HIP_TRY(uclbs::launch_fill(of + 4 * n, n, 1.0 / 3.0, c->stream));
    }
    if (c->keep_alpha)
      HIP_TRY(hipMemcpyAsync(of + 5 * n, L.alpha, (size_t)n * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
    else
      HIP_TRY(uclbs::launch_fill(of + 5 * n, n, 2.0, c->stream));
    const int nbox = L.multibox ? (int)L.hbox.size() : 1;
    for (int b = 0; b < nbox; ++b, ++bi) {
      const P3Block& k = job.blocks[bi];
      uclbs::Box vb{};
      vb.nx = k.dims[0];
      vb.ny = k.dims[1];
      vb.nz = k.dims[2];
      vb.n = (int64_t)k.nv;
      const int64_t base = L.multibox ? (int64_t)L.hbox[b].off : 0;
      const int ghost = L.multibox ? 1 : L.lo[0] - L.origin[0];
      HIP_TRY(uclbs::launch_fab_gather(of + base, ghost, n, vb, 6,
                                       reinterpret_cast<double*>(c->out_dev + k.dbl_off), c->stream));
      HIP_TRY(uclbs::launch_blank_gather(L.state ? L.state + base : nullptr, ghost, vb,
                                         reinterpret_cast<int32_t*>(c->out_dev + k.int_off),
                                         c->stream));
    }
  }
  HIP_TRY(hipMemcpyAsync(c->out_host, c->out_dev, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipEventRecord(c->out_ev, c->stream));
  char num[16];
  std::snprintf(num, sizeof num, "%09d", h->timestep);
  const std::string pre = dir && *dir ? std::string(dir) + "/" : std::string();
  job.grid = pre + "uclbs_grid_" + num + ".x";
  job.data = pre + "uclbs_data_" + num + ".q";
  c->out_rc = 0;
  c->out_msg.clear();
  try {
    c->out_thread = std::thread(write_plot3d, std::move(job));
  } catch (...) {
    (void)hipEventSynchronize(c->out_ev);
    return set_err(UCLBS_E_STATE, "plot3d: cannot start the writer thread");
  }
  return UCLBS_OK;
}

int uclbs_mic_setup(void* ctx, const char* dir, int n, const double* loc, const int32_t* type,
                    const int32_t* level, int32_t* valid) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (n < 0 || n > 99999) return set_err(UCLBS_E_ARG, "mic_setup: %d microphones (0..99999)", n);
  if (n && (!loc || !type || !level)) return set_err(UCLBS_E_ARG, "null argument");
  Level* L0;
  int rc = check_level(c, 0, &L0);
  if (rc) return rc;
  if ((rc = set_device(c))) return rc;
  for (int i = 0; i < n; ++i) {
    if (type[i] < 0 || type[i] > 63) return set_err(UCLBS_E_ARG, "mic %d: type %d outside 0..63", i + 1, type[i]);
    if (level[i] < 0) return set_err(UCLBS_E_ARG, "mic %d: level below 0", i + 1);
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_mics(c);
  int fine = 0;
  while (fine + 1 < (int)c->levels.size() && c->levels[fine + 1].defined) ++fine;
  double glo[3], ghi[3];
  {
    double dx0[3];
    int dlo0[3];
    level_geom(c, 0, glo, dx0, dlo0);
    for (int d = 0; d < 3; ++d)
      ghi[d] = c->placed ? c->probhi[d] : (double)(L0->hi[d] - L0->lo[d] + 1);
  }
  static const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                    {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  const std::string pre = dir && *dir ? std::string(dir) + "/" : std::string();
  std::vector<std::vector<uint32_t>> lists(fine + 1);
  for (int i = 0; i < n; ++i) {
    Mic m;
    m.number = i + 1;
    m.mask = type[i];
    const double* p = loc + 3 * i;
    bool in = true;
    for (int d = 0; d < 3; ++d) in = in && !(p[d] > ghi[d] || p[d] < glo[d]);
    if (!in) {  // mic_setup.f95:45-53
      c->mics.push_back(m);
      continue;
    }
    const int l = std::min(level[i], fine);
    m.level = l;
    const Level& L = c->levels[l];
    double plo[3], dx[3];
    int dlo[3], st[3];
    level_geom(c, l, plo, dx, dlo);
    const double rdx = dx[0];  // the level's rdx (x spacing) on every axis
    double lo[3];
    for (int d = 0; d < 3; ++d) {
      st[d] = dlo[d] + (int)std::floor((p[d] - plo[d]) / rdx);
      lo[d] = plo[d] + ((double)st[d] - (double)dlo[d]) * dx[d];
    }
    bool owned = true;
    for (int d = 0; d < 3; ++d) owned = owned && st[d] >= L.lo[d] && st[d] <= L.hi[d];
    owned = owned && valid_index(L, st) >= 0;
    if (!owned) {
      // outside every box of the level: a valid microphone that records nothing
      m.valid = 2;
      c->mics.push_back(m);
      continue;
    }
    double dist[8] = {}, total = 0.0;
    int bad = 0;
    for (int j = 0; j < 8; ++j) {
      int G[3];
      for (int d = 0; d < 3; ++d) G[d] = m.pos[j][d] = st[d] + kCorner[j][d];
      int64_t node = valid_index(L, G);
      if (node >= 0 && L.state) {  // a link table holds the node codes
        int32_t code = 0;
        HIP_TRY(hipMemcpy(&code, L.state + node, sizeof code, hipMemcpyDeviceToHost));
        if (code < 0) node = -1;
      }
      m.node[j] = node;
      if (node < 0) {
        ++bad;
        continue;
      }
      // distance with the reference's + rdx for the far corners
      // (mic_setup.f95:86-198): the mirror image of the corner about the cell's
      // first node
      double s = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double t = kCorner[j][d] ? p[d] - lo[d] + rdx : p[d] - lo[d];
        s += t * t;
      }
      dist[j] = std::sqrt(s);
      total += dist[j];
    }
    if (bad == 8) {  // mic_setup.f95:200-205
      c->mics.push_back(m);
      continue;
    }
    m.valid = 1;
    for (int j = 0; j < 8; ++j)  // mic_setup.f95:207-213: corners at x index <= 0 weigh 0
      m.w[j] = (m.node[j] >= 0 && m.pos[j][0] - dlo[0] > 0) ? dist[j] / total : 0.0;
    int64_t first = -1;
    for (int j = 0; j < 8 && first < 0; ++j) first = m.node[j];
    m.slot = (unsigned)lists[l].size();
    for (int j = 0; j < 8; ++j) lists[l].push_back((uint32_t)(m.node[j] >= 0 ? m.node[j] : first));
    c->mics.push_back(m);
  }
  // files with their header lines (mic_initialize.f95:17-70)
  static const char* kHead[6] = {"P,          ;", "Temp,       ;", "Density,    ;",
                                 "U-vel,      ;", "V-vel,      ;", "W-vel,      ;"};
  for (Mic& m : c->mics) {
    if (!m.valid) continue;
    char name[32];
    std::snprintf(name, sizeof name, "microphone%05d.txt", m.number);
    m.path = pre + name;
    std::string line = " ";
    for (int k = 0; k < 6; ++k)
      if (m.mask & (32 >> k)) line += kHead[k];
    line += "\n";
    std::FILE* fp = std::fopen(m.path.c_str(), "w");
    if (!fp || std::fputs(line.c_str(), fp) < 0 || std::fclose(fp) != 0) {
      std::string path = m.path;
      free_mics(c);
      return set_err(UCLBS_E_ARG, "mic_setup: cannot write %s", path.c_str());
    }
  }
  size_t off = 0;
  for (int l = 0; l <= fine; ++l) {
    if (lists[l].empty()) continue;
    MicLevel ml;
    ml.lvl = l;
    ml.m = (unsigned)lists[l].size();
    ml.off = off;
    off += 6 * (size_t)ml.m;
    hipError_t e = hipMalloc(&ml.d_nodes, lists[l].size() * sizeof(uint32_t));
    if (e == hipSuccess)
      e = hipMemcpy(ml.d_nodes, lists[l].data(), lists[l].size() * sizeof(uint32_t),
                    hipMemcpyHostToDevice);
    c->mic_levels.push_back(ml);
    if (e != hipSuccess) {
      free_mics(c);
      return set_err(UCLBS_E_HIP, "mic_setup: %s", hipGetErrorString(e));
    }
  }
  c->mic_stride = off;
  if (off) {
    // a ring of samples of at most 64 MiB (4096 samples at most)
    c->mic_cap = (int)std::max<size_t>(1, std::min<size_t>(4096, (64u << 20) / (off * sizeof(double))));
    hipError_t e = hipMalloc(&c->mic_ring, (size_t)c->mic_cap * off * sizeof(double));
    if (e != hipSuccess) {
      free_mics(c);
      return set_err(UCLBS_E_NOMEM, "mic_setup: %s", hipGetErrorString(e));
    }
  }
  c->mic_structure = c->structure;
  if (valid)
    for (int i = 0; i < n; ++i) valid[i] = c->mics[i].valid;
  return UCLBS_OK;
}

int uclbs_mic_weights(void* ctx, int i, int32_t pos[24], double w[8], int64_t node[8]) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (i < 0 || i >= (int)c->mics.size()) return set_err(UCLBS_E_ARG, "mic %d was not set up", i + 1);
  const Mic& m = c->mics[i];
  for (int j = 0; j < 8; ++j) {
    if (pos)
      for (int d = 0; d < 3; ++d) pos[3 * j + d] = m.pos[j][d];
    if (w) w[j] = m.w[j];
    if (node) node[j] = m.node[j];
  }
  return UCLBS_OK;
}

int uclbs_mic_flush(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (c->mic_count == 0) return UCLBS_OK;
  int rc = set_device(c);
  if (rc) return rc;
  const size_t cnt = (size_t)c->mic_count;
  std::vector<double> h(cnt * c->mic_stride);
  HIP_TRY(hipMemcpyAsync(h.data(), c->mic_ring, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->mic_count = 0;
  // probe columns rho, u, v, w, T, p; the file's order is p, T, rho, u, v, w
  static const int kCol[6] = {5, 4, 0, 1, 2, 3};
  std::string text;
  for (const Mic& m : c->mics) {
    if (m.valid != 1 || !m.mask) continue;
    const MicLevel* ml = nullptr;
    for (const MicLevel& x : c->mic_levels)
      if (x.lvl == m.level) ml = &x;
    text.clear();
    for (size_t s = 0; s < cnt; ++s) {
      const double* smp = h.data() + s * c->mic_stride + ml->off;
      bool firstcol = true;
      for (int k = 0; k < 6; ++k) {
        if (!(m.mask & (32 >> k))) continue;
        // the weighted sum over the eight corners in order (mic_write.f95:45-105)
        double v = 0.0;
        for (int j = 0; j < 8; ++j)
          if (m.node[j] >= 0) v = v + m.w[j] * smp[(size_t)kCol[k] * ml->m + m.slot + j];
        if (!firstcol) text += ' ';
        f12_5(v, text);
        firstcol = false;
      }
      text += '\n';
    }
    std::FILE* fp = std::fopen(m.path.c_str(), "a");
    if (!fp || std::fputs(text.c_str(), fp) < 0 || std::fclose(fp) != 0)
      return set_err(UCLBS_E_ARG, "mic_flush: cannot append to %s", m.path.c_str());
  }
  return UCLBS_OK;
}

int uclbs_mic_sample(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(UCLBS_E_ARG, "null context");
  if (c->mic_stride == 0) return UCLBS_OK;
  if (c->mic_structure != c->structure)
    return set_err(UCLBS_E_STATE, "mic_sample: the levels or link tables changed after "
                                  "uclbs_mic_setup (set the microphones up again)");
  int rc;
  if (c->mic_count == c->mic_cap && (rc = uclbs_mic_flush(ctx))) return rc;
  if ((rc = set_device(c))) return rc;
  const int fp32 = c->cfg.precision == UCLBS_PREC_FP32_SHIFTED;
  double* base = c->mic_ring + (size_t)c->mic_count * c->mic_stride;
  for (const MicLevel& ml : c->mic_levels) {
    const Level& L = c->levels[ml.lvl];
    if (c->cfg.lattice == UCLBS_LATTICE_D3Q39)
      HIP_TRY(uclbs::launch_d3q39_probe(L.f[L.cur], L.g[L.cur], L.box.n, fp32, c->shifted39,
                                        c->ref39, L.gas39.cv, ml.d_nodes, ml.m, base + ml.off,
                                        c->stream));
    else
      HIP_TRY(uclbs::launch_d3q27_probe(L.f[L.cur], L.box.n, fp32, ml.d_nodes, ml.m,
                                        base + ml.off, c->stream));
  }
  c->mic_count++;
  return UCLBS_OK;
}

}  // extern "C"

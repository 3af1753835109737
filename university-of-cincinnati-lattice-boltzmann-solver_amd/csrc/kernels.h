// Launch interface between the C-ABI layer (uclbs_abi.hip) and the kernel
// translation units.  Plain structs only; no torch or Fortran types.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "alpha_solve.h"
#include "d3q39.h"

namespace uclbs {

constexpr int kStatSlots = 64;   // per-wave counters are spread over this many slots
constexpr int kStatFields = 10;  // order of uclbs_stats
enum StatField : int { ST_NODES = 0, ST_ITERS, ST_TRIVIAL, ST_BOUND, ST_NONPOS, ST_NOCONV,
                       ST_FALLBACK, ST_LOGPATH, ST_IBB, ST_REPAIR };

constexpr int32_t kFacePeriodic = 2147483647;  // face code: wrap instead of a boundary

struct Box {
  int nx, ny, nz;
  int per[3];
  int64_t n;  // nx*ny*nz
};

struct StepArgs {
  double beta;
  double tol;
  int maxit;
  int alpha_mode;
  int shifted;  // 1: FP32 f - w storage, 0: FP64 full
  SeriesThr thr;
  int32_t face_code[6];  // x-, y-, z-, x+, y+, z+ (links leaving the box)
  // queue of far-from-equilibrium nodes (FP32 storage), in kFarShards + 1
  // shards: shard s holds defer_count[s] entries from defer_list[s * defer_cap];
  // streaming-kernel block b appends to shard b % kFarShards, the boundary
  // kernel to shard kFarShards.  Entry: {node, FP32 bits of the alpha seed}.
  // The lengths are double-buffered per step: the far kernel of a step reads
  // defer_count and zeroes defer_next, which the next step appends to.
  unsigned* defer_count;
  unsigned* defer_next;
  uint2* defer_list;
  unsigned defer_cap;
};
constexpr int kFarShards = 256;

// node flags of a level with link tables
enum : uint8_t { kFlagInterior = 0, kFlagBoundary = 1, kFlagSolid = 2, kFlagFace = 3,
                 kFlagGhost = 4 /* ghost shell of a refined level, filled from the coarser one */,
                 kFlagMirror = 5 /* ghost node of a box that shows a valid node of the same
                                    level: updated as a copy of that node's update */ };
// far-queue entries of mirror nodes carry this bit (solved, not counted)
constexpr unsigned kMirrorBit = 0x80000000u;

struct BcTables {
  const double* feq_in;
  const double* feq_out;
  const double* feq_fs;
};

hipError_t launch_d3q27_step_periodic(const Box& b, const void* fin, void* fout, double* alpha,
                                      int keep_alpha, unsigned long long* stats,
                                      const StepArgs& a, hipStream_t s);
// Boundary data of a level whose links are defined: the reference-form link
// table, the list of fluid nodes with a boundary link, and the equilibrium
// populations of the inlets / outlets / freestream (full FP64 values).
struct BoxDesc;
struct BcArgs {
  const int32_t* state;   // [27][n]
  const uint8_t* flags;   // [n]: kFlag* values
  const uint32_t* list;   // boundary fluid nodes
  int64_t nlist;
  int n_in, n_out;
  const double* feq_in;   // [n_in][27]
  const double* feq_out;  // [n_out][27]
  const double* feq_fs;   // [27] or null
  // a level of many boxes: the boxes and the box of each listed node (its
  // neighbours lie in that box or its ghost shell); null for a single box
  const BoxDesc* boxes;
  const uint32_t* box_of;
};

// Node of the bounding-volume hierarchy over the triangles: bb is the exact
// union (max / min) of the boxes of the triangles below it; an inner node has
// children a, b; a leaf (b < 0) lists -b triangles from idx[a].
struct BvhNode {
  double bb[6];  // xmax, xmin, ymax, ymin, zmax, zmin
  int32_t a, b;
};

struct LinkGeom {
  Box box;
  double origin[3];
  double dx;
  int32_t face_code[6];  // x-, y-, z-, x+, y+, z+
  int ntri;
  const double* tri;  // [ntri][9]
  const double* nrm;  // [ntri][3]
  const double* tbb;  // [ntri][6]: xmax, xmin, ymax, ymin, zmax, zmin
  double gbb[6];      // geometry bounding box, same order
  const BvhNode* bvh;     // hierarchy (node 0 the root), null for no triangles
  const int32_t* bvh_idx; // triangle indices of the leaves
  // ray_len > 0: the solid test casts every node's first ray along the unit
  // vector ray, ray_len long (re-casts ray_len long too), instead of away from
  // the box centre
  double ray[3];
  double ray_len;
};

// Builds the hierarchy on the host (median split of the longest centroid axis,
// leaves of at most 4 triangles) from the triangle boxes tbb; returns its
// depth (the device walk keeps a stack of 40 nodes, enough for depth 38: a
// median split of 2^26 triangles has depth 25).
int build_bvh(const double* tbb, int ntri, std::vector<BvhNode>* nodes,
              std::vector<int32_t>* idx);
constexpr int kBvhMaxDepth = 38;

hipError_t launch_build_links27(const LinkGeom& g, int32_t* state, unsigned long long* counts,
                                hipStream_t s);
hipError_t launch_build_links39(const LinkGeom& g, int32_t* state, unsigned long long* counts,
                                hipStream_t s);
// shifted lattice (c_x + 1): links keyed on their upstream node
hipError_t launch_build_links39s(const LinkGeom& g, int32_t* state, unsigned long long* counts,
                                 hipStream_t s);
hipError_t launch_flags27(const int32_t* state, const Box& b, const int32_t face_code[6],
                          uint8_t* flags, hipStream_t s);
hipError_t launch_d3q27_step_links(const Box& b, const void* fin, void* fout, double* alpha,
                                   int keep_alpha, unsigned long long* stats, const StepArgs& a,
                                   const BcArgs& bc, hipStream_t s);
// interior + face nodes (streaming kernel) and the boundary-list kernel; the
// two write disjoint nodes and may run concurrently
hipError_t launch_d3q27_boundary(const Box& b, const void* fin, void* fout, double* alpha,
                                 int keep_alpha, unsigned long long* stats, const StepArgs& a,
                                 const BcArgs& bc, hipStream_t s);
// queued far-from-equilibrium nodes, after the streaming kernel
hipError_t launch_d3q27_far(const Box& b, void* fout, double* alpha, int keep_alpha,
                            unsigned long long* stats, const StepArgs& a, hipStream_t s);
hipError_t launch_d3q27_init_uniform(const Box& b, void* f, int shifted, const uint8_t* flags,
                                     double rho, double ux, double uy, double uz, hipStream_t s);

hipError_t launch_d3q27_from_fp64(const Box& b, const double* src, void* dst, int shifted,
                                  hipStream_t s);
hipError_t launch_d3q27_to_fp64(const Box& b, const void* src, double* dst, int shifted,
                                hipStream_t s);
hipError_t launch_d3q27_macro(const Box& b, const void* f, int shifted, double* rho,
                              double* ux, double* uy, double* uz, hipStream_t s);
hipError_t launch_d3q27_init_tgv_region(const Box& dom, void* f, int64_t nn, int64_t off,
                                        const int ext[3], const int org[3], int shifted, double u0,
                                        double noise_amp, uint64_t seed, hipStream_t s);
hipError_t launch_d3q27_init_tgv(const Box& b, void* f, int shifted, double u0,
                                 double noise_amp, uint64_t seed, hipStream_t s);

// ---- coarse-fine exchange -------------------------------------------------------
struct LevelMap {   // a level's stored box in its global (level) index space
  Box box;
  int origin[3];    // global index of stored node (0, 0, 0)
  int per[3];
};
struct Range3 {
  int lo[3], hi[3];
};
hipError_t launch_coarse_to_fine27(const void* cprev, const void* cnew, const LevelMap& cm, void* ff,
                                   const LevelMap& fm, const unsigned* ghosts, unsigned nghost,
                                   double theta, int shifted, unsigned long long* stats,
                                   hipStream_t s);
hipError_t launch_fine_to_coarse27(const void* ff, const LevelMap& fm, void* fc, const LevelMap& cm,
                                   const Range3& cover, int shifted, hipStream_t s);

// ---- levels of many boxes ---------------------------------------------------------
// A level may hold many disjoint boxes, each stored with a one-node ghost shell
// in one stretch of the level's population arrays (SoA over the whole level).
struct BoxDesc {
  unsigned off;     // first stored node of the box
  unsigned n;       // stored nodes (ghost shell included)
  unsigned nx, ny;  // stored extents on x and y (nz = n / (nx ny))
};
struct BoxTile {    // one block of the batched streaming kernel
  unsigned box;
  unsigned start;   // first node of the tile within the box
};
struct C2FStencil { // a ghost node interpolated from the coarser level
  unsigned dst;     // fine stored node
  unsigned mask;    // populations stored (bit i; all 27 are interpolated)
  unsigned src[8];  // coarse stored nodes, tensor order (z outer, x inner)
  float w[8];       // their weights (products of 1 and 1/2; 0 = unused)
};
// alias[k] = {valid node mirrored by node k, its box's row and plane strides,
// 0} for kFlagMirror nodes
hipError_t launch_d3q27_step_boxes(unsigned nn, const void* fin, void* fout, double* alpha,
                                   int keep_alpha, unsigned long long* stats, const StepArgs& a,
                                   const uint8_t* flags, const uint4* alias, const BoxTile* tiles,
                                   unsigned ntiles, const BoxDesc* boxes, int tile_nodes,
                                   int links, hipStream_t s);
// flags of the valid nodes of a level of many boxes from its link table:
// solid (node code < 0), boundary (a link code < 0), interior; ghost-shell
// flags are kept
hipError_t launch_flags_boxes27(const int32_t* state, int64_t n, uint8_t* flags, hipStream_t s);
// dst[i][p.x] = src[i][p.y] for each pair p and each population i in p.z
// (bit i; ghost copies carry only the populations the box pulls from a ghost
// node, injections all 27)
constexpr unsigned kAllPops27 = (1u << 27) - 1u;
hipError_t launch_copy_pairs27(const void* src, unsigned nn_src, void* dst, unsigned nn_dst,
                               const uint3* pairs, unsigned npairs, int shifted, hipStream_t s);
// time-interpolated coarse stencils (theta between cprev and cnew), negative
// results replaced by the equilibrium of the interpolated moments
hipError_t launch_c2f_stencils27(const void* cprev, const void* cnew, unsigned nn_c, void* ff,
                                 unsigned nn_f, const C2FStencil* st, unsigned nst, double theta,
                                 int shifted, unsigned long long* stats, hipStream_t s,
                                 int all_pops = 0);

// ---- learned alpha -------------------------------------------------------------
constexpr int kMlpParams = 9217;  // 12-64-64-64-1
struct MlpSamples {               // device-resident training pairs
  float* X = nullptr;             // [cap][12]
  float* y = nullptr;             // [cap] (alpha - 2) / m
  unsigned* node_of = nullptr;    // node of a pending sample
  float* m_of = nullptr;          // its m
  unsigned long long* count = nullptr;
  int64_t cap = 0;
  // ordered placement of one collection (per wavefront ballots of taken
  // nodes, per block counts then slot offsets), sized for blk_n blocks
  unsigned long long* ballots = nullptr;  // [4 blk_n]
  unsigned long long* blk = nullptr;      // [blk_n]
  size_t blk_n = 0;
};
struct MlpWork {                  // activations of one training batch
  float *Z1, *A1, *Z2, *A2, *Z3, *A3, *Z4, *A4, *D1, *D2, *D3, *D4;
  float* part;                    // per-chunk weight-gradient sums: ceil(B / 256) x 4160
};
hipError_t launch_mlp_collect(const Box& b, const float* fin, const uint8_t* flags, int stride,
                              int phase, MlpSamples& s, hipStream_t st);
hipError_t launch_mlp_targets(const double* alpha, MlpSamples& s, unsigned long long first,
                              unsigned long long last, hipStream_t st);
hipError_t launch_mlp_loss_grad(const float* P, const float* X, const float* y,
                                const unsigned* idx, int B, MlpWork& w, float* g, double* loss,
                                hipStream_t st);
hipError_t launch_mlp_forward_fp32(const float* P, const float* X, int B, MlpWork& w, float* yhat,
                                   hipStream_t st);
hipError_t launch_mlp_adam(float* P, const float* g, float* m, float* v, int t, float lr,
                           hipStream_t st);
hipError_t launch_mlp_gather(const float* X, unsigned long long n, uint64_t seed, uint64_t step,
                             int B, float* Xb, unsigned* idx, hipStream_t st);
hipError_t launch_mlp_normalise(float* y, unsigned long long n, double* ms, hipStream_t st);
hipError_t launch_mlp_pack(const float* P, short* Wbf, hipStream_t st);
hipError_t launch_mlp_infer(const float* X, int B, const short* Wbf, const float* P, float* yhat,
                            hipStream_t st);
hipError_t launch_alpha_guess27(const Box& b, const float* fin, const uint8_t* flags,
                                const short* Wbf, const float* P, float ymean, float ystd,
                                double* alpha, hipStream_t st);

// ---- D3Q39 ------------------------------------------------------------------
// reference-state populations subtracted from the FP32 storage (zero for FP64)
struct Ref39 {
  double f[39];
  double g[39];
};
// Faces of a D3Q39 box (x-, y-, z-, x+, y+, z+): kFacePeriodic, -1000
// freestream, -6 pure outflow, -10 bounce-back, -11 inviscid; f / g the
// freestream equilibria (full values).
struct Faces39 {
  int code[6];
  double f[39];
  double g[39];
  // inlet / outlet faces (codes -101..-199, -201..-299): the row of bc holding
  // the equilibrium the face supplies (f then g, 78 doubles a row), -2 for the
  // freestream pair f, g above (type 3), or -1 for a pure-outflow outlet (the
  // node's own population, as -6)
  int tab[6];
  const double* bc;
};
// A node of a D3Q39 box whose upstream nodes x - c_i may lie across a
// non-periodic face: x - c_d < 0 needs x < max c_d, x - c_d >= n needs
// x >= n + min c_d (c_x in -3..3, on the shifted lattice -2..4; c_y, c_z in -3..3).
__host__ __device__ inline bool near_face39(int x, int y, int z, const int code[6], int nx, int ny,
                                            int nz, bool shifted) {
  const int c3[3] = {x, y, z}, n3[3] = {nx, ny, nz};
  bool near = false;
  for (int d = 0; d < 3; ++d) {
    const int lo_reach = (d == 0 && shifted) ? 4 : 3;
    const int hi_reach = (d == 0 && shifted) ? 2 : 3;
    near = near || (code[d] != kFacePeriodic && c3[d] < lo_reach) ||
           (code[d + 3] != kFacePeriodic && c3[d] >= n3[d] - hi_reach);
  }
  return near;
}
// A shifted-lattice D3Q39 level with a link table: node classes (0 plain
// pull, 1 near a face, 2 wall links, 3 solid), the table, the wall nodes and
// the rebuilt f and g of their wall links ([n][39] FP64, full values).
struct Walls39 {
  const uint8_t* cls;     // [n] or null (no table)
  const int32_t* state;   // [39][n]
  const unsigned* list;   // wall nodes
  unsigned n;
  double* wf;
  double* wg;
  long long lvl_step;     // the level's own step counter (start-up multiplier guesses)
  double mach;
};
struct Step39 {
  const void* fin;
  const void* gin;
  void* fout;
  void* gout;
  double* lm;      // [13][n] multipliers, updated in place
  double* alpha;   // [n] or null
  Gas39 gas;
  Ref39 ref;
  long long step;  // reference timestep(amr_max_lvl) seen by the start-up ramps
  int fp32, shifted;
  unsigned long long* stats;
  unsigned* neg_count;  // nodes with a negative post-collision population
  unsigned* neg_list;
  double* rep;          // [n][39] repair values
  int has_faces;        // any non-periodic face: the main launch skips near-face nodes
  Faces39 faces;
  const unsigned* near_list;  // near-face nodes, stepped by a second launch
  unsigned n_near;
  int ghost;            // depth of a refined level's ghost layer (not stepped), else 0
  Walls39 walls;        // cls == null: no link table
};
hipError_t launch_d3q39_step(const Box& b, const Step39& a, hipStream_t s);
hipError_t launch_coarse_to_fine39(const void* cpf, const void* cnf, const void* cpg,
                                   const void* cng, const double* clm, const LevelMap& cm, void* ff,
                                   void* fg, double* flm, const LevelMap& fm,
                                   const unsigned* ghosts, unsigned nghost, double theta, int fp32,
                                   int shifted, const Ref39& rf, const Gas39& gs,
                                   unsigned long long* stats, hipStream_t s);
hipError_t launch_fine_to_coarse39(const void* ff, const void* fg, const double* flm,
                                   const LevelMap& fm, void* cf, void* cg, double* clm,
                                   const LevelMap& cm, const Range3& cover, int fp32, hipStream_t s);
hipError_t launch_d3q39_init(const Box& b, const double* fields, void* f, void* g, double* lm,
                             int fp32, int shifted, const Gas39& gs, const Ref39& rf,
                             unsigned long long* bad, hipStream_t s);
hipError_t launch_d3q39_from_fp64(const Box& b, const double* src, void* dst, int fp32,
                                  const double* ref, hipStream_t s);
// AMReX FArrayBox (device, FP64, component-major, ghost width, component
// stride in elements) <-> [ncomp][n] in the box's node order
hipError_t launch_fab_gather(const double* fab, int ghost, int64_t cstride, const Box& b,
                             int ncomp, double* dense, hipStream_t s);
hipError_t launch_fab_scatter(const double* dense, const Box& b, int ncomp, int ghost,
                              int64_t cstride, double* fab, hipStream_t s);
hipError_t launch_blank_gather(const int32_t* code, int ghost, const Box& b, int32_t* blank,
                               hipStream_t s);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);
hipError_t launch_d3q27_probe(const void* f, int64_t n, int shifted, const uint32_t* nodes,
                              unsigned m, double* out, hipStream_t s);
// gradient-of-gradient-magnitude indicator of a one-box D3Q27 level (scheme 0
// central, 1 upwind, 2 downwind); work 5 n doubles, ggm n doubles, tags n bytes
hipError_t launch_d3q27_ggm(const Box& b, const void* f, int shifted, const int32_t* state,
                            int quantity, int scheme, double limit, const int vlo[3],
                            const int vhi[3], double* work, double* ggm, uint8_t* tags,
                            hipStream_t s);
// the same on a refined level of many boxes, through a dense array over its
// stored nodes' bounding box d (origin dlo)
hipError_t launch_d3q27_ggm_boxes(const void* f, long long ns, int shifted, const int32_t* state,
                                  const BoxDesc* boxes, const int* blo, unsigned nbox,
                                  unsigned max_box_n, const int dlo[3], const Box& d, int quantity,
                                  int scheme, double limit, double* work_s, double* work_d,
                                  uint8_t* filled, uint8_t* tag_d, int32_t* state_d,
                                  double* ggm_s, uint8_t* tag_s, hipStream_t s);
hipError_t launch_d3q39_probe(const void* f, const void* g, int64_t n, int fp32, int shifted,
                              const Ref39& rf, double cv, const uint32_t* nodes, unsigned m,
                              double* out, hipStream_t s);
hipError_t launch_d3q39_macro(const Box& b, const void* f, const void* g, int fp32, int shifted,
                              const Ref39& rf, double cv, double* out, hipStream_t s);
hipError_t launch_d3q39_to_fp64(const Box& b, const void* src, double* dst, int fp32,
                                const double* ref, hipStream_t s);

}  // namespace uclbs

// Launch interface between the C-ABI layer (uclbs_abi.hip) and the kernel
// translation units.  Plain structs only; no torch or Fortran types.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "alpha_solve.h"

namespace uclbs {

constexpr int kStatSlots = 64;   // per-wave counters are spread over this many slots
constexpr int kStatFields = 8;   // order of uclbs_stats
enum StatField : int { ST_NODES = 0, ST_ITERS, ST_TRIVIAL, ST_BOUND, ST_NONPOS, ST_NOCONV,
                       ST_FALLBACK, ST_LOGPATH };

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
  // deferred-node queue (FP32 storage): counters[2] ping-pong by step parity
  unsigned int* defer_count;
  unsigned int* defer_list;
  int parity;
};

hipError_t launch_d3q27_step_periodic(const Box& b, const void* fin, void* fout, double* alpha,
                                      unsigned long long* stats, const StepArgs& a,
                                      hipStream_t s);
// general-path fixup for the nodes the step queued (FP32 storage only)
hipError_t launch_d3q27_fixup(const Box& b, void* fout, double* alpha,
                              unsigned long long* stats, const StepArgs& a, hipStream_t s);
hipError_t launch_d3q27_from_fp64(const Box& b, const double* src, void* dst, int shifted,
                                  hipStream_t s);
hipError_t launch_d3q27_to_fp64(const Box& b, const void* src, double* dst, int shifted,
                                hipStream_t s);
hipError_t launch_d3q27_macro(const Box& b, const void* f, int shifted, double* rho,
                              double* ux, double* uy, double* uz, hipStream_t s);
hipError_t launch_d3q27_init_tgv(const Box& b, void* f, int shifted, double u0,
                                 double noise_amp, uint64_t seed, hipStream_t s);

}  // namespace uclbs

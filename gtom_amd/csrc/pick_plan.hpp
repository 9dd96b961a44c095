// The search plan shared by the exhaustive search (pick.hip) and its
// split-precision form (split.hip): batch buffers, orientation matrices and
// their pinned staging, the local-std scratch of the finish, the learned
// search's tile masks, and per-stage kernel timing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "dev_project.hpp"
#include "runtime.hpp"

struct gtomx_pick_plan {
  int n = 0, p = 0;
  int nz = 0;  // n for n^3 volumes, 1 for n x n images (the 2D branch)
  unsigned B = 0;
  int device = 0;
  size_t H = 0, V = 0;
  int ntiles = 0;        // pass-1 column tiles inside the disc (partial sums per orientation)
  int* tiles = nullptr;  // their workgroup indices
  int grid1[2] = {0, 0}; // resident pass-1 workgroups: reference layout, centred layout
  float2* G = nullptr;
  float* partial = nullptr;
  float* corr = nullptr;  // B x HERM_BLOCKS plane corrections of the unit-RMS sums
  float* scale = nullptr;
  float* mats = nullptr;
  size_t mats_cap = 0;   // orientations
  float* h_mats = nullptr;  // pinned staging
  size_t h_mats_cap = 0;
  hipEvent_t mats_ev = nullptr;
  bool mats_ev_recorded = false;
  void* ls_mem = nullptr;
  gx::LocalStdScratch ls{};
  float* ebuf = nullptr;
  float2* cproj = nullptr;  // centred projector layout
  float* gbest = nullptr;   // 2D: per-orientation-group maxima (groups x V)
  int32_t* gidx = nullptr;
  int og = 0;               // 2D: orientations per group
  bool gen3d = false;       // 3D volumes whose side has no fused or long-row search: chirp-z transforms
  int nparts = 0;           // 2D above 256: pass-1 workgroups per orientation
  float2* scr = nullptr;    // 2D above 256: pass-1 output (B spectra)
  size_t device_bytes = 0;
  // tile masks of the grouped (learned) search: device copy and pinned staging
  uint8_t* tmasks = nullptr;
  size_t tmasks_cap = 0;
  uint8_t* h_tmasks = nullptr;
  size_t h_tmasks_cap = 0;
  // proposals of the learned search (per tile: top-k directions, top-1 probability, class)
  int32_t* ldirs = nullptr;
  float* ltop1 = nullptr;
  int32_t* lcls = nullptr;
  float* lmass = nullptr;  // summed probability of each tile's top-k directions
  float* lpnp = nullptr;   // probability of "no peak"
  size_t lcap = 0;  // tiles x topk
  gx::StreamFence fence;  // the plan's buffers are reused by calls on any stream
  // kernel timing
  bool timing = false;
  int rows_variant = 0;  // GTOMX_ROWS_*: the row-pass kernel of images above 256
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Pending { int stage; size_t a, b; };
  std::vector<Pending> pending;
  double ms[3] = {0, 0, 0};
  unsigned long long launches[3] = {0, 0, 0};
};

namespace gx {

// dimsvolume and nvolumes against the plan, and the current device.
int check_plan_dims(const gtomx_pick_plan* pl, gtomx_int3 dimsvolume, unsigned nvolumes);
// Matrices M_o = Euler(angle_o)^T * oversample of nangles orientations into pl->mats (stream-ordered).
int upload_matrices(gtomx_pick_plan* pl, const gtomx_float3* h_angles, unsigned nangles, float oversample,
                    hipStream_t s);
// Tile masks into pl->tmasks (after upload_matrices in the same call).
int upload_masks(gtomx_pick_plan* pl, const uint8_t* h, size_t bytes, hipStream_t s);
// The projector view the sampler reads: *C (centred copy, returns true) or *P only.
bool setup_projector(gtomx_pick_plan* pl, const gtomx_complex* d_projectorft, float oversample, hipStream_t s,
                     Projector* P, CentredProjector* C, int* status);
// Arguments shared by the update entry points.
int check_update_args(const gtomx_pick_plan* plan, const gtomx_complex* d_projectorft, float oversample,
                      gtomx_int3 dimsprojector, const gtomx_complex* d_expft, gtomx_int3 dimsvolume,
                      unsigned nvolumes, const float* d_best, const int32_t* d_idx);

// The orientation loop of one volume or image: batches of pl->B orientations
// of mats (nangles matrices already on the device) through the search's kernels
// with indices angle0 + o. Kernel launches only (no host synchronisation, no
// allocation), so it can be captured in a graph.
int launch_update_batches(gtomx_pick_plan* pl, const Projector& P, const CentredProjector* C, const float* mats,
                          unsigned nangles, const float2* E, const float* ctf, float* best, int32_t* idx,
                          int angle0, hipStream_t s);
}  // namespace gx

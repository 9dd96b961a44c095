// Micrographs of uneven sizes through the 2D branch of the search in one call.
//
// The reference searches n x n images of one size at a time, 240 orientations
// per pass (src/Correlation/SubTomograms.cu:34-36, Project3DtoNDKernel<2>).
// Here one object holds a search plan per image side (each with its own batch,
// normally the autotuner's choice) and, per side, fixed device slots for one
// image's spectrum, CTF and running maxima. The whole orientation loop of a
// side (fills of the maxima, then every batch's kernels) is captured once as
// a hipGraph on the object's own stream and replayed for each image of that
// side after its spectrum and CTF are copied into the slots; the per-image
// local-std division (the reference's d_LocalStd + d_DivideSafeByVector,
// SubTomograms.cu:103-118) runs after the replay on the image's own buffers.
// A graph is captured again when the projector or the orientation list
// changes. The autotuner entry times candidate batches for one side.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "dev_project.hpp"
#include "pick_plan.hpp"
#include "runtime.hpp"

namespace gx {
namespace {

struct SizeSlot {
  int n = 0;
  gtomx_pick_plan* plan = nullptr;
  float2* E = nullptr;
  float* ctf = nullptr;
  float* best = nullptr;
  int32_t* idx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // what the uploaded matrices, the projector view and the graph were built for
  const void* key_proj = nullptr;
  uint64_t key_angles = 0;
  unsigned key_nangles = 0;
  bool prepared = false;
  Projector P{};
  CentredProjector C{};
  bool centred = false;
};

uint64_t fnv1a(const void* data, size_t bytes) {
  const unsigned char* b = static_cast<const unsigned char*>(data);
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < bytes; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

void drop_graph(SizeSlot& sl) {
  if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
  if (sl.graph) (void)hipGraphDestroy(sl.graph);
  sl.exec = nullptr;
  sl.graph = nullptr;
}

}  // namespace
}  // namespace gx

struct gtomx_mpick {
  int p = 0;
  int device = 0;
  hipStream_t ws = nullptr;  // the object's own (capturable) stream, default priority
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::vector<gx::SizeSlot> slots;
  unsigned long long replays = 0, captures = 0;
};

namespace gx {
namespace {

void mpick_release(gtomx_mpick* mp) {
  if (!mp) return;
  if (mp->ws) (void)hipStreamSynchronize(mp->ws);
  for (auto& sl : mp->slots) {
    drop_graph(sl);
    if (sl.plan) (void)gtomx_pick_plan_destroy(sl.plan);
    dev_free(sl.E);
    dev_free(sl.ctf);
    dev_free(sl.best);
    dev_free(sl.idx);
  }
  if (mp->ev_in) (void)hipEventDestroy(mp->ev_in);
  if (mp->ev_out) (void)hipEventDestroy(mp->ev_out);
  if (mp->ws) (void)hipStreamDestroy(mp->ws);
  delete mp;
}

// Matrices for oversample p/n and the projector view of one side, when the
// projector or the orientation list differ from what the slot holds (the
// graph, which bakes both in, is then dropped).
int prepare_slot(gtomx_mpick* mp, SizeSlot& sl, const gtomx_complex* d_proj, const gtomx_float3* h_angles,
                 unsigned nangles, uint64_t key) {
  if (sl.prepared && sl.key_proj == d_proj && sl.key_angles == key && sl.key_nangles == nangles) return GTOMX_OK;
  drop_graph(sl);
  const float oversample = (float)mp->p / (float)sl.n;
  GX_TRY(upload_matrices(sl.plan, h_angles, nangles, oversample, mp->ws));
  int st = GTOMX_OK;
  sl.centred = setup_projector(sl.plan, d_proj, oversample, mp->ws, &sl.P, &sl.C, &st);
  GX_TRY(st);
  sl.key_proj = d_proj;
  sl.key_angles = key;
  sl.key_nangles = nangles;
  sl.prepared = true;
  return GTOMX_OK;
}

// The orientation loop of one image in the slot: maxima reset, then every batch.
int slot_loop(SizeSlot& sl, unsigned nangles, hipStream_t s) {
  const size_t V = (size_t)sl.n * sl.n;
  GX_TRY(fill_f32(sl.best, V, -1e30f, s));  // d_ValueFill (SubTomograms.cu:41-42)
  GX_TRY(fill_i32(sl.idx, V, 0, s));
  return launch_update_batches(sl.plan, sl.P, sl.centred ? &sl.C : nullptr, sl.plan->mats, nangles, sl.E, sl.ctf,
                               sl.best, sl.idx, 0, s);
}

int capture_slot(gtomx_mpick* mp, SizeSlot& sl, unsigned nangles) {
  GX_HIP(hipStreamBeginCapture(mp->ws, hipStreamCaptureModeRelaxed));
  const int st = slot_loop(sl, nangles, mp->ws);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(mp->ws, &g);
  if (st != GTOMX_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  sl.graph = g;
  GX_HIP(hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0));
  mp->captures++;
  return GTOMX_OK;
}

SizeSlot* slot_for(gtomx_mpick* mp, int n) {
  for (auto& sl : mp->slots)
    if (sl.n == n) return &sl;
  return nullptr;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gtomx_mpick_create(gtomx_mpick** out, int p, const int* h_sides, const unsigned* h_batches,
                       const int* h_variants, unsigned nsides) {
  if (!out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "object pointer is NULL");
  *out = nullptr;
  if (!h_sides || nsides == 0 || nsides > 64) return fail(GTOMX_ERR_INVALID_ARGUMENT, "1 to 64 image sides");
  auto* mp = new gtomx_mpick();
  mp->p = p;
  hipError_t he = hipGetDevice(&mp->device);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&mp->ws, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&mp->ev_in, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&mp->ev_out, hipEventDisableTiming);
  if (he != hipSuccess) {
    mpick_release(mp);
    return hip_fail(he, "micrograph batch set-up");
  }
  for (unsigned k = 0; k < nsides; ++k) {
    const int n = h_sides[k];
    bool dup = false;
    for (unsigned j = 0; j < k; ++j) dup |= h_sides[j] == n;
    if (dup) {
      mpick_release(mp);
      return fail(GTOMX_ERR_INVALID_ARGUMENT, "image sides must be distinct");
    }
    mp->slots.emplace_back();
    SizeSlot& sl = mp->slots.back();
    sl.n = n;
    int st = gtomx_pick_plan_create(&sl.plan, gtomx_int3{p, p, p}, gtomx_int3{n, n, 1}, h_batches ? h_batches[k] : 0);
    if (st == GTOMX_OK && h_variants) st = gtomx_pick_plan_set_rows_variant(sl.plan, h_variants[k]);
    const size_t H = (size_t)(n / 2 + 1) * n, V = (size_t)n * n;
    void* d = nullptr;
    if (st == GTOMX_OK && (st = dev_alloc(&d, H * sizeof(float2))) == GTOMX_OK) sl.E = static_cast<float2*>(d);
    if (st == GTOMX_OK && (st = dev_alloc(&d, H * sizeof(float))) == GTOMX_OK) sl.ctf = static_cast<float*>(d);
    if (st == GTOMX_OK && (st = dev_alloc(&d, V * sizeof(float))) == GTOMX_OK) sl.best = static_cast<float*>(d);
    if (st == GTOMX_OK && (st = dev_alloc(&d, V * sizeof(int32_t))) == GTOMX_OK) sl.idx = static_cast<int32_t*>(d);
    if (st != GTOMX_OK) {
      std::string msg = gtomx_last_error();
      mpick_release(mp);
      return fail(st, msg);
    }
  }
  *out = mp;
  return GTOMX_OK;
}

int gtomx_mpick_destroy(gtomx_mpick* mp) {
  mpick_release(mp);
  return GTOMX_OK;
}

int gtomx_mpick_info(const gtomx_mpick* mp, unsigned* h_batches, int* h_variants, unsigned long long* captures,
                     unsigned long long* replays) {
  if (!mp) return fail(GTOMX_ERR_INVALID_ARGUMENT, "object is NULL");
  for (size_t k = 0; k < mp->slots.size(); ++k) {
    if (h_batches) h_batches[k] = mp->slots[k].plan->B;
    if (h_variants) h_variants[k] = mp->slots[k].plan->rows_variant;
  }
  if (captures) *captures = mp->captures;
  if (replays) *replays = mp->replays;
  return GTOMX_OK;
}

int gtomx_mpick_run(gtomx_mpick* mp, const gtomx_complex* d_projectorft, const gtomx_micrograph* h_images,
                    unsigned nimages, const gtomx_float3* h_angles, unsigned nangles, int use_graphs, void* stream) {
  if (!mp) return fail(GTOMX_ERR_INVALID_ARGUMENT, "object is NULL");
  if (!d_projectorft || (nimages && !h_images)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL projector or images");
  if (nangles == 0 || !h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "no orientations");
  if (nangles > 0x7fffffffu) return fail(GTOMX_ERR_INVALID_ARGUMENT, "orientation index exceeds int32");
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev != mp->device) return fail(GTOMX_ERR_INVALID_ARGUMENT, "object was created on another device");
  for (unsigned i = 0; i < nimages; ++i) {
    const gtomx_micrograph& im = h_images[i];
    if (!slot_for(mp, im.n)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "image side without a plan in this object");
    if (!im.d_spectrum || !im.d_ctf || !im.d_best || !im.d_index)
      return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL image buffer");
    if (!(im.maskradius >= 0.f) || !std::isfinite(im.maskradius))
      return fail(GTOMX_ERR_INVALID_ARGUMENT, "maskradius must be finite and non-negative");
  }
  if (nimages == 0) return GTOMX_OK;
  const uint64_t key = fnv1a(h_angles, (size_t)nangles * sizeof(gtomx_float3));
  hipStream_t us = as_stream(stream);
  GX_HIP(hipEventRecord(mp->ev_in, us));
  GX_HIP(hipStreamWaitEvent(mp->ws, mp->ev_in, 0));
  // images grouped by side, each group in its given order
  for (auto& sl : mp->slots) {
    bool any = false;
    for (unsigned i = 0; i < nimages && !any; ++i) any = h_images[i].n == sl.n;
    if (!any) continue;
    GX_TRY(prepare_slot(mp, sl, d_projectorft, h_angles, nangles, key));
    if (use_graphs && !sl.exec) GX_TRY(capture_slot(mp, sl, nangles));
    const size_t H = (size_t)(sl.n / 2 + 1) * sl.n, V = (size_t)sl.n * sl.n;
    for (unsigned i = 0; i < nimages; ++i) {
      const gtomx_micrograph& im = h_images[i];
      if (im.n != sl.n) continue;
      GX_HIP(hipMemcpyAsync(sl.E, im.d_spectrum, H * sizeof(float2), hipMemcpyDeviceToDevice, mp->ws));
      GX_HIP(hipMemcpyAsync(sl.ctf, im.d_ctf, H * sizeof(float), hipMemcpyDeviceToDevice, mp->ws));
      if (use_graphs) {
        GX_HIP(hipGraphLaunch(sl.exec, mp->ws));
        mp->replays++;
      } else {
        GX_TRY(slot_loop(sl, nangles, mp->ws));
      }
      GX_HIP(hipMemcpyAsync(im.d_best, sl.best, V * sizeof(float), hipMemcpyDeviceToDevice, mp->ws));
      GX_HIP(hipMemcpyAsync(im.d_index, sl.idx, V * sizeof(int32_t), hipMemcpyDeviceToDevice, mp->ws));
      GX_TRY(gtomx_pick_finish(sl.plan, im.d_spectrum, gtomx_int3{sl.n, sl.n, 1}, 1, im.maskradius, im.d_best,
                               im.d_index, nullptr, mp->ws));
    }
  }
  GX_HIP(hipEventRecord(mp->ev_out, mp->ws));
  GX_HIP(hipStreamWaitEvent(us, mp->ev_out, 0));
  return GTOMX_OK;
}

int gtomx_mpick_tune(int p, int n, const gtomx_complex* d_projectorft, const unsigned* h_batches,
                     const int* h_variants, unsigned ncand, const gtomx_float3* h_angles, unsigned nangles,
                     int repeats, float* h_ms_per_orientation, void* stream) {
  if (!d_projectorft || !h_batches || !h_angles || !h_ms_per_orientation || ncand == 0 || nangles == 0 ||
      repeats < 1)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "bad autotuner arguments");
  hipStream_t s = as_stream(stream);
  const size_t H = (size_t)(n / 2 + 1) * n, V = (size_t)n * n;
  DevBuf E, best, idx;
  GX_TRY(E.alloc(H * sizeof(float2)));
  GX_TRY(best.alloc(V * sizeof(float)));
  GX_TRY(idx.alloc(V * sizeof(int32_t)));
  GX_HIP(hipMemsetAsync(E.p, 0, H * sizeof(float2), s));
  hipEvent_t a = nullptr, b = nullptr;
  GX_HIP(hipEventCreate(&a));
  GX_HIP(hipEventCreate(&b));
  int result = GTOMX_OK;
  for (unsigned c = 0; c < ncand; ++c) {
    h_ms_per_orientation[c] = std::numeric_limits<float>::infinity();
    gtomx_pick_plan* plan = nullptr;
    if (gtomx_pick_plan_create(&plan, gtomx_int3{p, p, p}, gtomx_int3{n, n, 1}, h_batches[c]) != GTOMX_OK)
      continue;  // a batch that does not fit is skipped (its time stays infinite)
    if (h_variants && gtomx_pick_plan_set_rows_variant(plan, h_variants[c]) != GTOMX_OK) {
      gtomx_pick_plan_destroy(plan);
      result = GTOMX_ERR_INVALID_ARGUMENT;
      break;
    }
    const float oversample = (float)p / (float)n;
    Projector P;
    CentredProjector C{};
    int st = upload_matrices(plan, h_angles, nangles, oversample, s);
    const bool centred = st == GTOMX_OK && setup_projector(plan, d_projectorft, oversample, s, &P, &C, &st);
    auto loop = [&]() {
      return launch_update_batches(plan, P, centred ? &C : nullptr, plan->mats, nangles, E.as<float2>(), nullptr,
                                   best.as<float>(), idx.as<int32_t>(), 0, s);
    };
    if (st == GTOMX_OK) st = loop();  // warm-up
    if (st == GTOMX_OK && hipEventRecord(a, s) != hipSuccess) st = GTOMX_ERR_DEVICE;
    for (int r = 0; r < repeats && st == GTOMX_OK; ++r) st = loop();
    float ms = 0.f;
    if (st == GTOMX_OK && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
        hipEventElapsedTime(&ms, a, b) == hipSuccess)
      h_ms_per_orientation[c] = ms / (float)(repeats * (double)nangles);
    else if (st != GTOMX_OK)
      result = st;
    (void)hipStreamSynchronize(s);
    gtomx_pick_plan_destroy(plan);
    if (result != GTOMX_OK) break;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return result;
}

}  // extern "C"

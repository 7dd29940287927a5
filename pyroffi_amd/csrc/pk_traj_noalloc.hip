// pk_traj_lbfgs_noalloc_f32: the direct L-BFGS trajectory optimiser for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// pk_traj_lbfgs_f32 keeps its solver state (current and previous iterate,
// previous gradient, the s/y history and per-trajectory scalars) in a
// workspace the caller passes in.  This variant takes that workspace from
// `scratch`, a file-scope device array of kScratchFloats floats (16 MiB) that
// every GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel (it is a host-side stand-in, not a
// device address): the device address is queried once per device with
// hipGetSymbolAddress, with that device current, and kept per device.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents.  The optimiser reads no state it has not written in the same
// solve (iteration 0 initialises the iterate, history counters and scalars),
// so nothing has to be cleared between calls; a solve split over several
// calls (iter_begin > 0) resumes from the array and must not be interleaved
// with another solve on the same GPU.  Calls that overlap on one GPU (two
// streams, two host threads) race on the array.
#include <mutex>

#include "pk_common.h"

namespace {

constexpr long long kScratchFloats = 1LL << 22;  // 16 MiB: ~40,000 waypoints of a 7-joint robot at m = 6
constexpr int kMaxDevices = 64;

__device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, queried on the first call
// on that device.
int scratch_on_current_device(float** out) {
  int dev = -1;
  const hipError_t ge = hipGetDevice(&dev);
  if (ge != hipSuccess) return pk::set_error(PK_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(ge));
  if (dev < 0 || dev >= kMaxDevices) return pk::set_error(PK_ERR_INVALID, "device %d outside [0, %d)", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return pk::set_error(PK_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PK_OK;
}

}  // namespace

extern "C" int pk_traj_lbfgs_noalloc_f32(void* stream, const pk_traj_model* model, const float* x_init,
                                         const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                                         int32_t max_len, int32_t block_threads, const int32_t* traj_subset,
                                         int32_t n_subset, const pk_lbfgs_config* cfg, int32_t iter_begin,
                                         int32_t iter_end, float* out_best, float* out_best_cost) {
  PK_REQUIRE(model != nullptr && cfg != nullptr, "null model or config");
  const int64_t need = pk_traj_lbfgs_state_floats(n_waypoints, model->n_act, n_traj, cfg->m_hist);
  PK_REQUIRE(need >= 0, "invalid sizes (n_waypoints=%lld, n_act=%d, n_traj=%d, m_hist=%d)",
             (long long)n_waypoints, model->n_act, n_traj, cfg->m_hist);
  PK_REQUIRE(need <= kScratchFloats,
             "solver state of %lld floats exceeds the %lld-float scratch of the no-allocation entry point; "
             "use pk_traj_lbfgs_f32 with a workspace",
             (long long)need, kScratchFloats);
  if (n_traj == 0 || iter_begin == iter_end) return pk_traj_lbfgs_f32(stream, model, x_init, traj_start, n_traj,
                                                                      n_waypoints, max_len, block_threads,
                                                                      traj_subset, n_subset, cfg, iter_begin,
                                                                      iter_end, nullptr, out_best, out_best_cost);
  float* ptr = nullptr;
  const int st = scratch_on_current_device(&ptr);
  if (st != PK_OK) return st;
  return pk_traj_lbfgs_f32(stream, model, x_init, traj_start, n_traj, n_waypoints, max_len, block_threads,
                           traj_subset, n_subset, cfg, iter_begin, iter_end, ptr, out_best, out_best_cost);
}

// Side streams for independent launch sequences inside one library call
// (the top, bottom and central partitions of PSR / RGF-2sided): one GPU's
// counterpart of the reference's concurrently running ranks.
#include <mutex>

#include "sinv_core.hpp"

namespace sinv {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_side_mu;
hipStream_t g_side[kMaxDevices][StreamFork::kMaxSide] = {};

hipError_t side_stream(int i, hipStream_t* out) {
  int dev = 0;
  SINV_HIP_RET(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_side_mu);
  if (g_side[dev][i] == nullptr)
    SINV_HIP_RET(hipStreamCreateWithFlags(&g_side[dev][i], hipStreamNonBlocking));
  *out = g_side[dev][i];
  return hipSuccess;
}

}  // namespace

StreamFork::StreamFork(hipStream_t main, int nside) : main_(main) {
  if (nside < 0 || nside > kMaxSide) {
    err_ = hipErrorInvalidValue;
    joined_ = true;
    return;
  }
  err_ = hipEventCreateWithFlags(&fork_, hipEventDisableTiming);
  if (err_ == hipSuccess) err_ = hipEventRecord(fork_, main_);
  for (int i = 0; i < nside && err_ == hipSuccess; ++i) {
    err_ = side_stream(i, &side_[i]);
    if (err_ == hipSuccess) err_ = hipStreamWaitEvent(side_[i], fork_, 0);
    if (err_ == hipSuccess) n_ = i + 1;
  }
}

hipError_t StreamFork::join() {
  if (joined_) return err_;
  joined_ = true;
  for (int i = 0; i < n_; ++i) {
    hipEvent_t done = nullptr;
    hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(done, side_[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(main_, done, 0);
    // resources of a pending event are released once it completes
    if (done != nullptr) (void)hipEventDestroy(done);
    if (e != hipSuccess && err_ == hipSuccess) err_ = e;
  }
  if (fork_ != nullptr) (void)hipEventDestroy(fork_);
  fork_ = nullptr;
  return err_;
}

}  // namespace sinv

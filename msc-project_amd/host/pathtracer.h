// pathtracer.h -- the reference's host facade (include/core/Pathtracer.h:26-69) over the C ABI.
//
//   Pathtracer(filename)  loads a two-document YAML scene and builds it on the device;
//   image(&pixels, &w, &h) exposes the progressive framebuffer (sum over iterations, as the
//                          reference's m_image->pixels, Pathtracer.cpp:394-400);
//   process()             renders one iteration and returns the iteration count, or 0 when
//                          terminate() was called (Pathtracer.cpp:486-487);
//   clear(), active(), terminate() as in Pathtracer.cpp:402-423.
#pragma once

#include <atomic>
#include <stdexcept>
#include <string>
#include <vector>

#include "swr.h"

namespace swr {

class Pathtracer {
 public:
  explicit Pathtracer(const std::string& filename, int device = 0) {
    if (swr_create(device, &ctx_) != SWR_OK) throw std::runtime_error("no HIP device available");
    if (swr_scene_load_yaml(ctx_, filename.c_str(), &scene_, &desc_) != SWR_OK) {
      std::string msg = swr_last_error(ctx_);
      swr_destroy(ctx_);
      throw std::runtime_error(msg);
    }
    pixels_.assign(3 * (size_t)desc_.width * desc_.height, 0.f);
    frame_.assign(pixels_.size(), 0.f);
  }
  ~Pathtracer() {
    swr_scene_destroy(scene_);
    swr_destroy(ctx_);
  }
  Pathtracer(const Pathtracer&) = delete;
  Pathtracer& operator=(const Pathtracer&) = delete;

  void image(float** pixels, int* width, int* height) {
    *pixels = pixels_.data();
    *width = desc_.width;
    *height = desc_.height;
  }
  void clear() {
    std::fill(pixels_.begin(), pixels_.end(), 0.f);
    iteration_ = 0;
  }
  // terminate() may come from another thread than process() (the reference's viewer thread,
  // main.cpp:83-102, through boost::atomic<bool> m_terminate, Pathtracer.h:98).
  bool active() const { return !terminated_.load(std::memory_order_acquire); }
  void terminate() {
    terminated_.store(true, std::memory_order_release);
    swr_cancel(ctx_);
  }
  int process() {
    desc_.iteration = (uint32_t)iteration_;
    const int rc = swr_render(ctx_, scene_, &desc_, frame_.data(), nullptr, &stats_);
    if (rc == SWR_E_CANCELLED) return 0;
    if (rc != SWR_OK) throw std::runtime_error(swr_last_error(ctx_));
    for (size_t i = 0; i < pixels_.size(); ++i) pixels_[i] += frame_[i];
    return ++iteration_;
  }
  const swr_stats& stats() const { return stats_; }
  const swr_render_desc& desc() const { return desc_; }

 private:
  swr_context* ctx_ = nullptr;
  swr_scene* scene_ = nullptr;
  swr_render_desc desc_{};
  swr_stats stats_{};
  std::vector<float> pixels_, frame_;
  int iteration_ = 0;
  std::atomic<bool> terminated_{false};
};

}  // namespace swr

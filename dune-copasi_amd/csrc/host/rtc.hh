// Runtime compilation of a model's generated assembly kernels with hipRTC,
// with an on-disk cache of code objects keyed by a hash of the source, the
// options and the hipRTC / HIP versions; cache files carry a size and a
// checksum, and a file that fails them or that the runtime will not load is
// compiled again.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

namespace dcopasi {

// Compile `source` for gfx950 (or take it from the cache when use_cache and a
// valid file exists); returns the code object bytes and refreshes the cache.
std::vector<char> rtc_compile(const std::string& source, const std::string& name, bool use_cache = true);

class RtcModule {
 public:
  RtcModule() = default;
  RtcModule(const RtcModule&) = delete;
  RtcModule& operator=(const RtcModule&) = delete;
  ~RtcModule();
  void load(const std::vector<char>& code);
  bool try_load(const std::vector<char>& code);  // false (nothing loaded) when the runtime refuses it
  hipFunction_t function(const std::string& name) const;
  bool loaded() const { return mod_ != nullptr; }

 private:
  hipModule_t mod_ = nullptr;
};

// Process-wide cache of loaded modules: models with identical generated
// source on one device share one module.
std::shared_ptr<RtcModule> rtc_module(const std::string& source, const std::string& name, int device);

}  // namespace dcopasi

#include "rtc.hh"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "error.hh"

namespace dcopasi {

namespace {

const std::vector<std::string>& options() {
  static const std::vector<std::string> o = {
      "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  return o;
}

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) h = (h ^ c) * 1099511628211ull;
  return h;
}

std::string cache_dir() {
  const char* env = std::getenv("DCOPASI_RTC_CACHE");
  std::string d;
  if (env && *env) {
    d = env;
  } else {
    const char* tmp = std::getenv("TMPDIR");
    d = std::string(tmp && *tmp ? tmp : "/tmp") + "/dcopasi-rtc-" + std::to_string(getuid());
  }
  mkdir(d.c_str(), 0700);
  return d;
}

#define DCG_RTC(call)                                                                                  \
  do {                                                                                                 \
    hiprtcResult r_ = (call);                                                                          \
    if (r_ != HIPRTC_SUCCESS) throw DeviceError(std::string(#call) + ": " + hiprtcGetErrorString(r_)); \
  } while (0)

}  // namespace

std::vector<char> rtc_compile(const std::string& source, const std::string& name) {
  std::string key = source;
  for (auto& o : options()) key += "\n" + o;
  char hex[32];
  std::snprintf(hex, sizeof hex, "%016llx", static_cast<unsigned long long>(fnv1a(key)));
  const std::string path = cache_dir() + "/" + name + "-" + hex + ".co";
  {
    std::ifstream in(path, std::ios::binary);
    if (in) {
      std::vector<char> code((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      if (!code.empty()) return code;
    }
  }
  hiprtcProgram prog;
  DCG_RTC(hiprtcCreateProgram(&prog, source.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr));
  std::vector<const char*> opts;
  for (auto& o : options()) opts.push_back(o.c_str());
  hiprtcResult res = hiprtcCompileProgram(prog, static_cast<int>(opts.size()), opts.data());
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  std::string log(log_size, '\0');
  if (log_size) hiprtcGetProgramLog(prog, &log[0]);
  if (res != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    throw DeviceError("hipRTC compilation of the assembly kernels failed:\n" + log);
  }
  size_t size = 0;
  DCG_RTC(hiprtcGetCodeSize(prog, &size));
  std::vector<char> code(size);
  DCG_RTC(hiprtcGetCode(prog, code.data()));
  hiprtcDestroyProgram(&prog);
  // write-then-rename so concurrent processes never read a partial file
  const std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream out(tmp, std::ios::binary);
    out.write(code.data(), static_cast<std::streamsize>(code.size()));
  }
  std::rename(tmp.c_str(), path.c_str());
  return code;
}

RtcModule::~RtcModule() {
  if (mod_) (void)hipModuleUnload(mod_);
}

void RtcModule::load(const std::vector<char>& code) {
  DCG_HIP(hipModuleLoadData(&mod_, code.data()));
}

std::shared_ptr<RtcModule> rtc_module(const std::string& source, const std::string& name, int device) {
  static std::mutex mu;
  static std::map<std::pair<int, std::string>, std::weak_ptr<RtcModule>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(device, source);
  if (auto it = cache.find(key); it != cache.end())
    if (auto m = it->second.lock()) return m;
  auto m = std::make_shared<RtcModule>();
  m->load(rtc_compile(source, name));
  cache[key] = m;
  return m;
}

hipFunction_t RtcModule::function(const std::string& name) const {
  hipFunction_t f = nullptr;
  DCG_HIP(hipModuleGetFunction(&f, mod_, name.c_str()));
  return f;
}

}  // namespace dcopasi

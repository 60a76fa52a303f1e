#include "rtc.hh"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// Cache files: an 8-byte magic, the code size and an FNV-1a checksum of the
// code (both uint64), then the code object.  A file that fails any of these
// checks (truncated, overwritten, from an older layout) is ignored and the
// source compiled again.
constexpr char kMagic[8] = {'D', 'C', 'G', 'C', 'O', 'v', '2', '\0'};

std::string cache_path(const std::string& source, const std::string& name) {
  // the key: source, options, and the compiler and runtime that made the code
  std::string key = source;
  for (auto& o : options()) key += "\n" + o;
  int major = 0, minor = 0;
  if (hiprtcVersion(&major, &minor) == HIPRTC_SUCCESS)
    key += "\nhiprtc " + std::to_string(major) + "." + std::to_string(minor);
  key += "\nhip " + std::to_string(HIP_VERSION);
  char hex[32];
  std::snprintf(hex, sizeof hex, "%016llx", static_cast<unsigned long long>(fnv1a(key)));
  return cache_dir() + "/" + name + "-" + hex + ".co";
}

bool read_cached(const std::string& path, std::vector<char>* code) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  std::vector<char> all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const size_t head = sizeof kMagic + 2 * sizeof(uint64_t);
  if (all.size() < head || std::memcmp(all.data(), kMagic, sizeof kMagic) != 0) return false;
  uint64_t size = 0, sum = 0;
  std::memcpy(&size, all.data() + sizeof kMagic, sizeof size);
  std::memcpy(&sum, all.data() + sizeof kMagic + sizeof size, sizeof sum);
  if (size == 0 || all.size() != head + size) return false;
  code->assign(all.begin() + static_cast<long>(head), all.end());
  return fnv1a(std::string(code->data(), code->size())) == sum;
}

void write_cached(const std::string& path, const std::vector<char>& code) {
  const uint64_t size = code.size(), sum = fnv1a(std::string(code.data(), code.size()));
  // write-then-rename so concurrent processes never read a partial file
  const std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream out(tmp, std::ios::binary);
    out.write(kMagic, sizeof kMagic);
    out.write(reinterpret_cast<const char*>(&size), sizeof size);
    out.write(reinterpret_cast<const char*>(&sum), sizeof sum);
    out.write(code.data(), static_cast<std::streamsize>(code.size()));
    if (!out) {
      std::remove(tmp.c_str());
      return;  // an unwritable cache only costs the next compile
    }
  }
  std::rename(tmp.c_str(), path.c_str());
}

std::vector<char> compile(const std::string& source, const std::string& name) {
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
  return code;
}

}  // namespace

std::vector<char> rtc_compile(const std::string& source, const std::string& name, bool use_cache) {
  const std::string path = cache_path(source, name);
  std::vector<char> code;
  if (use_cache && read_cached(path, &code)) return code;
  code = compile(source, name);
  write_cached(path, code);
  return code;
}

RtcModule::~RtcModule() {
  if (mod_) (void)hipModuleUnload(mod_);
}

void RtcModule::load(const std::vector<char>& code) {
  DCG_HIP(hipModuleLoadData(&mod_, code.data()));
}

bool RtcModule::try_load(const std::vector<char>& code) {
  hipModule_t m = nullptr;
  if (hipModuleLoadData(&m, code.data()) != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky error of the failed load
    return false;
  }
  mod_ = m;
  return true;
}

std::shared_ptr<RtcModule> rtc_module(const std::string& source, const std::string& name, int device) {
  static std::mutex mu;
  static std::map<std::pair<int, std::string>, std::weak_ptr<RtcModule>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(device, source);
  if (auto it = cache.find(key); it != cache.end())
    if (auto m = it->second.lock()) return m;
  auto m = std::make_shared<RtcModule>();
  // a cached code object that the runtime refuses (a stale or damaged file
  // that passed the checks) is compiled again and replaced
  if (!m->try_load(rtc_compile(source, name, true))) m->load(rtc_compile(source, name, false));
  cache[key] = m;
  return m;
}

hipFunction_t RtcModule::function(const std::string& name) const {
  hipFunction_t f = nullptr;
  DCG_HIP(hipModuleGetFunction(&f, mod_, name.c_str()));
  return f;
}

}  // namespace dcopasi

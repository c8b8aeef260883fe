// C-ABI runtime: device, queues, buffers, hiprtc programs, launches, events,
// graphs, SHA1.  Replaces the reference's OpenCL runtime classes for one GPU
// (include/SkelCL/detail/{Device,DeviceList,DeviceBuffer,Program}.h,
// src/{Device,DeviceList,DeviceBuffer,Program,Util}.cpp).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "skelcl_rt.h"

// fill / generator kernels (native/fill_gen.hip)
extern "C" int skl_internal_fill(void* dst, const void* pattern, size_t elem,
                                 size_t count, hipStream_t s);
extern "C" int skl_internal_gen_f32_sym(float* out, size_t n, uint64_t seed,
                                        hipStream_t s);
extern "C" int skl_internal_gen_f32_unit(float* out, size_t n, uint64_t seed,
                                         float scale, hipStream_t s);
extern "C" int skl_internal_gen_i32_below(int32_t* out, size_t n,
                                          uint64_t seed, uint32_t m,
                                          hipStream_t s);

namespace skl {

namespace {
thread_local std::string g_error;
std::mutex g_mutex;
bool g_init = false;
int g_device = 0;
hipDeviceProp_t g_props;
hipStream_t g_stream = nullptr;
}  // namespace

void set_error(const std::string& msg) { g_error = msg; }

int fail(int status, const char* fmt, ...)
{
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  if (log_level() >= 1) fprintf(stderr, "[skelcl] error: %s\n", buf);
  return status;
}

int log_level()
{
  static int lvl = [] {
    const char* e = getenv("SKELCL_LOG");
    return e ? atoi(e) : 0;
  }();
  return lvl;
}

void logmsg(int level, const char* fmt, ...)
{
  if (level > log_level()) return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "[skelcl] ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
}

bool initialized() { return g_init; }

int ensure_initialized()
{
  if (!g_init)
    return fail(SKL_ERR_NOT_INIT, "skelcl is not initialised: call skl_init");
  return SKL_OK;
}

hipStream_t resolve_stream(skl_stream s)
{
  return s ? reinterpret_cast<hipStream_t>(s) : g_stream;
}

const hipDeviceProp_t& device_props() { return g_props; }
int device_id() { return g_device; }

std::string target_arch()
{
  const char* e = getenv("SKELCL_ARCH");
  return e && *e ? std::string(e) : std::string("gfx950");
}

// ---------------------------------------------------------------------------
// SHA1 (FIPS 180-4), used as the program cache key like util::hash
// (src/Util.cpp:159-180).
// ---------------------------------------------------------------------------
namespace {
struct Sha1 {
  uint32_t h[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u,
                   0xC3D2E1F0u};
  static uint32_t rol(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }
  void block(const unsigned char* p)
  {
    uint32_t w[80];
    for (int t = 0; t < 16; ++t)
      w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 |
             (uint32_t)p[4 * t + 2] << 8 | (uint32_t)p[4 * t + 3];
    for (int t = 16; t < 80; ++t)
      w[t] = rol(w[t - 3] ^ w[t - 8] ^ w[t - 14] ^ w[t - 16], 1);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
    for (int t = 0; t < 80; ++t) {
      uint32_t f, k;
      if (t < 20) {
        f = (b & c) | (~b & d);
        k = 0x5A827999u;
      } else if (t < 40) {
        f = b ^ c ^ d;
        k = 0x6ED9EBA1u;
      } else if (t < 60) {
        f = (b & c) | (b & d) | (c & d);
        k = 0x8F1BBCDCu;
      } else {
        f = b ^ c ^ d;
        k = 0xCA62C1D6u;
      }
      uint32_t tmp = rol(a, 5) + f + e + k + w[t];
      e = d;
      d = c;
      c = rol(b, 30);
      b = a;
      a = tmp;
    }
    h[0] += a;
    h[1] += b;
    h[2] += c;
    h[3] += d;
    h[4] += e;
  }
};
}  // namespace

std::string sha1_hex(const std::string& data)
{
  Sha1 s;
  const unsigned char* p = reinterpret_cast<const unsigned char*>(data.data());
  size_t n = data.size(), i = 0;
  for (; i + 64 <= n; i += 64) s.block(p + i);
  unsigned char tail[128];
  size_t rem = n - i;
  memcpy(tail, p + i, rem);
  tail[rem] = 0x80;
  size_t padlen = (rem + 1 + 8 <= 64) ? 64 : 128;
  memset(tail + rem + 1, 0, padlen - rem - 1);
  uint64_t bits = (uint64_t)n * 8u;
  for (int k = 0; k < 8; ++k)
    tail[padlen - 1 - k] = (unsigned char)(bits >> (8 * k));
  s.block(tail);
  if (padlen == 128) s.block(tail + 64);
  char out[41];
  for (int k = 0; k < 5; ++k) snprintf(out + 8 * k, 9, "%08x", s.h[k]);
  return std::string(out, 40);
}

}  // namespace skl

using namespace skl;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
extern "C" const char* skl_last_error(void) { return g_error.c_str(); }
extern "C" void skl_free_string(char* s) { free(s); }

static char* dup_string(const std::string& s)
{
  char* p = (char*)malloc(s.size() + 1);
  if (p) memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

// ---------------------------------------------------------------------------
// init / devices
// ---------------------------------------------------------------------------
extern "C" int skl_init(int ndevices, int device_type)
{
  std::lock_guard<std::mutex> lk(g_mutex);
  if (g_init) return SKL_OK;
  (void)ndevices;
  if (device_type != SKL_DEVICE_ANY && device_type != SKL_DEVICE_GPU)
    return fail(SKL_ERR_NO_DEVICE, "device type %d is not available: only GPUs",
                device_type);
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return fail(SKL_ERR_NO_DEVICE, "no HIP GPU device found (%s)",
                hipGetErrorString(e));
  g_device = 0;
  SKL_HIP_TRY(hipSetDevice(g_device));
  SKL_HIP_TRY(hipGetDeviceProperties(&g_props, g_device));
  SKL_HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  SKL_TRY(noalloc_setup(g_device));
  g_init = true;
  logmsg(2, "init: %s (%s), %d CUs", g_props.name, g_props.gcnArchName,
      g_props.multiProcessorCount);
  return SKL_OK;
}

extern "C" int skl_terminate(void)
{
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_init) return SKL_OK;
  hipDeviceSynchronize();
  if (g_stream) hipStreamDestroy(g_stream);
  g_stream = nullptr;
  g_init = false;
  return SKL_OK;
}

extern "C" int skl_is_initialized(void) { return g_init ? 1 : 0; }

extern "C" int skl_device_count(int* count)
{
  if (!count) return fail(SKL_ERR_INVALID, "count is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c > 0 ? 1 : 0;  // the build drives one GPU
  return SKL_OK;
}

extern "C" int skl_device_get_info(int device, skl_device_info* info)
{
  SKL_TRY(ensure_initialized());
  if (device != 0)
    return fail(SKL_ERR_NO_DEVICE, "device %d does not exist (one GPU)", device);
  if (!info) return fail(SKL_ERR_INVALID, "info is NULL");
  memset(info, 0, sizeof(*info));
  snprintf(info->name, sizeof(info->name), "%s", g_props.name);
  if (!info->name[0]) {  // some runtimes leave the marketing name empty
    char nm[256] = {0};
    if (hipDeviceGetName(nm, (int)sizeof(nm) - 1, g_device) == hipSuccess && nm[0])
      snprintf(info->name, sizeof(info->name), "%s", nm);
    else
      snprintf(info->name, sizeof(info->name), "AMD %s", g_props.gcnArchName);
  }
  snprintf(info->arch, sizeof(info->arch), "%s", g_props.gcnArchName);
  info->compute_units = g_props.multiProcessorCount;
  info->wavefront_size = g_props.warpSize;
  info->max_work_group_size = g_props.maxThreadsPerBlock;
  info->lds_per_block = g_props.sharedMemPerBlock;
  info->global_mem = g_props.totalGlobalMem;
  info->clock_khz = g_props.clockRate;
  info->memory_clock_khz = g_props.memoryClockRate;
  info->memory_bus_width = g_props.memoryBusWidth;
  info->supports_double = 1;
  info->l2_cache_bytes = g_props.l2CacheSize;
  return SKL_OK;
}

extern "C" int skl_device_synchronize(void)
{
  SKL_TRY(ensure_initialized());
  SKL_HIP_TRY(hipDeviceSynchronize());
  return SKL_OK;
}

// ---------------------------------------------------------------------------
// streams
// ---------------------------------------------------------------------------
extern "C" int skl_stream_create(skl_stream* stream)
{
  SKL_TRY(ensure_initialized());
  if (!stream) return fail(SKL_ERR_INVALID, "stream is NULL");
  hipStream_t s;
  SKL_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = reinterpret_cast<skl_stream>(s);
  return SKL_OK;
}

extern "C" int skl_stream_destroy(skl_stream stream)
{
  if (!stream) return SKL_OK;
  SKL_HIP_TRY(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)));
  return SKL_OK;
}

extern "C" int skl_stream_synchronize(skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  SKL_HIP_TRY(hipStreamSynchronize(resolve_stream(stream)));
  return SKL_OK;
}

namespace {
struct CallbackBox {
  void (*fn)(void*);
  void* user;
};
void host_tramp(void* p)
{
  CallbackBox* b = static_cast<CallbackBox*>(p);
  b->fn(b->user);
  delete b;
}
}  // namespace

extern "C" int skl_stream_add_callback(skl_stream stream, void (*fn)(void*),
                                       void* user)
{
  SKL_TRY(ensure_initialized());
  if (!fn) return fail(SKL_ERR_INVALID, "callback is NULL");
  CallbackBox* b = new CallbackBox{fn, user};
  hipError_t e = hipLaunchHostFunc(resolve_stream(stream), host_tramp, b);
  if (e != hipSuccess) {
    delete b;
    return fail(SKL_ERR_HIP, "hipLaunchHostFunc: %s", hipGetErrorString(e));
  }
  return SKL_OK;
}

// ---------------------------------------------------------------------------
// buffers
// ---------------------------------------------------------------------------
extern "C" int skl_malloc(void** dptr, size_t bytes)
{
  SKL_TRY(ensure_initialized());
  if (!dptr) return fail(SKL_ERR_INVALID, "dptr is NULL");
  *dptr = nullptr;
  if (bytes == 0) return SKL_OK;
  SKL_HIP_TRY(hipMalloc(dptr, bytes));
  return SKL_OK;
}

extern "C" int skl_free(void* dptr)
{
  if (!dptr) return SKL_OK;
  SKL_HIP_TRY(hipFree(dptr));
  return SKL_OK;
}

extern "C" int skl_host_malloc(void** hptr, size_t bytes)
{
  SKL_TRY(ensure_initialized());
  if (!hptr) return fail(SKL_ERR_INVALID, "hptr is NULL");
  SKL_HIP_TRY(hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault));
  return SKL_OK;
}

extern "C" int skl_host_free(void* hptr)
{
  if (!hptr) return SKL_OK;
  SKL_HIP_TRY(hipHostFree(hptr));
  return SKL_OK;
}

extern "C" int skl_memcpy_h2d(void* dst, size_t dst_offset, const void* src,
                              size_t bytes, skl_stream stream, int blocking)
{
  SKL_TRY(ensure_initialized());
  if (bytes == 0) return SKL_OK;
  if (!dst || !src) return fail(SKL_ERR_INVALID, "NULL pointer in h2d copy");
  hipStream_t s = resolve_stream(stream);
  SKL_HIP_TRY(hipMemcpyAsync(static_cast<char*>(dst) + dst_offset, src, bytes,
                             hipMemcpyHostToDevice, s));
  if (blocking) SKL_HIP_TRY(hipStreamSynchronize(s));
  return SKL_OK;
}

extern "C" int skl_memcpy_d2h(void* dst, const void* src, size_t src_offset,
                              size_t bytes, skl_stream stream, int blocking)
{
  SKL_TRY(ensure_initialized());
  if (bytes == 0) return SKL_OK;
  if (!dst || !src) return fail(SKL_ERR_INVALID, "NULL pointer in d2h copy");
  hipStream_t s = resolve_stream(stream);
  SKL_HIP_TRY(hipMemcpyAsync(dst, static_cast<const char*>(src) + src_offset,
                             bytes, hipMemcpyDeviceToHost, s));
  if (blocking) SKL_HIP_TRY(hipStreamSynchronize(s));
  return SKL_OK;
}

extern "C" int skl_memcpy_d2d(void* dst, size_t dst_offset, const void* src,
                              size_t src_offset, size_t bytes,
                              skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (bytes == 0) return SKL_OK;
  if (!dst || !src) return fail(SKL_ERR_INVALID, "NULL pointer in d2d copy");
  SKL_HIP_TRY(hipMemcpyAsync(static_cast<char*>(dst) + dst_offset,
                             static_cast<const char*>(src) + src_offset, bytes,
                             hipMemcpyDeviceToDevice, resolve_stream(stream)));
  return SKL_OK;
}

extern "C" int skl_fill(void* dst, const void* pattern, size_t elem_size,
                        size_t count, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (count == 0) return SKL_OK;
  if (!dst || !pattern) return fail(SKL_ERR_INVALID, "NULL pointer in fill");
  if (elem_size != 1 && elem_size != 2 && elem_size != 4 && elem_size != 8 &&
      elem_size != 16)
    return fail(SKL_ERR_INVALID, "fill element size %zu not in {1,2,4,8,16}",
                elem_size);
  return skl_internal_fill(dst, pattern, elem_size, count,
                           resolve_stream(stream));
}

// ---------------------------------------------------------------------------
// events
// ---------------------------------------------------------------------------
extern "C" int skl_event_create(skl_event* ev)
{
  SKL_TRY(ensure_initialized());
  if (!ev) return fail(SKL_ERR_INVALID, "ev is NULL");
  hipEvent_t e;
  SKL_HIP_TRY(hipEventCreate(&e));
  *ev = reinterpret_cast<skl_event>(e);
  return SKL_OK;
}

extern "C" int skl_event_destroy(skl_event ev)
{
  if (!ev) return SKL_OK;
  SKL_HIP_TRY(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
  return SKL_OK;
}

extern "C" int skl_event_record(skl_event ev, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  SKL_HIP_TRY(
      hipEventRecord(reinterpret_cast<hipEvent_t>(ev), resolve_stream(stream)));
  return SKL_OK;
}

extern "C" int skl_event_synchronize(skl_event ev)
{
  SKL_HIP_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(ev)));
  return SKL_OK;
}

extern "C" int skl_event_elapsed_ms(skl_event start, skl_event stop, float* ms)
{
  if (!ms) return fail(SKL_ERR_INVALID, "ms is NULL");
  SKL_HIP_TRY(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                  reinterpret_cast<hipEvent_t>(stop)));
  return SKL_OK;
}

// ---------------------------------------------------------------------------
// programs: hiprtc + code-object cache (Program.cpp:83-281)
// ---------------------------------------------------------------------------
struct skl_program_s {
  hipModule_t module = nullptr;
  std::string hash;
  std::map<std::string, hipFunction_t> kernels;
  int refs = 1;
};

struct skl_kernel_s {
  hipFunction_t fn;
};

namespace {
std::mutex g_prog_mutex;
std::map<std::string, skl_program_s*> g_programs;  // hash -> loaded module
std::map<std::string, std::unique_ptr<skl_kernel_s>> g_kernel_handles;

bool env_is(const char* name, const char* value)
{
  const char* e = getenv(name);
  return e && strcmp(e, value) == 0;
}

std::string cache_dir()
{
  const char* e = getenv("SKELCL_CACHE_DIR");
  if (e && *e) return e;
  const char* home = getenv("HOME");
  std::string base = home && *home ? std::string(home) + "/.cache" : "/tmp";
  return base + "/skelcl_amd";
}

void mkdirs(const std::string& path)
{
  std::string cur;
  for (size_t i = 0; i < path.size(); ++i) {
    cur += path[i];
    if (path[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
  }
  mkdir(path.c_str(), 0755);
}

std::vector<std::string> split_options(const char* options)
{
  std::vector<std::string> out;
  if (!options) return out;
  std::istringstream is(options);
  std::string tok;
  while (is >> tok) out.push_back(tok);
  return out;
}

int rtc_compile(const std::string& src, const char* options,
                std::string* code, std::string* log)
{
  hiprtcProgram prog;
  hiprtcResult r =
      hiprtcCreateProgram(&prog, src.c_str(), "skelcl_kernel.hip", 0, nullptr,
                          nullptr);
  if (r != HIPRTC_SUCCESS)
    return fail(SKL_ERR_COMPILE, "hiprtcCreateProgram: %s",
                hiprtcGetErrorString(r));
  std::vector<std::string> opts = {"--offload-arch=" + target_arch(), "-O3",
                                   "-std=c++17", "-fwrapv",
                                   // packed FP32 instructions stay off
                                   "-Xclang", "-target-feature", "-Xclang",
                                   "-packed-fp32-ops"};
  for (auto& o : split_options(options)) opts.push_back(o);
  std::vector<const char*> copts;
  for (auto& o : opts) copts.push_back(o.c_str());
  r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  size_t logsz = 0;
  hiprtcGetProgramLogSize(prog, &logsz);
  std::string lg(logsz, '\0');
  if (logsz) hiprtcGetProgramLog(prog, &lg[0]);
  while (!lg.empty() && lg.back() == '\0') lg.pop_back();
  if (log) *log = lg;
  if (env_is("SKELCL_PRINT_BUILD_LOG", "YES") && !lg.empty())
    fprintf(stderr, "[skelcl] build log:\n%s\n", lg.c_str());
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(SKL_ERR_COMPILE, "hiprtc compile failed: %s\n%s",
                hiprtcGetErrorString(r), lg.c_str());
  }
  size_t sz = 0;
  hiprtcGetCodeSize(prog, &sz);
  code->assign(sz, '\0');
  hiprtcGetCode(prog, &(*code)[0]);
  hiprtcDestroyProgram(&prog);
  return SKL_OK;
}

// Code-object cache on disk, as the reference's binary cache: a code object
// is loaded only when SKELCL_LOAD_BINARY is "YES" and saved only when
// SKELCL_SAVE_BINARY is "YES" (Program.cpp:139,256), under
// $SKELCL_CACHE_DIR/<key>.co.  The key covers the source, the options, the
// target and the hiprtc version, so a toolchain change never loads a stale
// object.
std::string program_key(const char* source, const char* options)
{
  int major = 0, minor = 0;
  hiprtcVersion(&major, &minor);
  return sha1_hex(std::string(source) + "\n//options: " + (options ? options : "") +
                  "\n//arch: " + target_arch() + "\n//hiprtc: " + std::to_string(major) + "." +
                  std::to_string(minor) + "\n//rt: 2");
}

bool cache_load(const std::string& key, std::string* code)
{
  if (!env_is("SKELCL_LOAD_BINARY", "YES")) return false;
  const std::string path = cache_dir() + "/" + key + ".co";
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  *code = ss.str();
  if (code->empty()) return false;
  logmsg(3, "loaded code object %s", path.c_str());
  return true;
}

void cache_save(const std::string& key, const std::string& code)
{
  if (!env_is("SKELCL_SAVE_BINARY", "YES")) return;
  mkdirs(cache_dir());
  const std::string path = cache_dir() + "/" + key + ".co";
  const std::string tmp = path + ".tmp" + std::to_string(getpid());
  std::ofstream f(tmp, std::ios::binary);
  if (!f) return;
  f.write(code.data(), (std::streamsize)code.size());
  f.close();
  if (rename(tmp.c_str(), path.c_str()) != 0) unlink(tmp.c_str());
  else logmsg(3, "saved code object %s", path.c_str());
}

// Compiles `source`, or takes it from the disk cache; `from_cache` tells
// which.
int compile_cached(const char* source, const char* options, const std::string& key,
                   std::string* code, std::string* log, bool* from_cache)
{
  *from_cache = cache_load(key, code);
  if (*from_cache) return SKL_OK;
  SKL_TRY(rtc_compile(source, options, code, log));
  cache_save(key, *code);
  return SKL_OK;
}
}  // namespace

extern "C" int skl_program_compile(const char* source, const char* options,
                                   char** code, size_t* code_size, char** log)
{
  if (!source || !*source)
    return fail(SKL_ERR_INVALID,
                "Tried to create program with empty user source.");
  std::string c, lg;
  bool cached = false;
  int st = compile_cached(source, options, program_key(source, options), &c, &lg, &cached);
  if (log) *log = dup_string(lg);
  if (st != SKL_OK) return st;
  if (code) {
    *code = (char*)malloc(c.size() ? c.size() : 1);
    memcpy(*code, c.data(), c.size());
  }
  if (code_size) *code_size = c.size();
  return SKL_OK;
}

extern "C" int skl_program_build(const char* source, const char* options,
                                 skl_program* program, char** log)
{
  SKL_TRY(ensure_initialized());
  if (!program) return fail(SKL_ERR_INVALID, "program is NULL");
  if (!source || !*source)
    return fail(SKL_ERR_INVALID,
                "Tried to create program with empty user source.");
  if (log) *log = nullptr;
  const std::string key = program_key(source, options);
  {
    std::lock_guard<std::mutex> lk(g_prog_mutex);
    auto it = g_programs.find(key);
    if (it != g_programs.end()) {
      it->second->refs++;
      *program = it->second;
      return SKL_OK;
    }
  }
  std::string code, lg;
  bool cached = false;
  int st = compile_cached(source, options, key, &code, &lg, &cached);
  if (log) *log = dup_string(lg);
  if (st != SKL_OK) return st;
  hipModule_t mod;
  hipError_t e = hipModuleLoadData(&mod, code.data());
  if (e != hipSuccess && cached) {
    // a cached object this runtime rejects: compile afresh (and replace it)
    logmsg(1, "cached code object %s rejected (%s): recompiling", key.c_str(),
           hipGetErrorString(e));
    code.clear();
    lg.clear();
    st = rtc_compile(source, options, &code, &lg);
    if (log) {
      if (*log) free(*log);
      *log = dup_string(lg);
    }
    if (st != SKL_OK) return st;
    cache_save(key, code);
    e = hipModuleLoadData(&mod, code.data());
  }
  if (e != hipSuccess)
    return fail(SKL_ERR_HIP, "hipModuleLoadData: %s", hipGetErrorString(e));
  auto* p = new skl_program_s;
  p->module = mod;
  p->hash = key;
  {
    std::lock_guard<std::mutex> lk(g_prog_mutex);
    auto it = g_programs.find(key);
    if (it != g_programs.end()) {  // raced with another builder
      hipModuleUnload(mod);
      delete p;
      it->second->refs++;
      *program = it->second;
      return SKL_OK;
    }
    g_programs[key] = p;
  }
  *program = p;
  return SKL_OK;
}

extern "C" int skl_program_release(skl_program program)
{
  if (!program) return SKL_OK;
  std::lock_guard<std::mutex> lk(g_prog_mutex);
  // Modules stay loaded for the life of the process (kernels of released
  // programs may still be queued); only the reference count drops.
  if (program->refs > 0) program->refs--;
  return SKL_OK;
}

extern "C" int skl_kernel_get(skl_program program, const char* name,
                              skl_kernel* kernel)
{
  if (!program || !name || !kernel)
    return fail(SKL_ERR_INVALID, "NULL argument to skl_kernel_get");
  std::lock_guard<std::mutex> lk(g_prog_mutex);
  std::string k = program->hash + "/" + name;
  auto it = g_kernel_handles.find(k);
  if (it != g_kernel_handles.end()) {
    *kernel = it->second.get();
    return SKL_OK;
  }
  hipFunction_t fn;
  hipError_t e = hipModuleGetFunction(&fn, program->module, name);
  if (e != hipSuccess)
    return fail(SKL_ERR_INVALID, "kernel %s not found: %s", name,
                hipGetErrorString(e));
  auto h = std::make_unique<skl_kernel_s>();
  h->fn = fn;
  *kernel = h.get();
  g_kernel_handles[k] = std::move(h);
  return SKL_OK;
}

extern "C" int skl_kernel_launch(skl_kernel kernel, const unsigned int grid[3],
                                 const unsigned int block[3],
                                 size_t dynamic_lds, void** args,
                                 skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!kernel || !grid || !block)
    return fail(SKL_ERR_INVALID, "NULL argument to skl_kernel_launch");
  if ((size_t)block[0] * block[1] * block[2] >
      (size_t)g_props.maxThreadsPerBlock)
    return fail(SKL_ERR_INVALID, "work-group of %u threads exceeds %d",
                block[0] * block[1] * block[2], g_props.maxThreadsPerBlock);
  if (grid[0] == 0 || grid[1] == 0 || grid[2] == 0) return SKL_OK;
  SKL_HIP_TRY(hipModuleLaunchKernel(kernel->fn, grid[0], grid[1], grid[2],
                                    block[0], block[1], block[2],
                                    (unsigned)dynamic_lds,
                                    resolve_stream(stream), args, nullptr));
  return SKL_OK;
}

extern "C" int skl_sha1_hex(const void* data, size_t len, char out[41])
{
  if (!out || (!data && len)) return fail(SKL_ERR_INVALID, "NULL argument");
  std::string h =
      sha1_hex(std::string(static_cast<const char*>(data ? data : ""), len));
  memcpy(out, h.c_str(), 41);
  return SKL_OK;
}

// ---------------------------------------------------------------------------
// graphs
// ---------------------------------------------------------------------------
struct skl_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int skl_graph_begin_capture(skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!stream)
    return fail(SKL_ERR_INVALID, "graph capture needs a created stream");
  SKL_HIP_TRY(hipStreamBeginCapture(resolve_stream(stream),
                                    hipStreamCaptureModeThreadLocal));
  return SKL_OK;
}

extern "C" int skl_graph_end_capture(skl_stream stream, skl_graph* graph)
{
  SKL_TRY(ensure_initialized());
  if (!graph) return fail(SKL_ERR_INVALID, "graph is NULL");
  auto* g = new skl_graph_s;
  hipError_t e = hipStreamEndCapture(resolve_stream(stream), &g->graph);
  if (e != hipSuccess) {
    delete g;
    return fail(SKL_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g->graph);
    delete g;
    return fail(SKL_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  *graph = g;
  return SKL_OK;
}

extern "C" int skl_graph_launch(skl_graph graph, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!graph) return fail(SKL_ERR_INVALID, "graph is NULL");
  SKL_HIP_TRY(hipGraphLaunch(graph->exec, resolve_stream(stream)));
  return SKL_OK;
}

extern "C" int skl_graph_destroy(skl_graph graph)
{
  if (!graph) return SKL_OK;
  if (graph->exec) hipGraphExecDestroy(graph->exec);
  if (graph->graph) hipGraphDestroy(graph->graph);
  delete graph;
  return SKL_OK;
}

// ---------------------------------------------------------------------------
// synthetic inputs
// ---------------------------------------------------------------------------
extern "C" int skl_generate_f32_sym(float* out, size_t n, uint64_t seed,
                                    skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  return skl_internal_gen_f32_sym(out, n, seed, resolve_stream(stream));
}

extern "C" int skl_generate_f32_unit(float* out, size_t n, uint64_t seed,
                                     float scale, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  return skl_internal_gen_f32_unit(out, n, seed, scale, resolve_stream(stream));
}

extern "C" int skl_generate_i32_below(int32_t* out, size_t n, uint64_t seed,
                                      uint32_t m, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  return skl_internal_gen_i32_below(out, n, seed, m, resolve_stream(stream));
}

// Autotuner: times candidate launch shapes of a skeleton on scratch buffers
// and keeps the fastest, cached per (kind, user source, types, log2 size) in
// memory and in a text file in the code-object cache directory.  A shape is
// the work-group size and the per-kind parameters: packs in flight per
// thread and grid cap (Map/Zip/Reduce), packs per lane (Scan), rows per
// thread (MapOverlap), the MFMA work-group tile (AllPairs (+,*) products).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "common.h"
#include "skeleton.h"
#include "skelcl_skeletons.h"

using namespace skl;

namespace {

struct Shape {
  int unroll, blocks_per_cu, scan_j, mo_rb, wg, tile;
};

std::mutex g_mutex;
std::map<std::string, Shape> g_cache;
bool g_loaded = false;

std::string tune_path()
{
  const char* e = getenv("SKELCL_CACHE_DIR");
  std::string dir;
  if (e && *e) dir = e;
  else {
    const char* home = getenv("HOME");
    dir = std::string(home && *home ? home : "/tmp") + "/.cache/skelcl_amd";
  }
  return dir + "/autotune_v2.txt";
}

void load_cache()
{
  if (g_loaded) return;
  g_loaded = true;
  std::ifstream f(tune_path());
  std::string key;
  Shape s;
  while (f >> key >> s.unroll >> s.blocks_per_cu >> s.scan_j >> s.mo_rb >> s.wg >> s.tile)
    g_cache[key] = s;
}

void save_entry(const std::string& key, const Shape& s)
{
  std::ofstream f(tune_path(), std::ios::app);
  if (f)
    f << key << " " << s.unroll << " " << s.blocks_per_cu << " " << s.scan_j << " " << s.mo_rb
      << " " << s.wg << " " << s.tile << "\n";
}

std::string tune_key(const skl_skeleton_s* s, size_t n)
{
  std::ostringstream o;
  o << s->kind << "|" << s->src << "|" << s->src2 << "|" << s->identity << "|" << s->precision
    << "|" << s->overlap << "|" << s->padding << "|" << s->options;
  for (auto& t : s->types) o << "|" << t.name;
  int cls = n ? (int)std::floor(std::log2((double)n)) : 0;
  return sha1_hex(o.str()) + ":" + std::to_string(cls) + ":" + device_props().gcnArchName;
}

Shape current(const skl_skeleton_s* s)
{
  int cus = device_props().multiProcessorCount > 0 ? device_props().multiProcessorCount : 256;
  return Shape{s->unroll, s->max_blocks / cus, s->scan_j, s->mo_rb, s->wg, s->mm_tile};
}

void apply(skl_skeleton_s* s, const Shape& sh)
{
  int cus = device_props().multiProcessorCount > 0 ? device_props().multiProcessorCount : 256;
  bool changed = sh.unroll != s->unroll || sh.scan_j != s->scan_j || sh.mo_rb != s->mo_rb ||
                 sh.wg != s->wg;
  s->unroll = sh.unroll;
  s->scan_j = sh.scan_j;
  s->mo_rb = sh.mo_rb;
  s->wg = sh.wg;
  s->mm_tile = sh.tile;  // chosen at launch: nothing to rebuild
  s->max_blocks = cus * sh.blocks_per_cu;
  // the shape is compiled into the kernels; the native stencil kernel takes
  // its rows per thread at launch and never runs the generated one
  if (changed && s->stencil_radius < 0) rebuild_programs(s);
}

std::vector<Shape> candidates(const skl_skeleton_s* s)
{
  Shape d = current(s);
  std::vector<Shape> c;
  switch (s->kind) {
    case SKL_MAP: case SKL_ZIP: case SKL_REDUCE:
      // small grids keep 24-32 KB of loads in flight per CU with few
      // waves, where HBM streams peak on MI355X; larger ones suit short inputs
      for (int w : {256, 512})
        for (int u : {1, 2, 4, 8})
          for (int b : {1, 2, 3, 4, 8, 16, 64})
            c.push_back(Shape{u, b, d.scan_j, d.mo_rb, w, d.tile});
      break;
    case SKL_SCAN:
      for (int w : {256, 512, 1024})
        for (int j : {2, 4, 8, 16})
          c.push_back(Shape{d.unroll, d.blocks_per_cu, j, d.mo_rb, w, d.tile});
      break;
    case SKL_MAPOVERLAP:  // 64 x 4 work-groups: the tile layout fixes the size
      for (int rb : {1, 2, 4, 6, 8})
        c.push_back(Shape{d.unroll, d.blocks_per_cu, d.scan_j, rb, d.wg, d.tile});
      break;
    case SKL_ALLPAIRS:
      // the MFMA work-group tiles of the (+,*) products (skl_internal_matmul)
      if (s->precision == SKL_PREC_FP8) {
        for (int t : {1, 2, 3}) c.push_back(Shape{d.unroll, d.blocks_per_cu, d.scan_j, d.mo_rb, d.wg, t});
      } else if (s->precision == SKL_PREC_BF16X3) {
        for (int t : {1, 2}) c.push_back(Shape{d.unroll, d.blocks_per_cu, d.scan_j, d.mo_rb, d.wg, t});
      } else {
        c.push_back(d);
      }
      break;
    default:
      c.push_back(d);
  }
  return c;
}

struct Buf {
  void* p = nullptr;
  ~Buf()
  {
    if (p) hipFree(p);
  }
};

int alloc(Buf& b, size_t bytes)
{
  SKL_HIP_TRY(hipMalloc(&b.p, bytes ? bytes : 16));
  SKL_HIP_TRY(hipMemset(b.p, 0, bytes ? bytes : 16));
  return SKL_OK;
}

}  // namespace

static int tune(skl_skeleton s, size_t n, const skl_arg* args, int nargs)
{
  SKL_TRY(ensure_initialized());
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (n == 0) return fail(SKL_ERR_INVALID, "autotune size must be positive");
  if (s->kind == SKL_ALLPAIRS_GENERAL || s->kind == SKL_MAP_INDEX ||
      s->kind == SKL_MAP_INDEXPOINT)
    return SKL_OK;  // fixed shapes
  if (s->simple) return SKL_OK;  // one element per work-item: nothing to tune
  const std::vector<Shape> cands = candidates(s);
  if (cands.size() < 2) return SKL_OK;  // nothing to choose
  const std::string key = tune_key(s, n);
  {
    std::lock_guard<std::mutex> lk(g_mutex);
    load_cache();
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      apply(s, it->second);
      return SKL_OK;
    }
  }
  size_t es = 0;  // the widest element type
  for (const TypeInfo& t : s->types) es = std::max(es, t.size);
  size_t rows = 0, cols = 0;
  if (s->kind == SKL_MAPOVERLAP || s->kind == SKL_ALLPAIRS) {
    // an n-pixel image / an n-element product of square operands
    rows = cols = (size_t)std::ceil(std::sqrt((double)n));
    n = rows * cols;
  }
  Buf in, in2, out;
  SKL_TRY(alloc(in, n * es));
  SKL_TRY(alloc(in2, n * es));
  SKL_TRY(alloc(out, n * es + 64));
  hipStream_t st = resolve_stream(nullptr);
  {
    // the clocks depend on the data (zeros draw less power): time on inputs
    // like real ones where the element type allows (uniform [-1, 1) floats,
    // ints in [0, 1000)); other types stay zero
    const std::string& t1 = s->types.size() > 1 ? s->types[1].name : s->types[0].name;
    const std::string& t2 = s->types.size() > 2 ? s->types[2].name : t1;
    auto gen = [&](void* p, const std::string& t, uint64_t seed) -> int {
      if (t == "float") return skl_generate_f32_sym((float*)p, n, seed, nullptr);
      if (t == "int" || t == "unsigned int")
        return skl_generate_i32_below((int*)p, n, seed, 1000, nullptr);
      return SKL_OK;
    };
    SKL_TRY(gen(in.p, t1, 0x7A11));
    SKL_TRY(gen(in2.p, t2, 0x7A12));
  }
  auto run = [&]() -> int {
    switch (s->kind) {
      case SKL_MAP: return skl_map_run(s, out.p, in.p, n, args, nargs, nullptr);
      case SKL_ZIP: return skl_zip_run(s, out.p, in.p, in2.p, n, args, nargs, nullptr);
      case SKL_REDUCE: return skl_reduce_run(s, out.p, in.p, n, args, nargs, nullptr);
      case SKL_SCAN: return skl_scan_run(s, out.p, in.p, n, 1, args, nargs, nullptr);
      case SKL_MAPOVERLAP:
        return skl_mapoverlap_run(s, out.p, in.p, rows, cols, args, nargs, nullptr);
      case SKL_ALLPAIRS:
        return skl_allpairs_run(s, out.p, in.p, in2.p, rows, cols, rows, args, nargs, nullptr);
    }
    return SKL_OK;
  };
  hipEvent_t e0, e1;
  SKL_HIP_TRY(hipEventCreate(&e0));
  SKL_HIP_TRY(hipEventCreate(&e1));
  auto time_shape = [&](const Shape& c, int reps, float* ms) -> int {
    apply(s, c);
    int stt = run();  // compiles this shape, warms up
    if (stt != SKL_OK) return stt;
    SKL_HIP_TRY(hipStreamSynchronize(st));
    SKL_HIP_TRY(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) SKL_TRY(run());
    SKL_HIP_TRY(hipEventRecord(e1, st));
    SKL_HIP_TRY(hipEventSynchronize(e1));
    SKL_HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    logmsg(2, "autotune %s: wg %d unroll %d blocks/CU %d scan_j %d rb %d tile %d -> %.4f ms",
           key.c_str(), c.wg, c.unroll, c.blocks_per_cu, c.scan_j, c.mo_rb, c.tile, *ms);
    return SKL_OK;
  };
  auto same = [](const Shape& x, const Shape& y) {
    return x.unroll == y.unroll && x.blocks_per_cu == y.blocks_per_cu && x.scan_j == y.scan_j &&
           x.mo_rb == y.mo_rb && x.wg == y.wg && x.tile == y.tile;
  };
  // pass 1: every candidate, 5 runs.  Pass 2: the three fastest and the
  // shape in use when tuning started, 20 runs each.  Shapes within a few
  // per cent of one another swap places from run to run, so the shape in
  // use stays unless another is at least 2% faster in pass 2.
  const Shape incumbent = current(s);
  {
    // back-to-back runs of the shape in use for at least 50 ms of device
    // time first: under a sudden load the clocks dip for several ms before
    // they settle, and the candidates are compared on a settled device
    float done = 0.0f;
    for (int i = 0, batch = 1; done < 50.0f && i < 256; ++i, batch = batch < 32 ? 2 * batch : 32) {
      float ms = 0.0f;
      if (time_shape(incumbent, batch, &ms) != SKL_OK) break;  // skipped, as in the passes below
      done += ms * (float)batch;
    }
  }
  std::vector<std::pair<float, Shape>> scored;
  for (const Shape& c : cands) {
    float ms = 0;
    if (time_shape(c, 5, &ms) == SKL_OK) scored.push_back({ms, c});
  }
  std::sort(scored.begin(), scored.end(),
            [](const std::pair<float, Shape>& x, const std::pair<float, Shape>& y) {
              return x.first < y.first;
            });
  std::vector<Shape> fin;
  for (size_t i = 0; i < scored.size() && i < 3; ++i) fin.push_back(scored[i].second);
  bool have_inc = false;
  for (const Shape& f : fin) have_inc |= same(f, incumbent);
  if (!have_inc) fin.push_back(incumbent);
  Shape best = incumbent;
  float best_ms = 1e30f, inc_ms = 1e30f;
  for (const Shape& f : fin) {
    float ms = 0;
    if (time_shape(f, 20, &ms) != SKL_OK) continue;
    if (same(f, incumbent)) inc_ms = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = f;
    }
  }
  if (inc_ms < 1e30f && best_ms > 0.98f * inc_ms) best = incumbent;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  apply(s, best);
  std::lock_guard<std::mutex> lk(g_mutex);
  g_cache[key] = best;
  save_entry(key, best);
  return SKL_OK;
}

extern "C" int skl_autotune(skl_skeleton s, size_t n, int enable)
{
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (!enable) {
    int cus = device_props().multiProcessorCount > 0 ? device_props().multiProcessorCount : 256;
    skl_skeleton_s d;
    d.kind = s->kind;
    d.types = s->types;
    d.op = s->op;
    d.precision = s->precision;
    d.V = s->V;
    d.extras = s->extras;
    d.stencil_radius = s->stencil_radius;
    d.apply_default_shape(cus);
    d.wg = s->kind == SKL_SCAN ? 512 : 256;  // skl_skeleton_create's defaults
    apply(s, Shape{d.unroll, d.max_blocks / cus, d.scan_j, d.mo_rb, d.wg, 0});
    return SKL_OK;
  }
  if (!s->extras.empty())
    return fail(SKL_ERR_INVALID,
                "the user function takes extra arguments: use skl_autotune_args");
  return tune(s, n, nullptr, 0);
}

extern "C" int skl_autotune_args(skl_skeleton s, size_t n, const skl_arg* args, int nargs)
{
  return tune(s, n, args, nargs);
}

extern "C" int skl_skeleton_set_launch_shape(skl_skeleton s, const int shape[4])
{
  if (!s || !shape) return fail(SKL_ERR_INVALID, "NULL argument");
  if (shape[0] < 1 || shape[0] > 16 || shape[1] < 1 || shape[1] > 1024 || shape[2] < 1 ||
      shape[2] > 32 || shape[3] < 1 || shape[3] > 16)
    return fail(SKL_ERR_INVALID, "launch shape out of range");
  Shape c = current(s);
  c.unroll = shape[0];
  c.blocks_per_cu = shape[1];
  c.scan_j = shape[2];
  c.mo_rb = shape[3];
  apply(s, c);
  return SKL_OK;
}

extern "C" int skl_skeleton_launch_shape(skl_skeleton s, int shape[4])
{
  if (!s || !shape) return fail(SKL_ERR_INVALID, "NULL argument");
  Shape c = current(s);
  shape[0] = c.unroll;
  shape[1] = c.blocks_per_cu;
  shape[2] = c.scan_j;
  shape[3] = c.mo_rb;
  return SKL_OK;
}

extern "C" int skl_skeleton_set_mma_tile(skl_skeleton s, int tile)
{
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (tile < 0 || tile > 3) return fail(SKL_ERR_INVALID, "MFMA tile %d outside [0, 3]", tile);
  s->mm_tile = tile;
  return SKL_OK;
}

extern "C" int skl_skeleton_mma_tile(skl_skeleton s, int* tile)
{
  if (!s || !tile) return fail(SKL_ERR_INVALID, "NULL argument");
  *tile = s->mm_tile;
  return SKL_OK;
}

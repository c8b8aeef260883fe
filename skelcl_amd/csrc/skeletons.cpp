// Skeleton objects: user-function fusion into the kernel templates and the
// launches of Map / Zip / Reduce / Scan / MapOverlap / AllPairs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <sstream>

#include "common.h"
#include "fusion.h"
#include "skeleton.h"
#include "skelcl_skeletons.h"

// kernel templates embedded at build time (kernels/*.hip)
extern "C" const char* skl_tmpl_prelude;
extern "C" const char* skl_tmpl_map;
extern "C" const char* skl_tmpl_reduce;
extern "C" const char* skl_tmpl_scan;
extern "C" const char* skl_tmpl_mapoverlap;
extern "C" int skl_internal_stencil_max_radius();
extern "C" int skl_internal_stencil_forward(float* out, const float* in, const float* w,
                                            int rows, int cols, int R, int neutral_pad,
                                            float neutral, int fused, int rb, int cb,
                                            hipStream_t s);
extern "C" const char* skl_tmpl_allpairs;

using namespace skl;

// ---- common definitions (CommonDefinitions, src/Source.cpp:102-142) --------
namespace {
struct CommonDefs {
  std::mutex mu;
  std::vector<std::string> level[3];
};
CommonDefs& common_defs()
{
  static CommonDefs d;  // registered from static initialisers: built on first use
  return d;
}
std::string common_text()
{
  CommonDefs& d = common_defs();
  std::lock_guard<std::mutex> lk(d.mu);
  std::string t;
  for (auto& lv : d.level)
    for (auto& x : lv) t += x + "\n";
  return t;
}
}  // namespace

extern "C" int skl_common_definition_append(const char* text, int level)
{
  if (!text) return fail(SKL_ERR_INVALID, "common definition is NULL");
  if (level < 0 || level > 2) return fail(SKL_ERR_INVALID, "unknown definition level %d", level);
  CommonDefs& d = common_defs();
  std::lock_guard<std::mutex> lk(d.mu);
  d.level[level].push_back(text);
  return SKL_OK;
}

extern "C" int skl_common_definitions(char** text)
{
  if (!text) return fail(SKL_ERR_INVALID, "text is NULL");
  const std::string t = common_text();
  *text = (char*)malloc(t.size() + 1);
  memcpy(*text, t.c_str(), t.size() + 1);
  return SKL_OK;
}

bool skl_skeleton_s::float_add() const
{
  return op == UserOp::Add && extras.empty() && !types.empty() && types[0].name == "float";
}

bool skl_skeleton_s::precise_add() const { return precision == SKL_PREC_DEFAULT && float_add(); }

bool skl_skeleton_s::fp64_add() const { return precision == SKL_PREC_FP64 && float_add(); }

void skl_skeleton_s::apply_default_shape(int cus)
{
  // measured on MI355X at 2^28 elements: the streaming Reduce is fastest
  // with one work-group per CU and 8 packs per thread per stride (16 in
  // flight with the next stride's loads); Map/Zip with 8 per CU and 4 packs
  if (kind == SKL_REDUCE) {
    unroll = 8;
    max_blocks = cus;
  } else {
    unroll = 4;
    max_blocks = cus * 8;
  }
  if (kind == SKL_SCAN) scan_j = default_scan_j();
  // MapOverlap rows per thread: the native stencil kernel is fastest at 8
  // (8192^2, radius 5: 0.284 ms against 0.293 at 6 and 0.336 at 4)
  mo_rb = stencil_radius >= 0 ? 8 : 4;
}

int skl_skeleton_s::default_scan_j() const
{
  // float (+) Scan, FP64 prefixes: 16 packs per lane measured fastest (2^28
  // and 577.5 M elements, forward and reversed) although the FP64 register
  // tile is twice the float one.  Double-float pairs: 8 (256 bytes).
  if (fp64_add() && kind == SKL_SCAN) return 16;
  const size_t acc = precise_add() || fp64_add() ? 2 * sizeof(float)
                                                 : (types.empty() ? 4 : types[0].size);
  const size_t per_pack = (size_t)(V > 0 ? V : 1) * (acc > 0 ? acc : 1);
  size_t j = 256 / per_pack;
  return (int)(j < 2 ? 2 : (j > 16 ? 16 : j));
}

namespace {

std::string str_or(const char* s, const char* dflt) { return s && *s ? s : dflt; }

const char* kind_name(int k)
{
  switch (k) {
    case SKL_MAP: return "Map";
    case SKL_MAP_INDEX: return "Map<Index>";
    case SKL_MAP_INDEXPOINT: return "Map<IndexPoint>";
    case SKL_ZIP: return "Zip";
    case SKL_REDUCE: return "Reduce";
    case SKL_SCAN: return "Scan";
    case SKL_MAPOVERLAP: return "MapOverlap";
    case SKL_ALLPAIRS: return "AllPairs";
    case SKL_ALLPAIRS_GENERAL: return "AllPairs(general)";
  }
  return "?";
}

// Number of leading (skeleton-supplied) parameters of the user function.
int leading_params(int kind)
{
  switch (kind) {
    case SKL_MAP: case SKL_MAP_INDEX: case SKL_MAP_INDEXPOINT: case SKL_MAPOVERLAP:
      return 1;
    case SKL_ZIP: case SKL_REDUCE: case SKL_SCAN: case SKL_ALLPAIRS:
      return 2;
    case SKL_ALLPAIRS_GENERAL:
      return 3;
  }
  return 1;
}

int expected_types(int kind)
{
  switch (kind) {
    case SKL_MAP: case SKL_MAPOVERLAP: return 2;
    case SKL_MAP_INDEX: case SKL_MAP_INDEXPOINT: case SKL_REDUCE: case SKL_SCAN: return 1;
    case SKL_ZIP: case SKL_ALLPAIRS: case SKL_ALLPAIRS_GENERAL: return 3;
  }
  return 0;
}

int cu_count()
{
  int cus = device_props().multiProcessorCount;
  return cus > 0 ? cus : 256;
}

// 16-byte packs: V elements of the widest of the given types
int pack_width(std::initializer_list<size_t> sizes)
{
  size_t m = 0;
  for (size_t s : sizes) m = std::max(m, s);
  if (m == 0 || m > 16 || 16 % m != 0) return 1;
  return (int)(16 / m);
}

bool aligned(const void* p, size_t a) { return a <= 1 || (reinterpret_cast<uintptr_t>(p) % a) == 0; }

std::string rename_wrap(const std::string& src, const std::string& from, const std::string& to)
{
  std::ostringstream o;
  // OpenCL C functions are device functions: compile the user source as
  // __host__ __device__ (clang's force_cuda_host_device pragma)
  o << "#define " << from << " " << to << "\n#pragma clang force_cuda_host_device begin\n"
    << "#line 1 \"user_source\"\n" << src << "\n#pragma clang force_cuda_host_device end\n"
    << "#undef " << from << "\n";
  return o.str();
}

// ---- kernel-argument marshalling (KernelUtil.h:134-245) -------------------
struct ArgPack {
  std::deque<std::vector<char>> store;
  std::vector<void*> ptrs;
  size_t dyn_lds = 0;

  void push_raw(const void* p, size_t n)
  {
    store.emplace_back((const char*)p, (const char*)p + n);
    ptrs.push_back(store.back().data());
  }
  template <typename T> void push(const T& v) { push_raw(&v, sizeof(T)); }
};

int marshal_extras(const std::vector<Param>& extras, const skl_arg* args, int nargs,
                   ArgPack* pk, size_t lds_base = 0)
{
  if (nargs < 0 || (nargs > 0 && !args))
    return fail(SKL_ERR_INVALID, "invalid extra argument list");
  if ((size_t)nargs != extras.size())
    return fail(SKL_ERR_INVALID,
                "the user function takes %zu extra argument(s) but %d were given",
                extras.size(), nargs);
  size_t lds = lds_base;
  for (int i = 0; i < nargs; ++i) {
    const Param& p = extras[i];
    const skl_arg& a = args[i];
    switch (p.kind) {
      case ParamKind::Value:
        if (a.kind != SKL_ARG_VALUE || !a.value)
          return fail(SKL_ERR_INVALID, "extra argument %d ('%s') must be a value", i,
                      p.decl.c_str());
        if (p.value_size && a.size != p.value_size)
          return fail(SKL_ERR_INVALID,
                      "extra argument %d ('%s') needs %zu bytes but %zu were given", i,
                      p.decl.c_str(), p.value_size, a.size);
        pk->push_raw(a.value, a.size);
        break;
      case ParamKind::Buffer:
        if (a.kind != SKL_ARG_BUFFER)
          return fail(SKL_ERR_INVALID, "extra argument %d ('%s') must be a buffer", i,
                      p.decl.c_str());
        pk->push(a.value);
        break;
      case ParamKind::Matrix: {
        if (a.kind != SKL_ARG_MATRIX)
          return fail(SKL_ERR_INVALID, "extra argument %d ('%s') must be a matrix", i,
                      p.decl.c_str());
        struct {
          const void* data;
          unsigned int cols;
          unsigned int pad;
        } m = {a.value, a.cols, 0};
        pk->push(m);
        break;
      }
      case ParamKind::Local: {
        if (a.kind != SKL_ARG_LOCAL)
          return fail(SKL_ERR_INVALID, "extra argument %d ('%s') must be Local(bytes)", i,
                      p.decl.c_str());
        lds = (lds + 15) & ~(size_t)15;
        unsigned int ofs = (unsigned int)lds;
        pk->push(ofs);
        lds += a.size;
        break;
      }
    }
  }
  pk->dyn_lds = lds;
  return SKL_OK;
}

// ---- source generation ------------------------------------------------------
struct Gen {
  const skl_skeleton_s* s;
  std::ostringstream o;

  void prelude()
  {
    o << skl_tmpl_prelude << "\n";
    o << "#define SKL_WG " << s->wg << "\n";
    // the common definitions the skeleton was created with, ahead of the
    // element typedefs that may name their types (Source.cpp:118-126)
    if (!s->common.empty())
      o << "#pragma clang force_cuda_host_device begin\n#line 1 \"common_definitions\"\n"
        << s->common << "\n#pragma clang force_cuda_host_device end\n";
  }
  void typedefs(const std::vector<std::string>& scl)
  {
    for (size_t i = 0; i < scl.size(); ++i)
      o << "typedef " << scl[i] << " SCL_TYPE_" << i << ";\n";
  }
  void user(const std::string& src, const std::string& name, const std::string& as)
  {
    o << rename_wrap(src, name, as);
  }
  void plumbing(const ExtraPlumbing& e)
  {
    o << "#define SKL_EXTRA_PARAMS " << e.params << "\n";
    o << "#define SKL_EXTRA_ARGS " << e.args << "\n";
    o << "#define SKL_EXTRA_SETUP " << e.setup << "\n";
  }
  void identity()
  {
    o << "#define SCL_IDENTITY (" << s->identity << ")\n";
  }
};

std::vector<Param> extras_of(const UserFunction& f, int lead)
{
  std::vector<Param> v;
  for (size_t i = lead; i < f.params.size(); ++i) v.push_back(f.params[i]);
  return v;
}

// accumulator hooks for Reduce / Scan
void emit_reduce_scan_hooks(std::ostringstream& o, const skl_skeleton_s* s)
{
  if (s->fp64_add()) {
    // float (+), precision FP64: FP64 accumulators, tile prefixes and carry
    // (one conversion and one FP64 add per element); every Scan output is
    // one rounding of the FP64 prefix
    o << R"(
typedef double SKL_ACC_T;
typedef double SKL_CARRY_T;
typedef double SKL_PREFIX_T;
__device__ __forceinline__ double skl_lift(const float& x) { return (double)x; }
#define SKL_ACC_COMB(a, b) ((a) + (b))
#define SKL_CARRY_COMB(a, b) ((a) + (b))
__device__ __forceinline__ double skl_to_carry(const double& a) { return a; }
__device__ __forceinline__ float skl_carry_out(const double& c) { return (float)c; }
#define SKL_ACC_ID 0.0
#define SKL_CARRY_ID 0.0
#define skl_make_prefix(c, p) ((c) + (p))
#define SKL_EMIT(P, q) ((float)((P) + (q)))
__device__ __forceinline__ float skl_emit0(const double& P) { return (float)P; }
)";
  } else if (s->precise_add()) {
    // float (+), DEFAULT: double-float pairs inside the work-group / tile,
    // FP64 across blocks (Reduce partials, Scan carries)
    o << R"(
typedef skl_df SKL_ACC_T;
typedef double SKL_CARRY_T;
struct skl_prefix_t { double c; skl_df p; };
typedef skl_prefix_t SKL_PREFIX_T;
__device__ __forceinline__ SKL_ACC_T skl_lift(const float& x) { skl_df r; r.hi = x; r.lo = 0.0f; return r; }
#define SKL_ACC_COMB(a, b) skl_df_add((a), (b))
#define SKL_CARRY_COMB(a, b) ((a) + (b))
__device__ __forceinline__ SKL_CARRY_T skl_to_carry(const skl_df& a) { return (double)a.hi + (double)a.lo; }
__device__ __forceinline__ float skl_carry_out(const double& c) { return (float)c; }
__device__ __forceinline__ skl_df skl_acc_zero() { skl_df r; r.hi = 0.0f; r.lo = 0.0f; return r; }
#define SKL_ACC_ID skl_acc_zero()
#define SKL_CARRY_ID 0.0
__device__ __forceinline__ SKL_PREFIX_T skl_make_prefix_df(double c, const skl_df& p) { SKL_PREFIX_T r; r.c = c; r.p = p; return r; }
#define skl_make_prefix(c, p) skl_make_prefix_df((c), (p))
__device__ __forceinline__ float skl_emit_df(const SKL_PREFIX_T& P, const skl_df& q) {
  skl_df s = skl_df_add(P.p, q);
  return (float)(P.c + ((double)s.hi + (double)s.lo));
}
#define SKL_EMIT(P, q) skl_emit_df((P), (q))
__device__ __forceinline__ float skl_emit0(const SKL_PREFIX_T& P) { return (float)(P.c + ((double)P.p.hi + (double)P.p.lo)); }
)";
  } else {
    o << R"(
typedef SKL_T SKL_ACC_T;
typedef SKL_T SKL_CARRY_T;
typedef SKL_T SKL_PREFIX_T;
__device__ __forceinline__ SKL_T skl_lift(const SKL_T& x) { return x; }
#define SKL_ACC_COMB(a, b) skl_user_f0((a), (b) SKL_EXTRA_ARGS)
#define SKL_CARRY_COMB(a, b) skl_user_f0((a), (b) SKL_EXTRA_ARGS)
__device__ __forceinline__ SKL_T skl_to_carry(const SKL_T& a) { return a; }
__device__ __forceinline__ SKL_T skl_carry_out(const SKL_T& c) { return c; }
__device__ __forceinline__ SKL_T skl_identity() { SKL_T v = SKL_IDENTITY_INIT; return v; }
#define SKL_ACC_ID skl_identity()
#define SKL_CARRY_ID skl_identity()
#define skl_make_prefix(c, p) SKL_ACC_COMB((c), (p))
#define SKL_EMIT(P, q) SKL_ACC_COMB((P), (q))
#define skl_emit0(P) (P)
)";
  }
}

}  // namespace

// Generates the fused source of variant `key` of skeleton `s`.
static int generate(const skl_skeleton_s* s, const std::string& key, std::string* out,
                    const skl_skeleton_s* zip = nullptr)
{
  Gen g{s, {}};
  std::ostringstream& o = g.o;
  g.prelude();
  const int lead = leading_params(s->kind);
  switch (s->kind) {
    case SKL_MAP: {
      bool is_void = s->types[0].name == "void";
      g.typedefs({s->types[1].name, is_void ? "int" : s->types[0].name});
      o << "#define SKL_TIN SCL_TYPE_0\n#define SKL_TOUT SCL_TYPE_1\n";
      o << "#define SKL_V " << s->V << "\n#define SKL_UNROLL " << s->unroll << "\n";
      g.user(s->src, s->func, "skl_user_f0");
      g.plumbing(make_extra_plumbing(s->extras, ""));
      o << "#define SKL_CALL1(x) skl_user_f0((x) SKL_EXTRA_ARGS)\n";
      if (key == "backward") {
        g.user(s->deriv_src, s->deriv_func, "skl_user_d0");
        o << "#define SKL_DCALL1(x) skl_user_d0((x) SKL_EXTRA_ARGS)\n#define SKL_DERIV_MAP 1\n";
      } else {
        o << (is_void ? "#define SKL_KIND_MAP_VOID 1\n" : "#define SKL_KIND_MAP 1\n");
      }
      o << skl_tmpl_map;
      break;
    }
    case SKL_MAP_INDEX:
    case SKL_MAP_INDEXPOINT: {
      bool is_void = s->types[0].name == "void";
      g.typedefs({is_void ? "int" : s->types[0].name});
      o << "#define SKL_TOUT SCL_TYPE_0\n";
      o << "#define SKL_V " << s->V << "\n#define SKL_UNROLL " << s->unroll << "\n";
      g.user(s->src, s->func, "skl_user_f0");
      g.plumbing(make_extra_plumbing(s->extras, ""));
      o << "#define SKL_CALL1(x) skl_user_f0((x) SKL_EXTRA_ARGS)\n";
      if (is_void) o << "#define SKL_VOID_OUT 1\n";
      if (s->kind == SKL_MAP_INDEX)
        o << (is_void ? "#define SKL_KIND_MAP_INDEX_VOID 1\n" : "#define SKL_KIND_MAP_INDEX 1\n");
      else
        o << "#define SKL_KIND_MAP_INDEXPOINT 1\n";
      o << skl_tmpl_map;
      break;
    }
    case SKL_ZIP: {
      bool is_void = s->types[0].name == "void";
      g.typedefs({s->types[1].name, s->types[2].name, is_void ? "int" : s->types[0].name});
      o << "#define SKL_TL SCL_TYPE_0\n#define SKL_TR SCL_TYPE_1\n#define SKL_TOUT SCL_TYPE_2\n";
      o << "#define SKL_V " << s->V << "\n#define SKL_UNROLL " << s->unroll << "\n";
      g.user(s->src, s->func, "skl_user_f0");
      g.plumbing(make_extra_plumbing(s->extras, ""));
      o << "#define SKL_CALL2(a, b) skl_user_f0((a), (b) SKL_EXTRA_ARGS)\n";
      if (key == "backward") {
        g.user(s->deriv_src, s->deriv_func, "skl_user_d0");
        if (s->deriv_func2 != s->deriv_func)
          g.user(s->deriv_src, s->deriv_func2, "skl_user_d1");
        o << "#define SKL_DCALL2L(a, b) skl_user_d0((a), (b) SKL_EXTRA_ARGS)\n"
          << "#define SKL_DCALL2R(a, b) skl_user_d1((a), (b) SKL_EXTRA_ARGS)\n"
          << "#define SKL_DERIV_ZIP 1\n";
      } else {
        o << (is_void ? "#define SKL_KIND_ZIP_VOID 1\n" : "#define SKL_KIND_ZIP 1\n");
      }
      o << skl_tmpl_map;
      break;
    }
    case SKL_REDUCE:
    case SKL_SCAN: {
      g.typedefs({s->types[0].name});
      o << "#define SKL_T SCL_TYPE_0\n";
      o << "#define SKL_V " << s->V << "\n#define SKL_UNROLL " << s->unroll << "\n";
      o << "#define SKL_J " << s->scan_j << "\n";
      g.identity();
      o << "#define SKL_IDENTITY_INIT " << s->identity << "\n";
      g.user(s->src, s->func, "skl_user_f0");
      if (zip) {  // fused Zip -> Reduce
        g.user(zip->src, zip->func, "skl_user_z0");
        o << "#define SKL_TL " << zip->types[1].name << "\n#define SKL_TR "
          << zip->types[2].name << "\n";
        o << "#define SKL_ZCALL(a, b) skl_user_z0((a), (b))\n#define SKL_FUSED_ZIP 1\n";
      }
      g.plumbing(make_extra_plumbing(s->extras, ""));
      emit_reduce_scan_hooks(o, s);
      if (key == "reverse") o << "#define SKL_SCAN_REVERSE 1\n";
      o << (s->kind == SKL_REDUCE ? skl_tmpl_reduce : skl_tmpl_scan);
      break;
    }
    case SKL_MAPOVERLAP: {
      g.typedefs({s->types[1].name, s->types[0].name});
      o << "#define SKL_TIN SCL_TYPE_0\n#define SKL_TOUT SCL_TYPE_1\n";
      o << "#define SCL_OVERLAP_RANGE " << s->overlap << "\n";
      o << "#define SKL_MO_RB " << s->mo_rb << "\n";
      if (s->padding == SKL_PADDING_NEUTRAL)
        o << "#define SKL_PAD_NEUTRAL 1\n#define SKL_NEUTRAL ((SKL_TIN)(" << s->neutral
          << "))\n#define NEUTRAL (" << s->neutral << ")\n";
      // the user function refers to input_matrix_t / getData, defined by the
      // template: emit the template's declarations first
      std::string tmpl = skl_tmpl_mapoverlap;
      size_t cut = tmpl.find("extern \"C\"");
      o << tmpl.substr(0, cut);
      g.user(s->src, s->func, "skl_user_f0");
      g.plumbing(make_extra_plumbing(s->extras, ""));
      o << "#define SKL_CALL1(x) skl_user_f0((x) SKL_EXTRA_ARGS)\n";
      o << tmpl.substr(cut);
      break;
    }
    case SKL_ALLPAIRS: {
      g.typedefs({s->types[1].name, s->types[2].name, s->types[0].name});
      o << "#define SKL_TL SCL_TYPE_0\n#define SKL_TR SCL_TYPE_1\n#define SKL_TOUT SCL_TYPE_2\n";
      g.identity();
      g.user(s->src, s->func, "skl_user_f0");   // reduce
      g.user(s->src2, s->func2, "skl_user_f1"); // zip
      auto re = extras_of(s->uf, 2), ze = extras_of(s->uf2, 2);
      ExtraPlumbing pr = make_extra_plumbing(re, "r"), pz = make_extra_plumbing(ze, "z");
      ExtraPlumbing both{pr.params + pz.params, "", pr.setup + pz.setup};
      g.plumbing(both);
      if (key == "fma")
        o << "#define SKL_AP_STEP(acc, a, b) fmaf((a), (b), (acc))\n";
      else
        o << "#define SKL_AP_STEP(acc, a, b) skl_user_f0((acc), skl_user_f1((a), (b)" << pz.args
          << ")" << pr.args << ")\n";
      o << "#define SKL_KIND_ALLPAIRS 1\n" << skl_tmpl_allpairs;
      break;
    }
    case SKL_ALLPAIRS_GENERAL: {
      g.typedefs({s->types[1].name, s->types[2].name, s->types[0].name});
      o << "#define SKL_TL SCL_TYPE_0\n#define SKL_TR SCL_TYPE_1\n#define SKL_TOUT SCL_TYPE_2\n";
      // lmatrix_t / rmatrix_t must precede the user source
      std::string tmpl = skl_tmpl_allpairs;
      size_t a = tmpl.find("#ifdef SKL_KIND_ALLPAIRS_GENERAL");
      size_t b = tmpl.find("#define SKL_APG_B");
      o << "#define SKL_KIND_ALLPAIRS_GENERAL 1\n";
      o << tmpl.substr(a, b - a);
      g.user(s->src, s->func, "skl_user_f0");
      g.plumbing(make_extra_plumbing(s->extras, ""));
      o << "#define SKL_CALL3(a, b, c) skl_user_f0((a), (b), (c) SKL_EXTRA_ARGS)\n";
      o << "#define SKL_APG_LDS_BASE 0\n";
      std::string dim = key.rfind("dim", 0) == 0 ? key.substr(3) : "K";
      o << "#define SKL_APG_DIM " << (dim.empty() ? "K" : dim) << "\n";
      o << tmpl.substr(b);
      break;
    }
    default:
      return fail(SKL_ERR_INVALID, "unknown skeleton kind %d", s->kind);
  }
  (void)lead;
  *out = o.str();
  return SKL_OK;
}

static std::string build_options(const skl_skeleton_s* s)
{
  std::string o = "-ffp-contract=" + s->contract;
  if (!s->options.empty()) o += " " + s->options;
  return o;
}

// Returns the built program of variant `key`, generating it on first use.
static int get_program(skl_skeleton_s* s, const std::string& key, skl_program* prog,
                       const skl_skeleton_s* zip = nullptr)
{
  auto it = s->progs.find(key);
  if (it != s->progs.end() && it->second.program) {
    *prog = it->second.program;
    return SKL_OK;
  }
  skl_skeleton_s::Prog p;
  SKL_TRY(generate(s, key, &p.source, zip));
  char* log = nullptr;
  int st = skl_program_build(p.source.c_str(), build_options(s).c_str(), &p.program, &log);
  if (log) skl_free_string(log);
  if (st != SKL_OK) return st;
  s->progs[key] = p;
  *prog = p.program;
  return SKL_OK;
}

static int get_kernel(skl_skeleton_s* s, const std::string& key, const char* name,
                      skl_kernel* k, const skl_skeleton_s* zip = nullptr)
{
  skl_program p;
  SKL_TRY(get_program(s, key, &p, zip));
  return skl_kernel_get(p, name, k);
}

static int ensure_scratch(skl_skeleton_s* s, size_t bytes)
{
  if (s->scratch_bytes >= bytes) return SKL_OK;
  if (s->scratch) {  // kept until the skeleton is released (graphs may hold it)
    s->retired.push_back(s->scratch);
    s->scratch = nullptr;
    s->scratch_bytes = 0;
  }
  size_t sz = std::max<size_t>(bytes, 4096);
  SKL_HIP_TRY(hipMalloc(&s->scratch, sz));
  SKL_HIP_TRY(hipMemset(s->scratch, 0, sz));
  SKL_HIP_TRY(hipDeviceSynchronize());
  s->scratch_bytes = sz;
  return SKL_OK;
}

static int launch(skl_kernel k, unsigned gx, unsigned gy, unsigned bx, unsigned by,
                  size_t lds, std::vector<void*>& args, skl_stream stream)
{
  unsigned grid[3] = {gx, gy, 1}, block[3] = {bx, by, 1};
  return skl_kernel_launch(k, grid, block, lds, args.data(), stream);
}

static std::vector<void*> with_extras(std::initializer_list<void*> fixed, const ArgPack& pk)
{
  std::vector<void*> v(fixed);
  v.insert(v.end(), pk.ptrs.begin(), pk.ptrs.end());
  return v;
}

// ---------------------------------------------------------------------------
// creation
// ---------------------------------------------------------------------------
// A literal as the generated code rounds it to float: a decimal with an
// f suffix rounds once, to float; without, to double and then to float.
static bool float_literal(const std::string& text, float* out)
{
  std::string t;
  for (char c : text)
    if (!isspace((unsigned char)c)) t += c;
  bool f = !t.empty() && (t.back() == 'f' || t.back() == 'F');
  if (f) t.pop_back();
  if (t.empty() || t.find_first_not_of("0123456789.eE+-") != std::string::npos) return false;
  char* end = nullptr;
  if (f) {
    *out = std::strtof(t.c_str(), &end);
  } else {
    *out = (float)std::strtod(t.c_str(), &end);
  }
  return end && *end == '\0';
}

// A float MapOverlap whose user function is the weighted stencil, compiled
// with nothing that could change its arithmetic (no macros in the common
// definitions, no build options but -ffp-contract), runs on the native
// stencil kernel, which performs every output's operations in the loop's
// order (native/stencil.hip): same results as the generated kernel.
static void detect_native_stencil(skl_skeleton_s* s)
{
  s->stencil_radius = -1;
  if (s->types.size() != 2 || s->types[0].name != "float" || s->types[1].name != "float")
    return;
  if (s->common.find('#') != std::string::npos) return;
  {
    std::string o = s->options;
    size_t at;
    while ((at = o.find("-ffp-contract=")) != std::string::npos) {
      size_t e = o.find_first_of(" \t", at);
      o.erase(at, e == std::string::npos ? std::string::npos : e - at);
    }
    if (o.find_first_not_of(" \t") != std::string::npos) return;
  }
  const int n = match_weighted_stencil(s->src, s->uf);
  if (n < 0 || n > s->overlap || n > skl_internal_stencil_max_radius()) return;
  float neutral = 0.0f;
  if (s->padding == SKL_PADDING_NEUTRAL && !float_literal(s->neutral, &neutral)) return;
  s->stencil_radius = n;
  s->stencil_neutral = neutral;
  // the transposed stencil of this radius is the adjoint
  if (n == s->overlap) s->stencil_checked = 1;
}

static int setup_skeleton(const skl_skeleton_desc* d, skl_skeleton_s* s)
{
  if (!d) return fail(SKL_ERR_INVALID, "descriptor is NULL");
  s->kind = d->kind;
  if (d->kind < SKL_MAP || d->kind > SKL_ALLPAIRS_GENERAL)
    return fail(SKL_ERR_INVALID, "unknown skeleton kind %d", d->kind);
  if (!d->source || !*d->source)
    return fail(SKL_ERR_INVALID, "Tried to create program with empty user source.");
  s->src = d->source;
  s->common = common_text();
  s->func = str_or(d->func, "func");
  s->identity = str_or(d->identity, "0");
  s->neutral = str_or(d->neutral, "0");
  // Scan: 512-thread tiles (measured on MI355X at 2^28: int32 5.56 TB/s
  // against 4.83 with 256 threads, float 13% faster); everything else 256
  s->wg = d->work_group_size > 0 ? d->work_group_size : (s->kind == SKL_SCAN ? 512 : 256);
  if (s->wg > 1024) return fail(SKL_ERR_INVALID, "work-group size %d exceeds 1024", s->wg);
  if (d->precision < SKL_PREC_DEFAULT || d->precision > SKL_PREC_FP64)
    return fail(SKL_ERR_INVALID, "unknown precision %d", d->precision);
  s->precision = d->precision;
  s->overlap = d->overlap_range;
  s->padding = d->padding;
  if (s->kind == SKL_MAPOVERLAP) {
    if (s->overlap < 0 || s->overlap > 32)
      return fail(SKL_ERR_INVALID, "overlap range %d outside [0, 32]", s->overlap);
    if (s->padding != SKL_PADDING_NEUTRAL && s->padding != SKL_PADDING_NEAREST)
      return fail(SKL_ERR_INVALID, "unknown padding %d", s->padding);
  }
  if (d->build_options) s->options = d->build_options;
  // floating-point contraction of the user functions: SKELCL_FP_CONTRACT
  // (on | off | fast; default on), a build option's own -ffp-contract wins
  s->contract = "on";
  if (const char* e = getenv("SKELCL_FP_CONTRACT")) {
    const std::string v = e;
    if (v != "on" && v != "off" && v != "fast")
      return fail(SKL_ERR_INVALID, "SKELCL_FP_CONTRACT=%s: use on, off or fast", e);
    s->contract = v;
  }
  {
    const size_t at = s->options.rfind("-ffp-contract=");
    if (at != std::string::npos) {
      std::string v = s->options.substr(at + 14);
      v = v.substr(0, v.find_first_of(" \t"));
      if (v == "on" || v == "off" || v == "fast") s->contract = v;
    }
  }
  int nt = expected_types(s->kind);
  if (d->ntypes != nt)
    return fail(SKL_ERR_INVALID, "%s needs %d type(s), got %d", kind_name(s->kind), nt,
                d->ntypes);
  for (int i = 0; i < nt; ++i) {
    if (!d->types[i].name || !*d->types[i].name)
      return fail(SKL_ERR_INVALID, "type %d has no name", i);
    s->types.push_back({d->types[i].name, d->types[i].size});
  }
  std::string err;
  if (!parse_user_function(s->src, s->func, &s->uf, &err)) return fail(SKL_ERR_PARSE, "%s", err.c_str());
  int lead = leading_params(s->kind);
  if ((int)s->uf.params.size() < lead)
    return fail(SKL_ERR_PARSE, "%s user function '%s' needs at least %d parameter(s)",
                kind_name(s->kind), s->func.c_str(), lead);
  s->extras = extras_of(s->uf, lead);
  if (s->kind == SKL_ALLPAIRS) {
    if (!d->source2 || !*d->source2)
      return fail(SKL_ERR_INVALID, "AllPairs(Reduce, Zip) needs the zip source");
    s->src2 = d->source2;
    s->func2 = str_or(d->func2, "func");
    if (!parse_user_function(s->src2, s->func2, &s->uf2, &err))
      return fail(SKL_ERR_PARSE, "%s", err.c_str());
    if (s->uf2.params.size() < 2)
      return fail(SKL_ERR_PARSE, "AllPairs zip function needs two parameters");
    s->extras = extras_of(s->uf, 2);
    auto ze = extras_of(s->uf2, 2);
    s->extras.insert(s->extras.end(), ze.begin(), ze.end());
    s->op2 = classify_binary(s->uf2);
  }
  s->op = classify_binary(s->uf);
  s->simple = uses_work_item_builtins(s->src);
  if (s->kind == SKL_MAPOVERLAP) detect_native_stencil(s);
  if (d->deriv_source && *d->deriv_source) {
    s->deriv_src = d->deriv_source;
    s->deriv_func = str_or(d->deriv_func, "dfunc");
    s->deriv_func2 = str_or(d->deriv_func2, s->deriv_func.c_str());
    if (!parse_user_function(s->deriv_src, s->deriv_func, &s->duf, &err))
      return fail(SKL_ERR_PARSE, "derivative: %s", err.c_str());
    if (s->kind == SKL_ZIP &&
        !parse_user_function(s->deriv_src, s->deriv_func2, &s->duf2, &err))
      return fail(SKL_ERR_PARSE, "derivative: %s", err.c_str());
    s->has_deriv = true;
  }
  // pack width from the element types
  switch (s->kind) {
    case SKL_MAP:
      s->V = pack_width({s->types[1].size, s->types[0].name == "void" ? 0 : s->types[0].size});
      break;
    case SKL_MAP_INDEX: case SKL_MAP_INDEXPOINT:
      s->V = pack_width({s->types[0].size});
      break;
    case SKL_ZIP:
      s->V = pack_width({s->types[0].size, s->types[1].size, s->types[2].size});
      break;
    case SKL_REDUCE: case SKL_SCAN:
      s->V = pack_width({s->types[0].size});
      break;
    default:
      s->V = 1;
  }
  if (s->kind == SKL_SCAN) s->scan_j = s->default_scan_j();
  return SKL_OK;
}

// AllPairs(+, *) with float output AND float inputs: the FMA form (and the
// native products when the identity is zero).  Other input types take the
// user functions' own fold.
static bool float_addmul(const skl_skeleton_s* s)
{
  return s->kind == SKL_ALLPAIRS && s->op == UserOp::Add && s->op2 == UserOp::Mul &&
         s->extras.empty() && s->types[0].name == "float" && s->types[1].name == "float" &&
         s->types[2].name == "float";
}

// True when the identity text is a zero literal ("0", "0.0f", "-0.", "0e0").
static bool identity_is_zero(const std::string& id)
{
  std::string t;
  for (char c : id)
    if (!isspace((unsigned char)c)) t += c;
  if (!t.empty() && (t.back() == 'f' || t.back() == 'F')) t.pop_back();
  if (t.empty()) return false;
  char* end = nullptr;
  const double v = strtod(t.c_str(), &end);
  return end && *end == '\0' && v == 0.0;
}

// Without contraction (the parity setting) AllPairs(+, *) on float runs the
// user functions' own fold in k order, two roundings per step, which is what
// a sequential host loop computes (AllPairsKernel.cl:89-91).
static std::string first_key(const skl_skeleton_s* s)
{
  return float_addmul(s) && s->contract != "off" ? "fma" : "main";
}

// Drops every generated program and builds the first one again: a skeleton
// always holds the kernel of its current launch shape and options.
int skl::rebuild_programs(skl_skeleton_s* s)
{
  for (auto& kv : s->progs) skl_program_release(kv.second.program);
  s->progs.clear();
  if (s->kind == SKL_ALLPAIRS_GENERAL) return SKL_OK;  // built per dim on first run
  skl_program p;
  return get_program(s, first_key(s), &p);
}

extern "C" int skl_skeleton_create(const skl_skeleton_desc* desc, skl_skeleton* out)
{
  SKL_TRY(ensure_initialized());
  if (!out) return fail(SKL_ERR_INVALID, "out is NULL");
  auto* s = new skl_skeleton_s;
  int st = setup_skeleton(desc, s);
  // the launch shape is compiled into the kernels: set it before the first
  // program is generated
  if (st == SKL_OK) s->apply_default_shape(cu_count());
  if (st == SKL_OK && s->kind != SKL_ALLPAIRS_GENERAL) {
    skl_program p;
    st = get_program(s, first_key(s), &p);
  }
  if (st != SKL_OK) {
    delete s;
    return st;
  }
  *out = s;
  return SKL_OK;
}

extern "C" int skl_skeleton_check(const skl_skeleton_desc* desc, char** source, char** log)
{
  skl_skeleton_s s;
  if (source) *source = nullptr;
  if (log) *log = nullptr;
  SKL_TRY(setup_skeleton(desc, &s));
  s.apply_default_shape(cu_count());
  std::string src;
  SKL_TRY(generate(&s, s.kind == SKL_ALLPAIRS_GENERAL ? "dim256" : first_key(&s), &src));
  if (source) {
    *source = (char*)malloc(src.size() + 1);
    memcpy(*source, src.c_str(), src.size() + 1);
  }
  return skl_program_compile(src.c_str(), build_options(&s).c_str(), nullptr, nullptr, log);
}

extern "C" int skl_skeleton_release(skl_skeleton s)
{
  if (!s) return SKL_OK;
  for (auto& kv : s->progs) skl_program_release(kv.second.program);
  if (s->scratch || s->seg_scratch || !s->retired.empty()) hipDeviceSynchronize();
  if (s->scratch) hipFree(s->scratch);
  if (s->seg_scratch) hipFree(s->seg_scratch);
  for (void* p : s->retired) hipFree(p);
  delete s;
  return SKL_OK;
}

extern "C" int skl_skeleton_work_group_size(skl_skeleton s, int* wg)
{
  if (!s || !wg) return fail(SKL_ERR_INVALID, "NULL argument");
  *wg = s->wg;
  return SKL_OK;
}

extern "C" int skl_skeleton_set_work_group_size(skl_skeleton s, int wg)
{
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (wg <= 0 || wg > 1024) return fail(SKL_ERR_INVALID, "work-group size %d", wg);
  if (wg == s->wg) return SKL_OK;
  s->wg = wg;
  return skl::rebuild_programs(s);  // kernels are compiled for the work-group size
}

extern "C" int skl_skeleton_set_precision(skl_skeleton s, int precision)
{
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (precision < SKL_PREC_DEFAULT || precision > SKL_PREC_FP64)
    return fail(SKL_ERR_INVALID, "unknown precision %d", precision);
  if (precision != s->precision) {
    const bool default_j = s->kind == SKL_SCAN && s->scan_j == s->default_scan_j();
    s->precision = precision;
    if (default_j) s->scan_j = s->default_scan_j();
    return skl::rebuild_programs(s);
  }
  return SKL_OK;
}

extern "C" int skl_skeleton_source(skl_skeleton s, char** source)
{
  if (!s || !source) return fail(SKL_ERR_INVALID, "NULL argument");
  std::string src;
  auto it = s->progs.begin();
  if (it != s->progs.end()) src = it->second.source;
  else SKL_TRY(generate(s, first_key(s), &src));
  *source = (char*)malloc(src.size() + 1);
  memcpy(*source, src.c_str(), src.size() + 1);
  return SKL_OK;
}

extern "C" int skl_skeleton_extra_count(skl_skeleton s, int* count)
{
  if (!s || !count) return fail(SKL_ERR_INVALID, "NULL argument");
  *count = (int)s->extras.size();
  return SKL_OK;
}

extern "C" int skl_describe_user_function(const char* source, const char* func,
                                          char** description)
{
  if (!source || !func || !description) return fail(SKL_ERR_INVALID, "NULL argument");
  UserFunction f;
  std::string err;
  if (!parse_user_function(source, func, &f, &err)) return fail(SKL_ERR_PARSE, "%s", err.c_str());
  std::ostringstream o;
  o << "return " << f.return_type << "\n";
  for (auto& p : f.params) {
    int k = p.kind == ParamKind::Value ? SKL_ARG_VALUE
            : p.kind == ParamKind::Buffer ? SKL_ARG_BUFFER
            : p.kind == ParamKind::Matrix ? SKL_ARG_MATRIX : SKL_ARG_LOCAL;
    o << "param " << k << " " << p.value_size << " " << p.type_text << "|" << p.name << "\n";
  }
  const char* ops[] = {"other", "add", "mul", "max", "min"};
  o << "op " << ops[(int)classify_binary(f)] << "\n";
  o << "work_item " << (uses_work_item_builtins(source) ? 1 : 0) << "\n";
  o << "stencil " << match_weighted_stencil(source, f) << "\n";
  std::string s = o.str();
  *description = (char*)malloc(s.size() + 1);
  memcpy(*description, s.c_str(), s.size() + 1);
  return SKL_OK;
}

extern "C" int skl_skeleton_extra_info(skl_skeleton s, int index, int* kind, size_t* size,
                                       char* type, size_t type_len)
{
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (index < 0 || index >= (int)s->extras.size())
    return fail(SKL_ERR_INVALID, "extra parameter %d does not exist", index);
  const Param& p = s->extras[index];
  if (kind) {
    switch (p.kind) {
      case ParamKind::Value: *kind = SKL_ARG_VALUE; break;
      case ParamKind::Buffer: *kind = SKL_ARG_BUFFER; break;
      case ParamKind::Matrix: *kind = SKL_ARG_MATRIX; break;
      case ParamKind::Local: *kind = SKL_ARG_LOCAL; break;
    }
  }
  if (size) *size = p.value_size;
  if (type && type_len) snprintf(type, type_len, "%s", p.type_text.c_str());
  return SKL_OK;
}

// ---------------------------------------------------------------------------
// runs
// ---------------------------------------------------------------------------
static int check_kind(skl_skeleton s, int kind)
{
  SKL_TRY(ensure_initialized());
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (s->kind != kind)
    return fail(SKL_ERR_INVALID, "skeleton is a %s, not a %s", kind_name(s->kind),
                kind_name(kind));
  return SKL_OK;
}

static unsigned blocks_for(unsigned long long items, unsigned long long per_block,
                           unsigned long long cap)
{
  unsigned long long b = (items + per_block - 1) / per_block;
  if (b > cap) b = cap;
  if (b == 0) b = 1;
  return (unsigned)b;
}

extern "C" int skl_map_run(skl_skeleton s, void* out, const void* in, size_t n,
                           const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_MAP));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (n == 0) return SKL_OK;
  bool is_void = s->types[0].name == "void";
  if (!in || (!is_void && !out)) return fail(SKL_ERR_INVALID, "NULL buffer");
  unsigned long long nn = n;
  skl_kernel k;
  if (is_void) {
    SKL_TRY(get_kernel(s, "main", "skl_map_void", &k));
    auto a = with_extras({&in, &nn}, pk);
    return launch(k, blocks_for(n, s->wg, 0xFFFFFFFFull), 1, s->wg, 1, pk.dyn_lds, a, stream);
  }
  bool fast = !s->simple && pk.dyn_lds == 0 && aligned(in, s->V * s->types[1].size) &&
              aligned(out, s->V * s->types[0].size);
  auto a = with_extras({&in, &out, &nn}, pk);
  if (fast) {
    SKL_TRY(get_kernel(s, "main", "skl_map", &k));
    unsigned long long npk = n / s->V;
    return launch(k, blocks_for(npk, (unsigned long long)s->wg * s->unroll, s->max_blocks), 1,
                  s->wg, 1, 0, a, stream);
  }
  SKL_TRY(get_kernel(s, "main", "skl_map_simple", &k));
  return launch(k, blocks_for(n, s->wg, 0xFFFFFFFFull), 1, s->wg, 1, pk.dyn_lds, a, stream);
}

extern "C" int skl_map_index_run(skl_skeleton s, void* out, size_t n, size_t offset,
                                 const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_MAP_INDEX));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (n == 0) return SKL_OK;
  bool is_void = s->types[0].name == "void";
  if (!is_void && !out) return fail(SKL_ERR_INVALID, "NULL buffer");
  unsigned long long nn = n, off = offset;
  skl_kernel k;
  if (is_void) {
    SKL_TRY(get_kernel(s, "main", "skl_map_index_void", &k));
    auto a = with_extras({&nn, &off}, pk);
    return launch(k, blocks_for(n, s->wg, 0xFFFFFFFFull), 1, s->wg, 1, pk.dyn_lds, a, stream);
  }
  auto a = with_extras({&out, &nn, &off}, pk);
  if (!s->simple && pk.dyn_lds == 0 && aligned(out, s->V * s->types[0].size)) {
    SKL_TRY(get_kernel(s, "main", "skl_map_index", &k));
    return launch(k, blocks_for(n / s->V, s->wg, s->max_blocks), 1, s->wg, 1, 0, a, stream);
  }
  SKL_TRY(get_kernel(s, "main", "skl_map_index_simple", &k));
  return launch(k, blocks_for(n, s->wg, 0xFFFFFFFFull), 1, s->wg, 1, pk.dyn_lds, a, stream);
}

extern "C" int skl_map_indexpoint_run(skl_skeleton s, void* out, size_t rows, size_t cols,
                                      size_t row_offset, const skl_arg* args, int nargs,
                                      skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_MAP_INDEXPOINT));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (rows == 0 || cols == 0) return SKL_OK;
  bool is_void = s->types[0].name == "void";
  if (!is_void && !out) return fail(SKL_ERR_INVALID, "NULL buffer");
  unsigned long long r = rows, c = cols, ro = row_offset;
  auto a = with_extras({&out, &r, &c, &ro}, pk);
  skl_kernel k;
  if (!s->simple && pk.dyn_lds == 0) {
    SKL_TRY(get_kernel(s, "main", "skl_map_indexpoint_fast", &k));
    return launch(k, blocks_for(cols, 64, 0xFFFFFFFFull), blocks_for(rows, 4, 65535), 256, 1, 0,
                  a, stream);
  }
  // reference shape: square local size floorPow2(sqrt(wg)) (MapDef.h:587-595)
  unsigned l = 1;
  while ((l * 2) * (l * 2) <= (unsigned)s->wg) l *= 2;
  SKL_TRY(get_kernel(s, "main", "skl_map_indexpoint", &k));
  if ((rows + l - 1) / l > 0xFFFFFFFFull || (cols + l - 1) / l > 65535)
    return fail(SKL_ERR_INVALID, "index matrix too large for the work-item form");
  return launch(k, blocks_for(rows, l, 0xFFFFFFFFull), blocks_for(cols, l, 65535), l, l,
                pk.dyn_lds, a, stream);
}

extern "C" int skl_zip_run(skl_skeleton s, void* out, const void* left, const void* right,
                           size_t n, const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_ZIP));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (n == 0) return SKL_OK;
  bool is_void = s->types[0].name == "void";
  if (!left || !right || (!is_void && !out)) return fail(SKL_ERR_INVALID, "NULL buffer");
  unsigned long long nn = n;
  skl_kernel k;
  if (is_void) {
    SKL_TRY(get_kernel(s, "main", "skl_zip_void", &k));
    auto a = with_extras({&left, &right, &nn}, pk);
    return launch(k, blocks_for(n, s->wg, 0xFFFFFFFFull), 1, s->wg, 1, pk.dyn_lds, a, stream);
  }
  bool fast = !s->simple && pk.dyn_lds == 0 && aligned(left, s->V * s->types[1].size) &&
              aligned(right, s->V * s->types[2].size) && aligned(out, s->V * s->types[0].size);
  auto a = with_extras({&left, &right, &out, &nn}, pk);
  if (fast) {
    SKL_TRY(get_kernel(s, "main", "skl_zip", &k));
    return launch(k, blocks_for(n / s->V, (unsigned long long)s->wg * s->unroll, s->max_blocks),
                  1, s->wg, 1, 0, a, stream);
  }
  SKL_TRY(get_kernel(s, "main", "skl_zip_simple", &k));
  return launch(k, blocks_for(n, s->wg, 0xFFFFFFFFull), 1, s->wg, 1, pk.dyn_lds, a, stream);
}

static size_t carry_size(const skl_skeleton_s* s)
{
  return s->precise_add() || s->fp64_add() ? sizeof(double) : s->types[0].size;
}

// Reduce scratch: the top completion counter at 0, 64 sub-counters 256 bytes
// apart, then one carry-sized partial per work-group.
static constexpr size_t kReduceCounterBytes = 256 * 65;

static int reduce_grid(const skl_skeleton_s* s, size_t n, unsigned* blocks, size_t* scratch)
{
  if (s->wg > 64 && s->wg % 64 != 0)
    return fail(SKL_ERR_INVALID, "Reduce work-group size %d: use at most 64 or a multiple of 64",
                s->wg);
  const size_t csz = (carry_size(s) + 7) & ~(size_t)7;
  const unsigned long long npk = (n + s->V - 1) / s->V;
  *blocks = blocks_for(npk, s->wg, (unsigned long long)s->max_blocks);
  *scratch = kReduceCounterBytes + (size_t)*blocks * csz;
  return SKL_OK;
}

// `fixed`: caller-provided scratch of `fixed_bytes` (no allocation, and the
// kernel must already be built); NULL = the skeleton's own scratch.
static int reduce_launch(skl_skeleton_s* s, const std::string& key, void* out,
                         std::initializer_list<const void*> ins, size_t n, const ArgPack& pk,
                         skl_stream stream, const skl_skeleton_s* zip = nullptr,
                         void* fixed = nullptr, size_t fixed_bytes = 0)
{
  if (n == 0) return fail(SKL_ERR_INVALID, "Reduce of an empty input");
  if (!out) return fail(SKL_ERR_INVALID, "NULL output buffer");
  unsigned blocks = 0;
  size_t need = 0;
  SKL_TRY(reduce_grid(s, n, &blocks, &need));
  void* base = fixed;
  skl_kernel k;
  if (fixed) {
    if (need > fixed_bytes)
      return fail(SKL_ERR_INVALID,
                  "Reduce of %zu elements on %u work-groups needs %zu bytes of scratch; the "
                  "fixed scratch holds %zu",
                  n, blocks, need, fixed_bytes);
    auto it = s->progs.find(key);
    if (it == s->progs.end() || !it->second.program)
      return fail(SKL_ERR_INVALID,
                  "the Reduce kernel for this launch shape is not built yet (building loads a "
                  "code object, which allocates): run skl_reduce_run once first");
    SKL_TRY(skl_kernel_get(it->second.program, "skl_reduce", &k));
  } else {
    const size_t csz = (carry_size(s) + 7) & ~(size_t)7;
    SKL_TRY(ensure_scratch(s, kReduceCounterBytes + (size_t)s->max_blocks * csz));
    base = s->scratch;
    SKL_TRY(get_kernel(s, key, "skl_reduce", &k, zip));
  }
  void* partials = static_cast<char*>(base) + kReduceCounterBytes;
  void* ticket = base;
  unsigned long long nn = n;
  std::vector<const void*> in(ins);
  std::vector<void*> a;
  for (auto& p : in) a.push_back((void*)&p);
  a.push_back(&out);
  a.push_back(&nn);
  a.push_back(&partials);
  a.push_back(&ticket);
  a.insert(a.end(), pk.ptrs.begin(), pk.ptrs.end());
  return launch(k, blocks, 1, s->wg, 1, pk.dyn_lds, a, stream);
}

extern "C" int skl_reduce_run(skl_skeleton s, void* out, const void* in, size_t n,
                              const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_REDUCE));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (!in) return fail(SKL_ERR_INVALID, "NULL input buffer");
  if (pk.dyn_lds) return fail(SKL_ERR_UNSUPPORTED, "Reduce user functions cannot take Local");
  if (!aligned(in, s->V * s->types[0].size))
    return fail(SKL_ERR_INVALID, "Reduce input must be aligned to %zu bytes",
                s->V * s->types[0].size);
  return reduce_launch(s, "main", out, {in}, n, pk, stream);
}

namespace skl {
int reduce_scratch_need(skl_skeleton s, size_t n, size_t* counter_bytes, size_t* total_bytes)
{
  SKL_TRY(check_kind(s, SKL_REDUCE));
  unsigned blocks = 0;
  SKL_TRY(reduce_grid(s, n ? n : 1, &blocks, total_bytes));
  *counter_bytes = kReduceCounterBytes;
  return SKL_OK;
}

int reduce_run_fixed_scratch(skl_skeleton s, void* out, const void* in, size_t n,
                             const skl_arg* args, int nargs, skl_stream stream, void* scratch,
                             size_t scratch_bytes)
{
  SKL_TRY(check_kind(s, SKL_REDUCE));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (!in) return fail(SKL_ERR_INVALID, "NULL input buffer");
  if (pk.dyn_lds) return fail(SKL_ERR_UNSUPPORTED, "Reduce user functions cannot take Local");
  if (!aligned(in, s->V * s->types[0].size))
    return fail(SKL_ERR_INVALID, "Reduce input must be aligned to %zu bytes",
                s->V * s->types[0].size);
  return reduce_launch(s, "main", out, {in}, n, pk, stream, nullptr, scratch, scratch_bytes);
}
}  // namespace skl

extern "C" int skl_zip_reduce_run(skl_skeleton zip, skl_skeleton reduce, void* out,
                                  const void* left, const void* right, size_t n,
                                  skl_stream stream)
{
  SKL_TRY(check_kind(zip, SKL_ZIP));
  SKL_TRY(check_kind(reduce, SKL_REDUCE));
  if (!zip->extras.empty() || !reduce->extras.empty())
    return fail(SKL_ERR_UNSUPPORTED, "fused Zip->Reduce takes no extra arguments");
  if (zip->types[0].name != reduce->types[0].name)
    return fail(SKL_ERR_INVALID, "Zip output type %s differs from Reduce type %s",
                zip->types[0].name.c_str(), reduce->types[0].name.c_str());
  if (zip->types[1].size != reduce->types[0].size || zip->types[2].size != reduce->types[0].size)
    return fail(SKL_ERR_UNSUPPORTED, "fused Zip->Reduce needs equal element sizes");
  if (!left || !right) return fail(SKL_ERR_INVALID, "NULL input buffer");
  size_t al = reduce->V * reduce->types[0].size;
  if (!aligned(left, al) || !aligned(right, al))
    return fail(SKL_ERR_INVALID, "inputs must be aligned to %zu bytes", al);
  ArgPack pk;
  // the fused program is keyed on what the Zip contributes to its source,
  // never on the Zip object (a released Zip's address can come back)
  std::ostringstream z;
  z << zip->src << "\n//func " << zip->func << "\n//options " << zip->options;
  for (auto& t : zip->types) z << "\n//type " << t.name << " " << t.size;
  const std::string key = "fused:" + sha1_hex(z.str());
  return reduce_launch(reduce, key, out, {left, right}, n, pk, stream, zip);
}

static int scan_launch(skl_skeleton s, const std::string& key, void* out, const void* in,
                       size_t n, int inclusive, const skl_arg* args, int nargs,
                       skl_stream stream);

extern "C" int skl_scan_run(skl_skeleton s, void* out, const void* in, size_t n, int inclusive,
                            const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_SCAN));
  return scan_launch(s, "main", out, in, n, inclusive, args, nargs, stream);
}

// The scan of the array read from its end (adjoint of Scan(+)).
int skl_scan_run_reverse(skl_skeleton s, void* out, const void* in, size_t n, int inclusive,
                         skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_SCAN));
  return scan_launch(s, "reverse", out, in, n, inclusive, nullptr, 0, stream);
}

int skl_run_backward_map(skl_skeleton s, void* d_in, const void* in, const void* d_out,
                         size_t n, const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_MAP));
  if (!s->has_deriv)
    return fail(SKL_ERR_INVALID, "Map backward needs a derivative function (deriv_source)");
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (n == 0) return SKL_OK;
  if (!d_in || !in || !d_out) return fail(SKL_ERR_INVALID, "NULL buffer");
  unsigned long long nn = n;
  skl_kernel k;
  SKL_TRY(get_kernel(s, "backward", "skl_map_backward", &k));
  auto a = with_extras({&in, &d_out, &d_in, &nn}, pk);
  return launch(k, blocks_for(n, (unsigned long long)s->wg * 4, s->max_blocks), 1, s->wg, 1,
                pk.dyn_lds, a, stream);
}

int skl_run_backward_zip(skl_skeleton s, void* d_left, void* d_right, const void* left,
                         const void* right, const void* d_out, size_t n, const skl_arg* args,
                         int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_ZIP));
  if (!s->has_deriv)
    return fail(SKL_ERR_INVALID, "Zip backward needs derivative functions (deriv_source)");
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (n == 0) return SKL_OK;
  if (!left || !right || !d_out) return fail(SKL_ERR_INVALID, "NULL buffer");
  unsigned long long nn = n;
  skl_kernel k;
  SKL_TRY(get_kernel(s, "backward", "skl_zip_backward", &k));
  auto a = with_extras({&left, &right, &d_out, &d_left, &d_right, &nn}, pk);
  return launch(k, blocks_for(n, (unsigned long long)s->wg * 4, s->max_blocks), 1, s->wg, 1,
                pk.dyn_lds, a, stream);
}

static int scan_launch(skl_skeleton s, const std::string& key, void* out, const void* in,
                       size_t n, int inclusive, const skl_arg* args, int nargs,
                       skl_stream stream)
{
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (n == 0) return SKL_OK;
  if (!in || !out) return fail(SKL_ERR_INVALID, "NULL buffer");
  if (pk.dyn_lds) return fail(SKL_ERR_UNSUPPORTED, "Scan user functions cannot take Local");
  size_t al = s->V * s->types[0].size;
  if (!aligned(in, al) || !aligned(out, al))
    return fail(SKL_ERR_INVALID, "Scan buffers must be aligned to %zu bytes", al);
  if (s->wg % 64 != 0)
    return fail(SKL_ERR_INVALID, "Scan work-group size %d is not a multiple of 64", s->wg);
  unsigned long long tile = (unsigned long long)s->wg * s->scan_j * s->V;
  unsigned long long tiles = (n + tile - 1) / tile;
  if (tiles > 0xFFFFFFFFull) return fail(SKL_ERR_INVALID, "Scan input too large");
  size_t words = (carry_size(s) + 3) / 4;
  SKL_TRY(ensure_scratch(s, 256 + tiles * words * 8));
  void* lb = static_cast<char*>(s->scratch) + 256;
  void* ctl = s->scratch;
  unsigned long long nn = n;
  int incl = inclusive ? 1 : 0;
  skl_kernel k;
  SKL_TRY(get_kernel(s, key, "skl_scan", &k));
  auto a = with_extras({&in, &out, &nn, &incl, &lb, &ctl}, pk);
  return launch(k, (unsigned)tiles, 1, (unsigned)s->wg, 1, 0, a, stream);
}

extern "C" int skl_mapoverlap_run(skl_skeleton s, void* out, const void* in, size_t rows,
                                  size_t cols, const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(check_kind(s, SKL_MAPOVERLAP));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (rows == 0 || cols == 0) return SKL_OK;
  if (!in || !out) return fail(SKL_ERR_INVALID, "NULL buffer");
  if (pk.dyn_lds) return fail(SKL_ERR_UNSUPPORTED, "MapOverlap user functions cannot take Local");
  if (rows > 0x7FFFFFFF || cols > 0x7FFFFFFF || rows * cols > (1ull << 40))
    return fail(SKL_ERR_INVALID, "matrix too large");
  int r = (int)rows, c = (int)cols;
  if (s->stencil_radius >= 0) {
    // rows per thread: the tuned value, on the instantiated 4 / 6 / 8
    const int rb = s->mo_rb >= 8 ? 8 : (s->mo_rb >= 6 ? 6 : 4);
    if ((rows + 4 * rb - 1) / (4 * rb) > 65535)
      return fail(SKL_ERR_INVALID, "matrix has too many rows");
    if (!args[0].value) return fail(SKL_ERR_INVALID, "NULL weight buffer");
    const int st = skl_internal_stencil_forward(
        (float*)out, (const float*)in, (const float*)args[0].value, r, c, s->stencil_radius,
        s->padding == SKL_PADDING_NEUTRAL ? 1 : 0, s->stencil_neutral,
        s->contract == "off" ? 0 : 1, rb, 2, resolve_stream(stream));
    if (st != SKL_OK) return fail(st, "MapOverlap: stencil kernel launch failed");
    return SKL_OK;
  }
  skl_kernel k;
  SKL_TRY(get_kernel(s, "main", "skl_mapoverlap", &k));
  auto a = with_extras({&in, &out, &r, &c}, pk);
  unsigned th = 4u * s->mo_rb;
  unsigned gy = (unsigned)((rows + th - 1) / th);
  if (gy > 65535) return fail(SKL_ERR_INVALID, "matrix has too many rows");
  return launch(k, (unsigned)((cols + 63) / 64), gy, 256, 1, 0, a, stream);
}

extern "C" int skl_allpairs_run(skl_skeleton s, void* out, const void* left, const void* right,
                                size_t rows, size_t cols, size_t dim, const skl_arg* args,
                                int nargs, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (s->kind != SKL_ALLPAIRS && s->kind != SKL_ALLPAIRS_GENERAL)
    return fail(SKL_ERR_INVALID, "skeleton is a %s, not an AllPairs", kind_name(s->kind));
  ArgPack pk;
  SKL_TRY(marshal_extras(s->extras, args, nargs, &pk));
  if (rows == 0 || cols == 0) return SKL_OK;
  if (dim == 0) return fail(SKL_ERR_INVALID, "AllPairs needs dim > 0 (AllPairsDef.h:127-129)");
  if (!left || !right || !out) return fail(SKL_ERR_INVALID, "NULL buffer");
  if (rows > 0x7FFFFFFF || cols > 0x7FFFFFFF || dim > 0x7FFFFFFF)
    return fail(SKL_ERR_INVALID, "matrix too large");
  int M = (int)rows, N = (int)cols, K = (int)dim;
  skl_kernel k;
  auto a = with_extras({&left, &right, &out, &M, &N, &K}, pk);
  if (s->kind == SKL_ALLPAIRS) {
    std::string key = first_key(s);
    // the (+,*) contraction on float from a zero identity is a matmul: the
    // native kernels (FP32 VALU with 8x8 register blocks, or the MFMA forms).
    // A non-zero identity starts the fold (AllPairsKernel.cl:63), so it
    // takes the in-order FMA fold at every precision.
    const bool product_form = s->precision == SKL_PREC_BF16X3 || s->precision == SKL_PREC_FP8;
    if ((key == "fma" || (float_addmul(s) && product_form)) && identity_is_zero(s->identity))
      return matmul_run((float*)out, (const float*)left, (const float*)right, rows, cols, dim,
                        0, 0, s->precision == SKL_PREC_FP64 ? SKL_PREC_DEFAULT : s->precision,
                        s->mm_tile, stream);
    if (pk.dyn_lds) return fail(SKL_ERR_UNSUPPORTED, "AllPairs user functions cannot take Local");
    SKL_TRY(get_kernel(s, key, "skl_allpairs", &k));
    unsigned gy = (unsigned)((rows + 63) / 64);
    if (gy > 65535) return fail(SKL_ERR_INVALID, "too many rows");
    return launch(k, (unsigned)((cols + 63) / 64), gy, 256, 1, 0, a, stream);
  }
  // general form: 64 columns in LDS with the rows read as wave-uniform
  // scalar loads, else 32 rows and 32 columns in LDS, when they fit
  const size_t lds_cols = 64 * dim * s->types[2].size;
  if (pk.dyn_lds == 0 && lds_cols <= 64 * 1024 && !s->simple) {
    SKL_TRY(get_kernel(s, "dim" + std::to_string(dim), "skl_allpairs_general_cols", &k));
    unsigned gy = (unsigned)((rows + 63) / 64);
    if (gy > 65535) return fail(SKL_ERR_INVALID, "too many rows");
    return launch(k, (unsigned)((cols + 63) / 64), gy, 1024, 1, lds_cols, a, stream);
  }
  size_t lds = ((32 * dim * s->types[1].size + 15) & ~(size_t)15) + 32 * dim * s->types[2].size;
  if (pk.dyn_lds == 0 && lds <= 64 * 1024 && !s->simple) {
    SKL_TRY(get_kernel(s, "dim" + std::to_string(dim), "skl_allpairs_general_lds", &k));
    unsigned gy = (unsigned)((rows + 31) / 32);
    if (gy > 65535) return fail(SKL_ERR_INVALID, "too many rows");
    return launch(k, (unsigned)((cols + 31) / 32), gy, 256, 1, lds, a, stream);
  }
  SKL_TRY(get_kernel(s, "main", "skl_allpairs_general", &k));
  unsigned gy = (unsigned)((rows + 15) / 16);
  if (gy > 65535) return fail(SKL_ERR_INVALID, "too many rows");
  return launch(k, (unsigned)((cols + 15) / 16), gy, 16, 16, pk.dyn_lds, a, stream);
}

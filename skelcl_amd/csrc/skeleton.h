// Internal state of a skeleton object (behind the opaque skl_skeleton).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fusion.h"
#include "skelcl_skeletons.h"

struct TypeInfo {
  std::string name;
  size_t size = 0;
};

struct skl_skeleton_s {
  int kind = 0;
  std::string src, func, src2, func2, identity, neutral;
  std::string deriv_src, deriv_func, deriv_func2, options;
  std::string common;  // common definitions registered when it was created
  std::string contract = "on";  // -ffp-contract of the user functions
  int stencil_checked = 0;  // MapOverlap: 1 = a weighted stencil, -1 = not, 0 = unknown
  // MapOverlap whose user function is the weighted stencil of this radius
  // (match_weighted_stencil), run by the native stencil kernel; -1 = not
  int stencil_radius = -1;
  float stencil_neutral = 0.0f;  // the neutral value, as the generated code rounds it
  std::vector<TypeInfo> types;  // kind-specific order, see skelcl_skeletons.h
  int overlap = 1, padding = SKL_PADDING_NEAREST;
  int wg = 256;
  int precision = SKL_PREC_DEFAULT;

  skl::UserFunction uf, uf2;          // main (Reduce for AllPairs) / Zip
  skl::UserFunction duf, duf2;        // derivative functions
  bool has_deriv = false;
  std::vector<skl::Param> extras;     // extra parameters, in call order
  bool simple = false;                // needs one element per work-item
  skl::UserOp op = skl::UserOp::Other, op2 = skl::UserOp::Other;

  // generated programs: key -> (source, program)
  struct Prog {
    std::string source;
    skl_program program = nullptr;
  };
  std::map<std::string, Prog> progs;

  // launch shape (autotuned or defaults)
  int V = 1;        // elements per 16-byte pack
  int unroll = 4;   // packs in flight per thread (Map/Zip/Reduce)
  int scan_j = 4;   // packs per lane (Scan tile = 256*J*V)
  int mo_rb = 4;    // MapOverlap rows per thread
  int mm_tile = 0;  // AllPairs (+,*) MFMA tile (skl_internal_matmul), 0 = size rule
  int max_blocks = 0;

  // scratch: [0,8) ticket, [8,12) epoch, partials / look-back words at 256
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // scratch of the batched segmented forms (zeroed when allocated)
  void* seg_scratch = nullptr;
  size_t seg_scratch_bytes = 0;
  // scratch of one-launch batches (skl_batch_run; zeroed when allocated)
  void* batch_scratch = nullptr;
  size_t batch_scratch_bytes = 0;
  // scratch buffers outgrown by later calls, freed with the skeleton: a
  // captured graph may still use them
  std::vector<void*> retired;
  // float (+) without extra arguments
  bool float_add() const;
  // its DEFAULT form: compensated pairs in a block, FP64 across blocks
  bool precise_add() const;
  // its FP64 form (precision FP64): FP64 throughout
  bool fp64_add() const;
  // default Scan tile depth: about 256 bytes of accumulators per lane
  int default_scan_j() const;
  // launch shape defaults per kind (unroll, max_blocks, scan_j, mo_rb)
  void apply_default_shape(int cus);
};

namespace skl {
// Drops the generated programs and builds the first one for the current
// launch shape and options (skeletons.cpp).
int rebuild_programs(skl_skeleton_s* s);
// Reduce with caller-provided scratch: no allocation and no build
// (skeletons.cpp); the layout needs `counter_bytes` of zeroed counters at
// the start of `total_bytes`.
// AllPairs (+,*) / matmul on float with the MFMA tile `tile` (0: by size)
// (extensions.cpp)
int matmul_run(float* C, const float* A, const float* B, size_t M, size_t N, size_t K,
               int trans_a, int trans_b, int precision, int tile, skl_stream stream);
int reduce_scratch_need(skl_skeleton_s* s, size_t n, size_t* counter_bytes, size_t* total_bytes);
int reduce_run_fixed_scratch(skl_skeleton_s* s, void* out, const void* in, size_t n,
                             const skl_arg* args, int nargs, skl_stream stream, void* scratch,
                             size_t scratch_bytes);
}  // namespace skl

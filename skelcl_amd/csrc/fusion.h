// User-function analysis and fused-kernel source generation.
//
// Replaces what the reference does with libclang source rewriting
// (libraries/stooling/src/SourceCode.cpp:98-283): finding the user function,
// appending its extra parameters to the skeleton kernel
// (TransferParametersCallback.cpp:54-114), forwarding them in the call
// (TransferArgumentsCallback.cpp:54-112), renaming the function
// (RenameFunctionCallback.cpp:57-107), retyping SCL_TYPE_i
// (RedefineTypedefCallback.cpp:58-70) and expanding T_matrix_t parameters
// (FixKernelParameterCallback.cpp:62-116).  Here a small tokenizer reads the
// signature and the generator emits macros the hand-written kernel templates
// expand, so the user source itself is compiled unchanged.
#pragma once

#include <string>
#include <vector>

namespace skl {

enum class ParamKind { Value, Buffer, Matrix, Local };

struct Param {
  std::string decl;       // as written, e.g. "__global float* out"
  std::string type_text;  // declaration without the name
  std::string name;
  ParamKind kind = ParamKind::Value;
  size_t value_size = 0;  // known size of a by-value scalar/vector type, else 0
  std::string pointee;    // element type of a pointer / __local parameter
};

struct UserFunction {
  std::string name;
  std::string return_type;
  std::vector<Param> params;
  std::string body;       // text between the braces, comments removed
};

// Operation a user function denotes, for dispatch to specialised kernels.
enum class UserOp { Other, Add, Mul, Max, Min };

// Removes // and /* */ comments (string/char literals kept).
std::string strip_comments(const std::string& src);

// Finds the definition of `name` in `src` and reads its signature.
bool parse_user_function(const std::string& src, const std::string& name,
                         UserFunction* out, std::string* err);

// Recognises `return a+b;`-style bodies of two-parameter functions.
UserOp classify_binary(const UserFunction& f);

// Recognises the weighted-stencil MapOverlap user function, token for token:
//   float F(input_matrix_t IMG, __global float* W) {
//     float S = 0.0f;
//     for (int Y = -N; Y <= N; ++Y)
//       for (int X = -N; X <= N; ++X)
//         S += getData(IMG, X, Y) * W[(Y + N) * (2N+1) + (X + N)];
//     return S;
//   }
// (any names; braces around the loop bodies optional; the product's
// operands in either order; `S = S + ...` for `+=`; `Y++` for `++Y`; a zero
// literal of any spelling; `const` on the weights).  Returns N, or -1 for any
// other source -- also for one with preprocessor lines, which could change
// what the tokens mean.
int match_weighted_stencil(const std::string& src, const UserFunction& f);

// True when the source calls work-item functions, barriers or uses __local,
// i.e. it needs the reference's one-element-per-work-item launch.
bool uses_work_item_builtins(const std::string& src);

// Size of a by-value type name, 0 when unknown.
size_t known_type_size(const std::string& type_text);

// Helpers for the generated extra-parameter plumbing.
struct ExtraPlumbing {
  std::string params;  // ", T0 skl_xa0, ..."  (kernel parameter list tail)
  std::string args;    // ", skl_xa0, ..."     (call argument tail)
  std::string setup;   // declarations of __local pointers at kernel entry
};
ExtraPlumbing make_extra_plumbing(const std::vector<Param>& extras,
                                  const std::string& prefix);

}  // namespace skl

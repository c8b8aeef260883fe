// Minimal test harness with gtest-style macros (gtest is not available):
// TEST(Suite, Name) registers a test, EXPECT_* record failures, main() runs
// the selected tests and returns non-zero when any failed.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

namespace skltest {
struct Case {
  const char* suite;
  const char* name;
  std::function<void()> fn;
};
inline std::vector<Case>& registry()
{
  static std::vector<Case> r;
  return r;
}
inline int& failures()
{
  static int f = 0;
  return f;
}
struct Reg {
  Reg(const char* s, const char* n, std::function<void()> f) { registry().push_back({s, n, f}); }
};
inline int run_all(int argc, char** argv)
{
  int failed_tests = 0, ran = 0;
  for (auto& c : registry()) {
    std::string full = std::string(c.suite) + "." + c.name;
    if (argc > 1 && full.find(argv[1]) == std::string::npos) continue;
    int before = failures();
    std::printf("[ RUN      ] %s\n", full.c_str());
    std::fflush(stdout);
    c.fn();
    ++ran;
    if (failures() != before) {
      ++failed_tests;
      std::printf("[  FAILED  ] %s\n", full.c_str());
    } else {
      std::printf("[       OK ] %s\n", full.c_str());
    }
    std::fflush(stdout);
  }
  std::printf("[==========] %d tests ran, %d failed\n", ran, failed_tests);
  return failed_tests ? 1 : 0;
}
}  // namespace skltest

#define TEST(S, N)                                                             \
  static void S##_##N##_impl();                                                \
  static skltest::Reg S##_##N##_reg(#S, #N, S##_##N##_impl);                   \
  static void S##_##N##_impl()

#define SKL_FAIL_AT(msg)                                                       \
  do {                                                                         \
    if (skltest::failures()++ < 20)                                            \
      std::printf("%s:%d: failure: %s\n", __FILE__, __LINE__, (msg).c_str());  \
  } while (0)

#define EXPECT_EQ(a, b)                                                        \
  do {                                                                         \
    auto _va = (a);                                                            \
    auto _vb = (b);                                                            \
    if (!(_va == _vb)) {                                                       \
      std::ostringstream _o;                                                   \
      _o << #a << " == " << #b << " (" << _va << " vs " << _vb << ")";         \
      SKL_FAIL_AT(_o.str());                                                   \
    }                                                                          \
  } while (0)
#define EXPECT_TRUE(c)                                                         \
  do {                                                                         \
    if (!(c)) SKL_FAIL_AT(std::string(#c));                                    \
  } while (0)
#define EXPECT_FALSE(c) EXPECT_TRUE(!(c))
#define EXPECT_LE(a, b) EXPECT_TRUE((a) <= (b))
#define EXPECT_NEAR(a, b, tol) EXPECT_TRUE(std::fabs((double)(a) - (double)(b)) <= (tol))
#define ASSERT_EQ(a, b) EXPECT_EQ(a, b)

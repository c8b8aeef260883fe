// The reference's skeleton tests (test/MapTests.cpp, ZipTests.cpp,
// ReduceTests.cpp, ScanTests.cpp, MapOverlapTests.cpp, AllPairsTests.cpp,
// IndexVectorTests.cpp, IndexMatrixTests.cpp, VectorTests.cpp) restated
// against the skelcl_amd C++ API, which compiles the same host code.
#include <SkelCL/Autodiff.h>
#include <SkelCL/SkelCL.h>

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "skelcl_test.h"

// ------------------------------------------------------------------ Map
TEST(MapTest, SimpleMap)
{
  skelcl::Map<float(float)> m{"float func(float f){ return -f; }"};
  skelcl::Vector<float> input(100);
  for (size_t i = 0; i < input.size(); ++i) input[i] = i * 2.5f;
  skelcl::Vector<float> output = m(input);
  EXPECT_EQ(input.size(), output.size());
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(-input[i], output[i]);
}

TEST(MapTest, CreateMapWithFile)
{
  const char* fn = "/tmp/skelcl_amd_map_test.cl";
  {
    std::ofstream f(fn);
    f << "float func(float f) { return -f; }";
  }
  skelcl::Map<float(float)> m{std::ifstream{fn}};
  skelcl::Vector<float> input(10, 1.5f);
  skelcl::Vector<float> output = m(input);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(-1.5f, output[i]);
}

TEST(MapTest, SimpleMapWithOut)
{
  skelcl::Map<float(float)> m{"float func(float f){ return -f; }"};
  skelcl::Vector<float> input(10);
  for (size_t i = 0; i < input.size(); ++i) input[i] = i * 2.5f;
  skelcl::Vector<float> output(10);
  m(skelcl::out(output), input);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(-input[i], output[i]);
}

TEST(MapTest, DeviceId)
{
  skelcl::Map<int(float)> m{"int func(float f) { return skelcl_get_device_id(); }"};
  skelcl::Vector<float> input(10);
  skelcl::Vector<int> output = m(input);
  EXPECT_EQ((size_t)10, output.size());
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(0, output[i]);
}

TEST(MapTest, AddArgs)
{
  skelcl::Map<float(float)> m{"float func(float f, float add, float add2)\
                                 { return f+add+add2; }"};
  skelcl::Vector<float> input(10);
  for (size_t i = 0; i < input.size(); ++i) input[i] = i * 2.5f;
  float add = 5.25f, add2 = 7.75f;
  skelcl::Vector<float> output = m(input, add, add2);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(input[i] + add + add2, output[i]);
}

TEST(MapTest, MapVoid)
{
  skelcl::Map<void(float)> m{"void func(float f, __global float* out) \
    { out[get_global_id(0)] = f; }"};
  skelcl::Vector<float> input(10);
  for (size_t i = 0; i < input.size(); ++i) input[i] = i * 2.5f;
  skelcl::Vector<float> output(10);
  m(input, skelcl::out(output));
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(input[i], output[i]);
}

TEST(MapTest, MatrixAddArgsMatrix)
{
  skelcl::Map<float(float)> m(R"(
float func( float f, float_matrix_t mat, float add2, int_matrix_t mat2 )
{
  return f + get(mat, 1, 1) + add2 + get(mat2, 1, 2);
}
)");
  std::vector<float> vec(10);
  for (size_t i = 0; i < 10; ++i) vec[i] = static_cast<float>(i * 2.5f);
  skelcl::Matrix<float> input(vec, 3);
  EXPECT_TRUE(skelcl::Matrix<float>::size_type(4, 3) == input.size());
  std::vector<float> vec2(10);
  for (size_t i = 0; i < 10; ++i) vec2[i] = static_cast<float>(i);
  skelcl::Matrix<float> mat(vec2, 3);
  skelcl::Matrix<int> mat2(skelcl::MatrixSize{2, 3});
  mat2[1][2] = 5;
  float add2 = 7.75f;
  skelcl::Matrix<float> output = m(input, mat, add2, mat2);
  EXPECT_TRUE(input.size() == output.size());
  for (size_t i = 0; i < output.rowCount(); ++i)
    for (size_t j = 0; j < output.columnCount(); ++j)
      EXPECT_EQ(input[i][j] + mat({1, 1}) + add2 + mat2[1][2], output[i][j]);
}

TEST(MapTest, MapWithLocalMemory)
{
  skelcl::Map<int(int)> map{"int func(int i, __local int* lp) \
    { lp[0] = i; barrier(CLK_LOCAL_MEM_FENCE); return -lp[0]; }"};
  map.setWorkGroupSize(16);
  skelcl::Vector<int> input(32);
  for (size_t i = 0; i < input.size(); ++i) input[i] = 5;
  skelcl::Vector<int> output = map(input, skelcl::Local(sizeof(int)));
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(-input[i], output[i]);
}

TEST(MapTest, MapFloat4)
{
  skelcl::Map<skelcl::float4(skelcl::float4)> m{"float4 func(float4 f) { return -f; }"};
  skelcl::Vector<skelcl::float4> input(10);
  for (size_t i = 0; i < input.size(); ++i) {
    input[i].s0 = 1.0f;
    input[i].s1 = 2.0f;
    input[i].s2 = 3.0f;
    input[i].s3 = 4.0f;
  }
  skelcl::Vector<skelcl::float4> output = m(input);
  EXPECT_EQ(input.size(), output.size());
  for (size_t i = 0; i < output.size(); ++i) {
    EXPECT_EQ(-input[i].s0, output[i].s0);
    EXPECT_EQ(-input[i].s1, output[i].s1);
    EXPECT_EQ(-input[i].s2, output[i].s2);
    EXPECT_EQ(-input[i].s3, output[i].s3);
  }
}

TEST(MapTest, MapDouble4)
{
  skelcl::Map<skelcl::double4(skelcl::double4)> m{"double4 func(double4 f) { return -f; }"};
  skelcl::Vector<skelcl::double4> input(10);
  for (size_t i = 0; i < input.size(); ++i) input[i] = skelcl::double4{1.0, 2.0, 3.0, 4.0};
  skelcl::Vector<skelcl::double4> output = m(input);
  for (size_t i = 0; i < output.size(); ++i) {
    EXPECT_EQ(-1.0, output[i].s0);
    EXPECT_EQ(-4.0, output[i].s3);
  }
}

TEST(MapTest, TempNotDestroyedEarly)  // MapTests.cpp:197-224
{
  auto execute = [](const skelcl::Vector<float>& input) {
    skelcl::Map<float(float)> m{"float func(float f) { return -f; }"};
    return m(input);
  };
  skelcl::Vector<float> input(1000, 2.0f);
  skelcl::Vector<float> output = execute(input);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(-2.0f, output[i]);
}

// ------------------------------------------------------------------ Zip
TEST(ZipTest, SimpleZip)
{
  skelcl::Zip<float(float, float)> z("float func(float x, float y){ return x+y; }");
  skelcl::Vector<float> left(100), right(100);
  for (size_t i = 0; i < 100; ++i) {
    left[i] = i;
    right[i] = 2 * i;
  }
  skelcl::Vector<float> output = z(left, right);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ(left[i] + right[i], output[i]);
}

TEST(ZipTest, AddArgs)
{
  skelcl::Zip<float(float, float)> z(
      "float func(float x, float y, float a, float b){ return x*y+a+b; }");
  skelcl::Vector<float> left(100), right(100);
  for (size_t i = 0; i < 100; ++i) {
    left[i] = i;
    right[i] = 2 * i;
  }
  float a = 5.5f, b = 3.25f;
  skelcl::Vector<float> output = z(left, right, a, b);
  for (size_t i = 0; i < output.size(); ++i)
    EXPECT_EQ(left[i] * right[i] + a + b, output[i]);
}

TEST(ZipTest, SimpleZipWithOutput)
{
  skelcl::Zip<float(float, float)> z("float func(float x, float y) { return x; }");
  skelcl::Vector<float> left(50), right(50);
  for (size_t i = 0; i < 50; ++i) {
    left[i] = i;
    right[i] = 7;
  }
  z(skelcl::out(left), left, right);
  for (size_t i = 0; i < left.size(); ++i) EXPECT_EQ((float)i, left[i]);
}

TEST(ZipTest, SimpleMatrixZip)
{
  skelcl::Zip<float(float, float)> z("float func(float mat1,float mat2)"
                                     "{ return mat1 + mat2; }");
  skelcl::Matrix<float> a(skelcl::MatrixSize{8, 8});
  for (size_t i = 0; i < 8; ++i)
    for (size_t j = 0; j < 8; ++j) a[i][j] = i * 8 + j;
  skelcl::Matrix<float> c = z(a, a);
  for (size_t i = 0; i < 8; ++i)
    for (size_t j = 0; j < 8; ++j) EXPECT_EQ(2.0f * (i * 8 + j), c[i][j]);
}

TEST(ZipTest, VoidZip)
{
  skelcl::Zip<void(float, float)> z("void func(float x, float y, __global float* out)"
                                    "{ out[get_global_id(0)] = x+y; }");
  skelcl::Vector<float> left(20), right(20), output(20);
  for (size_t i = 0; i < 20; ++i) {
    left[i] = i;
    right[i] = i;
  }
  z(left, right, skelcl::out(output));
  for (size_t i = 0; i < 20; ++i) EXPECT_EQ(2.0f * i, output[i]);
}

// ------------------------------------------------------------------ Reduce
TEST(ReduceTest, SimpleReduce)
{
  skelcl::Reduce<float(float)> r("float func(float x, float y){ return x+y; }");
  skelcl::Vector<float> input(100);
  for (size_t i = 0; i < input.size(); ++i) input[i] = i;
  skelcl::Vector<float> output = r(input);
  EXPECT_LE((size_t)1, output.size());
  EXPECT_EQ(4950, output[0]);
}

TEST(ReduceTest, AdditionalArg)
{
  skelcl::Reduce<float(float)> r("float func(float x, float y, float a){ return a*(x+y); }");
  skelcl::Vector<float> input(100);
  for (size_t i = 0; i < input.size(); ++i) input[i] = i;
  skelcl::Vector<float> output = r(input, 1.0f);
  EXPECT_EQ(4950, output[0]);
}

TEST(ReduceTest, SimpleReduce2)
{
  skelcl::Reduce<int(int)> r("int func(int x, int y){ return x+y; }");
  skelcl::Vector<int> input(1587);
  for (unsigned int i = 0; i < input.size(); ++i) input[i] = i;
  skelcl::Vector<int> output = r(input);
  EXPECT_EQ(1258491, output[0]);
}

TEST(ReduceTest, LongReduce)
{
  skelcl::Reduce<int(int)> r("int func(int x, int y){ return x+y; }");
  skelcl::Vector<int> input(100000000, 1);
  skelcl::Vector<int> output = r(input);
  EXPECT_EQ(100000000, output[0]);
}

TEST(ReduceTest, SmallReduces)
{
  skelcl::Reduce<float(float)> r("float func(float x, float y){ return x+y; }");
  for (size_t size = 1; size < 256; ++size) {
    skelcl::Vector<float> input(size, 1.0f);
    skelcl::Vector<float> output = r(input);
    EXPECT_EQ((float)size, output[0]);
  }
}

TEST(ReduceTest, MidSizeReduces)
{
  skelcl::Reduce<float(float)> r("float func(float x, float y){ return x+y; }");
  for (size_t size = 8192; size < 8448; ++size) {
    skelcl::Vector<float> input(size, 1.0f);
    skelcl::Vector<float> output = r(input);
    EXPECT_EQ((float)size, output[0]);
  }
}

TEST(ReduceTest, FusedZipReduce)
{
  skelcl::Zip<float(float, float)> mult("float func(float x, float y){ return x*y; }");
  skelcl::Reduce<float(float)> sum("float func(float x, float y){ return x+y; }", "0");
  skelcl::Vector<float> a(1000), b(1000);
  for (size_t i = 0; i < 1000; ++i) {
    a[i] = i % 7;
    b[i] = i % 5;
  }
  float exp = 0;
  for (size_t i = 0; i < 1000; ++i) exp += (i % 7) * (i % 5);
  EXPECT_EQ(exp, sum(mult(a, b))[0]);       // examples/dot_product/main.cpp:105-116
  EXPECT_EQ(exp, sum.fused(mult, a, b)[0]);  // fused extension
}

// ------------------------------------------------------------------ Scan
TEST(ScanTest, SimpleScan)
{
  skelcl::Scan<int(int)> s{"int func(int x, int y){ return x+y; }"};
  skelcl::Vector<int> input(1024);
  for (size_t i = 0; i < input.size(); ++i) input[i] = 1;
  skelcl::Vector<int> output = s(input);
  input.dataOnDeviceModified();
  EXPECT_EQ((size_t)1024, output.size());
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ((int)i, output[i]);
}

TEST(ScanTest, SmallScan)
{
  skelcl::Scan<int(int)> s{"int func(int x, int y){ return x+y; }"};
  skelcl::Vector<int> input(10);
  for (size_t i = 0; i < input.size(); ++i) input[i] = 1;
  skelcl::Vector<int> output = s(input);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ((int)i, output[i]);
}

TEST(ScanTest, InclusiveExtension)
{
  skelcl::Scan<float(float)> s{"float func(float x, float y){ return x+y; }", "0.0f"};
  skelcl::Vector<float> input(5000, 1.0f);
  skelcl::Vector<float> output = s.inclusive(input);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ((float)(i + 1), output[i]);
}

// ------------------------------------------------------------------ MapOverlap
template <typename F>
static void shiftTest(const char* src, unsigned range, int dx, int dy, size_t size, F init)
{
  skelcl::MapOverlap<int(int)> m{src, range};
  skelcl::Matrix<int> input(skelcl::MatrixSize{size, size});
  for (size_t i = 0; i < size; ++i)
    for (size_t j = 0; j < size; ++j) input[i][j] = init(i, j);
  skelcl::Matrix<int> output = m(input);
  EXPECT_EQ(size, output.rowCount());
  EXPECT_EQ(size, output.columnCount());
  for (long i = 0; i < (long)size; ++i)
    for (long j = 0; j < (long)size; ++j) {
      long r = std::min(std::max(i + dy, 0L), (long)size - 1);
      long c = std::min(std::max(j + dx, 0L), (long)size - 1);
      EXPECT_EQ(input[r][c], output[i][j]);
    }
}

TEST(MapOverlapTest, SimpleMatrixRightShift)
{
  shiftTest("int func(input_matrix_t f){ return getData(f, -1, 0); }", 1, -1, 0, 1024,
            [](size_t i, size_t j) { return (int)(i * j + i); });
}
TEST(MapOverlapTest, SimpleMatrixTwoRightShift)
{
  shiftTest("int func(input_matrix_t f){ return getData(f, -2, 0); }", 2, -2, 0, 100,
            [](size_t, size_t j) { return (int)j; });
}
TEST(MapOverlapTest, SimpleMatrixLeftShift)
{
  shiftTest("int func(input_matrix_t f){ return getData(f, +1, 0); }", 1, 1, 0, 100,
            [](size_t, size_t j) { return (int)j; });
}
TEST(MapOverlapTest, SimpleMatrixTwoLeftShift)
{
  shiftTest("int func(input_matrix_t f){ return getData(f, +2, 0); }", 2, 2, 0, 100,
            [](size_t, size_t j) { return (int)j; });
}
TEST(MapOverlapTest, SimpleMatrixDownShift)
{
  shiftTest("int func(input_matrix_t f){ return getData(f, 0, -1); }", 1, 0, -1, 100,
            [](size_t i, size_t) { return (int)i; });
}
TEST(MapOverlapTest, SimpleMatrixTwoUpShift)
{
  shiftTest("int func(input_matrix_t f){ return getData(f, 0, +2); }", 2, 0, 2, 100,
            [](size_t i, size_t) { return (int)i; });
}

TEST(MapOverlapTest, NeutralPadding)
{
  skelcl::MapOverlap<int(int)> m{"int func(input_matrix_t f){ return getData(f, -1, -1); }", 1,
                                 skelcl::detail::Padding::NEUTRAL, -7};
  skelcl::Matrix<int> input(skelcl::MatrixSize{37, 53});
  for (size_t i = 0; i < 37; ++i)
    for (size_t j = 0; j < 53; ++j) input[i][j] = (int)(i * 1000 + j);
  skelcl::Matrix<int> output = m(input);
  for (size_t i = 0; i < 37; ++i)
    for (size_t j = 0; j < 53; ++j)
      EXPECT_EQ((i == 0 || j == 0) ? -7 : input[i - 1][j - 1], output[i][j]);
}

// ------------------------------------------------------------------ AllPairs
static void fillMatrix(skelcl::Matrix<float>& m, unsigned seed)
{
  std::srand(seed);
  for (size_t i = 0; i < m.rowCount(); ++i)
    for (size_t j = 0; j < m.columnCount(); ++j) m[i][j] = (float)(std::rand() % 10);
}

static void allpairsTest(size_t M, size_t K, size_t N)
{
  skelcl::Zip<float(float, float)> zip("float func(float x, float y){ return x*y; }");
  skelcl::Reduce<float(float)> reduce("float func(float x, float y){ return x+y; }");
  skelcl::AllPairs<float(float, float)> allpairs(reduce, zip);
  skelcl::Matrix<float> left(skelcl::MatrixSize{M, K}), right(skelcl::MatrixSize{K, N});
  fillMatrix(left, 1);
  fillMatrix(right, 2);
  skelcl::Matrix<float> output = allpairs(left, right);
  EXPECT_EQ(M, output.rowCount());
  EXPECT_EQ(N, output.columnCount());
  for (size_t i = 0; i < M; ++i)
    for (size_t j = 0; j < N; ++j) {
      float s = 0;
      for (size_t k = 0; k < K; ++k) s += left[i][k] * right[k][j];
      EXPECT_EQ(s, output[i][j]);
    }
}

TEST(AllPairsTest, SquareMatrices) { allpairsTest(64, 64, 64); }
TEST(AllPairsTest, RectangularMatrices) { allpairsTest(32, 128, 32); }
TEST(AllPairsTest, OddMatrices) { allpairsTest(100, 60, 11); }

TEST(AllPairsTest, ZipReduceArgs)  // AllPairsTests.cpp:136-173
{
  skelcl::Zip<float(float, float)> zip("float func(float x, float y, float a){ return x*y+a; }");
  skelcl::Reduce<float(float)> reduce("float func(float x, float y, float b){ return x+y+b; }");
  skelcl::AllPairs<float(float, float)> allpairs(reduce, zip);
  skelcl::Matrix<float> left(skelcl::MatrixSize{16, 8}), right(skelcl::MatrixSize{8, 12});
  fillMatrix(left, 3);
  fillMatrix(right, 4);
  float a = 1, b = 2;
  skelcl::Matrix<float> output = allpairs(left, right, b, a);
  for (size_t i = 0; i < 16; ++i)
    for (size_t j = 0; j < 12; ++j) {
      float s = 0;
      for (size_t k = 0; k < 8; ++k) s = s + (left[i][k] * right[k][j] + a) + b;
      EXPECT_EQ(s, output[i][j]);
    }
}

TEST(AllPairsTest, GeneralAllPairs)  // AllPairsTests.cpp:181-223
{
  skelcl::AllPairs<float(float, float)> allpairs(
      "float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) {"
      "float res = 0;"
      "for (int i = 0; i < dim; ++i) { res += getElementFromRow(row, i) *"
      " getElementFromColumn(col, i); }"
      "return res; }");
  skelcl::Matrix<float> left(skelcl::MatrixSize{50, 20}), right(skelcl::MatrixSize{20, 40});
  fillMatrix(left, 5);
  fillMatrix(right, 6);
  skelcl::Matrix<float> output = allpairs(left, right);
  for (size_t i = 0; i < 50; ++i)
    for (size_t j = 0; j < 40; ++j) {
      float s = 0;
      for (size_t k = 0; k < 20; ++k) s += left[i][k] * right[k][j];
      EXPECT_EQ(s, output[i][j]);
    }
}

TEST(AllPairsTest, GeneralAllPairsExtraArg)  // AllPairsTests.cpp:226-269
{
  skelcl::AllPairs<float(float, float)> allpairs(
      "float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim, int start) {"
      "float res = 0;"
      "for (int i = start; i < dim; ++i) { res += getElementFromRow(row, i) *"
      " getElementFromColumn(col, i); }"
      "return res; }");
  skelcl::Matrix<float> left(skelcl::MatrixSize{20, 10}), right(skelcl::MatrixSize{10, 30});
  fillMatrix(left, 7);
  fillMatrix(right, 8);
  skelcl::Matrix<float> output = allpairs(left, right, 3);
  for (size_t i = 0; i < 20; ++i)
    for (size_t j = 0; j < 30; ++j) {
      float s = 0;
      for (size_t k = 3; k < 10; ++k) s += left[i][k] * right[k][j];
      EXPECT_EQ(s, output[i][j]);
    }
}

// ------------------------------------------------------------------ Index containers
TEST(IndexVectorTest, MapIndex)
{
  skelcl::IndexVector index(1000);
  skelcl::Map<int(skelcl::Index)> m("int func(Index i){ return (int)i * 2; }");
  skelcl::Vector<int> output = m(index);
  for (size_t i = 0; i < output.size(); ++i) EXPECT_EQ((int)(2 * i), output[i]);
}

TEST(IndexMatrixTest, MapIndexPoint)
{
  skelcl::IndexMatrix index(skelcl::MatrixSize{21, 33});
  skelcl::Map<int(skelcl::IndexPoint)> m("int func(IndexPoint p){ return p.x + 100 * p.y; }");
  skelcl::Matrix<int> output = m(index);
  for (size_t r = 0; r < 21; ++r)
    for (size_t c = 0; c < 33; ++c) EXPECT_EQ((int)(c + 100 * r), output[r][c]);
}

// ------------------------------------------------------------------ containers
TEST(VectorTest, LazyCoherence)
{
  skelcl::Map<int(int)> inc("int func(int x){ return x + 1; }");
  skelcl::Vector<int> v(100, 1);
  skelcl::Vector<int> w = inc(v);  // v uploaded
  EXPECT_EQ(2, w[0]);
  v[0] = 41;                 // host write...
  v.dataOnHostModified();    // ...announced, as the reference requires
  skelcl::Vector<int> x = inc(v);
  EXPECT_EQ(42, x[0]);
  EXPECT_EQ(2, x[99]);
  skelcl::Vector<int> copy = x;  // deep copy through the host
  EXPECT_EQ(42, copy[0]);
}

// ------------------------------------------------------------------ autodiff
TEST(AutodiffTest, FunctionPairs)
{
  using namespace skelcl::autodiff;
  // Map: y = x*x, dy/dx = 2x
  skelcl::Map<float(float)> sq("float func(float x){ return x*x; }", "func",
                               "float d(float x){ return 2.0f*x; }", "d");
  skelcl::Vector<float> x(1000), g(1000, 1.0f);
  for (size_t i = 0; i < 1000; ++i) x[i] = 0.001f * i;
  Context c1;
  skelcl::Vector<float> y = MapFn<float>::forward(c1, sq, x);
  skelcl::Vector<float> dx = MapFn<float>::backward(c1, sq, g);
  for (size_t i = 0; i < 1000; ++i) EXPECT_EQ(2.0f * x[i], dx[i]);
  // Reduce(+): broadcast
  skelcl::Reduce<float(float)> sum("float func(float x, float y){ return x+y; }");
  Context c2;
  skelcl::Vector<float> s = ReduceSumFn<float>::forward(c2, sum, x);
  skelcl::Vector<float> seed(1, 3.0f);
  skelcl::Vector<float> ds = ReduceSumFn<float>::backward(c2, sum, seed);
  EXPECT_EQ((size_t)1000, ds.size());
  EXPECT_EQ(3.0f, ds[999]);
  // Scan(+) exclusive: dx_i = sum_{j>i} dy_j
  skelcl::Scan<float(float)> scan("float func(float x, float y){ return x+y; }", "0.0f");
  Context c3;
  skelcl::Vector<float> sc = ScanSumFn<float>::forward(c3, scan, x, false);
  skelcl::Vector<float> dsc = ScanSumFn<float>::backward(c3, scan, g);
  for (size_t i = 0; i < 1000; ++i) EXPECT_EQ((float)(999 - i), dsc[i]);
  // MatMul: dA = dC . B^T
  skelcl::Zip<float(float, float)> mul("float func(float x, float y){ return x*y; }");
  skelcl::AllPairs<float(float, float)> mm(sum, mul);
  skelcl::Matrix<float> A(skelcl::MatrixSize{4, 3}, 1.0f), B(skelcl::MatrixSize{3, 5}, 2.0f);
  Context c4;
  skelcl::Matrix<float> C = MatMulFn::forward(c4, mm, A, B);
  EXPECT_EQ(6.0f, C[3][4]);
  skelcl::Matrix<float> dC(skelcl::MatrixSize{4, 5}, 1.0f);
  auto grads = MatMulFn::backward(c4, mm, dC);
  EXPECT_EQ(10.0f, grads.first[0][0]);   // sum_j dC[i][j] * B[k][j] = 5 * 2
  EXPECT_EQ(4.0f, grads.second[2][4]);   // sum_i A[i][k] * dC[i][j] = 4 * 1
  // Zip: f(a, b) = a*b + 3a, df/da = b + 3, df/db = a
  skelcl::Zip<float(float, float)> z("float func(float a, float b){ return a*b + 3.0f*a; }",
                                     "func",
                                     "float da(float a, float b){ return b + 3.0f; }\n"
                                     "float db(float a, float b){ return a; }",
                                     "da", "db");
  skelcl::Vector<float> a(300), b(300), gz(300);
  for (size_t i = 0; i < 300; ++i) {
    a[i] = 0.5f * (float)i;
    b[i] = 2.0f - 0.25f * (float)i;
    gz[i] = (float)(i % 7) - 3.0f;
  }
  Context c5;
  skelcl::Vector<float> zz = ZipFn<float>::forward(c5, z, a, b);
  EXPECT_EQ(a[10] * b[10] + 3.0f * a[10], zz[10]);
  auto zg = ZipFn<float>::backward(c5, z, gz);
  for (size_t i = 0; i < 300; ++i) {
    EXPECT_EQ(gz[i] * (b[i] + 3.0f), zg.first[i]);
    EXPECT_EQ(gz[i] * a[i], zg.second[i]);
  }
  // Stencil: a 3x3 weighted sum under NEUTRAL zero padding; its adjoint is
  // the transposed stencil (each weight reaches the mirrored neighbour)
  skelcl::MapOverlap<float(float)> st(
      "float func(input_matrix_t m, __global float* w) { float s = 0.0f;"
      " for (int dy = -1; dy <= 1; ++dy) for (int dx = -1; dx <= 1; ++dx)"
      " s += getData(m, dx, dy) * w[(dy + 1) * 3 + (dx + 1)]; return s; }",
      1, skelcl::detail::Padding::NEUTRAL, 0.0f);
  skelcl::Vector<float> w(9);
  for (size_t i = 0; i < 9; ++i) w[i] = (float)(i + 1);
  skelcl::Matrix<float> img(skelcl::MatrixSize{6, 7}, 1.0f);
  Context c6;
  skelcl::Matrix<float> so = StencilFn::forward(c6, st, img, w);
  EXPECT_EQ(45.0f, so[2][3]);                        // interior: sum of all weights
  EXPECT_EQ(5.0f + 6.0f + 8.0f + 9.0f, so[0][0]);    // corner: only dy, dx >= 0
  skelcl::Matrix<float> dy(skelcl::MatrixSize{6, 7}, 0.0f);
  dy[3][3] = 1.0f;  // gradient of out[3][3] alone
  skelcl::Matrix<float> dimg = StencilFn::backward(c6, st, dy);
  // out[3][3] reads img[3+dy][3+dx] with weight w[(dy+1)*3 + dx+1]
  EXPECT_EQ(1.0f, dimg[2][2]);
  EXPECT_EQ(6.0f, dimg[3][4]);
  EXPECT_EQ(9.0f, dimg[4][4]);
  EXPECT_EQ(0.0f, dimg[0][0]);
}

TEST(SegmentedTest, ReduceScanAndOneLaunchBatch)  // extension, no reference test
{
  // uneven segments with an empty one and a tail outside the last segment
  const long offs[] = {0, 3, 3, 1000, 1001, 40000};
  const size_t n = 40007;
  skelcl::Vector<long> offsets(offs, offs + 6);
  skelcl::Vector<int> in(n);
  for (size_t i = 0; i < n; ++i) in[i] = (int)(i % 7) - 3;
  skelcl::Reduce<int(int)> r("int func(int x, int y){ return x+y; }");
  skelcl::Scan<int(int)> sc("int func(int x, int y){ return x+y; }", "0");
  skelcl::Vector<int> sums = r.segmented(in, offsets);
  skelcl::Vector<int> scan = sc.segmented(in, offsets, false);
  auto b = skelcl::batch(r, sc, in, offsets, true);
  for (int s = 0; s < 5; ++s) {
    int acc = 0;
    for (long i = offs[s]; i < offs[s + 1]; ++i) {
      EXPECT_EQ(acc, scan[i]);
      acc += in[i];
      EXPECT_EQ(acc, b.scan[i]);
    }
    EXPECT_EQ(acc, sums[s]);
    EXPECT_EQ(acc, b.sums[s]);
  }
  for (size_t i = 40000; i < n; ++i) EXPECT_EQ(0, b.scan[i]);

  // float: both forms against a double-precision host scan (compensated
  // FP32 inside a tile: far below 1e-5 of the running absolute sum)
  skelcl::Vector<float> fin(n);
  for (size_t i = 0; i < n; ++i) fin[i] = 0.001f * (float)((i * 37) % 1001) - 0.5f;
  skelcl::Reduce<float(float)> fr("float func(float x, float y){ return x+y; }");
  skelcl::Scan<float(float)> fs("float func(float x, float y){ return x+y; }", "0.0f");
  skelcl::Vector<float> fsums = fr.segmented(fin, offsets);
  skelcl::Vector<float> fscan = fs.segmented(fin, offsets, true);
  auto fb = skelcl::batch(fr, fs, fin, offsets, true);
  for (int s = 0; s < 5; ++s) {
    double acc = 0.0, mag = 0.0;
    bool ok = true;
    for (long i = offs[s]; i < offs[s + 1]; ++i) {
      acc += fin[i];
      mag += std::fabs((double)fin[i]);
      ok = ok && std::fabs(fscan[i] - acc) <= 1e-5 * mag + 1e-30 &&
           std::fabs(fb.scan[i] - acc) <= 1e-5 * mag + 1e-30;
    }
    EXPECT_TRUE(ok);
    EXPECT_TRUE(std::fabs(fsums[s] - acc) <= 1e-5 * mag + 1e-30);
    EXPECT_TRUE(std::fabs(fb.sums[s] - acc) <= 1e-5 * mag + 1e-30);
  }
}

int main(int argc, char** argv)
{
  skelcl::init(skelcl::nDevices(1));
  int rc = skltest::run_all(argc, argv);
  skelcl::terminate();
  return rc;
}

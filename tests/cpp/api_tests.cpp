// Container, device, distribution and common-definition tests of the C++
// drop-in, restating the reference's test/VectorTests.cpp:66-198,
// MatrixTests.cpp:60-99, DevicesTests.cpp:55-78, DeviceSelectionTests.cpp,
// IndexVectorTests.cpp:61-98, IndexMatrixTests.cpp:61-107 and the
// distribution rules of the skeletons' prepareInput; plus user structs and
// macros as common definitions (detail/Macros.h:55-61), as
// examples/mandelbrot/main.cpp uses them.
#include <SkelCL/SkelCL.h>

#include <sstream>
#include <stdexcept>
#include <vector>

#include "skelcl_test.h"

// ---- common definitions: a struct and a function shared by host and device
SKELCL_COMMON_DEFINITION(typedef struct {
  unsigned char r;
  unsigned char g;
  unsigned char b;
} Pixel;)
SKELCL_COMMON_DEFINITION(unsigned char shade(unsigned int v) { return (unsigned char)(v * 3u + 1u); })
#define SHADE_WIDTH 96
#define SHADE_STEP 5
SKELCL_ADD_DEFINE(SHADE_WIDTH)
SKELCL_ADD_DEFINE(SHADE_STEP)

namespace {
struct A {
  A() : _a(0) {}
  A(int a) : _a(a) {}
  int _a;
};
struct Position {
  Position() : _x(0), _y(0) {}
  Position(int x, int y) : _x(x), _y(y) {}
  int _x, _y;
};
}  // namespace

// ---- VectorTests.cpp ------------------------------------------------------
TEST(VectorTest, CreateEmptyVector)
{
  skelcl::Vector<int> vi;
  EXPECT_TRUE(vi.empty());
  EXPECT_EQ((size_t)0, vi.size());
}

TEST(VectorTest, Copy)
{
  skelcl::Vector<int> vi(5);
  for (unsigned i = 0; i < 5; ++i) vi[i] = 5;
  skelcl::Vector<int> copy(vi);
  for (unsigned i = 0; i < 5; ++i) vi[i] = 0;
  EXPECT_EQ(vi.size(), copy.size());
  for (unsigned i = 0; i < copy.size(); ++i) EXPECT_EQ(5, copy[i]);
}

TEST(VectorTest, CopyWithDataOnDevice)
{
  skelcl::Vector<int> vi(5);
  for (unsigned i = 0; i < 5; ++i) vi[i] = 5;
  vi.setDistribution(skelcl::detail::SingleDistribution<skelcl::Vector<int>>());
  vi.createDeviceBuffers();
  vi.copyDataToDevices();
  vi.dataOnDeviceModified();  // the device copy is now the current one
  skelcl::Vector<int> copy(vi);
  for (unsigned i = 0; i < 5; ++i) vi[i] = 0;
  EXPECT_EQ(vi.size(), copy.size());
  for (unsigned i = 0; i < copy.size(); ++i) EXPECT_EQ(5, copy[i]);
  EXPECT_TRUE(copy.distribution().isSingle());
}

TEST(VectorTest, CreateVector)
{
  skelcl::Vector<int> vi(10);
  EXPECT_EQ((size_t)10, vi.size());
  EXPECT_EQ(0, *vi.begin());
  skelcl::Vector<A> va(5, A(1));
  EXPECT_EQ((size_t)5, va.size());
  EXPECT_EQ(1, va.begin()->_a);
}

TEST(VectorTest, Resize)
{
  skelcl::Vector<int> vi;
  vi.resize(10);
  EXPECT_EQ((size_t)10, vi.size());
  vi.push_back(5);
  EXPECT_EQ((size_t)11, vi.size());
  vi.pop_back();
  EXPECT_EQ((size_t)10, vi.size());
  vi.insert(vi.begin(), 5);
  EXPECT_EQ((size_t)11, vi.size());
  vi.insert(vi.begin(), 2, 5);
  EXPECT_EQ((size_t)13, vi.size());
  std::vector<int> other(2);
  vi.insert(vi.begin(), other.begin(), other.end());
  EXPECT_EQ((size_t)15, vi.size());
  vi.erase(vi.begin());
  EXPECT_EQ((size_t)14, vi.size());
  vi.erase(vi.begin(), vi.begin() + 2);
  EXPECT_EQ((size_t)12, vi.size());
  vi.clear();
  EXPECT_EQ((size_t)0, vi.size());
}

TEST(VectorTest, CompleteInitialization)
{
  skelcl::Vector<Position> positions(1024 * 768);
  auto inserter = positions.begin();
  for (int x = 0; x < 1024; ++x)
    for (int y = 0; y < 768; ++y) *(inserter++) = Position(x, y);
  EXPECT_EQ((size_t)(1024 * 768), positions.size());
  EXPECT_EQ(23, positions[23 * 768 + 42]._x);
  EXPECT_EQ(42, positions[23 * 768 + 42]._y);
}

TEST(VectorTest, ModifiedHostDataReachesTheDevice)
{
  // a host modifier after a skeleton run drops the device copy, so the next
  // run uploads the new contents
  skelcl::Map<int(int)> neg("int func(int x){ return -x; }");
  skelcl::Vector<int> v(100, 3);
  skelcl::Vector<int> a = neg(v);
  v.push_back(7);
  skelcl::Vector<int> b = neg(v);
  EXPECT_EQ((size_t)101, b.size());
  EXPECT_EQ(-3, a[99]);
  EXPECT_EQ(-7, b[100]);
}

// ---- MatrixTests.cpp ------------------------------------------------------
TEST(MatrixTest, CreateEmptyMatrix)
{
  skelcl::Matrix<float> mi;
  EXPECT_TRUE(mi.empty());
  EXPECT_EQ(skelcl::MatrixSize(0, 0), mi.size());
}

TEST(MatrixTest, CreateMatrix)
{
  skelcl::Matrix<double> mi({10, 10});
  EXPECT_FALSE(mi.empty());
  EXPECT_EQ((size_t)100, mi.size().elemCount());
  EXPECT_EQ(0.0, *mi.begin());
  for (size_t i = 0; i < 10; ++i)
    for (size_t j = 0; j < 10; ++j) EXPECT_EQ(0.0, mi({i, j}));
}

TEST(MatrixTest, CreateMatrixFrom2DVector)
{
  std::vector<std::vector<int>> vec(10);
  for (unsigned i = 0; i < 10; ++i) {
    vec[i].resize(10);
    for (unsigned j = 0; j < 10; ++j) vec[i][j] = (int)(i * j);
  }
  skelcl::Matrix<int> mi = skelcl::Matrix<int>::from2DVector(vec);
  EXPECT_EQ((size_t)100, mi.size().elemCount());
  for (size_t i = 0; i < 10; ++i)
    for (size_t j = 0; j < 10; ++j) EXPECT_EQ((int)(i * j), mi[i][j]);
}

TEST(MatrixTest, IteratorConstructors)
{
  std::vector<int> v(23);
  for (int i = 0; i < 23; ++i) v[i] = i + 1;
  skelcl::Matrix<int> byCols(v.begin(), v.end(), 5);  // 5 rows, the last one padded
  EXPECT_EQ(skelcl::MatrixSize(5, 5), byCols.size());
  EXPECT_EQ(23, byCols[4][2]);
  EXPECT_EQ(0, byCols[4][3]);
  skelcl::Matrix<int> bySize(v.begin(), v.begin() + 6, skelcl::MatrixSize(2, 3));
  EXPECT_EQ(6, bySize[1][2]);
  skelcl::Matrix<int> fromVec(v, 4);
  EXPECT_EQ(skelcl::MatrixSize(6, 4), fromVec.size());
  EXPECT_EQ(21, fromVec[5][0]);
  skelcl::Matrix<int> withDist({3, 3}, 7, skelcl::detail::BlockDistribution<skelcl::Matrix<int>>());
  EXPECT_TRUE(withDist.distribution().isBlock());
  EXPECT_EQ(7, withDist[2][2]);
}

TEST(MatrixTest, RowOperations)
{
  skelcl::Matrix<int> m({2, 3}, 1);
  m.push_back_row({7, 8, 9, 10});  // cut to 3 columns
  EXPECT_EQ(skelcl::MatrixSize(3, 3), m.size());
  EXPECT_EQ(9, m[2][2]);
  m.insert_row(0, {4});  // filled up with 0
  EXPECT_EQ(4, m[0][0]);
  EXPECT_EQ(0, m[0][1]);
  EXPECT_EQ(1, m[1][0]);
  m.erase_row(1);
  EXPECT_EQ(skelcl::MatrixSize(3, 3), m.size());
  EXPECT_EQ(1, m[1][0]);
  m.pop_back_row();
  EXPECT_EQ(skelcl::MatrixSize(2, 3), m.size());
  EXPECT_EQ(1, *(m.row_end(1) - 1));
  m.resize({1, 4});  // keeps the row-major prefix (MatrixDef.h:370-395)
  EXPECT_EQ(4, m[0][0]);
  EXPECT_EQ(1, m[0][3]);
}

// ---- Index containers (IndexVectorTests.cpp, IndexMatrixTests.cpp) --------
TEST(IndexVectorTest, CreateIndexVector)
{
  skelcl::Vector<skelcl::Index> vi(1024);
  EXPECT_EQ((size_t)1024, vi.size());
  EXPECT_EQ((size_t)0, (size_t)vi.front());
  EXPECT_EQ((size_t)1023, (size_t)vi.back());
}

TEST(IndexVectorTest, SimpleVoidMap)
{
  skelcl::IndexVector index(1024);
  skelcl::Map<void(skelcl::Index)> m("void func(Index i, __global int* out) { out[i] = i; }");
  skelcl::Vector<int> v(1024);
  m(index, skelcl::out(v));
  EXPECT_EQ(index.size(), v.size());
  for (size_t i = 0; i < v.size(); ++i) EXPECT_EQ((int)i, v[i]);
}

TEST(IndexMatrixTest, SimpleVoidMap)
{
  skelcl::IndexMatrix index({128, 32});
  skelcl::Map<void(skelcl::IndexPoint)> map(
      "void func(IndexPoint ip, int_matrix_t out) { set(out, ip.y, ip.x, ip.y+ip.x); }");
  skelcl::Matrix<int> m({128, 32});
  map(index, skelcl::out(m));
  for (size_t y = 0; y < m.size().rowCount(); ++y)
    for (size_t x = 0; x < m.size().columnCount(); ++x) EXPECT_EQ((int)(y + x), m[y][x]);
}

// ---- devices (DevicesTests.cpp, DeviceSelectionTests.cpp, Device.h) -------
TEST(DeviceTest, GlobalDeviceList)
{
  EXPECT_EQ((size_t)1, skelcl::detail::globalDeviceList.size());
  auto& dev = skelcl::detail::globalDeviceList.front();
  EXPECT_TRUE(dev->isType(skelcl::device_type::GPU));
  EXPECT_TRUE(dev->isType(skelcl::device_type::ANY));
  EXPECT_TRUE(!dev->isType(skelcl::device_type::CPU));
  EXPECT_EQ(std::string("GPU"), dev->typeAsString());
  EXPECT_TRUE(dev->supportsDouble());
  EXPECT_TRUE(dev->maxWorkGroupSize() >= 256);
  EXPECT_TRUE(dev->maxComputeUnits() > 0);
  EXPECT_EQ((size_t)0, dev->id());
}

TEST(DeviceTest, DeviceListInstances)
{
  skelcl::detail::DeviceList devices;
  EXPECT_TRUE(devices.empty());
  devices.init(skelcl::nDevices(1));
  EXPECT_EQ((size_t)1, devices.size());
  skelcl::detail::DeviceList all;
  all.init(skelcl::allDevices());
  EXPECT_TRUE(all.size() > 0);
}

TEST(DeviceTest, DeviceTypeText)
{
  std::istringstream in("GPU");
  skelcl::device_type t = skelcl::device_type::ANY;
  in >> t;
  EXPECT_TRUE(t == skelcl::device_type::GPU);
  std::ostringstream o;
  o << skelcl::device_type::CPU;
  EXPECT_EQ(std::string("CPU"), o.str());
  bool threw = false;
  try {
    std::istringstream bad("TPU");
    bad >> t;
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
}

TEST(DeviceTest, CopiesUseElementAndByteOffsets)
{
  // enqueueWrite/Read count in elements, enqueueCopy in bytes (Device.h:142-245)
  auto dev = skelcl::detail::globalDeviceList.front();
  skelcl::detail::DeviceBuffer buf(dev, 16, sizeof(int));
  EXPECT_EQ((size_t)64, buf.sizeInBytes());
  std::vector<int> host(16);
  for (int i = 0; i < 16; ++i) host[i] = 100 + i;
  dev->enqueueWrite(buf, host.begin()).wait();
  std::vector<int> part(4, -1);
  dev->enqueueWrite(buf, part.data(), 2, 5, 1).wait();  // host[1..2] -> elements 5, 6
  std::vector<int> back(16, 0);
  dev->enqueueRead(buf, back.begin()).wait();
  EXPECT_EQ(104, back[4]);
  EXPECT_EQ(-1, back[5]);
  EXPECT_EQ(-1, back[6]);
  EXPECT_EQ(107, back[7]);
  std::vector<int> two(3, 0);
  dev->enqueueRead(buf, two.data(), 2, 10, 1).wait();  // elements 10, 11 -> two[1..2]
  EXPECT_EQ(0, two[0]);
  EXPECT_EQ(110, two[1]);
  EXPECT_EQ(111, two[2]);
  skelcl::detail::DeviceBuffer dst(dev, 16, sizeof(int));
  dev->enqueueWrite(dst, std::vector<int>(16, 0).data()).wait();
  dev->enqueueCopy(buf, dst, 8 * sizeof(int), 0).wait();  // bytes: the last 8 ints
  dev->enqueueRead(dst, back.begin()).wait();
  EXPECT_EQ(108, back[0]);
  EXPECT_EQ(115, back[7]);
  EXPECT_EQ(0, back[8]);
  skelcl::detail::DeviceBuffer deep(buf);  // a copy is a new device buffer
  EXPECT_TRUE(deep.devicePointer() != buf.devicePointer());
  dev->enqueueRead(deep, back.begin()).wait();
  EXPECT_EQ(-1, back[5]);
}

// ---- distributions (Distributions.h, the skeletons' prepareInput rules) ---
TEST(DistributionTest, Kinds)
{
  using skelcl::Vector;
  skelcl::detail::Distribution<Vector<int>> none;
  EXPECT_TRUE(!none.isValid());
  skelcl::detail::SingleDistribution<Vector<int>> single;
  EXPECT_TRUE(single.isValid());
  EXPECT_EQ((size_t)1, single.devices().size());
  EXPECT_TRUE(single.device(0) == skelcl::detail::globalDeviceList[0]);
  skelcl::detail::BlockDistribution<Vector<int>> block;
  EXPECT_EQ(skelcl::detail::globalDeviceList.size(), block.devices().size());
  Vector<int> v(10);
  skelcl::distribution::setCopy(v);
  EXPECT_TRUE(v.distribution().isCopy());
  EXPECT_EQ((size_t)10, v.distribution().sizeForDevice(v, skelcl::detail::globalDeviceList[0]));
  v.setDistribution(skelcl::distribution::Single(v));
  EXPECT_TRUE(v.distribution().isSingle());
  auto b = skelcl::distribution::Block(v);
  EXPECT_TRUE(b != nullptr && b->isBlock());
  // converting to another element type keeps the kind
  Vector<float> f(3);
  f.setDistribution(v.distribution());
  EXPECT_TRUE(f.distribution().isSingle());
}

TEST(DistributionTest, SkeletonDefaults)
{
  skelcl::Map<int(int)> neg("int func(int x){ return -x; }");
  skelcl::Vector<int> in(64, 2);
  skelcl::Vector<int> out = neg(in);
  EXPECT_TRUE(in.distribution().isBlock());   // MapHelperDef.h:70-71
  EXPECT_TRUE(out.distribution().isBlock());  // MapHelperDef.h:92
  skelcl::Reduce<int(int)> sum("int func(int x, int y){ return x+y; }");
  skelcl::Vector<int> r(64, 1);
  EXPECT_EQ(64, sum(r)[0]);
  EXPECT_TRUE(r.distribution().isSingle());  // ReduceDef.h:135-136
  skelcl::Zip<int(int, int)> add("int func(int x, int y){ return x+y; }");
  skelcl::Vector<int> l(8, 1), rr(8, 2);
  skelcl::distribution::setCopy(rr);
  skelcl::Vector<int> z = add(l, rr);
  EXPECT_TRUE(l.distribution().isCopy());  // ZipDef.h:236-237
  EXPECT_TRUE(z.distribution().isCopy());
  EXPECT_EQ(3, z[7]);
  skelcl::Zip<float(float, float)> mul("float func(float x, float y){ return x*y; }");
  skelcl::Reduce<float(float)> fsum("float func(float x, float y){ return x+y; }");
  skelcl::AllPairs<float(float, float)> mm(fsum, mul);
  skelcl::Matrix<float> A({4, 3}, 1.0f), B({3, 5}, 2.0f);
  skelcl::Matrix<float> C = mm(A, B);
  EXPECT_TRUE(A.distribution().isBlock());  // AllPairsDef.h:361-363
  EXPECT_TRUE(B.distribution().isCopy());
  EXPECT_EQ(6.0f, C[3][4]);
  skelcl::MapOverlap<int(int)> ov("int func(input_matrix_t m){ return getData(m, 0, 0); }", 2);
  skelcl::Matrix<int> M({6, 6}, 4);
  skelcl::Matrix<int> O = ov(M);
  EXPECT_TRUE(M.distribution().isOverlap());  // MapOverlapDef.h:254
  auto* ol = dynamic_cast<skelcl::detail::OLDistribution<skelcl::Matrix<int>>*>(&M.distribution());
  EXPECT_TRUE(ol != nullptr && ol->overlapRadius() == 2u);
  EXPECT_EQ(4, O[5][5]);
}

// ---- util (detail/Util.h) -------------------------------------------------
TEST(UtilTest, TypeNamesAndRounding)
{
  using namespace skelcl::detail::util;
  EXPECT_EQ(std::string("float4"), typeToString<skelcl::float4>());
  EXPECT_EQ(std::string("Pixel"), typeToString<Pixel>());
  EXPECT_EQ(std::string("float"), typeToString<float>());
  EXPECT_EQ(std::string("2fd4e1c67a2d28fced849ee1bb76e7391b93eb12"),
            hash("The quick brown fox jumps over the lazy dog"));
  EXPECT_EQ((size_t)3, devideAndRoundUp(9, 4));
  EXPECT_EQ((size_t)4, devideAndAlign(9, 4, 4));
  EXPECT_EQ((size_t)8, ceilToMultipleOf(5, 4));
  EXPECT_EQ((size_t)4, ceilToMultipleOf(0, 4));
  EXPECT_TRUE(isPowerOfTwo(64) && !isPowerOfTwo(65));
  EXPECT_EQ(16, floorPow2(17));
  EXPECT_EQ(32, ceilPow2(17));
  EXPECT_EQ(32, ceilPow2(16));
  EXPECT_EQ(std::string(""), envVarValue("SKELCL_SURELY_UNSET_VARIABLE"));
  skelcl::Local l = skelcl::local(256);
  EXPECT_EQ((size_t)256, l.getSizeInBytes());
}

// ---- a user struct as element type (examples/mandelbrot/main.cpp:49-88) ---
TEST(CommonDefinitionTest, MapPixelOfIndexPoint)
{
  const std::string text = skelcl::detail::CommonDefinitions::getSource();
  EXPECT_TRUE(text.find("Pixel;") != std::string::npos);
  EXPECT_TRUE(text.find("#define SHADE_WIDTH 96") != std::string::npos);
  EXPECT_TRUE(text.find("Pixel_matrix_t") != std::string::npos);
  skelcl::IndexMatrix positions({40, SHADE_WIDTH});
  skelcl::Map<Pixel(skelcl::IndexPoint)> m(
      "Pixel func(IndexPoint p, unsigned int bias) {"
      "  Pixel px; px.r = shade(p.x * SHADE_STEP); px.g = shade(p.y + bias);"
      "  px.b = (unsigned char)((p.x + p.y * SHADE_WIDTH) & 255); return px; }");
  skelcl::Matrix<Pixel> img = m(positions, 11u);
  EXPECT_EQ(skelcl::MatrixSize(40, SHADE_WIDTH), img.size());
  int bad = 0;
  for (unsigned y = 0; y < 40; ++y)
    for (unsigned x = 0; x < SHADE_WIDTH; ++x) {
      const Pixel& p = img[y][x];
      if (p.r != shade(x * SHADE_STEP) || p.g != shade(y + 11u) ||
          p.b != (unsigned char)((x + y * SHADE_WIDTH) & 255))
        ++bad;
    }
  EXPECT_EQ(0, bad);
}

TEST(CommonDefinitionTest, MatrixOfUserStructAsExtraArgument)
{
  // Matrix<Pixel> registered Pixel_matrix_t: a user function reads it with get()
  skelcl::Matrix<Pixel> lut({1, 4});
  for (unsigned i = 0; i < 4; ++i) {
    lut[0][i].r = (unsigned char)(10 * i);
    lut[0][i].g = (unsigned char)(i + 1);
    lut[0][i].b = 0;
  }
  skelcl::Map<int(int)> m(
      "int func(int x, Pixel_matrix_t t) { Pixel p = get(t, 0, x & 3); return p.r + p.g; }");
  skelcl::Vector<int> in(16);
  for (int i = 0; i < 16; ++i) in[i] = i;
  skelcl::Vector<int> out = m(in, lut);
  for (int i = 0; i < 16; ++i) EXPECT_EQ(10 * (i & 3) + (i & 3) + 1, out[i]);
}

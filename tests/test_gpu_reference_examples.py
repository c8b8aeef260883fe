"""The reference's own example programs (examples/*/main.cpp, compiled
unchanged against this library by skelcl_amd/build.py into
tests/cpp/_ref_examples/) run on the GPU and pass their own checks:
dot_product prints the same value as its sequential "gold" loop, saxpy finds
no errors (at its 1e-12 bound, so with contraction off, as its host
arithmetic), matrix_mult's result check succeeds for both its AllPairs forms,
montecarlo's check finds no errors.  gauss and mandelbrot read their user
function from a file in the working directory; these tests write their own
files of the same signature, and check the outputs against numpy."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_ref_examples")


def _run(name, args, cwd=None, env_extra=None, timeout=120):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (the reference tree was absent at build time)")
    env = dict(os.environ)
    env.update(env_extra or {})
    print(f"run {name} {' '.join(args)}", flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd,
                       env=env)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-3000:]
    return out


def test_dot_product_matches_its_gold_loop():
    out = _run("dot_product", ["-n", "1000003", "-c", "--device_count", "1"])
    got = re.search(r"skelcl:\s*(-?\d+)", out)
    gold = re.search(r"gold:\s*(-?\d+)", out)
    assert got and gold, out
    assert got.group(1) == gold.group(1)


def test_saxpy_finds_no_errors():
    out = _run("saxpy", ["-n", "4194305", "-c", "--device_count", "1"],
               env_extra={"SKELCL_FP_CONTRACT": "off"})
    assert "errors detected" not in out
    assert "Time:" in out


@pytest.mark.parametrize("extra", [[], ["--alt_kernel"], ["--use_double"],
                                   ["--use_double", "--alt_kernel"]])
def test_matrix_mult_result_check_succeeds(extra):
    out = _run("matrix_mult", ["-d", "300", "-m", "257", "-n", "190", "-c",
                               "--device_count", "1", *extra])
    assert "result check succeeded" in out
    assert "result check failed" not in out


def test_montecarlo_finds_no_errors():
    out = _run("montecarlo", ["-n", "20000", "-c", "--device_count", "1"])
    assert "no errors detected" in out


GAUSS_FUNC = """
int func(input_matrix_t img, __global float* w, int range)
{
  float sum = 0.0f;
  float norm = 0.0f;
  for (int dy = -range; dy <= range; ++dy) {
    for (int dx = -range; dx <= range; ++dx) {
      float k = w[dx + range] * w[dy + range];
      sum += getData(img, dx, dy) * k;
      norm += k;
    }
  }
  float v = sum / norm;
  return v > 255.0f ? 255 : (v < 0.0f ? 0 : (int)v);
}
"""


def test_gauss_blurs_a_pgm(tmp_path):
    """gauss reads ./gauss2D.cl and a P2 image, runs MapOverlap with NEAREST
    padding and two extra arguments (a weight Vector and the range), and
    writes <name>_<range>_devs_<n>.pgm."""
    rows, cols, R = 70, 93, 5
    img = oracle.gen_i32_below(0x3101, rows * cols, 256).reshape(rows, cols)
    with open(tmp_path / "img.pgm", "w") as f:
        f.write(f"P2\n# synthetic\n{cols} {rows}\n255\n")
        f.write("\n".join(str(v) for v in img.ravel()) + "\n")
    (tmp_path / "gauss2D.cl").write_text(GAUSS_FUNC)
    _run("gauss", ["--inFile", "img.pgm", "-r", str(R), "--device_count", "1"], cwd=tmp_path)
    lines = (tmp_path / "img_5_devs_1.pgm").read_text().split("\n")
    assert lines[0] == "P2" and lines[2] == f"{cols} {rows}"
    got = np.array([int(v) for v in lines[4:] if v.strip()]).reshape(rows, cols)
    # the example's weights: exp(-i*i / (2 a^2)), a = 5 / 2.354, stored as float
    a = 5 / 2.354
    w = np.array([np.float32(np.exp(-(i * i) / (2 * a * a))) for i in range(-R, R + 1)],
                 np.float64)
    pad = np.pad(img.astype(np.float64), R, mode="edge")
    k = np.outer(w, w)
    s = np.zeros((rows, cols))
    for dy in range(-R, R + 1):
        for dx in range(-R, R + 1):
            s += pad[R + dy:R + dy + rows, R + dx:R + dx + cols] * k[dy + R, dx + R]
    ref = np.clip(np.floor(s / k.sum()), 0, 255)
    assert np.abs(got - ref).max() <= 1  # float sums against FP64: +-1 at a truncation edge


MANDEL_FUNC = """
Pixel func(IndexPoint p, float startX, float startY, float dx, float dy)
{
  float cx = startX + p.x * dx;
  float cy = startY + p.y * dy;
  float x = 0.0f;
  float y = 0.0f;
  int n = 0;
  while (x * x + y * y <= 4.0f && n < ITERATIONS) {
    float t = x * x - y * y + cx;
    y = 2.0f * x * y + cy;
    x = t;
    ++n;
  }
  Pixel px;
  px.r = (unsigned char)(n * 2);
  px.g = (unsigned char)(n == ITERATIONS ? 0 : 255 - n);
  px.b = (unsigned char)(HEIGHT / 256 + n % 7);
  return px;
}
"""


def test_mandelbrot_writes_the_set(tmp_path):
    """mandelbrot builds Map<Pixel(IndexPoint)> over a 3072 x 4096 index
    matrix with Pixel from SKELCL_COMMON_DEFINITION and ITERATIONS / HEIGHT
    from SKELCL_ADD_DEFINE, and writes a P6 image.  With contraction off the
    escape counts are those of float32 arithmetic in the same order, so
    sampled rows match numpy exactly."""
    (tmp_path / "mandelbrot.cl").write_text(MANDEL_FUNC)
    _run("mandelbrot", ["--device_count", "1"], cwd=tmp_path,
         env_extra={"SKELCL_FP_CONTRACT": "off"})
    data = (tmp_path / "mandelbrot.ppm").read_bytes()
    W, H, IT = 4096, 3072, 100
    head = f"P6\n{W} {H}\n255\n".encode()
    assert data.startswith(head) and len(data) == len(head) + W * H * 3
    pix = np.frombuffer(data[len(head):], np.uint8).reshape(H, W, 3)
    f32 = np.float32
    zoom, cxc, cyc = 27615, -0.73, -0.16
    # the example's bounds: a float quotient, then double arithmetic rounded
    # to float (main.cpp computes CENTER_X - (float)WIDTH / (ZOOM * 2))
    sx = f32(cxc - float(f32(W) / f32(zoom * 2)))
    ex = f32(cxc + float(f32(W) / f32(zoom * 2)))
    sy = f32(cyc - float(f32(H) / f32(zoom * 2)))
    ey = f32(cyc + float(f32(H) / f32(zoom * 2)))
    ddx, ddy = f32((ex - sx) / f32(W)), f32((ey - sy) / f32(H))
    for row in (0, 777, 1536, 2900, H - 1):
        cx = (sx + np.arange(W, dtype=np.float32) * ddx).astype(np.float32)
        cy = f32(sy + f32(row) * ddy)
        x = np.zeros(W, np.float32)
        y = np.zeros(W, np.float32)
        n = np.zeros(W, np.int32)
        for _ in range(IT):
            live = (x * x + y * y <= f32(4.0)) & (n < IT)
            t = (x * x - y * y + cx).astype(np.float32)
            y = np.where(live, f32(2.0) * x * y + cy, y).astype(np.float32)
            x = np.where(live, t, x).astype(np.float32)
            n = n + live
        assert np.array_equal(pix[row, :, 0], (n * 2).astype(np.uint8)), row
        assert np.array_equal(pix[row, :, 2], (H // 256 + n % 7).astype(np.uint8)), row

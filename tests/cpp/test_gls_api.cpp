// C++ parity test of the gls:: API (include/glsCV.h) through libglsx.so,
// following the reference's own test cases (UnitTest_glsCV/unittest_glsFft.cpp:
// FFT_16x16 .. FFT_256x256, SHIFT, SCALE, IFFT, 1 x N, N x 1;
// UnitTest_glsPhaseCorrelate.cpp) with the FP64 oracle as the expected value
// instead of cv::dft.  Exit status 0 = all cases pass.
#include <cmath>
#include <cstring>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "glsCV.h"
#include "../../oracle/oracle.h"

static int g_fail = 0;

static double rel_rms(const std::vector<float>& y, const std::vector<double>& ref) {
  double num = 0, den = 0;
  for (size_t i = 0; i < ref.size(); ++i) {
    const double d = (double)y[i] - ref[i];
    num += d * d;
    den += ref[i] * ref[i];
  }
  return std::sqrt(num / den);
}

static void test_fft(const char* name, int rows, int cols, int flags) {
  std::mt19937_64 rng(20261203 + rows * 7 + cols + flags);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);  // cv::randu range of FillRandU
  std::vector<float> x((size_t)rows * cols * 2);
  for (auto& v : x) v = u(rng);
  gls::GlsMat src(gls::Size(cols, rows), CV_32FC2), dst;
  src.upload(x.data());
  gls::fft(src, dst, flags);
  std::vector<float> y = dst.download<float>();
  std::vector<double> xd(x.begin(), x.end()), ref(x.size());
  orc_fft2d_f64(xd.data(), ref.data(), rows, cols, flags);
  const double e = rel_rms(y, ref);
  const bool ok = e <= 1e-5 && dst.type() == CV_32FC2 && dst.rows == rows && dst.cols == cols;
  std::printf("%s %s %dx%d flags=%d relrms=%.3e\n", ok ? "PASS" : "FAIL", name, rows, cols, flags, e);
  g_fail += !ok;
}

static void test_glsmat_semantics() {
  gls::GlsMat a(4, 8, CV_32FC1);
  std::vector<float> v(32);
  for (int i = 0; i < 32; ++i) v[i] = (float)i;
  a.upload(v.data());
  gls::GlsMat b = a;                 // shallow: same storage
  gls::GlsMat c = a.clone();         // deep copy
  bool ok = b.id() == a.id() && c.id() != a.id() && c.download<float>() == v;
  std::vector<float> z(32, 0.0f);
  b.upload(z.data());                // writes through the shared handle
  ok = ok && a.download<float>() == z && c.download<float>() == v;
  gls::GlsMat d(2, 2, CV_32FC2);
  const void* before = d.id();
  gls::GlsMat s(2, 2, CV_32FC2);
  s.upload(std::vector<float>(8, 1.0f).data());
  gls::fft(s, d);                    // matching dst is reused (getDstMat)
  ok = ok && d.id() == before;
  bool threw = false;
  try {
    gls::GlsMat bad(3, 5, CV_32FC2);
    gls::fft(bad, d);
  } catch (const gls::Exception& e) {
    threw = e.code != 0;
  }
  ok = ok && threw;
  std::printf("%s GlsMat semantics (shallow copy, clone, dst reuse, throw on bad shape)\n", ok ? "PASS" : "FAIL");
  g_fail += !ok;
}

static void test_phase_correlate(int n, int dx, int dy) {
  std::mt19937_64 rng(99 + n);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::vector<float> a((size_t)n * n), b((size_t)n * n);
  for (auto& v : a) v = u(rng);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) b[(size_t)((y + dy + n) % n) * n + (x + dx + n) % n] = a[(size_t)y * n + x];
  gls::GlsMat m1(n, n, CV_32FC1), m2(n, n, CV_32FC1), win;
  m1.upload(a.data());
  m2.upload(b.data());
  double response = 0;
  gls::Point peak;
  gls::Point2d s = gls::phaseCorrelate(m1, m2, win, &response, &peak);
  orc_poc_result ref;
  orc_phase_correlate(a.data(), b.data(), nullptr, n, n, 1, &ref, nullptr);
  const bool ok = peak.x == ref.peak_x && peak.y == ref.peak_y && peak.x == n / 2 - dx && peak.y == n / 2 - dy &&
                  std::fabs(s.x - dx) < 1e-3 && std::fabs(s.y - dy) < 1e-3 && std::fabs(response - 1.0) < 1e-3;
  std::printf("%s phaseCorrelate %dx%d shift=(%d,%d) got=(%.4f,%.4f) peak=(%d,%d) response=%.5f\n",
              ok ? "PASS" : "FAIL", n, n, dx, dy, s.x, s.y, peak.x, peak.y, response);
  g_fail += !ok;
}

// glsCam's FFT view (glsCam/glsCam.cpp:199-202): fft(SHIFT), logMagSpectrums(1.0),
// normalize(0, 1, NORM_MINMAX), in place on one handle, against the oracle chain
// on the FP64 spectrum.
static void test_display_chain(int n) {
  std::mt19937_64 rng(7 + n);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::vector<float> x((size_t)n * n);
  for (auto& v : x) v = u(rng);
  gls::GlsMat frame(n, n, CV_32FC1);
  frame.upload(x.data());
  gls::fft(frame, frame, gls::GLS_FFT_SHIFT);
  gls::logMagSpectrums(frame, frame, 1.0f);
  gls::normalize(frame, frame, 0, 1, gls::NORM_MINMAX);
  const bool shape_ok = frame.type() == CV_32FC1 && frame.rows == n && frame.cols == n;
  std::vector<float> got = frame.download<float>();
  std::vector<double> xd((size_t)n * n * 2, 0.0), spec((size_t)n * n * 2);
  for (size_t i = 0; i < x.size(); ++i) xd[2 * i] = x[i];
  orc_fft2d_f64(xd.data(), spec.data(), n, n, ORC_FFT_SHIFT);
  std::vector<float> c(spec.begin(), spec.end()), lm(x.size()), ref(x.size());
  orc_log_mag_spectrums_f32(c.data(), lm.data(), (long)x.size(), 1.0f);
  orc_normalize_f32(lm.data(), ref.data(), (long)x.size(), 0.0, 1.0, ORC_NORM_MINMAX);
  double worst = 0;
  for (size_t i = 0; i < x.size(); ++i) worst = std::fmax(worst, std::fabs((double)got[i] - ref[i]));
  // fused form: one call for fft + logMagSpectrums
  gls::GlsMat src(n, n, CV_32FC1), fused;
  src.upload(x.data());
  gls::fftMagSpectrums(src, fused, gls::GLS_FFT_SHIFT, true, 1.0f);
  gls::normalize(fused, fused, 0, 1, gls::NORM_MINMAX);
  gls::GlsMat view;
  gls::spectrumView(src, view, gls::GLS_FFT_SHIFT, 1.0f, 0.0, 1.0);
  const bool same = fused.download<float>() == got && view.download<float>() == got;
  // normalised display values: the FP32 FFT's error near empty bins, scaled by
  // 1/(max - min) (tests/test_spectrum_gpu.py checks the per-bin bound)
  const bool ok = shape_ok && worst < 1e-4 && same;
  std::printf("%s display chain %dx%d max|err|=%.2e fused==view==unfused:%d\n", ok ? "PASS" : "FAIL", n, n, worst,
              same);
  g_fail += !ok;
}

// ---- the reference's own call shapes, with only cv:: type names read as gls:: --
using gls::GlsMat;
using gls::Mat;
using gls::Point2d;
using gls::Size;
using gls::Point;

// cv::randu over [0, 1) of every float of the image (FillRandU, UnitTest_Common)
static void FillRandU(Mat& m, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  for (int r = 0; r < m.rows; ++r) {
    float* p = m.ptr<float>(r);
    for (size_t i = 0; i < (size_t)m.cols * m.channels(); ++i) p[i] = u(rng);
  }
}

static double rel_rms_mat(const Mat& y, const std::vector<double>& ref) {
  double num = 0, den = 0;
  size_t k = 0;
  for (int r = 0; r < y.rows; ++r)
    for (int i = 0; i < y.cols * y.channels(); ++i, ++k) {
      const double d = (double)y.ptr<float>(r)[i] - ref[k];
      num += d * d;
      den += ref[k] * ref[k];
    }
  return std::sqrt(num / den);
}

// unittest_glsFft.cpp:100-200 (test_glsFft): Mat in, GlsMat _src(imgSrc),
// gls::fft(_src, _dst, _flags), _dst.download(imgFft); the FP64 oracle stands
// in for cv::dft + fftShift.
static int test_glsFft(const int N, const int flags, const int dir = -1) {
  Size size;
  if (dir < 0) size = Size(N, N);
  else if (dir == 0) size = Size(N, 1);
  else size = Size(1, N);
  Mat imgSrc = Mat(size, CV_32FC2);
  Mat imgFft = Mat::zeros(imgSrc.size(), imgSrc.type());
  FillRandU(imgSrc, 1000u + N * 3u + flags + (dir + 1) * 7u);
  std::vector<double> xd((size_t)imgSrc.rows * imgSrc.cols * 2), ref(xd.size());
  for (size_t i = 0; i < xd.size(); ++i) xd[i] = imgSrc.ptr<float>(0)[i];
  orc_fft2d_f64(xd.data(), ref.data(), imgSrc.rows, imgSrc.cols, flags);
  {
    int _flags = 0;
    if (flags & gls::DFT_SCALE) _flags |= gls::GLS_FFT_SCALE;
    if (flags & gls::DFT_INVERSE) _flags |= gls::GLS_FFT_INVERSE;
    if (flags & gls::GLS_FFT_SHIFT) _flags |= gls::GLS_FFT_SHIFT;
    GlsMat _src(imgSrc);
    GlsMat _dst;
    gls::fft(_src, _dst, _flags);
    _dst.download(imgFft);
  }
  const double e = rel_rms_mat(imgFft, ref);
  const int errNum = e <= 1e-5 && imgFft.type() == CV_32FC2 && imgFft.size() == size ? 0 : -1;
  std::printf("%s test_glsFft N=%d flags=%d dir=%d relrms=%.3e\n", errNum ? "FAIL" : "PASS", N, flags, dir, e);
  g_fail += errNum != 0;
  return errNum;
}

// UnitTest_glsPhaseCorrelate.cpp:40-140 (test_glsPhaseCorrelatet): images
// converted with (GlsMat)imgSrc, a Hann window, gls::phaseCorrelate(_imgSrc,
// _imgSft, _window, &res_gl).  The shift is an integer circular shift here
// (cv::warpAffine is not available), so the reference's bar -- within 0.25 px
// of OpenCV's error -- becomes: integer peak equal to the oracle's, sub-pixel
// shift and response within 1e-3 of the oracle's chain on the same inputs.
static int test_glsPhaseCorrelatet(Size size, int dx, int dy) {
  int errNum = 0;
  Mat imgSrc = Mat::zeros(size, CV_32FC1);
  FillRandU(imgSrc, 77u + size.width);
  Mat imgSft = Mat::zeros(size, CV_32FC1);
  for (int y = 0; y < size.height; ++y)
    for (int x = 0; x < size.width; ++x)
      imgSft.at<float>((y + dy + size.height) % size.height, (x + dx + size.width) % size.width) =
          imgSrc.at<float>(y, x);
  Mat window(size, CV_32FC1);  // cv::createHanningWindow
  for (int y = 0; y < size.height; ++y)
    for (int x = 0; x < size.width; ++x) {
      const double wy = 0.5 - 0.5 * std::cos(2.0 * M_PI * y / (size.height - 1));
      const double wx = 0.5 - 0.5 * std::cos(2.0 * M_PI * x / (size.width - 1));
      window.at<float>(y, x) = (float)std::sqrt(wx * wy);
    }
  GlsMat _imgSrc = (GlsMat)imgSrc;
  GlsMat _imgSft = (GlsMat)imgSft;
  GlsMat _window = (GlsMat)window;
  double res_gl;
  Point2d mv_gl;
  mv_gl = gls::phaseCorrelate(_imgSrc, _imgSft, _window, &res_gl);
  orc_poc_result ref;
  orc_phase_correlate(imgSrc.ptr<float>(), imgSft.ptr<float>(), window.ptr<float>(), size.height, size.width, 1,
                      &ref, nullptr);
  if (std::fabs(mv_gl.x - ref.shift_x) > 1e-3) errNum++;
  if (std::fabs(mv_gl.y - ref.shift_y) > 1e-3) errNum++;
  if (std::fabs(res_gl - ref.response) > 1e-3 * std::fmax(1.0, std::fabs(ref.response))) errNum++;
  // the integer shift is recovered (peak = centre - shift)
  if (std::lround(mv_gl.x) != dx || std::lround(mv_gl.y) != dy) errNum++;
  std::printf("%s test_glsPhaseCorrelatet %dx%d shift=(%d,%d) mv_gl=(%.4f,%.4f) oracle=(%.4f,%.4f) res=%.4f/%.4f\n",
              errNum ? "FAIL" : "PASS", size.width, size.height, dx, dy, mv_gl.x, mv_gl.y, ref.shift_x, ref.shift_y,
              res_gl, ref.response);
  g_fail += errNum != 0;
  return errNum;
}

// GlsMat <-> Mat transfers (glsMat.cpp:90-96, 374-419, 493-552): upload from
// and download into images with a row pitch, dst reuse vs replacement, the
// conversion operator (used as (Mat)C at glsPhaseCorrelate.cpp:219).
static void test_mat_transfers() {
  bool ok = true;
  Mat a(5, 7, CV_32FC2);
  FillRandU(a, 5);
  GlsMat g(a);
  Mat b = (Mat)g;  // operator Mat
  ok = ok && b.size() == a.size() && b.type() == a.type() && b.isContinuous();
  for (int r = 0; r < 5; ++r) ok = ok && std::memcmp(b.ptr<float>(r), a.ptr<float>(r), 7 * 8) == 0;
  // a borrowed image with a padded row pitch
  std::vector<float> padded(5 * 20, -1.0f);
  Mat view(5, 7, CV_32FC2, padded.data(), 20 * sizeof(float));
  ok = ok && !view.isContinuous();
  g.download(view);  // same size and type: written in place, padding kept
  for (int r = 0; r < 5; ++r) {
    ok = ok && std::memcmp(view.ptr<float>(r), a.ptr<float>(r), 7 * 8) == 0;
    ok = ok && padded[r * 20 + 14] == -1.0f && padded[r * 20 + 19] == -1.0f;
  }
  ok = ok && view.data == (unsigned char*)padded.data();
  // upload from the padded view into a second GlsMat
  GlsMat g2(5, 7, CV_32FC2);
  g2.upload(view);
  Mat c;
  g2.download(c);  // empty dst: a new image
  for (int r = 0; r < 5; ++r) ok = ok && std::memcmp(c.ptr<float>(r), a.ptr<float>(r), 7 * 8) == 0;
  // dst of another size is replaced, the borrowed memory is left alone
  std::vector<float> small(4, 3.0f);
  Mat other(1, 2, CV_32FC2, small.data());
  g.download(other);
  ok = ok && other.rows == 5 && other.cols == 7 && other.data != (unsigned char*)small.data() && small[0] == 3.0f;
  bool threw = false;
  try {
    Mat wrong(5, 7, CV_32FC1);
    g.upload(wrong);  // type differs: CV_Assert in the reference
  } catch (const gls::Exception&) {
    threw = true;
  }
  ok = ok && threw;
  std::printf("%s GlsMat <-> Mat transfers (pitch, reuse, conversion, type check)\n", ok ? "PASS" : "FAIL");
  g_fail += !ok;
}

// glsPhaseCorrelate.cpp:174-210 as the reference writes the chain, with the
// standalone operations: window multiply into the inputs, two FFTs,
// mulSpectrumsPhaseOnly, the in-place inverse FFT with SHIFT, split, and
// minMaxLoc of the real plane.  Its peak must equal the fused
// gls::phaseCorrelate's and the oracle chain's; split / merge / minMaxLoc
// are also checked against the values they must return.
static void test_phase_correlate_by_operations(int n, int dx, int dy) {
  int errNum = 0;
  Mat img1 = Mat::zeros(Size(n, n), CV_32FC1);
  FillRandU(img1, 91u + n);
  Mat img2 = Mat::zeros(Size(n, n), CV_32FC1);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) img2.at<float>((y + dy + n) % n, (x + dx + n) % n) = img1.at<float>(y, x);
  Mat win(n, n, CV_32FC1);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x)
      win.at<float>(y, x) = (float)((0.5 - 0.5 * std::cos(2.0 * M_PI * y / n)) * (0.5 - 0.5 * std::cos(2.0 * M_PI * x / n)));
  GlsMat padded1 = (GlsMat)img1, padded2 = (GlsMat)img2, paddedWin = (GlsMat)win;
  gls::multiply(paddedWin, padded1, padded1);
  gls::multiply(paddedWin, padded2, padded2);
  GlsMat FFT1, FFT2, C;
  gls::fft(padded1, FFT1);
  gls::fft(padded2, FFT2);
  gls::mulSpectrumsPhaseOnly(FFT1, FFT2, C);
  gls::fft(C, C, gls::GLS_FFT_INVERSE | gls::GLS_FFT_SHIFT);
  std::vector<GlsMat> cpx(2);
  gls::split(C, cpx);
  C = cpx[0];
  Point peakLoc;
  double maxVal = 0;
  gls::minMaxLoc(C, NULL, &maxVal, NULL, &peakLoc);
  Point fused;
  GlsMat _img1 = (GlsMat)img1, _img2 = (GlsMat)img2, _win = (GlsMat)win;
  gls::phaseCorrelate(_img1, _img2, _win, nullptr, &fused);
  orc_poc_result ref;
  orc_phase_correlate(img1.ptr<float>(), img2.ptr<float>(), win.ptr<float>(), n, n, 1, &ref, nullptr);
  if (peakLoc.x != fused.x || peakLoc.y != fused.y) errNum++;
  if (peakLoc.x != ref.peak_x || peakLoc.y != ref.peak_y) errNum++;
  if (peakLoc.x != n / 2 - dx || peakLoc.y != n / 2 - dy) errNum++;
  // the surface's maximum is the value at the peak, read back through (Mat)C
  Mat surf = (Mat)C;
  if ((float)maxVal != surf.at<float>(peakLoc.y, peakLoc.x)) errNum++;
  // merge(split(x)) == x, bit for bit
  GlsMat spec = FFT1, back;
  std::vector<GlsMat> planes;
  gls::split(spec, planes);
  gls::merge(planes, back);
  Mat a = (Mat)spec, b = (Mat)back;
  if (planes.size() != 2 || std::memcmp(a.data, b.data, (size_t)n * n * 8) != 0) errNum++;
  // a mask is refused, as the reference asserts an empty one
  bool threw = false;
  try {
    gls::minMaxLoc(C, NULL, NULL, NULL, NULL, C);
  } catch (const gls::Exception&) {
    threw = true;
  }
  if (!threw) errNum++;
  std::printf("%s test_phase_correlate_by_operations n=%d shift=(%d,%d) ops=(%d,%d) fused=(%d,%d) oracle=(%d,%d)\n",
              errNum ? "FAIL" : "PASS", n, dx, dy, peakLoc.x, peakLoc.y, fused.x, fused.y, ref.peak_x, ref.peak_y);
  g_fail += errNum != 0;
}

int main() {
  gls::glsCvInit();
  test_fft("FFT_256x256", 256, 256, 0);
  test_fft("FFT_32x32", 32, 32, 0);
  test_fft("FFT_16x16", 16, 16, 0);
  test_fft("FFT_16x16_SHIFT", 16, 16, gls::GLS_FFT_SHIFT);
  test_fft("FFT_128x128_SCALE", 128, 128, gls::GLS_FFT_SCALE);
  test_fft("IFFT_64x64", 64, 64, gls::GLS_FFT_INVERSE);
  test_fft("IFFT_32x32_SCALE", 32, 32, gls::GLS_FFT_INVERSE | gls::GLS_FFT_SCALE);
  for (int n = 64; n <= 2048; n *= 2) test_fft("FFT_SCALE_benchmark", n, n, gls::GLS_FFT_SCALE);
  test_fft("FFT_32H", 1, 32, 0);
  test_fft("FFT_64V", 64, 1, 0);
  test_fft("FFT_4096x4096", 4096, 4096, 0);
  test_fft("FFT_128x2048_nonsquare", 128, 2048, gls::GLS_FFT_SHIFT);
  test_glsmat_semantics();
  test_phase_correlate(32, 3, -2);
  test_phase_correlate(256, -40, 17);
  test_phase_correlate(1024, 200, -255);
  test_display_chain(256);
  // the reference's test cases in its own call shape (unittest_glsFft.cpp:202-386,
  // UnitTest_glsPhaseCorrelate.cpp:142-256)
  test_mat_transfers();
  test_glsFft(16, 0);
  test_glsFft(256, 0);
  test_glsFft(64, gls::GLS_FFT_SHIFT);
  test_glsFft(128, gls::DFT_SCALE);
  test_glsFft(32, gls::DFT_INVERSE | gls::DFT_SCALE);
  test_glsFft(1024, gls::DFT_SCALE);
  test_glsFft(32, 0, 0);
  test_glsFft(64, 0, 1);
  test_glsPhaseCorrelatet(Size(32, 32), 3, -5);
  test_glsPhaseCorrelatet(Size(256, 256), -17, 29);
  test_glsPhaseCorrelatet(Size(512, 512), 100, -61);
  test_phase_correlate_by_operations(64, 5, -9);
  test_phase_correlate_by_operations(512, -120, 37);
  gls::glsCvTerminate();
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}

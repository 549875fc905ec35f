// gls_api.cpp -- the reference's C++ API for this path (namespace gls, same
// names and signatures as glsCV/glsMat.h, glsFft.h, glsPhaseCorrelate.h,
// glsCV.h) implemented over the C ABI.  A non-zero status becomes a thrown
// gls::Exception, as GLS_Assert throws cv::Exception in the reference.
#include "../../include/glsCV.h"

#include <string>

#define GLS_API __attribute__((visibility("default")))

namespace gls {

static void check(int status) {
  if (status != GLSX_OK) throw Exception(status, glsx_last_error());
}

GLS_API void glsCvInit(int device) { check(glsx_init(device)); }
GLS_API void glsCvTerminate(void) { check(glsx_shutdown()); }

// ---- GlsMat -------------------------------------------------------------------
GLS_API GlsMat::GlsMat(const Size& size, int type) { create(size.height, size.width, type); }
GLS_API GlsMat::GlsMat(int r, int c, int type) { create(r, c, type); }
GLS_API GlsMat::GlsMat(const Mat& host) {
  if (host.empty()) throw Exception(GLSX_ERR_ARG, "GlsMat(const Mat&): empty image");
  create(host.rows, host.cols, host.type());
  upload(host);
}
GLS_API GlsMat::GlsMat(const GlsMat& o) : rows(o.rows), cols(o.cols), buf_(o.buf_), type_(o.type_) {
  if (buf_) check(glsx_buf_retain(buf_));
}
GLS_API GlsMat& GlsMat::operator=(const GlsMat& o) {
  if (this == &o) return *this;
  if (o.buf_) check(glsx_buf_retain(o.buf_));
  reset();
  buf_ = o.buf_;
  rows = o.rows;
  cols = o.cols;
  type_ = o.type_;
  return *this;
}
GLS_API GlsMat::~GlsMat() {
  if (buf_) glsx_buf_release(buf_);
}
void GlsMat::reset() {
  if (buf_) check(glsx_buf_release(buf_));
  buf_ = nullptr;
  rows = cols = 0;
  type_ = -1;
}
void GlsMat::create(int r, int c, int type) {
  glsx_buf b = nullptr;
  check(glsx_buf_create(r, c, type, &b));
  reset();
  buf_ = b;
  rows = r;
  cols = c;
  type_ = type;
}
GLS_API size_t GlsMat::elemSize() const {
  static const int depth_bytes[8] = {1, 1, 2, 2, 4, 4, 8, 2};
  return (size_t)depth_bytes[depth()] * channels();
}
GLS_API Size GlsMat::size() const { return Size(cols, rows); }
GLS_API void GlsMat::upload(const void* host, size_t pitch) { check(glsx_buf_upload(buf_, host, pitch, nullptr)); }
// glsMat.cpp:374-419: the host image must have this size and type
GLS_API void GlsMat::upload(const Mat& src) {
  if (empty()) throw Exception(GLSX_ERR_ARG, "upload: empty GlsMat");
  if (src.type() != type_ || src.rows != rows || src.cols != cols)
    throw Exception(GLSX_ERR_ARG, "upload: host image size or type differs");
  check(glsx_buf_upload(buf_, src.data, src.step, nullptr));
}
// glsMat.cpp:493-552: reuse dst when size and type match, else a new image
GLS_API void GlsMat::download(Mat& dst) const {
  if (empty()) throw Exception(GLSX_ERR_ARG, "download: empty GlsMat");
  if (dst.empty() || dst.type() != type_ || dst.rows != rows || dst.cols != cols) dst = Mat(rows, cols, type_);
  check(glsx_buf_download(buf_, dst.data, dst.step, nullptr));
}
GLS_API GlsMat::operator Mat() const {
  Mat m;
  download(m);
  return m;
}
GLS_API void GlsMat::download(void* host, size_t pitch) const {
  check(glsx_buf_download(buf_, host, pitch, nullptr));
}
GLS_API void GlsMat::copyTo(GlsMat& m) const {
  if (m.buf_ == buf_) return;
  if (m.empty() || m.rows != rows || m.cols != cols || m.type_ != type_) m.create(rows, cols, type_);
  check(glsx_buf_copy(buf_, m.buf_, nullptr));
}
GLS_API GlsMat GlsMat::clone() const {
  GlsMat m(rows, cols, type_);
  check(glsx_buf_copy(buf_, m.buf_, nullptr));
  return m;
}
GLS_API void* GlsMat::devPtr() const {
  void* p = nullptr;
  check(glsx_buf_info(buf_, nullptr, nullptr, nullptr, &p, nullptr));
  return p;
}

// ---- fft (glsCV/glsFft.cpp:675-708) -------------------------------------------
GLS_API void fft(const GlsMat& src, GlsMat& dst, int flag) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "fft: empty source");
  // getDstMat (glsMat.h:173-214): reuse dst when size and type fit; the HIP
  // passes run in place, so dst may alias src.
  if (dst.empty() || dst.rows != src.rows || dst.cols != src.cols || dst.type() != CV_32FC2) {
    GlsMat fresh(src.rows, src.cols, CV_32FC2);
    check(glsx_fft2d(src.handle(), fresh.handle(), flag, nullptr));
    dst = fresh;
    return;
  }
  check(glsx_fft2d(src.handle(), dst.handle(), flag, nullptr));
}

GLS_API void ifft(const GlsMat& src, GlsMat& dst, int flag) { fft(src, dst, flag | GLS_FFT_INVERSE); }

// ---- spectrum display chain (glsBasicOperation.cpp:337-351, glsNormalize.cpp:52-75)
namespace {
// getDstMat for a CV_32FC1 plane of src's size; a dst that aliases src (or
// has another size/type) gets a fresh buffer, as the reference reassigns it.
GlsMat plane_for(const GlsMat& src, GlsMat& dst, bool may_alias) {
  const bool fits = !dst.empty() && dst.rows == src.rows && dst.cols == src.cols && dst.type() == CV_32FC1;
  if (fits && (may_alias || dst.id() != src.id())) return dst;
  return GlsMat(src.rows, src.cols, CV_32FC1);
}
}  // namespace

GLS_API void magSpectrums(const GlsMat& src, GlsMat& dst) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "magSpectrums: empty source");
  GlsMat out = plane_for(src, dst, false);
  check(glsx_mag_spectrums(src.handle(), out.handle(), nullptr));
  dst = out;
}

GLS_API void logMagSpectrums(const GlsMat& src, GlsMat& dst, float offset) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "logMagSpectrums: empty source");
  GlsMat out = plane_for(src, dst, false);
  check(glsx_log_mag_spectrums(src.handle(), out.handle(), offset, nullptr));
  dst = out;
}

GLS_API void fftMagSpectrums(const GlsMat& src, GlsMat& dst, int flag, bool logMag, float offset) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "fftMagSpectrums: empty source");
  GlsMat out = plane_for(src, dst, false);
  check(glsx_fft2d_mag(src.handle(), out.handle(), flag, logMag ? 1 : 0, offset, nullptr));
  dst = out;
}

GLS_API void spectrumView(const GlsMat& src, GlsMat& dst, int flag, float offset, double alpha, double beta) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "spectrumView: empty source");
  GlsMat out = plane_for(src, dst, false);
  check(glsx_spectrum_view(src.handle(), out.handle(), flag, offset, alpha, beta, nullptr));
  dst = out;
}

GLS_API void normalize(const GlsMat& src, GlsMat& dst, double alpha, double beta, int normType) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "normalize: empty source");
  if (src.type() != CV_32FC1) throw Exception(GLSX_ERR_ARG, "normalize: source must be CV_32FC1");
  GlsMat out = plane_for(src, dst, true);
  check(glsx_normalize(src.handle(), out.handle(), alpha, beta, normType, nullptr));
  dst = out;
}

// ---- the chain's operations one at a time (glsBasicOperation.cpp:265-335,
// glsSplit.cpp:192-210, glsMerge.cpp:193-211, glsMinMaxLoc.cpp:149-199) --------------
namespace {
// getDstMat(src0, src1, dst) (glsMat.h:185-196): a new image when dst is one
// of the sources or differs from src0 in size or type
GlsMat dst_for(const GlsMat& src0, const GlsMat& src1, GlsMat& dst) {
  const bool fits = !dst.empty() && dst.rows == src0.rows && dst.cols == src0.cols && dst.type() == src0.type();
  if (fits && dst.id() != src0.id() && dst.id() != src1.id()) return dst;
  return GlsMat(src0.rows, src0.cols, src0.type());
}
}  // namespace

GLS_API void multiply(const GlsMat& src0, const GlsMat& src1, GlsMat& dst) {
  if (src0.empty() || src1.empty()) throw Exception(GLSX_ERR_ARG, "multiply: empty source");
  GlsMat out = dst_for(src0, src1, dst);
  check(glsx_multiply(src0.handle(), src1.handle(), out.handle(), nullptr));
  dst = out;
}

GLS_API void mulSpectrums(const GlsMat& src0, const GlsMat& src1, GlsMat& dst, bool conj) {
  if (src0.empty() || src1.empty()) throw Exception(GLSX_ERR_ARG, "mulSpectrums: empty source");
  GlsMat out = dst_for(src0, src1, dst);
  check(glsx_mul_spectrums(src0.handle(), src1.handle(), out.handle(), conj ? 1 : 0, nullptr));
  dst = out;
}

GLS_API void mulSpectrumsPhaseOnly(const GlsMat& src0, const GlsMat& src1, GlsMat& dst) {
  if (src0.empty() || src1.empty()) throw Exception(GLSX_ERR_ARG, "mulSpectrumsPhaseOnly: empty source");
  GlsMat out = dst_for(src0, src1, dst);
  check(glsx_mul_spectrums_phase_only(src0.handle(), src1.handle(), out.handle(), nullptr));
  dst = out;
}

GLS_API void split(const GlsMat& src, std::vector<GlsMat>& dst) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "split: empty source");
  if (src.type() != CV_32FC1 && src.type() != CV_32FC2)
    throw Exception(GLSX_ERR_ARG, "split: source must be CV_32FC1 or CV_32FC2");
  const int cn = src.channels();
  dst.resize(cn);
  std::vector<glsx_buf> planes(cn);
  for (int c = 0; c < cn; ++c) {
    dst[c] = GlsMat(src.rows, src.cols, CV_32FC1);
    planes[c] = dst[c].handle();
  }
  check(glsx_split(src.handle(), planes.data(), cn, nullptr));
}

GLS_API void merge(const std::vector<GlsMat>& src, GlsMat& dst) {
  const int cn = (int)src.size();
  if (cn < 1 || cn > 2) throw Exception(GLSX_ERR_ARG, "merge: 1 or 2 planes");
  std::vector<glsx_buf> planes(cn);
  for (int c = 0; c < cn; ++c) {
    if (src[c].empty()) throw Exception(GLSX_ERR_ARG, "merge: empty plane");
    planes[c] = src[c].handle();
  }
  const int type = cn == 1 ? CV_32FC1 : CV_32FC2;
  // getDstMat(size, type, dst) (glsMat.h:205-214)
  const bool fits = !dst.empty() && dst.rows == src[0].rows && dst.cols == src[0].cols && dst.type() == type;
  GlsMat out = fits ? dst : GlsMat(src[0].rows, src[0].cols, type);
  check(glsx_merge(planes.data(), cn, out.handle(), nullptr));
  dst = out;
}

GLS_API void minMaxLoc(const GlsMat& src, double* minVal, double* maxVal, Point* minLoc, Point* maxLoc,
                       const GlsMat& mask) {
  if (src.empty()) throw Exception(GLSX_ERR_ARG, "minMaxLoc: empty source");
  if (!mask.empty()) throw Exception(GLSX_ERR_ARG, "minMaxLoc: masks are not supported (the reference asserts an empty mask)");
  glsx_minmax_result r;
  check(glsx_min_max_loc(src.handle(), &r, nullptr));
  if (minVal) *minVal = r.min_val;
  if (maxVal) *maxVal = r.max_val;
  if (minLoc) *minLoc = Point(r.min_x, r.min_y);
  if (maxLoc) *maxLoc = Point(r.max_x, r.max_y);
}

// ---- phaseCorrelate (glsCV/glsPhaseCorrelate.cpp:142-236) -----------------------
GLS_API Point2d phaseCorrelate(const GlsMat& src1, const GlsMat& src2, GlsMat& window, double* response,
                               Point* peak) {
  if (src1.empty() || src2.empty()) throw Exception(GLSX_ERR_ARG, "phaseCorrelate: empty input");
  glsx_poc_result r;
  check(glsx_phase_correlate(src1.handle(), src2.handle(), window.empty() ? nullptr : window.handle(), &r,
                             nullptr));
  if (response) *response = r.response;
  if (peak) *peak = Point(r.peak_x, r.peak_y);
  return Point2d(r.shift_x, r.shift_y);
}

GLS_API Point2d phaseCorrelate(const GlsMat& src1, const GlsMat& src2, GlsMat& window, double* response) {
  return phaseCorrelate(src1, src2, window, response, nullptr);
}

}  // namespace gls

/*
 * glsCV.h -- umbrella header of the MI355X implementation of glsCV's FFT and
 * phase-correlation path (reference: glsCV/glsCV.h).  Same namespace and
 * function names as the reference; no OpenCV dependency: the few cv types the
 * path's signatures use (Size, Point, Point2d, type codes, DFT flags) are
 * provided here with OpenCV's numeric values.
 */
#ifndef GLSX_GLSCV_H
#define GLSX_GLSCV_H

#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "glsx.h"

/* OpenCV type codes and DFT flags with OpenCV's values (kept if OpenCV is
 * already included). */
#ifndef CV_32FC1
#define CV_8U 0
#define CV_32S 4
#define CV_32F 5
#define CV_8UC1 0
#define CV_32SC1 4
#define CV_32FC1 5
#define CV_32FC2 13
#define CV_MAT_DEPTH(flags) ((flags) & 7)
#define CV_MAT_CN(flags) ((((flags) >> 3) & 63) + 1)
#endif
namespace gls {

/* cv::DFT_INVERSE / cv::DFT_SCALE */
enum { DFT_INVERSE = 1, DFT_SCALE = 2 };

struct Size {
    int width = 0, height = 0;
    Size() = default;
    Size(int w, int h) : width(w), height(h) {}
    bool operator==(const Size& o) const { return width == o.width && height == o.height; }
    bool operator!=(const Size& o) const { return !(*this == o); }
};
struct Point {
    int x = 0, y = 0;
    Point() = default;
    Point(int x_, int y_) : x(x_), y(y_) {}
};
struct Point2d {
    double x = 0, y = 0;
    Point2d() = default;
    Point2d(double x_, double y_) : x(x_), y(y_) {}
};

/* GLS_Assert throws cv::Exception in the reference (glsCV/glsMacro.h:40). */
class Exception : public std::runtime_error {
  public:
    Exception(int status, const std::string& msg) : std::runtime_error(msg), code(status) {}
    int code;
};

/* The host image the reference's GlsMat exchanges with the CPU (cv::Mat in
 * glsCV/glsMat.h:95,153,156-157): rows x cols elements of a cv type, row
 * pitch `step` bytes, 2D only.  Copies share the pixels (cv::Mat's header
 * semantics); clone() copies them.  A Mat made from a caller's pointer
 * borrows that memory and never frees it. */
class Mat {
  public:
    int rows = 0;
    int cols = 0;
    unsigned char* data = nullptr;
    size_t step = 0; /* bytes from one row to the next */

    Mat() = default;
    Mat(int rows_, int cols_, int type) { create(rows_, cols_, type); }
    Mat(Size size, int type) { create(size.height, size.width, type); }
    /* borrowed pixels; step 0 = tightly packed rows */
    Mat(int rows_, int cols_, int type, void* pixels, size_t step_ = 0)
        : rows(rows_), cols(cols_), data((unsigned char*)pixels), type_(type) {
        check_type(type);
        step = step_ ? step_ : (size_t)cols_ * elemSize();
        if (step < (size_t)cols_ * elemSize()) throw Exception(GLSX_ERR_ARG, "Mat: step smaller than a row");
    }

    /* (re)allocates unless the size and type already match (cv::Mat::create) */
    void create(int rows_, int cols_, int type) {
        check_type(type);
        if (rows_ <= 0 || cols_ <= 0) throw Exception(GLSX_ERR_ARG, "Mat: rows and cols must be positive");
        if (mem_ && rows == rows_ && cols == cols_ && type_ == type && isContinuous()) return;
        type_ = type;
        rows = rows_;
        cols = cols_;
        step = (size_t)cols_ * elemSize();
        mem_ = std::shared_ptr<unsigned char[]>(new unsigned char[step * (size_t)rows_]());
        data = mem_.get();
    }
    void create(Size size, int type) { create(size.height, size.width, type); }
    /* a new image of zeros (cv::Mat::zeros) */
    static Mat zeros(int rows_, int cols_, int type) { return Mat(rows_, cols_, type); }
    static Mat zeros(Size size, int type) { return Mat(size, type); }

    int type() const { return type_; }
    int depth() const { return CV_MAT_DEPTH(type_); }
    int channels() const { return CV_MAT_CN(type_); }
    size_t elemSize() const {
        static const int depth_bytes[8] = {1, 1, 2, 2, 4, 4, 8, 2};
        return (size_t)depth_bytes[depth()] * channels();
    }
    size_t total() const { return (size_t)rows * cols; }
    bool empty() const { return data == nullptr || total() == 0; }
    bool isContinuous() const { return rows <= 1 || step == (size_t)cols * elemSize(); }
    Size size() const { return Size(cols, rows); }

    template <class T>
    T* ptr(int row = 0) { return (T*)(data + (size_t)row * step); }
    template <class T>
    const T* ptr(int row = 0) const { return (const T*)(data + (size_t)row * step); }
    template <class T>
    T& at(int row, int col) { return ptr<T>(row)[col]; }
    template <class T>
    const T& at(int row, int col) const { return ptr<T>(row)[col]; }

    void copyTo(Mat& m) const {
        if (m.data == data && m.step == step && m.rows == rows && m.cols == cols) return;
        m.create(rows, cols, type_);
        for (int r = 0; r < rows; ++r) std::memcpy(m.ptr<unsigned char>(r), ptr<unsigned char>(r), cols * elemSize());
    }
    Mat clone() const {
        Mat m;
        if (!empty()) copyTo(m);
        return m;
    }

  private:
    static void check_type(int type) {
        if (type < 0 || type >= 512 || CV_MAT_DEPTH(type) == 7) throw Exception(GLSX_ERR_ARG, "Mat: bad type");
    }
    std::shared_ptr<unsigned char[]> mem_;
    int type_ = 0;
};

/* glsCV/glsCV.h:67-68 */
void glsCvInit(int device = 0);
void glsCvTerminate(void);

}  // namespace gls

#include "glsMat.h"
#include "glsFft.h"
#include "glsPhaseCorrelate.h"
#include "glsBasicOperation.h"
#include "glsNormalize.h"
#include "glsSplit.h"
#include "glsMerge.h"
#include "glsMinMaxLoc.h"

#endif

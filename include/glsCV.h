/*
 * glsCV.h -- umbrella header of the MI355X implementation of glsCV's FFT and
 * phase-correlation path (reference: glsCV/glsCV.h).  Same namespace and
 * function names as the reference; no OpenCV dependency: the few cv types the
 * path's signatures use (Size, Point, Point2d, type codes, DFT flags) are
 * provided here with OpenCV's numeric values.
 */
#ifndef GLSX_GLSCV_H
#define GLSX_GLSCV_H

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

/* glsCV/glsCV.h:67-68 */
void glsCvInit(int device = 0);
void glsCvTerminate(void);

}  // namespace gls

#include "glsMat.h"
#include "glsFft.h"
#include "glsPhaseCorrelate.h"
#include "glsBasicOperation.h"
#include "glsNormalize.h"

#endif

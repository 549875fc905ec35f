/*
 * glsMinMaxLoc.h -- gls::minMaxLoc (reference: glsCV/glsMinMaxLoc.h:42,
 * glsMinMaxLoc.cpp:149-199).  Channel 0 of a CV_32FC1 / CV_32FC2 image: the
 * first minimum and maximum in row-major order, locations as Point(x, y).
 * NaN is never chosen; with no value below FLT_MAX (above -FLT_MAX) the
 * result is FLT_MAX (-FLT_MAX) at (0, 0), as the reference's strict scans
 * give.  Any output pointer may be null.  mask must be empty (the reference
 * asserts that too); the call waits for the result, as the reference
 * downloads it.
 */
#ifndef GLSX_GLSMINMAXLOC_H
#define GLSX_GLSMINMAXLOC_H
#include "glsMat.h"
namespace gls {
struct Point;
void minMaxLoc(const GlsMat& src, double* minVal, double* maxVal = 0, Point* minLoc = 0, Point* maxLoc = 0,
               const GlsMat& mask = GlsMat());
}  // namespace gls
#endif

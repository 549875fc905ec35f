/*
 * glsPhaseCorrelate.h -- gls::phaseCorrelate (reference:
 * glsCV/glsPhaseCorrelate.h:58).  Returns the position of src2 relative to
 * src1 ((cols/2, rows/2) minus the 5x5 weighted centroid of the fft-shifted
 * correlation peak); *response = 5x5 sum / (rows*cols).  window may be an
 * empty GlsMat.
 */
#ifndef GLSX_GLSPHASECORRELATE_H
#define GLSX_GLSPHASECORRELATE_H

namespace gls {

class GlsMat;
struct Point2d;
struct Point;

Point2d phaseCorrelate(const GlsMat& src1, const GlsMat& src2, GlsMat& window, double* response = 0);

/* Same, also reporting the integer peak of the correlation surface. */
Point2d phaseCorrelate(const GlsMat& src1, const GlsMat& src2, GlsMat& window, double* response, Point* peak);

}  // namespace gls

#endif

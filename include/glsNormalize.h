/*
 * glsNormalize.h -- gls::normalize (reference: glsCV/glsNormalize.h:47,
 * glsNormalize.cpp:52-75).  src: CV_32FC1; dst: CV_32FC1 of the same size
 * (reused or reallocated, in place allowed).  NORM_MINMAX maps the minimum to
 * alpha and the maximum to beta; NORM_L1 / NORM_L2 scale the plane to norm
 * alpha.  Other norm types throw, as the reference's GLS_Assert does.
 */
#ifndef GLSX_GLSNORMALIZE_H
#define GLSX_GLSNORMALIZE_H
namespace gls {
class GlsMat;
/* cv::NormTypes values */
enum { NORM_INF = 1, NORM_L1 = 2, NORM_L2 = 4, NORM_MINMAX = 32 };
void normalize(const GlsMat& src, GlsMat& dst, double alpha = 1, double beta = 0, int normType = NORM_L2);
}  // namespace gls
#endif

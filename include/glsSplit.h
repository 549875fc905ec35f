/*
 * glsSplit.h -- gls::split (reference: glsCV/glsSplit.h:42,
 * glsSplit.cpp:192-210).  dst is resized to the source's channel count and
 * every element assigned a new CV_32FC1 image of the source's size, as the
 * reference does.  Source: CV_32FC1 or CV_32FC2.
 */
#ifndef GLSX_GLSSPLIT_H
#define GLSX_GLSSPLIT_H
#include <vector>
namespace gls {
class GlsMat;
void split(const GlsMat& src, std::vector<GlsMat>& dst);
}  // namespace gls
#endif

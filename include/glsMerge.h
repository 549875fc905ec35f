/*
 * glsMerge.h -- gls::merge (reference: glsCV/glsMerge.h:42,
 * glsMerge.cpp:193-211).  One or two CV_32FC1 planes of one size into a
 * CV_32FC1 / CV_32FC2 image; dst is reused when it already has that size and
 * type, otherwise assigned a new image.  Every plane must be non-empty (the
 * reference reads an unset sampler for an empty one).
 */
#ifndef GLSX_GLSMERGE_H
#define GLSX_GLSMERGE_H
#include <vector>
namespace gls {
class GlsMat;
void merge(const std::vector<GlsMat>& src, GlsMat& dst);
}  // namespace gls
#endif

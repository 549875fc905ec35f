/*
 * glsBasicOperation.h -- the spectrum operations of glsCV's basic-operation
 * set that consume the FFT output (reference: glsCV/glsBasicOperation.h,
 * glsBasicOperation.cpp:337-351).  src: CV_32FC2 spectrum; dst: CV_32FC1 of
 * the same size, reused when it already has that size and type (getDstMat),
 * otherwise reallocated (so dst may be the src handle, as glsCam calls it).
 */
#ifndef GLSX_GLSBASICOPERATION_H
#define GLSX_GLSBASICOPERATION_H
namespace gls {
class GlsMat;
/* dst = sqrt(re^2 + im^2) */
void magSpectrums(const GlsMat& src, GlsMat& dst);
/* dst = log(abs(sqrt(re^2 + im^2) + offset)) */
void logMagSpectrums(const GlsMat& src, GlsMat& dst, float offset);
/* Extension: gls::fft(src, tmp, flag) followed by magSpectrums /
 * logMagSpectrums(tmp, dst, offset), fused into the FFT's last pass. */
void fftMagSpectrums(const GlsMat& src, GlsMat& dst, int flag, bool logMag, float offset = 1.0f);
/* Extension: glsCam's FFT view, fft(src, flag) -> logMagSpectrums(offset) ->
 * normalize(alpha, beta, NORM_MINMAX), in three launches. */
void spectrumView(const GlsMat& src, GlsMat& dst, int flag, float offset = 1.0f, double alpha = 0.0,
                  double beta = 1.0);
}  // namespace gls
#endif

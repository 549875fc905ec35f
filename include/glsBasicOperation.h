/*
 * glsBasicOperation.h -- the operations of glsCV's basic-operation set on the
 * FFT / phase-correlation path (reference: glsCV/glsBasicOperation.h:56-88).
 * dst is reused when it already has the result's size and type and is not
 * one of the sources (getDstMat, glsMat.h:173-214), otherwise it is assigned
 * a new image; so dst may be a source handle, as the reference's callers
 * pass it.
 */
#ifndef GLSX_GLSBASICOPERATION_H
#define GLSX_GLSBASICOPERATION_H
namespace gls {
class GlsMat;
/* dst = src0 * src1 per element and channel; CV_32FC1 or CV_32FC2, sources
 * of one size and type (glsBasicOperation.cpp:265-272). */
void multiply(const GlsMat& src0, const GlsMat& src1, GlsMat& dst);
/* CV_32FC2 spectra: dst = src0 * src1, or src0 * conj(src1) with conj
 * (glsBasicOperation.cpp:320-327). */
void mulSpectrums(const GlsMat& src0, const GlsMat& src1, GlsMat& dst, bool conj = false);
/* dst = c / |c|, c = src0 * conj(src1); 0 where |c| <= 1e-30 (the reference
 * divides 0/0 into NaN) (glsBasicOperation.cpp:329-335). */
void mulSpectrumsPhaseOnly(const GlsMat& src0, const GlsMat& src1, GlsMat& dst);
/* src: CV_32FC2 spectrum, dst: CV_32FC1.  dst = sqrt(re^2 + im^2) */
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

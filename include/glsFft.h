/*
 * glsFft.h -- gls::fft (reference: glsCV/glsFft.h:43-55), same flags and
 * signature.  src: CV_32FC2 (or CV_32FC1, imag = 0), power-of-two rows and
 * cols up to 16384 (1 x N and N x 1 are 1D); dst: CV_32FC2 of the same size,
 * reused when it already has that size and type (getDstMat), in place allowed.
 */
#ifndef GLSX_GLSFFT_H
#define GLSX_GLSFFT_H

namespace gls {

class GlsMat;

enum {
    GLS_FFT_SCALE = 2,    /* DFT_SCALE: 1/(rows*cols) */
    GLS_FFT_INVERSE = 1,  /* DFT_INVERSE */
    GLS_FFT_SHIFT = 64    /* fftshift of the output */
};

void fft(const GlsMat& src, GlsMat& dst, int flag = 0);

/* sugar: fft(src, dst, GLS_FFT_INVERSE | GLS_FFT_SCALE) */
void ifft(const GlsMat& src, GlsMat& dst, int flag = GLS_FFT_SCALE);

}  // namespace gls

#endif

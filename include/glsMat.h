/*
 * glsMat.h -- gls::GlsMat, a refcounted handle to a device buffer
 * (reference: glsCV/glsMat.h:45-214, where it holds a GL texture).
 * Shallow copy on assignment, device memory freed with the last reference,
 * clone() is a deep copy, data is row major with pitch cols * elemSize.
 */
#ifndef GLSX_GLSMAT_H
#define GLSX_GLSMAT_H

#include <cstddef>
#include <vector>

#include "glsx.h"

namespace gls {

struct Size;
class Mat;

class GlsMat {
  public:
    int rows = 0;
    int cols = 0;

    GlsMat() = default;
    GlsMat(const Size& size, int type);
    GlsMat(int rows, int cols, int type);
    /* a buffer of the host image's size and type, uploaded (glsMat.cpp:90-96) */
    explicit GlsMat(const Mat& host);
    GlsMat(const GlsMat& o);
    GlsMat& operator=(const GlsMat& o);
    ~GlsMat();

    int type() const { return type_; }
    int channels() const { return ((type_ >> 3) & 63) + 1; }
    int depth() const { return type_ & 7; }
    bool empty() const { return buf_ == nullptr; }
    bool isContinuous() const { return true; }
    size_t total() const { return (size_t)rows * cols; }
    size_t elemSize() const;
    Size size() const;

    /* GlsMat::upload / download (glsMat.cpp:374-419, 493-552) with the
     * reference's host image: upload requires the same size and type;
     * download writes into `dst` when its size and type match (any step),
     * otherwise `dst` becomes a new image; the conversion downloads into a
     * new image. */
    void upload(const Mat& src);
    void download(Mat& dst) const;
    operator Mat() const;
    /* the same on raw host memory; pitch in bytes, 0 = tightly packed rows */
    void upload(const void* host, size_t host_pitch = 0);
    void download(void* host, size_t host_pitch = 0) const;
    template <class T>
    std::vector<T> download() const {
        std::vector<T> v(total() * elemSize() / sizeof(T));
        download(v.data());
        return v;
    }

    void copyTo(GlsMat& m) const;
    GlsMat clone() const;

    glsx_buf handle() const { return buf_; }
    void* devPtr() const;
    /* identity of the underlying storage (the reference compares texid()) */
    const void* id() const { return buf_; }

  private:
    void create(int rows, int cols, int type);
    void reset();
    glsx_buf buf_ = nullptr;
    int type_ = -1;
};

}  // namespace gls

using gls::GlsMat;

#endif

// Shared device-side building blocks of the LES library: complex arithmetic,
// the LDS Stockham FFT used by every horizontal transform, launch helpers.
//
// FFT algorithm (fixed so results are reproducible across devices and against
// the CPU restatement used in testing): complex length-n transforms with
// n = 2^a 3^b as a Stockham autosort sweep of radix-4 stages, one radix-2
// stage when a is odd, then radix-3 stages.  Stage s (radix R, span p = the
// product of earlier radices): for i < n/R with k = i mod p the inputs
// x_r = in[i + r n/R] (r >= 1) are multiplied by exp(sign 2 pi i r k/(pR)),
// combined by a size-R DFT and written to out[(i - k) R + k + r p].  The
// transforms are unnormalised like FFTW's (src/decomp_2d_fft_2d.f90:397-577).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LES_MAX_STAGES 10

template <typename T>
struct Cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ Cplx<T> c_add(Cplx<T> a, Cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cplx<T> c_sub(Cplx<T> a, Cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ Cplx<T> c_mul(Cplx<T> a, Cplx<T> w) {
  return {a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}
template <typename T>
__device__ __forceinline__ Cplx<T> c_scale(Cplx<T> a, T s) { return {a.x * s, a.y * s}; }
// i*k*a with k real, as the reference's eye*kkx*f_hat
template <typename T>
__device__ __forceinline__ Cplx<T> c_ik(T k, Cplx<T> a) { return {-(k * a.y), k * a.x}; }

// Plan of one transform length (host-built, passed by value to kernels).
struct FftPlan {
  int n, nst;
  int R[LES_MAX_STAGES], P[LES_MAX_STAGES], off[LES_MAX_STAGES];
  int twf, twi;  // offsets of the forward / inverse twiddle tables in the twiddle pool
};

template <typename T>
__device__ __forceinline__ void bfly(Cplx<T>* x, int R, int sgn) {
  if (R == 4) {
    Cplx<T> t0 = c_add(x[0], x[2]), t1 = c_sub(x[0], x[2]);
    Cplx<T> t2 = c_add(x[1], x[3]), t3 = c_sub(x[1], x[3]);
    Cplx<T> t3r = (sgn < 0) ? Cplx<T>{t3.y, -t3.x} : Cplx<T>{-t3.y, t3.x};
    x[0] = c_add(t0, t2);
    x[2] = c_sub(t0, t2);
    x[1] = c_add(t1, t3r);
    x[3] = c_sub(t1, t3r);
  } else if (R == 2) {
    Cplx<T> y0 = c_add(x[0], x[1]), y1 = c_sub(x[0], x[1]);
    x[0] = y0;
    x[1] = y1;
  } else {
    const T h = (T)0.86602540378443864676372317075294;
    Cplx<T> t1 = c_add(x[1], x[2]);
    Cplx<T> t2 = {x[0].x - (T)0.5 * t1.x, x[0].y - (T)0.5 * t1.y};
    Cplx<T> t3 = c_sub(x[1], x[2]);
    Cplx<T> t3r = (sgn < 0) ? Cplx<T>{h * t3.y, -(h * t3.x)} : Cplx<T>{-(h * t3.y), h * t3.x};
    x[0] = c_add(x[0], t1);
    x[1] = c_add(t2, t3r);
    x[2] = c_sub(t2, t3r);
  }
}

// Transform `nlines` lines held in LDS at a + l*stride (l < nlines) with the
// block's threads; b is scratch of the same shape.  Returns the buffer that
// holds the result (a or b).  Ends with a barrier.
template <typename T>
__device__ Cplx<T>* lds_fft(Cplx<T>* a, Cplx<T>* b, int nlines, int stride, const FftPlan& pl,
                            const Cplx<T>* __restrict__ twpool, bool inverse) {
  const int n = pl.n;
  const int sgn = inverse ? 1 : -1;
  const Cplx<T>* tw0 = twpool + (inverse ? pl.twi : pl.twf);
  Cplx<T>* in = a;
  Cplx<T>* out = b;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.R[s], p = pl.P[s], m = n / R;
    const Cplx<T>* tw = tw0 + pl.off[s];
    const int total = nlines * m;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int l = idx / m;
      const int i = idx - l * m;
      const int k = i % p;
      const Cplx<T>* src = in + l * stride;
      Cplx<T> x[4];
      x[0] = src[i];
      for (int r = 1; r < R; ++r) x[r] = c_mul(src[i + r * m], tw[(r - 1) * p + k]);
      bfly(x, R, sgn);
      Cplx<T>* dst = out + l * stride;
      const int j = (i - k) * R + k;
      for (int r = 0; r < R; ++r) dst[j + r * p] = x[r];
    }
    __syncthreads();
    Cplx<T>* t = in;
    in = out;
    out = t;
  }
  return in;
}

// Twiddle pool of one precision, plus plans of the four lengths in use.
template <typename T>
struct FftSet {
  FftPlan nx, ny, nx2, ny2;
  const Cplx<T>* tw;
};

#define LES_CHECK(x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      les_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      return 5;                                                                          \
    }                                                                                    \
  } while (0)

void les_set_error(const char* fmt, ...);

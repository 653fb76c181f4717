// Oracle FFT core -- TEST INFRASTRUCTURE ONLY.
//
// This file belongs to the CPU oracle of the LES time step. It is used by the
// tests, the smoke check and the CPU-baseline leg of the benchmark, never by
// the product library.
//
// It restates, on the CPU, the discrete Fourier transforms the reference gets
// from FFTW (src/decomp_2d_fft_2d.f90:397-577, src/fft_engine.f90:47-83):
// unnormalised forward (sign -1) and backward (sign +1) transforms.  The
// algorithm is fixed (and documented) so that the GPU kernels can perform the
// same floating-point operations in the same order:
//   * complex length-n transforms, n = 2^a 3^b, as a Stockham autosort sweep of
//     radix-4 stages, then one radix-2 stage if a is odd, then radix-3 stages;
//   * stage s with radix R and span p (product of earlier radices): for every
//     i < n/R with k = i mod p, inputs x_r = in[i + r n/R] are multiplied by
//     the twiddle exp(sign 2 pi i r k / (p R)) (r >= 1), combined by a size-R
//     DFT and stored at out[(i - k) R + k + r p];
//   * the twiddle product a w is (fma(a.x, w.x, -(a.y w.y)), fma(a.x, w.y,
//     a.y w.x)): fused multiply-adds, each rounded once;
//   * twiddles are computed once on the host in double with cos/sin of
//     sign * 2 pi (r k) / (p R);
//   * real rows are transformed two at a time: z = a + i b, Z = FFT(z),
//     A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i); the
//     inverse packs Z_k = A_k + i B_k with the imaginary parts of A_0, A_{n/2},
//     B_0, B_{n/2} dropped (FFTW's c2r semantics).
// Parity of this restatement is pinned against numpy.fft in tests/.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

namespace orc {

struct Cx {
  double x, y;
};

static inline Cx cadd(Cx a, Cx b) { return {a.x + b.x, a.y + b.y}; }
static inline Cx csub(Cx a, Cx b) { return {a.x - b.x, a.y - b.y}; }
static inline Cx cmul(Cx a, Cx w) { return {std::fma(a.x, w.x, -(a.y * w.y)), std::fma(a.x, w.y, a.y * w.x)}; }
static inline Cx cscale(Cx a, double s) { return {a.x * s, a.y * s}; }

static const double kTwoPi = 6.283185307179586476925286766559;
static const double kSqrt3Half = 0.86602540378443864676372317075294;

struct Plan {
  int n = 0, nst = 0;
  int R[16], P[16], off[16];
  std::vector<Cx> twf, twi;  // forward (sign -1) and inverse (sign +1) twiddles

  void build(int n_) {
    n = n_;
    int a = 0, b = 0, m = n;
    while (m % 2 == 0) { m /= 2; ++a; }
    while (m % 3 == 0) { m /= 3; ++b; }
    if (m != 1 || n < 2) { n = 0; return; }
    nst = 0;
    while (a >= 2) { R[nst++] = 4; a -= 2; }
    if (a == 1) R[nst++] = 2;
    for (int t = 0; t < b; ++t) R[nst++] = 3;
    int p = 1, o = 0;
    for (int s = 0; s < nst; ++s) {
      P[s] = p;
      off[s] = o;
      o += (R[s] - 1) * p;
      p *= R[s];
    }
    twf.resize(o > 0 ? o : 1);
    twi.resize(o > 0 ? o : 1);
    for (int s = 0; s < nst; ++s) {
      for (int r = 1; r < R[s]; ++r)
        for (int k = 0; k < P[s]; ++k) {
          double ang = kTwoPi * (double)(r * k) / (double)(P[s] * R[s]);
          double c = std::cos(ang), sn = std::sin(ang);
          twf[off[s] + (r - 1) * P[s] + k] = {c, -sn};
          twi[off[s] + (r - 1) * P[s] + k] = {c, sn};
        }
    }
  }

  // one radix-R butterfly on x[0..R-1]; sgn = -1 forward, +1 inverse
  static inline void bfly(Cx* x, int Rr, int sgn) {
    if (Rr == 2) {
      Cx y0 = cadd(x[0], x[1]), y1 = csub(x[0], x[1]);
      x[0] = y0; x[1] = y1;
    } else if (Rr == 4) {
      Cx t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
      Cx t2 = cadd(x[1], x[3]), t3 = csub(x[1], x[3]);
      Cx t3r = (sgn < 0) ? Cx{t3.y, -t3.x} : Cx{-t3.y, t3.x};
      x[0] = cadd(t0, t2);
      x[2] = csub(t0, t2);
      x[1] = cadd(t1, t3r);
      x[3] = csub(t1, t3r);
    } else {  // 3
      Cx t1 = cadd(x[1], x[2]);
      Cx t2 = {x[0].x - 0.5 * t1.x, x[0].y - 0.5 * t1.y};
      Cx t3 = csub(x[1], x[2]);
      Cx t3r = (sgn < 0) ? Cx{kSqrt3Half * t3.y, -(kSqrt3Half * t3.x)}
                         : Cx{-(kSqrt3Half * t3.y), kSqrt3Half * t3.x};
      x[0] = cadd(x[0], t1);
      x[1] = cadd(t2, t3r);
      x[2] = csub(t2, t3r);
    }
  }

  // in-place transform of data[0..n-1] using scratch[0..n-1]
  void exec(Cx* data, Cx* scratch, bool inverse) const {
    const std::vector<Cx>& tw = inverse ? twi : twf;
    int sgn = inverse ? 1 : -1;
    Cx* in = data;
    Cx* out = scratch;
    for (int s = 0; s < nst; ++s) {
      int Rr = R[s], p = P[s], m = n / Rr;
      const Cx* t = tw.data() + off[s];
      for (int i = 0; i < m; ++i) {
        int k = i % p;
        Cx x[4];
        x[0] = in[i];
        for (int r = 1; r < Rr; ++r) x[r] = cmul(in[i + r * m], t[(r - 1) * p + k]);
        bfly(x, Rr, sgn);
        int j = (i - k) * Rr + k;
        for (int r = 0; r < Rr; ++r) out[j + r * p] = x[r];
      }
      Cx* tmp = in;
      in = out;
      out = tmp;
    }
    if (in != data) std::memcpy(data, in, sizeof(Cx) * n);
  }
};

// Split Z = FFT(a + i b) into the half spectra A, B (k = 0..n/2).
static inline void unpack_pair(const Cx* Z, int n, Cx* A, Cx* B) {
  for (int k = 0; k <= n / 2; ++k) {
    Cx zk = Z[k], zm = Z[(n - k) % n];
    A[k] = {0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y)};
    B[k] = {0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x)};
  }
}

// Pack half spectra A, B (k = 0..n/2) into Z = A + i B over k = 0..n-1.
static inline void pack_pair(const Cx* A, const Cx* B, int n, Cx* Z) {
  for (int k = 0; k <= n / 2; ++k) {
    Cx a = A[k], b = B[k];
    if (k == 0 || k == n / 2) { a.y = 0.0; b.y = 0.0; }
    Z[k] = {a.x - b.y, a.y + b.x};
    if (k != 0 && k != n / 2) Z[n - k] = {a.x + b.y, b.x - a.y};
  }
}

}  // namespace orc

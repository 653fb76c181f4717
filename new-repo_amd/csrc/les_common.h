// Shared device-side building blocks of the LES library: complex arithmetic,
// the wave-level LDS Stockham FFT used by every horizontal transform.
//
// FFT algorithm (fixed so results are reproducible across devices and against
// the CPU restatement used in testing): complex length-n transforms with
// n = 2^a 3^b as a Stockham autosort sweep of radix-4 stages, one radix-2
// stage when a is odd, then radix-3 stages.  Stage s (radix R, span p = the
// product of earlier radices): for i < n/R with k = i mod p the inputs
// x_r = in[i + r n/R] (r >= 1) are multiplied by exp(sign 2 pi i r k/(pR))
// (fused multiply-adds: (fma(a.x, w.x, -(a.y w.y)), fma(a.x, w.y, a.y w.x))),
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
// Twiddle product with fused multiply-adds (one rounding each), four
// instructions instead of six.
template <typename T>
__device__ __forceinline__ Cplx<T> c_mul(Cplx<T> a, Cplx<T> w) {
  return {fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x)};
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

// Wave-level transform of one line of n <= LES_MAX_N points held in LDS,
// in place.  The 64 lanes of a wavefront own the line: every stage first
// loads the inputs of all of the lane's butterflies into registers, then
// writes the outputs, so no other wave and no block barrier is involved.
// Within a wavefront LDS instructions complete in issue order; wave_sync()
// keeps the compiler from moving LDS accesses across the phase boundaries.
#define LES_WAVE 64
#define LES_MAX_N 1024

// LDS placement inside a transform.  The Stockham writes out[(i-k)R + k +
// r p] of the early stages hit few banks (stride R, or runs of p < 16), so
// between the first stage's reads and the last stage's writes (both
// contiguous across lanes, hence conflict-free) element e of a line lives at
// lds_sw(e): within each aligned block of 16 elements, e ^ (5 (e >> 4) mod
// 16).  This spreads every stage's 16 consecutive lanes over distinct bank
// pairs for 8-byte elements (one group of 8 lanes per 16-byte element).  The
// line is in natural order outside the transform.  Needs n % 16 == 0.
__device__ __forceinline__ int lds_sw(int e) { return e ^ (((e >> 4) * 5) & 15); }

// Alternative placement for the compile-time transforms: lds_pad(e) = e +
// 2 (e >> 4) + (e >> 5), two pad slots after every 16 elements and one more
// after every 32 (a line then holds lds_line_cap(n) elements).  Being
// additive, it makes each access of a stage one per-lane base plus an
// immediate offset, where the XOR costs a few integer instructions per
// access; it leaves slightly more bank conflicts (about 1.1-1.25 of the
// ideal cycles against 1.0-1.1).  The row kernels in FP32, integer-bound,
// use it; everything else uses lds_sw.
__host__ __device__ constexpr int lds_pad(int e) { return e + 2 * (e >> 4) + (e >> 5); }
__host__ __device__ constexpr int lds_line_cap(int n) { return lds_pad(n - 1) + 1; }
enum { LES_LAY_XOR = 0, LES_LAY_PAD = 1 };

// One 8- or 16-byte LDS access per complex value (ds_read_b64 / _b128)
// instead of two 4- or 8-byte ones.
template <typename T>
struct Vec2T;
template <>
struct Vec2T<float> { using type = float2; };
template <>
struct Vec2T<double> { using type = double2; };
template <typename T>
__device__ __forceinline__ Cplx<T> lds_ld(const Cplx<T>* p) {
  const typename Vec2T<T>::type v = *reinterpret_cast<const typename Vec2T<T>::type*>(p);
  return {v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void lds_st(Cplx<T>* p, Cplx<T> a) {
  typename Vec2T<T>::type v;
  v.x = a.x;
  v.y = a.y;
  *reinterpret_cast<typename Vec2T<T>::type*>(p) = v;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// k = i mod p of butterfly i: KM 0 mask (p a power of two: every radix-4
// and radix-2 stage, which precede the radix-3 ones, and the first radix-3
// stage), KM 1 identity (last stage: i < n/R = p), KM 2 true modulo (middle
// radix-3 stages).  A compile-time choice, so no division is ever evaluated
// where it is not needed.
template <int KM>
__device__ __forceinline__ int stage_k(int i, int p) {
  if constexpr (KM == 0) return i & (p - 1);
  else if constexpr (KM == 1) return i;
  else return i % p;
}

// swr / sww: the stage reads / writes through lds_sw.
template <typename T, int R, int KM>
__device__ __forceinline__ void wave_stage(Cplx<T>* z, int n, int p, const Cplx<T>* __restrict__ tw, int sgn, int lane,
                                           bool swr, bool sww) {
  constexpr int MAXB = (LES_MAX_N / R + LES_WAVE - 1) / LES_WAVE;
  const int m = n / R;
  Cplx<T> x[MAXB][4];
  int kk[MAXB];
  // addresses are recomputed in each stage, not kept live across stages
  __asm__ volatile("" : "+v"(lane));
#pragma unroll
  for (int t = 0; t < MAXB; ++t) {
    const int i = lane + LES_WAVE * t;
    if (i < m) {
      const int k = stage_k<KM>(i, p);
      kk[t] = k;
      x[t][0] = lds_ld(z + (swr ? lds_sw(i) : i));
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int e = i + r * m;
        x[t][r] = c_mul(lds_ld(z + (swr ? lds_sw(e) : e)), tw[(r - 1) * p + k]);
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int t = 0; t < MAXB; ++t) {
    const int i = lane + LES_WAVE * t;
    if (i < m) {
      const int k = kk[t];
      bfly(x[t], R, sgn);
      const int j = (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = j + r * p;
        lds_st(z + (sww ? lds_sw(e) : e), x[t][r]);
      }
    }
  }
  wave_sync();
}

template <typename T>
__device__ __forceinline__ void wave_fft(Cplx<T>* z, const FftPlan& pl, const Cplx<T>* __restrict__ twpool, bool inverse,
                                         int lane) {
  const int sgn = inverse ? 1 : -1;
  const Cplx<T>* tw0 = twpool + (inverse ? pl.twi : pl.twf);
  const bool swz = (pl.n & 15) == 0;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.R[s], p = pl.P[s];
    const Cplx<T>* tw = tw0 + pl.off[s];
    const bool swr = swz && s > 0, sww = swz && s < pl.nst - 1;
    if (R == 4) {
      wave_stage<T, 4, 0>(z, pl.n, p, tw, sgn, lane, swr, sww);
    } else if (R == 2) {
      wave_stage<T, 2, 0>(z, pl.n, p, tw, sgn, lane, swr, sww);
    } else if (p * 3 == pl.n) {
      wave_stage<T, 3, 1>(z, pl.n, p, tw, sgn, lane, swr, sww);
    } else if ((p & (p - 1)) == 0) {
      wave_stage<T, 3, 0>(z, pl.n, p, tw, sgn, lane, swr, sww);
    } else {
      wave_stage<T, 3, 2>(z, pl.n, p, tw, sgn, lane, swr, sww);
    }
  }
}

// ---------------------------------------------------------------- fixed lengths
// Compile-time plans for the production lengths, built by the host's rule
// (radix-4 stages, one radix-2 stage when the power of two is odd, then
// radix-3 stages).  A wavefront loads its lanes' forward twiddles of every
// stage into registers once (TwReg), so the stages touch only LDS; inverse
// transforms use the conjugates, which equal the inverse table bit for bit
// (the host stores (cos, -sin) and (cos, sin)).
constexpr int cp_a(int n) { int a = 0; while (n % 2 == 0) { n /= 2; ++a; } return a; }
constexpr int cp_b(int n) { while (n % 2 == 0) n /= 2; int b = 0; while (n % 3 == 0) { n /= 3; ++b; } return b; }
constexpr int cp_nst(int n) { return cp_a(n) / 2 + (cp_a(n) & 1) + cp_b(n); }
constexpr int cp_R(int n, int s) {
  return s < cp_a(n) / 2 ? 4 : ((cp_a(n) & 1) && s == cp_a(n) / 2 ? 2 : 3);
}
constexpr int cp_P(int n, int s) { int p = 1; for (int t = 0; t < s; ++t) p *= cp_R(n, t); return p; }
constexpr int cp_OFF(int n, int s) { int o = 0; for (int t = 0; t < s; ++t) o += (cp_R(n, t) - 1) * cp_P(n, t); return o; }
constexpr int cp_MB(int n, int s) { return (n / cp_R(n, s) + LES_WAVE - 1) / LES_WAVE; }
constexpr int cp_BASE(int n, int s) { int c = 0; for (int t = 0; t < s; ++t) c += cp_MB(n, t) * (cp_R(n, t) - 1); return c; }
constexpr int cp_NTW(int n) { return cp_BASE(n, cp_nst(n)); }

template <typename T, int N>
struct TwReg {
  Cplx<T> w[cp_NTW(N)];
};

template <typename T, int N, int S>
__device__ __forceinline__ void tw_load_s(Cplx<T>* w, const Cplx<T>* __restrict__ twf, int lane) {
  if constexpr (S < cp_nst(N)) {
    constexpr int R = cp_R(N, S), P = cp_P(N, S), OFF = cp_OFF(N, S), MB = cp_MB(N, S), BASE = cp_BASE(N, S);
    constexpr int M = N / R;
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int i = lane + LES_WAVE * b;
      if (i < M) {
        const int k = i % P;
#pragma unroll
        for (int r = 1; r < R; ++r) w[BASE + b * (R - 1) + r - 1] = twf[OFF + (r - 1) * P + k];
      }
    }
    tw_load_s<T, N, S + 1>(w, twf, lane);
  }
}

// the forward table of the plan of length N (twpool + plan.twf)
template <typename T, int N>
__device__ __forceinline__ void tw_load(TwReg<T, N>& t, const Cplx<T>* __restrict__ twf, int lane) {
  tw_load_s<T, N, 0>(t.w, twf, lane);
}

// Offsets lds_pad(W + r P) - lds_pad(W) of the writes of stage S are the
// same for every butterfly W (checked at compile time for each plan used).
template <int N, int S>
constexpr bool cp_wsplit() {
  constexpr int R = cp_R(N, S), P = cp_P(N, S), M = N / R;
  for (int i = 0; i < M; ++i) {
    const int k = i % P, W = (i - k) * R + k;
    for (int r = 0; r < R; ++r)
      if (lds_pad(W + r * P) - lds_pad(W) != lds_pad(r * P)) return false;
  }
  return true;
}

template <typename T, int N, int S, int LAY>
__device__ __forceinline__ void wave_fft_s(Cplx<T>* z, const Cplx<T>* w, bool inverse, int lane) {
  static_assert(N % 16 == 0, "the LDS placements work on blocks of 16 elements");
  if constexpr (S < cp_nst(N)) {
    constexpr int R = cp_R(N, S), P = cp_P(N, S), MB = cp_MB(N, S), BASE = cp_BASE(N, S);
    constexpr int M = N / R;
    constexpr bool SWR = S > 0, SWW = S < cp_nst(N) - 1;
    constexpr bool PAD = LAY == LES_LAY_PAD;
    // padded reads lane + 64 b + r M split into lds_pad(lane) + lds_pad(64 b +
    // r M) when the constant is a multiple of 32
    static_assert(!PAD || !SWR || M % 32 == 0, "padded reads need M % 32 == 0");
    static_assert(!PAD || !SWW || cp_wsplit<N, S>(), "padded write offsets must not depend on the lane");
    const int sgn = inverse ? 1 : -1;
    Cplx<T> x[MB][4];
    if constexpr (!PAD) {
      // XOR addresses are recomputed in each stage, not kept live across stages
      __asm__ volatile("" : "+v"(lane));
    }
    const int rb = (PAD && SWR) ? lds_pad(lane) : lane;
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int i = lane + LES_WAVE * b;
      if (i < M) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int c = LES_WAVE * b + r * M;
          int a;
          if constexpr (PAD) a = rb + (SWR ? lds_pad(c) : c);
          else a = SWR ? lds_sw(i + r * M) : i + r * M;
          Cplx<T> v = lds_ld(z + a);
          if (r > 0) {
            Cplx<T> tw = w[BASE + b * (R - 1) + r - 1];
            if (inverse) tw.y = -tw.y;
            v = c_mul(v, tw);
          }
          x[b][r] = v;
        }
      }
    }
    wave_sync();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int i = lane + LES_WAVE * b;
      if (i < M) {
        const int k = (int)((unsigned)i % (unsigned)P);
        bfly(x[b], R, sgn);
        const int j = (i - k) * R + k;
        const int wb = (PAD && SWW) ? lds_pad(j) : j;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int a;
          if constexpr (PAD) a = wb + (SWW ? lds_pad(r * P) : r * P);
          else a = SWW ? lds_sw(j + r * P) : j + r * P;
          lds_st(z + a, x[b][r]);
        }
      }
    }
    wave_sync();
    wave_fft_s<T, N, S + 1, LAY>(z, w, inverse, lane);
  }
}

// Same stages, arithmetic and order as wave_fft, with register twiddles; the
// line holds n (LES_LAY_XOR) or lds_line_cap(n) (LES_LAY_PAD) elements.
template <typename T, int N, int LAY = LES_LAY_XOR>
__device__ __forceinline__ void wave_fft_t(Cplx<T>* z, const TwReg<T, N>& t, bool inverse, int lane) {
  wave_fft_s<T, N, 0, LAY>(z, t.w, inverse, lane);
}

// Block barrier for LDS exchange only: waits for the wavefront's LDS
// operations, not for its outstanding global loads (a prefetch stays in
// flight across it, unlike __syncthreads, whose workgroup fence can wait for
// them).  The memory clobber keeps the compiler from moving memory accesses
// across it.
__device__ __forceinline__ void block_sync_lds() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (block b on XCD b % 8), each with its own L2.  Tiles that share
// 128-byte lines (adjacent column tiles of narrower rows) must run on one XCD
// at nearly the same time, so physical block b = 8 m + x runs logical tile
// s_x + m, where XCD x owns the contiguous range [s_x, s_x + size_x) of the
// n tiles (sizes q + 1 for x < n % 8, q = n / 8 otherwise).  A bijection on
// [0, n).
#define LES_NXCD 8
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n / LES_NXCD, r = n % LES_NXCD;
  const int x = b % LES_NXCD, m = b / LES_NXCD;
  return x * q + (x < r ? x : r) + m;
}

// 3/2 padding / truncation of one transformed line in place (buffer of at
// least max(n1, n2) points): modes |k| < min(n1, n2)/2 are kept, the
// Nyquist mode of the shorter length is dropped, the rest is zero
// (src/compute_convective_padd.f90:288-319, 352-382).  For n2 = 3 n1 / 2 and
// n1 = 3 n2 / 2 the moved high modes [n1-h+1, n1) and their targets
// [n2-h+1, n2) never overlap, so they are copied directly; the zeroed band
// [h, n2-h] overlaps the sources and is written after the copies.
template <typename T>
__device__ __forceinline__ void wave_remap(Cplx<T>* z, int n1, int n2, int lane) {
  const int h = (n1 < n2 ? n1 : n2) / 2;
  for (int q = lane; q < h - 1; q += LES_WAVE) z[n2 - h + 1 + q] = z[n1 - h + 1 + q];
  wave_sync();
  for (int kk = h + lane; kk <= n2 - h; kk += LES_WAVE) z[kk] = Cplx<T>{(T)0, (T)0};
  wave_sync();
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

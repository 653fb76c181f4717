// Batched horizontal transform kernels (device code + launch helpers).
//
// Replaces the reference's FFTW/2DECOMP path (src/decomp_2d_fft_2d.f90:397-577
// r2c_1d_x / c2c_1d_y / c2r_1d_x, src/compute_convective_padd.f90:274-388
// transpose_to_big / dealias): the pencil transposes become strided column
// passes over a single device allocation.
//
// Layout: a real field is (nx, ny, planes) with x fastest; a spectral field is
// (lhx, ny, planes) complex, lhx = nx/2 + 1.  Row kernels take row pairs
// (j, j+1) of a plane as one complex line z = a + i b; column kernels take a
// block of CB adjacent kx columns of a plane and transform along y in LDS.
#pragma once
#include "les_common.h"

#define LES_BLOCK 256
#define LES_MAX_FIELDS 6

// ------------------------------------------------------------------ row r2c
// Load modes of the real input.
enum { LOAD_SCALE = 0, LOAD_SCALE_DIV = 1 };

template <typename TI, typename T>
struct RowR2CArgs {
  const TI* src[LES_MAX_FIELDS];
  Cplx<T>* dst[LES_MAX_FIELDS];
  int nx, ny, lhx;
  int k0, nplanes;  // planes [k0, k0+nplanes)
  int lpb;          // row pairs per block
  int mode;
  T scale, div;
};

template <typename TI, typename T>
__global__ void __launch_bounds__(LES_BLOCK) k_rows_r2c(RowR2CArgs<TI, T> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* bufA = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int nx = a.nx, lhx = a.lhx;
  Cplx<T>* bufB = bufA + a.lpb * nx;
  const int f = blockIdx.y;
  const TI* __restrict__ src = a.src[f];
  Cplx<T>* __restrict__ dst = a.dst[f];
  const int npairs = a.ny / 2;
  const long job0 = (long)blockIdx.x * a.lpb;
  const long njobs = (long)a.nplanes * npairs;
  int nl = a.lpb;
  if (job0 + nl > njobs) nl = (int)(njobs - job0);
  for (int idx = threadIdx.x; idx < nl * nx; idx += blockDim.x) {
    const int l = idx / nx, n = idx - l * nx;
    const long job = job0 + l;
    const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
    const long base = ((long)k * a.ny + 2 * jp) * nx + n;
    T v0 = (T)src[base], v1 = (T)src[base + nx];
    if (a.mode == LOAD_SCALE_DIV) {
      v0 = a.scale * (v0 / a.div);
      v1 = a.scale * (v1 / a.div);
    } else {
      v0 = a.scale * v0;
      v1 = a.scale * v1;
    }
    bufA[l * nx + n] = {v0, v1};
  }
  __syncthreads();
  Cplx<T>* Z = lds_fft(bufA, bufB, nl, nx, pl, tw, false);
  for (int idx = threadIdx.x; idx < nl * lhx; idx += blockDim.x) {
    const int l = idx / lhx, kk = idx - l * lhx;
    const long job = job0 + l;
    const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
    const Cplx<T> zk = Z[l * nx + kk], zm = Z[l * nx + (kk == 0 ? 0 : nx - kk)];
    const Cplx<T> A = {(T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y)};
    const Cplx<T> B = {(T)0.5 * (zk.y + zm.y), (T)0.5 * (zm.x - zk.x)};
    const long ob = ((long)k * a.ny + 2 * jp) * lhx + kk;
    dst[ob] = A;
    dst[ob + lhx] = B;
  }
}

// ------------------------------------------------------------------ row c2r
template <typename T, typename TO>
struct RowC2RArgs {
  const Cplx<T>* src[LES_MAX_FIELDS];
  TO* dst[LES_MAX_FIELDS];
  int nx, ny, lhx;
  int k0, nplanes;
  int lpb;
};

template <typename T, typename TO>
__global__ void __launch_bounds__(LES_BLOCK) k_rows_c2r(RowC2RArgs<T, TO> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* bufA = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int nx = a.nx, lhx = a.lhx, h = nx / 2;
  Cplx<T>* bufB = bufA + a.lpb * nx;
  const int f = blockIdx.y;
  const Cplx<T>* __restrict__ src = a.src[f];
  TO* __restrict__ dst = a.dst[f];
  const int npairs = a.ny / 2;
  const long job0 = (long)blockIdx.x * a.lpb;
  const long njobs = (long)a.nplanes * npairs;
  int nl = a.lpb;
  if (job0 + nl > njobs) nl = (int)(njobs - job0);
  for (int idx = threadIdx.x; idx < nl * lhx; idx += blockDim.x) {
    const int l = idx / lhx, kk = idx - l * lhx;
    const long job = job0 + l;
    const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
    const long ib = ((long)k * a.ny + 2 * jp) * lhx + kk;
    Cplx<T> A = src[ib], B = src[ib + lhx];
    if (kk == 0 || kk == h) {
      A.y = (T)0;
      B.y = (T)0;
    }
    bufA[l * nx + kk] = {A.x - B.y, A.y + B.x};
    if (kk != 0 && kk != h) bufA[l * nx + nx - kk] = {A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  Cplx<T>* Z = lds_fft(bufA, bufB, nl, nx, pl, tw, true);
  for (int idx = threadIdx.x; idx < nl * nx; idx += blockDim.x) {
    const int l = idx / nx, n = idx - l * nx;
    const long job = job0 + l;
    const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
    const long ob = ((long)k * a.ny + 2 * jp) * nx + n;
    const Cplx<T> z = Z[l * nx + n];
    dst[ob] = (TO)z.x;
    dst[ob + nx] = (TO)z.y;
  }
}

// ------------------------------------------------------------------ columns
// Spectral-space operations applied between the forward and inverse y passes.
enum { COP_GRAD = 0, COP_DIV = 1, COP_FILT2 = 2, COP_FILT1 = 3, COP_COPY = 4, COP_PGRAD = 5 };

template <typename T>
struct ColArgs {
  const Cplx<T>* in[2];
  Cplx<T>* out[3];
  const T* kkx;
  const T* kky;
  const T* g0;
  const T* g1;
  int lhx, ny;
  int k0, nplanes;
  int cb;        // columns per block
  int ncb;       // column blocks per plane
  int nin, nout;
  int op;
  int fwd, inv;  // apply forward y transform to inputs / inverse to outputs
};

template <typename T>
__device__ __forceinline__ Cplx<T> col_op(const ColArgs<T>& a, int o, int kx, int j, Cplx<T> x0, Cplx<T> x1) {
  const long q = (long)j * a.lhx + kx;
  const bool nyq = (kx == a.lhx - 1) || (j == a.ny / 2);
  switch (a.op) {
    case COP_GRAD: {
      Cplx<T> v = nyq ? Cplx<T>{(T)0, (T)0} : x0;
      if (o == 0) return c_ik(a.kkx[q], v);
      if (o == 1) return c_ik(a.kky[q], v);
      return v;
    }
    case COP_DIV: {
      Cplx<T> v0 = nyq ? Cplx<T>{(T)0, (T)0} : x0;
      Cplx<T> v1 = nyq ? Cplx<T>{(T)0, (T)0} : x1;
      return c_add(c_ik(a.kkx[q], v0), c_ik(a.kky[q], v1));
    }
    case COP_FILT2:
      return c_scale(x0, o == 0 ? a.g0[q] : a.g1[q]);
    case COP_FILT1:
      return c_scale(x0, a.g0[q]);
    case COP_PGRAD:
      if (o == 0) return c_ik(a.kkx[q], x0);
      if (o == 1) return c_ik(a.kky[q], x0);
      return x0;
    default:
      return x0;
  }
}

template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) k_cols(ColArgs<T> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  const int ny = a.ny, cb = a.cb;
  const int L = cb * ny;
  Cplx<T>* X0 = reinterpret_cast<Cplx<T>*>(smem_raw);
  Cplx<T>* X1 = X0 + L;
  Cplx<T>* Y = X0 + 2 * L;
  Cplx<T>* W = X0 + 3 * L;
  const int k = a.k0 + blockIdx.x / a.ncb;
  const int c0 = (blockIdx.x % a.ncb) * cb;
  int nc = a.lhx - c0;
  if (nc > cb) nc = cb;
  const long pbase = (long)k * ny * a.lhx;
  // load inputs transposed: X[c * ny + j]
  for (int t = 0; t < a.nin; ++t) {
    const Cplx<T>* __restrict__ src = a.in[t] + pbase;
    Cplx<T>* X = t == 0 ? X0 : X1;
    for (int idx = threadIdx.x; idx < nc * ny; idx += blockDim.x) {
      const int j = idx / nc, c = idx - j * nc;
      X[c * ny + j] = src[(long)j * a.lhx + c0 + c];
    }
  }
  __syncthreads();
  Cplx<T>* R0 = X0;
  Cplx<T>* R1 = X1;
  if (a.fwd) {
    // forward transforms of the inputs; Y and W serve as scratch
    R0 = lds_fft(X0, Y, nc, ny, pl, tw, false);
    if (R0 == Y) {  // keep inputs out of the output buffers
      for (int idx = threadIdx.x; idx < nc * ny; idx += blockDim.x) X0[idx] = Y[idx];
      __syncthreads();
      R0 = X0;
    }
    if (a.nin > 1) {
      R1 = lds_fft(X1, W, nc, ny, pl, tw, false);
      if (R1 == W) {
        for (int idx = threadIdx.x; idx < nc * ny; idx += blockDim.x) X1[idx] = W[idx];
        __syncthreads();
        R1 = X1;
      }
    }
  }
  for (int o = 0; o < a.nout; ++o) {
    for (int idx = threadIdx.x; idx < nc * ny; idx += blockDim.x) {
      const int c = idx / ny, j = idx - c * ny;
      Y[idx] = col_op(a, o, c0 + c, j, R0[idx], a.nin > 1 ? R1[idx] : R0[idx]);
    }
    __syncthreads();
    Cplx<T>* Rs = a.inv ? lds_fft(Y, W, nc, ny, pl, tw, true) : Y;
    Cplx<T>* __restrict__ dst = a.out[o] + pbase;
    for (int idx = threadIdx.x; idx < nc * ny; idx += blockDim.x) {
      const int j = idx / nc, c = idx - j * nc;
      dst[(long)j * a.lhx + c0 + c] = Rs[c * ny + j];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ 3/2 resampling
// Padding keeps modes |k| < n/2 (Nyquist dropped) of a length-n1 line in a
// length-n2 spectrum; truncation is the mirror image
// (src/compute_convective_padd.f90:288-319, 352-382).
template <typename T>
__device__ __forceinline__ void remap_modes(const Cplx<T>* z, int n1, Cplx<T>* zb, int n2, int nlines, int s1, int s2) {
  const int h = (n1 < n2 ? n1 : n2) / 2;
  for (int idx = threadIdx.x; idx < nlines * n2; idx += blockDim.x) {
    const int l = idx / n2, kk = idx - l * n2;
    Cplx<T> v = {(T)0, (T)0};
    if (kk < h) v = z[l * s1 + kk];
    else if (kk > n2 - h) v = z[l * s1 + n1 - (n2 - kk)];
    zb[l * s2 + kk] = v;
  }
}

// Convection source planes padded in x: which = 0..5 -> -c u, -c v, -c w,
// c (dw/dy - dv/dz), c (du/dz - dw/dx), c (dv/dx - du/dy), with the wall-law
// first level of the vorticities (src/compute_convective_padd.f90:133-169).
template <typename TI>
struct ConvSrc {
  const double *u, *v, *w;
  const TI *dudy, *dudz, *dvdx, *dvdz, *dwdx, *dwdy;
};

template <typename T, typename TI>
__device__ __forceinline__ T conv_src(const ConvSrc<TI>& s, int which, int k, long q, long np, T c) {
  const long a = (long)k * np + q;
  switch (which) {
    case 0: return (T)(-((double)c * s.u[a]));
    case 1: return (T)(-((double)c * s.v[a]));
    case 2: return (T)(-((double)c * s.w[a]));
    case 3:
      if (k == 1) return c * ((T)0.5 * ((T)s.dwdy[a] + (T)s.dwdy[a + np]) - (T)s.dvdz[a]);
      return c * ((T)s.dwdy[a] - (T)s.dvdz[a]);
    case 4:
      if (k == 1) return c * ((T)s.dudz[a] - (T)0.5 * ((T)s.dwdx[a] + (T)s.dwdx[a + np]));
      return c * ((T)s.dudz[a] - (T)s.dwdx[a]);
    default: return c * ((T)s.dvdx[a] - (T)s.dudy[a]);
  }
}

// Row resampling: real rows of length n1 -> real rows of length n2.
// src_mode 0: plain real array (nx = n1) src[f]; src_mode 1: convection
// sources (field = blockIdx.y) computed from ConvSrc.
template <typename T, typename TI, typename TO>
struct RowResArgs {
  const T* src[LES_MAX_FIELDS];
  TO* dst[LES_MAX_FIELDS];
  ConvSrc<TI> cs;
  T cconv;
  int src_mode;
  int n1, n2, ny;
  int k0, nplanes;
  long src_plane, dst_plane;  // plane strides (elements)
  int lpb;
};

template <typename T, typename TI, typename TO>
__global__ void __launch_bounds__(LES_BLOCK) k_rows_resample(RowResArgs<T, TI, TO> a, FftPlan p1, FftPlan p2,
                                                             const Cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  const int n1 = a.n1, n2 = a.n2;
  const int nm = n1 > n2 ? n1 : n2;
  Cplx<T>* bufA = reinterpret_cast<Cplx<T>*>(smem_raw);
  Cplx<T>* bufB = bufA + a.lpb * nm;
  const int f = blockIdx.y;
  const int npairs = a.ny / 2;
  const long job0 = (long)blockIdx.x * a.lpb;
  const long njobs = (long)a.nplanes * npairs;
  int nl = a.lpb;
  if (job0 + nl > njobs) nl = (int)(njobs - job0);
  const long np = (long)n1 * a.ny;
  for (int idx = threadIdx.x; idx < nl * n1; idx += blockDim.x) {
    const int l = idx / n1, n = idx - l * n1;
    const long job = job0 + l;
    const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
    T v0, v1;
    if (a.src_mode == 1) {
      const long q = (long)(2 * jp) * n1 + n;
      v0 = conv_src<T, TI>(a.cs, f, k, q, np, a.cconv);
      v1 = conv_src<T, TI>(a.cs, f, k, q + n1, np, a.cconv);
    } else {
      const long b = (long)k * a.src_plane + (long)(2 * jp) * n1 + n;
      v0 = a.src[f][b];
      v1 = a.src[f][b + n1];
    }
    bufA[l * nm + n] = {v0, v1};
  }
  __syncthreads();
  Cplx<T>* Z = lds_fft(bufA, bufB, nl, nm, p1, tw, false);
  Cplx<T>* Zb = (Z == bufA) ? bufB : bufA;
  remap_modes(Z, n1, Zb, n2, nl, nm, nm);
  __syncthreads();
  Cplx<T>* R = lds_fft(Zb, Z, nl, nm, p2, tw, true);
  for (int idx = threadIdx.x; idx < nl * n2; idx += blockDim.x) {
    const int l = idx / n2, n = idx - l * n2;
    const long job = job0 + l;
    const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
    const long ob = (long)k * a.dst_plane + (long)(2 * jp) * n2 + n;
    const Cplx<T> z = R[l * nm + n];
    a.dst[f][ob] = (TO)z.x;
    a.dst[f][ob + n2] = (TO)z.y;
  }
}

// Column resampling: columns (i, i+1) of a real (W, h1) plane -> (W, h2).
template <typename T>
struct ColResArgs {
  const T* src[LES_MAX_FIELDS];
  T* dst[LES_MAX_FIELDS];
  int W, h1, h2;
  int k0, nplanes;
  long src_plane, dst_plane;
  int cp;   // column pairs per block
  int ncp;  // column-pair blocks per plane
};

template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) k_cols_resample(ColResArgs<T> a, FftPlan p1, FftPlan p2,
                                                             const Cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  const int h1 = a.h1, h2 = a.h2;
  const int nm = h1 > h2 ? h1 : h2;
  Cplx<T>* bufA = reinterpret_cast<Cplx<T>*>(smem_raw);
  Cplx<T>* bufB = bufA + a.cp * nm;
  const int f = blockIdx.y;
  const int k = a.k0 + blockIdx.x / a.ncp;
  const int p0 = (blockIdx.x % a.ncp) * a.cp;  // first column pair
  int nl = a.W / 2 - p0;
  if (nl > a.cp) nl = a.cp;
  const T* __restrict__ src = a.src[f] + (long)k * a.src_plane;
  for (int idx = threadIdx.x; idx < nl * 2 * h1; idx += blockDim.x) {
    const int j = idx / (2 * nl), r = idx - j * (2 * nl);
    const int l = r >> 1, comp = r & 1;
    const T v = src[(long)j * a.W + 2 * (p0 + l) + comp];
    if (comp == 0) bufA[l * nm + j].x = v;
    else bufA[l * nm + j].y = v;
  }
  __syncthreads();
  Cplx<T>* Z = lds_fft(bufA, bufB, nl, nm, p1, tw, false);
  Cplx<T>* Zb = (Z == bufA) ? bufB : bufA;
  remap_modes(Z, h1, Zb, h2, nl, nm, nm);
  __syncthreads();
  Cplx<T>* R = lds_fft(Zb, Z, nl, nm, p2, tw, true);
  T* __restrict__ dst = a.dst[f] + (long)k * a.dst_plane;
  for (int idx = threadIdx.x; idx < nl * 2 * h2; idx += blockDim.x) {
    const int j = idx / (2 * nl), r = idx - j * (2 * nl);
    const int l = r >> 1, comp = r & 1;
    const Cplx<T> z = R[l * nm + j];
    dst[(long)j * a.W + 2 * (p0 + l) + comp] = comp == 0 ? z.x : z.y;
  }
}

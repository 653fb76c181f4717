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
#define LES_COL_JOBS 3

// ------------------------------------------------------------------ row r2c
// Load modes of the real input.
// LOAD_AB2_SCALE_DIV: the AB2 predictor u + dt (t1 r + t2 r_f) of each
// value (the arithmetic of k_ab2, src/LES_engine.f90:253-283), then
// LOAD_SCALE_DIV: the pressure right-hand side straight from u, rhs, rhs_f.
// LOAD_PCLEAN (the PCL instance only): the TKE budget's cleaned pressure
// p - (u u + v v + W W) / 2, W = (w(k) + w(k+1)) / 2, from p (src) and u, v, w
// (ab_r), 0 on the planes 0 and nz+1.
enum { LOAD_SCALE = 0, LOAD_SCALE_DIV = 1, LOAD_AB2_SCALE_DIV = 2, LOAD_PCLEAN = 3 };

template <typename TI, typename T>
struct RowR2CArgs {
  const TI* src[LES_MAX_FIELDS];
  Cplx<T>* dst[LES_MAX_FIELDS];
  int nx, ny, lhx;
  int k0, nplanes;  // planes [k0, k0+nplanes)
  int mode;
  T scale, div;
  int ls;           // LDS line stride per wavefront (row_line<T>(nx))
  const TI* ab_r[3];  // LOAD_AB2_SCALE_DIV: rhs and rhs_f of field f
  const TI* ab_f[3];
  T ab_dt, ab_t1, ab_t2;
  int pc_nz;          // LOAD_PCLEAN: nz (planes 0 and nz+1 are ghosts)
  // k_rows_r2c_nyq: the FP64 Nyquist sums of the raw input (row sums A,
  // column partials P of each band of 8 rows)
  double *nyq_a, *nyq_p;
};

// LDS layout of the row transforms: the register-twiddle 256-point
// transform in FP32 uses the padded placement (these kernels are bound by
// integer address work), everything else the XOR placement on nx points.
template <typename T>
__host__ __device__ constexpr int row_layout(int n) {
  return (sizeof(T) == 4 && n == 256) ? LES_LAY_PAD : LES_LAY_XOR;
}
template <typename T>
__host__ __device__ constexpr int row_line(int n) {
  return row_layout<T>(n) == LES_LAY_PAD ? lds_line_cap(n) : n;
}

// One wavefront per row pair: load, transform and unpack without any
// block-level barrier.  LDS: one line of a.ls points per wavefront.
template <bool NYQ = false, bool PCL = false, typename TI, typename T, typename FFT>
__device__ __forceinline__ void rows_r2c_body(const RowR2CArgs<TI, T>& a, Cplx<T>* z, int lane, long job, FFT&& fft,
                                              double* part = nullptr) {
  const int nx = a.nx, lhx = a.lhx;
  const int npairs = a.ny / 2;
  const int f = blockIdx.y;
  const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
  const TI* __restrict__ src = a.src[f] + ((long)k * a.ny + 2 * jp) * nx;
  double s0 = 0.0, s1 = 0.0;
  for (int n = lane; n < nx; n += LES_WAVE) {
    if constexpr (NYQ) {
      // alternating row sums and the pair's column partial (-1)^j f, in FP64
      const double r0 = src[n], r1 = src[n + nx];
      s0 += (n & 1) ? -r0 : r0;
      s1 += (n & 1) ? -r1 : r1;
      part[n] = r0 - r1;
    }
    T v0 = (T)src[n], v1 = (T)src[n + nx];
    if (a.mode == LOAD_AB2_SCALE_DIV) {
      const long o = ((long)k * a.ny + 2 * jp) * nx + n;
      const TI* __restrict__ r = a.ab_r[f];
      const TI* __restrict__ rf = a.ab_f[f];
      v0 = v0 + a.ab_dt * (a.ab_t1 * (T)r[o] + a.ab_t2 * (T)rf[o]);
      v1 = v1 + a.ab_dt * (a.ab_t1 * (T)r[o + nx] + a.ab_t2 * (T)rf[o + nx]);
    }
    if constexpr (PCL) {
      if (k == 0 || k == a.pc_nz + 1) {
        v0 = (T)0;
        v1 = (T)0;
      } else {
        const long o = ((long)k * a.ny + 2 * jp) * nx + n, pl = (long)a.ny * nx;
        const TI* __restrict__ u = a.ab_r[0];
        const TI* __restrict__ v = a.ab_r[1];
        const TI* __restrict__ w = a.ab_r[2];
        const T h0 = (T)0.5 * ((T)w[o] + (T)w[o + pl]);
        const T h1 = (T)0.5 * ((T)w[o + nx] + (T)w[o + nx + pl]);
        v0 = v0 - (T)0.5 * (((T)u[o] * (T)u[o] + (T)v[o] * (T)v[o]) + h0 * h0);
        v1 = v1 - (T)0.5 * (((T)u[o + nx] * (T)u[o + nx] + (T)v[o + nx] * (T)v[o + nx]) + h1 * h1);
      }
    }
    if (a.mode == LOAD_SCALE_DIV || a.mode == LOAD_AB2_SCALE_DIV) {
      v0 = a.scale * (v0 / a.div);
      v1 = a.scale * (v1 / a.div);
    } else {
      v0 = a.scale * v0;
      v1 = a.scale * v1;
    }
    z[n] = {v0, v1};
  }
  if constexpr (NYQ) {
    for (int o = LES_WAVE / 2; o > 0; o >>= 1) {
      s0 += __shfl_xor(s0, o, LES_WAVE);
      s1 += __shfl_xor(s1, o, LES_WAVE);
    }
    if (lane == 0) {
      double* A = a.nyq_a + (long)k * a.ny + 2 * jp;
      A[0] = s0 / (double)nx;
      A[1] = s1 / (double)nx;
    }
  }
  wave_sync();
  fft(z, false);
  Cplx<T>* __restrict__ dst = a.dst[f] + ((long)k * a.ny + 2 * jp) * lhx;
  for (int kk = lane; kk < lhx; kk += LES_WAVE) {
    const Cplx<T> zk = z[kk], zm = z[kk == 0 ? 0 : nx - kk];
    const Cplx<T> A = {(T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y)};
    const Cplx<T> B = {(T)0.5 * (zk.y + zm.y), (T)0.5 * (zm.x - zk.x)};
    dst[kk] = A;
    dst[kk + lhx] = B;
  }
}

// FP32: held to 64 VGPRs (8 waves per SIMD instead of 7; two VGPRs spill).
template <typename TI, typename T, bool PCL = false>
__global__ void __launch_bounds__(LES_BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 1)))
k_rows_r2c(RowR2CArgs<TI, T> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  Cplx<T>* z = reinterpret_cast<Cplx<T>*>(smem_raw) + wave * a.ls;
  const long job = (long)blockIdx.x * (blockDim.x / LES_WAVE) + wave;
  if (job >= (long)a.nplanes * (a.ny / 2)) return;
  if (pl.n == 256) {
    TwReg<T, 256> tr;
    tw_load(tr, tw + pl.twf, lane);
    rows_r2c_body<false, PCL>(a, z, lane, job, [&](Cplx<T>* p, bool inv) { wave_fft_t<T, 256, row_layout<T>(256)>(p, tr, inv, lane); });
  } else {
    rows_r2c_body<false, PCL>(a, z, lane, job, [&](Cplx<T>* p, bool inv) { wave_fft(p, pl, tw, inv, lane); });
  }
}

// k_rows_r2c with the first pass of the mixed mode's FP64 Nyquist removal
// folded into the load (the sums k_nyq_reduce forms, from the values this
// kernel reads anyway): A_j = sum_i (-1)^i f_ij / nx of both rows (the same
// lane-strided sums and butterfly as k_nyq_reduce), and per block of 4 row
// pairs the band's column partial sum_j (-1)^j f_ij, the 4 pairs added in
// order after a block barrier.  The host launches it only when ny % 8 == 0,
// so every block holds 4 row pairs of one plane and no wavefront leaves
// early.  LDS: the transform lines, then 4 nx doubles of pair partials.
// The FP64 sums beside the FP32 transform's registers need 84 VGPRs (5
// waves per SIMD); held to 64 the kernel spills.
#ifndef LES_NYQ_WAVES
#define LES_NYQ_WAVES 5
#endif
template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) __attribute__((amdgpu_waves_per_eu(LES_NYQ_WAVES)))
k_rows_r2c_nyq(RowR2CArgs<double, T> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  Cplx<T>* z = reinterpret_cast<Cplx<T>*>(smem_raw) + wave * a.ls;
  double* part = reinterpret_cast<double*>(reinterpret_cast<Cplx<T>*>(smem_raw) + (LES_BLOCK / LES_WAVE) * a.ls);
  const long job = (long)blockIdx.x * (blockDim.x / LES_WAVE) + wave;
  if (pl.n == 256) {
    TwReg<T, 256> tr;
    tw_load(tr, tw + pl.twf, lane);
    rows_r2c_body<true>(a, z, lane, job, [&](Cplx<T>* p, bool inv) { wave_fft_t<T, 256, row_layout<T>(256)>(p, tr, inv, lane); },
                        part + wave * a.nx);
  } else {
    rows_r2c_body<true>(a, z, lane, job, [&](Cplx<T>* p, bool inv) { wave_fft(p, pl, tw, inv, lane); }, part + wave * a.nx);
  }
  __syncthreads();
  const int nx = a.nx;
  double* __restrict__ P = a.nyq_p + (long)blockIdx.x * nx;
  for (int i = threadIdx.x; i < nx; i += LES_BLOCK)
    P[i] = ((part[i] + part[nx + i]) + part[2 * nx + i]) + part[3 * nx + i];
}

// ------------------------------------------------------------------ row c2r
template <typename T, typename TO>
struct RowC2RArgs {
  const Cplx<T>* src[LES_MAX_FIELDS];
  TO* dst[LES_MAX_FIELDS];
  int nx, ny, lhx;
  int k0, nplanes;
  int ls;  // LDS line stride per wavefront (row_line<T>(nx))
};

// Load the half spectra of rows (2jp, 2jp+1) of plane k as one complex line
// and run the inverse transform: z then holds row 2jp in .x, row 2jp+1 in .y.
template <typename T, typename TO, typename FFT>
__device__ __forceinline__ void rows_c2r_line(const RowC2RArgs<T, TO>& a, Cplx<T>* z, int lane, int k, int jp, FFT&& fft) {
  const int nx = a.nx, lhx = a.lhx, h = nx / 2;
  const int f = blockIdx.y;
  const Cplx<T>* __restrict__ src = a.src[f] + ((long)k * a.ny + 2 * jp) * lhx;
  for (int kk = lane; kk < lhx; kk += LES_WAVE) {
    Cplx<T> A = src[kk], B = src[kk + lhx];
    if (kk == 0 || kk == h) {
      A.y = (T)0;
      B.y = (T)0;
    }
    z[kk] = {A.x - B.y, A.y + B.x};
    if (kk != 0 && kk != h) z[nx - kk] = {A.x + B.y, B.x - A.y};
  }
  wave_sync();
  fft(z, true);
}

template <typename T, typename TO, typename FFT>
__device__ __forceinline__ void rows_c2r_body(const RowC2RArgs<T, TO>& a, Cplx<T>* z, int lane, long job, FFT&& fft) {
  const int nx = a.nx;
  const int npairs = a.ny / 2;
  const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
  rows_c2r_line(a, z, lane, k, jp, fft);
  TO* __restrict__ dst = a.dst[blockIdx.y] + ((long)k * a.ny + 2 * jp) * nx;
  for (int n = lane; n < nx; n += LES_WAVE) {
    const Cplx<T> v = z[n];
    dst[n] = (TO)v.x;
    dst[n + nx] = (TO)v.y;
  }
}

// FP32: held to 64 VGPRs, as the r2c kernel.
template <typename T, typename TO>
__global__ void __launch_bounds__(LES_BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 1)))
k_rows_c2r(RowC2RArgs<T, TO> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  Cplx<T>* z = reinterpret_cast<Cplx<T>*>(smem_raw) + wave * a.ls;
  const long job = (long)blockIdx.x * (blockDim.x / LES_WAVE) + wave;
  if (job >= (long)a.nplanes * (a.ny / 2)) return;
  if (pl.n == 256) {
    TwReg<T, 256> tr;
    tw_load(tr, tw + pl.twf, lane);
    rows_c2r_body(a, z, lane, job, [&](Cplx<T>* p, bool inv) { wave_fft_t<T, 256, row_layout<T>(256)>(p, tr, inv, lane); });
  } else {
    rows_c2r_body(a, z, lane, job, [&](Cplx<T>* p, bool inv) { wave_fft(p, pl, tw, inv, lane); });
  }
}

// ------------------------------------------------------------------ columns
// Spectral-space operations applied between the forward and inverse y passes.
enum { COP_GRAD = 0, COP_DIV = 1, COP_FILT2 = 2, COP_FILT1 = 3, COP_COPY = 4, COP_PGRAD = 5 };

template <typename T>
struct ColArgs {
  // per job (blockIdx.y of the register-resident kernel; job 0 elsewhere):
  // the same op on up to LES_COL_JOBS independent input/output sets
  const Cplx<T>* in[LES_COL_JOBS][2];
  Cplx<T>* out[LES_COL_JOBS][3];
  const T* kkx;
  const T* kky;
  const T* g0;
  const T* g1;
  int lhx, ny;
  int k0, nplanes;
  int ls;        // LDS line stride (at least ny, chosen so the tile transposes hit distinct banks)
  int ncb;       // column blocks per plane
  int nin, nout;
  int op;
  int fwd, inv;  // apply forward y transform to inputs / inverse to outputs
  int ntiles;    // ncb * nplanes
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

template <typename T, int CB, typename FFT>
__device__ __forceinline__ void cols_body(const ColArgs<T>& a, unsigned char* smem, FFT&& fft) {
  const int ny = a.ny, LS = a.ls;
  Cplx<T>* X0 = reinterpret_cast<Cplx<T>*>(smem);
  Cplx<T>* X1 = X0 + CB * LS;
  Cplx<T>* Y = X0 + (a.nin > 1 ? 2 : 1) * CB * LS;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int k = a.k0 + tile / a.ncb;
  const int c0 = (tile % a.ncb) * CB;
  int nc = a.lhx - c0;
  if (nc > CB) nc = CB;
  const long pbase = (long)k * ny * a.lhx + c0;
  for (int t = 0; t < a.nin; ++t) {
    const Cplx<T>* __restrict__ src = a.in[0][t] + pbase;
    Cplx<T>* X = t == 0 ? X0 : X1;
    for (int idx = threadIdx.x; idx < CB * ny; idx += blockDim.x) {
      const int j = idx / CB, c = idx % CB;
      if (c < nc) X[c * LS + j] = src[(long)j * a.lhx + c];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE, nw = blockDim.x / LES_WAVE;
  if (a.fwd) {
    for (int c = wave; c < nc; c += nw) {
      fft(X0 + c * LS, false);
      if (a.nin > 1) fft(X1 + c * LS, false);
    }
  }
  for (int o = 0; o < a.nout; ++o) {
    for (int c = wave; c < nc; c += nw) {
      const Cplx<T>* x0 = X0 + c * LS;
      const Cplx<T>* x1 = (a.nin > 1 ? X1 : X0) + c * LS;
      Cplx<T>* y = Y + c * LS;
      for (int j = lane; j < ny; j += LES_WAVE) y[j] = col_op(a, o, c0 + c, j, x0[j], x1[j]);
      wave_sync();
      if (a.inv) fft(y, true);
    }
    __syncthreads();
    Cplx<T>* __restrict__ dst = a.out[0][o] + pbase;
    for (int idx = threadIdx.x; idx < CB * ny; idx += blockDim.x) {
      const int j = idx / CB, c = idx % CB;
      if (c < nc) dst[(long)j * a.lhx + c] = Y[c * LS + j];
    }
    __syncthreads();
  }
}

// Persistent, software-pipelined version for ny = 256 (tiles of CB columns,
// E = CB * 256 / LES_BLOCK elements per thread and input): the block walks a
// contiguous run of tiles of its XCD (stride = blocks of that XCD, so the
// concurrently running blocks of one XCD take adjacent tiles) and loads the
// next tile into registers while it transforms the current one.
template <typename T, int CB, typename FFT>
__device__ __forceinline__ void cols_body_pipe(const ColArgs<T>& a, unsigned char* smem, FFT&& fft) {
  constexpr int NY = 256, E = CB * NY / LES_BLOCK;
  static_assert(E * LES_BLOCK == CB * NY, "tile must split evenly over the block");
  const int LS = a.ls;
  Cplx<T>* X0 = reinterpret_cast<Cplx<T>*>(smem);
  Cplx<T>* X1 = X0 + CB * LS;
  Cplx<T>* Y = X0 + (a.nin > 1 ? 2 : 1) * CB * LS;
  const int x = blockIdx.x % LES_NXCD, m = blockIdx.x / LES_NXCD, nb = gridDim.x / LES_NXCD;
  const int q = a.ntiles / LES_NXCD, r = a.ntiles % LES_NXCD;
  const int tbeg = x * q + (x < r ? x : r), tend = tbeg + q + (x < r ? 1 : 0);
  int t = tbeg + m;
  if (t >= tend) return;
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE, nw = blockDim.x / LES_WAVE;
  Cplx<T> pre[2][E];
  auto fetch = [&](int tt) {
    const int k = a.k0 + tt / a.ncb, c0 = (tt % a.ncb) * CB;
    int nc = a.lhx - c0;
    if (nc > CB) nc = CB;
    const long pb = (long)k * NY * a.lhx + c0;
#pragma unroll
    for (int in = 0; in < 2; ++in) {
      if (in < a.nin) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int idx = threadIdx.x + e * LES_BLOCK, j = idx / CB, c = idx % CB;
          pre[in][e] = c < nc ? a.in[0][in][pb + (long)j * a.lhx + c] : Cplx<T>{(T)0, (T)0};
        }
      }
    }
  };
  fetch(t);
  while (true) {
    const int k = a.k0 + t / a.ncb, c0 = (t % a.ncb) * CB;
    int nc = a.lhx - c0;
    if (nc > CB) nc = CB;
    const long pbase = (long)k * NY * a.lhx + c0;
#pragma unroll
    for (int in = 0; in < 2; ++in) {
      if (in < a.nin) {
        Cplx<T>* X = in == 0 ? X0 : X1;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int idx = threadIdx.x + e * LES_BLOCK, j = idx / CB, c = idx % CB;
          X[c * LS + j] = pre[in][e];
        }
      }
    }
    block_sync_lds();
    const int tn = t + nb;
    if (tn < tend) fetch(tn);  // in flight during this tile's transforms
    if (a.fwd) {
      for (int c = wave; c < nc; c += nw) {
        fft(X0 + c * LS, false);
        if (a.nin > 1) fft(X1 + c * LS, false);
      }
    }
    for (int o = 0; o < a.nout; ++o) {
      for (int c = wave; c < nc; c += nw) {
        const Cplx<T>* x0 = X0 + c * LS;
        const Cplx<T>* x1 = (a.nin > 1 ? X1 : X0) + c * LS;
        Cplx<T>* y = Y + c * LS;
        for (int j = lane; j < NY; j += LES_WAVE) y[j] = col_op(a, o, c0 + c, j, x0[j], x1[j]);
        wave_sync();
        if (a.inv) fft(y, true);
      }
      block_sync_lds();
      Cplx<T>* __restrict__ dst = a.out[0][o] + pbase;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int idx = threadIdx.x + e * LES_BLOCK, j = idx / CB, c = idx % CB;
        if (c < nc) dst[(long)j * a.lhx + c] = Y[c * LS + j];
      }
      block_sync_lds();
    }
    if (tn >= tend) break;
    t = tn;
  }
}

// A tile of CB adjacent kx columns of one plane: loaded row by row (CB
// contiguous complex values per row), transposed into LDS so every column is
// a contiguous line, transformed along y by one wavefront per column,
// combined by the spectral op and transformed back, then stored row by row.
// Length 256 runs on register twiddles and the pipelined persistent body.
template <typename T, int CB>
__global__ void __launch_bounds__(LES_BLOCK) k_cols(ColArgs<T> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1);
  cols_body<T, CB>(a, smem_raw, [&](Cplx<T>* p, bool inv) { wave_fft(p, pl, tw, inv, lane); });
}

// ny = 256: a kernel of its own, so its register budget is not set by the
// runtime-plan path (whose radix-2 stages need twice the butterflies).
template <typename T, int CB>
__global__ void __launch_bounds__(LES_BLOCK) k_cols_256(ColArgs<T> a, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1);
  TwReg<T, 256> tr;
  tw_load(tr, tw + pl.twf, lane);
  cols_body_pipe<T, CB>(a, smem_raw, [&](Cplx<T>* p, bool inv) { wave_fft_t(p, tr, inv, lane); });
}

// ------------------------------------------------------------------ columns, ny = 256, in registers
// The 256-point column transform as two register-resident halves.  The four
// radix-4 Stockham stages (engine above) fall into 16 closed groups twice:
// stages 0-1 map inputs a + 16 t (t < 16) of thread a to its outputs 16 a + c
// (c < 16), and stages 2-3 map inputs k + 16 t of thread k to outputs
// k + 16 t.  So a thread loads its column's rows a + 16 t straight from
// global memory, runs stages 0-1 in registers, trades through LDS (16 x 16
// transpose per column, row e at 17 (e >> 4) + (e & 15): conflict-free), runs
// stages 2-3 in registers and holds spectral rows a + 16 t, which is again
// the input layout of stages 0-1, so the spectral op and the inverse
// transform continue in registers and the result is stored from them.  The
// butterflies, twiddles and their order are those of the Stockham stages.
// Block: 16 columns (128-byte row segments) x 16 row classes; one LDS
// exchange per transform instead of one per stage plus two tile transposes.
#define LES_C16_LS 273  // column stride in the exchange buffer (= 1 mod 16)
static_assert(cp_OFF(256, 1) == 3 && cp_OFF(256, 2) == 15 && cp_OFF(256, 3) == 63, "256-point twiddle layout");

// Twiddles are read from the plan's table (4 KB at most, cache-resident)
// where they are used: stages 0-1 read the same entries in every thread,
// stages 2-3 the thread's own.
template <typename T>
__device__ __forceinline__ Cplx<T> tw16(const Cplx<T>* __restrict__ twf, int idx, bool inverse) {
  Cplx<T> w = twf[idx];
  if (inverse) w.y = -w.y;
  return w;
}

// stages 0-1: X[t] = z[a + 16 t] -> X[c] = z'[16 a + c]
template <typename T>
__device__ __forceinline__ void c16_pair_a(Cplx<T>* X, const Cplx<T>* __restrict__ twf, bool inverse) {
  const int sgn = inverse ? 1 : -1;
  Cplx<T> Y[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    Cplx<T> x[4] = {X[m], X[m + 4], X[m + 8], X[m + 12]};
#pragma unroll
    for (int r = 1; r < 4; ++r) x[r] = c_mul(x[r], tw16(twf, r - 1, inverse));
    bfly(x, 4, sgn);
#pragma unroll
    for (int r = 0; r < 4; ++r) Y[m][r] = x[r];
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    Cplx<T> x[4] = {Y[0][b], Y[1][b], Y[2][b], Y[3][b]};
#pragma unroll
    for (int r = 1; r < 4; ++r) x[r] = c_mul(x[r], tw16(twf, 3 + (r - 1) * 4 + b, inverse));
    bfly(x, 4, sgn);
#pragma unroll
    for (int r = 0; r < 4; ++r) X[b + 4 * r] = x[r];
  }
}

// stages 2-3: X[t] = z'[k + 16 t] -> X[t] = Z[k + 16 t]
template <typename T>
__device__ __forceinline__ void c16_pair_b(Cplx<T>* X, const Cplx<T>* __restrict__ twf, int k, bool inverse) {
  const int sgn = inverse ? 1 : -1;
  Cplx<T> V[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    Cplx<T> x[4] = {X[q], X[q + 4], X[q + 8], X[q + 12]};
#pragma unroll
    for (int r = 1; r < 4; ++r) x[r] = c_mul(x[r], tw16(twf, 15 + (r - 1) * 16 + k, inverse));
    bfly(x, 4, sgn);
#pragma unroll
    for (int r = 0; r < 4; ++r) V[q][r] = x[r];
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Cplx<T> x[4] = {V[0][r], V[1][r], V[2][r], V[3][r]};
#pragma unroll
    for (int rp = 1; rp < 4; ++rp) x[rp] = c_mul(x[rp], tw16(twf, 63 + (rp - 1) * 64 + k + 16 * r, inverse));
    bfly(x, 4, sgn);
#pragma unroll
    for (int r2 = 0; r2 < 4; ++r2) X[r + 4 * r2] = x[r2];
  }
}

// Whole transform of thread (c, a)'s 16 values through the exchange buffer
// (ex: this block's buffer; the caller's barrier protects earlier readers).
template <typename T>
__device__ __forceinline__ void c16_fft(Cplx<T>* X, const Cplx<T>* __restrict__ twf, bool inverse, Cplx<T>* ex, int c,
                                        int a) {
  c16_pair_a(X, twf, inverse);
  Cplx<T>* col = ex + c * LES_C16_LS;
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) lds_st(col + 17 * a + cc, X[cc]);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 16; ++t) X[t] = lds_ld(col + a + 17 * t);
  __syncthreads();
  c16_pair_b(X, twf, a, inverse);
}

// col_op with the operation fixed at compile time (same arithmetic).
template <typename T, int OP>
__device__ __forceinline__ Cplx<T> c16_op(const ColArgs<T>& a, int o, int kx, int j, Cplx<T> x0, Cplx<T> x1) {
  const long q = (long)j * a.lhx + kx;
  const bool nyq = (kx == a.lhx - 1) || (j == a.ny / 2);
  const Cplx<T> zero = {(T)0, (T)0};
  if constexpr (OP == COP_GRAD) {
    const Cplx<T> v = nyq ? zero : x0;
    if (o == 0) return c_ik(a.kkx[q], v);
    if (o == 1) return c_ik(a.kky[q], v);
    return v;
  } else if constexpr (OP == COP_DIV) {
    const Cplx<T> v0 = nyq ? zero : x0;
    const Cplx<T> v1 = nyq ? zero : x1;
    return c_add(c_ik(a.kkx[q], v0), c_ik(a.kky[q], v1));
  } else if constexpr (OP == COP_FILT2) {
    return c_scale(x0, o == 0 ? a.g0[q] : a.g1[q]);
  } else if constexpr (OP == COP_FILT1) {
    return c_scale(x0, a.g0[q]);
  } else if constexpr (OP == COP_PGRAD) {
    if (o == 0) return c_ik(a.kkx[q], x0);
    if (o == 1) return c_ik(a.kky[q], x0);
    return x0;
  } else {
    return x0;
  }
}

template <typename T, int NIN, int OP, int NOUT>
__global__ void __launch_bounds__(LES_BLOCK) k_cols_c16(ColArgs<T> a, const Cplx<T>* __restrict__ twf) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<T>* ex = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int c = threadIdx.x & 15, ar = threadIdx.x >> 4;  // column in the tile, row class
  const int job = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int k = a.k0 + tile / a.ncb;
  const int c0 = (tile % a.ncb) * 16;
  const bool valid = c0 + c < a.lhx;
  const int kx = valid ? c0 + c : a.lhx - 1;
  const long base = (long)k * 256 * a.lhx + kx;
  Cplx<T> X0[16], X1[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) X0[t] = valid ? a.in[job][0][base + (long)(ar + 16 * t) * a.lhx] : Cplx<T>{(T)0, (T)0};
  if constexpr (NIN == 2 && OP == COP_DIV && NOUT == 1) {
    // divergence: i kx X0 is formed before the second input is loaded (the
    // transforms' barriers keep its loads after), so only one input line
    // and the partial sum are live at a time; the same operations in the
    // same order as c16_op's COP_DIV
    if (a.fwd) c16_fft(X0, twf, false, ex, c, ar);
    const Cplx<T> zero = {(T)0, (T)0};
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int j = ar + 16 * t;
      const bool nyq = (kx == a.lhx - 1) || (j == a.ny / 2);
      X0[t] = c_ik(a.kkx[(long)j * a.lhx + kx], nyq ? zero : X0[t]);
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) X1[t] = valid ? a.in[job][1][base + (long)(ar + 16 * t) * a.lhx] : zero;
    if (a.fwd) c16_fft(X1, twf, false, ex, c, ar);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int j = ar + 16 * t;
      const bool nyq = (kx == a.lhx - 1) || (j == a.ny / 2);
      X0[t] = c_add(X0[t], c_ik(a.kky[(long)j * a.lhx + kx], nyq ? zero : X1[t]));
    }
    if (a.inv) c16_fft(X0, twf, true, ex, c, ar);
    if (valid) {
      Cplx<T>* __restrict__ dst = a.out[job][0] + base;
#pragma unroll
      for (int t = 0; t < 16; ++t) dst[(long)(ar + 16 * t) * a.lhx] = X0[t];
    }
    return;
  }
  if constexpr (NIN > 1) {
#pragma unroll
    for (int t = 0; t < 16; ++t) X1[t] = valid ? a.in[job][1][base + (long)(ar + 16 * t) * a.lhx] : Cplx<T>{(T)0, (T)0};
  }
  if (a.fwd) {
    c16_fft(X0, twf, false, ex, c, ar);
    if constexpr (NIN > 1) c16_fft(X1, twf, false, ex, c, ar);
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    Cplx<T> Y[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) Y[t] = c16_op<T, OP>(a, o, kx, ar + 16 * t, X0[t], NIN > 1 ? X1[t] : X0[t]);
    if (a.inv) c16_fft(Y, twf, true, ex, c, ar);
    if (valid) {
      Cplx<T>* __restrict__ dst = a.out[job][o] + base;
#pragma unroll
      for (int t = 0; t < 16; ++t) dst[(long)(ar + 16 * t) * a.lhx] = Y[t];
    }
  }
}

// ------------------------------------------------------------------ 3/2 resampling
// Padding keeps modes |k| < n/2 (Nyquist dropped) of a length-n1 line in a
// length-n2 spectrum; truncation is the mirror image
// (src/compute_convective_padd.f90:288-319, 352-382).
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

// Row resampling: real rows of length n1 -> real rows of length n2, one
// wavefront per row pair (transform, pad or truncate in place, transform
// back).  src_mode 0: plain real array (nx = n1) src[f]; src_mode 1:
// convection sources (field = blockIdx.y) computed from ConvSrc.
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
  int ls;                     // LDS line stride per wavefront (res_line<T>(n1, n2))
};

// The FP32 256 <-> 384 pair runs on register twiddles with the padded
// placement; other lengths on the generic transform (XOR placement).
template <typename T>
__host__ __device__ constexpr bool res_fixed(int n1, int n2) {
  return sizeof(T) == 4 && ((n1 == 256 && n2 == 384) || (n1 == 384 && n2 == 256));
}
template <typename T>
__host__ __device__ constexpr int res_line(int n1, int n2) {
  return res_fixed<T>(n1, n2) ? lds_line_cap(n1 > n2 ? n1 : n2) : (n1 > n2 ? n1 : n2);
}

template <typename T, typename TI, typename TO, typename RES>
__device__ __forceinline__ void rows_resample_body(const RowResArgs<T, TI, TO>& a, Cplx<T>* z, int lane, long job,
                                                   RES&& res) {
  const int n1 = a.n1, n2 = a.n2;
  const int npairs = a.ny / 2;
  const int f = blockIdx.y;
  const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
  const long np = (long)n1 * a.ny;
  if (a.src_mode == 1) {
    for (int n = lane; n < n1; n += LES_WAVE) {
      const long q = (long)(2 * jp) * n1 + n;
      z[n] = {conv_src<T, TI>(a.cs, f, k, q, np, a.cconv), conv_src<T, TI>(a.cs, f, k, q + n1, np, a.cconv)};
    }
  } else {
    const T* __restrict__ src = a.src[f] + (long)k * a.src_plane + (long)(2 * jp) * n1;
    for (int n = lane; n < n1; n += LES_WAVE) z[n] = {src[n], src[n + n1]};
  }
  wave_sync();
  res(z);
  TO* __restrict__ dst = a.dst[f] + (long)k * a.dst_plane + (long)(2 * jp) * n2;
  for (int n = lane; n < n2; n += LES_WAVE) {
    const Cplx<T> v = z[n];
    dst[n] = (TO)v.x;
    dst[n + n2] = (TO)v.y;
  }
}

// The FP32 256 <-> 384 pair runs on register twiddles (37 per lane); FP32
// instances are held to 80 VGPRs (6 waves per SIMD, no spills).
template <typename T, typename TI, typename TO>
__global__ void __launch_bounds__(LES_BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
k_rows_resample(RowResArgs<T, TI, TO> a, FftPlan p1, FftPlan p2, const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  Cplx<T>* z = reinterpret_cast<Cplx<T>*>(smem_raw) + wave * a.ls;
  const long job = (long)blockIdx.x * (blockDim.x / LES_WAVE) + wave;
  if (job >= (long)a.nplanes * (a.ny / 2)) return;
  if constexpr (sizeof(T) == 4) {
    // each transform's twiddles are loaded where it starts, so the two sets
    // are never live together (fewer VGPRs, more waves per SIMD)
    if (res_fixed<T>(p1.n, p2.n) && p1.n == 256) {
      rows_resample_body(a, z, lane, job, [&](Cplx<T>* q) {
        {
          TwReg<T, 256> ta;
          tw_load(ta, tw + p1.twf, lane);
          wave_fft_t<T, 256, LES_LAY_PAD>(q, ta, false, lane);
        }
        wave_remap(q, 256, 384, lane);
        TwReg<T, 384> tb;
        tw_load(tb, tw + p2.twf, lane);
        wave_fft_t<T, 384, LES_LAY_PAD>(q, tb, true, lane);
      });
      return;
    }
    if (res_fixed<T>(p1.n, p2.n) && p1.n == 384) {
      rows_resample_body(a, z, lane, job, [&](Cplx<T>* q) {
        {
          TwReg<T, 384> ta;
          tw_load(ta, tw + p1.twf, lane);
          wave_fft_t<T, 384, LES_LAY_PAD>(q, ta, false, lane);
        }
        wave_remap(q, 384, 256, lane);
        TwReg<T, 256> tb;
        tw_load(tb, tw + p2.twf, lane);
        wave_fft_t<T, 256, LES_LAY_PAD>(q, tb, true, lane);
      });
      return;
    }
  }
#ifndef LES_RES64_FIXED
#define LES_RES64_FIXED 1
#endif
  if constexpr (sizeof(T) == 8 && LES_RES64_FIXED) {
    // FP64 256 <-> 384: register twiddles with the XOR placement (line of n2)
    if ((p1.n == 256 && p2.n == 384) || (p1.n == 384 && p2.n == 256)) {
      if (p1.n == 256) {
        rows_resample_body(a, z, lane, job, [&](Cplx<T>* q) {
          {
            TwReg<T, 256> ta;
            tw_load(ta, tw + p1.twf, lane);
            wave_fft_t<T, 256, LES_LAY_XOR>(q, ta, false, lane);
          }
          wave_remap(q, 256, 384, lane);
          TwReg<T, 384> tb;
          tw_load(tb, tw + p2.twf, lane);
          wave_fft_t<T, 384, LES_LAY_XOR>(q, tb, true, lane);
        });
      } else {
        rows_resample_body(a, z, lane, job, [&](Cplx<T>* q) {
          {
            TwReg<T, 384> ta;
            tw_load(ta, tw + p1.twf, lane);
            wave_fft_t<T, 384, LES_LAY_XOR>(q, ta, false, lane);
          }
          wave_remap(q, 384, 256, lane);
          TwReg<T, 256> tb;
          tw_load(tb, tw + p2.twf, lane);
          wave_fft_t<T, 256, LES_LAY_XOR>(q, tb, true, lane);
        });
      }
      return;
    }
  }
  rows_resample_body(a, z, lane, job, [&](Cplx<T>* q) {
    wave_fft(q, p1, tw, false, lane);
    wave_remap(q, a.n1, a.n2, lane);
    wave_fft(q, p2, tw, true, lane);
  });
}

// Column resampling: column pairs (i, i+1) of a real (W, h1) plane -> (W, h2).
// A tile of CP pairs is loaded row by row (2 CP contiguous reals), held as
// CP complex lines in LDS, resampled by one wavefront per line, stored row by
// row.
template <typename T>
struct ColResArgs {
  const T* src[LES_MAX_FIELDS];
  T* dst[LES_MAX_FIELDS];
  int W, h1, h2;
  int k0, nplanes;
  long src_plane, dst_plane;
  int ls;   // LDS line stride (at least max(h1, h2))
  int ncp;  // column-pair tiles per plane
};

template <typename T, int CP, typename RES>
__device__ __forceinline__ void cols_resample_body(const ColResArgs<T>& a, unsigned char* smem, RES&& res) {
  const int h1 = a.h1, h2 = a.h2, LS = a.ls;
  Cplx<T>* buf = reinterpret_cast<Cplx<T>*>(smem);
  const int f = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int k = a.k0 + tile / a.ncp;
  const int p0 = (tile % a.ncp) * CP;  // first column pair
  int nl = a.W / 2 - p0;
  if (nl > CP) nl = CP;
  const T* __restrict__ src = a.src[f] + (long)k * a.src_plane + 2 * p0;
  for (int idx = threadIdx.x; idx < 2 * CP * h1; idx += blockDim.x) {
    const int j = idx / (2 * CP), r = idx % (2 * CP);
    const int l = r >> 1;
    if (l < nl) {
      const T v = src[(long)j * a.W + r];
      if (r & 1) buf[l * LS + j].y = v;
      else buf[l * LS + j].x = v;
    }
  }
  __syncthreads();
  const int wave = threadIdx.x / LES_WAVE, nw = blockDim.x / LES_WAVE;
  for (int l = wave; l < nl; l += nw) res(buf + l * LS);
  __syncthreads();
  T* __restrict__ dst = a.dst[f] + (long)k * a.dst_plane + 2 * p0;
  for (int idx = threadIdx.x; idx < 2 * CP * h2; idx += blockDim.x) {
    const int j = idx / (2 * CP), r = idx % (2 * CP);
    const int l = r >> 1;
    if (l < nl) {
      const Cplx<T> z = buf[l * LS + j];
      dst[(long)j * a.W + r] = (r & 1) ? z.y : z.x;
    }
  }
}

// ------------------------------------------------------------------ column resampling 256 <-> 384, in registers
// The 384-point transform (stages radix 4, 4, 4, 2, 3) splits like the
// 256-point one: stages 0-1 map inputs a + 24 t (t < 16) of thread a (a < 24)
// to outputs 16 a + c, stages 2-4 stay in the residue class k mod 16 (thread
// k < 16 holds k + 16 t, t < 24).  A block resamples 16 complex lines (32
// real columns, 128-byte row segments) with 384 threads (line l = tid % 16,
// row class r = tid / 16): the 256-point halves use row classes 0..15, the
// 384-point first half all 24.  Between the transforms the spectrum is
// re-laid through LDS with the padding / truncation of wave_remap.
#define LES_C24_LS 417  // line stride in the exchange buffer (>= c16_pad(383) + 1, = 1 mod 16)
// exchange-buffer row e of a line: one pad slot after every 16 rows
__device__ __forceinline__ int c16_pad(int e) { return e + (e >> 4); }
static_assert(cp_OFF(384, 1) == 3 && cp_OFF(384, 2) == 15 && cp_OFF(384, 3) == 63 && cp_OFF(384, 4) == 127,
              "384-point twiddle layout");

// stages 2-4 of the 384-point transform: X[t] = z''[k + 16 t] -> X[t] = Z[k + 16 t]
template <typename T>
__device__ __forceinline__ void c24_pair_b(Cplx<T>* X, const Cplx<T>* __restrict__ twf, int k, bool inverse) {
  const int sgn = inverse ? 1 : -1;
  Cplx<T> V[24];
#pragma unroll
  for (int q = 0; q < 6; ++q) {  // stage 2 (radix 4, span 16): butterfly 16 q + k
    Cplx<T> x[4] = {X[q], X[q + 6], X[q + 12], X[q + 18]};
#pragma unroll
    for (int r = 1; r < 4; ++r) x[r] = c_mul(x[r], tw16(twf, 15 + (r - 1) * 16 + k, inverse));
    bfly(x, 4, sgn);
#pragma unroll
    for (int r = 0; r < 4; ++r) V[4 * q + r] = x[r];
  }
#pragma unroll
  for (int u = 0; u < 12; ++u) {  // stage 3 (radix 2, span 64): butterfly 16 u + k
    Cplx<T> x[4] = {V[u], c_mul(V[u + 12], tw16(twf, 63 + 16 * (u % 4) + k, inverse)), V[u], V[u]};
    bfly(x, 2, sgn);
    X[8 * (u / 4) + (u % 4)] = x[0];
    X[8 * (u / 4) + (u % 4) + 4] = x[1];
  }
#pragma unroll
  for (int v = 0; v < 8; ++v) {  // stage 4 (radix 3, span 128): butterfly k + 16 v
    Cplx<T> x[4] = {X[v], X[v + 8], X[v + 16], X[v]};
#pragma unroll
    for (int r = 1; r < 3; ++r) x[r] = c_mul(x[r], tw16(twf, 127 + (r - 1) * 128 + k + 16 * v, inverse));
    bfly(x, 3, sgn);
    V[v] = x[0];
    V[v + 8] = x[1];
    V[v + 16] = x[2];
  }
#pragma unroll
  for (int t = 0; t < 24; ++t) X[t] = V[t];
}

// PAD: 256 -> 384 rows (forward 256, pad, inverse 384); else 384 -> 256.
// FP32: held to 96 VGPRs (5 waves per SIMD instead of 4; two VGPRs spill
// in the padding direction).
template <typename T, bool PAD>
__global__ void __launch_bounds__(384) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 1)))
k_colres_c16(ColResArgs<T> a, const Cplx<T>* __restrict__ tw256,
                                                    const Cplx<T>* __restrict__ tw384) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<T>* ex = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int l = threadIdx.x & 15, rc = threadIdx.x >> 4;  // line, row class (0..23)
  const int f = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int k = a.k0 + tile / a.ncp;
  const int p0 = (tile % a.ncp) * 16;  // first line (column pair)
  const bool valid = p0 + l < a.W / 2;
  const long col = 2L * (valid ? p0 + l : 0);
  const Cplx<T>* __restrict__ src = reinterpret_cast<const Cplx<T>*>(a.src[f] + (long)k * a.src_plane + col);
  Cplx<T>* __restrict__ dst = reinterpret_cast<Cplx<T>*>(a.dst[f] + (long)k * a.dst_plane + col);
  const long rs = a.W / 2;  // row stride in complex elements
  Cplx<T>* line = ex + l * LES_C24_LS;
  const Cplx<T> zero = {(T)0, (T)0};
  Cplx<T> X[24];
  if constexpr (PAD) {
    // forward 256: rows rc + 16 t
    if (rc < 16) {
#pragma unroll
      for (int t = 0; t < 16; ++t) X[t] = valid ? src[(long)(rc + 16 * t) * rs] : zero;
      c16_pair_a(X, tw256, false);
#pragma unroll
      for (int c = 0; c < 16; ++c) lds_st(line + c16_pad(16 * rc + c), X[c]);
    }
    __syncthreads();
    if (rc < 16) {
#pragma unroll
      for (int t = 0; t < 16; ++t) X[t] = lds_ld(line + c16_pad(rc + 16 * t));
    }
    __syncthreads();
    if (rc < 16) {
      c16_pair_b(X, tw256, rc, false);
      // spectrum rows m = rc + 16 t of 256 -> rows of 384: m < 128 kept, 128
      // (Nyquist) dropped, m > 128 -> m + 128; rows 128..256 zero
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int m = rc + 16 * t;
        if (m != 128) {
          const int e = m < 128 ? m : m + 128;
          lds_st(line + c16_pad(e), X[t]);
        }
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) lds_st(line + c16_pad(128 + rc + 16 * t), zero);
      if (rc == 0) lds_st(line + c16_pad(256), zero);
    }
    __syncthreads();
    // inverse 384, first half: rows rc + 24 t
#pragma unroll
    for (int t = 0; t < 16; ++t) X[t] = lds_ld(line + c16_pad(rc + 24 * t));
    __syncthreads();
    c16_pair_a(X, tw384, true);
#pragma unroll
    for (int c = 0; c < 16; ++c) lds_st(line + c16_pad(16 * rc + c), X[c]);
    __syncthreads();
    if (rc < 16) {
#pragma unroll
      for (int t = 0; t < 24; ++t) X[t] = lds_ld(line + c16_pad(rc + 16 * t));
      c24_pair_b(X, tw384, rc, true);
      if (valid) {
#pragma unroll
        for (int t = 0; t < 24; ++t) dst[(long)(rc + 16 * t) * rs] = X[t];
      }
    }
  } else {
    // forward 384, first half: rows rc + 24 t
#pragma unroll
    for (int t = 0; t < 16; ++t) X[t] = valid ? src[(long)(rc + 24 * t) * rs] : zero;
    c16_pair_a(X, tw384, false);
#pragma unroll
    for (int c = 0; c < 16; ++c) lds_st(line + c16_pad(16 * rc + c), X[c]);
    __syncthreads();
    if (rc < 16) {
#pragma unroll
      for (int t = 0; t < 24; ++t) X[t] = lds_ld(line + c16_pad(rc + 16 * t));
    }
    __syncthreads();
    if (rc < 16) {
      c24_pair_b(X, tw384, rc, false);
      // spectrum rows m = rc + 16 t of 384 -> rows of 256: m < 128 kept,
      // m > 256 -> m - 128, the rest dropped; row 128 zero
#pragma unroll
      for (int t = 0; t < 24; ++t) {
        const int m = rc + 16 * t;
        if (m < 128) lds_st(line + c16_pad(m), X[t]);
        else if (m > 256) lds_st(line + c16_pad(m - 128), X[t]);
      }
      if (rc == 0) lds_st(line + c16_pad(128), zero);
    }
    __syncthreads();
    // inverse 256
    if (rc < 16) {
#pragma unroll
      for (int t = 0; t < 16; ++t) X[t] = lds_ld(line + c16_pad(rc + 16 * t));
    }
    __syncthreads();
    if (rc < 16) {
      c16_pair_a(X, tw256, true);
#pragma unroll
      for (int c = 0; c < 16; ++c) lds_st(line + c16_pad(16 * rc + c), X[c]);
    }
    __syncthreads();
    if (rc < 16) {
#pragma unroll
      for (int t = 0; t < 16; ++t) X[t] = lds_ld(line + c16_pad(rc + 16 * t));
      c16_pair_b(X, tw256, rc, true);
      if (valid) {
#pragma unroll
        for (int t = 0; t < 16; ++t) dst[(long)(rc + 16 * t) * rs] = X[t];
      }
    }
  }
}

template <typename T, int CP>
__global__ void __launch_bounds__(LES_BLOCK) k_cols_resample(ColResArgs<T> a, FftPlan p1, FftPlan p2,
                                                             const Cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1);
  if constexpr (sizeof(T) == 4) {
    if (p1.n == 256 && p2.n == 384) {
      TwReg<T, 256> ta;
      TwReg<T, 384> tb;
      tw_load(ta, tw + p1.twf, lane);
      tw_load(tb, tw + p2.twf, lane);
      cols_resample_body<T, CP>(a, smem_raw, [&](Cplx<T>* q) {
        wave_fft_t(q, ta, false, lane);
        wave_remap(q, 256, 384, lane);
        wave_fft_t(q, tb, true, lane);
      });
      return;
    }
    if (p1.n == 384 && p2.n == 256) {
      TwReg<T, 384> ta;
      TwReg<T, 256> tb;
      tw_load(ta, tw + p1.twf, lane);
      tw_load(tb, tw + p2.twf, lane);
      cols_resample_body<T, CP>(a, smem_raw, [&](Cplx<T>* q) {
        wave_fft_t(q, ta, false, lane);
        wave_remap(q, 384, 256, lane);
        wave_fft_t(q, tb, true, lane);
      });
      return;
    }
  }
  cols_resample_body<T, CP>(a, smem_raw, [&](Cplx<T>* q) {
    wave_fft(q, p1, tw, false, lane);
    wave_remap(q, a.h1, a.h2, lane);
    wave_fft(q, p2, tw, true, lane);
  });
}

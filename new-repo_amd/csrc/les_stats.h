// Running-average statistics and running diagnostics on the device.
//
// Statistics (src/IO_statistics.f90:266-513): per point of levels 1..nz, FP64
// running sums of 26 quantities in the order of the reference's RAV_*_3D.dat
// file, plus the sum of u* over the surface (RAV_*_2D.dat), and the number of
// accumulated steps.  The sums are accumulated in every precision mode; in
// the mixed modes the FP32 fields are widened first.
//
// Profiles (matlab_src/get_data_struct_RAV.m:21-35,
// matlab_src/get_profile_struct.m:9-59, matlab_src/compute_Cij.m:1-36):
// plane means of the time averages (sum / steps, SGS stresses negated), with
// the resolved stresses R_i3 = <u_i w> - u_i(w node) w, the total stresses
// and the dispersive stresses C_i3.  A two-pass reduction with a fixed order:
// slices of each plane into partial sums, then one thread per level.
//
// Diagnostics (src/running_diagnostics.f90:80-305): CFL, mean kinetic
// energy of the domain-mean velocity, plane-mean u*, rms of the divergence
// of the velocity gradients, by a fixed-grid reduction then one block.
#pragma once

// sums of the 3D running averages, reference file order
// (src/IO_statistics.f90:458-490)
enum {
  ST_U = 0, ST_V, ST_W, ST_P, ST_U2, ST_V2, ST_W2, ST_U3, ST_V3, ST_W3, ST_U4, ST_V4, ST_W4,
  ST_UV, ST_UW, ST_VW, ST_TXX, ST_TYY, ST_TZZ, ST_TXY, ST_TXZ, ST_TYZ, ST_DUDZ, ST_DVDZ, ST_NUT, ST_CS, ST_N
};
static_assert(ST_N == LES_NSTAT, "statistics count");

template <typename T>
struct StatAccArgs {
  const double *u, *v, *w, *p, *ustar;
  const T *txx, *tyy, *tzz, *txy, *txz, *tyz, *dudz, *dvdz, *nut, *cs;
  double* s;    // ST_N blocks of n sums; point (i, j, k) at (k - 1) np + i + nx j
  double* sus;  // np sums of u*
  long np, n;   // n = np nz
};

// One step of io_3D_stat_averages + io_2D_stat_averages
// (src/IO_statistics.f90:319-324,372-440).  Powers as products: u**2 = u u,
// u**3 = (u u) u, u**4 = (u u)(u u).  <uw> at level 1 is reset to zero.
template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) k_stat_accum(StatAccArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const long c = a.np + t;  // field index of level k = t / np + 1
  const long m = a.n;
  double* s = a.s + t;
  const double u = a.u[c], v = a.v[c], w = a.w[c];
  if (t < a.np) {
    s[ST_UW * m] = 0.0;
    s[ST_VW * m] = 0.0;
    a.sus[t] = a.sus[t] + a.ustar[t];
  } else {
    const double um = a.u[c - a.np], vm = a.v[c - a.np];
    s[ST_UW * m] = s[ST_UW * m] + w * (0.5 * (um + u));
    s[ST_VW * m] = s[ST_VW * m] + w * (0.5 * (vm + v));
  }
  s[ST_U * m] = s[ST_U * m] + u;
  s[ST_V * m] = s[ST_V * m] + v;
  s[ST_W * m] = s[ST_W * m] + w;
  s[ST_P * m] = s[ST_P * m] + a.p[c];
  const double u2 = u * u, v2 = v * v, w2 = w * w;
  s[ST_U2 * m] = s[ST_U2 * m] + u2;
  s[ST_V2 * m] = s[ST_V2 * m] + v2;
  s[ST_W2 * m] = s[ST_W2 * m] + w2;
  s[ST_U3 * m] = s[ST_U3 * m] + u2 * u;
  s[ST_V3 * m] = s[ST_V3 * m] + v2 * v;
  s[ST_W3 * m] = s[ST_W3 * m] + w2 * w;
  s[ST_U4 * m] = s[ST_U4 * m] + u2 * u2;
  s[ST_V4 * m] = s[ST_V4 * m] + v2 * v2;
  s[ST_W4 * m] = s[ST_W4 * m] + w2 * w2;
  s[ST_UV * m] = s[ST_UV * m] + u * v;
  s[ST_TXX * m] = s[ST_TXX * m] + (double)a.txx[c];
  s[ST_TYY * m] = s[ST_TYY * m] + (double)a.tyy[c];
  s[ST_TZZ * m] = s[ST_TZZ * m] + (double)a.tzz[c];
  s[ST_TXY * m] = s[ST_TXY * m] + (double)a.txy[c];
  s[ST_TXZ * m] = s[ST_TXZ * m] + (double)a.txz[c];
  s[ST_TYZ * m] = s[ST_TYZ * m] + (double)a.tyz[c];
  s[ST_DUDZ * m] = s[ST_DUDZ * m] + (double)a.dudz[c];
  s[ST_DVDZ * m] = s[ST_DVDZ * m] + (double)a.dvdz[c];
  s[ST_NUT * m] = s[ST_NUT * m] + (double)a.nut[c];
  s[ST_CS * m] = s[ST_CS * m] + sqrt((double)a.cs[c]);
}

__global__ void k_fill_zero(double* p, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// ---------------------------------------------------------------- profiles
// Per-level sums of the profile reduction (X = time average of X):
enum {
  PS_U = 0, PS_V, PS_W, PS_P, PS_VARU, PS_VARV, PS_VARW, PS_VARW_UVP, PS_TXX, PS_TYY, PS_TZZ,
  PS_R13, PS_TXZ, PS_ST13, PS_R23, PS_TYZ, PS_ST23, PS_UCW, PS_UC, PS_VCW, PS_VC, PS_CS, PS_NUT, PS_N
};
#define LES_PROF_SLICES 16

struct StatPlaneArgs {
  const double* s;
  double* part;    // [nz][LES_PROF_SLICES][PS_N]
  long np, n;
  int nx, nz;
  double nt;       // accumulated steps
};

// time average of sum `f` at in-plane point q of level k (1-based)
__device__ __forceinline__ double st_avg(const StatPlaneArgs& a, int f, long q, int k) {
  return a.s[(long)f * a.n + (long)(k - 1) * a.np + q] / a.nt;
}

// One block per (slice, level): the slice's rows j, each row's points i in
// order, combined across the block's threads in a fixed tree.
__global__ void __launch_bounds__(LES_BLOCK) k_stat_planes(StatPlaneArgs a) {
  const int slice = blockIdx.x, k = blockIdx.y + 1;
  const long q0 = a.np * slice / LES_PROF_SLICES, q1 = a.np * (slice + 1) / LES_PROF_SLICES;
  double acc[PS_N];
#pragma unroll
  for (int f = 0; f < PS_N; ++f) acc[f] = 0.0;
  for (long q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const double u = st_avg(a, ST_U, q, k), v = st_avg(a, ST_V, q, k), w = st_avg(a, ST_W, q, k);
    const double vu = st_avg(a, ST_U2, q, k) - u * u;
    const double vv = st_avg(a, ST_V2, q, k) - v * v;
    const double vw = st_avg(a, ST_W2, q, k) - w * w;
    double vw_uvp = 0.0;
    if (k < a.nz) {
      const double wp = st_avg(a, ST_W, q, k + 1);
      vw_uvp = 0.5 * (vw + (st_avg(a, ST_W2, q, k + 1) - wp * wp));
    }
    // u, v on the w node: 0 at the surface level, else the mean of k-1 and k
    double uc = 0.0, vc = 0.0;
    if (k > 1) {
      uc = 0.5 * (st_avg(a, ST_U, q, k - 1) + u);
      vc = 0.5 * (st_avg(a, ST_V, q, k - 1) + v);
    }
    const double txz = -st_avg(a, ST_TXZ, q, k), tyz = -st_avg(a, ST_TYZ, q, k);
    const double r13 = st_avg(a, ST_UW, q, k) - uc * w;
    const double r23 = st_avg(a, ST_VW, q, k) - vc * w;
    acc[PS_U] += u;
    acc[PS_V] += v;
    acc[PS_W] += w;
    acc[PS_P] += st_avg(a, ST_P, q, k);
    acc[PS_VARU] += vu;
    acc[PS_VARV] += vv;
    acc[PS_VARW] += vw;
    acc[PS_VARW_UVP] += vw_uvp;
    acc[PS_TXX] += -st_avg(a, ST_TXX, q, k);
    acc[PS_TYY] += -st_avg(a, ST_TYY, q, k);
    acc[PS_TZZ] += -st_avg(a, ST_TZZ, q, k);
    acc[PS_R13] += r13;
    acc[PS_TXZ] += txz;
    acc[PS_ST13] += r13 + txz;
    acc[PS_R23] += r23;
    acc[PS_TYZ] += tyz;
    acc[PS_ST23] += r23 + tyz;
    acc[PS_UCW] += uc * w;
    acc[PS_UC] += uc;
    acc[PS_VCW] += vc * w;
    acc[PS_VC] += vc;
    acc[PS_CS] += st_avg(a, ST_CS, q, k);
    acc[PS_NUT] += st_avg(a, ST_NUT, q, k);
  }
  __shared__ double red[LES_BLOCK];
  double* out = a.part + ((long)(k - 1) * LES_PROF_SLICES + slice) * PS_N;
  for (int f = 0; f < PS_N; ++f) {
    red[threadIdx.x] = acc[f];
    __syncthreads();
    for (int h = LES_BLOCK / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[f] = red[0];
    __syncthreads();
  }
}

// Output rows of les_stat_profiles (LES_NPROF x nz, row-major by quantity).
enum {
  PR_U = 0, PR_V, PR_W, PR_P, PR_VARU, PR_VARV, PR_VARW, PR_TXX, PR_TYY, PR_TZZ, PR_R13, PR_TXZ, PR_ST13,
  PR_R23, PR_TYZ, PR_ST23, PR_TKE, PR_CS, PR_NUT, PR_C13, PR_C23, PR_N
};
static_assert(PR_N == LES_NPROF, "profile count");

__global__ void k_stat_profiles(const double* part, double* prof, long np, int nz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  double m[PS_N];
  for (int f = 0; f < PS_N; ++f) {
    double acc = 0.0;
    for (int s = 0; s < LES_PROF_SLICES; ++s) acc += part[((long)k * LES_PROF_SLICES + s) * PS_N + f];
    m[f] = acc / (double)np;
  }
  const double c13 = m[PS_UCW] - m[PS_UC] * m[PS_W];
  const double c23 = m[PS_VCW] - m[PS_VC] * m[PS_W];
  double* o = prof + k;
  o[PR_U * nz] = m[PS_U];
  o[PR_V * nz] = m[PS_V];
  o[PR_W * nz] = m[PS_W];
  o[PR_P * nz] = m[PS_P];
  o[PR_VARU * nz] = m[PS_VARU];
  o[PR_VARV * nz] = m[PS_VARV];
  o[PR_VARW * nz] = m[PS_VARW];
  o[PR_TXX * nz] = m[PS_TXX];
  o[PR_TYY * nz] = m[PS_TYY];
  o[PR_TZZ * nz] = m[PS_TZZ];
  o[PR_R13 * nz] = m[PS_R13];
  o[PR_TXZ * nz] = m[PS_TXZ];
  o[PR_ST13 * nz] = m[PS_ST13] + c13;
  o[PR_R23 * nz] = m[PS_R23];
  o[PR_TYZ * nz] = m[PS_TYZ];
  o[PR_ST23 * nz] = m[PS_ST23] + c23;
  o[PR_TKE * nz] = 0.5 * (m[PS_VARU] + m[PS_VARV] + m[PS_VARW_UVP]);
  o[PR_CS * nz] = m[PS_CS];
  o[PR_NUT * nz] = m[PS_NUT];
  o[PR_C13 * nz] = c13;
  o[PR_C23 * nz] = c23;
}

// ---------------------------------------------------------------- diagnostics
enum { DG_UMAX = 0, DG_VMAX, DG_WMAX, DG_USUM, DG_VSUM, DG_WSUM, DG_DIV2, DG_USTAR, DG_N };
#define LES_DIAG_BLOCKS 1024

template <typename T>
struct DiagArgs {
  const double *u, *v, *w, *ustar;
  const T *dudx, *dvdy, *dwdz;
  double* part;  // [LES_DIAG_BLOCKS][DG_N]
  long np, n;    // n = np nz
};

// Fixed grid of LES_DIAG_BLOCKS blocks striding over levels 1..nz; each
// block writes its partial maxima and sums (src/running_diagnostics.f90:
// 148-181 max |u|, |v|, |w|; 193-234 sums of u, v, w on the uvp node;
// 246-269 u*; 281-305 (du/dx + dv/dy + dw/dz)**2).
template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) k_diag(DiagArgs<T> a) {
  double acc[DG_N];
#pragma unroll
  for (int f = 0; f < DG_N; ++f) acc[f] = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < a.n; t += stride) {
    const long c = a.np + t;
    const double u = a.u[c], v = a.v[c], w = a.w[c];
    acc[DG_UMAX] = fmax(acc[DG_UMAX], fabs(u));
    acc[DG_VMAX] = fmax(acc[DG_VMAX], fabs(v));
    acc[DG_WMAX] = fmax(acc[DG_WMAX], fabs(w));
    acc[DG_USUM] += u;
    acc[DG_VSUM] += v;
    acc[DG_WSUM] += 0.5 * (w + a.w[c + a.np]);
    const double d = ((double)a.dudx[c] + (double)a.dvdy[c]) + (double)a.dwdz[c];
    acc[DG_DIV2] += d * d;
    if (t < a.np) acc[DG_USTAR] += a.ustar[t];
  }
  __shared__ double red[LES_BLOCK];
  for (int f = 0; f < DG_N; ++f) {
    red[threadIdx.x] = acc[f];
    __syncthreads();
    for (int h = LES_BLOCK / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) {
        const double o = red[threadIdx.x + h];
        red[threadIdx.x] = f < DG_USUM ? fmax(red[threadIdx.x], o) : red[threadIdx.x] + o;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) a.part[(long)blockIdx.x * DG_N + f] = red[0];
    __syncthreads();
  }
}

struct DiagFinalArgs {
  const double* part;
  double* out;  // cfl, mke, u_star, rms div
  double dt, dx, dy, dz, nxyz, nxy;
};

// One block: partials of block b summed in block order by thread b mod
// LES_BLOCK, then a fixed tree; the formulas of the reference.
__global__ void __launch_bounds__(LES_BLOCK) k_diag_final(DiagFinalArgs a) {
  double acc[DG_N];
#pragma unroll
  for (int f = 0; f < DG_N; ++f) acc[f] = 0.0;
  for (int b = threadIdx.x; b < LES_DIAG_BLOCKS; b += LES_BLOCK)
    for (int f = 0; f < DG_N; ++f) {
      const double o = a.part[(long)b * DG_N + f];
      acc[f] = f < DG_USUM ? fmax(acc[f], o) : acc[f] + o;
    }
  __shared__ double red[LES_BLOCK];
  __shared__ double tot[DG_N];
  for (int f = 0; f < DG_N; ++f) {
    red[threadIdx.x] = acc[f];
    __syncthreads();
    for (int h = LES_BLOCK / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) {
        const double o = red[threadIdx.x + h];
        red[threadIdx.x] = f < DG_USUM ? fmax(red[threadIdx.x], o) : red[threadIdx.x] + o;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) tot[f] = red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.out[0] = a.dt * (tot[DG_UMAX] / a.dx + tot[DG_VMAX] / a.dy + tot[DG_WMAX] / a.dz);
    const double us = tot[DG_USUM] / a.nxyz, vs = tot[DG_VSUM] / a.nxyz, ws = tot[DG_WSUM] / a.nxyz;
    a.out[1] = 0.5 * (us * us + vs * vs + ws * ws);
    a.out[2] = tot[DG_USTAR] / a.nxy;
    a.out[3] = sqrt(tot[DG_DIV2] / a.nxyz);
  }
}

// Lagrangian scale-dependent dynamic model: test filtering of the Germano
// inputs, Lagrangian pathline interpolation of F_LM/F_MM/F_QN/F_NN, the
// relaxation update and the dynamic coefficients
// (src/compute_LagScaleDep.f90:204-707).  Included by les_gpu.hip.
#pragma once

// FF(i, j, k) of the level-2 halo array built by update_halo3 and the z
// fix-ups of compute_interpolag_coeff (src/decomp_2d_custom.f90:902-1024,
// src/compute_LagScaleDep.f90:439-452) for 1-based i in [0, nx+1],
// j in [0, ny+1], k in [0, nz+2].  halo_mode 0 reproduces a one-rank run:
// y halo rows are shifted copies (row 0 <- row 2, row ny+1 <- row ny-1),
// plane 0 is plane 1 and the halo cells of planes nz+1, nz+2 are those of
// planes nz-1, nz.  halo_mode 1 is periodic in x and y with clamped z.
__device__ __forceinline__ long ff_index(int i, int j, int k, int nx, int ny, int nz, int mode) {
  int kk = k;
  if (mode == 0) {
    const bool inter = (i >= 1 && i <= nx && j >= 1 && j <= ny);
    if (k == 0) kk = 1;
    if (k >= nz + 1) {
      if (inter) return ((long)(nz + 1) * ny + (j - 1)) * nx + (i - 1);
      kk = k - 2;
    }
    int jj = j == 0 ? 2 : (j == ny + 1 ? ny - 1 : j);
    int ii = i == 0 ? nx : (i == nx + 1 ? 1 : i);
    return ((long)kk * ny + (jj - 1)) * nx + (ii - 1);
  }
  if (k == 0) kk = 1;
  if (k >= nz + 1) kk = nz + 1;
  int jj = j == 0 ? ny : (j == ny + 1 ? 1 : j);
  int ii = i == 0 ? nx : (i == nx + 1 ? 1 : i);
  return ((long)kk * ny + (jj - 1)) * nx + (ii - 1);
}

__device__ __forceinline__ double fmax32(double a, float b) {
  const float fa = (float)a;
  return (double)(fa > b ? fa : b);
}

// tmp**(-1/8) from correctly rounded operations only
__device__ __forceinline__ double pow_m18(double t) { return 1.0 / sqrt(sqrt(sqrt(t))); }

template <typename T>
struct LagrArgs {
  const T *ulag, *vlag, *wlag;
  const double* Fo[4];  // previous F_LM, F_MM, F_QN, F_NN
  double* Fn[4];        // updated
  const T *LM, *MM, *QN, *NN;
  T *cs, *beta;
  double dt, csc, dx, dy, dz, opftdelta, lagran_dt;
  int nx, ny, nz;
  int nlev;  // levels 1..nlev are updated
  long np;
  int interp;  // 0 on the first update (F initialised from MM, NN)
  int halo_mode;
};

template <typename T>
__global__ void k_lagr(LagrArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nlev) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const int i = 1 + (int)(q % a.nx), j = 1 + (int)(q / a.nx);
  const long c = (long)k * a.np + q;
  double F[4];
  if (a.interp) {
    // compute_interpolag_weight (src/compute_LagScaleDep.f90:309-411)
    const double cs = a.csc;
    double uw, vw, ww;
    if (k == 1) {
      uw = (double)a.ulag[c] / cs;
      vw = (double)a.vlag[c] / cs;
      ww = 0.25 * ((double)a.wlag[c + a.np] / cs);
    } else {
      uw = 0.5 * ((double)a.ulag[c] / cs + (double)a.ulag[c - a.np] / cs);
      vw = 0.5 * ((double)a.vlag[c] / cs + (double)a.vlag[c - a.np] / cs);
      ww = (double)a.wlag[c] / cs;
    }
    const double xp = -(uw * a.dt * cs / a.dx);
    const double yp = -(vw * a.dt * cs / a.dy);
    double zp;
    if (k <= 2) {
      const double w2 = (double)a.wlag[2 * a.np + q] / cs;
      double z2 = -(w2 * a.dt * cs / a.dz);
      z2 = z2 + fmin(z2, 0.0);
      z2 = fmin(1.0, z2);
      zp = fmax(-1.0, z2);
    } else {
      zp = -(ww * a.dt * cs / a.dz);
    }
    const int ax = i + (signbit(xp) ? -1 : 1);
    const int ay = j + (signbit(yp) ? -1 : 1);
    const int az = k + (signbit(zp) ? -1 : 1);
    const double cx = fabs(xp), cy = fabs(yp), cz = fabs(zp);
    const double fx = 1.0 - cx, fy = 1.0 - cy, fz = 1.0 - cz;
    const int nx = a.nx, ny = a.ny, nz = a.nz, hm = a.halo_mode;
    const long i000 = ff_index(i, j, k, nx, ny, nz, hm), i001 = ff_index(i, j, az, nx, ny, nz, hm);
    const long i010 = ff_index(i, ay, k, nx, ny, nz, hm), i011 = ff_index(i, ay, az, nx, ny, nz, hm);
    const long i100 = ff_index(ax, j, k, nx, ny, nz, hm), i101 = ff_index(ax, j, az, nx, ny, nz, hm);
    const long i110 = ff_index(ax, ay, k, nx, ny, nz, hm), i111 = ff_index(ax, ay, az, nx, ny, nz, hm);
    // compute_interpolag_coeff (src/compute_LagScaleDep.f90:455-469)
    for (int f = 0; f < 4; ++f) {
      const double* G = a.Fo[f];
      F[f] = fx * fy * (G[i000] * fz + G[i001] * cz) + fx * cy * (G[i010] * fz + G[i011] * cz) +
             cx * fy * (G[i100] * fz + G[i101] * cz) + cx * cy * (G[i110] * fz + G[i111] * cz);
    }
  } else {
    // first update (src/compute_LagScaleDep.f90:264-277)
    const double mm = (double)a.MM[c], nn = (double)a.NN[c];
    F[1] = mm; F[0] = 0.03 * mm;
    F[3] = nn; F[2] = 0.03 * nn;
  }
  // update_interpolag_coeff (src/compute_LagScaleDep.f90:607-640)
  for (int s = 0; s < 2; ++s) {
    const double ab = (double)(s == 0 ? a.LM[c] : a.QN[c]);
    const double bb = (double)(s == 0 ? a.MM[c] : a.NN[c]);
    double fab = F[2 * s], fbb = F[2 * s + 1];
    double tmp = fmax32(fab * fbb, 1e-24f);
    tmp = a.opftdelta * pow_m18(tmp);
    tmp = fmax32(tmp, 1e-24f);
    const double dumfac = a.lagran_dt / tmp;
    const double epsi = dumfac / (1.0 + dumfac);
    const double nab = epsi * ab + (1.0 - epsi) * fab;
    const double nbb = epsi * bb + (1.0 - epsi) * fbb;
    F[2 * s] = fmax32(nab, 1e-24f);
    F[2 * s + 1] = nbb;
  }
  for (int f = 0; f < 4; ++f) a.Fn[f][c] = F[f];
  // compute_dynamic_coeff (src/compute_LagScaleDep.f90:651-707); the beta
  // exponent log(2)/(log(4)-log(2)) is exactly 1 in IEEE double.
  const double c2 = fmax32(F[0] / F[1], 1e-24f);
  const double c4 = fmax32(F[2] / F[3], 1e-24f);
  double b = fmax32(c4 / c2, 0.125f);
  if (k == a.nz + 1) b = 1.0;
  a.beta[c] = (T)b;
  a.cs[c] = (T)fmax32(c2 / b, 1e-24f);
}

// compute_LagScaleDep_coeff on an update step (src/compute_LagScaleDep.f90:237-292)
// on levels 1..nlev (nlev = nz + 1: the whole column).  The procedure is
// local in z apart from the Lagrangian gather, which reads F_* one level
// above: with nlev < nz + 1 the levels above keep their F_*, so level nlev
// gathers from the last full update.  The physical
// nz still decides the boundary semantics (halo planes, beta = 1 at nz + 1).
template <typename T>
static void lasd_update(Les& L, int jtt, int nlev) {
  (void)jtt;
  if (nlev > L.nz + 1) nlev = L.nz + 1;
  const long npt = L.np * nlev;
  // 15 materialised inputs + the 6 S_ij fields
  T* in[21];
  for (int i = 0; i < 9; ++i) in[i] = (T*)L.wk_real[i];
  in[9] = F<T>(L, LES_S11); in[10] = F<T>(L, LES_S12); in[11] = F<T>(L, LES_S13);
  in[12] = F<T>(L, LES_S22); in[13] = F<T>(L, LES_S23); in[14] = F<T>(L, LES_S33);
  for (int i = 15; i < 21; ++i) in[i] = (T*)L.wk_real[i - 6];
  LasdInArgs<T> ia;
  for (int i = 0; i < 9; ++i) ia.o[i] = in[i];
  for (int i = 0; i < 6; ++i) ia.o[9 + i] = in[15 + i];
  ia.u = D(L, LES_U); ia.v = D(L, LES_V); ia.w = D(L, LES_W);
  for (int i = 0; i < 6; ++i) ia.s[i] = in[9 + i];
  ia.smod = (T*)L.wk_real[63];
  ia.np = L.np; ia.nz = L.nz; ia.nlev = nlev;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_lasd_inputs<T>), dim3(nblk(npt)), dim3(LES_BLOCK), 0, L.stream, ia);
  kt_end(L, KT_POINTWISE, (double)npt * (3 * 8.0 + 22.0 * sizeof(T)), "k_lasd_inputs", sizeof(T));
  // test filters at 2 delta and 4 delta from one forward transform per field
  const T* g2 = sizeof(T) == 8 ? (const T*)L.g2_64 : (const T*)L.g2_32;
  const T* g4 = sizeof(T) == 8 ? (const T*)L.g4_64 : (const T*)L.g4_32;
  T* fl2[21];
  T* fl4[21];
  for (int f = 0; f < 21; ++f) { fl2[f] = (T*)L.wk_real[15 + f]; fl4[f] = (T*)L.wk_real[36 + f]; }
  // three fields at a time: inputs in spectral slots 0..2, filtered outputs in 3..8
  static_assert(LES_NSPEC >= 9, "Germano filtering needs 9 spectral workspaces");
  for (int f0 = 0; f0 < 21; f0 += 3) {
    const T* src[3];
    Cplx<T>* sp[3];
    for (int t = 0; t < 3; ++t) { src[t] = in[f0 + t]; sp[t] = (Cplx<T>*)L.wk_spec[t]; }
    rows_r2c<T, T>(L, src, sp, 3, 1, nlev, (T)1);
    Cplx<T>* o2[3];
    Cplx<T>* o4[3];
    for (int t = 0; t < 3; ++t) { o2[t] = (Cplx<T>*)L.wk_spec[3 + 2 * t]; o4[t] = (Cplx<T>*)L.wk_spec[4 + 2 * t]; }
    cols_jobs<T>(L, COP_FILT2, 3, sp, nullptr, o2, o4, nullptr, 2, 1, 1, 1, nlev, g2, g4);
    const Cplx<T>* cs[6];
    T* outs[6];
    for (int t = 0; t < 3; ++t) {
      cs[2 * t] = (Cplx<T>*)L.wk_spec[3 + 2 * t];
      cs[2 * t + 1] = (Cplx<T>*)L.wk_spec[4 + 2 * t];
      outs[2 * t] = fl2[f0 + t];
      outs[2 * t + 1] = fl4[f0 + t];
    }
    rows_c2r<T, T>(L, cs, outs, 6, 1, nlev);
  }
  T* LM = (T*)L.wk_real[57];
  T* MM = (T*)L.wk_real[58];
  T* QN = (T*)L.wk_real[59];
  T* NN = (T*)L.wk_real[60];
  for (int s = 0; s < 2; ++s) {
    GermanoArgs<T> ga;
    for (int f = 0; f < 21; ++f) ga.fl[f] = s == 0 ? fl2[f] : fl4[f];
    ga.AB = s == 0 ? LM : QN;
    ga.BB = s == 0 ? MM : NN;
    ga.tf2 = (T)(s == 0 ? 4.0 : 16.0);
    ga.cst = (T)(2.0 * (L.delta * L.delta));
    ga.np = L.np; ga.nz = L.nz; ga.nlev = nlev;
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_germano<T>), dim3(nblk(npt)), dim3(LES_BLOCK), 0, L.stream, ga);
    kt_end(L, KT_POINTWISE, (double)npt * 23.0 * sizeof(T), "k_germano", sizeof(T));
  }
  // Lagrangian interpolation, relaxation and coefficients.  The gather reads
  // the previous F_* one level around the point, so the update reads one
  // buffer and writes the other: the whole column swaps the two buffers;
  // the near-wall band (nlev < nz + 1) updates F_* in place after copying the
  // levels its gather reads (0..nlev+1) to the second buffer, and the levels
  // above keep their values.
  if (nlev < L.nz + 1) {
    const size_t cnt = (size_t)(nlev + 2) * L.np;
    kt_begin(L, KT_POINTWISE);
    for (int f = 0; f < 4; ++f)
      hipMemcpyAsync(L.Fold[f], D(L, LES_F_LM + f), 8 * cnt, hipMemcpyDeviceToDevice, L.stream);
    kt_end(L, KT_POINTWISE, 4.0 * 16 * cnt, "hipMemcpyAsync F_* band");
  } else {
    for (int f = 0; f < 4; ++f) {
      void* tmp = L.f[LES_F_LM + f];
      L.f[LES_F_LM + f] = (void*)L.Fold[f];
      L.Fold[f] = (double*)tmp;
    }
  }
  LagrArgs<T> la;
  la.ulag = F<T>(L, LES_U_LAG); la.vlag = F<T>(L, LES_V_LAG); la.wlag = F<T>(L, LES_W_LAG);
  for (int f = 0; f < 4; ++f) { la.Fo[f] = L.Fold[f]; la.Fn[f] = D(L, LES_F_LM + f); }
  la.LM = LM; la.MM = MM; la.QN = QN; la.NN = NN;
  la.cs = F<T>(L, LES_CS_OPT2); la.beta = F<T>(L, LES_BETA);
  la.dt = L.P.dt; la.csc = (double)L.P.cs_count;
  la.dx = L.dx; la.dy = L.dy; la.dz = L.dz;
  la.opftdelta = 1.5 * L.delta;
  la.lagran_dt = L.P.dt * (double)L.P.cs_count;
  la.nx = L.nx; la.ny = L.ny; la.nz = L.nz; la.nlev = nlev; la.np = L.np;
  la.interp = L.F_coeff_init;
  la.halo_mode = L.P.halo_mode;
  kt_begin(L, KT_LAGR);
  hipLaunchKernelGGL((k_lagr<T>), dim3(nblk(npt)), dim3(LES_BLOCK), 0, L.stream, la);
  kt_end(L, KT_LAGR, (double)npt * (3.0 * sizeof(T) + 4 * 8.0 + 4.0 * sizeof(T) + 4 * 8.0 + 2.0 * sizeof(T)), "k_lagr",
         sizeof(T));
  L.F_coeff_init = 1;
  const size_t tb = sizeof(T) * (size_t)L.nf;
  kt_begin(L, KT_POINTWISE);
  hipMemsetAsync(L.f[LES_U_LAG], 0, tb, L.stream);
  hipMemsetAsync(L.f[LES_V_LAG], 0, tb, L.stream);
  hipMemsetAsync(L.f[LES_W_LAG], 0, tb, L.stream);
  kt_end(L, KT_POINTWISE, 3.0 * tb, "hipMemsetAsync u_lag");
}

// ---------------------------------------------------------------- plane-averaged dynamic model
// compute_dynamic_mum's plane averages and coefficient
// (src/compute_DynSmag.f90:225-243 with compute_planar_average_wghost,
// src/utility_tools.f90:100-132): for level k, the sums of L:M and M:M over
// the plane in storage order (i fastest, then j) as one running FP64 sum
// each, the order of the reference's loop, divided by nx*ny; CS_Z = num/den
// (0 where den = 0), clipped at 0 below, written over the whole level.
// One block per level.  All threads stage chunks of both planes in LDS with
// coalesced loads, double-buffered so the next chunk is in flight while the
// current one is summed; lane 0 of wave 0 adds the L:M chunk and lane 0 of
// wave 1 the M:M chunk, in order, so the sums equal the sequential ones bit
// for bit.  The add chain (np dependent FP64 adds per level) bounds the
// kernel; it runs once per cs_count steps.
#define LES_PS_CHUNK 2048
template <typename T>
struct DynPlaneArgs {
  const T* LM;
  const T* MM;
  T* cs;
  long np;
  double npl;  // nx*ny as a real
};

template <typename T>
__global__ __launch_bounds__(256) void k_dynsmag_plane(DynPlaneArgs<T> a) {
  constexpr int PER = LES_PS_CHUNK / 256;
  __shared__ double buf[2][2][LES_PS_CHUNK];
  __shared__ double res[2];
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const long base = (long)(blockIdx.x + 1) * a.np;
  const T* s0 = a.LM + base;
  const T* s1 = a.MM + base;
  const long nch = (a.np + LES_PS_CHUNK - 1) / LES_PS_CHUNK;
  double r0[PER], r1[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const long i = (long)m * 256 + t;
    r0[m] = i < a.np ? (double)s0[i] : 0.0;
    r1[m] = i < a.np ? (double)s1[i] : 0.0;
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) { buf[0][0][m * 256 + t] = r0[m]; buf[0][1][m * 256 + t] = r1[m]; }
  __syncthreads();
  double s = 0.0;
  for (long c = 0; c < nch; ++c) {
    const bool more = c + 1 < nch;
    if (more) {
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        const long i = (c + 1) * LES_PS_CHUNK + (long)m * 256 + t;
        r0[m] = i < a.np ? (double)s0[i] : 0.0;
        r1[m] = i < a.np ? (double)s1[i] : 0.0;
      }
    }
    if (lane == 0 && wv < 2) {
      const double* p = buf[c & 1][wv];
      const long rem = a.np - c * LES_PS_CHUNK;
      const int cnt = rem < LES_PS_CHUNK ? (int)rem : LES_PS_CHUNK;
#pragma unroll 16
      for (int i = 0; i < cnt; ++i) s = s + p[i];
    }
    if (more) {
      const int b = (int)((c + 1) & 1);
#pragma unroll
      for (int m = 0; m < PER; ++m) { buf[b][0][m * 256 + t] = r0[m]; buf[b][1][m * 256 + t] = r1[m]; }
    }
    __syncthreads();
  }
  if (lane == 0 && wv < 2) res[wv] = s;
  __syncthreads();
  const double num = res[0] / a.npl, den = res[1] / a.npl;
  double cz = den != 0.0 ? num / den : 0.0;
  cz = cz > 0.0 ? cz : 0.0;
  const T v = (T)cz;
  T* out = a.cs + base;
  for (long q = t; q < a.np; q += 256) out[q] = v;
}

// compute_DynSmag_coeff update (src/compute_DynSmag.f90:76-80) and
// compute_dynamic_mum (:96-243): the Germano identity at the 2-delta test
// filter only.  Inputs, filtering and the pointwise L:M, M:M are those of the
// Lagrangian model's 2-delta half (the reference forms them identically);
// the plane average replaces the Lagrangian relaxation.
template <typename T>
static void dynsmag_update(Les& L) {
  const int nlev = L.nz + 1;
  const long npt = L.np * nlev;
  T* in[21];
  for (int i = 0; i < 9; ++i) in[i] = (T*)L.wk_real[i];
  in[9] = F<T>(L, LES_S11); in[10] = F<T>(L, LES_S12); in[11] = F<T>(L, LES_S13);
  in[12] = F<T>(L, LES_S22); in[13] = F<T>(L, LES_S23); in[14] = F<T>(L, LES_S33);
  for (int i = 15; i < 21; ++i) in[i] = (T*)L.wk_real[i - 6];
  LasdInArgs<T> ia;
  for (int i = 0; i < 9; ++i) ia.o[i] = in[i];
  for (int i = 0; i < 6; ++i) ia.o[9 + i] = in[15 + i];
  ia.u = D(L, LES_U); ia.v = D(L, LES_V); ia.w = D(L, LES_W);
  for (int i = 0; i < 6; ++i) ia.s[i] = in[9 + i];
  ia.smod = (T*)L.wk_real[63];
  ia.np = L.np; ia.nz = L.nz; ia.nlev = nlev;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_lasd_inputs<T>), dim3(nblk(npt)), dim3(LES_BLOCK), 0, L.stream, ia);
  kt_end(L, KT_POINTWISE, (double)npt * (3 * 8.0 + 22.0 * sizeof(T)), "k_lasd_inputs", sizeof(T));
  const T* g2 = sizeof(T) == 8 ? (const T*)L.g2_64 : (const T*)L.g2_32;
  T* fl[21];
  for (int f = 0; f < 21; ++f) fl[f] = (T*)L.wk_real[15 + f];
  for (int f0 = 0; f0 < 21; f0 += 3) {
    const T* src[3];
    Cplx<T>* sp[3];
    const Cplx<T>* cs[3];
    T* outs[3];
    for (int t = 0; t < 3; ++t) {
      src[t] = in[f0 + t];
      sp[t] = (Cplx<T>*)L.wk_spec[t];
      cs[t] = (const Cplx<T>*)L.wk_spec[3 + t];
      outs[t] = fl[f0 + t];
    }
    rows_r2c<T, T>(L, src, sp, 3, 1, nlev, (T)1);
    Cplx<T>* fo[3];
    for (int t = 0; t < 3; ++t) fo[t] = (Cplx<T>*)L.wk_spec[3 + t];
    cols_jobs<T>(L, COP_FILT1, 3, sp, nullptr, fo, nullptr, nullptr, 1, 1, 1, 1, nlev, g2);
    rows_c2r<T, T>(L, cs, outs, 3, 1, nlev);
  }
  GermanoArgs<T> ga;
  for (int f = 0; f < 21; ++f) ga.fl[f] = fl[f];
  ga.AB = (T*)L.wk_real[57];
  ga.BB = (T*)L.wk_real[58];
  ga.tf2 = (T)4.0;
  ga.cst = (T)(2.0 * (L.delta * L.delta));
  ga.np = L.np; ga.nz = L.nz; ga.nlev = nlev;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_germano<T>), dim3(nblk(npt)), dim3(LES_BLOCK), 0, L.stream, ga);
  kt_end(L, KT_POINTWISE, (double)npt * 23.0 * sizeof(T), "k_germano", sizeof(T));
  DynPlaneArgs<T> pa;
  pa.LM = ga.AB; pa.MM = ga.BB; pa.cs = F<T>(L, LES_CS_OPT2);
  pa.np = L.np; pa.npl = (double)(L.nx * L.ny);
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_dynsmag_plane<T>), dim3(nlev), dim3(256), 0, L.stream, pa);
  kt_end(L, KT_POINTWISE, (double)npt * 3.0 * sizeof(T), "k_dynsmag_plane", sizeof(T));
}

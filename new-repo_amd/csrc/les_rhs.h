// Right-hand side of the momentum equations: 3/2-padded rotational
// convection and the divergence of the SGS stress (src/LES_engine.f90:509-572).
// Included by les_gpu.hip.
#pragma once

// Products on the padded grid (src/compute_convective_padd.f90:179-249);
// velocities on the padded grid are the negated ones, as in the reference.
// The nx x ny variants (TRUNC, FSMTHG, NODLSG) form the same products with
// nbp = nx ny and c2 = 1/(nx ny) (src/compute_convective_trunc.f90:177-244).
template <typename T, typename TO>
struct ProdArgs {
  const T *ub, *vb, *wb, *o1, *o2, *o3;
  TO *px, *py, *pz;
  long nbp;
  int nz;
  T c2;
};

template <typename T, typename TO>
__global__ void k_conv_prod(ProdArgs<T, TO> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nbp * a.nz) return;
  const int k = 1 + (int)(t / a.nbp);
  const long q = t % a.nbp;
  const long c = (long)k * a.nbp + q, up = c + a.nbp, dn = c - a.nbp;
  const T c2 = a.c2, h = (T)0.5;
  T x, y, z;
  if (k == 1) {
    x = c2 * (-(a.vb[c] * a.o3[c]) + h * a.wb[up] * a.o2[up]);
    y = c2 * (a.ub[c] * a.o3[c] + h * a.wb[up] * (-a.o1[up]));
    z = (T)0;
  } else {
    x = c2 * (-(a.vb[c] * a.o3[c]) + h * (a.wb[up] * a.o2[up] + a.wb[c] * a.o2[c]));
    y = c2 * (a.ub[c] * a.o3[c] + h * (a.wb[c] * (-a.o1[c]) + a.wb[up] * (-a.o1[up])));
    z = c2 * h * ((a.ub[c] + a.ub[dn]) * (-a.o2[c]) + (a.vb[c] + a.vb[dn]) * a.o1[c]);
  }
  a.px[c] = (TO)x;
  a.py[c] = (TO)y;
  a.pz[c] = (TO)z;
}

// The six convection sources of conv_src on levels 1..nz+1 of the nx x ny
// grid (field = blockIdx.y).
template <typename T>
struct ConvInArgs {
  ConvSrc<T> cs;
  T* dst[6];
  T c;
  long np;
  int nlev;
};

template <typename T>
__global__ void k_conv_sources(ConvInArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nlev) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  a.dst[blockIdx.y][(long)k * a.np + q] = conv_src<T, T>(a.cs, blockIdx.y, k, q, a.np, a.c);
}

struct SetPlaneArgs {
  double *a, *b, *c;
  long np;
  int k0, k1;
  double val;
};

__global__ void k_set_planes(SetPlaneArgs s) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.np) return;
  s.a[(long)s.k0 * s.np + q] = s.val; s.b[(long)s.k0 * s.np + q] = s.val; s.c[(long)s.k0 * s.np + q] = s.val;
  s.a[(long)s.k1 * s.np + q] = s.val; s.b[(long)s.k1 * s.np + q] = s.val; s.c[(long)s.k1 * s.np + q] = s.val;
}

// compute_dealias_convec (src/compute_convective_padd.f90:115-262)
// Column-pair tiles of LES_COLRES_TILE_BYTES contiguous bytes per row.
#ifndef LES_COLRES_TILE_BYTES
#define LES_COLRES_TILE_BYTES 64
#endif
// 256 <-> 384: register-resident kernel (k_colres_c16); LES_COLRES16=0
// selects the LDS-engine kernel (k_cols_resample) instead, LES_COLRES16_FP64=0
// for the FP64 path only.
#ifndef LES_COLRES16
#define LES_COLRES16 1
#endif
#ifndef LES_COLRES16_FP64
#define LES_COLRES16_FP64 1
#endif
// the kernel launch_cols_resample runs for the column lengths h1 -> h2
static const char* colres_name(int h1, int h2) {
  return (LES_COLRES16 && ((h1 == 256 && h2 == 384) || (h1 == 384 && h2 == 256))) ? "k_colres_c16" : "k_cols_resample";
}

template <typename T>
static void launch_cols_resample(Les& L, ColResArgs<T>& a, int nf, const FftPlan& p1, const FftPlan& p2) {
  if constexpr (LES_COLRES16 && (sizeof(T) == 4 || LES_COLRES16_FP64)) {
   if ((a.h1 == 256 && a.h2 == 384) || (a.h1 == 384 && a.h2 == 256)) {
    // register-resident 256 <-> 384 (k_colres_c16)
    a.ncp = (a.W / 2 + 15) / 16;
    a.ls = LES_C24_LS;
    const size_t sh = (size_t)16 * LES_C24_LS * sizeof(Cplx<T>);
    const Cplx<T>* t256 = twpool<T>(L) + (p1.n == 256 ? p1.twf : p2.twf);
    const Cplx<T>* t384 = twpool<T>(L) + (p1.n == 384 ? p1.twf : p2.twf);
    const dim3 g((unsigned)(a.ncp * a.nplanes), nf);
    if (a.h1 == 256)
      hipLaunchKernelGGL((k_colres_c16<T, true>), g, dim3(384), sh, L.stream, a, t256, t384);
    else
      hipLaunchKernelGGL((k_colres_c16<T, false>), g, dim3(384), sh, L.stream, a, t256, t384);
    return;
   }
  }
  constexpr int CP = LES_COLRES_TILE_BYTES / 2 / (int)sizeof(T);
  const int nm = a.h1 > a.h2 ? a.h1 : a.h2;
  a.ls = lds_stride(nm, CP, (int)sizeof(Cplx<T>));
  a.ncp = (a.W / 2 + CP - 1) / CP;
  const size_t sh = (size_t)CP * a.ls * sizeof(Cplx<T>);
  hipLaunchKernelGGL((k_cols_resample<T, CP>), dim3((unsigned)(a.ncp * a.nplanes), nf), dim3(LES_BLOCK), sh, L.stream, a,
                     p1, p2, twpool<T>(L));
}

// Real multiplier g on the 2D spectrum of planes 1..nlev of nf real fields:
// forward transforms, g, inverse transforms (unnormalised, as the test
// filter), three fields at a time; dst may alias src.
template <typename T, typename TO>
static void spectral_multiply(Les& L, T* const* src, TO* const* dst, int nf, int nlev, const T* g) {
  for (int f0 = 0; f0 < nf; f0 += 3) {
    const int m = nf - f0 < 3 ? nf - f0 : 3;
    const T* s[3];
    Cplx<T>* sp[3];
    const Cplx<T>* cs[3];
    TO* o[3];
    for (int t = 0; t < m; ++t) {
      s[t] = src[f0 + t];
      sp[t] = (Cplx<T>*)L.wk_spec[t];
      cs[t] = (const Cplx<T>*)L.wk_spec[3 + t];
      o[t] = dst[f0 + t];
    }
    rows_r2c<T, T>(L, s, sp, m, 1, nlev, (T)1);
    Cplx<T>* fo[3];
    for (int t = 0; t < m; ++t) fo[t] = (Cplx<T>*)L.wk_spec[3 + t];
    cols_jobs<T>(L, COP_FILT1, m, sp, nullptr, fo, nullptr, nullptr, 1, 1, 1, 1, nlev, g);
    rows_c2r<T, TO>(L, cs, o, m, 1, nlev);
  }
}

// compute_dealias_convec of the other dealiasing options, chosen at compile
// time in the reference (DSLG, src/Makefile:17-40): 1 TRUNC
// (src/compute_convective_trunc.f90:115-247), 2 FSMTHG
// (src/compute_convective_Fsmthg.f90:141-264), 3 NODLSG
// (src/compute_convective_nodlsg.f90:88-205).  Sources and products stay on
// the nx x ny grid with const = 1/(nx ny) in both places.  TRUNC and FSMTHG
// multiply the spectra of the sources (levels 1..nz+1; level 0 enters no
// product) by ga and of the products (levels 1..nz) by gb; NODLSG forms the
// products of the scaled sources directly, so its result carries the factor
// const^3 of the reference's code.  The sources use the padded path's
// buffers (wk_big, plane stride nx ny), the products wk_prod.
template <typename T>
static void convection_nxny(Les& L, double* cx, double* cy, double* cz) {
  const int nz = L.nz;
  const long np = L.np;
  const T c = (T)(1.0 / (double)(L.nx * L.ny));
  const bool filt = L.P.dealias != 3;
  T* srcf[6];
  for (int f = 0; f < 6; ++f) srcf[f] = (T*)L.wk_big[f];
  {
    ConvInArgs<T> a;
    a.cs.u = D(L, LES_U); a.cs.v = D(L, LES_V); a.cs.w = D(L, LES_W);
    a.cs.dudy = F<T>(L, LES_DUDY); a.cs.dudz = F<T>(L, LES_DUDZ);
    a.cs.dvdx = F<T>(L, LES_DVDX); a.cs.dvdz = F<T>(L, LES_DVDZ);
    a.cs.dwdx = F<T>(L, LES_DWDX); a.cs.dwdy = F<T>(L, LES_DWDY);
    for (int f = 0; f < 6; ++f) a.dst[f] = srcf[f];
    a.c = c; a.np = np; a.nlev = nz + 1;
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_conv_sources<T>), dim3(nblk(np * (nz + 1)), 6), dim3(LES_BLOCK), 0, L.stream, a);
    kt_end(L, KT_POINTWISE, (double)np * (nz + 1) * (3.0 * 8 + 8.0 * sizeof(T) + 6.0 * sizeof(T)), "k_conv_sources", sizeof(T));
  }
  const T* ga = sizeof(T) == 8 ? (const T*)L.ga_64 : (const T*)L.ga_32;
  const T* gb = sizeof(T) == 8 ? (const T*)L.gb_64 : (const T*)L.gb_32;
  if (filt) spectral_multiply<T, T>(L, srcf, srcf, 6, nz + 1, ga);
  T* pr[3] = {(T*)L.wk_prod[0], (T*)L.wk_prod[1], (T*)L.wk_prod[2]};
  if (filt) {
    ProdArgs<T, T> a;
    a.ub = srcf[0]; a.vb = srcf[1]; a.wb = srcf[2]; a.o1 = srcf[3]; a.o2 = srcf[4]; a.o3 = srcf[5];
    a.px = pr[0]; a.py = pr[1]; a.pz = pr[2];
    a.nbp = np; a.nz = nz; a.c2 = c;
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_conv_prod<T, T>), dim3(nblk(np * nz)), dim3(LES_BLOCK), 0, L.stream, a);
    kt_end(L, KT_POINTWISE, 9.0 * np * nz * sizeof(T), "k_conv_prod", sizeof(T), sizeof(T), "nx x ny");
    double* out[3] = {cx, cy, cz};
    spectral_multiply<T, double>(L, pr, out, 3, nz, gb);
  } else {
    ProdArgs<T, double> a;
    a.ub = srcf[0]; a.vb = srcf[1]; a.wb = srcf[2]; a.o1 = srcf[3]; a.o2 = srcf[4]; a.o3 = srcf[5];
    a.px = cx; a.py = cy; a.pz = cz;
    a.nbp = np; a.nz = nz; a.c2 = c;
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_conv_prod<T, double>), dim3(nblk(np * nz)), dim3(LES_BLOCK), 0, L.stream, a);
    kt_end(L, KT_POINTWISE, np * nz * (6.0 * sizeof(T) + 3.0 * 8), "k_conv_prod", sizeof(T), 8, "nx x ny");
  }
  SetPlaneArgs sp;
  sp.a = cx; sp.b = cy; sp.c = cz; sp.np = np; sp.k0 = 0; sp.k1 = nz + 1; sp.val = BOGUS;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_set_planes, dim3(nblk(np)), dim3(LES_BLOCK), 0, L.stream, sp);
  kt_end(L, KT_POINTWISE, 6.0 * 8 * np, "k_set_planes");
}

template <typename T>
static void convection(Les& L, double* cx, double* cy, double* cz) {
  if (L.P.dealias != 0) {
    convection_nxny<T>(L, cx, cy, cz);
    return;
  }
  const int nx = L.nx, ny = L.ny, nz = L.nz, nx2 = L.nx2, ny2 = L.ny2;
  const T c = (T)(1.0 / (double)(nx * ny));
  // pad in x: planes 1..nz+1 of u, v, w and the vorticities
  {
    RowResArgs<T, T, T> a;
    a.cs.u = D(L, LES_U); a.cs.v = D(L, LES_V); a.cs.w = D(L, LES_W);
    a.cs.dudy = F<T>(L, LES_DUDY); a.cs.dudz = F<T>(L, LES_DUDZ);
    a.cs.dvdx = F<T>(L, LES_DVDX); a.cs.dvdz = F<T>(L, LES_DVDZ);
    a.cs.dwdx = F<T>(L, LES_DWDX); a.cs.dwdy = F<T>(L, LES_DWDY);
    a.cconv = c;
    a.src_mode = 1;
    for (int f = 0; f < 6; ++f) { a.src[f] = nullptr; a.dst[f] = (T*)L.wk_hyb[f]; }
    a.n1 = nx; a.n2 = nx2; a.ny = ny; a.k0 = 1; a.nplanes = nz + 1;
    a.src_plane = (long)nx * ny; a.dst_plane = (long)nx2 * ny;
    a.ls = res_line<T>(nx, nx2);
    long jobs = (long)a.nplanes * (ny / 2);
    size_t sh = (size_t)kWavesPerBlock * a.ls * sizeof(Cplx<T>);
    kt_begin(L, KT_ROWS_RES);
    hipLaunchKernelGGL((k_rows_resample<T, T, T>), dim3((unsigned)((jobs + kWavesPerBlock - 1) / kWavesPerBlock), 6),
                       dim3(LES_BLOCK), sh, L.stream, a, L.px, L.px2, twpool<T>(L));
    kt_end(L, KT_ROWS_RES, (double)a.nplanes * L.np * (3.0 * 8 + 6.0 * sizeof(T)) + 6.0 * a.nplanes * nx2 * ny * sizeof(T),
           "k_rows_resample", sizeof(T), sizeof(T), "pad");
  }
  // pad in y
  {
    ColResArgs<T> a;
    for (int f = 0; f < 6; ++f) { a.src[f] = (const T*)L.wk_hyb[f]; a.dst[f] = (T*)L.wk_big[f]; }
    a.W = nx2; a.h1 = ny; a.h2 = ny2; a.k0 = 1; a.nplanes = nz + 1;
    a.src_plane = (long)nx2 * ny; a.dst_plane = L.nbp;
    kt_begin(L, KT_COLS_RES);
    launch_cols_resample<T>(L, a, 6, L.py, L.py2);
    kt_end(L, KT_COLS_RES, 6.0 * a.nplanes * nx2 * (double)(ny + ny2) * sizeof(T), colres_name(ny, ny2), sizeof(T), 0, "pad");
  }
  // products
  {
    ProdArgs<T, T> a;
    a.ub = (const T*)L.wk_big[0]; a.vb = (const T*)L.wk_big[1]; a.wb = (const T*)L.wk_big[2];
    a.o1 = (const T*)L.wk_big[3]; a.o2 = (const T*)L.wk_big[4]; a.o3 = (const T*)L.wk_big[5];
    a.px = (T*)L.wk_prod[0]; a.py = (T*)L.wk_prod[1]; a.pz = (T*)L.wk_prod[2];
    a.nbp = L.nbp; a.nz = nz;
    a.c2 = (T)(1.0 / (double)(nx2 * ny2));
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_conv_prod<T, T>), dim3(nblk(L.nbp * nz)), dim3(LES_BLOCK), 0, L.stream, a);
    kt_end(L, KT_POINTWISE, 9.0 * L.nbp * nz * sizeof(T), "k_conv_prod", sizeof(T), sizeof(T), "padded");
  }
  // truncate in y
  {
    ColResArgs<T> a;
    for (int f = 0; f < 3; ++f) { a.src[f] = (const T*)L.wk_prod[f]; a.dst[f] = (T*)L.wk_hyb3[f]; }
    a.W = nx2; a.h1 = ny2; a.h2 = ny; a.k0 = 1; a.nplanes = nz;
    a.src_plane = L.nbp; a.dst_plane = (long)nx2 * ny;
    kt_begin(L, KT_COLS_RES);
    launch_cols_resample<T>(L, a, 3, L.py2, L.py);
    kt_end(L, KT_COLS_RES, 3.0 * a.nplanes * nx2 * (double)(ny + ny2) * sizeof(T), colres_name(ny2, ny), sizeof(T), 0, "truncate");
  }
  // truncate in x, straight into rhs
  {
    RowResArgs<T, T, double> a;
    a.src[0] = (const T*)L.wk_hyb3[0]; a.src[1] = (const T*)L.wk_hyb3[1]; a.src[2] = (const T*)L.wk_hyb3[2];
    a.dst[0] = cx; a.dst[1] = cy; a.dst[2] = cz;
    a.cconv = (T)0;
    a.src_mode = 0;
    a.n1 = nx2; a.n2 = nx; a.ny = ny; a.k0 = 1; a.nplanes = nz;
    a.src_plane = (long)nx2 * ny; a.dst_plane = L.np;
    a.ls = res_line<T>(nx2, nx);
    long jobs = (long)nz * (ny / 2);
    size_t sh = (size_t)kWavesPerBlock * a.ls * sizeof(Cplx<T>);
    kt_begin(L, KT_ROWS_RES);
    hipLaunchKernelGGL((k_rows_resample<T, T, double>), dim3((unsigned)((jobs + kWavesPerBlock - 1) / kWavesPerBlock), 3),
                       dim3(LES_BLOCK), sh, L.stream, a, L.px2, L.px, twpool<T>(L));
    kt_end(L, KT_ROWS_RES, 3.0 * nz * ny * ((double)nx2 * sizeof(T) + (double)nx * 8), "k_rows_resample", sizeof(T),
           sizeof(T), "truncate");
  }
  SetPlaneArgs sp;
  sp.a = cx; sp.b = cy; sp.c = cz; sp.np = L.np; sp.k0 = 0; sp.k1 = nz + 1; sp.val = BOGUS;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_set_planes, dim3(nblk(L.np)), dim3(LES_BLOCK), 0, L.stream, sp);
  kt_end(L, KT_POINTWISE, 6.0 * 8 * L.np, "k_set_planes");
}

// rhs(k) += d/dx tx + d/dy ty + d/dz tz for k = 1..nz, plus the forcing of
// compute_RHS_adv; optionally the pressure BC layers
// (src/compute_derivatives.f90:545-574; src/LES_engine.f90:547-562).
template <typename T>
struct DivEpiArgs {
  double* rhs;
  const T* tz;
  int stag, save_wk;
  double *bottom_wk, *top_wk;
  double cz;
  long np;
  int nz;
  int forcing;  // 1: x forcing, 2: y forcing
  double pf, coriol, geo;
  const double* vel;  // v for x, u for y (Coriolis)
  int coriolis;
};

// rhs(k) += (horizontal divergence) + dz(tz) + forcing, applied as the store
// of the inverse x transform of the divergence spectrum, so the divergence
// never goes through HBM; levels 1..nz+1 (nz+1 only for top_wk).
template <typename T>
__device__ __forceinline__ void div_epi_point(const DivEpiArgs<T>& a, int k, long q, double dh) {
  const long c = (long)k * a.np + q;
  if (k <= a.nz) {
    const double dzt = a.stag ? a.cz * ((double)a.tz[c + a.np] - (double)a.tz[c])
                              : a.cz * ((double)a.tz[c] - (double)a.tz[c - a.np]);
    double r = a.rhs[c] + dh + dzt;
    if (a.forcing == 1) {
      r = r + a.pf;
      if (a.coriolis) r = r + a.coriol * a.vel[c] - a.coriol * a.geo;
    } else if (a.forcing == 2) {
      r = r + a.pf;
      if (a.coriolis) r = r - a.coriol * a.vel[c] + a.coriol * a.geo;
    }
    a.rhs[c] = r;
    if (a.save_wk && k == 1) a.bottom_wk[q] = dh + a.cz * ((double)a.tz[c] - (double)a.tz[q]);
  } else if (a.save_wk) {
    a.top_wk[q] = dh;
  }
}

template <typename T>
struct DivEpi3 {
  DivEpiArgs<T> c[3];  // per component (blockIdx.y)
};

// FP32: held to 80 VGPRs (6 waves per SIMD instead of 5, no spills).
template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
k_rows_c2r_div(RowC2RArgs<T, T> a, DivEpi3<T> ep, FftPlan pl, const Cplx<T>* __restrict__ tw) {
  const DivEpiArgs<T>& e = ep.c[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  Cplx<T>* z = reinterpret_cast<Cplx<T>*>(smem_raw) + wave * a.ls;
  const long job = (long)blockIdx.x * (blockDim.x / LES_WAVE) + wave;
  const int npairs = a.ny / 2;
  if (job >= (long)a.nplanes * npairs) return;
  const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
  if (pl.n == 256) {
    TwReg<T, 256> tr;
    tw_load(tr, tw + pl.twf, lane);
    rows_c2r_line(a, z, lane, k, jp, [&](Cplx<T>* p, bool inv) { wave_fft_t<T, 256, row_layout<T>(256)>(p, tr, inv, lane); });
  } else {
    rows_c2r_line(a, z, lane, k, jp, [&](Cplx<T>* p, bool inv) { wave_fft(p, pl, tw, inv, lane); });
  }
  const long q0 = (long)(2 * jp) * a.nx;
  for (int n = lane; n < a.nx; n += LES_WAVE) {
    const Cplx<T> v = z[n];
    div_epi_point(e, k, q0 + n, (double)v.x);
    div_epi_point(e, k, q0 + a.nx + n, (double)v.y);
  }
}

template <typename T>
static DivEpiArgs<T> div_epi_args(Les& L, int rhs_id, int tz, int stag, int save_wk, int forcing) {
  DivEpiArgs<T> a;
  a.rhs = D(L, rhs_id); a.tz = F<T>(L, tz);
  a.stag = stag; a.save_wk = save_wk;
  a.bottom_wk = L.layer[LES_BOTTOM_WK]; a.top_wk = L.layer[LES_TOP_WK];
  a.cz = 1.0 / L.dz; a.np = L.np; a.nz = L.nz;
  a.forcing = forcing;
  a.pf = forcing == 1 ? L.P.pfx : L.P.pfy;
  a.coriol = L.P.coriol;
  a.geo = forcing == 1 ? L.P.vg : L.P.ug;
  a.vel = forcing == 1 ? D(L, LES_V) : D(L, LES_U);
  a.coriolis = L.P.coriolis;
  return a;
}

// Inverse x transform of nf divergence spectra (planes 1..nz+1) with the
// rhs update of each component as the store.
template <typename T>
static void rows_c2r_div(Les& L, const Cplx<T>* const* src, const DivEpi3<T>& ep, int nf) {
  RowC2RArgs<T, T> ca;
  for (int f = 0; f < nf; ++f) { ca.src[f] = src[f]; ca.dst[f] = nullptr; }
  ca.nx = L.nx; ca.ny = L.ny; ca.lhx = L.lhx; ca.k0 = 1; ca.nplanes = L.nz + 1;
  ca.ls = row_line<T>(L.nx);
  const long jobs = (long)ca.nplanes * (L.ny / 2);
  const size_t sh = (size_t)kWavesPerBlock * ca.ls * sizeof(Cplx<T>);
  double extra = 0.0;
  for (int f = 0; f < nf; ++f) extra += (ep.c[f].coriolis && ep.c[f].forcing) ? 8.0 : 0.0;
  kt_begin(L, KT_ROWS_C2R);
  hipLaunchKernelGGL((k_rows_c2r_div<T>), dim3((unsigned)((jobs + kWavesPerBlock - 1) / kWavesPerBlock), nf),
                     dim3(LES_BLOCK), sh, L.stream, ca, ep, L.px, twpool<T>(L));
  kt_end(L, KT_ROWS_C2R, (double)nf * (L.nz + 1) * L.ny * ((double)L.lhx * sizeof(Cplx<T>)) +
                             (double)L.np * L.nz * (nf * (16.0 + sizeof(T)) + extra),
         "k_rows_c2r_div", sizeof(T));
}

template <typename T>
static void div_stress(Les& L, int rhs_id, int tx, int ty, int tz, int stag, int save_wk, int forcing) {
  Cplx<T>* s0 = (Cplx<T>*)L.wk_spec[0];
  Cplx<T>* s1 = (Cplx<T>*)L.wk_spec[1];
  Cplx<T>* s2 = (Cplx<T>*)L.wk_spec[2];
  const T* src[2] = {F<T>(L, tx), F<T>(L, ty)};
  Cplx<T>* dst[2] = {s0, s1};
  const T c = (T)(1.0 / (double)(L.nx * L.ny));
  rows_r2c<T, T>(L, src, dst, 2, 1, L.nz + 1, c);
  cols<T>(L, COP_DIV, s0, s1, s2, nullptr, nullptr, 1, 1, 1, 1, L.nz + 1);
  DivEpi3<T> ep;
  ep.c[0] = div_epi_args<T>(L, rhs_id, tz, stag, save_wk, forcing);
  const Cplx<T>* cs[1] = {s2};
  rows_c2r_div<T>(L, cs, ep, 1);
}

// The three components together: the five distinct stress fields are
// transformed once (txy enters the x and y divergences), then one inverse
// launch updates rhsx, rhsy, rhsz.  The arithmetic per component is that of
// div_stress.
// First half: the spectra of the three horizontal divergences (needs the
// stresses only).
template <typename T>
static void div_stress_spectra(Les& L) {
  Cplx<T>* sp[8];
  for (int i = 0; i < 8; ++i) sp[i] = (Cplx<T>*)L.wk_spec[i];
  static_assert(LES_NSPEC >= 8, "div_stress_xyz needs 8 spectral workspaces");
  const T* src[5] = {F<T>(L, LES_TXX), F<T>(L, LES_TXY), F<T>(L, LES_TYY), F<T>(L, LES_TXZ), F<T>(L, LES_TYZ)};
  const T c = (T)(1.0 / (double)(L.nx * L.ny));
  rows_r2c<T, T>(L, src, sp, 5, 1, L.nz + 1, c);
  const Cplx<T>* a0[3] = {sp[0], sp[1], sp[3]};
  const Cplx<T>* a1[3] = {sp[1], sp[2], sp[4]};
  Cplx<T>* d[3] = {sp[5], sp[6], sp[7]};
  cols_jobs<T>(L, COP_DIV, 3, a0, a1, d, nullptr, nullptr, 1, 1, 1, 1, L.nz + 1);
}

// Second half: inverse transforms with the rhs update as their store.
template <typename T>
static void div_stress_apply(Les& L) {
  Cplx<T>* sp[8];
  for (int i = 0; i < 8; ++i) sp[i] = (Cplx<T>*)L.wk_spec[i];
  DivEpi3<T> ep;
  const int rhs[3] = {LES_RHSX, LES_RHSY, LES_RHSZ}, tz[3] = {LES_TXZ, LES_TYZ, LES_TZZ};
  for (int f = 0; f < 3; ++f) ep.c[f] = div_epi_args<T>(L, rhs[f], tz[f], f < 2 ? 1 : 0, f == 2 ? 1 : 0, f < 2 ? f + 1 : 0);
  const Cplx<T>* cs[3] = {sp[5], sp[6], sp[7]};
  rows_c2r_div<T>(L, cs, ep, 3);
}

template <typename T>
static void div_stress_xyz(Les& L) {
  div_stress_spectra<T>(L);
  div_stress_apply<T>(L);
}

// First part of compute_RHS_adv: rhs_f <- rhs (src/LES_engine.f90:517-522)
// as a buffer swap, the new rhs being fully overwritten by the convective term.
template <typename T>
static void rhs_convection(Les& L) {
  for (int c = 0; c < 3; ++c) {
    void* t = L.f[LES_RHSX + c];
    L.f[LES_RHSX + c] = L.f[LES_RHSX_F + c];
    L.f[LES_RHSX_F + c] = t;
  }
  convection<T>(L, D(L, LES_RHSX), D(L, LES_RHSY), D(L, LES_RHSZ));
}

template <typename T>
static void compute_rhs_adv(Les& L) {
  rhs_convection<T>(L);
  div_stress_xyz<T>(L);
}

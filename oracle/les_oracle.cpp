// LES time-step oracle -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement, in plain C++17 + OpenMP and FP64 throughout, of one time
// step of the reference pseudo-spectral LES (fmargairaz/new-repo), 1-rank
// semantics.  Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark load this library; the product (the HIP library under
// new-repo_amd/) never links or calls it.
//
// What it follows (reference file:line):
//   step order, AB2, projection, BCs ........ src/LES_engine.f90:189-353,449-572
//   wavenumbers (Nyquist zeroed) .............. src/fft_engine.f90:97-145
//   2D FFTs over z-planes ..................... src/decomp_2d_fft_2d.f90:397-577
//   grad_velocity_filt, ddz_uv, ddz_w ......... src/compute_derivatives.f90:94-137,375-441
//   div_stress (+ pressure BC layers) ......... src/compute_derivatives.f90:464-577
//   padded convection ......................... src/compute_convective_padd.f90:115-388
//   pressure + tridiagonal + gradient ......... src/compute_pressure.f90:116-372,448-493,544-619
//   test filter (spectral cutoff) ............. src/testfilter.f90:101-210
//   S_ij, nu_t, tau_ij ........................ src/compute_sgs.f90:105-335
//   LASD schedule, Germano, Lagrangian ........ src/compute_LagScaleDep.f90:204-707
//   1-rank halo of the Lagrangian stencil ..... src/decomp_2d_custom.f90:902-1024
//   wall law (obukhov, wallstress) ............ src/compute_wall_law.f90:107-254
//   set_bc ..................................... src/boundary_condition_mpi.f90:33-63
//   running statistics ........................ src/IO_statistics.f90:266-513
//   profiles of the running averages .......... matlab_src/get_data_struct_RAV.m:21-35,
//                                               matlab_src/get_profile_struct.m:9-59,
//                                               matlab_src/compute_Cij.m:1-36
//   running diagnostics ....................... src/running_diagnostics.f90:80-305
//   TKE budget sums ........................... src/tke_budget1D.f90:209-440
//
// Deliberate restatement choices (documented in the repository's design notes):
//   * FFTs use the algorithm in fftcore.h, not FFTW; they are unnormalised like
//     FFTW's and pinned against numpy in tests/.
//   * In div_stress the x- and y-derivative spectra are summed before one
//     inverse transform (linear, same result up to rounding).
//   * The LASD relaxation time uses tmp**(-1/8) = 1/sqrt(sqrt(sqrt(tmp))),
//     built from correctly rounded operations only.
//   * log(0.5 dz / z0) is evaluated once per surface point at set-up.
// Everything else reproduces the reference's operation order, including the
// float32 casts inside the FP64 LASD (src/compute_LagScaleDep.f90:624-690).
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "fftcore.h"

// Planted mistakes, for the oracle's own pinning tests only.  Every library
// that the tests, the smoke check and the benchmark load is built with
// ORC_PLANT = 0, where each PLANT(n) below is the constant false and its
// branch is compiled out.  tests/test_oracle_pins.py builds this file again
// with -DORC_PLANT=n, one library per n, and checks that at least one of the
// pinning tests named for n fails on that build: each n is a plausible
// transcription mistake (a wrong exponent, constant, sign, index or order).
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(n) (ORC_PLANT == (n))

namespace orc {

static const double BOGUS = -1.23456789e20;
static const double PI = 3.1415926535897932384626433;
static const double KVONK = 0.4;

struct Params {
  int nx, ny, nz;
  int sgs_model;    // 0 constant Smagorinsky, 1 plane-averaged dynamic Smagorinsky, 2 Lagrangian scale-dependent, 3 dns
  int cs_count, dyn_init, lagran_init;
  int wall_model;   // 0 wall_filter (2-delta filtered first level), 1 wall_inst
  int ubc_u_neu, ubc_v_neu;  // top BC type for u, v (1 Neumann, 0 Dirichlet)
  int halo_mode;    // 0 one-rank reference halo, 1 periodic
  int coriolis;
  int dealias;      // convection: 0 PADD (3/2 rule), 1 TRUNC, 2 FSMTHG, 3 NODLSG
  double lx, ly, lz, dt, z0, pfx, pfy, cs0, coriol, ug, vg;
  double ubc_u, ubc_v, ubc_w;
  double re;        // Reynolds number (sgs_model 3, dns)
};

struct Oracle {
  Params P;
  int nx, ny, nz, nzg, lhx;
  size_t np;   // nx*ny
  size_t nf;   // nx*ny*nzg
  double dx, dy, dz, delta;
  Plan px, py, px2, py2;
  std::vector<double> kkx, kky, kk2, G2, G4;
  std::vector<double> GA, GB;   // TRUNC / FSMTHG multipliers of the sources and of the products
  std::vector<double> logterm;  // log(0.5 dz / z0) per surface point
  // fields (nx, ny, 0:nz+1), x fastest
  std::vector<double> u, v, w, dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz;
  std::vector<double> p, dpdx, dpdy, dpdz, rhsx, rhsy, rhsz, rhsx_f, rhsy_f, rhsz_f;
  std::vector<double> txx, txy, txz, tyy, tyz, tzz;
  std::vector<double> cs_opt2, nu_t, F_LM, F_MM, F_QN, F_NN, beta, u_lag, v_lag, w_lag;
  std::vector<double> lbc_u, lbc_v, bottom_wk, top_wk, ustar;
  std::vector<double> s11, s12, s13, s22, s23, s33;
  int jt_total = 0;
  int jt = 0;  // steps since initialisation (the reference's jt, reset at every start)
  bool F_coeff_init = false;
  bool lagran_flag = false;
  // running statistics: 26 sums of (nx, ny, nz) and the u* sum
  std::vector<double> st_sum, st_ustar;
  int st_iters = 0;
  // TKE budget: 61 per-level sums of plane averages (nz each)
  std::vector<double> tk_sum;
  std::vector<double> tk_abs;  // the same sums of |value|: the scale of each sum's rounding

  size_t I(int i, int j, int k) const { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k); }
  size_t S(int kx, int j, int k) const { return (size_t)kx + (size_t)lhx * ((size_t)j + (size_t)ny * (size_t)k); }
};

static Oracle* g = nullptr;

// ---------------------------------------------------------------- 2D transforms
// Forward 2D transform of one real plane (nx, ny) scaled by `scale` into the
// half spectrum (lhx, ny): x rows in pairs, then y columns.
static void fft2_fwd(const Oracle& o, const double* in, Cx* out, double scale) {
  int nx = o.nx, ny = o.ny, lhx = o.lhx;
  std::vector<Cx> z(nx > ny ? nx : ny), s(z.size()), A(lhx), B(lhx);
  for (int j = 0; j < ny; j += 2) {
    for (int n = 0; n < nx; ++n) z[n] = {scale * in[n + (size_t)nx * j], scale * in[n + (size_t)nx * (j + 1)]};
    o.px.exec(z.data(), s.data(), false);
    unpack_pair(z.data(), nx, A.data(), B.data());
    for (int k = 0; k < lhx; ++k) {
      out[k + (size_t)lhx * j] = A[k];
      out[k + (size_t)lhx * (j + 1)] = B[k];
    }
  }
  for (int k = 0; k < lhx; ++k) {
    for (int j = 0; j < ny; ++j) z[j] = out[k + (size_t)lhx * j];
    o.py.exec(z.data(), s.data(), false);
    for (int j = 0; j < ny; ++j) out[k + (size_t)lhx * j] = z[j];
  }
}

// Inverse y pass over all columns of a half spectrum (in place).
static void fft2_inv_y(const Oracle& o, Cx* sp) {
  int ny = o.ny, lhx = o.lhx;
  std::vector<Cx> z(ny), s(ny);
  for (int k = 0; k < lhx; ++k) {
    for (int j = 0; j < ny; ++j) z[j] = sp[k + (size_t)lhx * j];
    o.py.exec(z.data(), s.data(), true);
    for (int j = 0; j < ny; ++j) sp[k + (size_t)lhx * j] = z[j];
  }
}

// Inverse x pass (rows in pairs, c2r semantics) of a y-inverted half spectrum.
static void fft2_inv_x(const Oracle& o, const Cx* sp, double* out) {
  int nx = o.nx, ny = o.ny, lhx = o.lhx;
  std::vector<Cx> z(nx), s(nx);
  for (int j = 0; j < ny; j += 2) {
    pack_pair(sp + (size_t)lhx * j, sp + (size_t)lhx * (j + 1), nx, z.data());
    o.px.exec(z.data(), s.data(), true);
    for (int n = 0; n < nx; ++n) {
      out[n + (size_t)nx * j] = z[n].x;
      out[n + (size_t)nx * (j + 1)] = z[n].y;
    }
  }
}

static void fft2_inv(const Oracle& o, Cx* sp, double* out) {
  fft2_inv_y(o, sp);
  fft2_inv_x(o, sp, out);
}

static inline bool nyq(const Oracle& o, int kx, int j) { return kx == o.lhx - 1 || j == o.ny / 2; }
static inline Cx ik(double k, Cx a) { return {-(k * a.y), k * a.x}; }

// ---------------------------------------------------------------- set-up
static void build_tables(Oracle& o) {
  int nx = o.nx, ny = o.ny, lhx = o.lhx;
  o.kkx.assign((size_t)lhx * ny, 0.0);
  o.kky.assign((size_t)lhx * ny, 0.0);
  o.kk2.assign((size_t)lhx * ny, 0.0);
  double cx = 2.0 * PI / o.P.lx, cy = 2.0 * PI / o.P.ly;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < lhx; ++i) {
      double kx = (double)i;
      int m = (j + ny / 2) % ny;
      double ky = (double)(m - ny / 2);
      if (i == lhx - 1) { kx = 0.0; ky = 0.0; }
      if (j == ny / 2) { kx = 0.0; ky = 0.0; }
      kx = cx * kx;
      ky = cy * ky;
      o.kkx[i + (size_t)lhx * j] = kx;
      o.kky[i + (size_t)lhx * j] = ky;
      o.kk2[i + (size_t)lhx * j] = kx * kx + ky * ky;
    }
  // test filters (src/testfilter.f90:101-131), alpha = 2 and 4
  for (int a = 0; a < 2; ++a) {
    double alpha = a == 0 ? 2.0 : 4.0;
    std::vector<double>& G = a == 0 ? o.G2 : o.G4;
    G.assign((size_t)lhx * ny, 1.0 / (double)(nx * ny));
    double delta = alpha * std::sqrt(o.dx * o.dy);
    double t = PI / delta;
    double kc2 = t * t;
    for (size_t q = 0; q < G.size(); ++q)
      if (o.kk2[q] >= kc2) G[q] = 0.0;
    for (int j = 0; j < ny; ++j) G[(lhx - 1) + (size_t)lhx * j] = 0.0;
    for (int i = 0; i < lhx; ++i) G[i + (size_t)lhx * (ny / 2)] = 0.0;
  }
  // Dealiasing multipliers of TRUNC and FSMTHG on the half spectrum.  The
  // reference filters each real line along x, then along y, with 1D real
  // transforms, zeroing or scaling entry m of the half spectrum of the line;
  // the two stages together multiply the 2D spectrum by Fx(i) Fy(|ky|).
  //  TRUNC sources: scr(tt:lh) = 0 with tt = ceiling(n/3), i.e. modes m >= tt-1
  //    (src/compute_convective_trunc.f90:279-304);
  //  TRUNC products: tt = ceiling(2 lh / 3) (:322-348), lh = n/2 + 1;
  //  FSMTHG sources and products: F(m) = exp(-36 (m / (lh-1))^36), F(lh-1) = 0
  //    (src/compute_convective_Fsmthg.f90:82-100, 281-299, 339-359).
  if (o.P.dealias == 1 || o.P.dealias == 2) {
    int lhy = ny / 2 + 1;
    for (int t = 0; t < 2; ++t) {
      std::vector<double>& G = t == 0 ? o.GA : o.GB;
      G.assign((size_t)lhx * ny, 0.0);
      std::vector<double> fx(lhx), fy(lhy);
      if (o.P.dealias == 1) {
        int ttx = t == 0 ? (nx + 2) / 3 : (2 * lhx + 2) / 3;
        int tty = t == 0 ? (ny + 2) / 3 : (2 * lhy + 2) / 3;
        for (int m = 0; m < lhx; ++m) fx[m] = m <= ttx - 2 ? 1.0 : 0.0;
        for (int m = 0; m < lhy; ++m) fy[m] = m <= tty - 2 ? 1.0 : 0.0;
      } else {
        for (int m = 0; m < lhx; ++m) fx[m] = std::exp(-36.0 * std::pow((double)m / (double)(lhx - 1), 36.0));
        for (int m = 0; m < lhy; ++m) fy[m] = std::exp(-36.0 * std::pow((double)m / (double)(lhy - 1), 36.0));
        fx[lhx - 1] = 0.0;
        fy[lhy - 1] = 0.0;
      }
      for (int j = 0; j < ny; ++j) {
        int my = j <= ny / 2 ? j : ny - j;
        for (int i = 0; i < lhx; ++i) G[i + (size_t)lhx * j] = fx[i] * fy[my];
      }
    }
  }
}

// ---------------------------------------------------------------- operators
// Filter one real plane with G in spectral space (test_filter_layer).
static void filter_plane(const Oracle& o, const double* in, double* out, const std::vector<double>& G) {
  std::vector<Cx> sp((size_t)o.lhx * o.ny);
  fft2_fwd(o, in, sp.data(), 1.0);
  for (size_t q = 0; q < sp.size(); ++q) sp[q] = cscale(sp[q], G[q]);
  fft2_inv(o, sp.data(), out);
}

// obukhov + wallstress (src/compute_wall_law.f90:107-254)
static void wall_law(Oracle& o) {
  size_t np = o.np;
  std::vector<double> u1(np), v1(np);
  if (o.P.wall_model == 0 || PLANT(35)) {
    filter_plane(o, &o.u[o.I(0, 0, 1)], u1.data(), o.G2);
    filter_plane(o, &o.v[o.I(0, 0, 1)], v1.data(), o.G2);
  } else {
    for (size_t q = 0; q < np; ++q) { u1[q] = o.u[o.I(0, 0, 1) + q]; v1[q] = o.v[o.I(0, 0, 1) + q]; }
  }
  double c1 = 0.5 * o.dz * KVONK;
  for (size_t q = 0; q < np; ++q) {
    double uavg = std::sqrt(u1[q] * u1[q] + v1[q] * v1[q]);
    double us = KVONK * uavg / o.logterm[q];
    o.ustar[q] = us;
    size_t k1 = o.I(0, 0, 1) + q;
    if (uavg == 0.0) {
      o.dudz[k1] = 0.0; o.dvdz[k1] = 0.0;
      o.lbc_u[q] = 0.0; o.lbc_v[q] = 0.0;
      o.txz[k1] = 0.0; o.tyz[k1] = 0.0;
    } else {
      o.lbc_u[q] = us / c1 * u1[q] / uavg;
      o.lbc_v[q] = us / c1 * v1[q] / uavg;
      o.dudz[k1] = o.lbc_u[q];
      o.dvdz[k1] = o.lbc_v[q];
      o.txz[k1] = (us * us) * u1[q] / uavg;
      o.tyz[k1] = (us * us) * v1[q] / uavg;
    }
  }
}

// grad_velocity_filt (src/compute_derivatives.f90:94-137) on all nz+2 planes
static void grad_filt(Oracle& o, std::vector<double>& f, std::vector<double>& fx, std::vector<double>& fy) {
  double c = 1.0 / (double)(o.nx * o.ny);
#pragma omp parallel
  {
    std::vector<Cx> sp((size_t)o.lhx * o.ny), d1(sp.size()), d2(sp.size());
#pragma omp for schedule(static)
    for (int k = 0; k < o.nzg; ++k) {
      fft2_fwd(o, &f[o.I(0, 0, k)], sp.data(), c);
      for (int j = 0; j < o.ny; ++j)
        for (int i = 0; i < o.lhx; ++i) {
          size_t q = i + (size_t)o.lhx * j;
          Cx a = nyq(o, i, j) ? Cx{0.0, 0.0} : sp[q];
          sp[q] = a;
          d1[q] = ik(o.kkx[q], a);
          d2[q] = ik(o.kky[q], a);
        }
      fft2_inv(o, d1.data(), &fx[o.I(0, 0, k)]);
      fft2_inv(o, d2.data(), &fy[o.I(0, 0, k)]);
      fft2_inv(o, sp.data(), &f[o.I(0, 0, k)]);
    }
  }
}

// set_bc bottom/top for u, v (staggered) and w (not staggered)
static void set_bc_all(Oracle& o) {
  size_t np = o.np;
  int nz = o.nz;
  for (size_t q = 0; q < np; ++q) {
    // bottom: wall law switches u, v to Neumann (src/LES_engine.f90:168-171)
    o.u[q] = o.u[o.I(0, 0, 1) + q] - o.lbc_u[q] * o.dz;
    o.v[q] = o.v[o.I(0, 0, 1) + q] - o.lbc_v[q] * o.dz;
    o.w[o.I(0, 0, 1) + q] = 0.0;  // lbc_w = lbc%v = 0
    // top
    size_t t = o.I(0, 0, nz + 1) + q, b = o.I(0, 0, nz) + q;
    if (o.P.ubc_u_neu) o.u[t] = o.u[b] + o.P.ubc_u * o.dz; else o.u[t] = 2.0 * o.P.ubc_u - o.u[b];
    if (o.P.ubc_v_neu) o.v[t] = o.v[b] + o.P.ubc_v * o.dz; else o.v[t] = 2.0 * o.P.ubc_v - o.v[b];
    o.w[t] = o.P.ubc_w;
  }
}

static void compute_vel_grad(Oracle& o) {
  grad_filt(o, o.u, o.dudx, o.dudy);
  grad_filt(o, o.v, o.dvdx, o.dvdy);
  grad_filt(o, o.w, o.dwdx, o.dwdy);
  set_bc_all(o);
  double c = 1.0 / o.dz;
  int nz = o.nz;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q, b = o.I(0, 0, k - 1) + q;
      o.dudz[a] = c * (o.u[a] - o.u[b]);
      o.dvdz[a] = c * (o.v[a] - o.v[b]);
      if (k <= nz) o.dwdz[a] = c * (o.w[o.I(0, 0, k + 1) + q] - o.w[a]);
      else o.dwdz[a] = 0.0;
    }
}

// compute_sij (src/compute_sgs.f90:260-335): wall-law first level on uvp nodes
static void compute_sij(Oracle& o) {
  int nz = o.nz;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q, b = o.I(0, 0, k - 1) + q;
      double ux, uy, uz, vx, vy, vz, wx, wy, wz;
      if (k == 1) {
        size_t c2 = o.I(0, 0, 2) + q;
        ux = o.dudx[a]; uy = o.dudy[a];
        uz = o.dudz[a]; vz = o.dvdz[a];
        vx = o.dvdx[a]; vy = o.dvdy[a];
        wx = 0.5 * (o.dwdx[a] + o.dwdx[c2]);
        wy = 0.5 * (o.dwdy[a] + o.dwdy[c2]);
        wz = o.dwdz[a];
      } else {
        ux = 0.5 * (o.dudx[a] + o.dudx[b]);
        uy = 0.5 * (o.dudy[a] + o.dudy[b]);
        uz = o.dudz[a];
        vx = 0.5 * (o.dvdx[a] + o.dvdx[b]);
        vy = 0.5 * (o.dvdy[a] + o.dvdy[b]);
        vz = o.dvdz[a];
        wx = o.dwdx[a];
        wy = o.dwdy[a];
        wz = 0.5 * (o.dwdz[a] + o.dwdz[b]);
      }
      o.s11[a] = ux;
      o.s12[a] = 0.5 * (uy + vx);
      o.s13[a] = 0.5 * (uz + wx);
      o.s22[a] = vy;
      o.s23[a] = 0.5 * (vz + wy);
      o.s33[a] = wz;
    }
}

static inline double smag(double s11, double s12, double s13, double s22, double s23, double s33) {
  return std::sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
}

static inline double f32(double x) { return (double)(float)x; }
static inline double fmax32(double a, float b) {
  float fa = (float)a;
  return (double)(fa > b ? fa : b);
}

// test_filter of planes 1..nz+1 of `spec` (precomputed forward spectra) with G
static void filter_from_spec(const Oracle& o, const std::vector<Cx>& spec, const std::vector<double>& G,
                             std::vector<double>& out) {
  size_t ns = (size_t)o.lhx * o.ny;
#pragma omp parallel
  {
    std::vector<Cx> sp(ns);
#pragma omp for schedule(static)
    for (int k = 1; k <= o.nz + 1; ++k) {
      for (size_t q = 0; q < ns; ++q) sp[q] = cscale(spec[ns * k + q], G[q]);
      fft2_inv(o, sp.data(), &out[o.I(0, 0, k)]);
    }
  }
}

static void fwd_spec(const Oracle& o, const std::vector<double>& f, std::vector<Cx>& spec) {
  size_t ns = (size_t)o.lhx * o.ny;
  spec.assign(ns * o.nzg, Cx{0.0, 0.0});
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= o.nz + 1; ++k) fft2_fwd(o, &f[o.I(0, 0, k)], &spec[ns * k], 1.0);
}

// compute_dynamic_mum for both filter widths (src/compute_LagScaleDep.f90:485-596);
// with nscale = 1 only the 2-delta part (LM, MM), which is the pointwise half of
// compute_dynamic_mum of the plane-averaged model (src/compute_DynSmag.f90:96-221:
// same u_ft/v_ft/w_ft at the first node, same L_ij, M_ij = 2 delta^2
// (filt(|S| S_ij) - 4 |S~| S~_ij), same L:M and M:M sums).
static void dynamic_mum(Oracle& o, const std::vector<double>& Smod, std::vector<double>& LM, std::vector<double>& MM,
                        std::vector<double>& QN, std::vector<double>& NN, int nscale = 2) {
  int nz = o.nz;
  size_t nf = o.nf;
  // 21 inputs: u_ft v_ft w_ft, A11 A12 A13 A22 A23 A33, S11..S33, S*Sij
  std::vector<std::vector<double>> in(21, std::vector<double>(nf, 0.0));
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      double uf, vf, wf;
      if (k == 1) {
        uf = o.u[a]; vf = o.v[a]; wf = (PLANT(13) ? 0.5 : 0.25) * o.w[o.I(0, 0, 2) + q];
      } else {
        size_t b = o.I(0, 0, k - 1) + q;
        uf = 0.5 * (o.u[a] + o.u[b]);
        vf = 0.5 * (o.v[a] + o.v[b]);
        wf = o.w[a];
      }
      in[0][a] = uf; in[1][a] = vf; in[2][a] = wf;
      in[3][a] = uf * uf; in[4][a] = uf * vf; in[5][a] = uf * wf;
      in[6][a] = vf * vf; in[7][a] = vf * wf; in[8][a] = wf * wf;
      in[9][a] = o.s11[a]; in[10][a] = o.s12[a]; in[11][a] = o.s13[a];
      in[12][a] = o.s22[a]; in[13][a] = o.s23[a]; in[14][a] = o.s33[a];
      double S = Smod[a];
      in[15][a] = S * o.s11[a]; in[16][a] = S * o.s12[a]; in[17][a] = S * o.s13[a];
      in[18][a] = S * o.s22[a]; in[19][a] = S * o.s23[a]; in[20][a] = S * o.s33[a];
    }
  std::vector<std::vector<Cx>> spec(21);
  for (int f = 0; f < 21; ++f) fwd_spec(o, in[f], spec[f]);
  double cst = (PLANT(12) ? 1.0 : 2.0) * (o.delta * o.delta);
  for (int sc = 0; sc < nscale; ++sc) {
    const std::vector<double>& G = (sc == 0 || PLANT(11)) ? o.G2 : o.G4;
    double tf2 = (sc == 0 || PLANT(10)) ? 4.0 : 16.0;
    std::vector<std::vector<double>> fl(21, std::vector<double>(nf, 0.0));
    for (int f = 0; f < 21; ++f) filter_from_spec(o, spec[f], G, fl[f]);
    std::vector<double>& AB = sc == 0 ? LM : QN;
    std::vector<double>& BB = sc == 0 ? MM : NN;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
      for (size_t q = 0; q < o.np; ++q) {
        size_t a = o.I(0, 0, k) + q;
        double uf = fl[0][a], vf = fl[1][a], wf = fl[2][a];
        double A11 = fl[3][a] - uf * uf, A12 = fl[4][a] - uf * vf, A13 = fl[5][a] - uf * wf;
        double A22 = fl[6][a] - vf * vf, A23 = fl[7][a] - vf * wf, A33 = fl[8][a] - wf * wf;
        double S11 = fl[9][a], S12 = fl[10][a], S13 = fl[11][a], S22 = fl[12][a], S23 = fl[13][a], S33 = fl[14][a];
        double Sft = smag(S11, S12, S13, S22, S23, S33);
        double B11 = cst * (fl[15][a] - tf2 * Sft * S11);
        double B12 = cst * (fl[16][a] - tf2 * Sft * S12);
        double B13 = cst * (fl[17][a] - tf2 * Sft * S13);
        double B22 = cst * (fl[18][a] - tf2 * Sft * S22);
        double B23 = cst * (fl[19][a] - tf2 * Sft * S23);
        double B33 = cst * (fl[20][a] - tf2 * Sft * S33);
        AB[a] = A11 * B11 + A22 * B22 + A33 * B33 + 2.0 * (A12 * B12 + A13 * B13 + A23 * B23);
        BB[a] = B11 * B11 + B22 * B22 + B33 * B33 + 2.0 * (B12 * B12 + B13 * B13 + B23 * B23);
      }
  }
}

// value of the level-2 halo array FF built by update_halo3 + the z fix-ups in
// compute_interpolag_coeff (src/decomp_2d_custom.f90:902-1024,
// src/compute_LagScaleDep.f90:439-452), one-rank semantics; i in [0,nx+1],
// j in [0,ny+1], k in [0,nz+2] (1-based interior indices).
static inline double ffval(const Oracle& o, const std::vector<double>& F, int i, int j, int k) {
  int nx = o.nx, ny = o.ny, nz = o.nz;
  bool inter = (i >= 1 && i <= nx && j >= 1 && j <= ny);
  int kk = k;
  if (o.P.halo_mode == 0) {
    if (k == 0) kk = 1;
    if (k >= nz + 1) {
      if (inter) return F[o.I(i - 1, j - 1, nz + 1)];
      kk = k - 2;
    }
    int jj = j;
    if (j == 0) jj = 2;
    else if (j == ny + 1) jj = ny - 1;
    int ii = i;
    if (i == 0) ii = PLANT(18) ? 2 : nx;
    else if (i == nx + 1) ii = PLANT(18) ? nx - 1 : 1;
    return F[o.I(ii - 1, jj - 1, kk)];
  } else {
    if (k == 0) kk = 1;
    if (k >= nz + 1) kk = nz + 1;
    int jj = j == 0 ? ny : (j == ny + 1 ? 1 : j);
    int ii = i == 0 ? nx : (i == nx + 1 ? 1 : i);
    return F[o.I(ii - 1, jj - 1, kk)];
  }
}

// compute_interpolag_weight + compute_interpolag_coeff (x4)
// (src/compute_LagScaleDep.f90:309-474)
static void lagrangian_interp(Oracle& o) {
  int nx = o.nx, ny = o.ny, nz = o.nz;
  double cs = (double)o.P.cs_count;
  std::vector<double> ul(o.nf), vl(o.nf), wl(o.nf);
  const double csd = PLANT(15) ? 1.0 : cs;
  for (size_t q = 0; q < o.nf; ++q) { ul[q] = o.u_lag[q] / csd; vl[q] = o.v_lag[q] / csd; wl[q] = o.w_lag[q] / csd; }
  std::vector<double> xp(o.nf), yp(o.nf), zp(o.nf);
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      double uw, vw, ww;
      if (k == 1) {
        uw = ul[a]; vw = vl[a];
        ww = 0.25 * wl[o.I(0, 0, 2) + q];
      } else {
        size_t b = o.I(0, 0, k - 1) + q;
        uw = 0.5 * (ul[a] + ul[b]);
        vw = 0.5 * (vl[a] + vl[b]);
        ww = wl[a];
      }
      const double sg = PLANT(14) ? -1.0 : 1.0;
      xp[a] = -(sg * uw * o.P.dt * cs / o.dx);
      yp[a] = -(sg * vw * o.P.dt * cs / o.dy);
      zp[a] = -(sg * ww * o.P.dt * cs / o.dz);
    }
  for (size_t q = 0; q < o.np; ++q) {
    size_t a1 = o.I(0, 0, 1) + q, a2 = o.I(0, 0, 2) + q;
    if (!PLANT(17)) zp[a2] = zp[a2] + std::fmin(zp[a2], 0.0);
    zp[a1] = zp[a1] + std::fmax(zp[a1], 0.0);
    zp[a2] = std::fmin(1.0, zp[a2]);
    zp[a1] = std::fmin(1.0, PLANT(16) ? zp[a1] : zp[a2]);
    zp[a2] = std::fmax(-1.0, zp[a2]);
    zp[a1] = std::fmax(-1.0, PLANT(16) ? zp[a1] : zp[a2]);
  }
  std::vector<double>* Fs[4] = {&o.F_LM, &o.F_MM, &o.F_QN, &o.F_NN};
  for (int f = 0; f < 4; ++f) {
    const std::vector<double> Fold = *Fs[f];
    std::vector<double>& Fn = *Fs[f];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t a = o.I(i - 1, j - 1, k);
          int ax = i + (std::signbit(xp[a]) ? -1 : 1);
          int ay = j + (std::signbit(yp[a]) ? -1 : 1);
          int az = k + (std::signbit(zp[a]) ? -1 : 1);
          double cx = std::fabs(xp[a]), cy = std::fabs(yp[a]), cz = std::fabs(zp[a]);
          double fx = 1.0 - cx, fy = 1.0 - cy, fz = 1.0 - cz;
          double v = fx * fy * (ffval(o, Fold, i, j, k) * fz + ffval(o, Fold, i, j, az) * cz) +
                     fx * cy * (ffval(o, Fold, i, ay, k) * fz + ffval(o, Fold, i, ay, az) * cz) +
                     cx * fy * (ffval(o, Fold, ax, j, k) * fz + ffval(o, Fold, ax, j, az) * cz) +
                     cx * cy * (ffval(o, Fold, ax, ay, k) * fz + ffval(o, Fold, ax, ay, az) * cz);
          Fn[a] = v;
        }
  }
}

// tmp**(-1/8) from correctly rounded operations
static inline double pow_m18(double t) {
  if (PLANT(1)) return 1.0 / std::sqrt(std::sqrt(t));
  return 1.0 / std::sqrt(std::sqrt(std::sqrt(t)));
}

// update_interpolag_coeff (src/compute_LagScaleDep.f90:607-640)
static void update_F(Oracle& o, std::vector<double>& FAB, std::vector<double>& FBB, const std::vector<double>& AB,
                     const std::vector<double>& BB) {
  double opftdelta = (PLANT(2) ? 1.0 : 1.5) * o.delta;
  double lagran_dt = o.P.dt * (PLANT(3) ? 1.0 : (double)o.P.cs_count);
  for (int k = 1; k <= o.nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      double tmp = fmax32(FAB[a] * FBB[a], 1e-24f);
      tmp = opftdelta * pow_m18(tmp);
      tmp = fmax32(tmp, 1e-24f);
      double dumfac = lagran_dt / tmp;
      double epsi = dumfac / (1.0 + dumfac);
      double fab = PLANT(4) ? (1.0 - epsi) * AB[a] + epsi * FAB[a] : epsi * AB[a] + (1.0 - epsi) * FAB[a];
      double fbb = PLANT(4) ? (1.0 - epsi) * BB[a] + epsi * FBB[a] : epsi * BB[a] + (1.0 - epsi) * FBB[a];
      FAB[a] = PLANT(5) ? fab : fmax32(fab, 1e-24f);
      FBB[a] = fbb;
    }
}

// compute_dynamic_coeff (src/compute_LagScaleDep.f90:651-707); beta exponent
// log(2)/(log(4)-log(2)) evaluates to exactly 1 in IEEE double.
static void dynamic_coeff(Oracle& o) {
  for (int k = 1; k <= o.nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      double c2 = fmax32(o.F_LM[a] / o.F_MM[a], 1e-24f);
      double c4 = fmax32(o.F_QN[a] / o.F_NN[a], 1e-24f);
      double b = PLANT(6) ? c2 / c4 : c4 / c2;
      b = fmax32(b, PLANT(7) ? 0.25f : 0.125f);
      if (k == o.nz + 1 && !PLANT(8)) b = 1.0;
      o.beta[a] = b;
      o.cs_opt2[a] = fmax32(PLANT(9) ? c2 * b : c2 / b, 1e-24f);
    }
}

// compute_LagScaleDep_coeff (src/compute_LagScaleDep.f90:204-298)
static void lasd(Oracle& o, int jtt) {
  if (jtt >= o.P.dyn_init) {
    for (size_t q = 0; q < o.nf; ++q) { o.u_lag[q] += o.u[q]; o.v_lag[q] += o.v[q]; o.w_lag[q] += o.w[q]; }
  }
  if (jtt < o.P.dyn_init) {
    for (int k = 1; k <= o.nz + 1; ++k)
      for (size_t q = 0; q < o.np; ++q) o.cs_opt2[o.I(0, 0, k) + q] = 0.03;
  } else if (jtt % o.P.cs_count == 0) {
    std::vector<double> Smod(o.nf, 0.0), LM(o.nf, 0.0), MM(o.nf, 0.0), QN(o.nf, 0.0), NN(o.nf, 0.0);
    for (int k = 1; k <= o.nz + 1; ++k)
      for (size_t q = 0; q < o.np; ++q) {
        size_t a = o.I(0, 0, k) + q;
        Smod[a] = smag(o.s11[a], o.s12[a], o.s13[a], o.s22[a], o.s23[a], o.s33[a]);
      }
    dynamic_mum(o, Smod, LM, MM, QN, NN);
    bool interp_after = false;
    if (o.F_coeff_init) {
      if (PLANT(20)) interp_after = true;
      else lagrangian_interp(o);
    } else {
      const double c0 = PLANT(19) ? 1.0 : 0.03;
      for (int k = 1; k <= o.nz + 1; ++k)
        for (size_t q = 0; q < o.np; ++q) {
          size_t a = o.I(0, 0, k) + q;
          o.F_MM[a] = MM[a]; o.F_LM[a] = c0 * MM[a];
          o.F_NN[a] = NN[a]; o.F_QN[a] = c0 * NN[a];
        }
      o.F_coeff_init = true;
    }
    update_F(o, o.F_LM, o.F_MM, LM, MM);
    update_F(o, o.F_QN, o.F_NN, QN, NN);
    if (interp_after) lagrangian_interp(o);
    dynamic_coeff(o);
    o.lagran_flag = true;
    std::fill(o.u_lag.begin(), o.u_lag.end(), 0.0);
    std::fill(o.v_lag.begin(), o.v_lag.end(), 0.0);
    std::fill(o.w_lag.begin(), o.w_lag.end(), 0.0);
  }
}

// compute_DynSmag_coeff (src/compute_DynSmag.f90:61-82) with the plane
// averages of compute_dynamic_mum (:225-243): num(k), den(k) are the level sums
// of L:M and M:M in i-fastest, then j order divided by nx*ny
// (compute_planar_average_wghost, src/utility_tools.f90:100-132), CS_Z(k) =
// num/den (0 where den = 0), clipped at 0 below, and Cs^2 = CS_Z(k) on the
// whole level.  The update runs on the local step count jt, the reset to 0.03
// on the global count jtt.
static void dynsmag(Oracle& o, int jt, int jtt) {
  int nz = o.nz;
  if (jtt < o.P.dyn_init) {
    for (int k = 1; k <= nz + 1; ++k)
      for (size_t q = 0; q < o.np; ++q) o.cs_opt2[o.I(0, 0, k) + q] = 0.03;
  } else if (jt % o.P.cs_count == 0) {
    std::vector<double> Smod(o.nf, 0.0), LM(o.nf, 0.0), MM(o.nf, 0.0);
    for (int k = 1; k <= nz + 1; ++k)
      for (size_t q = 0; q < o.np; ++q) {
        size_t a = o.I(0, 0, k) + q;
        Smod[a] = smag(o.s11[a], o.s12[a], o.s13[a], o.s22[a], o.s23[a], o.s33[a]);
      }
    dynamic_mum(o, Smod, LM, MM, LM, MM, 1);
    const double npl = (double)(o.nx * o.ny);
    for (int k = 1; k <= nz + 1; ++k) {
      double num = 0.0, den = 0.0;
      for (size_t q = 0; q < o.np; ++q) num = num + LM[o.I(0, 0, k) + q];
      for (size_t q = 0; q < o.np; ++q) den = den + MM[o.I(0, 0, k) + q];
      num = num / npl;
      den = den / npl;
      double c = den != 0.0 ? num / den : 0.0;
      c = c > 0.0 ? c : 0.0;
      for (size_t q = 0; q < o.np; ++q) o.cs_opt2[o.I(0, 0, k) + q] = c;
    }
  }
}

// sgs_stress (src/compute_sgs.f90:105-247)
static void sgs_stress(Oracle& o, int jt, int jtt) {
  int nz = o.nz;
  if (o.P.sgs_model == 3) {
    // molecular stress 2/Re S_ij on levels 1..nz, every component on the
    // gradients' own nodes (src/compute_sgs.f90:124-140); S_ij, nu_t, Cs^2
    // and the ghost levels are left as they are
    const double c = 2.0 / o.P.re;
    for (int k = 1; k <= nz; ++k)
      for (size_t q = 0; q < o.np; ++q) {
        size_t a = o.I(0, 0, k) + q;
        o.txx[a] = c * o.dudx[a];
        o.txy[a] = c * 0.5 * (o.dudy[a] + o.dvdx[a]);
        o.tyy[a] = c * o.dvdy[a];
        o.tzz[a] = c * o.dwdz[a];
        o.txz[a] = c * 0.5 * (o.dudz[a] + o.dwdx[a]);
        o.tyz[a] = c * 0.5 * (o.dvdz[a] + o.dwdy[a]);
      }
    return;
  }
  compute_sij(o);
  std::vector<double> l(nz + 2, o.delta);
  if (o.P.sgs_model == 0) {
    double n = 2.0;
    double z = 0.5 * o.dz;
    l[1] = std::pow(std::pow(o.P.cs0, n) * std::pow(KVONK * z, -n) + std::pow(o.delta, -n), -1.0 / n);
    for (int k = 2; k <= nz + 1; ++k) {
      z = ((double)k - (PLANT(24) ? 0.0 : 1.0)) * o.dz;
      l[k] = std::pow(std::pow(o.P.cs0, n) * std::pow(KVONK * z, -n) + std::pow(o.delta, -n), -1.0 / n);
    }
    for (int k = 1; k <= nz + 1; ++k)
      for (size_t q = 0; q < o.np; ++q) o.cs_opt2[o.I(0, 0, k) + q] = PLANT(25) ? o.P.cs0 : o.P.cs0 * o.P.cs0;
  } else if (o.P.sgs_model == 1) {
    dynsmag(o, jt, jtt);
  } else {
    lasd(o, jtt);
  }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      double st = smag(o.s11[a], o.s12[a], o.s13[a], o.s22[a], o.s23[a], o.s33[a]);
      o.nu_t[a] = st * o.cs_opt2[a] * (l[k] * l[k]);
    }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      if (k <= nz) {
        double nu = (k == 1 && !PLANT(22)) ? o.nu_t[a] : 0.5 * (o.nu_t[a] + o.nu_t[o.I(0, 0, k + 1) + q]);
        if (PLANT(21)) nu = o.nu_t[a];
        o.txx[a] = 2.0 * nu * o.dudx[a];
        o.txy[a] = 2.0 * nu * 0.5 * (o.dudy[a] + o.dvdx[a]);
        o.tyy[a] = 2.0 * nu * o.dvdy[a];
        o.tzz[a] = 2.0 * nu * o.dwdz[a];
      }
      if (k >= 2) {
        double nw = PLANT(23) ? 0.5 * (o.nu_t[a] + o.nu_t[o.I(0, 0, k - 1) + q]) : o.nu_t[a];
        o.txz[a] = 2.0 * nw * o.s13[a];
        o.tyz[a] = 2.0 * nw * o.s23[a];
      }
    }
  for (size_t q = 0; q < o.np; ++q) o.tzz[q] = o.tzz[o.I(0, 0, 1) + q];
}

// 3/2-padded rotational convection (src/compute_convective_padd.f90:115-388).
// Padding: each real row pair (a, b) -> Z = FFT(a + i b), copy modes |k| < n/2
// into the longer spectrum (Nyquist dropped), inverse; truncation is the
// mirror image with the Nyquist slot zeroed.
static void pad_line(const Plan& pn, const Plan& pN, Cx* z, Cx* zb, Cx* s, int n, int N) {
  pn.exec(z, s, false);
  for (int k = 0; k < N; ++k) zb[k] = {0.0, 0.0};
  for (int k = 0; k < n / 2; ++k) zb[k] = z[k];
  for (int k = 1; k < n / 2; ++k) zb[N - k] = z[n - k];
  pN.exec(zb, s, true);
}
static void trunc_line(const Plan& pN, const Plan& pn, Cx* zb, Cx* z, Cx* s, int N, int n) {
  pN.exec(zb, s, false);
  for (int k = 0; k < n; ++k) z[k] = {0.0, 0.0};
  for (int k = 0; k < n / 2; ++k) z[k] = zb[k];
  for (int k = 1; k < n / 2; ++k) z[n - k] = zb[N - k];
  pn.exec(z, s, true);
}

// real plane (nx, ny) -> padded real plane (nx2, ny2)
static void pad_plane(const Oracle& o, const double* in, double* out) {
  int nx = o.nx, ny = o.ny, nx2 = 3 * nx / 2, ny2 = 3 * ny / 2;
  std::vector<double> hyb((size_t)nx2 * ny);
  int m = nx2 > ny2 ? nx2 : ny2;
  std::vector<Cx> z(m), zb(m), s(m);
  for (int j = 0; j < ny; j += 2) {
    for (int n = 0; n < nx; ++n) z[n] = {in[n + (size_t)nx * j], in[n + (size_t)nx * (j + 1)]};
    pad_line(o.px, o.px2, z.data(), zb.data(), s.data(), nx, nx2);
    for (int n = 0; n < nx2; ++n) { hyb[n + (size_t)nx2 * j] = zb[n].x; hyb[n + (size_t)nx2 * (j + 1)] = zb[n].y; }
  }
  for (int i = 0; i < nx2; i += 2) {
    for (int j = 0; j < ny; ++j) z[j] = {hyb[i + (size_t)nx2 * j], hyb[i + 1 + (size_t)nx2 * j]};
    pad_line(o.py, o.py2, z.data(), zb.data(), s.data(), ny, ny2);
    for (int j = 0; j < ny2; ++j) { out[i + (size_t)nx2 * j] = zb[j].x; out[i + 1 + (size_t)nx2 * j] = zb[j].y; }
  }
}

// padded real plane (nx2, ny2) -> truncated real plane (nx, ny)
static void trunc_plane(const Oracle& o, const double* in, double* out) {
  int nx = o.nx, ny = o.ny, nx2 = 3 * nx / 2, ny2 = 3 * ny / 2;
  std::vector<double> hyb((size_t)nx2 * ny);
  int m = nx2 > ny2 ? nx2 : ny2;
  std::vector<Cx> z(m), zb(m), s(m);
  for (int i = 0; i < nx2; i += 2) {
    for (int j = 0; j < ny2; ++j) zb[j] = {in[i + (size_t)nx2 * j], in[i + 1 + (size_t)nx2 * j]};
    trunc_line(o.py2, o.py, zb.data(), z.data(), s.data(), ny2, ny);
    for (int j = 0; j < ny; ++j) { hyb[i + (size_t)nx2 * j] = z[j].x; hyb[i + 1 + (size_t)nx2 * j] = z[j].y; }
  }
  for (int j = 0; j < ny; j += 2) {
    for (int n = 0; n < nx2; ++n) zb[n] = {hyb[n + (size_t)nx2 * j], hyb[n + (size_t)nx2 * (j + 1)]};
    trunc_line(o.px2, o.px, zb.data(), z.data(), s.data(), nx2, nx);
    for (int n = 0; n < nx; ++n) { out[n + (size_t)nx * j] = z[n].x; out[n + (size_t)nx * (j + 1)] = z[n].y; }
  }
}

static void convection_nxny(Oracle& o, std::vector<double>& cx, std::vector<double>& cy, std::vector<double>& cz);

static void convection(Oracle& o, std::vector<double>& cx, std::vector<double>& cy, std::vector<double>& cz) {
  if (o.P.dealias != 0) {
    convection_nxny(o, cx, cy, cz);
    return;
  }
  int nx = o.nx, ny = o.ny, nz = o.nz, nx2 = 3 * nx / 2, ny2 = 3 * ny / 2;
  size_t nb = (size_t)nx2 * ny2;
  double c = 1.0 / (double)(nx * ny);
  // padded planes 1..nz+1 of u, v, w, vort1, vort2, vort3
  std::vector<std::vector<double>> big(6, std::vector<double>(nb * (nz + 2), 0.0));
#pragma omp parallel
  {
    std::vector<double> tmp(o.np);
#pragma omp for schedule(static)
    for (int k = 1; k <= nz + 1; ++k) {
      for (int f = 0; f < 6; ++f) {
        for (size_t q = 0; q < o.np; ++q) {
          size_t a = o.I(0, 0, k) + q;
          double val;
          if (f == 0) val = -(c * o.u[a]);
          else if (f == 1) val = -(c * o.v[a]);
          else if (f == 2) val = -(c * o.w[a]);
          else if (f == 3) {
            if (k == 1) val = c * (0.5 * (o.dwdy[a] + o.dwdy[o.I(0, 0, 2) + q]) - o.dvdz[a]);
            else val = c * (o.dwdy[a] - o.dvdz[a]);
          } else if (f == 4) {
            if (k == 1) val = c * (o.dudz[a] - 0.5 * (o.dwdx[a] + o.dwdx[o.I(0, 0, 2) + q]));
            else val = c * (o.dudz[a] - o.dwdx[a]);
          } else val = c * (o.dvdx[a] - o.dudy[a]);
          tmp[q] = val;
        }
        pad_plane(o, tmp.data(), &big[f][nb * k]);
      }
    }
  }
  double c2 = 1.0 / (double)(nx2 * ny2);
#pragma omp parallel
  {
    std::vector<double> prod(nb), res(o.np);
#pragma omp for schedule(static)
    for (int k = 1; k <= nz; ++k) {
      for (int t = 0; t < 3; ++t) {
        const double *ub = &big[0][nb * k], *vb = &big[1][nb * k], *wb = &big[2][nb * k];
        const double *o1 = &big[3][nb * k], *o2 = &big[4][nb * k], *o3 = &big[5][nb * k];
        const double *ubm = &big[0][nb * (k - 1)], *vbm = &big[1][nb * (k - 1)];
        const double *wbp = &big[2][nb * (k + 1)], *o1p = &big[3][nb * (k + 1)], *o2p = &big[4][nb * (k + 1)];
        for (size_t q = 0; q < nb; ++q) {
          double val;
          if (t == 0) {
            if (k == 1) val = c2 * (-(vb[q] * o3[q]) + (PLANT(26) ? 1.0 : 0.5) * wbp[q] * o2p[q]);
            else if (PLANT(28)) val = c2 * (-(vb[q] * o3[q]) + 0.5 * (wb[q] * o2[q] + big[2][nb * (k - 1) + q] * big[4][nb * (k - 1) + q]));
            else val = c2 * (-(vb[q] * o3[q]) + 0.5 * (wbp[q] * o2p[q] + wb[q] * o2[q]));
          } else if (t == 1) {
            if (k == 1) val = c2 * (ub[q] * o3[q] + 0.5 * wbp[q] * (-o1p[q]));
            else val = c2 * (ub[q] * o3[q] + 0.5 * (wb[q] * (-o1[q]) + wbp[q] * (-o1p[q])));
          } else {
            if (k == 1) val = 0.0;
            else if (PLANT(27)) val = c2 * 0.5 * ((ub[q] + ub[q]) * (-o2[q]) + (vb[q] + vb[q]) * o1[q]);
            else val = c2 * 0.5 * ((ub[q] + ubm[q]) * (-o2[q]) + (vb[q] + vbm[q]) * o1[q]);
          }
          prod[q] = val;
        }
        trunc_plane(o, prod.data(), res.data());
        std::vector<double>& dst = t == 0 ? cx : (t == 1 ? cy : cz);
        std::memcpy(&dst[o.I(0, 0, k)], res.data(), sizeof(double) * o.np);
      }
    }
  }
  for (size_t q = 0; q < o.np; ++q) {
    cx[q] = BOGUS; cy[q] = BOGUS; cz[q] = BOGUS;
    cx[o.I(0, 0, nz + 1) + q] = BOGUS; cy[o.I(0, 0, nz + 1) + q] = BOGUS; cz[o.I(0, 0, nz + 1) + q] = BOGUS;
  }
}

// compute_dealias_convec of the other dealiasing options, chosen at compile
// time in the reference (DSLG, src/Makefile:17-40): 1 TRUNC
// (src/compute_convective_trunc.f90:115-247), 2 FSMTHG
// (src/compute_convective_Fsmthg.f90:141-264), 3 NODLSG
// (src/compute_convective_nodlsg.f90:88-205).  The six sources (the same as
// the padded version's, levels 1..nz+1; level 0 enters no product) and the
// products stay on the nx x ny grid with const = 1/(nx ny) in both places;
// TRUNC and FSMTHG filter the sources (levels 1..nz+1) with GA and the
// products (levels 1..nz) with GB through the unnormalised 2D transforms.
static void convection_nxny(Oracle& o, std::vector<double>& cx, std::vector<double>& cy, std::vector<double>& cz) {
  int nz = o.nz;
  double c = 1.0 / (double)(o.nx * o.ny);
  bool filt = o.P.dealias != 3;
  std::vector<std::vector<double>> src(6, std::vector<double>(o.nf, 0.0));
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q, a2 = o.I(0, 0, 2) + q;
      src[0][a] = -(c * o.u[a]);
      src[1][a] = -(c * o.v[a]);
      src[2][a] = -(c * o.w[a]);
      if (k == 1) {
        src[3][a] = c * (0.5 * (o.dwdy[a] + o.dwdy[a2]) - o.dvdz[a]);
        src[4][a] = c * (o.dudz[a] - 0.5 * (o.dwdx[a] + o.dwdx[a2]));
      } else {
        src[3][a] = c * (o.dwdy[a] - o.dvdz[a]);
        src[4][a] = c * (o.dudz[a] - o.dwdx[a]);
      }
      src[5][a] = c * (o.dvdx[a] - o.dudy[a]);
    }
  if (filt) {
    std::vector<Cx> spec;
    for (int f = 0; f < 6; ++f) {
      fwd_spec(o, src[f], spec);
      filter_from_spec(o, spec, o.GA, src[f]);
    }
  }
  std::vector<double>* dst[3] = {&cx, &cy, &cz};
  std::vector<std::vector<double>> prod(3, std::vector<double>(o.nf, 0.0));
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q, up = o.I(0, 0, k + 1) + q, dn = o.I(0, 0, k - 1) + q;
      const std::vector<double>&ub = src[0], &vb = src[1], &wb = src[2], &o1 = src[3], &o2 = src[4], &o3 = src[5];
      double x, y, z;
      if (k == 1) {
        x = c * (-(vb[a] * o3[a]) + 0.5 * wb[up] * o2[up]);
        y = c * (ub[a] * o3[a] + 0.5 * wb[up] * (-o1[up]));
        z = 0.0;
      } else {
        x = c * (-(vb[a] * o3[a]) + 0.5 * (wb[up] * o2[up] + wb[a] * o2[a]));
        y = c * (ub[a] * o3[a] + 0.5 * (wb[a] * (-o1[a]) + wb[up] * (-o1[up])));
        z = c * 0.5 * ((ub[a] + ub[dn]) * (-o2[a]) + (vb[a] + vb[dn]) * o1[a]);
      }
      prod[0][a] = x; prod[1][a] = y; prod[2][a] = z;
    }
  for (int t = 0; t < 3; ++t) {
    if (filt) {
      std::vector<Cx> spec;
      fwd_spec(o, prod[t], spec);
      filter_from_spec(o, spec, o.GB, *dst[t]);
    } else {
      for (int k = 1; k <= nz; ++k)
        std::memcpy(&(*dst[t])[o.I(0, 0, k)], &prod[t][o.I(0, 0, k)], sizeof(double) * o.np);
    }
  }
  for (size_t q = 0; q < o.np; ++q) {
    cx[q] = BOGUS; cy[q] = BOGUS; cz[q] = BOGUS;
    cx[o.I(0, 0, nz + 1) + q] = BOGUS; cy[o.I(0, 0, nz + 1) + q] = BOGUS; cz[o.I(0, 0, nz + 1) + q] = BOGUS;
  }
}

// div_stress (src/compute_derivatives.f90:464-577): rhs += d/dx tx + d/dy ty +
// d/dz tz; stag selects ddz_w (uvp output) or ddz_uv (w output).
static void div_stress(Oracle& o, std::vector<double>& rhs, const std::vector<double>& tx, const std::vector<double>& ty,
                       const std::vector<double>& tz, bool stag, bool save_wk) {
  int nz = o.nz;
  double c = 1.0 / (double)(o.nx * o.ny);
  double cz = 1.0 / o.dz;
  std::vector<double> divh(o.nf, 0.0);
#pragma omp parallel
  {
    std::vector<Cx> sx((size_t)o.lhx * o.ny), sy(sx.size());
#pragma omp for schedule(static)
    for (int k = 0; k <= nz + 1; ++k) {
      fft2_fwd(o, &tx[o.I(0, 0, k)], sx.data(), c);
      fft2_fwd(o, &ty[o.I(0, 0, k)], sy.data(), c);
      for (int j = 0; j < o.ny; ++j)
        for (int i = 0; i < o.lhx; ++i) {
          size_t q = i + (size_t)o.lhx * j;
          Cx a = nyq(o, i, j) ? Cx{0.0, 0.0} : sx[q];
          Cx b = nyq(o, i, j) ? Cx{0.0, 0.0} : sy[q];
          sx[q] = cadd(ik(o.kkx[q], a), ik(o.kky[q], b));
        }
      fft2_inv(o, sx.data(), &divh[o.I(0, 0, k)]);
    }
  }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      bool st = PLANT(38) ? !stag : stag;
      double dz_ = st ? cz * (tz[o.I(0, 0, k + 1) + q] - tz[a]) : cz * (tz[a] - tz[o.I(0, 0, k - 1) + q]);
      rhs[a] = rhs[a] + divh[a] + dz_;
    }
  if (save_wk) {
    for (size_t q = 0; q < o.np; ++q) {
      size_t a1 = o.I(0, 0, 1) + q;
      double dz1 = cz * (tz[a1] - tz[q]);
      o.bottom_wk[q] = PLANT(37) ? divh[a1] : divh[a1] + dz1;
      o.top_wk[q] = divh[o.I(0, 0, nz + 1) + q];
    }
  }
}

static void compute_rhs(Oracle& o) {
  int nz = o.nz;
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      o.rhsx_f[a] = o.rhsx[a]; o.rhsy_f[a] = o.rhsy[a]; o.rhsz_f[a] = o.rhsz[a];
    }
  convection(o, o.rhsx, o.rhsy, o.rhsz);
  div_stress(o, o.rhsx, o.txx, o.txy, o.txz, true, false);
  div_stress(o, o.rhsy, o.txy, o.tyy, o.tyz, true, false);
  div_stress(o, o.rhsz, o.txz, o.tyz, o.tzz, false, true);
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      o.rhsx[a] = o.rhsx[a] + o.P.pfx;
      o.rhsy[a] = o.rhsy[a] + o.P.pfy;
      if (o.P.coriolis) {
        const double fs = PLANT(39) ? -1.0 : 1.0;
        const double ug = PLANT(40) ? o.P.vg : o.P.ug, vg = PLANT(40) ? o.P.ug : o.P.vg;
        o.rhsx[a] = o.rhsx[a] + fs * o.P.coriol * o.v[a] - o.P.coriol * vg;
        o.rhsy[a] = o.rhsy[a] - fs * o.P.coriol * o.u[a] + o.P.coriol * ug;
      }
    }
}

// Thomas algorithm for one (kx, ky) system of length nz+2
// (src/compute_pressure.f90:448-493, 544-619)
static void tridiag(const Oracle& o, int i, int j, const Cx* r, Cx* x) {
  int nz = o.nz;
  size_t q = i + (size_t)o.lhx * j;
  double idz2 = 1.0 / (o.dz * o.dz);
  double kk = o.kkx[q] * o.kkx[q] + o.kky[q] * o.kky[q];
  double bmid = -(kk + 2.0 / (o.dz * o.dz));
  bool mean = (i == 0 && j == 0);
  std::vector<double> gam(nz + 3);
  // rows 1..nz+2 (1-based); b(1) = -1, c(1) = 1 except the mean mode (1, 0)
  double b1 = mean ? 1.0 : -1.0, c1 = mean ? 0.0 : 1.0;
  double bet = b1;
  x[1] = {r[1].x / bet, r[1].y / bet};
  for (int k = 2; k <= nz + 2; ++k) {
    double a = (k == nz + 2) ? -1.0 : idz2;
    double b = (k == nz + 2) ? 1.0 : bmid;
    double cprev = (k == 2) ? c1 : idz2;
    gam[k] = cprev / bet;
    bet = b - a * gam[k];
    x[k] = {(r[k].x - a * x[k - 1].x) / bet, (r[k].y - a * x[k - 1].y) / bet};
  }
  for (int k = nz + 1; k >= 1; --k) x[k] = {x[k].x - gam[k + 1] * x[k + 1].x, x[k].y - gam[k + 1] * x[k + 1].y};
}

// press_stag_array + poisson_solver + compute_press_grad
// (src/compute_pressure.f90:116-372)
static void pressure(Oracle& o, double tadv1) {
  int nz = o.nz, lhx = o.lhx, ny = o.ny;
  size_t ns = (size_t)lhx * ny;
  double c = 1.0 / (double)(o.nx * o.ny);
  double cr = c / tadv1;
  std::vector<Cx> Hx(ns * o.nzg, Cx{0, 0}), Hy(Hx.size(), Cx{0, 0}), Hz(Hx.size(), Cx{0, 0});
#pragma omp parallel
  {
    std::vector<double> tmp(o.np);
#pragma omp for schedule(static)
    for (int k = 1; k <= nz; ++k) {
      for (int f = 0; f < 3; ++f) {
        const std::vector<double>& src = f == 0 ? o.u : (f == 1 ? o.v : o.w);
        for (size_t q = 0; q < o.np; ++q) tmp[q] = cr * (src[o.I(0, 0, k) + q] / o.P.dt);
        std::vector<Cx>& H = f == 0 ? Hx : (f == 1 ? Hy : Hz);
        fft2_fwd(o, tmp.data(), &H[ns * k], 1.0);
      }
    }
  }
  std::vector<Cx> bot(ns), top(ns);
  {
    std::vector<double> tmp(o.np);
    for (size_t q = 0; q < o.np; ++q) tmp[q] = c * o.bottom_wk[q];
    fft2_fwd(o, tmp.data(), bot.data(), 1.0);
    for (size_t q = 0; q < o.np; ++q) tmp[q] = c * o.top_wk[q];
    fft2_fwd(o, tmp.data(), top.data(), 1.0);
  }
  std::vector<Cx> phat(ns * o.nzg, Cx{0, 0});
  double dz = o.dz;
#pragma omp parallel
  {
    std::vector<Cx> r(nz + 3), x(nz + 3);
#pragma omp for schedule(static)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < lhx; ++i) {
        if (nyq(o, i, j)) continue;
        size_t q = i + (size_t)lhx * j;
        Cx bl = cscale(bot[q], PLANT(29) ? 1.0 : dz), tl = cscale(PLANT(30) ? bot[q] : top[q], dz);
        r[1] = (i == 0 && j == 0 && !PLANT(31)) ? Cx{0.0, 0.0} : bl;
        for (int k = 1; k <= nz; ++k) {
          Cx hx = Hx[ns * k + q], hy = Hy[ns * k + q], hz = Hz[ns * k + q], hzp = Hz[ns * (k + 1) + q];
          double sr = o.kkx[q] * hx.x + o.kky[q] * hy.x;
          double si = o.kkx[q] * hx.y + o.kky[q] * hy.y;
          r[k + 1] = {-si + (hzp.x - hz.x) / dz, sr + (hzp.y - hz.y) / dz};
        }
        r[nz + 2] = tl;
        tridiag(o, i, j, r.data(), x.data());
        phat[q] = x[1];
        for (int k = 1; k <= nz; ++k) phat[ns * k + q] = x[k + 1];
      }
  }
  double cz = 1.0 / dz;
#pragma omp parallel
  {
    std::vector<Cx> a(ns), b(ns), pp(ns);
#pragma omp for schedule(static)
    for (int k = 0; k <= nz + 1; ++k) {
      for (size_t q = 0; q < ns; ++q) {
        pp[q] = phat[ns * k + q];
        a[q] = ik(o.kkx[q], pp[q]);
        b[q] = ik(o.kky[q], pp[q]);
      }
      if (k >= 1 && k <= nz) {
        fft2_inv(o, a.data(), &o.dpdx[o.I(0, 0, k)]);
        fft2_inv(o, b.data(), &o.dpdy[o.I(0, 0, k)]);
      }
      fft2_inv(o, pp.data(), &o.p[o.I(0, 0, k)]);
    }
  }
  for (size_t q = 0; q < o.np; ++q) o.p[o.I(0, 0, nz + 1) + q] = BOGUS;
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) o.dpdz[o.I(0, 0, k) + q] = cz * (o.p[o.I(0, 0, k) + q] - o.p[o.I(0, 0, k - 1) + q]);
}

static void step(Oracle& o) {
  o.jt_total += 1;
  o.jt += 1;
  int jtt = o.jt_total, nz = o.nz;
  wall_law(o);
  compute_vel_grad(o);
  sgs_stress(o, o.jt, jtt);
  compute_rhs(o);
  double tadv1 = (jtt == 1 && !PLANT(32)) ? 1.0 : 1.5;
  double tadv2 = (jtt == 1) ? 0.0 : (PLANT(34) ? 0.5 : 1.0 - tadv1);
  double dt = o.P.dt;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      o.u[a] = o.u[a] + dt * (tadv1 * o.rhsx[a] + tadv2 * o.rhsx_f[a]);
      o.v[a] = o.v[a] + dt * (tadv1 * o.rhsy[a] + tadv2 * o.rhsy_f[a]);
      o.w[a] = o.w[a] + dt * (tadv1 * o.rhsz[a] + tadv2 * o.rhsz_f[a]);
    }
  for (size_t q = 0; q < o.np; ++q) { o.w[q] = 0.0; o.w[o.I(0, 0, nz + 1) + q] = 0.0; }
  pressure(o, tadv1);
  double dtt = dt * tadv1;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      o.u[a] = o.u[a] - dtt * o.dpdx[a];
      o.v[a] = o.v[a] - dtt * o.dpdy[a];
      o.w[a] = o.w[a] - dtt * o.dpdz[a];
      if (!PLANT(33)) {
        o.rhsx[a] = o.rhsx[a] - o.dpdx[a];
        o.rhsy[a] = o.rhsy[a] - o.dpdy[a];
        o.rhsz[a] = o.rhsz[a] - o.dpdz[a];
      }
    }
  for (size_t q = 0; q < o.np; ++q) { o.w[o.I(0, 0, 1) + q] = 0.0; o.w[o.I(0, 0, nz + 1) + q] = 0.0; }
  set_bc_all(o);
}

// ---------------------------------------------------------------- statistics
// Sums in the order of RAV_<jt>_3D.dat (src/IO_statistics.f90:458-490).
enum {
  SU = 0, SV, SW, SP, SU2, SV2, SW2, SU3, SV3, SW3, SU4, SV4, SW4, SUV, SUW, SVW,
  STXX, STYY, STZZ, STXY, STXZ, STYZ, SDUDZ, SDVDZ, SNUT, SCS, SN
};

static void stat_reset(Oracle& o) {
  o.st_sum.assign((size_t)SN * o.np * o.nz, 0.0);
  o.st_ustar.assign(o.np, 0.0);
  o.st_iters = 0;
}

// One call of io_stat_averages (src/IO_statistics.f90:273, 319-324,
// 372-440), 1-rank: Mav_uw, Mav_vw reset at level 1, then k = 2..nz with u, v
// averaged to the w node; the other sums over k = 1..nz.  Integer powers as
// products: u**2 = u*u, u**3 = (u*u)*u, u**4 = (u*u)*(u*u).
static void stat_accumulate(Oracle& o) {
  if (o.st_sum.empty()) stat_reset(o);
  o.st_iters += 1;
  const size_t n = o.np * o.nz;
  double* S = o.st_sum.data();
  for (size_t q = 0; q < o.np; ++q) o.st_ustar[q] = o.st_ustar[q] + o.ustar[q];
  for (size_t q = 0; q < o.np; ++q) { S[SUW * n + q] = 0.0; S[SVW * n + q] = 0.0; }
  for (int k = 2; k <= o.nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      const size_t a = o.I(0, 0, k) + q, t = (size_t)(k - 1) * o.np + q;
      S[SUW * n + t] = S[SUW * n + t] + o.w[a] * (0.5 * (o.u[a - o.np] + o.u[a]));
      S[SVW * n + t] = S[SVW * n + t] + o.w[a] * (0.5 * (o.v[a - o.np] + o.v[a]));
    }
  for (int k = 1; k <= o.nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      const size_t a = o.I(0, 0, k) + q, t = (size_t)(k - 1) * o.np + q;
      const double u = o.u[a], v = o.v[a], w = o.w[a];
      auto add = [&](int f, double x) { S[f * n + t] = S[f * n + t] + x; };
      add(SU, u); add(SV, v); add(SW, w); add(SP, o.p[a]);
      add(SU2, u * u); add(SV2, v * v); add(SW2, w * w);
      add(SU3, (u * u) * u); add(SV3, (v * v) * v); add(SW3, (w * w) * w);
      add(SU4, (u * u) * (u * u)); add(SV4, (v * v) * (v * v)); add(SW4, (w * w) * (w * w));
      add(SUV, u * v);
      add(STXX, o.txx[a]); add(STYY, o.tyy[a]); add(STZZ, o.tzz[a]);
      add(STXY, o.txy[a]); add(STXZ, o.txz[a]); add(STYZ, o.tyz[a]);
      add(SDUDZ, o.dudz[a]); add(SDVDZ, o.dvdz[a]);
      add(SNUT, o.nu_t[a]); add(SCS, std::sqrt(o.cs_opt2[a]));
    }
}

// Plane mean as squeeze(mean(mean(X, 2), 1)): mean over y for each x, then
// over x.
template <typename Fn>
static double plane_mean(const Oracle& o, Fn X) {
  double sx = 0.0;
  for (int i = 0; i < o.nx; ++i) {
    double sy = 0.0;
    for (int j = 0; j < o.ny; ++j) sy += X((size_t)i + (size_t)o.nx * j);
    sx += sy / o.ny;
  }
  return sx / o.nx;
}

// Profiles of the time averages (sum / nt; SGS stresses negated by
// get_data_struct_RAV.m:31-34), as get_profile_struct.m:9-59 and
// compute_Cij.m:1-36 on one output window: colocate_var(X, 'w') is 0 at
// level 1 and (X(k-1) + X(k)) / 2 above; colocate_var(X, 'uvp') is
// (X(k) + X(k+1)) / 2 below level nz and 0 at nz.  Rows: u v w p var_u
// var_v var_w txx tyy tzz R13 txz St13 R23 tyz St23 tke cs nut C13 C23.
static void stat_profiles(const Oracle& o, double* prof) {
  const int nz = o.nz;
  const size_t n = o.np * o.nz;
  const double nt = (double)o.st_iters;
  auto A = [&](int f, int k, size_t q) { return o.st_sum[f * n + (size_t)(k - 1) * o.np + q] / nt; };
  for (int k = 1; k <= nz; ++k) {
    auto col = [&](int f, size_t q) { return k == 1 ? 0.0 : 0.5 * (A(f, k - 1, q) + A(f, k, q)); };
    auto var = [&](int f, int f2, int kk, size_t q) { const double m = A(f, kk, q); return A(f2, kk, q) - m * m; };
    const double u = plane_mean(o, [&](size_t q) { return A(SU, k, q); });
    const double v = plane_mean(o, [&](size_t q) { return A(SV, k, q); });
    const double w = plane_mean(o, [&](size_t q) { return A(SW, k, q); });
    const double p = plane_mean(o, [&](size_t q) { return A(SP, k, q); });
    const double vu = plane_mean(o, [&](size_t q) { return var(SU, SU2, k, q); });
    const double vv = plane_mean(o, [&](size_t q) { return var(SV, SV2, k, q); });
    const double vw = plane_mean(o, [&](size_t q) { return var(SW, SW2, k, q); });
    const double vw_uvp = plane_mean(o, [&](size_t q) {
      return k < nz ? 0.5 * (var(SW, SW2, k, q) + var(SW, SW2, k + 1, q)) : 0.0; });
    const double txx = plane_mean(o, [&](size_t q) { return -A(STXX, k, q); });
    const double tyy = plane_mean(o, [&](size_t q) { return -A(STYY, k, q); });
    const double tzz = plane_mean(o, [&](size_t q) { return -A(STZZ, k, q); });
    auto r13 = [&](size_t q) { return A(SUW, k, q) - col(SU, q) * A(SW, k, q); };
    auto r23 = [&](size_t q) { return A(SVW, k, q) - col(SV, q) * A(SW, k, q); };
    const double R13 = plane_mean(o, r13), R23 = plane_mean(o, r23);
    const double txz = plane_mean(o, [&](size_t q) { return -A(STXZ, k, q); });
    const double tyz = plane_mean(o, [&](size_t q) { return -A(STYZ, k, q); });
    const double st13 = plane_mean(o, [&](size_t q) { return r13(q) + (-A(STXZ, k, q)); });
    const double st23 = plane_mean(o, [&](size_t q) { return r23(q) + (-A(STYZ, k, q)); });
    const double C13 = plane_mean(o, [&](size_t q) { return col(SU, q) * A(SW, k, q); }) -
                       plane_mean(o, [&](size_t q) { return col(SU, q); }) * w;
    const double C23 = plane_mean(o, [&](size_t q) { return col(SV, q) * A(SW, k, q); }) -
                       plane_mean(o, [&](size_t q) { return col(SV, q); }) * w;
    const double cs = plane_mean(o, [&](size_t q) { return A(SCS, k, q); });
    const double nut = plane_mean(o, [&](size_t q) { return A(SNUT, k, q); });
    const double row[21] = {u, v, w, p, vu, vv, vw, txx, tyy, tzz, R13, txz, st13 + C13,
                            R23, tyz, st23 + C23, 0.5 * (vu + vv + vw_uvp), cs, nut, C13, C23};
    for (int r = 0; r < 21; ++r) prof[(size_t)r * nz + (k - 1)] = row[r];
  }
}

// ---------------------------------------------------------------- TKE budget
// compute_tke_budget (src/tke_budget1D.f90:209-390), 1 rank: 61 quantities
// on the uvp nodes of levels 1..nz, each reduced to its plane average
// (compute_planar_average, src/utility_tools.f90:59-91: one running sum per
// level over j, then i, divided by nx*ny) and added to its per-level sum
// (avg_var, src/tke_budget1D.f90:397-440).  I(f) = (f(k) + f(k+1)) / 2 is
// int_var (:448-468); with the wall law the level-1 txz, tyz and the
// level-1 dudz + dwdx, dvdz + dwdy sums use the reference's special forms
// (:284-313).  Integer powers as products: x**2 = x*x, x**3 = (x*x)*x.
// The cleaned pressure p - (u*u + v*v + W*W) / 2 on levels 1..nz goes
// through grad_velocity_filt (filtered field, d/dx, d/dy) and ddz_uv; its
// ghost planes 0 and nz+1 are 0 here (the reference leaves them
// uninitialised: update_ghost does nothing on one rank in z,
// src/decomp_2d_custom.f90:364-366), which reaches dpdz at levels 1 and nz.
enum {
  TK_U = 0, TK_V, TK_W, TK_DUDX, TK_DUDY, TK_DUDZ, TK_DVDX, TK_DVDY, TK_DVDZ, TK_DWDX, TK_DWDY, TK_DWDZ,
  TK_UU, TK_VV, TK_WW, TK_UV, TK_UW, TK_VW, TK_TXX, TK_TYY, TK_TZZ, TK_TXY, TK_TXZ, TK_TYZ, TK_CS,
  TK_UUW, TK_VVW, TK_WWW, TK_UWW, TK_WTZZ, TK_UTXZ, TK_VTYZ, TK_DXX, TK_DYY, TK_DZZ, TK_DXY, TK_DXZ,
  TK_DYZ, TK_DUDX2, TK_DUDY2, TK_DUDZ2, TK_DVDX2, TK_DVDY2, TK_DVDZ2, TK_DWDX2, TK_DWDY2, TK_DWDZ2,
  TK_DUDXDWDX, TK_DUDYDWDY, TK_DUDZDWDZ, TK_P, TK_DPDX, TK_DPDY, TK_DPDZ, TK_PW, TK_PU, TK_PDUDX,
  TK_PDVDY, TK_PDWDZ, TK_PDWDX, TK_PDUDZ, TK_N
};
static_assert(TK_N == 61, "TKE budget quantities");

static void tke_reset(Oracle& o) {
  o.tk_sum.assign((size_t)TK_N * o.nz, 0.0);
  o.tk_abs.assign((size_t)TK_N * o.nz, 0.0);
}

// the 61 values at point q of level k; pf, px, py: the filtered cleaned
// pressure and its x, y derivatives (all planes 0..nz+1)
static void tke_point(const Oracle& o, int k, size_t q, const std::vector<double>& pf, const std::vector<double>& px,
                      const std::vector<double>& py, double* x) {
  const size_t a = o.I(0, 0, k) + q, b = a + o.np;
  auto I = [&](const std::vector<double>& f) { return 0.5 * (f[a] + f[b]); };
  const double u = o.u[a], v = o.v[a], W = PLANT(43) ? o.w[a] : I(o.w);
  const double Idudz = I(o.dudz), Idvdz = I(o.dvdz), Idwdx = I(o.dwdx), Idwdy = I(o.dwdy);
  // txz, tyz on the uvp node; the wall law keeps level 1's own values
  const double Txz = k == 1 && !PLANT(45) ? o.txz[a] : I(o.txz);
  const double Tyz = k == 1 ? o.tyz[a] : I(o.tyz);
  const double S13 = k == 1 ? o.dudz[a] + 0.5 * (o.dwdx[a] + o.dwdx[b])
                            : 0.5 * ((o.dudz[a] + o.dwdx[a]) + (o.dudz[b] + o.dwdx[b]));
  const double S23 = k == 1 ? o.dvdz[a] + 0.5 * (o.dwdy[a] + o.dwdy[b])
                            : 0.5 * ((o.dvdz[a] + o.dwdy[a]) + (o.dvdz[b] + o.dwdy[b]));
  const double c = 1.0 / o.dz;
  const double Pf = pf[a];
  const double dpdz = PLANT(42) ? c * (pf[a] - pf[a - o.np]) : 0.5 * (c * (pf[a] - pf[a - o.np]) + c * (pf[b] - pf[a]));
  x[TK_U] = u; x[TK_V] = v; x[TK_W] = W;
  x[TK_DUDX] = o.dudx[a]; x[TK_DUDY] = o.dudy[a]; x[TK_DUDZ] = Idudz;
  x[TK_DVDX] = o.dvdx[a]; x[TK_DVDY] = o.dvdy[a]; x[TK_DVDZ] = Idvdz;
  x[TK_DWDX] = Idwdx; x[TK_DWDY] = Idwdy; x[TK_DWDZ] = o.dwdz[a];
  x[TK_UU] = u * u; x[TK_VV] = v * v; x[TK_WW] = W * W;
  x[TK_UV] = u * v; x[TK_UW] = W * u; x[TK_VW] = W * v;
  x[TK_TXX] = o.txx[a]; x[TK_TYY] = o.tyy[a]; x[TK_TZZ] = o.tzz[a];
  x[TK_TXY] = o.txy[a]; x[TK_TXZ] = I(o.txz); x[TK_TYZ] = I(o.tyz);
  x[TK_CS] = o.cs_opt2[a];
  x[TK_UUW] = (u * u) * W; x[TK_VVW] = (v * v) * W; x[TK_WWW] = (W * W) * W; x[TK_UWW] = PLANT(44) ? (u * u) * W : u * (W * W);
  x[TK_WTZZ] = W * o.tzz[a]; x[TK_UTXZ] = Txz * u; x[TK_VTYZ] = Tyz * v;
  x[TK_DXX] = o.dudx[a] * o.txx[a]; x[TK_DYY] = o.dvdy[a] * o.tyy[a]; x[TK_DZZ] = o.dwdz[a] * o.tzz[a];
  x[TK_DXY] = (PLANT(46) ? 1.0 : 0.5) * (o.dudy[a] + o.dvdx[a]) * o.txy[a];
  x[TK_DXZ] = (0.5 * S13) * Txz; x[TK_DYZ] = (0.5 * Tyz) * S23;
  x[TK_DUDX2] = o.dudx[a] * o.dudx[a]; x[TK_DUDY2] = o.dudy[a] * o.dudy[a]; x[TK_DUDZ2] = Idudz * Idudz;
  x[TK_DVDX2] = o.dvdx[a] * o.dvdx[a]; x[TK_DVDY2] = o.dvdy[a] * o.dvdy[a]; x[TK_DVDZ2] = Idvdz * Idvdz;
  x[TK_DWDX2] = Idwdx * Idwdx; x[TK_DWDY2] = Idwdy * Idwdy; x[TK_DWDZ2] = o.dwdz[a] * o.dwdz[a];
  x[TK_DUDXDWDX] = o.dudx[a] * Idwdx; x[TK_DUDYDWDY] = o.dudy[a] * Idwdy; x[TK_DUDZDWDZ] = Idudz * o.dwdz[a];
  x[TK_P] = Pf; x[TK_DPDX] = px[a]; x[TK_DPDY] = py[a]; x[TK_DPDZ] = dpdz;
  x[TK_PW] = Pf * W; x[TK_PU] = Pf * u;
  x[TK_PDUDX] = Pf * o.dudx[a]; x[TK_PDVDY] = Pf * o.dvdy[a]; x[TK_PDWDZ] = Pf * o.dwdz[a];
  x[TK_PDWDX] = Pf * Idwdx; x[TK_PDUDZ] = Pf * Idudz;
}

static void tke_accumulate(Oracle& o) {
  if (o.tk_sum.empty()) tke_reset(o);
  std::vector<double> pf(o.nf, 0.0), px(o.nf), py(o.nf);
  for (int k = 1; k <= o.nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      const size_t a = o.I(0, 0, k) + q;
      const double h = 0.5 * (o.w[a] + o.w[a + o.np]);
      pf[a] = o.p[a] - 0.5 * ((o.u[a] * o.u[a] + o.v[a] * o.v[a]) + (PLANT(41) ? 0.0 : h * h));
    }
  grad_filt(o, pf, px, py);
  const double nxy = (double)(o.nx * o.ny);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= o.nz; ++k) {
    double sum[TK_N] = {0.0}, asum[TK_N] = {0.0}, x[TK_N];
    for (size_t q = 0; q < o.np; ++q) {
      tke_point(o, k, q, pf, px, py, x);
      for (int f = 0; f < TK_N; ++f) {
        sum[f] = sum[f] + x[f];
        asum[f] = asum[f] + std::fabs(x[f]);
      }
    }
    for (int f = 0; f < TK_N; ++f) o.tk_abs[(size_t)f * o.nz + (k - 1)] += asum[f] / nxy;
    for (int f = 0; f < TK_N; ++f)
      o.tk_sum[(size_t)f * o.nz + (k - 1)] = (PLANT(47) ? 0.0 : o.tk_sum[(size_t)f * o.nz + (k - 1)]) + sum[f] / nxy;
  }
}

// compute_running_diagnistics (src/running_diagnostics.f90:80-305), 1 rank:
// CFL = dt (max|u|/dx + max|v|/dy + max|w|/dz) over levels 1..nz; MKE =
// (U^2 + V^2 + W^2) / 2 of the domain means (w averaged to the uvp node);
// plane mean of u*; rms of (du/dx + dv/dy + dw/dz)**2 over levels 1..nz.
static void diagnostics(const Oracle& o, double* out) {
  double um = -1234567890.0, vm = -1234567890.0, wm = -1234567890.0;
  double us = 0.0, vs = 0.0, ws = 0.0, rms = 0.0;
  for (int k = 1; k <= o.nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      const size_t a = o.I(0, 0, k) + q;
      if (std::fabs(o.u[a]) > um) um = std::fabs(o.u[a]);
      if (std::fabs(o.v[a]) > vm) vm = std::fabs(o.v[a]);
      if (std::fabs(o.w[a]) > wm) wm = std::fabs(o.w[a]);
      us = us + o.u[a];
      vs = vs + o.v[a];
      ws = ws + 0.5 * (o.w[a] + o.w[a + o.np]);
      const double d = o.dudx[a] + o.dvdy[a] + o.dwdz[a];
      rms = rms + d * d;
    }
  const double nxyz = (double)o.nx * (double)o.ny * (double)o.nz;
  out[0] = o.P.dt * (um / o.dx + vm / o.dy + wm / o.dz);
  us = us / nxyz; vs = vs / nxyz; ws = ws / nxyz;
  out[1] = 0.5 * (us * us + vs * vs + ws * ws);
  double s = 0.0;
  for (size_t q = 0; q < o.np; ++q) s = s + o.ustar[q];
  out[2] = s / (double)(o.nx * o.ny);
  out[3] = std::sqrt(rms / nxyz);
}

}  // namespace orc

using namespace orc;

// ------------------------------------------------------------------ C API
extern "C" {

// Field ids shared with the tests (same numbering as the product's les_field).
enum {
  F_U = 0, F_V, F_W, F_DUDX, F_DUDY, F_DUDZ, F_DVDX, F_DVDY, F_DVDZ, F_DWDX, F_DWDY, F_DWDZ,
  F_P, F_DPDX, F_DPDY, F_DPDZ, F_RHSX, F_RHSY, F_RHSZ, F_RHSXF, F_RHSYF, F_RHSZF,
  F_TXX, F_TXY, F_TXZ, F_TYY, F_TYZ, F_TZZ, F_CS2, F_NUT, F_FLM, F_FMM, F_FQN, F_FNN, F_BETA,
  F_ULAG, F_VLAG, F_WLAG, F_S11, F_S12, F_S13, F_S22, F_S23, F_S33, F_COUNT
};

static std::vector<double>* field(Oracle& o, int id) {
  switch (id) {
    case F_U: return &o.u; case F_V: return &o.v; case F_W: return &o.w;
    case F_DUDX: return &o.dudx; case F_DUDY: return &o.dudy; case F_DUDZ: return &o.dudz;
    case F_DVDX: return &o.dvdx; case F_DVDY: return &o.dvdy; case F_DVDZ: return &o.dvdz;
    case F_DWDX: return &o.dwdx; case F_DWDY: return &o.dwdy; case F_DWDZ: return &o.dwdz;
    case F_P: return &o.p; case F_DPDX: return &o.dpdx; case F_DPDY: return &o.dpdy; case F_DPDZ: return &o.dpdz;
    case F_RHSX: return &o.rhsx; case F_RHSY: return &o.rhsy; case F_RHSZ: return &o.rhsz;
    case F_RHSXF: return &o.rhsx_f; case F_RHSYF: return &o.rhsy_f; case F_RHSZF: return &o.rhsz_f;
    case F_TXX: return &o.txx; case F_TXY: return &o.txy; case F_TXZ: return &o.txz;
    case F_TYY: return &o.tyy; case F_TYZ: return &o.tyz; case F_TZZ: return &o.tzz;
    case F_CS2: return &o.cs_opt2; case F_NUT: return &o.nu_t;
    case F_FLM: return &o.F_LM; case F_FMM: return &o.F_MM; case F_FQN: return &o.F_QN; case F_FNN: return &o.F_NN;
    case F_BETA: return &o.beta; case F_ULAG: return &o.u_lag; case F_VLAG: return &o.v_lag; case F_WLAG: return &o.w_lag;
    case F_S11: return &o.s11; case F_S12: return &o.s12; case F_S13: return &o.s13;
    case F_S22: return &o.s22; case F_S23: return &o.s23; case F_S33: return &o.s33;
  }
  return nullptr;
}

// ip: nx ny nz sgs_model cs_count dyn_init lagran_init wall_model ubc_u_neu ubc_v_neu halo_mode coriolis dealias
// rp: lx ly lz dt z0 pfx pfy cs0 coriol ug vg ubc_u ubc_v ubc_w
int orc_init(const int* ip, const double* rp) {
  delete g;
  g = new Oracle();
  Oracle& o = *g;
  Params& P = o.P;
  P.nx = ip[0]; P.ny = ip[1]; P.nz = ip[2]; P.sgs_model = ip[3]; P.cs_count = ip[4]; P.dyn_init = ip[5];
  P.lagran_init = ip[6]; P.wall_model = ip[7]; P.ubc_u_neu = ip[8]; P.ubc_v_neu = ip[9]; P.halo_mode = ip[10];
  P.coriolis = ip[11]; P.dealias = ip[12];
  if (P.dealias < 0 || P.dealias > 3) return 1;
  P.lx = rp[0]; P.ly = rp[1]; P.lz = rp[2]; P.dt = rp[3]; P.z0 = rp[4]; P.pfx = rp[5]; P.pfy = rp[6];
  P.cs0 = rp[7]; P.coriol = rp[8]; P.ug = rp[9]; P.vg = rp[10]; P.ubc_u = rp[11]; P.ubc_v = rp[12]; P.ubc_w = rp[13];
  P.re = rp[14];
  o.nx = P.nx; o.ny = P.ny; o.nz = P.nz; o.nzg = P.nz + 2; o.lhx = P.nx / 2 + 1;
  o.np = (size_t)o.nx * o.ny;
  o.nf = o.np * o.nzg;
  o.dx = P.lx / (double)P.nx; o.dy = P.ly / (double)P.ny; o.dz = P.lz / (double)P.nz;
  o.delta = std::pow(o.dx * o.dy * o.dz, 1.0 / 3.0);
  o.px.build(o.nx); o.py.build(o.ny); o.px2.build(3 * o.nx / 2); o.py2.build(3 * o.ny / 2);
  if (!o.px.n || !o.py.n || !o.px2.n || !o.py2.n || (o.nx % 4) || (o.ny % 4)) return 1;
  build_tables(o);
  o.logterm.assign(o.np, std::log(((PLANT(36) ? 1.0 : 0.5) * o.dz) / P.z0));
  std::vector<double>* all[] = {&o.u, &o.v, &o.w, &o.dudx, &o.dudy, &o.dudz, &o.dvdx, &o.dvdy, &o.dvdz, &o.dwdx,
                                &o.dwdy, &o.dwdz, &o.p, &o.dpdx, &o.dpdy, &o.dpdz, &o.rhsx, &o.rhsy, &o.rhsz,
                                &o.rhsx_f, &o.rhsy_f, &o.rhsz_f, &o.txx, &o.txy, &o.txz, &o.tyy, &o.tyz, &o.tzz,
                                &o.cs_opt2, &o.nu_t, &o.F_LM, &o.F_MM, &o.F_QN, &o.F_NN, &o.beta, &o.u_lag,
                                &o.v_lag, &o.w_lag, &o.s11, &o.s12, &o.s13, &o.s22, &o.s23, &o.s33};
  for (auto* f : all) f->assign(o.nf, 0.0);
  o.lbc_u.assign(o.np, 0.0); o.lbc_v.assign(o.np, 0.0);
  o.bottom_wk.assign(o.np, 0.0); o.top_wk.assign(o.np, 0.0); o.ustar.assign(o.np, 0.0);
  o.jt_total = 0;
  o.jt = 0;
  o.F_coeff_init = P.lagran_init != 0;
  o.lagran_flag = P.lagran_init != 0;
  return 0;
}

int orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); return omp_get_max_threads(); }

// Initial state from interior velocities (nx, ny, nz) as read by ic_type='files'
// (src/initial_conditions.f90:130-218): BCs with lbc = 0 (Dirichlet bottom for
// u, v before the wall-law switch), then w(0) = BOGUS.
int orc_set_ic(const double* u0, const double* v0, const double* w0) {
  if (!g) return 1;
  Oracle& o = *g;
  int nz = o.nz;
  for (int k = 1; k <= nz; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t s = (size_t)(k - 1) * o.np + q;
      o.u[o.I(0, 0, k) + q] = u0[s]; o.v[o.I(0, 0, k) + q] = v0[s]; o.w[o.I(0, 0, k) + q] = w0[s];
    }
  for (size_t q = 0; q < o.np; ++q) {
    o.u[q] = 2.0 * 0.0 - o.u[o.I(0, 0, 1) + q];
    o.v[q] = 2.0 * 0.0 - o.v[o.I(0, 0, 1) + q];
    o.w[o.I(0, 0, 1) + q] = 0.0;
    size_t t = o.I(0, 0, nz + 1) + q, b = o.I(0, 0, nz) + q;
    o.u[t] = o.P.ubc_u_neu ? o.u[b] + o.P.ubc_u * o.dz : 2.0 * o.P.ubc_u - o.u[b];
    o.v[t] = o.P.ubc_v_neu ? o.v[b] + o.P.ubc_v * o.dz : 2.0 * o.P.ubc_v - o.v[b];
    o.w[t] = o.P.ubc_w;
    o.w[q] = BOGUS;
  }
  return 0;
}

int orc_step(int n) {
  if (!g) return 1;
  for (int s = 0; s < n; ++s) step(*g);
  return 0;
}

int orc_get(int id, double* out) {
  if (!g) return 1;
  std::vector<double>* f = field(*g, id);
  if (!f) return 2;
  std::memcpy(out, f->data(), sizeof(double) * g->nf);
  return 0;
}

int orc_put(int id, const double* in) {
  if (!g) return 1;
  std::vector<double>* f = field(*g, id);
  if (!f) return 2;
  std::memcpy(f->data(), in, sizeof(double) * g->nf);
  return 0;
}

int orc_get_2d(int which, double* out) {
  if (!g) return 1;
  std::vector<double>* f = which == 0 ? &g->lbc_u : which == 1 ? &g->lbc_v : which == 2 ? &g->ustar
                         : which == 3 ? &g->bottom_wk : &g->top_wk;
  std::memcpy(out, f->data(), sizeof(double) * g->np);
  return 0;
}

int orc_jt_total(void) { return g ? g->jt_total : -1; }

// ---- unit-level entry points used by the tests
// forward 2D transform of nplanes real planes (nx, ny) -> (lhx, ny) complex (interleaved)
int orc_fft2_fwd(int nx, int ny, int nplanes, const double* in, double* out, double scale) {
  Oracle o;
  o.nx = nx; o.ny = ny; o.lhx = nx / 2 + 1;
  o.px.build(nx); o.py.build(ny);
  if (!o.px.n || !o.py.n || (ny % 2)) return 1;
  for (int k = 0; k < nplanes; ++k)
    fft2_fwd(o, in + (size_t)k * nx * ny, (Cx*)out + (size_t)k * o.lhx * ny, scale);
  return 0;
}

int orc_fft2_inv(int nx, int ny, int nplanes, const double* in, double* out) {
  Oracle o;
  o.nx = nx; o.ny = ny; o.lhx = nx / 2 + 1;
  o.px.build(nx); o.py.build(ny);
  if (!o.px.n || !o.py.n || (ny % 2)) return 1;
  std::vector<Cx> sp((size_t)o.lhx * ny);
  for (int k = 0; k < nplanes; ++k) {
    std::memcpy(sp.data(), (const Cx*)in + (size_t)k * o.lhx * ny, sizeof(Cx) * sp.size());
    fft2_inv(o, sp.data(), out + (size_t)k * nx * ny);
  }
  return 0;
}

// complex 1D transform of length n (interleaved re/im)
int orc_fft1(int n, int inverse, double* data) {
  Plan p;
  p.build(n);
  if (!p.n) return 1;
  std::vector<Cx> s(n);
  p.exec((Cx*)data, s.data(), inverse != 0);
  return 0;
}

// 3/2 padding and truncation of single planes
int orc_pad_plane(int nx, int ny, const double* in, double* out) {
  Oracle o;
  o.nx = nx; o.ny = ny; o.lhx = nx / 2 + 1;
  o.px.build(nx); o.py.build(ny); o.px2.build(3 * nx / 2); o.py2.build(3 * ny / 2);
  pad_plane(o, in, out);
  return 0;
}
int orc_trunc_plane(int nx, int ny, const double* in, double* out) {
  Oracle o;
  o.nx = nx; o.ny = ny; o.lhx = nx / 2 + 1;
  o.px.build(nx); o.py.build(ny); o.px2.build(3 * nx / 2); o.py2.build(3 * ny / 2);
  trunc_plane(o, in, out);
  return 0;
}

// one tridiagonal solve of the pressure system for mode (i, j) of the current
// oracle grid; r, x are (nz+3) complex, entries 1..nz+2 used
int orc_tridiag(int i, int j, const double* r, double* x) {
  if (!g) return 1;
  tridiag(*g, i, j, (const Cx*)r, (Cx*)x);
  return 0;
}

// spectral tables of the current grid: which 0 kkx, 1 kky, 2 kk2, 3 G(2 delta),
// 4 G(4 delta), 5 dealiasing multiplier of the sources, 6 of the products
int orc_table(int which, double* out) {
  if (!g) return 1;
  std::vector<double>* t = which == 0 ? &g->kkx : which == 1 ? &g->kky : which == 2 ? &g->kk2 : which == 3 ? &g->G2
                         : which == 4 ? &g->G4 : which == 5 ? &g->GA : &g->GB;
  if (t->empty()) return 2;
  std::memcpy(out, t->data(), sizeof(double) * t->size());
  return 0;
}

// stand-alone operators on the current state (for operator-level tests)
int orc_op(int which, double arg) {
  if (!g) return 1;
  Oracle& o = *g;
  switch (which) {
    case 0: wall_law(o); break;
    case 1: compute_vel_grad(o); break;
    case 2: sgs_stress(o, (int)arg, (int)arg); break;
    case 3: compute_rhs(o); break;
    case 4: pressure(o, arg); break;
    case 5: compute_sij(o); break;
    case 6: convection(o, o.rhsx, o.rhsy, o.rhsz); break;
    case 7: lagrangian_interp(o); break;
    case 8: set_bc_all(o); break;
    case 9: dynamic_coeff(o); break;
    default: return 2;
  }
  return 0;
}

int orc_set_jt_total(int jt) { if (!g) return 1; g->jt_total = jt; return 0; }
int orc_jt(void) { return g ? g->jt : -1; }
// sgs_stress(jt, jtt) with the two step counts given separately
int orc_sgs_stress(int jt, int jtt) { if (!g) return 1; sgs_stress(*g, jt, jtt); return 0; }

// The LASD update's parts on their own, for the tests that pin each of them:
// update_interpolag_coeff of (F_LM, F_MM) (which = 0) or (F_QN, F_NN)
// (which = 1) with the given AB, BB fields (nx, ny, nz+2); and
// compute_dynamic_mum at both filter widths from the current u, v, w and S_ij
// (LM, MM, QN, NN out, (nx, ny, nz+2), levels 1..nz+1 written).
int orc_update_F(int which, const double* AB, const double* BB) {
  if (!g) return 1;
  std::vector<double> ab(AB, AB + g->nf), bb(BB, BB + g->nf);
  if (which == 0) update_F(*g, g->F_LM, g->F_MM, ab, bb);
  else update_F(*g, g->F_QN, g->F_NN, ab, bb);
  return 0;
}
int orc_dynamic_mum(double* LM, double* MM, double* QN, double* NN) {
  if (!g) return 1;
  Oracle& o = *g;
  std::vector<double> Smod(o.nf, 0.0), lm(o.nf, 0.0), mm(o.nf, 0.0), qn(o.nf, 0.0), nn(o.nf, 0.0);
  for (int k = 1; k <= o.nz + 1; ++k)
    for (size_t q = 0; q < o.np; ++q) {
      size_t a = o.I(0, 0, k) + q;
      Smod[a] = smag(o.s11[a], o.s12[a], o.s13[a], o.s22[a], o.s23[a], o.s33[a]);
    }
  dynamic_mum(o, Smod, lm, mm, qn, nn);
  std::memcpy(LM, lm.data(), sizeof(double) * o.nf);
  std::memcpy(MM, mm.data(), sizeof(double) * o.nf);
  std::memcpy(QN, qn.data(), sizeof(double) * o.nf);
  std::memcpy(NN, nn.data(), sizeof(double) * o.nf);
  return 0;
}

// running statistics and diagnostics
int orc_stat_reset(void) { if (!g) return 1; stat_reset(*g); return 0; }
int orc_stat_accumulate(void) { if (!g) return 1; stat_accumulate(*g); return 0; }
int orc_stat_get(double* sums3d, double* ustar2d, int* iterations) {
  if (!g) return 1;
  if (g->st_sum.empty()) stat_reset(*g);
  if (sums3d) std::memcpy(sums3d, g->st_sum.data(), sizeof(double) * g->st_sum.size());
  if (ustar2d) std::memcpy(ustar2d, g->st_ustar.data(), sizeof(double) * g->np);
  if (iterations) *iterations = g->st_iters;
  return 0;
}
int orc_stat_profiles(double* prof) {
  if (!g || g->st_iters < 1) return 1;
  stat_profiles(*g, prof);
  return 0;
}
int orc_diagnostics(double* out) { if (!g) return 1; diagnostics(*g, out); return 0; }
int orc_set_geostrophic(double ug, double vg) {
  if (!g) return 1;
  g->P.ug = ug;
  g->P.vg = vg;
  return 0;
}
// TKE budget sums (TK_N x nz, row-major by quantity)
int orc_tke_reset(void) { if (!g) return 1; tke_reset(*g); return 0; }
int orc_tke_accumulate(void) { if (!g) return 1; tke_accumulate(*g); return 0; }
int orc_tke_get(double* sums, double* abs_sums) {
  if (!g) return 1;
  if (g->tk_sum.empty()) tke_reset(*g);
  std::memcpy(sums, g->tk_sum.data(), sizeof(double) * g->tk_sum.size());
  if (abs_sums) std::memcpy(abs_sums, g->tk_abs.data(), sizeof(double) * g->tk_abs.size());
  return 0;
}
int orc_set_2d(int which, const double* in) {
  if (!g) return 1;
  if (which == 5) {  // a surface roughness map z0(x, y): the wall law's log(0.5 dz / z0) per point
    for (size_t q = 0; q < g->np; ++q) g->logterm[q] = std::log((0.5 * g->dz) / in[q]);
    return 0;
  }
  std::vector<double>* f = which == 0 ? &g->lbc_u : which == 1 ? &g->lbc_v : which == 2 ? &g->ustar
                         : which == 3 ? &g->bottom_wk : &g->top_wk;
  std::memcpy(f->data(), in, sizeof(double) * g->np);
  return 0;
}

}  // extern "C"

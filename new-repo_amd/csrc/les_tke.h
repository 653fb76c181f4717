// TKE budget sums (src/tke_budget1D.f90:209-440): per call, the plane
// averages of 61 quantities on the uvp nodes of levels 1..nz, added to
// per-level window sums (avg_var).  The quantities, in the order of the
// reference's avg_var calls (and of its ta1_profiles/<name>.out files):
//   u v W dudx dudy I(dudz) dvdx dvdy I(dvdz) I(dwdx) I(dwdy) dwdz
//   uu vv WW uv Wu Wv txx tyy tzz txy I(txz) I(tyz) cs
//   (uu)W (vv)W (WW)W u(WW) W tzz Txz u Tyz v
//   dudx txx, dvdy tyy, dwdz tzz, (S12) txy, (S13 / 2) Txz, (Tyz / 2) S23
//   the nine squared gradients on the uvp node, dudx I(dwdx), dudy I(dwdy),
//   I(dudz) dwdz, then of the cleaned pressure Pf: Pf, dPf/dx, dPf/dy,
//   I(dPf/dz), Pf W, Pf u, Pf dudx, Pf dvdy, Pf dwdz, Pf I(dwdx), Pf I(dudz)
// with I(f) = (f(k) + f(k+1)) / 2 (int_var), W = I(w), Txz, Tyz = I(txz),
// I(tyz) except at level 1 (the wall law's own values), S13 = I(dudz + dwdx)
// and S23 = I(dvdz + dwdy) except dudz(1) + (dwdx(1) + dwdx(2)) / 2 and
// dvdz(1) + (dwdy(1) + dwdy(2)) / 2 at level 1 (:284-313).  Pf is
// p - (u u + v v + W W) / 2 on levels 1..nz, 0 on the ghost planes, through
// grad_velocity_filt (filtered, d/dx, d/dy); dPf/dz = (Pf(k) - Pf(k-1)) / dz
// (ddz_uv).  The reference leaves those ghost planes uninitialised (its
// update_ghost does nothing on one rank in z), which reaches I(dPf/dz) at
// levels 1 and nz.  Every product in the reference's operand order; the
// plane sums in a fixed order (slices of the plane, each summed per thread
// then across the block's threads in thread order, the slices in order).
#pragma once
#include "les_common.h"

enum {
  TK_U = 0, TK_V, TK_W, TK_DUDX, TK_DUDY, TK_DUDZ, TK_DVDX, TK_DVDY, TK_DVDZ, TK_DWDX, TK_DWDY, TK_DWDZ,
  TK_UU, TK_VV, TK_WW, TK_UV, TK_UW, TK_VW, TK_TXX, TK_TYY, TK_TZZ, TK_TXY, TK_TXZ, TK_TYZ, TK_CS,
  TK_UUW, TK_VVW, TK_WWW, TK_UWW, TK_WTZZ, TK_UTXZ, TK_VTYZ, TK_DXX, TK_DYY, TK_DZZ, TK_DXY, TK_DXZ,
  TK_DYZ, TK_DUDX2, TK_DUDY2, TK_DUDZ2, TK_DVDX2, TK_DVDY2, TK_DVDZ2, TK_DWDX2, TK_DWDY2, TK_DWDZ2,
  TK_DUDXDWDX, TK_DUDYDWDY, TK_DUDZDWDZ, TK_P, TK_DPDX, TK_DPDY, TK_DPDZ, TK_PW, TK_PU, TK_PDUDX,
  TK_PDVDY, TK_PDWDZ, TK_PDWDX, TK_PDUDZ, TK_N
};
static_assert(TK_N == LES_NTKE, "TKE budget quantities");
#define LES_TKE_SLICES 16

template <typename T>
struct TkeArgs {
  const double *u, *v, *w;
  const T *dudx, *dudy, *dudz, *dvdx, *dvdy, *dvdz, *dwdx, *dwdy, *dwdz;
  const T *txx, *txy, *txz, *tyy, *tyz, *tzz, *cs;
  const double *pf, *px, *py;  // filtered cleaned pressure, its x and y derivatives
  double* part;                // [nz][LES_TKE_SLICES][TK_N]
  long np;
  int nz;
  double cz;  // 1 / dz
};

// the 61 values at in-plane point q of level k (1..nz)
template <typename T>
__device__ __forceinline__ void tke_point(const TkeArgs<T>& A, int k, long q, double* x) {
  const long a = (long)k * A.np + q, b = a + A.np;
  auto I = [&](const T* f) { return 0.5 * ((double)f[a] + (double)f[b]); };
  const double u = A.u[a], v = A.v[a], W = 0.5 * (A.w[a] + A.w[b]);
  const double dudx = A.dudx[a], dudy = A.dudy[a], dvdx = A.dvdx[a], dvdy = A.dvdy[a], dwdz = A.dwdz[a];
  const double Idudz = I(A.dudz), Idvdz = I(A.dvdz), Idwdx = I(A.dwdx), Idwdy = I(A.dwdy);
  const double txx = A.txx[a], tyy = A.tyy[a], tzz = A.tzz[a], txy = A.txy[a];
  const double Itxz = I(A.txz), Ityz = I(A.tyz);
  const double Txz = k == 1 ? (double)A.txz[a] : Itxz;
  const double Tyz = k == 1 ? (double)A.tyz[a] : Ityz;
  const double S13 = k == 1 ? (double)A.dudz[a] + 0.5 * ((double)A.dwdx[a] + (double)A.dwdx[b])
                            : 0.5 * (((double)A.dudz[a] + (double)A.dwdx[a]) + ((double)A.dudz[b] + (double)A.dwdx[b]));
  const double S23 = k == 1 ? (double)A.dvdz[a] + 0.5 * ((double)A.dwdy[a] + (double)A.dwdy[b])
                            : 0.5 * (((double)A.dvdz[a] + (double)A.dwdy[a]) + ((double)A.dvdz[b] + (double)A.dwdy[b]));
  const double Pf = A.pf[a];
  const double dpdz = 0.5 * (A.cz * (A.pf[a] - A.pf[a - A.np]) + A.cz * (A.pf[b] - A.pf[a]));
  x[TK_U] = u; x[TK_V] = v; x[TK_W] = W;
  x[TK_DUDX] = dudx; x[TK_DUDY] = dudy; x[TK_DUDZ] = Idudz;
  x[TK_DVDX] = dvdx; x[TK_DVDY] = dvdy; x[TK_DVDZ] = Idvdz;
  x[TK_DWDX] = Idwdx; x[TK_DWDY] = Idwdy; x[TK_DWDZ] = dwdz;
  x[TK_UU] = u * u; x[TK_VV] = v * v; x[TK_WW] = W * W;
  x[TK_UV] = u * v; x[TK_UW] = W * u; x[TK_VW] = W * v;
  x[TK_TXX] = txx; x[TK_TYY] = tyy; x[TK_TZZ] = tzz;
  x[TK_TXY] = txy; x[TK_TXZ] = Itxz; x[TK_TYZ] = Ityz;
  x[TK_CS] = A.cs[a];
  x[TK_UUW] = (u * u) * W; x[TK_VVW] = (v * v) * W; x[TK_WWW] = (W * W) * W; x[TK_UWW] = u * (W * W);
  x[TK_WTZZ] = W * tzz; x[TK_UTXZ] = Txz * u; x[TK_VTYZ] = Tyz * v;
  x[TK_DXX] = dudx * txx; x[TK_DYY] = dvdy * tyy; x[TK_DZZ] = dwdz * tzz;
  x[TK_DXY] = (0.5 * (dudy + dvdx)) * txy;
  x[TK_DXZ] = (0.5 * S13) * Txz; x[TK_DYZ] = (0.5 * Tyz) * S23;
  x[TK_DUDX2] = dudx * dudx; x[TK_DUDY2] = dudy * dudy; x[TK_DUDZ2] = Idudz * Idudz;
  x[TK_DVDX2] = dvdx * dvdx; x[TK_DVDY2] = dvdy * dvdy; x[TK_DVDZ2] = Idvdz * Idvdz;
  x[TK_DWDX2] = Idwdx * Idwdx; x[TK_DWDY2] = Idwdy * Idwdy; x[TK_DWDZ2] = dwdz * dwdz;
  x[TK_DUDXDWDX] = dudx * Idwdx; x[TK_DUDYDWDY] = dudy * Idwdy; x[TK_DUDZDWDZ] = Idudz * dwdz;
  x[TK_P] = Pf; x[TK_DPDX] = A.px[a]; x[TK_DPDY] = A.py[a]; x[TK_DPDZ] = dpdz;
  x[TK_PW] = Pf * W; x[TK_PU] = Pf * u;
  x[TK_PDUDX] = Pf * dudx; x[TK_PDVDY] = Pf * dvdy; x[TK_PDWDZ] = Pf * dwdz;
  x[TK_PDWDX] = Pf * Idwdx; x[TK_PDUDZ] = Pf * Idudz;
}

// One block per (slice, level), all quantities in one pass over the fields:
// each thread sums its points of the slice, then the block's sums are
// combined in thread order, quantities 0..30 and then 31..60 through one
// LDS buffer (31 x 256 doubles, 63.5 KB).
#define LES_TKE_RED 31
template <typename T>
__global__ void __launch_bounds__(LES_BLOCK) k_tke_sums(TkeArgs<T> a) {
  __shared__ double red[LES_TKE_RED * LES_BLOCK];
  const int slice = blockIdx.x, k = blockIdx.y + 1;
  const long q0 = a.np * slice / LES_TKE_SLICES, q1 = a.np * (slice + 1) / LES_TKE_SLICES;
  double acc[TK_N];
#pragma unroll
  for (int f = 0; f < TK_N; ++f) acc[f] = 0.0;
  for (long q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    double x[TK_N];
    tke_point(a, k, q, x);
#pragma unroll
    for (int f = 0; f < TK_N; ++f) acc[f] += x[f];
  }
  double* out = a.part + ((long)(k - 1) * LES_TKE_SLICES + slice) * TK_N;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int f0 = h * LES_TKE_RED, nq = h == 0 ? LES_TKE_RED : TK_N - LES_TKE_RED;
    if (h) __syncthreads();  // the first half's sums have been read
#pragma unroll
    for (int f = 0; f < LES_TKE_RED; ++f)
      if (f < nq) red[f * LES_BLOCK + threadIdx.x] = acc[f0 + f];
    __syncthreads();
    if ((int)threadIdx.x < nq) {
      double s = 0.0;
      for (int t = 0; t < LES_BLOCK; ++t) s += red[threadIdx.x * LES_BLOCK + t];
      out[f0 + threadIdx.x] = s;
    }
  }
}

// sums[f][k-1] += (sum over the slices, in order) / (nx ny)
__global__ void k_tke_add(const double* part, double* sums, int nz, double nxy) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)TK_N * nz) return;
  const int f = (int)(t / nz), k = (int)(t % nz);
  double s = 0.0;
  for (int sl = 0; sl < LES_TKE_SLICES; ++sl) s += part[((long)k * LES_TKE_SLICES + sl) * TK_N + f];
  sums[t] = sums[t] + s / nxy;
}

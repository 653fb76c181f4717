// Point-wise and column-local operator kernels of the LES step.
// Each kernel cites the reference loop it replaces; the floating-point
// expressions keep the reference's operation order.
#pragma once
#include "les_common.h"

// ---------------------------------------------------------------- BCs + ddz
// set_bc bottom/top (src/boundary_condition_mpi.f90:33-63) fused with
// ddz_uv / ddz_w (src/compute_derivatives.f90:375-441), as called from
// compute_vel_grad (src/LES_engine.f90:464-478).  Bottom u, v are Neumann
// with the wall-law gradient (src/LES_engine.f90:168-171); w(1) = lbc_w = 0.
template <typename T>
struct BcDdzArgs {
  double *u, *v, *w;
  T *dudz, *dvdz, *dwdz;
  const double *lbc_u, *lbc_v;
  long np;
  int nz;
  double dz, cz;
  int ubc_u_neu, ubc_v_neu;
  double ubc_u, ubc_v, ubc_w;
};

template <typename T>
__global__ void k_bc_ddz(BcDdzArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * (a.nz + 1)) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const long c = (long)k * a.np + q;
  const double cz = a.cz;
  // u and v on uvp nodes -> w nodes
  for (int comp = 0; comp < 2; ++comp) {
    double* f = comp == 0 ? a.u : a.v;
    T* d = comp == 0 ? a.dudz : a.dvdz;
    double fk, fm;
    if (k == 1) {
      fk = f[c];
      fm = fk - (comp == 0 ? a.lbc_u[q] : a.lbc_v[q]) * a.dz;
      f[q] = fm;
    } else if (k == a.nz + 1) {
      fm = f[c - a.np];
      const int neu = comp == 0 ? a.ubc_u_neu : a.ubc_v_neu;
      const double bc = comp == 0 ? a.ubc_u : a.ubc_v;
      fk = neu ? fm + bc * a.dz : 2.0 * bc - fm;
      f[c] = fk;
    } else {
      fk = f[c];
      fm = f[c - a.np];
    }
    d[c] = (T)(cz * (fk - fm));
  }
  // w on w nodes -> uvp nodes
  if (k <= a.nz) {
    double wk = (k == 1) ? 0.0 : a.w[c];
    double wp = (k == a.nz) ? a.ubc_w : a.w[c + a.np];
    if (k == 1) a.w[c] = 0.0;
    a.dwdz[c] = (T)(cz * (wp - wk));
  } else {
    a.w[c] = a.ubc_w;
    a.dwdz[c] = (T)0;
  }
}

// ---------------------------------------------------------------- wall law
// obukhov + wallstress, neutral log law (src/compute_wall_law.f90:107-254).
template <typename T>
struct WallArgs {
  const T *u1, *v1;          // 2-delta filtered first level
  const double *uraw, *vraw; // raw first level (wall_inst)
  int inst;
  const double* logterm;
  double *ustar, *lbc_u, *lbc_v;
  T *dudz1, *dvdz1, *txz1, *tyz1;
  double c1;
  long np;
};

template <typename T>
__global__ void k_wall(WallArgs<T> a) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.np) return;
  const double u1 = a.inst ? a.uraw[q] : (double)a.u1[q];
  const double v1 = a.inst ? a.vraw[q] : (double)a.v1[q];
  const double uavg = sqrt(u1 * u1 + v1 * v1);
  const double us = 0.4 * uavg / a.logterm[q];
  a.ustar[q] = us;
  if (uavg == 0.0) {
    a.dudz1[q] = (T)0; a.dvdz1[q] = (T)0;
    a.lbc_u[q] = 0.0; a.lbc_v[q] = 0.0;
    a.txz1[q] = (T)0; a.tyz1[q] = (T)0;
  } else {
    const double bu = us / a.c1 * u1 / uavg;
    const double bv = us / a.c1 * v1 / uavg;
    a.lbc_u[q] = bu; a.lbc_v[q] = bv;
    a.dudz1[q] = (T)bu; a.dvdz1[q] = (T)bv;
    a.txz1[q] = (T)((us * us) * u1 / uavg);
    a.tyz1[q] = (T)((us * us) * v1 / uavg);
  }
}

// ---------------------------------------------------------------- SGS
template <typename T>
struct SgsArgs {
  const T *dudx, *dudy, *dudz, *dvdx, *dvdy, *dvdz, *dwdx, *dwdy, *dwdz;
  const T* cs;
  T* nu_t;
  T *txx, *txy, *txz, *tyy, *tyz, *tzz;
  T *s11, *s12, *s13, *s22, *s23, *s33, *smod;
  const double *u, *v, *w;
  T *ulag, *vlag, *wlag;
  const double* lvec;
  long np;
  int nz;
  int accumulate, write_sij;
  int acc_kmax;  // accumulate u_lag on levels 0..acc_kmax only
};

// compute_sij (src/compute_sgs.f90:260-335): first level on uvp nodes with
// the wall-law du/dz, dv/dz; other levels averaged to w nodes.
template <typename T>
__device__ __forceinline__ void sij_at(const SgsArgs<T>& a, int k, long q, T* s) {
  const long c = (long)k * a.np + q;
  T ux, uy, uz, vx, vy, vz, wx, wy, wz;
  const T h = (T)0.5;
  if (k == 1) {
    ux = a.dudx[c]; uy = a.dudy[c]; uz = a.dudz[c]; vz = a.dvdz[c];
    vx = a.dvdx[c]; vy = a.dvdy[c];
    wx = h * (a.dwdx[c] + a.dwdx[c + a.np]);
    wy = h * (a.dwdy[c] + a.dwdy[c + a.np]);
    wz = a.dwdz[c];
  } else {
    const long b = c - a.np;
    ux = h * (a.dudx[c] + a.dudx[b]);
    uy = h * (a.dudy[c] + a.dudy[b]);
    uz = a.dudz[c];
    vx = h * (a.dvdx[c] + a.dvdx[b]);
    vy = h * (a.dvdy[c] + a.dvdy[b]);
    vz = a.dvdz[c];
    wx = a.dwdx[c];
    wy = a.dwdy[c];
    wz = h * (a.dwdz[c] + a.dwdz[b]);
  }
  s[0] = ux;
  s[1] = h * (uy + vx);
  s[2] = h * (uz + wx);
  s[3] = vy;
  s[4] = h * (vz + wy);
  s[5] = wz;
}

// |S| = sqrt(2 S:S) in the reference's order (src/compute_sgs.f90:181-182)
template <typename T>
__device__ __forceinline__ T smag(const T* s) {
  return sqrt((T)2 * (s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + (T)2 * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4])));
}

template <typename T>
__device__ __forceinline__ void lag_accumulate(const SgsArgs<T>& a, int k, long q) {
  // u_lag += u over k = 0..nz+1 (src/compute_LagScaleDep.f90:220-233)
  const long c = (long)k * a.np + q;
  if (k > a.acc_kmax) return;
  a.ulag[c] = (T)((double)a.ulag[c] + a.u[c]);
  a.vlag[c] = (T)((double)a.vlag[c] + a.v[c]);
  a.wlag[c] = (T)((double)a.wlag[c] + a.w[c]);
  if (k == 1) {
    a.ulag[q] = (T)((double)a.ulag[q] + a.u[q]);
    a.vlag[q] = (T)((double)a.vlag[q] + a.v[q]);
    a.wlag[q] = (T)((double)a.wlag[q] + a.w[q]);
  }
}

// S_ij and |S| to memory for the dynamic procedure (update steps)
template <typename T>
__global__ void k_sij(SgsArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * (a.nz + 1)) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const long c = (long)k * a.np + q;
  T s[6];
  sij_at(a, k, q, s);
  a.s11[c] = s[0]; a.s12[c] = s[1]; a.s13[c] = s[2];
  a.s22[c] = s[3]; a.s23[c] = s[4]; a.s33[c] = s[5];
  a.smod[c] = smag(s);
  if (a.accumulate) lag_accumulate(a, k, q);
}

// nu_t = |S| Cs^2 l^2 and tau_ij = 2 nu_t S_ij with the staggering of
// src/compute_sgs.f90:176-242 (txx..tzz on uvp with nu_t averaged, txz, tyz
// on w nodes, first level from the wall law, tzz(0) = tzz(1)).
template <typename T>
__global__ void k_sgs(SgsArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * (a.nz + 1)) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const long c = (long)k * a.np + q;
  T s[6];
  sij_at(a, k, q, s);
  const T lk = (T)a.lvec[k];
  const T nuk = smag(s) * a.cs[c] * (lk * lk);
  a.nu_t[c] = nuk;
  if (k <= a.nz) {
    T nu = nuk;
    if (k > 1) {
      T s2[6];
      sij_at(a, k + 1, q, s2);
      const T lk1 = (T)a.lvec[k + 1];
      const T nuk1 = smag(s2) * a.cs[c + a.np] * (lk1 * lk1);
      nu = (T)0.5 * (nuk + nuk1);
    }
    a.txx[c] = (T)2 * nu * a.dudx[c];
    a.txy[c] = (T)2 * nu * (T)0.5 * (a.dudy[c] + a.dvdx[c]);
    a.tyy[c] = (T)2 * nu * a.dvdy[c];
    const T tz = (T)2 * nu * a.dwdz[c];
    a.tzz[c] = tz;
    if (k == 1) a.tzz[q] = tz;
  }
  if (k >= 2) {
    a.txz[c] = (T)2 * nuk * s[2];
    a.tyz[c] = (T)2 * nuk * s[4];
  }
  if (a.accumulate) lag_accumulate(a, k, q);
}

// sgs_model dns (src/compute_sgs.f90:124-140): tau_ij = 2/Re S_ij on levels
// 1..nz, each component on the nodes of its gradients; S_ij, nu_t and Cs^2
// are not touched.
template <typename T>
__global__ void k_dns_stress(SgsArgs<T> a, T c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nz) return;
  const long i = a.np + t;  // levels 1..nz
  a.txx[i] = c * a.dudx[i];
  a.txy[i] = c * (T)0.5 * (a.dudy[i] + a.dvdx[i]);
  a.tyy[i] = c * a.dvdy[i];
  a.tzz[i] = c * a.dwdz[i];
  a.txz[i] = c * (T)0.5 * (a.dudz[i] + a.dwdx[i]);
  a.tyz[i] = c * (T)0.5 * (a.dvdz[i] + a.dwdy[i]);
}

template <typename T>
struct CsFillArgs {
  T* cs;
  T val;
  long np;
  int nz;
};

template <typename T>
__global__ void k_cs_fill(CsFillArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * (a.nz + 1)) return;
  a.cs[a.np + t] = a.val;
}

// ---------------------------------------------------------------- LASD inputs
// u_ft, v_ft, w_ft on w nodes (first level special, w_ft(1) = w(2)/4) and the
// products to be test-filtered (src/compute_LagScaleDep.f90:503-577).
template <typename T>
struct LasdInArgs {
  T* o[15];  // uf vf wf, A11 A12 A13 A22 A23 A33, S*S11 S*S12 S*S13 S*S22 S*S23 S*S33
  const double *u, *v, *w;
  const T* s[6];
  const T* smod;
  long np;
  int nz;
  int nlev;  // levels 1..nlev
};

template <typename T>
__global__ void k_lasd_inputs(LasdInArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nlev) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const long c = (long)k * a.np + q;
  T uf, vf, wf;
  if (k == 1) {
    uf = (T)a.u[c];
    vf = (T)a.v[c];
    wf = (T)(0.25 * a.w[c + a.np]);
  } else {
    uf = (T)(0.5 * (a.u[c] + a.u[c - a.np]));
    vf = (T)(0.5 * (a.v[c] + a.v[c - a.np]));
    wf = (T)a.w[c];
  }
  a.o[0][c] = uf; a.o[1][c] = vf; a.o[2][c] = wf;
  a.o[3][c] = uf * uf; a.o[4][c] = uf * vf; a.o[5][c] = uf * wf;
  a.o[6][c] = vf * vf; a.o[7][c] = vf * wf; a.o[8][c] = wf * wf;
  const T S = a.smod[c];
  for (int i = 0; i < 6; ++i) a.o[9 + i][c] = S * a.s[i][c];
}

// Germano contractions from the 21 filtered fields
// (src/compute_LagScaleDep.f90:542-594): AB = L:M, BB = M:M.
template <typename T>
struct GermanoArgs {
  const T* fl[21];
  T *AB, *BB;
  T tf2, cst;
  long np;
  int nz;
  int nlev;  // levels 1..nlev
};

template <typename T>
__global__ void k_germano(GermanoArgs<T> a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nlev) return;
  const long c = a.np + t;
  const T uf = a.fl[0][c], vf = a.fl[1][c], wf = a.fl[2][c];
  const T A11 = a.fl[3][c] - uf * uf, A12 = a.fl[4][c] - uf * vf, A13 = a.fl[5][c] - uf * wf;
  const T A22 = a.fl[6][c] - vf * vf, A23 = a.fl[7][c] - vf * wf, A33 = a.fl[8][c] - wf * wf;
  T s[6] = {a.fl[9][c], a.fl[10][c], a.fl[11][c], a.fl[12][c], a.fl[13][c], a.fl[14][c]};
  const T Sft = smag(s);
  const T B11 = a.cst * (a.fl[15][c] - a.tf2 * Sft * s[0]);
  const T B12 = a.cst * (a.fl[16][c] - a.tf2 * Sft * s[1]);
  const T B13 = a.cst * (a.fl[17][c] - a.tf2 * Sft * s[2]);
  const T B22 = a.cst * (a.fl[18][c] - a.tf2 * Sft * s[3]);
  const T B23 = a.cst * (a.fl[19][c] - a.tf2 * Sft * s[4]);
  const T B33 = a.cst * (a.fl[20][c] - a.tf2 * Sft * s[5]);
  a.AB[c] = A11 * B11 + A22 * B22 + A33 * B33 + (T)2 * (A12 * B12 + A13 * B13 + A23 * B23);
  a.BB[c] = B11 * B11 + B22 * B22 + B33 * B33 + (T)2 * (B12 * B12 + B13 * B13 + B23 * B23);
}

// ---------------------------------------------------------------- AB2
// (src/LES_engine.f90:253-283)
struct Ab2Args {
  double *u, *v, *w;
  const double *rx, *ry, *rz, *fx, *fy, *fz;
  double dt, t1, t2;
  long np;
  int nz;
};

__global__ void k_ab2(Ab2Args a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * (a.nz + 2)) return;
  const int k = (int)(t / a.np);
  if (k >= 1 && k <= a.nz) {
    a.u[t] = a.u[t] + a.dt * (a.t1 * a.rx[t] + a.t2 * a.fx[t]);
    a.v[t] = a.v[t] + a.dt * (a.t1 * a.ry[t] + a.t2 * a.fy[t]);
    a.w[t] = a.w[t] + a.dt * (a.t1 * a.rz[t] + a.t2 * a.fz[t]);
  } else {
    a.w[t] = 0.0;
  }
}

// ---------------------------------------------------------------- IC BCs
// set_initial_conditions tail (src/initial_conditions.f90:184-208): Dirichlet
// zero bottom for u, v (before the wall-law switch), w(1) = 0, top BCs, and
// w(:,:,0) = BOGUS.
struct IcBcArgs {
  double *u, *v, *w;
  long np;
  int nz;
  double dz;
  int ubc_u_neu, ubc_v_neu;
  double ubc_u, ubc_v, ubc_w, bogus;
};

__global__ void k_ic_bc(IcBcArgs a) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.np) return;
  const long k1 = a.np + q, kt = (long)(a.nz + 1) * a.np + q, kb = (long)a.nz * a.np + q;
  a.u[q] = 2.0 * 0.0 - a.u[k1];
  a.v[q] = 2.0 * 0.0 - a.v[k1];
  a.w[k1] = 0.0;
  a.u[kt] = a.ubc_u_neu ? a.u[kb] + a.ubc_u * a.dz : 2.0 * a.ubc_u - a.u[kb];
  a.v[kt] = a.ubc_v_neu ? a.v[kb] + a.ubc_v * a.dz : 2.0 * a.ubc_v - a.v[kb];
  a.w[kt] = a.ubc_w;
  a.w[q] = a.bogus;
}

// ---------------------------------------------------------------- NaN check
struct NanArgs {
  const double *u, *v, *w;
  long n, off;
  int* flag;
};

__global__ void k_nan(NanArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const long c = a.off + t;
  if (a.u[c] != a.u[c] || a.v[c] != a.v[c] || a.w[c] != a.w[c]) a.flag[0] = 1;
}

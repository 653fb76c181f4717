// Pressure Poisson solve and projection (src/compute_pressure.f90:116-372,
// src/LES_engine.f90:299-352).  The pressure path runs in FP64 in every
// precision mode.  Included by les_gpu.hip.
#pragma once

// Per-(kx, ky) tridiagonal system of length nz+2 solved by the Thomas
// algorithm, one thread per mode, with the divergence right-hand side built
// on the fly (src/compute_pressure.f90:205-216, 256-305, 448-493, 544-619).
struct TriArgs {
  const Cplx<double> *Hd, *Hz;        // y-transformed spectra, planes 1..nz: i kx Hx + i ky Hy, and Hz
  const Cplx<double> *bot, *top;      // BC layer spectra
  const double *kkx, *kky;
  Cplx<double>* phat;                 // planes 0..nz+1
  double* gam;                        // scratch: the bet each 16-level chunk starts from, rows at chunk * ns
  int lhx, ny, nz;
  long ns;
  double dz, idz2;
};

// Two lanes per mode, one per component: given the real factors bet and gam
// the recurrences of the real and imaginary parts are independent, so the
// pair computes the same factors and one component each (twice the
// parallelism of one thread per mode, whose few waves leave most SIMDs idle;
// adjacent lanes read and write the two halves of each complex value).  The
// sweeps are sequential in z but their loads are not: the loads of chunk n+1
// of LES_TRI_U levels are in flight while chunk n is computed.  Per
// component the arithmetic and its order are those of the plain sweep.
#ifndef LES_TRI_U
#define LES_TRI_U 16
#endif
// The sweep takes all of a wavefront's time and its registers allow one
// wavefront per SIMD, so only the modes that are solved get lanes: the
// (lhx-1) (ny-1) modes off the Nyquist row and column, 2 lanes each (at
// 256^2: 1020 wavefronts on 1024 SIMDs, one round; with every mode's lanes
// it took 1032, and the last few ran in a second round).  The first
// tri_zero_blocks blocks only zero the Nyquist modes' phat and finish at
// once.
__host__ __device__ inline long tri_nwork(int lhx, int ny) { return (long)(lhx - 1) * (ny - 1); }
__host__ __device__ inline long tri_zero_blocks(int lhx, int ny) { return ((long)lhx * ny - tri_nwork(lhx, ny) + 63) / 64; }
__global__ void __launch_bounds__(64) k_tridiag(TriArgs a) {
  const int nz = a.nz;
  const long ns = a.ns;
  const long nzb = tri_zero_blocks(a.lhx, a.ny);
  if (blockIdx.x < nzb) {
    // Nyquist modes, one lane each: the row j = ny/2 (every i), then the
    // column i = lhx-1 of the other rows
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (long)a.lhx + a.ny - 1) return;
    const int jn = (int)(m - a.lhx);
    const long q = m < a.lhx ? (long)(a.ny / 2) * a.lhx + m : (long)(jn < a.ny / 2 ? jn : jn + 1) * a.lhx + a.lhx - 1;
    for (int k = 0; k <= nz + 1; ++k) a.phat[(long)k * ns + q] = Cplx<double>{0.0, 0.0};
    return;
  }
  const long t = (long)(blockIdx.x - nzb) * blockDim.x + threadIdx.x;
  const long w = t >> 1;
  const int c = (int)(t & 1);
  if (w >= tri_nwork(a.lhx, a.ny)) return;
  const int i = (int)(w % (a.lhx - 1)), jw = (int)(w / (a.lhx - 1));
  const int j = jw < a.ny / 2 ? jw : jw + 1;
  const long q = (long)j * a.lhx + i;
  double* __restrict__ ph = reinterpret_cast<double*>(a.phat) + c;   // component c of phat
  double* __restrict__ gam = a.gam;
  const double* __restrict__ Hdc = reinterpret_cast<const double*>(a.Hd) + c;
  const double* __restrict__ Hzc = reinterpret_cast<const double*>(a.Hz) + c;
  const double dz = a.dz, idz2 = a.idz2;
  const double kx = a.kkx[q], ky = a.kky[q];
  const double bmid = -((kx * kx + ky * ky) + 2.0 / (dz * dz));
  const bool mean = (i == 0 && j == 0);
  // row 1: bottom layer (mean mode pinned)
  const double r1 = mean ? 0.0 : reinterpret_cast<const double*>(a.bot)[2 * q + c] * dz;
  double bet = mean ? 1.0 : -1.0;
  const double c1 = mean ? 0.0 : 1.0;
  double x = r1 / bet;
  ph[2 * q] = x;
  double hz = Hzc[2 * (ns + q)];
  // rows 2..nz+1 (interior levels 1..nz): r = (i kx Hx + i ky Hy) + dHz/dz,
  // the horizontal divergence formed by the column pass (real part
  // -(kx Hx.y) - ky Hy.y = -(kx Hx.y + ky Hy.y) exactly; imaginary part
  // kx Hx.x + ky Hy.x).  Software-pipelined: the loads of the next chunk of
  // LES_TRI_U levels are issued before the current chunk's recurrence runs.
  double hA[2][LES_TRI_U], hB[2][LES_TRI_U];
  auto fwd_load = [&](double (&h)[2][LES_TRI_U], int k0) {
#pragma unroll
    for (int u = 0; u < LES_TRI_U; ++u) {
      const int k = k0 + u;
      if (k <= nz + 1) {
        const long p = (long)(k - 1) * ns + q;
        h[0][u] = Hdc[2 * p];
        h[1][u] = (k - 1 < nz) ? Hzc[2 * (p + ns)] : 0.0;
      }
    }
  };
  // The factors g_k = c_{k-1} / bet_{k-1} depend on the mode only: instead of
  // storing them, each chunk saves the bet it starts from (one value per 16
  // levels; both lanes of the pair store the same value) and the back
  // substitution recomputes the chunk's g_k with the same operations.
  auto fwd_step = [&](const double (&h)[2][LES_TRI_U], int k0) {
    gam[(long)((k0 - 2) / LES_TRI_U) * ns + q] = bet;
#pragma unroll
    for (int u = 0; u < LES_TRI_U; ++u) {
      const int k = k0 + u;
      if (k <= nz + 1) {
        const double r = h[0][u] + (h[1][u] - hz) / dz;
        hz = h[1][u];
        const double cprev = (k == 2) ? c1 : idz2;
        const double g = cprev / bet;
        bet = bmid - idz2 * g;
        x = (r - idz2 * x) / bet;
        ph[2 * ((long)(k - 1) * ns + q)] = x;
      }
    }
  };
  fwd_load(hA, 2);
  for (int k0 = 2; k0 <= nz + 1; k0 += 2 * LES_TRI_U) {
    if (k0 + LES_TRI_U <= nz + 1) fwd_load(hB, k0 + LES_TRI_U);
    fwd_step(hA, k0);
    if (k0 + LES_TRI_U > nz + 1) break;
    if (k0 + 2 * LES_TRI_U <= nz + 1) fwd_load(hA, k0 + 2 * LES_TRI_U);
    fwd_step(hB, k0 + LES_TRI_U);
  }
  // row nz+2: top layer
  double gtop;
  {
    const double r = reinterpret_cast<const double*>(a.top)[2 * q + c] * dz;
    const double am = -1.0, bk = 1.0;
    gtop = idz2 / bet;
    bet = bk - am * gtop;
    x = (r - am * x) / bet;
  }
  // back substitution: x holds x(nz+2).  Row nz+1 uses g_{nz+2}; then chunk
  // c of the forward sweep (g_k, k = 2 + 16c .. 17 + 16c) serves rows k - 1,
  // its g_k recomputed from the saved bet; each lane reads back its own
  // stores, with the next chunk's loads issued first.
  {
    double* xp = &ph[2 * ((long)nz * ns + q)];
    x = *xp - gtop * x;
    *xp = x;
  }
  const int nchunk = (nz + LES_TRI_U - 1) / LES_TRI_U;
  double pA[LES_TRI_U], pB[LES_TRI_U], bA = 0.0, bB = 0.0;
  auto bwd_load = [&](double (&pv)[LES_TRI_U], double& bs, int ch) {
    const int k0 = 2 + LES_TRI_U * ch;
#pragma unroll
    for (int u = 0; u < LES_TRI_U; ++u) {
      const int kk = k0 - 1 + u;  // row
      if (kk <= nz) pv[u] = ph[2 * ((long)(kk - 1) * ns + q)];
    }
    bs = gam[(long)ch * ns + q];
  };
  auto bwd_step = [&](const double (&pv)[LES_TRI_U], double bs, int ch) {
    const int k0 = 2 + LES_TRI_U * ch;
    double G[LES_TRI_U];
    double bt = bs;
#pragma unroll
    for (int u = 0; u < LES_TRI_U; ++u) {
      const int k = k0 + u;
      if (k <= nz + 1) {
        const double cprev = (k == 2) ? c1 : idz2;
        G[u] = cprev / bt;
        bt = bmid - idz2 * G[u];
      }
    }
#pragma unroll
    for (int u = LES_TRI_U - 1; u >= 0; --u) {
      const int kk = k0 - 1 + u;  // row kk uses g_{kk+1}
      if (kk <= nz) {
        x = pv[u] - G[u] * x;
        ph[2 * ((long)(kk - 1) * ns + q)] = x;
      }
    }
  };
  bwd_load(pA, bA, nchunk - 1);
  for (int ch = nchunk - 1; ch >= 0; ch -= 2) {
    if (ch - 1 >= 0) bwd_load(pB, bB, ch - 1);
    bwd_step(pA, bA, ch);
    if (ch - 1 < 0) break;
    if (ch - 2 >= 0) bwd_load(pA, bA, ch - 2);
    bwd_step(pB, bB, ch - 1);
  }
  ph[2 * ((long)(nz + 1) * ns + q)] = 0.0;
}

// With ab2 set the right-hand side is formed from u + dt (tadv1 rhs + tadv2
// rhs_f) on load (u, v, w themselves are left as they are; the projection
// forms the same predictor again).
template <typename T>
static void press_stag_array(Les& L, double tadv1, bool ab2 = false, double tadv2 = 0.0, bool fused = false) {
  const int nz = L.nz;
  const double c = 1.0 / (double)(L.nx * L.ny);
  Cplx<double>* Hx = (Cplx<double>*)L.pspec[0];
  Cplx<double>* Hy = (Cplx<double>*)L.pspec[1];
  Cplx<double>* Hz = (Cplx<double>*)L.pspec[2];
  Cplx<double>* ph = (Cplx<double>*)L.pspec[3];
  // rH = const/tadv1 * (u/dt) on levels 1..nz (src/compute_pressure.f90:136-149)
  const double* src[3] = {D(L, LES_U), D(L, LES_V), D(L, LES_W)};
  Cplx<double>* dst[3] = {Hx, Hy, Hz};
  if (ab2) {
    const double* ab[6] = {D(L, LES_RHSX), D(L, LES_RHSY), D(L, LES_RHSZ),
                           D(L, LES_RHSX_F), D(L, LES_RHSY_F), D(L, LES_RHSZ_F)};
    const double abc[3] = {L.P.dt, tadv1, tadv2};
    rows_r2c<double, double>(L, src, dst, 3, 1, nz, c / tadv1, LOAD_AB2_SCALE_DIV, L.P.dt, 0, ab, abc);
  } else {
    rows_r2c<double, double>(L, src, dst, 3, 1, nz, c / tadv1, LOAD_SCALE_DIV, L.P.dt);
  }
  // forward y transforms; the horizontal divergence i kx Hx + i ky Hy is
  // formed by the column pass (into Hx), so the solve reads two spectra
  cols<double>(L, COP_DIV, Hx, Hy, Hx, nullptr, nullptr, 1, 1, 0, 1, nz);
  cols<double>(L, COP_COPY, Hz, nullptr, Hz, nullptr, nullptr, 1, 1, 0, 1, nz);
  // bottom and top layers (src/compute_pressure.f90:152-165, 196-203)
  Cplx<double>* lb = (Cplx<double>*)L.playr[0];
  Cplx<double>* lt = (Cplx<double>*)L.playr[1];
  const double* lsrc[2] = {L.layer[LES_BOTTOM_WK], L.layer[LES_TOP_WK]};
  Cplx<double>* ldst[2] = {lb, lt};
  rows_r2c<double, double>(L, lsrc, ldst, 2, 0, 1, c);
  cols<double>(L, COP_COPY, lb, nullptr, lb, nullptr, nullptr, 1, 1, 0, 0, 1);
  cols<double>(L, COP_COPY, lt, nullptr, lt, nullptr, nullptr, 1, 1, 0, 0, 1);
  // tridiagonal solve per mode
  TriArgs t;
  t.Hd = Hx; t.Hz = Hz; t.bot = lb; t.top = lt;
  t.kkx = L.kkx64; t.kky = L.kky64;
  t.phat = ph;
  t.gam = L.tri_gam;
  t.lhx = L.lhx; t.ny = L.ny; t.nz = nz; t.ns = L.ns;
  t.dz = L.dz; t.idz2 = 1.0 / (L.dz * L.dz);
  kt_begin(L, KT_TRIDIAG);
  const long tri_blocks = tri_zero_blocks(L.lhx, L.ny) + (2 * tri_nwork(L.lhx, L.ny) + 63) / 64;
  hipLaunchKernelGGL(k_tridiag, dim3((unsigned)tri_blocks), dim3(64), 0, L.stream, t);
  kt_end(L, KT_TRIDIAG, 16.0 * L.ns * (2.0 * nz + 2.0 + (nz + 2.0)), "k_tridiag");
  // gradient: i kx p, i ky p and p back to physical space (src/compute_pressure.f90:320-357)
  cols<double>(L, COP_PGRAD, ph, nullptr, Hx, Hy, Hz, 3, 0, 1, 0, nz + 2);
  if (fused) return;  // les_step: the projection is the inverse x transform's store (press_project)
  const Cplx<double>* cs[3] = {Hx, Hy, Hz};
  double* outs[3] = {D(L, LES_DPDX), D(L, LES_DPDY), D(L, LES_P)};
  rows_c2r<double, double>(L, cs, outs, 3, 0, nz + 2);
  L.gradp_lazy = 0;
}

// dpdz, projection, rhs -= grad p, w BCs and set_bc
// (src/compute_pressure.f90:353-369; src/LES_engine.f90:299-352)
struct ProjArgs {
  double *u, *v, *w, *p, *dpdx, *dpdy, *dpdz, *rx, *ry, *rz;
  const double *lbc_u, *lbc_v;
  double dtt, cz, dz;
  long np;
  int nz;
  int ubc_u_neu, ubc_v_neu;
  double ubc_u, ubc_v, ubc_w, bogus;
  int ab2;                      // u, v, w hold the old velocity: form the AB2 predictor first
  const double *fx, *fy, *fz;   // rhs_f
  double dt, t1, t2;
};

__global__ void k_proj(ProjArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nz) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const long c = (long)k * a.np + q;
  const double gz = a.cz * (a.p[c] - a.p[c - a.np]);
  a.dpdz[c] = gz;
  const double gx = a.dpdx[c], gy = a.dpdy[c];
  const double rxc = a.rx[c], ryc = a.ry[c], rzc = a.rz[c];
  double us = a.u[c], vs = a.v[c], ws = a.w[c];
  if (a.ab2) {
    us = us + a.dt * (a.t1 * rxc + a.t2 * a.fx[c]);
    vs = vs + a.dt * (a.t1 * ryc + a.t2 * a.fy[c]);
    ws = ws + a.dt * (a.t1 * rzc + a.t2 * a.fz[c]);
  }
  const double un = us - a.dtt * gx;
  const double vn = vs - a.dtt * gy;
  double wn = ws - a.dtt * gz;
  a.rx[c] = rxc - gx;
  a.ry[c] = ryc - gy;
  a.rz[c] = rzc - gz;
  if (k == 1) {
    wn = 0.0;
    a.u[q] = un - a.lbc_u[q] * a.dz;
    a.v[q] = vn - a.lbc_v[q] * a.dz;
    if (a.ab2) a.w[q] = 0.0;  // the AB2 pass sets w = 0 on the ghost levels
  }
  a.u[c] = un;
  a.v[c] = vn;
  a.w[c] = wn;
  if (k == a.nz) {
    const long tp = c + a.np;
    a.u[tp] = a.ubc_u_neu ? un + a.ubc_u * a.dz : 2.0 * a.ubc_u - un;
    a.v[tp] = a.ubc_v_neu ? vn + a.ubc_v * a.dz : 2.0 * a.ubc_v - vn;
    a.w[tp] = a.ubc_w;
    a.p[tp] = a.bogus;
  }
}

// ---- the step's projection, fused with the pressure-gradient transforms.
// The x and y components are projected as the store of the inverse x
// transform of i kx phat and i ky phat (blockIdx.y 0 and 1); blockIdx.y 2
// stores p.  The gradient then never goes through HBM: dpdx, dpdy and dpdz
// are formed again from phat and p only when the host asks for them
// (refresh_gradp).  The arithmetic per value is k_proj's.
struct ProjEpiArgs {
  double *vel, *r;          // u or v, rhsx or rhsy (null for the p component)
  const double *rf, *lbc;   // rhs_f, lbc_u or lbc_v
  double* p;                // p (component 2)
  int neu;                  // top BC Neumann
  double ubc;
};
struct ProjEpi3 {
  ProjEpiArgs c[3];
  double dt, t1, t2, dtt, dz;
  long np;
  int nz;
};

__device__ __forceinline__ void proj_epi_point(const ProjEpi3& e, const ProjEpiArgs& a, int k, long q, double g) {
  const long c = (long)k * e.np + q;
  if (a.p) {
    a.p[c] = g;
    return;
  }
  if (k < 1 || k > e.nz) return;
  const double rc = a.r[c];
  double us = a.vel[c];
  us = us + e.dt * (e.t1 * rc + e.t2 * a.rf[c]);
  const double un = us - e.dtt * g;
  a.r[c] = rc - g;
  a.vel[c] = un;
  if (k == 1) a.vel[q] = un - a.lbc[q] * e.dz;
  if (k == e.nz) a.vel[c + e.np] = a.neu ? un + a.ubc * e.dz : 2.0 * a.ubc - un;
}

__global__ void __launch_bounds__(LES_BLOCK) k_rows_c2r_proj(RowC2RArgs<double, double> a, ProjEpi3 ep, FftPlan pl,
                                                             const Cplx<double>* __restrict__ tw) {
  const ProjEpiArgs& e = ep.c[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  Cplx<double>* z = reinterpret_cast<Cplx<double>*>(smem_raw) + wave * a.ls;
  const long job = (long)blockIdx.x * (blockDim.x / LES_WAVE) + wave;
  const int npairs = a.ny / 2;
  if (job >= (long)a.nplanes * npairs) return;
  const int k = a.k0 + (int)(job / npairs), jp = (int)(job % npairs);
  if (pl.n == 256) {
    TwReg<double, 256> tr;
    tw_load(tr, tw + pl.twf, lane);
    rows_c2r_line(a, z, lane, k, jp, [&](Cplx<double>* p, bool inv) {
      wave_fft_t<double, 256, row_layout<double>(256)>(p, tr, inv, lane);
    });
  } else {
    rows_c2r_line(a, z, lane, k, jp, [&](Cplx<double>* p, bool inv) { wave_fft(p, pl, tw, inv, lane); });
  }
  const long q0 = (long)(2 * jp) * a.nx;
  for (int n = lane; n < a.nx; n += LES_WAVE) {
    const Cplx<double> v = z[n];
    proj_epi_point(ep, e, k, q0 + n, v.x);
    proj_epi_point(ep, e, k, q0 + a.nx + n, v.y);
  }
}

// w: dpdz from p, projection, rhsz -= dpdz, the w BCs and p(nz+1) = BOGUS
// (k_proj's arithmetic for the vertical component)
struct ProjWArgs {
  double *w, *rz, *p;
  const double* fz;
  double dt, t1, t2, dtt, cz, ubc_w, bogus;
  long np;
  int nz;
};

__global__ void k_proj_w(ProjWArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.np * a.nz) return;
  const int k = 1 + (int)(t / a.np);
  const long q = t % a.np;
  const long c = (long)k * a.np + q;
  const double gz = a.cz * (a.p[c] - a.p[c - a.np]);
  const double rzc = a.rz[c];
  double ws = a.w[c];
  ws = ws + a.dt * (a.t1 * rzc + a.t2 * a.fz[c]);
  double wn = ws - a.dtt * gz;
  a.rz[c] = rzc - gz;
  if (k == 1) {
    wn = 0.0;
    a.w[q] = 0.0;
  }
  a.w[c] = wn;
  if (k == a.nz) {
    a.w[c + a.np] = a.ubc_w;
    a.p[c + a.np] = a.bogus;
  }
}

// les_step's projection (AB2 predictor formed on load): after
// press_stag_array(..., ab2 = true), whose last pass leaves i kx phat,
// i ky phat and phat y-transformed in the pressure workspaces.
static void press_project(Les& L, double tadv1, double tadv2) {
  const int nz = L.nz;
  RowC2RArgs<double, double> ca;
  for (int f = 0; f < 3; ++f) { ca.src[f] = (const Cplx<double>*)L.pspec[f]; ca.dst[f] = nullptr; }
  ca.nx = L.nx; ca.ny = L.ny; ca.lhx = L.lhx; ca.k0 = 0; ca.nplanes = nz + 2;
  ca.ls = row_line<double>(L.nx);
  ProjEpi3 ep;
  ep.dt = L.P.dt; ep.t1 = tadv1; ep.t2 = tadv2; ep.dtt = L.P.dt * tadv1; ep.dz = L.dz;
  ep.np = L.np; ep.nz = nz;
  const int vel[2] = {LES_U, LES_V}, rr[2] = {LES_RHSX, LES_RHSY}, rf[2] = {LES_RHSX_F, LES_RHSY_F};
  const int lbc[2] = {LES_LBC_U, LES_LBC_V};
  for (int f = 0; f < 2; ++f) {
    ep.c[f].vel = D(L, vel[f]); ep.c[f].r = D(L, rr[f]); ep.c[f].rf = D(L, rf[f]);
    ep.c[f].lbc = L.layer[lbc[f]]; ep.c[f].p = nullptr;
    ep.c[f].neu = f == 0 ? L.P.ubc_u_neu : L.P.ubc_v_neu;
    ep.c[f].ubc = f == 0 ? L.P.ubc_u : L.P.ubc_v;
  }
  ep.c[2] = ProjEpiArgs{nullptr, nullptr, nullptr, nullptr, D(L, LES_P), 0, 0.0};
  const long jobs = (long)ca.nplanes * (L.ny / 2);
  const size_t sh = (size_t)kWavesPerBlock * ca.ls * sizeof(Cplx<double>);
  kt_begin(L, KT_ROWS_C2R);
  hipLaunchKernelGGL(k_rows_c2r_proj, dim3((unsigned)((jobs + kWavesPerBlock - 1) / kWavesPerBlock), 3),
                     dim3(LES_BLOCK), sh, L.stream, ca, ep, L.px, twpool<double>(L));
  // spectra in; p out; u, v, rhs, rhs_f in and u, v, rhs out on levels 1..nz
  kt_end(L, KT_ROWS_C2R, 3.0 * (nz + 2) * L.ny * ((double)L.lhx * 16) + 8.0 * L.np * (nz + 2) + 2.0 * 8 * L.np * nz * 5,
         "k_rows_c2r_proj");
  ProjWArgs w;
  w.w = D(L, LES_W); w.rz = D(L, LES_RHSZ); w.p = D(L, LES_P); w.fz = D(L, LES_RHSZ_F);
  w.dt = L.P.dt; w.t1 = tadv1; w.t2 = tadv2; w.dtt = L.P.dt * tadv1; w.cz = 1.0 / L.dz;
  w.ubc_w = L.P.ubc_w; w.bogus = BOGUS; w.np = L.np; w.nz = nz;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_proj_w, dim3(nblk(L.np * nz)), dim3(LES_BLOCK), 0, L.stream, w);
  // p (levels 0..nz), w, rhsz, rhsz_f in; w, rhsz out
  kt_end(L, KT_POINTWISE, 8.0 * L.np * nz * 6.0, "k_proj_w");
  L.gradp_lazy = 1;
}

// dpdx, dpdy (levels 0..nz+1) and dpdz (levels 1..nz) of the last step,
// from its phat (the pressure workspace keeps it until the next pressure
// solve) and p: the values the unfused projection stores, bit for bit.
static __global__ void k_dpdz(const double* p, double* dpdz, long np, int nz, double cz) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * nz) return;
  const long c = np + t;
  dpdz[c] = cz * (p[c] - p[c - np]);
}

static void refresh_gradp(Les& L) {
  if (!L.gradp_lazy) return;
  const int nz = L.nz;
  Cplx<double>* Hx = (Cplx<double>*)L.pspec[0];
  Cplx<double>* Hy = (Cplx<double>*)L.pspec[1];
  Cplx<double>* Hz = (Cplx<double>*)L.pspec[2];
  cols<double>(L, COP_PGRAD, (Cplx<double>*)L.pspec[3], nullptr, Hx, Hy, Hz, 3, 0, 1, 0, nz + 2);
  const Cplx<double>* cs[2] = {Hx, Hy};
  double* outs[2] = {D(L, LES_DPDX), D(L, LES_DPDY)};
  rows_c2r<double, double>(L, cs, outs, 2, 0, nz + 2);
  hipLaunchKernelGGL(k_dpdz, dim3(nblk(L.np * nz)), dim3(LES_BLOCK), 0, L.stream, (const double*)D(L, LES_P),
                     D(L, LES_DPDZ), L.np, nz, 1.0 / L.dz);
  L.gradp_lazy = 0;
}

static void projection(Les& L, double tadv1, bool ab2 = false, double tadv2 = 0.0) {
  refresh_gradp(L);
  ProjArgs a;
  a.ab2 = ab2 ? 1 : 0;
  a.fx = D(L, LES_RHSX_F); a.fy = D(L, LES_RHSY_F); a.fz = D(L, LES_RHSZ_F);
  a.dt = L.P.dt; a.t1 = tadv1; a.t2 = tadv2;
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W); a.p = D(L, LES_P);
  a.dpdx = D(L, LES_DPDX); a.dpdy = D(L, LES_DPDY); a.dpdz = D(L, LES_DPDZ);
  a.rx = D(L, LES_RHSX); a.ry = D(L, LES_RHSY); a.rz = D(L, LES_RHSZ);
  a.lbc_u = L.layer[LES_LBC_U]; a.lbc_v = L.layer[LES_LBC_V];
  a.dtt = L.P.dt * tadv1; a.cz = 1.0 / L.dz; a.dz = L.dz;
  a.np = L.np; a.nz = L.nz;
  a.ubc_u_neu = L.P.ubc_u_neu; a.ubc_v_neu = L.P.ubc_v_neu;
  a.ubc_u = L.P.ubc_u; a.ubc_v = L.P.ubc_v; a.ubc_w = L.P.ubc_w; a.bogus = BOGUS;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_proj, dim3(nblk(L.np * L.nz)), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, 8.0 * L.np * L.nz * (16.0 + (ab2 ? 3.0 : 0.0)), "k_proj");
}

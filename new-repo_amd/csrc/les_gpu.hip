// MI355X (gfx950) LES time step: device state, operator kernels, time-step
// driver and the C-ABI declared in include/les_gpu.h.
//
// One LES time step of the reference (src/LES_engine.f90:189-353) runs as a
// sequence of kernels on the library's HIP stream (les_step puts the SGS stage
// on a second stream beside the convection); the whole domain is resident in
// HBM.
// The 2DECOMP pencil transposes and MPI halo exchanges of the reference are
// replaced by strided column passes and by index arithmetic inside kernels.
#include <hip/hip_runtime.h>

#include <cassert>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "les_state.h"
#include "les_ops.h"

Les* g_les = nullptr;
static char g_err[512] = "";

void les_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

template <>
const Cplx<double>* twpool<double>(const Les& L) { return L.tw64; }
template <>
const Cplx<float>* twpool<float>(const Les& L) { return L.tw32; }

static const double BOGUS = -1.23456789e20;
static const double PI = 3.1415926535897932384626433;
static const double TWO_PI = 6.283185307179586476925286766559;
static const double KVONK = 0.4;

// ------------------------------------------------------------------ plans
struct HostPlan {
  FftPlan p;
  std::vector<double> re_f, im_f, re_i, im_i;
};

static bool build_plan(int n, HostPlan& hp, int pool_off, int* used) {
  int a = 0, b = 0, m = n;
  while (m % 2 == 0) { m /= 2; ++a; }
  while (m % 3 == 0) { m /= 3; ++b; }
  if (m != 1 || n < 2 || n > LES_MAX_N) return false;  // wave-level transforms hold n <= LES_MAX_N points
  FftPlan& p = hp.p;
  p.n = n;
  p.nst = 0;
  while (a >= 2) { p.R[p.nst++] = 4; a -= 2; }
  if (a == 1) p.R[p.nst++] = 2;
  for (int t = 0; t < b; ++t) p.R[p.nst++] = 3;
  if (p.nst > LES_MAX_STAGES) return false;
  int span = 1, o = 0;
  for (int s = 0; s < p.nst; ++s) {
    p.P[s] = span;
    p.off[s] = o;
    o += (p.R[s] - 1) * span;
    span *= p.R[s];
  }
  int sz = o > 0 ? o : 1;
  hp.re_f.assign(sz, 0.0); hp.im_f.assign(sz, 0.0); hp.re_i.assign(sz, 0.0); hp.im_i.assign(sz, 0.0);
  for (int s = 0; s < p.nst; ++s)
    for (int r = 1; r < p.R[s]; ++r)
      for (int k = 0; k < p.P[s]; ++k) {
        double ang = TWO_PI * (double)(r * k) / (double)(p.P[s] * p.R[s]);
        double c = std::cos(ang), sn = std::sin(ang);
        int q = p.off[s] + (r - 1) * p.P[s] + k;
        hp.re_f[q] = c; hp.im_f[q] = -sn;
        hp.re_i[q] = c; hp.im_i[q] = sn;
      }
  p.twf = pool_off;
  p.twi = pool_off + sz;
  *used = 2 * sz;
  return true;
}

// ------------------------------------------------------------------ memory
static void* dalloc(Les& L, size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 8;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    L.alloc_failed = 1;
    return nullptr;
  }
  hipMemset(p, 0, bytes);
  L.allocs.push_back(p);
  return p;
}

static bool field_is_fp64(int id) {
  switch (id) {
    case LES_U: case LES_V: case LES_W: case LES_P: case LES_DPDX: case LES_DPDY: case LES_DPDZ:
    case LES_RHSX: case LES_RHSY: case LES_RHSZ: case LES_RHSX_F: case LES_RHSY_F: case LES_RHSZ_F:
    case LES_F_LM: case LES_F_MM: case LES_F_QN: case LES_F_NN:
      return true;
  }
  return false;
}

// ------------------------------------------------------------------ timing
static void kt_begin(Les& L, int fam) {
  if (L.timing) hipEventRecord(L.ev_a[fam], L.stream);
}
// Ends the timed launch begun by kt_begin: adds its time, its algorithmic
// bytes (every logical input read once, every output written once, at the
// storage precision) and one launch to the kernel family `fam` and to the
// kernel named `base` (with the element sizes ta, tb of its template
// precisions, 0 for none, and an optional variant tag).
static void kt_end(Les& L, int fam, double bytes, const char* base, int ta = 0, int tb = 0,
                   const char* tag = nullptr) {
  if (L.timing) {
    L.kt_bytes[fam] += bytes;
    hipEventRecord(L.ev_b[fam], L.stream);
    hipEventSynchronize(L.ev_b[fam]);
    float ms = 0.f;
    hipEventElapsedTime(&ms, L.ev_a[fam], L.ev_b[fam]);
    L.kt_ms[fam] += ms;
    L.kt_calls[fam] += 1;
    int i = 0;
    for (; i < L.nkst; ++i) {
      const Les::KStat& k = L.kst[i];
      if (k.ta == ta && k.tb == tb && std::strcmp(k.base, base) == 0 &&
          ((!k.tag && !tag) || (k.tag && tag && std::strcmp(k.tag, tag) == 0)))
        break;
    }
    if (i == L.nkst) {
      if (L.nkst == LES_MAX_KSTAT) return;
      L.kst[i] = Les::KStat{base, ta, tb, tag, fam, 0.0, 0.0, 0};
      L.nkst += 1;
    }
    L.kst[i].ms += ms;
    L.kst[i].bytes += bytes;
    L.kst[i].calls += 1;
  }
}

// ------------------------------------------------------------------ launch helpers
// Row kernels: one wavefront per row pair, LES_BLOCK / 64 of them per block.
static const int kWavesPerBlock = LES_BLOCK / LES_WAVE;

template <typename TI, typename T>
static void rows_r2c(Les& L, const TI* const* src, Cplx<T>* const* dst, int nf, int k0, int npl, T scale, int mode = LOAD_SCALE,
                     T div = (T)1, int ny_override = 0, const TI* const* ab = nullptr, const T* abc = nullptr) {
  RowR2CArgs<TI, T> a;
  for (int f = 0; f < nf; ++f) { a.src[f] = src[f]; a.dst[f] = dst[f]; }
  a.nx = L.nx; a.ny = ny_override ? ny_override : L.ny; a.lhx = L.lhx; a.k0 = k0; a.nplanes = npl;
  a.mode = mode; a.scale = scale; a.div = div;
  a.ls = row_line<T>(L.nx);
  if (mode == LOAD_AB2_SCALE_DIV) {
    for (int f = 0; f < nf && f < 3; ++f) { a.ab_r[f] = ab[f]; a.ab_f[f] = ab[3 + f]; }
    a.ab_dt = abc[0]; a.ab_t1 = abc[1]; a.ab_t2 = abc[2];
  }
  if (mode == LOAD_PCLEAN) {  // one field; ab = u, v, w
    for (int f = 0; f < 3; ++f) a.ab_r[f] = ab[f];
    a.pc_nz = L.nz;
  }
  long jobs = (long)npl * (a.ny / 2);
  dim3 grid((unsigned)((jobs + kWavesPerBlock - 1) / kWavesPerBlock), nf);
  size_t sh = (size_t)kWavesPerBlock * a.ls * sizeof(Cplx<T>);
  kt_begin(L, KT_ROWS_R2C);
  if (mode == LOAD_PCLEAN)
    hipLaunchKernelGGL((k_rows_r2c<TI, T, true>), grid, dim3(LES_BLOCK), sh, L.stream, a, L.px, twpool<T>(L));
  else
    hipLaunchKernelGGL((k_rows_r2c<TI, T>), grid, dim3(LES_BLOCK), sh, L.stream, a, L.px, twpool<T>(L));
  const double in_fields = mode == LOAD_AB2_SCALE_DIV ? 3.0 : mode == LOAD_PCLEAN ? 4.0 : 1.0;
  kt_end(L, KT_ROWS_R2C, (double)nf * npl * a.ny * (in_fields * L.nx * sizeof(TI) + (double)L.lhx * sizeof(Cplx<T>)),
         "k_rows_r2c", sizeof(TI), sizeof(T),
         mode == LOAD_AB2_SCALE_DIV ? "AB2 predictor" : mode == LOAD_PCLEAN ? "cleaned pressure" : nullptr);
}

// The grad_velocity_filt forward transform of one velocity component in the
// mixed mode, with the FP64 Nyquist sums of field slot fld (0 u, 1 v, 2 w)
// formed on the way (k_rows_r2c_nyq).  Needs ny % 8 == 0.
static void rows_r2c_nyq(Les& L, const double* src, Cplx<float>* dst, int npl, float scale, int fld) {
  RowR2CArgs<double, float> a;
  a.src[0] = src; a.dst[0] = dst;
  a.nx = L.nx; a.ny = L.ny; a.lhx = L.lhx; a.k0 = 0; a.nplanes = npl;
  a.mode = LOAD_SCALE; a.scale = scale; a.div = 1.0f;
  a.ls = row_line<float>(L.nx);
  a.nyq_a = L.nyq_rows + (long)fld * L.ny * L.nzg;
  a.nyq_p = L.nyq_part + (long)fld * (L.ny / 4) * L.nx * L.nzg;
  const long jobs = (long)npl * (L.ny / 2);
  assert(L.ny % 8 == 0 && jobs % kWavesPerBlock == 0);
  const size_t sh = (size_t)kWavesPerBlock * a.ls * sizeof(Cplx<float>) + (size_t)kWavesPerBlock * L.nx * sizeof(double);
  kt_begin(L, KT_ROWS_R2C);
  hipLaunchKernelGGL(k_rows_r2c_nyq<float>, dim3((unsigned)(jobs / kWavesPerBlock), 1), dim3(LES_BLOCK), sh, L.stream, a, L.px,
                     twpool<float>(L));
  // the field read, its half spectrum written; A (1/nx) and P (1/(8 nx))
  // of a field not counted, as in k_nyq_reduce
  kt_end(L, KT_ROWS_R2C, (double)npl * L.ny * (L.nx * 8.0 + (double)L.lhx * sizeof(Cplx<float>)), "k_rows_r2c_nyq", 4);
}

template <typename T, typename TO>
static void rows_c2r(Les& L, const Cplx<T>* const* src, TO* const* dst, int nf, int k0, int npl) {
  RowC2RArgs<T, TO> a;
  for (int f = 0; f < nf; ++f) { a.src[f] = src[f]; a.dst[f] = dst[f]; }
  a.nx = L.nx; a.ny = L.ny; a.lhx = L.lhx; a.k0 = k0; a.nplanes = npl;
  a.ls = row_line<T>(L.nx);
  long jobs = (long)npl * (L.ny / 2);
  dim3 grid((unsigned)((jobs + kWavesPerBlock - 1) / kWavesPerBlock), nf);
  size_t sh = (size_t)kWavesPerBlock * a.ls * sizeof(Cplx<T>);
  kt_begin(L, KT_ROWS_C2R);
  hipLaunchKernelGGL((k_rows_c2r<T, TO>), grid, dim3(LES_BLOCK), sh, L.stream, a, L.px, twpool<T>(L));
  kt_end(L, KT_ROWS_C2R, (double)nf * npl * L.ny * ((double)L.lhx * sizeof(Cplx<T>) + (double)L.nx * sizeof(TO)),
         "k_rows_c2r", sizeof(T), sizeof(TO));
}

// Column tiles: LES_COL_TILE_BYTES contiguous bytes per row for
// single-input ops, half that for two-input ops (three LDS tiles instead of
// two), at least 4 columns.  With the XCD-aware tile order (xcd_tile) the
// narrow rows are not over-fetched; 32-byte rows (4 FP32 / 4 FP64 columns)
// measured fastest at 256^2 because the small tiles leave room for more
// blocks per CU.
#ifndef LES_COL_TILE_BYTES
#define LES_COL_TILE_BYTES 32
#endif

// LDS line stride (in line elements of eb bytes, at least n) for a tile of
// `lines` lines filled and drained row by row (consecutive lanes take
// consecutive lines, then the next row).  One LDS cycle serves 128 bytes:
// g = 128 / eb lanes spread over r = g / lines rows.  A stride of r (mod g)
// puts line c, row j at bank slot c r + j, all distinct.
static int lds_stride(int n, int lines, int eb) {
  const int g = 128 / eb;
  const int r = lines >= g ? 1 : g / lines;
  int st = n;
  while (st % g != r % g) ++st;
  return st;
}
template <typename T>
static int col_tile(int nin) {
  int cb = (nin > 1 ? LES_COL_TILE_BYTES / 2 : LES_COL_TILE_BYTES) / (int)sizeof(Cplx<T>);
  return cb < 4 ? 4 : cb;
}

// ny = 256 runs the persistent pipelined kernel body: a grid of (resident
// blocks per CU) x (CUs), a multiple of the 8 XCDs; other lengths one block
// per tile.
// ny = 256: register-resident transform (k_cols_c16); LES_COL16=0 selects
// the LDS-engine kernel (k_cols_256) instead.
#ifndef LES_COL16
#define LES_COL16 1
#endif
template <typename T, int CB>
static void launch_cols(Les& L, const ColArgs<T>& a, int npl, int njobs) {
  const size_t sh = (size_t)(a.nin + 1) * CB * a.ls * sizeof(Cplx<T>);
  if (!(L.ny == 256 && LES_COL16) && njobs > 1) {
    // the LDS-engine kernels take one job per launch
    for (int j = 0; j < njobs; ++j) {
      ColArgs<T> b = a;
      for (int t = 0; t < 2; ++t) b.in[0][t] = a.in[j][t];
      for (int o = 0; o < 3; ++o) b.out[0][o] = a.out[j][o];
      launch_cols<T, CB>(L, b, npl, 1);
    }
    return;
  }
  if (L.ny == 256 && LES_COL16) {
    ColArgs<T> b = a;
    b.ncb = (L.lhx + 15) / 16;
    b.ntiles = b.ncb * npl;
    const size_t sh16 = (size_t)16 * LES_C16_LS * sizeof(Cplx<T>);
    const dim3 g((unsigned)b.ntiles, (unsigned)njobs);
    const Cplx<T>* tw = twpool<T>(L) + L.py.twf;
#define LES_C16(NIN, OP, NOUT) \
  hipLaunchKernelGGL((k_cols_c16<T, NIN, OP, NOUT>), g, dim3(LES_BLOCK), sh16, L.stream, b, tw)
    if (a.op == COP_GRAD && a.nout == 3) LES_C16(1, COP_GRAD, 3);
    else if (a.op == COP_GRAD) LES_C16(1, COP_GRAD, 2);
    else if (a.op == COP_DIV) LES_C16(2, COP_DIV, 1);
    else if (a.op == COP_FILT2) LES_C16(1, COP_FILT2, 2);
    else if (a.op == COP_FILT1) LES_C16(1, COP_FILT1, 1);
    else if (a.op == COP_PGRAD) LES_C16(1, COP_PGRAD, 3);
    else LES_C16(1, COP_COPY, 1);
#undef LES_C16
  } else if (L.ny == 256) {
    static int occ[2] = {0, 0};  // per input count
    int& o = occ[a.nin > 1];
    if (o == 0) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_cols_256<T, CB>, LES_BLOCK, sh) != hipSuccess || o < 1) o = 1;
    }
    long g = (long)o * L.ncu;
    long need = ((long)a.ntiles + LES_NXCD - 1) / LES_NXCD * LES_NXCD;
    if (g > need) g = need;
    const unsigned grid = (unsigned)((g + LES_NXCD - 1) / LES_NXCD * LES_NXCD);
    hipLaunchKernelGGL((k_cols_256<T, CB>), dim3(grid), dim3(LES_BLOCK), sh, L.stream, a, L.py, twpool<T>(L));
  } else {
    hipLaunchKernelGGL((k_cols<T, CB>), dim3((unsigned)a.ntiles), dim3(LES_BLOCK), sh, L.stream, a, L.py, twpool<T>(L));
  }
}

// The same column op on njobs (at most LES_COL_JOBS) independent sets of
// inputs and outputs, one launch on the register-resident path: the sets'
// tiles share the grid, so the partly filled last round of blocks is paid
// once instead of once per set.
template <typename T>
static void cols_jobs(Les& L, int op, int njobs, const Cplx<T>* const* in0, const Cplx<T>* const* in1,
                      Cplx<T>* const* o0, Cplx<T>* const* o1, Cplx<T>* const* o2, int nout, int fwd, int inv, int k0,
                      int npl, const T* g0 = nullptr, const T* g1 = nullptr) {
  assert(njobs >= 1 && njobs <= LES_COL_JOBS);
  ColArgs<T> a;
  for (int j = 0; j < LES_COL_JOBS; ++j) {
    const int s = j < njobs ? j : 0;
    a.in[j][0] = in0[s]; a.in[j][1] = in1 ? in1[s] : nullptr;
    a.out[j][0] = o0[s]; a.out[j][1] = o1 ? o1[s] : nullptr; a.out[j][2] = o2 ? o2[s] : nullptr;
  }
  a.kkx = sizeof(T) == 8 ? (const T*)L.kkx64 : (const T*)L.kkx32;
  a.kky = sizeof(T) == 8 ? (const T*)L.kky64 : (const T*)L.kky32;
  a.g0 = g0; a.g1 = g1;
  a.lhx = L.lhx; a.ny = L.ny; a.k0 = k0; a.nplanes = npl;
  a.nin = in1 ? 2 : 1;
  a.nout = nout; a.op = op; a.fwd = fwd; a.inv = inv;
  const int cb = col_tile<T>(a.nin);
  a.ls = lds_stride(L.ny, cb, (int)sizeof(Cplx<T>));
  a.ncb = (L.lhx + cb - 1) / cb;
  a.ntiles = a.ncb * npl;
  kt_begin(L, KT_COLS);
  if (cb == 4) launch_cols<T, 4>(L, a, npl, njobs);
  else if (cb == 8) launch_cols<T, 8>(L, a, npl, njobs);
  else launch_cols<T, 16>(L, a, npl, njobs);
  static const char* const opn[] = {"GRAD", "DIV", "FILT2", "FILT1", "COPY", "PGRAD"};
  kt_end(L, KT_COLS, (double)njobs * npl * L.ns * sizeof(Cplx<T>) * (a.nin + nout),
         (L.ny == 256 && LES_COL16) ? "k_cols_c16" : L.ny == 256 ? "k_cols_256" : "k_cols", sizeof(T), 0, opn[op]);
}

template <typename T>
static void cols(Les& L, int op, const Cplx<T>* in0, const Cplx<T>* in1, Cplx<T>* o0, Cplx<T>* o1, Cplx<T>* o2,
                 int nout, int fwd, int inv, int k0, int npl, const T* g0 = nullptr, const T* g1 = nullptr) {
  const Cplx<T>* i0[1] = {in0};
  const Cplx<T>* i1[1] = {in1};
  Cplx<T>* p0[1] = {o0};
  Cplx<T>* p1[1] = {o1};
  Cplx<T>* p2[1] = {o2};
  cols_jobs<T>(L, op, 1, i0, in1 ? i1 : nullptr, p0, p1, p2, nout, fwd, inv, k0, npl, g0, g1);
}

// ------------------------------------------------------------------ field access
template <typename T>
static T* F(Les& L, int id) { return (T*)L.f[id]; }
static double* D(Les& L, int id) { return (double*)L.f[id]; }

static unsigned nblk(long n, int b = LES_BLOCK) { return (unsigned)((n + b - 1) / b); }

// Mixed mode with ny % 8 == 0: grad_filt's forward transforms of u, v, w
// also form the sums of the FP64 Nyquist removal (k_rows_r2c_nyq), which
// then starts at its second pass.  Every caller of the removal runs
// grad_filt on u, v and w first, on the same stream.
#ifndef LES_NYQ_SUMS
#define LES_NYQ_SUMS 1
#endif
static bool nyq_sums_fused(const Les& L) { return LES_NYQ_SUMS && L.mixed && L.ny % 8 == 0; }

// ------------------------------------------------------------------ operators
// grad_velocity_filt for one velocity component (FP64 path: the filtered
// field is written back as in src/compute_derivatives.f90:134).
template <typename T>
static void grad_filt(Les& L, int fid, int dx_id, int dy_id) {
  Cplx<T>* s0 = (Cplx<T>*)L.wk_spec[0];
  Cplx<T>* s1 = (Cplx<T>*)L.wk_spec[1];
  Cplx<T>* s2 = (Cplx<T>*)L.wk_spec[2];
  Cplx<T>* s3 = (Cplx<T>*)L.wk_spec[3];
  const double* src[1] = {D(L, fid)};
  Cplx<T>* dst[1] = {s0};
  T c = (T)(1.0 / (double)(L.nx * L.ny));
  if constexpr (std::is_same<T, float>::value) {
    if (nyq_sums_fused(L)) rows_r2c_nyq(L, src[0], s0, L.nzg, c, fid - LES_U);
    else rows_r2c<double, T>(L, src, dst, 1, 0, L.nzg, c);
  } else {
    rows_r2c<double, T>(L, src, dst, 1, 0, L.nzg, c);
  }
  cols<T>(L, COP_GRAD, s0, nullptr, s1, s2, s3, L.mixed ? 2 : 3, 1, 1, 0, L.nzg);
  const Cplx<T>* cs[2] = {s1, s2};
  T* outs[2] = {F<T>(L, dx_id), F<T>(L, dy_id)};
  rows_c2r<T, T>(L, cs, outs, 2, 0, L.nzg);
  if (!L.mixed) {
    const Cplx<T>* c3[1] = {s3};
    double* o3[1] = {D(L, fid)};
    rows_c2r<T, double>(L, c3, o3, 1, 0, L.nzg);
  }
}

// grad_velocity_filt for u, v and w together, with the arithmetic of
// grad_filt per field: the forward transforms (one launch of three fields,
// or one per field where the mixed mode forms the Nyquist sums), one column
// launch of three jobs, one inverse launch of the six derivatives and, in
// FP64, one of the three filtered fields.  The FP64 filtered spectra
// overwrite the inputs: k_cols_c16 reads its whole tile before its first
// store and every block owns its tile.
// Measured at 256^3 (MI355X): FP64 0.055 ms per step faster than three
// grad_filt chains; in FP32 0.025 ms slower, its six-field inverse launch
// reading spectra that the per-field chain still finds in the 256 MB
// last-level cache.  So the mixed modes keep the per-field chains.
template <typename T>
static void grad_filt3(Les& L) {
  const int fid[3] = {LES_U, LES_V, LES_W};
  const int dx_id[3] = {LES_DUDX, LES_DVDX, LES_DWDX}, dy_id[3] = {LES_DUDY, LES_DVDY, LES_DWDY};
  static_assert(LES_NSPEC >= 9, "grad_filt3 needs 9 spectral workspaces");
  Cplx<T>* in[3];
  Cplx<T>* ox[3];
  Cplx<T>* oy[3];
  const double* src[3];
  const Cplx<T>* cs[6];
  T* outs[6];
  for (int f = 0; f < 3; ++f) {
    in[f] = (Cplx<T>*)L.wk_spec[f];
    ox[f] = (Cplx<T>*)L.wk_spec[3 + f];
    oy[f] = (Cplx<T>*)L.wk_spec[6 + f];
    src[f] = D(L, fid[f]);
    cs[2 * f] = ox[f]; cs[2 * f + 1] = oy[f];
    outs[2 * f] = F<T>(L, dx_id[f]); outs[2 * f + 1] = F<T>(L, dy_id[f]);
  }
  T c = (T)(1.0 / (double)(L.nx * L.ny));
  if constexpr (std::is_same<T, float>::value) {
    if (nyq_sums_fused(L)) {
      for (int f = 0; f < 3; ++f) rows_r2c_nyq(L, src[f], in[f], L.nzg, c, f);
    } else {
      rows_r2c<double, T>(L, src, in, 3, 0, L.nzg, c);
    }
  } else {
    rows_r2c<double, T>(L, src, in, 3, 0, L.nzg, c);
  }
  cols_jobs<T>(L, COP_GRAD, 3, in, nullptr, ox, oy, L.mixed ? nullptr : in, L.mixed ? 2 : 3, 1, 1, 0, L.nzg);
  rows_c2r<T, T>(L, cs, outs, 6, 0, L.nzg);
  if (!L.mixed) {
    const Cplx<T>* c3[3] = {in[0], in[1], in[2]};
    double* o3[3] = {D(L, LES_U), D(L, LES_V), D(L, LES_W)};
    rows_c2r<T, double>(L, c3, o3, 3, 0, L.nzg);
  }
}

static void nyquist_remove_fp64(Les& L, bool with_ddz = false);
// mixed modes: the Nyquist removal's last pass also applies the boundary
// conditions and forms the vertical derivatives (k_nyq_apply_ddz)
#ifndef LES_NYQ_DDZ
#define LES_NYQ_DDZ 1
#endif
// column segments per block of k_nyq_apply_ddz (64 columns each)
#ifndef LES_NYQ_SEGS
#define LES_NYQ_SEGS 8
#endif

template <typename T>
static void vel_grad_bc_ddz(Les& L);

// the horizontal gradients of u, v, w (first part of compute_vel_grad)
template <typename T>
static void vel_grads(Les& L) {
  if (L.mixed) {
    grad_filt<T>(L, LES_U, LES_DUDX, LES_DUDY);
    grad_filt<T>(L, LES_V, LES_DVDX, LES_DVDY);
    grad_filt<T>(L, LES_W, LES_DWDX, LES_DWDY);
  } else {
    grad_filt3<T>(L);
  }
}

template <typename T>
static void compute_vel_grad(Les& L) {
  vel_grads<T>(L);
  vel_grad_bc_ddz<T>(L);
}

// Second part of compute_vel_grad: the FP64 Nyquist removal of the mixed
// mode, the boundary conditions and the vertical derivatives.
template <typename T>
static void vel_grad_bc_ddz(Les& L) {
  if (L.mixed && LES_NYQ_DDZ) {
    nyquist_remove_fp64(L, true);
    return;
  }
  if (L.mixed) nyquist_remove_fp64(L);
  BcDdzArgs<T> a;
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W);
  a.dudz = F<T>(L, LES_DUDZ); a.dvdz = F<T>(L, LES_DVDZ); a.dwdz = F<T>(L, LES_DWDZ);
  a.lbc_u = L.layer[LES_LBC_U]; a.lbc_v = L.layer[LES_LBC_V];
  a.np = L.np; a.nz = L.nz; a.dz = L.dz; a.cz = 1.0 / L.dz;
  a.ubc_u_neu = L.P.ubc_u_neu; a.ubc_v_neu = L.P.ubc_v_neu;
  a.ubc_u = L.P.ubc_u; a.ubc_v = L.P.ubc_v; a.ubc_w = L.P.ubc_w;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_bc_ddz<T>), dim3(nblk(L.np * (L.nz + 1))), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, (double)L.np * (L.nz + 1) * (3 * 8.0 + 3.0 * sizeof(T)), "k_bc_ddz", sizeof(T));
}

// 2-delta filtered first level (get_layer_filter, src/utility_tools.f90:261-307)
// then the wall law (src/compute_wall_law.f90:107-254).
template <typename T>
static void wall_law(Les& L) {
  T* u1 = (T*)L.wk_layer[0];
  T* v1 = (T*)L.wk_layer[1];
  if (L.P.wall_model == 0) {
    Cplx<T>* sa = (Cplx<T>*)L.wk_layer[2];
    Cplx<T>* sb = (Cplx<T>*)L.wk_layer[3];
    Cplx<T>* sc = (Cplx<T>*)L.wk_layer[4];
    Cplx<T>* sd = (Cplx<T>*)L.wk_layer[5];
    const double* src[2] = {D(L, LES_U) + L.np, D(L, LES_V) + L.np};
    Cplx<T>* dst[2] = {sa, sb};
    rows_r2c<double, T>(L, src, dst, 2, 0, 1, (T)1);
    const T* g2 = sizeof(T) == 8 ? (const T*)L.g2_64 : (const T*)L.g2_32;
    cols<T>(L, COP_FILT1, sa, nullptr, sc, nullptr, nullptr, 1, 1, 1, 0, 1, g2);
    cols<T>(L, COP_FILT1, sb, nullptr, sd, nullptr, nullptr, 1, 1, 1, 0, 1, g2);
    const Cplx<T>* cs[2] = {sc, sd};
    T* outs[2] = {u1, v1};
    rows_c2r<T, T>(L, cs, outs, 2, 0, 1);
  }
  WallArgs<T> a;
  a.u1 = u1; a.v1 = v1;
  a.uraw = D(L, LES_U) + L.np; a.vraw = D(L, LES_V) + L.np;
  a.inst = L.P.wall_model == 1;
  a.logterm = L.logterm;
  a.ustar = L.layer[LES_USTAR]; a.lbc_u = L.layer[LES_LBC_U]; a.lbc_v = L.layer[LES_LBC_V];
  a.dudz1 = F<T>(L, LES_DUDZ) + L.np; a.dvdz1 = F<T>(L, LES_DVDZ) + L.np;
  a.txz1 = F<T>(L, LES_TXZ) + L.np; a.tyz1 = F<T>(L, LES_TYZ) + L.np;
  a.c1 = 0.5 * L.dz * KVONK;
  a.np = L.np;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_wall<T>), dim3(nblk(L.np)), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, (double)L.np * (2.0 * sizeof(T) + 8 * 4.0 + 4.0 * sizeof(T)), "k_wall", sizeof(T));
}

template <typename T>
static SgsArgs<T> sgs_args(Les& L) {
  SgsArgs<T> a;
  a.dudx = F<T>(L, LES_DUDX); a.dudy = F<T>(L, LES_DUDY); a.dudz = F<T>(L, LES_DUDZ);
  a.dvdx = F<T>(L, LES_DVDX); a.dvdy = F<T>(L, LES_DVDY); a.dvdz = F<T>(L, LES_DVDZ);
  a.dwdx = F<T>(L, LES_DWDX); a.dwdy = F<T>(L, LES_DWDY); a.dwdz = F<T>(L, LES_DWDZ);
  a.cs = F<T>(L, LES_CS_OPT2); a.nu_t = F<T>(L, LES_NU_T);
  a.txx = F<T>(L, LES_TXX); a.txy = F<T>(L, LES_TXY); a.txz = F<T>(L, LES_TXZ);
  a.tyy = F<T>(L, LES_TYY); a.tyz = F<T>(L, LES_TYZ); a.tzz = F<T>(L, LES_TZZ);
  a.s11 = F<T>(L, LES_S11); a.s12 = F<T>(L, LES_S12); a.s13 = F<T>(L, LES_S13);
  a.s22 = F<T>(L, LES_S22); a.s23 = F<T>(L, LES_S23); a.s33 = F<T>(L, LES_S33);
  a.smod = (T*)L.wk_real[63];
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W);
  a.ulag = F<T>(L, LES_U_LAG); a.vlag = F<T>(L, LES_V_LAG); a.wlag = F<T>(L, LES_W_LAG);
  a.lvec = L.lvec;
  a.np = L.np; a.nz = L.nz;
  a.accumulate = 0; a.write_sij = 0;
  a.acc_kmax = L.nz + 1;
  return a;
}

static int lag_levels(const Les& L, int jtt);
template <typename T>
static void lasd_update(Les& L, int jtt, int nlev);
template <typename T>
static void learned_update(Les& L, int jtt);
template <typename T>
static void dynsmag_update(Les& L);

// sgs_stress (src/compute_sgs.f90:105-247).  Model 3 dns: the molecular
// stress only (k_dns_stress).  Model 0 cst_Smag, 1 dyn_Smag
// (updated when the local step count jt is a multiple of cs_count,
// src/compute_DynSmag.f90:70-80), 2 dyn_LagScaleDep (on the global count jtt,
// src/compute_LagScaleDep.f90:237); both dynamic models hold Cs^2 = 0.03
// while jtt < dyn_init.
template <typename T>
static void sgs_stress(Les& L, int jt, int jtt) {
  SgsArgs<T> a = sgs_args<T>(L);
  if (L.P.sgs_model == 3) {
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_dns_stress<T>), dim3(nblk(L.np * L.nz)), dim3(LES_BLOCK), 0, L.stream, a, (T)(2.0 / L.P.re));
    kt_end(L, KT_POINTWISE, (double)L.np * L.nz * 15.0 * sizeof(T), "k_dns_stress", sizeof(T));
    return;
  }
  bool lasd = L.P.sgs_model == 2;
  bool dyns = L.P.sgs_model == 1;
  bool acc = lasd && jtt >= L.P.dyn_init;
  bool update = lasd && jtt >= L.P.dyn_init && (jtt % L.P.cs_count) == 0;
  bool dyn_update = dyns && jtt >= L.P.dyn_init && (jt % L.P.cs_count) == 0;
  if (acc) a.acc_kmax = lag_levels(L, jtt);
  // levels whose u_lag sums a launch updates (0..acc_kmax, at most nz+1)
  auto lag_bytes = [&]() {
    if (!a.accumulate || a.acc_kmax < 1) return 0.0;
    return (double)L.np * ((a.acc_kmax < L.nz + 1 ? a.acc_kmax : L.nz + 1) + 1.0) * (3 * 8.0 + 6.0 * sizeof(T));
  };
  if (L.P.sgs_model == 0) {
    CsFillArgs<T> c;
    c.cs = F<T>(L, LES_CS_OPT2); c.val = (T)(L.P.cs0 * L.P.cs0); c.np = L.np; c.nz = L.nz;
    hipLaunchKernelGGL((k_cs_fill<T>), dim3(nblk(L.np * (L.nz + 1))), dim3(LES_BLOCK), 0, L.stream, c);
  } else if ((lasd || dyns) && jtt < L.P.dyn_init) {
    CsFillArgs<T> c;
    c.cs = F<T>(L, LES_CS_OPT2); c.val = (T)0.03; c.np = L.np; c.nz = L.nz;
    hipLaunchKernelGGL((k_cs_fill<T>), dim3(nblk(L.np * (L.nz + 1))), dim3(LES_BLOCK), 0, L.stream, c);
  }
  if (update) {
    a.accumulate = acc ? 1 : 0;
    a.write_sij = 1;
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_sij<T>), dim3(nblk(L.np * (L.nz + 1))), dim3(LES_BLOCK), 0, L.stream, a);
    kt_end(L, KT_POINTWISE, (double)L.np * (L.nz + 1) * 16.0 * sizeof(T) + lag_bytes(), "k_sij", sizeof(T), 0,
           a.accumulate ? "with u_lag sums" : nullptr);
    if constexpr (sizeof(T) == 4) {
      if (L.P.precision == 2) learned_update<T>(L, jtt);
      else lasd_update<T>(L, jtt, L.nz + 1);
    } else {
      lasd_update<T>(L, jtt, L.nz + 1);
    }
    a.accumulate = 0;
  } else if (dyn_update) {
    a.write_sij = 1;
    kt_begin(L, KT_POINTWISE);
    hipLaunchKernelGGL((k_sij<T>), dim3(nblk(L.np * (L.nz + 1))), dim3(LES_BLOCK), 0, L.stream, a);
    kt_end(L, KT_POINTWISE, (double)L.np * (L.nz + 1) * 16.0 * sizeof(T), "k_sij", sizeof(T));
    dynsmag_update<T>(L);
  } else {
    a.accumulate = acc ? 1 : 0;
  }
  a.write_sij = 0;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL((k_sgs<T>), dim3(nblk(L.np * (L.nz + 1))), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, (double)L.np * (L.nz + 1) * 17.0 * sizeof(T) + lag_bytes(), "k_sgs", sizeof(T), 0,
         a.accumulate ? "with u_lag sums" : nullptr);
}

#include "les_lasd.h"
#include "les_rhs.h"
#include "les_press.h"
#include "les_mlp.h"
#include "les_learn.h"
#include "les_stats.h"
#include "les_tke.h"

// ---------------------------------------------------------------- allocation-free profiles
// Device array behind les_stat_profiles_noalloc: the per-slice partial sums
// of k_stat_planes ([nz][LES_PROF_SLICES][PS_N]) followed by the profiles
// (LES_NPROF x nz), for nz <= LES_NOALLOC_MAX_NZ: 1024 x (16 x 23 + 21) x 8 B
// = 3.19 MB, held by every GPU the library runs on until the process ends.
// Host code never passes the name `scratch` to a kernel or copy (in host code
// it is a stand-in, not a device address): les_init looks up each device's
// address once with hipGetSymbolAddress and keeps it in g_scratch[device].
#define LES_NOALLOC_MAX_NZ 1024
#define LES_NOALLOC_PART ((size_t)LES_NOALLOC_MAX_NZ * LES_PROF_SLICES * PS_N)
#define LES_NOALLOC_DOUBLES (LES_NOALLOC_PART + (size_t)LES_NOALLOC_MAX_NZ * LES_NPROF)
static __device__ __align__(256) double scratch[LES_NOALLOC_DOUBLES];
#define LES_MAX_DEVICES 64
static double* g_scratch[LES_MAX_DEVICES] = {};

// Address of this device's `scratch` (the device must be current); 0 on success.
static int scratch_lookup(int dev) {
  if (dev < 0 || dev >= LES_MAX_DEVICES) { les_set_error("les_init: device id %d out of range", dev); return 10; }
  if (g_scratch[dev]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    les_set_error("les_init: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    return 5;
  }
  g_scratch[dev] = (double*)ptr;
  return 0;
}

// ---------------------------------------------------------------- mixed-mode Nyquist filter
// In the mixed-precision mode the velocity stays FP64, so the Nyquist-free
// velocity of grad_velocity_filt (src/compute_derivatives.f90:105-134) is
// formed directly in FP64 by removing the projection on the kx = nx/2 and
// ky = ny/2 modes: f -= (-1)^i A_j + (-1)^j B_i - (-1)^(i+j) C, with A, B, C
// the normalised alternating sums along x, y and both.
// Pass 1: one block per band of 4 RPW rows of a plane, RPW consecutive rows
// per wavefront (RPW = 4 when ny is a multiple of 16, else 1): A_j =
// sum_i (-1)^i f_ij / nx by wavefront reduction, and the band's alternating
// column partial sums P_i = sum_{j in band} (-1)^j f_ij (each wave's rows in
// order, then the four waves).  Every sum has a fixed order, so results are
// reproducible.  blockIdx.y selects the field (u, v, w in one launch).
struct NyqArgs {
  double* f[3];
  double *A, *B, *C, *P;  // per field: A ny*nzg, B nx*nzg, C nzg, P (ny/4)*nx*nzg
  int nx, ny, nzg;
};

// LDS: the four waves' column partials, 4 nx doubles (dynamic, sized at the
// launch so the block holds only what nx needs).
template <int RPW>
__global__ void __launch_bounds__(LES_BLOCK) k_nyq_reduce(NyqArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = a.nx, ny = a.ny, fld = blockIdx.y;
  double* part0 = reinterpret_cast<double*>(smem_raw);
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  const long row0 = ((long)blockIdx.x * (LES_BLOCK / LES_WAVE) + wave) * RPW;  // global row k * ny + j
  const double* __restrict__ fr = a.f[fld] + row0 * nx;
  double s[RPW], sj[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    s[r] = 0.0;
    sj[r] = ((row0 + r) % ny) & 1 ? -1.0 : 1.0;
  }
  for (int i = lane; i < nx; i += LES_WAVE) {
    double v[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) v[r] = fr[(long)r * nx + i];
    double acc = sj[0] * v[0];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      s[r] += (i & 1) ? -v[r] : v[r];
      if (r > 0) acc += sj[r] * v[r];
    }
    part0[wave * nx + i] = acc;
  }
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    for (int o = LES_WAVE / 2; o > 0; o >>= 1) s[r] += __shfl_xor(s[r], o, LES_WAVE);
    if (lane == 0) a.A[(long)fld * ny * a.nzg + row0 + r] = s[r] / (double)nx;
  }
  __syncthreads();
  double* __restrict__ P = a.P + (long)fld * (ny / 4) * nx * a.nzg;
  for (int i = threadIdx.x; i < nx; i += blockDim.x)
    P[(long)blockIdx.x * nx + i] = ((part0[i] + part0[nx + i]) + part0[2 * nx + i]) + part0[3 * nx + i];
}

// Pass 2, one block per plane and field: B_i = sum of the nb band partials /
// ny per column; C = sum_j (-1)^j A_j / ny, each thread summing its rows j in
// order, then a fixed tree.
__global__ void __launch_bounds__(LES_BLOCK) k_nyq_bc(NyqArgs a, int nb) {
  const int nx = a.nx, ny = a.ny, fld = blockIdx.y;
  const long k = blockIdx.x;
  const double* __restrict__ P = a.P + (long)fld * (ny / 4) * nx * a.nzg;
  const double* __restrict__ A = a.A + (long)fld * ny * a.nzg + k * ny;
  double* __restrict__ B = a.B + (long)fld * nx * a.nzg + k * nx;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += P[(k * nb + b) * nx + i];
    B[i] = s / (double)ny;
  }
  double c = 0.0;
  for (int jj = threadIdx.x; jj < ny; jj += blockDim.x) c += (jj & 1) ? -A[jj] : A[jj];
  __shared__ double red[LES_BLOCK];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int h = LES_BLOCK / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.C[(long)fld * a.nzg + k] = red[0] / (double)ny;
}

// Pass 3: f -= (-1)^i A_j + (-1)^j B_i - (-1)^(i+j) C, RPW rows per wavefront.
template <int RPW>
__global__ void __launch_bounds__(LES_BLOCK) k_nyq_apply(NyqArgs a) {
  const int nx = a.nx, ny = a.ny, fld = blockIdx.y;
  const int lane = threadIdx.x & (LES_WAVE - 1), wave = threadIdx.x / LES_WAVE;
  const long row0 = ((long)blockIdx.x * (LES_BLOCK / LES_WAVE) + wave) * RPW;
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const long row = row0 + r;
    const int j = (int)(row % ny);
    const long k = row / ny;
    const double sj = (j & 1) ? -1.0 : 1.0;
    const double av = a.A[(long)fld * ny * a.nzg + row], c = a.C[(long)fld * a.nzg + k];
    const double* __restrict__ B = a.B + (long)fld * nx * a.nzg + k * nx;
    double* __restrict__ f = a.f[fld] + row * nx;
    for (int i = lane; i < nx; i += LES_WAVE) {
      const double si = (i & 1) ? -1.0 : 1.0;
      f[i] -= si * av + sj * B[i] - si * sj * c;
    }
  }
}

// Pass 3 fused with the boundary conditions and the vertical derivatives of
// compute_vel_grad (k_bc_ddz), so that u, v, w are read and written once.
// The staggered differences need the Nyquist-free value of a neighbouring
// plane, so a thread walks a segment of one column upwards carrying the
// previous and next planes in registers.  A block holds 64 columns (one
// wavefront) times SEGS segments of the same columns; the raw values just
// outside a segment (planes k0-1 and k1, written by the neighbouring
// segments) are read before a block barrier and every store comes after it.
// Each output location has one writer and the same arithmetic as the two
// unfused passes: plane 0 of u, v takes the wall-law ghost u(1) - lbc dz,
// plane nz+1 the top condition, w(1) = 0 and w(nz+1) = ubc_w.
// The plane index is wavefront-uniform, and so is the row j when nx is a
// multiple of 64 (ROWU): each wavefront then stages its segment's A_jk and
// C_k in LDS (read back as broadcasts), leaving two vector loads per point
// (the raw value and B_ik).  LDS: 2 (len + 2) doubles per wavefront.
template <typename T, int SEGS, bool ROWU>
__global__ void __launch_bounds__(64 * SEGS) k_nyq_apply_ddz(NyqArgs a, BcDdzArgs<T> b) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = a.nx, ny = a.ny, nzg = a.nzg, nz = b.nz, fld = blockIdx.y;
  const long np = (long)nx * ny;
  const long q = (long)blockIdx.x * LES_WAVE + (threadIdx.x & (LES_WAVE - 1));
  const int seg = __builtin_amdgcn_readfirstlane(threadIdx.x / LES_WAVE), len = (nzg + SEGS - 1) / SEGS;
  const int k0 = seg * len, k1 = min(k0 + len, nzg);
  const bool live = q < np && k0 < k1;
  const int j = ROWU ? __builtin_amdgcn_readfirstlane((int)(q / nx)) : (int)(q / nx), i = (int)(q % nx);
  const double si = (i & 1) ? -1.0 : 1.0, sj = (j & 1) ? -1.0 : 1.0;
  const double* __restrict__ A = a.A + (long)fld * ny * nzg + j;
  const double* __restrict__ B = a.B + (long)fld * nx * nzg + i;
  const double* __restrict__ C = a.C + (long)fld * nzg;
  double* f = a.f[fld] + q;  // f[k * np]: plane k of this column, updated in place
  // planes k0-1 .. k1 of A_j and C, at index k - k0 + 1
  double* sA = reinterpret_cast<double*>(smem_raw) + (long)seg * 2 * (len + 2);
  double* sC = sA + (len + 2);
  if (ROWU) {
    for (int t = threadIdx.x & (LES_WAVE - 1); t < len + 2; t += LES_WAVE) {
      const int kk = k0 - 1 + t;
      if (kk >= 0 && kk < nzg) {
        sA[t] = A[(long)kk * ny];
        sC[t] = C[kk];
      }
    }
  }
  // the Nyquist-free value of plane k from its raw value (k_nyq_apply)
  auto filt = [&](int k, double raw) {
    const double av = ROWU ? sA[k - k0 + 1] : A[(long)k * ny], cv = ROWU ? sC[k - k0 + 1] : C[k];
    return raw - (si * av + sj * B[(long)k * nx] - si * sj * cv);
  };
  double below = 0.0, above = 0.0;
  if (live && k0 > 0) below = f[(long)(k0 - 1) * np];
  if (live && k1 < nzg) above = f[(long)k1 * np];
  __syncthreads();
  if (!live) return;
  const double cz = b.cz;
  T* d = fld == 0 ? b.dudz + q : fld == 1 ? b.dvdz + q : b.dwdz + q;
  const double lbc = fld < 2 ? (fld == 0 ? b.lbc_u[q] : b.lbc_v[q]) : 0.0;
  const int neu = fld == 0 ? b.ubc_u_neu : b.ubc_v_neu;
  const double tbc = fld == 0 ? b.ubc_u : b.ubc_v;
  // plane k from the Nyquist-free values of planes k-1, k, k+1
  auto emit = [&](int k, double prev, double cur, double nxt) {
    const long c = (long)k * np;
    if (fld < 2) {
      if (k == 0) {
        f[c] = nxt - lbc * b.dz;
      } else if (k <= nz) {
        const double fm = k == 1 ? cur - lbc * b.dz : prev;
        f[c] = cur;
        d[c] = (T)(cz * (cur - fm));
      } else {
        const double fk = neu ? prev + tbc * b.dz : 2.0 * tbc - prev;
        f[c] = fk;
        d[c] = (T)(cz * (fk - prev));
      }
    } else {
      if (k == 0) {
        f[c] = cur;
      } else if (k <= nz) {
        const double wk = k == 1 ? 0.0 : cur;
        const double wp = k == nz ? b.ubc_w : nxt;
        f[c] = wk;
        d[c] = (T)(cz * (wp - wk));
      } else {
        f[c] = b.ubc_w;
        d[c] = (T)0;
      }
    }
  };
  double prev = k0 > 0 ? filt(k0 - 1, below) : 0.0;
  double cur = filt(k0, f[(long)k0 * np]);
  // four planes at a time: their raw values are loaded (this thread's own
  // planes, stored later by this thread, or the value read before the
  // barrier) before the four stores
  for (int kb = k0; kb < k1; kb += 4) {
    double nxt[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int kk = kb + 1 + m;
      nxt[m] = kk < k1 ? filt(kk, f[(long)kk * np]) : (kk == k1 && kk < nzg) ? filt(kk, above) : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (kb + m < k1) {
        emit(kb + m, prev, cur, nxt[m]);
        prev = cur;
        cur = nxt[m];
      }
    }
  }
}

// Nyquist rows and columns of u, v, w removed in FP64 (three passes, the
// three fields together).
static void nyquist_remove_fp64(Les& L, bool with_ddz) {
  NyqArgs a;
  a.f[0] = D(L, LES_U); a.f[1] = D(L, LES_V); a.f[2] = D(L, LES_W);
  a.A = L.nyq_rows;
  a.B = a.A + 3L * L.ny * L.nzg;
  a.C = a.B + 3L * L.nx * L.nzg;
  a.P = L.nyq_part;
  a.nx = L.nx; a.ny = L.ny; a.nzg = L.nzg;
  const long nrows = (long)L.ny * L.nzg;
  const int rpw = L.ny % 16 == 0 ? 4 : 1;
  const unsigned nblocks = (unsigned)(nrows / (4 * rpw));
  // bands of column partials per plane: 8 rows each when the forward
  // transforms formed the sums, else 4 rpw rows
  const bool fused_sums = nyq_sums_fused(L);
  const int nb = fused_sums ? L.ny / 8 : L.ny / (4 * rpw);
  // algorithmic bytes: the reduction reads u, v, w; the apply pass reads
  // and writes them; the row sums A and column partials are not counted
  // (1/nx and 4/nx of a field)
  if (!fused_sums) {
    kt_begin(L, KT_POINTWISE);
    const size_t shp = sizeof(double) * (LES_BLOCK / LES_WAVE) * (size_t)L.nx;
    if (rpw == 4) hipLaunchKernelGGL(k_nyq_reduce<4>, dim3(nblocks, 3), dim3(LES_BLOCK), shp, L.stream, a);
    else hipLaunchKernelGGL(k_nyq_reduce<1>, dim3(nblocks, 3), dim3(LES_BLOCK), shp, L.stream, a);
    kt_end(L, KT_POINTWISE, 24.0 * L.nf, "k_nyq_reduce");
  }
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_nyq_bc, dim3((unsigned)L.nzg, 3), dim3(LES_BLOCK), 0, L.stream, a, nb);
  kt_end(L, KT_POINTWISE, 3.0 * 8.0 * (double)L.nzg * ((double)nb * L.nx + L.nx), "k_nyq_bc");
  kt_begin(L, KT_POINTWISE);
  if (with_ddz) {
    BcDdzArgs<float> b;
    b.u = D(L, LES_U); b.v = D(L, LES_V); b.w = D(L, LES_W);
    b.dudz = F<float>(L, LES_DUDZ); b.dvdz = F<float>(L, LES_DVDZ); b.dwdz = F<float>(L, LES_DWDZ);
    b.lbc_u = L.layer[LES_LBC_U]; b.lbc_v = L.layer[LES_LBC_V];
    b.np = L.np; b.nz = L.nz; b.dz = L.dz; b.cz = 1.0 / L.dz;
    b.ubc_u_neu = L.P.ubc_u_neu; b.ubc_v_neu = L.P.ubc_v_neu;
    b.ubc_u = L.P.ubc_u; b.ubc_v = L.P.ubc_v; b.ubc_w = L.P.ubc_w;
    const unsigned ncb = (unsigned)((L.np + LES_WAVE - 1) / LES_WAVE);
    constexpr int segs = LES_NYQ_SEGS;
    const size_t sh = sizeof(double) * 2 * segs * (size_t)((L.nzg + segs - 1) / segs + 2);
    if (L.nx % LES_WAVE == 0)
      hipLaunchKernelGGL((k_nyq_apply_ddz<float, segs, true>), dim3(ncb, 3), dim3(64 * segs), sh, L.stream, a, b);
    else
      hipLaunchKernelGGL((k_nyq_apply_ddz<float, segs, false>), dim3(ncb, 3), dim3(64 * segs), 0, L.stream, a, b);
    // u, v, w read and written, their vertical derivatives written
    kt_end(L, KT_POINTWISE, 48.0 * L.nf + 12.0 * L.np * (L.nz + 1), "k_nyq_apply_ddz", 4);
    return;
  }
  if (rpw == 4) hipLaunchKernelGGL(k_nyq_apply<4>, dim3(nblocks, 3), dim3(LES_BLOCK), 0, L.stream, a);
  else hipLaunchKernelGGL(k_nyq_apply<1>, dim3(nblocks, 3), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, 48.0 * L.nf, "k_nyq_apply");
}

// advection step (src/LES_engine.f90:253-283)
static void ab2(Les& L, double tadv1, double tadv2) {
  Ab2Args a;
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W);
  a.rx = D(L, LES_RHSX); a.ry = D(L, LES_RHSY); a.rz = D(L, LES_RHSZ);
  a.fx = D(L, LES_RHSX_F); a.fy = D(L, LES_RHSY_F); a.fz = D(L, LES_RHSZ_F);
  a.dt = L.P.dt; a.t1 = tadv1; a.t2 = tadv2; a.np = L.np; a.nz = L.nz;
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_ab2, dim3(nblk(L.np * (L.nz + 2))), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, 12.0 * 8 * L.np * L.nz, "k_ab2");
}

// ------------------------------------------------------------------ step
#ifndef LES_OVERLAP
#define LES_OVERLAP 1
#endif
// les_step's projection as the store of the pressure-gradient transforms
#ifndef LES_FUSED_PROJ
#define LES_FUSED_PROJ 1
#endif
template <typename T>
static int les_step_t(Les& L) {
  L.jt_total += 1;
  L.jt += 1;
  const int jtt = L.jt_total;
  if (L.timing) hipEventRecord(L.ph_ev[0], L.stream);
  if (L.overlap && !L.timing && L.mixed) {
    // Mixed mode: the horizontal gradients leave u, v, w unchanged, so the
    // wall law (first level only, its own workspaces) runs on the second
    // stream beside them; the Nyquist removal and the vertical derivatives
    // wait for it.  (In FP64 the gradient pass writes the filtered velocity
    // the wall law reads, so the order stays.)
    hipEventRecord(L.ev_fork, L.stream);
    hipStreamWaitEvent(L.stream2, L.ev_fork, 0);
    hipStream_t s1 = L.stream;
    L.stream = L.stream2;
    wall_law<T>(L);
    hipEventRecord(L.ev_join, L.stream2);
    L.stream = s1;
    vel_grads<T>(L);
    hipStreamWaitEvent(L.stream, L.ev_join, 0);
    vel_grad_bc_ddz<T>(L);
  } else {
    wall_law<T>(L);
    compute_vel_grad<T>(L);
  }
  if (L.timing) hipEventRecord(L.ph_ev[1], L.stream);
  if (L.overlap && !L.timing && L.P.dealias == 0) {
    // The SGS stage and the padded convection read the same velocity and
    // gradients and write disjoint buffers: the SGS stage and the spectra of
    // the stress divergence run on the second stream while the convection
    // runs on the first; the rhs update waits for both.  (The nx x ny
    // convections share the spectral workspaces with the dynamic models and
    // stay sequential, as does the per-kernel timing mode.)
    hipEventRecord(L.ev_fork, L.stream);
    hipStreamWaitEvent(L.stream2, L.ev_fork, 0);
    hipStream_t s1 = L.stream;
    L.stream = L.stream2;
    sgs_stress<T>(L, L.jt, jtt);
    div_stress_spectra<T>(L);
    hipEventRecord(L.ev_join, L.stream2);
    L.stream = s1;
    rhs_convection<T>(L);
    hipStreamWaitEvent(L.stream, L.ev_join, 0);
    div_stress_apply<T>(L);
  } else {
    sgs_stress<T>(L, L.jt, jtt);
    if (L.timing) hipEventRecord(L.ph_ev[2], L.stream);
    compute_rhs_adv<T>(L);
  }
  if (L.timing) hipEventRecord(L.ph_ev[3], L.stream);
  double tadv1 = (jtt == 1) ? 1.0 : 1.5;
  double tadv2 = (jtt == 1) ? 0.0 : 1.0 - tadv1;
  // AB2 folded into the pressure right-hand side and the projection: the
  // predictor is recomputed where it is read instead of being stored
  press_stag_array<T>(L, tadv1, true, tadv2, LES_FUSED_PROJ != 0);
  if (LES_FUSED_PROJ) press_project(L, tadv1, tadv2);
  else projection(L, tadv1, true, tadv2);
  if (L.timing) {
    hipEventRecord(L.ph_ev[4], L.stream);
    hipEventSynchronize(L.ph_ev[4]);
    float ms;
    hipEventElapsedTime(&ms, L.ph_ev[0], L.ph_ev[1]); L.phase_ms[0] += ms;
    hipEventElapsedTime(&ms, L.ph_ev[1], L.ph_ev[2]); L.phase_ms[1] += ms;
    hipEventElapsedTime(&ms, L.ph_ev[2], L.ph_ev[3]); L.phase_ms[2] += ms;
    hipEventElapsedTime(&ms, L.ph_ev[3], L.ph_ev[4]); L.phase_ms[3] += ms;
    hipEventElapsedTime(&ms, L.ph_ev[0], L.ph_ev[4]); L.phase_ms[4] += ms;
  }
  return 0;
}

// ------------------------------------------------------------------ set-up
template <typename T>
static void c16_attrs(int big) {
  const void* k[] = {(const void*)k_cols_c16<T, 1, COP_GRAD, 3>, (const void*)k_cols_c16<T, 1, COP_GRAD, 2>,
                     (const void*)k_cols_c16<T, 2, COP_DIV, 1>,  (const void*)k_cols_c16<T, 1, COP_FILT2, 2>,
                     (const void*)k_cols_c16<T, 1, COP_FILT1, 1>, (const void*)k_cols_c16<T, 1, COP_PGRAD, 3>,
                     (const void*)k_cols_c16<T, 1, COP_COPY, 1>};
  for (const void* f : k) hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, big);
}

static int setup_kernel_attrs() {
  const int big = 160 * 1024;
  hipFuncSetAttribute((const void*)k_rows_r2c<double, double>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_r2c<double, float>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_r2c<float, float>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_c2r<double, double>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_c2r<float, float>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_c2r<float, double>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols_256<double, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  c16_attrs<double>(big);
  c16_attrs<float>(big);
  hipFuncSetAttribute((const void*)k_colres_c16<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_colres_c16<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_colres_c16<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_colres_c16<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols_256<float, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols<double, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols<double, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols<float, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols<float, 16>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_resample<double, double, double>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_resample<float, double, float>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_resample<float, float, float>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_rows_resample<float, double, double>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols_resample<double, LES_COLRES_TILE_BYTES / 16>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_cols_resample<float, LES_COLRES_TILE_BYTES / 8>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, big);
  return 0;
}

static void free_all(Les* L) {
  if (!L) return;
  delete L->mlp;
  L->mlp = nullptr;
  for (void* p : L->allocs) hipFree(p);
  L->allocs.clear();
  for (int i = 0; i < KT_COUNT; ++i) { hipEventDestroy(L->ev_a[i]); hipEventDestroy(L->ev_b[i]); }
  for (int i = 0; i < 6; ++i) hipEventDestroy(L->ph_ev[i]);
  hipEventDestroy(L->ev_fork);
  hipEventDestroy(L->ev_join);
  hipEventDestroy(L->ev_feat0);
  hipEventDestroy(L->ev_feat1);
  hipStreamDestroy(L->stream3);
  hipStreamDestroy(L->stream2);
  hipStreamDestroy(L->stream);
}

extern "C" {

const char* les_last_error(void) { return g_err; }

int les_init(const les_params* p) {
  if (g_les) { les_set_error("les_init: already initialised"); return 4; }
  if (!p) { les_set_error("les_init: null params"); return 1; }
  const les_params P = *p;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { les_set_error("les_init: no HIP device"); return 5; }
  if (int rc = scratch_lookup(dev)) return rc;
  for (int n : {P.nx, P.ny}) {
    int m = n;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    if (m != 1 || n % 4 || n < 8) { les_set_error("les_init: nx, ny must be 2^a 3^b, multiples of 4, >= 8"); return 10; }
  }
  if (P.nz < 3) { les_set_error("les_init: nz must be >= 3"); return 10; }
  if (P.sgs_model < 0 || P.sgs_model > 3) { les_set_error("Invalid SGS model"); return 1; }
  if (P.sgs_model == 3 && !(P.re > 0.0)) { les_set_error("sgs_model dns needs Re > 0"); return 1; }
  if (P.precision == 2 && P.sgs_model != 2) {
    les_set_error("the learned SGS surrogate replaces dyn_LagScaleDep only");
    return 1;
  }
  if (P.cs_count < 1) { les_set_error("cs_count must be >= 1"); return 1; }
  if (P.dealias < 0 || P.dealias > 3) { les_set_error("dealias must be 0 PADD, 1 TRUNC, 2 FSMTHG or 3 NODLSG"); return 1; }
  if (P.wall_model != 0 && P.wall_model != 1) { les_set_error("Invalid wall law type"); return 1; }
  if (P.precision < 0 || P.precision > 2) { les_set_error("precision must be 0, 1 or 2"); return 1; }
  Les* L = new Les();
  std::memset((void*)L, 0, sizeof(Les) - sizeof(std::vector<void*>));
  new (&L->allocs) std::vector<void*>();
  L->P = P;
  L->dev = dev;
  L->nx = P.nx; L->ny = P.ny; L->nz = P.nz; L->nzg = P.nz + 2; L->lhx = P.nx / 2 + 1;
  L->nx2 = 3 * P.nx / 2; L->ny2 = 3 * P.ny / 2;
  L->np = (long)P.nx * P.ny; L->nf = L->np * L->nzg; L->ns = (long)L->lhx * P.ny; L->nsp = L->ns * L->nzg;
  L->nbp = (long)L->nx2 * L->ny2;
  L->dx = P.lx / (double)P.nx; L->dy = P.ly / (double)P.ny; L->dz = P.lz / (double)P.nz;
  L->delta = std::pow(L->dx * L->dy * L->dz, 1.0 / 3.0);
  // near-wall band of the learned mode: SGS level k sits at z = dz/2 (k = 1)
  // or (k-1) dz, the height the surrogate's log(z/delta) feature uses
  L->nband = 0;
  for (int k = 1; k <= P.nz + 1; ++k) {
    const double z = k == 1 ? 0.5 * L->dz : (double)(k - 1) * L->dz;
    if (z / L->delta < P.learned_zmin) L->nband = k;
  }
  L->jt_total = 0;
  L->jt = 0;
  L->overlap = LES_OVERLAP;
  L->F_coeff_init = P.lagran_init != 0;
  L->mixed = P.precision != 0;
  if (hipStreamCreate(&L->stream) != hipSuccess) { delete L; les_set_error("stream"); return 5; }
  if (hipStreamCreate(&L->stream2) != hipSuccess) {
    hipStreamDestroy(L->stream);
    delete L;
    les_set_error("stream");
    return 5;
  }
  if (hipStreamCreate(&L->stream3) != hipSuccess) {
    hipStreamDestroy(L->stream2);
    hipStreamDestroy(L->stream);
    delete L;
    les_set_error("stream");
    return 5;
  }
  hipEventCreateWithFlags(&L->ev_fork, hipEventDisableTiming);
  hipEventCreateWithFlags(&L->ev_join, hipEventDisableTiming);
  hipEventCreateWithFlags(&L->ev_feat0, hipEventDisableTiming);
  hipEventCreateWithFlags(&L->ev_feat1, hipEventDisableTiming);
  {
    int dev = 0;
    L->ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&L->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        L->ncu < 1)
      L->ncu = 256;
  }
  for (int i = 0; i < KT_COUNT; ++i) { hipEventCreate(&L->ev_a[i]); hipEventCreate(&L->ev_b[i]); }
  for (int i = 0; i < 6; ++i) hipEventCreate(&L->ph_ev[i]);
  g_les = L;
  // ---- plans and twiddles
  HostPlan hp[4];
  int sizes[4] = {L->nx, L->ny, L->nx2, L->ny2};
  int off = 0;
  for (int i = 0; i < 4; ++i) {
    int used = 0;
    if (!build_plan(sizes[i], hp[i], off, &used)) { les_set_error("fft plan %d", sizes[i]); les_finalize(); return 10; }
    off += used;
  }
  std::vector<Cplx<double>> t64(off);
  std::vector<Cplx<float>> t32(off);
  for (int i = 0; i < 4; ++i) {
    int sz = (int)hp[i].re_f.size();
    for (int q = 0; q < sz; ++q) {
      t64[hp[i].p.twf + q] = {hp[i].re_f[q], hp[i].im_f[q]};
      t64[hp[i].p.twi + q] = {hp[i].re_i[q], hp[i].im_i[q]};
    }
  }
  for (int q = 0; q < off; ++q) t32[q] = {(float)t64[q].x, (float)t64[q].y};
  L->px = hp[0].p; L->py = hp[1].p; L->px2 = hp[2].p; L->py2 = hp[3].p;
  L->tw64 = (Cplx<double>*)dalloc(*L, sizeof(Cplx<double>) * off);
  L->tw32 = (Cplx<float>*)dalloc(*L, sizeof(Cplx<float>) * off);
  if (!L->tw64 || !L->tw32) { les_finalize(); les_set_error("out of memory"); return 3; }
  hipMemcpy(L->tw64, t64.data(), sizeof(Cplx<double>) * off, hipMemcpyHostToDevice);
  hipMemcpy(L->tw32, t32.data(), sizeof(Cplx<float>) * off, hipMemcpyHostToDevice);
  // ---- spectral tables (src/fft_engine.f90:97-145; src/testfilter.f90:101-131)
  int nx = L->nx, ny = L->ny, lhx = L->lhx;
  std::vector<double> kkx(L->ns), kky(L->ns), kk2(L->ns), G2(L->ns), G4(L->ns);
  double cx = 2.0 * PI / P.lx, cy = 2.0 * PI / P.ly;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < lhx; ++i) {
      double kx = (double)i;
      int m = (j + ny / 2) % ny;
      double ky = (double)(m - ny / 2);
      if (i == lhx - 1 || j == ny / 2) { kx = 0.0; ky = 0.0; }
      kx = cx * kx;
      ky = cy * ky;
      long q = i + (long)lhx * j;
      kkx[q] = kx; kky[q] = ky; kk2[q] = kx * kx + ky * ky;
    }
  for (int a = 0; a < 2; ++a) {
    double alpha = a == 0 ? 2.0 : 4.0;
    std::vector<double>& G = a == 0 ? G2 : G4;
    double dl = alpha * std::sqrt(L->dx * L->dy);
    double t = PI / dl;
    double kc2 = t * t;
    for (long q = 0; q < L->ns; ++q) G[q] = kk2[q] >= kc2 ? 0.0 : 1.0 / (double)(nx * ny);
    for (int j = 0; j < ny; ++j) G[(lhx - 1) + (long)lhx * j] = 0.0;
    for (int i = 0; i < lhx; ++i) G[i + (long)lhx * (ny / 2)] = 0.0;
  }
  // Dealiasing multipliers of TRUNC and FSMTHG on the half spectrum.  The
  // reference filters each real line along x, then along y, with 1D real
  // transforms that zero or scale entry m of the line's half spectrum; the
  // two stages together multiply the 2D spectrum by Fx(i) Fy(|ky|).
  //  TRUNC sources: scr(tt:lh) = 0, tt = ceiling(n/3): modes m >= tt-1
  //    (src/compute_convective_trunc.f90:279-304); products: tt = ceiling(2 lh/3)
  //    (:322-348), lh = n/2 + 1;
  //  FSMTHG: F(m) = exp(-36 (m/(lh-1))^36), F(lh-1) = 0, sources and products
  //    (src/compute_convective_Fsmthg.f90:82-100).
  if (P.dealias == 1 || P.dealias == 2) {
    const int lhy = ny / 2 + 1;
    std::vector<double> gab[2] = {std::vector<double>(L->ns), std::vector<double>(L->ns)};
    for (int t = 0; t < 2; ++t) {
      std::vector<double> fx(lhx), fy(lhy);
      if (P.dealias == 1) {
        const int ttx = t == 0 ? (nx + 2) / 3 : (2 * lhx + 2) / 3;
        const int tty = t == 0 ? (ny + 2) / 3 : (2 * lhy + 2) / 3;
        for (int m = 0; m < lhx; ++m) fx[m] = m <= ttx - 2 ? 1.0 : 0.0;
        for (int m = 0; m < lhy; ++m) fy[m] = m <= tty - 2 ? 1.0 : 0.0;
      } else {
        for (int m = 0; m < lhx; ++m) fx[m] = std::exp(-36.0 * std::pow((double)m / (double)(lhx - 1), 36.0));
        for (int m = 0; m < lhy; ++m) fy[m] = std::exp(-36.0 * std::pow((double)m / (double)(lhy - 1), 36.0));
        fx[lhx - 1] = 0.0;
        fy[lhy - 1] = 0.0;
      }
      for (int j = 0; j < ny; ++j) {
        const int my = j <= ny / 2 ? j : ny - j;
        for (int i = 0; i < lhx; ++i) gab[t][i + (long)lhx * j] = fx[i] * fy[my];
      }
    }
    std::vector<float> g32(L->ns);
    double** d64g[2] = {&L->ga_64, &L->gb_64};
    float** d32g[2] = {&L->ga_32, &L->gb_32};
    for (int t = 0; t < 2; ++t) {
      *d64g[t] = (double*)dalloc(*L, sizeof(double) * L->ns);
      *d32g[t] = (float*)dalloc(*L, sizeof(float) * L->ns);
      if (!*d64g[t] || !*d32g[t]) { les_finalize(); les_set_error("out of memory"); return 3; }
      hipMemcpy(*d64g[t], gab[t].data(), sizeof(double) * L->ns, hipMemcpyHostToDevice);
      for (long q = 0; q < L->ns; ++q) g32[q] = (float)gab[t][q];
      hipMemcpy(*d32g[t], g32.data(), sizeof(float) * L->ns, hipMemcpyHostToDevice);
    }
  }
  std::vector<float> f32(L->ns);
  double** d64[4] = {&L->kkx64, &L->kky64, &L->g2_64, &L->g4_64};
  float** d32[4] = {&L->kkx32, &L->kky32, &L->g2_32, &L->g4_32};
  std::vector<double>* hsrc[4] = {&kkx, &kky, &G2, &G4};
  for (int t = 0; t < 4; ++t) {
    *d64[t] = (double*)dalloc(*L, sizeof(double) * L->ns);
    *d32[t] = (float*)dalloc(*L, sizeof(float) * L->ns);
    if (!*d64[t] || !*d32[t]) { les_finalize(); les_set_error("out of memory"); return 3; }
    hipMemcpy(*d64[t], hsrc[t]->data(), sizeof(double) * L->ns, hipMemcpyHostToDevice);
    for (long q = 0; q < L->ns; ++q) f32[q] = (float)(*hsrc[t])[q];
    hipMemcpy(*d32[t], f32.data(), sizeof(float) * L->ns, hipMemcpyHostToDevice);
  }
  // ---- surface: z0 and log(0.5 dz / z0)
  std::vector<double> lt(L->np, std::log((0.5 * L->dz) / P.z0)), z0(L->np, P.z0);
  L->logterm = (double*)dalloc(*L, sizeof(double) * L->np);
  for (int i = 0; i < LES_NLAYERS; ++i) L->layer[i] = (double*)dalloc(*L, sizeof(double) * L->np);
  hipMemcpy(L->logterm, lt.data(), sizeof(double) * L->np, hipMemcpyHostToDevice);
  hipMemcpy(L->layer[LES_Z0], z0.data(), sizeof(double) * L->np, hipMemcpyHostToDevice);
  // ---- mixing length per level (src/compute_sgs.f90:142-163)
  std::vector<double> lv(L->nzg, L->delta);
  if (P.sgs_model == 0) {
    double n = 2.0, z = 0.5 * L->dz;
    lv[1] = std::pow(std::pow(P.cs0, n) * std::pow(KVONK * z, -n) + std::pow(L->delta, -n), -1.0 / n);
    for (int k = 2; k <= L->nz + 1; ++k) {
      z = ((double)k - 1.0) * L->dz;
      lv[k] = std::pow(std::pow(P.cs0, n) * std::pow(KVONK * z, -n) + std::pow(L->delta, -n), -1.0 / n);
    }
  }
  L->lvec = (double*)dalloc(*L, sizeof(double) * L->nzg);
  hipMemcpy(L->lvec, lv.data(), sizeof(double) * L->nzg, hipMemcpyHostToDevice);
  // ---- fields
  size_t tb = L->mixed ? 4 : 8;
  for (int id = 0; id < LES_NFIELDS; ++id) {
    L->fbytes[id] = field_is_fp64(id) ? 8 : (int)tb;
    L->f[id] = dalloc(*L, (size_t)L->fbytes[id] * L->nf);
    if (!L->f[id]) { les_finalize(); les_set_error("out of memory (fields)"); return 3; }
  }
  // w(:,:,0) = BOGUS (src/system.f90:196-198) is set by les_set_ic.
  // ---- workspaces
  for (int i = 0; i < LES_NSPEC; ++i) L->wk_spec[i] = dalloc(*L, 2 * tb * L->nsp);
  for (int i = 0; i < 64; ++i) L->wk_real[i] = dalloc(*L, tb * L->nf);
  for (int i = 0; i < 6; ++i) L->wk_hyb[i] = dalloc(*L, tb * (size_t)L->nx2 * L->ny * L->nzg);
  for (int i = 0; i < 6; ++i) L->wk_big[i] = dalloc(*L, tb * (size_t)L->nbp * L->nzg);
  for (int i = 0; i < 3; ++i) L->wk_prod[i] = dalloc(*L, tb * (size_t)L->nbp * L->nzg);
  for (int i = 0; i < 3; ++i) L->wk_hyb3[i] = dalloc(*L, tb * (size_t)L->nx2 * L->ny * L->nzg);
  for (int i = 0; i < 4; ++i) L->pspec[i] = (double*)dalloc(*L, 16 * (size_t)L->nsp);
  L->tri_gam = (double*)dalloc(*L, 8 * (size_t)L->ns * (L->nz + 3));
  for (int i = 0; i < 4; ++i) L->playr[i] = (double*)dalloc(*L, 16 * (size_t)L->ns);
  for (int i = 0; i < 6; ++i) L->wk_layer[i] = dalloc(*L, 2 * tb * (size_t)L->ns + 2 * tb * L->np);
  for (int i = 0; i < 4; ++i) L->Fold[i] = (double*)dalloc(*L, 8 * L->nf);
  L->nyq_rows = (double*)dalloc(*L, 8 * 3 * (size_t)(L->nx + L->ny + 2) * L->nzg);
  L->nyq_part = (double*)dalloc(*L, 8 * 3 * (size_t)(L->ny / 4) * L->nx * L->nzg);
  L->nanflag = (int*)dalloc(*L, 64);
  L->dg_part = (double*)dalloc(*L, sizeof(double) * LES_DIAG_BLOCKS * DG_N);
  L->dg_out = (double*)dalloc(*L, sizeof(double) * 4);
  if (L->alloc_failed) { les_finalize(); les_set_error("out of memory (workspaces)"); return 3; }
  for (int i = 0; i < 64; ++i) if (!L->wk_real[i]) { les_finalize(); les_set_error("out of memory"); return 3; }
  setup_kernel_attrs();
  if (hipDeviceSynchronize() != hipSuccess) { les_set_error("init sync failed"); les_finalize(); return 5; }
  return 0;
}

int les_finalize(void) {
  if (!g_les) return 0;
  hipStreamSynchronize(g_les->stream);
  free_all(g_les);
  g_les->allocs.~vector();
  ::operator delete((void*)g_les);
  g_les = nullptr;
  return 0;
}

int les_sync(void) {
  if (!g_les) return 1;
  LES_CHECK(hipStreamSynchronize(g_les->stream));
  return 0;
}

int les_jt_total(void) { return g_les ? g_les->jt_total : -1; }
int les_set_jt_total(int jt) {
  if (!g_les) return 1;
  g_les->jt_total = jt;
  return 0;
}

// Host <-> device copies with precision conversion.
__global__ void k_cvt_d2f(const double* a, float* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_cvt_f2d(const float* a, double* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

static int copy_field(int id, double* host, int up) {
  Les* L = g_les;
  if (!L) { les_set_error("not initialised"); return 1; }
  if (id < 0 || id >= LES_NFIELDS) { les_set_error("bad field id %d", id); return 10; }
  // the step leaves the pressure gradient to be formed on request
  if (id == LES_DPDX || id == LES_DPDY || id == LES_DPDZ || (up && id == LES_P)) refresh_gradp(*L);
  size_t n = (size_t)L->nf;
  if (L->fbytes[id] == 8) {
    if (up) { LES_CHECK(hipMemcpyAsync(L->f[id], host, 8 * n, hipMemcpyHostToDevice, L->stream)); }
    else { LES_CHECK(hipMemcpyAsync(host, L->f[id], 8 * n, hipMemcpyDeviceToHost, L->stream)); }
  } else {
    double* stage = nullptr;
    LES_CHECK(hipMallocAsync((void**)&stage, 8 * n, L->stream));
    if (up) {
      LES_CHECK(hipMemcpyAsync(stage, host, 8 * n, hipMemcpyHostToDevice, L->stream));
      hipLaunchKernelGGL(k_cvt_d2f, dim3(nblk((long)n)), dim3(LES_BLOCK), 0, L->stream, stage, (float*)L->f[id], (long)n);
    } else {
      hipLaunchKernelGGL(k_cvt_f2d, dim3(nblk((long)n)), dim3(LES_BLOCK), 0, L->stream, (const float*)L->f[id], stage, (long)n);
      LES_CHECK(hipMemcpyAsync(host, stage, 8 * n, hipMemcpyDeviceToHost, L->stream));
    }
    LES_CHECK(hipFreeAsync(stage, L->stream));
  }
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

int les_upload_field(int id, const double* host) { return copy_field(id, (double*)host, 1); }
int les_download_field(int id, double* host) { return copy_field(id, host, 0); }

int les_field_ptr(int id, void** dptr, int* elem_bytes) {
  if (!g_les) return 1;
  if (id < 0 || id >= LES_NFIELDS) return 10;
  if (id == LES_DPDX || id == LES_DPDY || id == LES_DPDZ || id == LES_P) {
    refresh_gradp(*g_les);
    LES_CHECK(hipStreamSynchronize(g_les->stream));
  }
  *dptr = g_les->f[id];
  *elem_bytes = g_les->fbytes[id];
  return 0;
}

int les_upload_layer(int id, const double* host) {
  Les* L = g_les;
  if (!L) return 1;
  if (id < 0 || id >= LES_NLAYERS) return 10;
  LES_CHECK(hipMemcpy(L->layer[id], host, 8 * L->np, hipMemcpyHostToDevice));
  if (id == LES_Z0) {
    std::vector<double> z0(L->np), lt(L->np);
    std::memcpy(z0.data(), host, 8 * L->np);
    for (long q = 0; q < L->np; ++q) lt[q] = std::log((0.5 * L->dz) / z0[q]);
    LES_CHECK(hipMemcpy(L->logterm, lt.data(), 8 * L->np, hipMemcpyHostToDevice));
  }
  return 0;
}

int les_download_layer(int id, double* host) {
  Les* L = g_les;
  if (!L) return 1;
  if (id < 0 || id >= LES_NLAYERS) return 10;
  LES_CHECK(hipStreamSynchronize(L->stream));
  LES_CHECK(hipMemcpy(host, L->layer[id], 8 * L->np, hipMemcpyDeviceToHost));
  return 0;
}

int les_set_ic(const double* u0, const double* v0, const double* w0) {
  Les* L = g_les;
  if (!L) return 1;
  std::vector<double> h(L->nf, 0.0);
  const double* src[3] = {u0, v0, w0};
  for (int c = 0; c < 3; ++c) {
    std::fill(h.begin(), h.end(), 0.0);
    for (int k = 1; k <= L->nz; ++k) std::memcpy(&h[(size_t)k * L->np], src[c] + (size_t)(k - 1) * L->np, 8 * L->np);
    int rc = les_upload_field(LES_U + c, h.data());
    if (rc) return rc;
  }
  IcBcArgs a;
  a.u = D(*L, LES_U); a.v = D(*L, LES_V); a.w = D(*L, LES_W);
  a.np = L->np; a.nz = L->nz; a.dz = L->dz;
  a.ubc_u_neu = L->P.ubc_u_neu; a.ubc_v_neu = L->P.ubc_v_neu;
  a.ubc_u = L->P.ubc_u; a.ubc_v = L->P.ubc_v; a.ubc_w = L->P.ubc_w;
  a.bogus = BOGUS;
  hipLaunchKernelGGL(k_ic_bc, dim3(nblk(L->np)), dim3(LES_BLOCK), 0, L->stream, a);
  LES_CHECK(hipGetLastError());
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

int les_step(int nsteps) {
  Les* L = g_les;
  if (!L) { les_set_error("not initialised"); return 1; }
  for (int s = 0; s < nsteps; ++s) {
    int rc = L->mixed ? les_step_t<float>(*L) : les_step_t<double>(*L);
    if (rc) return rc;
  }
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_check_nan(int* found) {
  Les* L = g_les;
  if (!L) return 1;
  LES_CHECK(hipMemsetAsync(L->nanflag, 0, 4, L->stream));
  NanArgs a;
  a.u = D(*L, LES_U); a.v = D(*L, LES_V); a.w = D(*L, LES_W);
  a.n = L->np * L->nz; a.off = L->np; a.flag = L->nanflag;
  hipLaunchKernelGGL(k_nan, dim3(nblk(a.n)), dim3(LES_BLOCK), 0, L->stream, a);
  int h = 0;
  LES_CHECK(hipMemcpyAsync(&h, L->nanflag, 4, hipMemcpyDeviceToHost, L->stream));
  LES_CHECK(hipStreamSynchronize(L->stream));
  if (found) *found = h;
  if (h) { les_set_error("I got a NaN"); return 9; }
  return 0;
}

int les_timers(double* t, int reset) {
  Les* L = g_les;
  if (!L) return 1;
  if (t) {
    for (int i = 0; i < 5; ++i) t[i] = L->phase_ms[i];
    for (int i = 0; i < KT_COUNT; ++i) {
      t[5 + 3 * i] = L->kt_ms[i];
      t[6 + 3 * i] = (double)L->kt_calls[i];
      t[7 + 3 * i] = L->kt_bytes[i];
    }
  }
  if (reset) {
    for (int i = 0; i < 5; ++i) L->phase_ms[i] = 0;
    for (int i = 0; i < KT_COUNT; ++i) { L->kt_ms[i] = 0; L->kt_calls[i] = 0; L->kt_bytes[i] = 0; }
    L->nkst = 0;
  }
  return 0;
}

int les_kernel_stat(int i, int* count, char* name, int name_cap, double* v) {
  Les* L = g_les;
  if (!L) return 1;
  if (count) *count = L->nkst;
  if (i < 0 && !name && !v) return 0;
  if (i < 0 || i >= L->nkst) { les_set_error("les_kernel_stat: index %d of %d", i, L->nkst); return 10; }
  const Les::KStat& k = L->kst[i];
  if (name && name_cap > 0) {
    auto tn = [](int b) { return b == 8 ? "double" : b == 4 ? "float" : b == 2 ? "bf16" : "?"; };
    char buf[160];
    if (k.ta && k.tb) snprintf(buf, sizeof(buf), "%s<%s, %s>", k.base, tn(k.ta), tn(k.tb));
    else if (k.ta) snprintf(buf, sizeof(buf), "%s<%s>", k.base, tn(k.ta));
    else snprintf(buf, sizeof(buf), "%s", k.base);
    if (k.tag) {
      const size_t n = std::strlen(buf);
      snprintf(buf + n, sizeof(buf) - n, " [%s]", k.tag);
    }
    snprintf(name, (size_t)name_cap, "%s", buf);
  }
  if (v) { v[0] = k.ms; v[1] = (double)k.calls; v[2] = k.bytes; v[3] = (double)k.fam; }
  return 0;
}

int les_set_timing(int on) {
  if (!g_les) return 1;
  g_les->timing = on;
  return 0;
}

int les_set_overlap(int on) {
  if (!g_les) return 1;
  g_les->overlap = on ? 1 : 0;
  return 0;
}

// ---- operator-level entry points
int les_wall_law(void) {
  if (!g_les) return 1;
  if (g_les->mixed) wall_law<float>(*g_les); else wall_law<double>(*g_les);
  LES_CHECK(hipGetLastError());
  return 0;
}
int les_compute_vel_grad(void) {
  if (!g_les) return 1;
  if (g_les->mixed) compute_vel_grad<float>(*g_les); else compute_vel_grad<double>(*g_les);
  LES_CHECK(hipGetLastError());
  return 0;
}
int les_grad_velocity_filt(int f, int dfdx, int dfdy) {
  if (!g_les) return 1;
  if (f < LES_U || f > LES_W) { les_set_error("grad_velocity_filt: f must be u, v or w"); return 10; }
  if (g_les->mixed) grad_filt<float>(*g_les, f, dfdx, dfdy); else grad_filt<double>(*g_les, f, dfdx, dfdy);
  LES_CHECK(hipGetLastError());
  return 0;
}
int les_sgs_stress(int jt, int jtt) {
  if (!g_les) return 1;
  if (g_les->mixed) sgs_stress<float>(*g_les, jt, jtt); else sgs_stress<double>(*g_les, jt, jtt);
  LES_CHECK(hipGetLastError());
  return 0;
}
int les_compute_rhs_adv(void) {
  if (!g_les) return 1;
  if (g_les->mixed) compute_rhs_adv<float>(*g_les); else compute_rhs_adv<double>(*g_les);
  LES_CHECK(hipGetLastError());
  return 0;
}
int les_press_stag_array(double tadv1) {
  if (!g_les) return 1;
  if (g_les->mixed) press_stag_array<float>(*g_les, tadv1); else press_stag_array<double>(*g_les, tadv1);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_fft2d_r2c(const void* d_in, void* d_out, int nplanes, double scale) {
  Les* L = g_les;
  if (!L) return 1;
  if (L->mixed) {
    const float* s[1] = {(const float*)d_in};
    Cplx<float>* d[1] = {(Cplx<float>*)d_out};
    rows_r2c<float, float>(*L, s, d, 1, 0, nplanes, (float)scale);
    cols<float>(*L, COP_COPY, d[0], nullptr, d[0], nullptr, nullptr, 1, 1, 0, 0, nplanes);
  } else {
    const double* s[1] = {(const double*)d_in};
    Cplx<double>* d[1] = {(Cplx<double>*)d_out};
    rows_r2c<double, double>(*L, s, d, 1, 0, nplanes, scale);
    cols<double>(*L, COP_COPY, d[0], nullptr, d[0], nullptr, nullptr, 1, 1, 0, 0, nplanes);
  }
  LES_CHECK(hipGetLastError());
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

int les_fft2d_c2r(const void* d_in, void* d_out, int nplanes) {
  Les* L = g_les;
  if (!L) return 1;
  // the input is preserved (unlike FFTW's c2r): the y pass goes to a workspace
  size_t need = (size_t)nplanes * L->ns;
  if (need > (size_t)L->nsp) { les_set_error("les_fft2d_c2r: at most nz+2 planes"); return 10; }
  if (L->mixed) {
    Cplx<float>* w = (Cplx<float>*)L->wk_spec[7];
    cols<float>(*L, COP_COPY, (const Cplx<float>*)d_in, nullptr, w, nullptr, nullptr, 1, 0, 1, 0, nplanes);
    const Cplx<float>* s[1] = {w};
    float* d[1] = {(float*)d_out};
    rows_c2r<float, float>(*L, s, d, 1, 0, nplanes);
  } else {
    Cplx<double>* w = (Cplx<double>*)L->wk_spec[7];
    cols<double>(*L, COP_COPY, (const Cplx<double>*)d_in, nullptr, w, nullptr, nullptr, 1, 0, 1, 0, nplanes);
    const Cplx<double>* s[1] = {w};
    double* d[1] = {(double*)d_out};
    rows_c2r<double, double>(*L, s, d, 1, 0, nplanes);
  }
  LES_CHECK(hipGetLastError());
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- generic ddz
// ddz_uv (uvp -> w, levels 1..nz+1) and ddz_w (w -> uvp, levels 1..nz, top
// level zero), src/compute_derivatives.f90:375-441, on fields of either
// precision.
template <typename TI, typename TO>
__global__ void k_ddz(const TI* f, TO* d, long np, int nz, double cz, int w_nodes) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * (nz + 1)) return;
  const long c = np + t;
  const int k = 1 + (int)(t / np);
  if (w_nodes) {
    d[c] = (k <= nz) ? (TO)(cz * ((double)f[c + np] - (double)f[c])) : (TO)0;
  } else {
    d[c] = (TO)(cz * ((double)f[c] - (double)f[c - np]));
  }
}

static int ddz_any(int dfdz, int f, int w_nodes) {
  Les* L = g_les;
  if (!L) return 1;
  if (dfdz < 0 || dfdz >= LES_NFIELDS || f < 0 || f >= LES_NFIELDS) return 10;
  const double cz = 1.0 / L->dz;
  dim3 g(nblk(L->np * (L->nz + 1)));
  const bool fi = L->fbytes[f] == 8, fo = L->fbytes[dfdz] == 8;
  if (fi && fo) hipLaunchKernelGGL((k_ddz<double, double>), g, dim3(LES_BLOCK), 0, L->stream, (const double*)L->f[f], (double*)L->f[dfdz], L->np, L->nz, cz, w_nodes);
  else if (fi) hipLaunchKernelGGL((k_ddz<double, float>), g, dim3(LES_BLOCK), 0, L->stream, (const double*)L->f[f], (float*)L->f[dfdz], L->np, L->nz, cz, w_nodes);
  else if (fo) hipLaunchKernelGGL((k_ddz<float, double>), g, dim3(LES_BLOCK), 0, L->stream, (const float*)L->f[f], (double*)L->f[dfdz], L->np, L->nz, cz, w_nodes);
  else hipLaunchKernelGGL((k_ddz<float, float>), g, dim3(LES_BLOCK), 0, L->stream, (const float*)L->f[f], (float*)L->f[dfdz], L->np, L->nz, cz, w_nodes);
  LES_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
static int test_filter_t(Les& L, int fid, int alpha) {
  Cplx<T>* s0 = (Cplx<T>*)L.wk_spec[0];
  Cplx<T>* s1 = (Cplx<T>*)L.wk_spec[1];
  const T* src[1] = {F<T>(L, fid)};
  Cplx<T>* dst[1] = {s0};
  rows_r2c<T, T>(L, src, dst, 1, 1, L.nz + 1, (T)1);
  const T* g = alpha == 2 ? (sizeof(T) == 8 ? (const T*)L.g2_64 : (const T*)L.g2_32)
                          : (sizeof(T) == 8 ? (const T*)L.g4_64 : (const T*)L.g4_32);
  cols<T>(L, COP_FILT1, s0, nullptr, s1, nullptr, nullptr, 1, 1, 1, 1, L.nz + 1, g);
  const Cplx<T>* cs[1] = {s1};
  T* outs[1] = {F<T>(L, fid)};
  rows_c2r<T, T>(L, cs, outs, 1, 1, L.nz + 1);
  return 0;
}

extern "C" {
int les_ddz_uv(int dfdz, int f) { return ddz_any(dfdz, f, 0); }
int les_ddz_w(int dfdz, int f) { return ddz_any(dfdz, f, 1); }

int les_compute_dealias_convec(int cx, int cy, int cz) {
  Les* L = g_les;
  if (!L) return 1;
  for (int id : {cx, cy, cz})
    if (id < 0 || id >= LES_NFIELDS || L->fbytes[id] != 8) { les_set_error("convective output must be an FP64 field"); return 10; }
  if (L->mixed) convection<float>(*L, D(*L, cx), D(*L, cy), D(*L, cz));
  else convection<double>(*L, D(*L, cx), D(*L, cy), D(*L, cz));
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_div_stress(int rhs, int tx, int ty, int tz, int stag, int save_wk) {
  Les* L = g_les;
  if (!L) return 1;
  if (L->fbytes[rhs] != 8) { les_set_error("div_stress: rhs must be an FP64 field"); return 10; }
  for (int id : {tx, ty, tz})
    if (id < 0 || id >= LES_NFIELDS || field_is_fp64(id)) { les_set_error("div_stress: stress fields expected"); return 10; }
  if (L->mixed) div_stress<float>(*L, rhs, tx, ty, tz, stag, save_wk, 0);
  else div_stress<double>(*L, rhs, tx, ty, tz, stag, save_wk, 0);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_ab2(double tadv1, double tadv2) {
  if (!g_les) return 1;
  ab2(*g_les, tadv1, tadv2);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_projection(double tadv1) {
  if (!g_les) return 1;
  projection(*g_les, tadv1);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_test_filter(int f, int alpha) {
  Les* L = g_les;
  if (!L) return 1;
  if (f < 0 || f >= LES_NFIELDS || field_is_fp64(f)) { les_set_error("test_filter: flow-precision field expected"); return 10; }
  if (alpha != 2 && alpha != 4) { les_set_error("test_filter: alpha must be 2 or 4"); return 1; }
  if (L->mixed) test_filter_t<float>(*L, f, alpha); else test_filter_t<double>(*L, f, alpha);
  LES_CHECK(hipGetLastError());
  return 0;
}
}  // extern "C"

// ---- running statistics and diagnostics
template <typename T>
static void stat_accumulate(Les& L) {
  StatAccArgs<T> a;
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W); a.p = D(L, LES_P);
  a.ustar = L.layer[LES_USTAR];
  a.txx = F<T>(L, LES_TXX); a.tyy = F<T>(L, LES_TYY); a.tzz = F<T>(L, LES_TZZ);
  a.txy = F<T>(L, LES_TXY); a.txz = F<T>(L, LES_TXZ); a.tyz = F<T>(L, LES_TYZ);
  a.dudz = F<T>(L, LES_DUDZ); a.dvdz = F<T>(L, LES_DVDZ); a.nut = F<T>(L, LES_NU_T); a.cs = F<T>(L, LES_CS_OPT2);
  a.s = L.st_sum; a.sus = L.st_ustar;
  a.np = L.np; a.n = L.np * L.nz;
  hipLaunchKernelGGL(k_stat_accum<T>, dim3(nblk(a.n)), dim3(LES_BLOCK), 0, L.stream, a);
}

template <typename T>
static void diagnostics(Les& L) {
  DiagArgs<T> a;
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W); a.ustar = L.layer[LES_USTAR];
  a.dudx = F<T>(L, LES_DUDX); a.dvdy = F<T>(L, LES_DVDY); a.dwdz = F<T>(L, LES_DWDZ);
  a.part = L.dg_part; a.np = L.np; a.n = L.np * L.nz;
  hipLaunchKernelGGL(k_diag<T>, dim3(LES_DIAG_BLOCKS), dim3(LES_BLOCK), 0, L.stream, a);
  DiagFinalArgs b;
  b.part = L.dg_part; b.out = L.dg_out;
  b.dt = L.P.dt; b.dx = L.dx; b.dy = L.dy; b.dz = L.dz;
  b.nxyz = (double)L.nx * (double)L.ny * (double)L.nz;
  b.nxy = (double)L.nx * (double)L.ny;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(LES_BLOCK), 0, L.stream, b);
}

// compute_tke_budget: the cleaned pressure (formed on load by the forward
// row transform) through the FP64 grad_velocity_filt chain (the filtered
// spectrum in place in the column pass like grad_filt3, the filtered field
// into tk_pf), then the plane sums of the 61 quantities in one pass over the
// fields, then the window sums.
template <typename T>
static void tke_accumulate(Les& L) {
  Cplx<double>* s0 = (Cplx<double>*)L.tk_spec[0];
  Cplx<double>* s1 = (Cplx<double>*)L.tk_spec[1];
  Cplx<double>* s2 = (Cplx<double>*)L.tk_spec[2];
  // the cleaned pressure formed on load (LOAD_PCLEAN)
  const double* src[1] = {D(L, LES_P)};
  const double* uvw[3] = {D(L, LES_U), D(L, LES_V), D(L, LES_W)};
  Cplx<double>* dst[1] = {s0};
  rows_r2c<double, double>(L, src, dst, 1, 0, L.nzg, 1.0 / (double)(L.nx * L.ny), LOAD_PCLEAN, 1.0, 0, uvw);
  cols<double>(L, COP_GRAD, s0, nullptr, s1, s2, s0, 3, 1, 1, 0, L.nzg);
  const Cplx<double>* cs[3] = {s1, s2, s0};
  double* outs[3] = {L.tk_px, L.tk_py, L.tk_pf};
  rows_c2r<double, double>(L, cs, outs, 3, 0, L.nzg);
  TkeArgs<T> a;
  a.u = D(L, LES_U); a.v = D(L, LES_V); a.w = D(L, LES_W);
  a.dudx = F<T>(L, LES_DUDX); a.dudy = F<T>(L, LES_DUDY); a.dudz = F<T>(L, LES_DUDZ);
  a.dvdx = F<T>(L, LES_DVDX); a.dvdy = F<T>(L, LES_DVDY); a.dvdz = F<T>(L, LES_DVDZ);
  a.dwdx = F<T>(L, LES_DWDX); a.dwdy = F<T>(L, LES_DWDY); a.dwdz = F<T>(L, LES_DWDZ);
  a.txx = F<T>(L, LES_TXX); a.txy = F<T>(L, LES_TXY); a.txz = F<T>(L, LES_TXZ);
  a.tyy = F<T>(L, LES_TYY); a.tyz = F<T>(L, LES_TYZ); a.tzz = F<T>(L, LES_TZZ);
  a.cs = F<T>(L, LES_CS_OPT2);
  a.pf = L.tk_pf; a.px = L.tk_px; a.py = L.tk_py;
  a.part = L.tk_part; a.np = L.np; a.nz = L.nz; a.cz = 1.0 / L.dz;
  const double rd = (double)L.np * L.nz * (4.0 * 8 + 19.0 * sizeof(T) + 3.0 * 8);
  kt_begin(L, KT_POINTWISE);
  hipLaunchKernelGGL(k_tke_sums<T>, dim3(LES_TKE_SLICES, L.nz), dim3(LES_BLOCK), 0, L.stream, a);
  kt_end(L, KT_POINTWISE, rd, "k_tke_sums", sizeof(T));
  hipLaunchKernelGGL(k_tke_add, dim3(nblk((long)TK_N * L.nz)), dim3(LES_BLOCK), 0, L.stream, (const double*)L.tk_part,
                     L.tk_sum, L.nz, (double)(L.nx * L.ny));
}

static void stat_zero(Les& L) {
  const long n = (long)LES_NSTAT * L.np * L.nz;
  hipLaunchKernelGGL(k_fill_zero, dim3(nblk(n)), dim3(LES_BLOCK), 0, L.stream, L.st_sum, n);
  hipLaunchKernelGGL(k_fill_zero, dim3(nblk(L.np)), dim3(LES_BLOCK), 0, L.stream, L.st_ustar, L.np);
  L.st_iters = 0;
}

extern "C" {

int les_stat_init(void) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->st_sum) {
    L->st_sum = (double*)dalloc(*L, sizeof(double) * LES_NSTAT * (size_t)L->np * L->nz);
    L->st_ustar = (double*)dalloc(*L, sizeof(double) * L->np);
    L->st_part = (double*)dalloc(*L, sizeof(double) * (size_t)L->nz * LES_PROF_SLICES * PS_N);
    L->st_prof = (double*)dalloc(*L, sizeof(double) * (size_t)LES_NPROF * L->nz);
    if (!L->st_sum || !L->st_ustar || !L->st_part || !L->st_prof) {
      les_set_error("Out of memory when initialising io module variables");
      return 3;
    }
  }
  stat_zero(*L);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_set_geostrophic(double ug, double vg) {
  Les* L = g_les;
  if (!L) return 1;
  L->P.ug = ug;
  L->P.vg = vg;
  return 0;
}

int les_tke_init(void) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->tk_sum) {
    L->tk_sum = (double*)dalloc(*L, sizeof(double) * LES_NTKE * (size_t)L->nz);
    L->tk_part = (double*)dalloc(*L, sizeof(double) * (size_t)L->nz * LES_TKE_SLICES * LES_NTKE);
    L->tk_pf = (double*)dalloc(*L, sizeof(double) * L->nf);
    L->tk_px = (double*)dalloc(*L, sizeof(double) * L->nf);
    L->tk_py = (double*)dalloc(*L, sizeof(double) * L->nf);
    for (int i = 0; i < 3; ++i) L->tk_spec[i] = (double*)dalloc(*L, 2 * sizeof(double) * L->nsp);
    if (!L->tk_sum || !L->tk_part || !L->tk_pf || !L->tk_px || !L->tk_py || !L->tk_spec[0] || !L->tk_spec[1] ||
        !L->tk_spec[2]) {
      les_set_error("Out of memory when initialising the tke budget");
      return 3;
    }
  }
  return les_tke_reset();
}

int les_tke_reset(void) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->tk_sum) { les_set_error("tke budget not initialised (les_tke_init)"); return 1; }
  hipLaunchKernelGGL(k_fill_zero, dim3(nblk((long)LES_NTKE * L->nz)), dim3(LES_BLOCK), 0, L->stream, L->tk_sum,
                     (long)LES_NTKE * L->nz);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_tke_accumulate(void) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->tk_sum) { les_set_error("tke budget not initialised (les_tke_init)"); return 1; }
  if (L->mixed) tke_accumulate<float>(*L); else tke_accumulate<double>(*L);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_tke_download(double* sums) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->tk_sum) { les_set_error("tke budget not initialised (les_tke_init)"); return 1; }
  if (!sums) { les_set_error("les_tke_download: null output"); return 10; }
  LES_CHECK(hipStreamSynchronize(L->stream));
  LES_CHECK(hipMemcpy(sums, L->tk_sum, sizeof(double) * LES_NTKE * (size_t)L->nz, hipMemcpyDeviceToHost));
  return 0;
}

int les_stat_reset(void) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->st_sum) { les_set_error("statistics not initialised (les_stat_init)"); return 1; }
  stat_zero(*L);
  LES_CHECK(hipGetLastError());
  return 0;
}

int les_stat_accumulate(void) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->st_sum) { les_set_error("statistics not initialised (les_stat_init)"); return 1; }
  if (L->mixed) stat_accumulate<float>(*L); else stat_accumulate<double>(*L);
  LES_CHECK(hipGetLastError());
  L->st_iters += 1;
  return 0;
}

int les_stat_download(double* sums3d, double* ustar2d, int* iterations) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->st_sum) { les_set_error("statistics not initialised (les_stat_init)"); return 1; }
  LES_CHECK(hipStreamSynchronize(L->stream));
  if (sums3d)
    LES_CHECK(hipMemcpy(sums3d, L->st_sum, sizeof(double) * LES_NSTAT * (size_t)L->np * L->nz, hipMemcpyDeviceToHost));
  if (ustar2d) LES_CHECK(hipMemcpy(ustar2d, L->st_ustar, sizeof(double) * L->np, hipMemcpyDeviceToHost));
  if (iterations) *iterations = L->st_iters;
  return 0;
}

int les_stat_profiles(double* prof) {
  Les* L = g_les;
  if (!L) return 1;
  if (!L->st_sum) { les_set_error("statistics not initialised (les_stat_init)"); return 1; }
  if (L->st_iters < 1) { les_set_error("no accumulated statistics"); return 1; }
  StatPlaneArgs a;
  a.s = L->st_sum; a.part = L->st_part; a.np = L->np; a.n = L->np * L->nz;
  a.nx = L->nx; a.nz = L->nz; a.nt = (double)L->st_iters;
  hipLaunchKernelGGL(k_stat_planes, dim3(LES_PROF_SLICES, L->nz), dim3(LES_BLOCK), 0, L->stream, a);
  hipLaunchKernelGGL(k_stat_profiles, dim3(nblk(L->nz, 64)), dim3(64), 0, L->stream, (const double*)L->st_part,
                     L->st_prof, L->np, L->nz);
  LES_CHECK(hipGetLastError());
  LES_CHECK(hipMemcpyAsync(prof, L->st_prof, sizeof(double) * LES_NPROF * (size_t)L->nz, hipMemcpyDeviceToHost,
                           L->stream));
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

// les_stat_profiles with its partial sums and profiles in this device's
// `scratch` (address from les_init) instead of les_stat_init's buffers: the
// same kernels in the same order, so the same values bit for bit.  Nothing
// is allocated: the launches read the running sums and write the scratch,
// and one asynchronous copy brings the profiles to `prof`.  The scratch
// needs no clearing: k_stat_planes writes every partial k_stat_profiles
// reads, and k_stat_profiles every profile value copied out.
int les_stat_profiles_noalloc(double* prof) {
  Les* L = g_les;
  if (!L) return 1;
  if (L->nz > LES_NOALLOC_MAX_NZ) {
    les_set_error("les_stat_profiles_noalloc: nz = %d exceeds the scratch limit of %d levels", L->nz,
                  LES_NOALLOC_MAX_NZ);
    return 10;
  }
  if (!prof) { les_set_error("les_stat_profiles_noalloc: null output"); return 10; }
  if (!L->st_sum) { les_set_error("statistics not initialised (les_stat_init)"); return 1; }
  if (L->st_iters < 1) { les_set_error("no accumulated statistics"); return 1; }
  double* ptr = g_scratch[L->dev];
  if (!ptr) { les_set_error("les_stat_profiles_noalloc: no scratch address for device %d", L->dev); return 5; }
  StatPlaneArgs a;
  a.s = L->st_sum; a.part = ptr; a.np = L->np; a.n = L->np * L->nz;
  a.nx = L->nx; a.nz = L->nz; a.nt = (double)L->st_iters;
  double* dprof = ptr + LES_NOALLOC_PART;
  hipLaunchKernelGGL(k_stat_planes, dim3(LES_PROF_SLICES, L->nz), dim3(LES_BLOCK), 0, L->stream, a);
  hipLaunchKernelGGL(k_stat_profiles, dim3(nblk(L->nz, 64)), dim3(64), 0, L->stream, (const double*)ptr, dprof,
                     L->np, L->nz);
  LES_CHECK(hipGetLastError());
  LES_CHECK(hipMemcpyAsync(prof, dprof, sizeof(double) * LES_NPROF * (size_t)L->nz, hipMemcpyDeviceToHost,
                           L->stream));
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

int les_running_diagnostics(double* out) {
  Les* L = g_les;
  if (!L) return 1;
  if (L->mixed) diagnostics<float>(*L); else diagnostics<double>(*L);
  LES_CHECK(hipGetLastError());
  LES_CHECK(hipMemcpyAsync(out, L->dg_out, sizeof(double) * 4, hipMemcpyDeviceToHost, L->stream));
  LES_CHECK(hipStreamSynchronize(L->stream));
  return 0;
}

}  // extern "C"

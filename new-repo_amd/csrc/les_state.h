// Device state of the LES library: one global instance (the reference keeps
// the same data as Fortran module globals, src/system.f90:21-47).
#pragma once
#include <vector>

#include "../../include/les_gpu.h"
#include "fft_kernels.h"

// Number of spectral workspaces (the Germano filtering uses slots 0..8).
#define LES_NSPEC 12

// Kernel families timed with HIP events when timing is enabled.
enum {
  KT_ROWS_R2C = 0, KT_ROWS_C2R, KT_COLS, KT_ROWS_RES, KT_COLS_RES, KT_POINTWISE, KT_TRIDIAG, KT_LAGR, KT_MLP, KT_COUNT
};

#define LES_MAX_KSTAT 96

struct Les {
  les_params P;
  int nx, ny, nz, nzg, lhx, nx2, ny2;
  long np, nf, ns, nsp;  // nx*ny, nx*ny*nzg, lhx*ny, lhx*ny*nzg
  long nbp;              // nx2*ny2
  double dx, dy, dz, delta;
  int jt_total;
  int jt;                // steps since les_init (the reference's local count)
  int F_coeff_init;
  hipStream_t stream;
  hipStream_t stream2;   // second stream of the step: the SGS stage beside the convection
  hipEvent_t ev_fork, ev_join;
  hipStream_t stream3;   // learned mode: the surrogate's chain beside the band's plain update
  hipEvent_t ev_feat0, ev_feat1;
  int overlap;           // les_step on two streams (les_set_overlap)
  int mixed;             // flow kernels in FP32

  // host-built FFT plans and twiddles
  FftPlan px, py, px2, py2;
  Cplx<double>* tw64;
  Cplx<float>* tw32;

  // tables (double and float copies)
  double *kkx64, *kky64, *g2_64, *g4_64;
  float *kkx32, *kky32, *g2_32, *g4_32;
  // TRUNC / FSMTHG multipliers of the convection sources (a) and products (b)
  double *ga_64, *gb_64;
  float *ga_32, *gb_32;
  double* logterm;  // log(0.5 dz / z0) per surface point
  double* lvec;     // mixing length per level (nz+2)

  // fields: index by les_field; elem size per field
  void* f[LES_NFIELDS];
  int fbytes[LES_NFIELDS];
  double* layer[LES_NLAYERS];
  double* z0;

  // workspaces
  void* wk_spec[LES_NSPEC]; // spectral (lhx, ny, nzg) of flow precision
  void* wk_real[64];      // real (nx, ny, nzg) of flow precision
  void* wk_hyb[6];        // (nx2, ny, nzg)
  void* wk_big[6];        // (nx2, ny2, nzg)
  void* wk_prod[3];       // (nx2, ny2, nzg)
  void* wk_hyb3[3];       // (nx2, ny, nzg)
  double* pspec[4];       // pressure spectra (double): Hx, Hy, Hz, phat
  double* tri_gam;        // Thomas scratch factors, (nz+3) rows of ns
  double* playr[4];       // bottom/top layer spectra and scratch
  void* wk_layer[6];      // (nx, ny) planes of flow precision and spectra
  double* Fold[4];        // previous F_* for the Lagrangian gather
  double* nyq_rows;       // mixed-mode FP64 Nyquist removal sums (A, B, C of u, v, w)
  double* nyq_part;       // their band partial sums
  int* nanflag;
  int alloc_failed;

  // learned SGS surrogate (allocated when weights or samples are first used)
  struct MlpState* mlp;
  double judge[4];        // rel. L2 of Cs^2 and beta vs the plain procedure, -, step
  int nband;              // levels 1..nband (z / delta < learned_zmin) keep the plain procedure
  int ncu;                // compute units of the device (persistent grids)
  int dev;                // the HIP device current at les_init
  int gradp_lazy;         // dpdx, dpdy, dpdz not stored by the last step (formed on request)

  // running statistics (allocated by les_stat_init) and diagnostics
  double* st_sum;         // LES_NSTAT sums of nx*ny*nz
  double* st_ustar;       // sum of u* (nx*ny)
  double* st_part;        // profile partial sums
  double* st_prof;        // LES_NPROF x nz profiles
  int st_iters;           // accumulated steps
  // TKE budget (allocated by les_tke_init)
  double* tk_sum;         // LES_NTKE x nz window sums
  double* tk_part;        // plane partial sums
  double* tk_pf;          // filtered cleaned pressure (nx, ny, nzg)
  double* tk_px;          // its x derivative
  double* tk_py;          // its y derivative
  double* tk_spec[3];     // FP64 spectra of the chain
  double* dg_part;        // diagnostics partials
  double* dg_out;         // cfl, mke, u*, rms div

  // timing
  int timing;
  hipEvent_t ev_a[KT_COUNT], ev_b[KT_COUNT];
  double kt_ms[KT_COUNT];
  double kt_bytes[KT_COUNT];
  long kt_calls[KT_COUNT];
  hipEvent_t ph_ev[6];
  double phase_ms[5];
  // per kernel (one entry per kernel instance and launch site name)
  struct KStat {
    const char* base;     // kernel name (string literal)
    int ta, tb;           // element sizes of its template precisions (0: none)
    const char* tag;      // variant (string literal or null)
    int fam;
    double ms, bytes;
    long calls;
  } kst[LES_MAX_KSTAT];
  int nkst;
  std::vector<void*> allocs;
};

extern Les* g_les;

// helpers implemented in les_gpu.hip
template <typename T>
const Cplx<T>* twpool(const Les& L);

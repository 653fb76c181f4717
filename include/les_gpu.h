/*
 * C-ABI of the MI355X LES time-step library (liblesgpu.so).
 *
 * Drop-in boundary for the reference's LES_engine / compute_* Fortran module
 * API (fmargairaz/new-repo).  The reference's procedures take Fortran arrays
 * by reference and keep module state; here the whole domain is device
 * resident, fields are named by `les_field` ids, and every entry point returns
 * an int status that mirrors the reference's decomp_2d_abort codes:
 *   0 ok, 1 invalid option/model, 3 out of memory, 4 double initialisation,
 *   5 HIP runtime error, 9 NaN found / singular solve, 10 bad shape or id.
 * Calls come from one host thread; every entry point runs its kernels on the
 * library's HIP stream (les_step also uses a second one inside a step, joined
 * before it returns, see les_set_overlap); host buffers are synchronised only
 * by download/status calls.  Not reentrant,
 * like the reference's module state.
 *
 * Sizes: nx, ny multiples of 4 of the form 2^a 3^b with 3 nx / 2, 3 ny / 2 <= 1024
 * (every transform is held by one wavefront); nz >= 3.
 *
 * Layout of every 3D field: (nx, ny, nz+2) float64 on the host, Fortran
 * order (x fastest), z levels 0..nz+1 -- the reference's
 * dimension(nx, ny, 0:nz+1) arrays with their ghost planes.  Fields the
 * reference dimensions 1:nz+1 (cs_opt2, nu_t, F_*, Beta_mum, S_ij) use the
 * same shape with level 0 unused.
 */
#ifndef LES_GPU_H
#define LES_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters after the reference's nondimensionalisation
 * (src/parameters_IO.f90:240-298); replaces the namelist globals. */
typedef struct les_params {
  int nx, ny, nz;
  int sgs_model;   /* 0 cst_Smag, 1 dyn_Smag, 2 dyn_LagScaleDep, 3 dns (src/compute_sgs.f90:124-175) */
  int cs_count;    /* LASD update period (src/compute_LagScaleDep.f90:237) */
  int dyn_init;    /* first LASD step */
  int lagran_init; /* restart with F_* present (src/parameters_IO.f90:77) */
  int wall_model;  /* 0 wall_filter, 1 wall_inst (src/compute_wall_law.f90:119-151) */
  int ubc_u_neu, ubc_v_neu; /* top BC type of u, v: 1 Neumann, 0 Dirichlet */
  int halo_mode;   /* Lagrangian stencil halo: 0 one-rank reference semantics, 1 periodic */
  int coriolis;    /* coriolis_flag */
  int precision;   /* 0 FP64, 1 mixed (FP32 flow + FP64 pressure/averages/AB2), 2 mixed + learned SGS */
  int judge_every; /* learned SGS: steps between plain-LASD judge updates */
  int dealias;     /* convection dealiasing, the reference's compile-time DSLG
                      (src/Makefile:17-40): 0 PADD, 1 TRUNC, 2 FSMTHG, 3 NODLSG */
  double lx, ly, lz, dt, z0, pfx, pfy, cs0, coriol, ug, vg, ubc_u, ubc_v, ubc_w;
  double re;       /* Reynolds number (physical_param Re), used by sgs_model dns */
  double learned_zmin; /* learned SGS: levels with z / delta < learned_zmin (the
                          near-wall band) keep the plain procedure on every update
                          (16 in the parameter modules: the band whose long-run
                          mean profiles stay inside the FP64 noise floor) */
} les_params;

/* Field ids (system_variables, compute_sgs, compute_LagScaleDep globals). */
enum les_field {
  LES_U = 0, LES_V, LES_W,
  LES_DUDX, LES_DUDY, LES_DUDZ, LES_DVDX, LES_DVDY, LES_DVDZ, LES_DWDX, LES_DWDY, LES_DWDZ,
  LES_P, LES_DPDX, LES_DPDY, LES_DPDZ,
  LES_RHSX, LES_RHSY, LES_RHSZ, LES_RHSX_F, LES_RHSY_F, LES_RHSZ_F,
  LES_TXX, LES_TXY, LES_TXZ, LES_TYY, LES_TYZ, LES_TZZ,
  LES_CS_OPT2, LES_NU_T, LES_F_LM, LES_F_MM, LES_F_QN, LES_F_NN, LES_BETA,
  LES_U_LAG, LES_V_LAG, LES_W_LAG,
  LES_S11, LES_S12, LES_S13, LES_S22, LES_S23, LES_S33,
  LES_NFIELDS
};

/* Surface (nx, ny) arrays. */
enum les_layer { LES_LBC_U = 0, LES_LBC_V, LES_USTAR, LES_BOTTOM_WK, LES_TOP_WK, LES_Z0, LES_NLAYERS };

/* --- set-up / tear-down -------------------------------------------------- */
/* Replaces create_system + fft_init + decomp_2d_fft_2d_init + deriv_init +
 * conv_init + press_init + wall_init + test_filter_init/create + sgs_init
 * (src/LES_engine.f90:125-153; src/system.f90:66-203). */
int les_init(const les_params* p);
/* clean_system + *_finalize (src/LES_engine.f90:413-422). */
int les_finalize(void);
const char* les_last_error(void);

/* set_initial_conditions(ic_type='files') + BCs (src/initial_conditions.f90:130-208):
 * u, v, w are (nx, ny, nz) float64 interiors. */
int les_set_ic(const double* u, const double* v, const double* w);

/* Host mirrors: whole (nx, ny, nz+2) field / (nx, ny) layer. */
int les_upload_field(int id, const double* host);
int les_download_field(int id, double* host);
int les_upload_layer(int id, const double* host);
int les_download_layer(int id, double* host);
/* Device pointer and element size (4 or 8) of a field, for zero-copy use. */
int les_field_ptr(int id, void** dptr, int* elem_bytes);

/* --- the time step -------------------------------------------------------- */
/* nsteps iterations of the body of LES_main_engine's time loop
 * (src/LES_engine.f90:189-353). */
int les_step(int nsteps);
int les_sync(void);
int les_jt_total(void);
int les_set_jt_total(int jt_total);
/* check_NaN (src/LES_engine.f90:583-616): *found = 1 if u, v or w holds a NaN;
 * returns 9 in that case. */
int les_check_nan(int* found);
/* Per-phase device times in ms since les_init or the last reset:
 * t[0] time_grad, t[1] time_stress, t[2] time_rhs, t[3] time_press,
 * t[4] time_loop (src/LES_engine.f90:100-107). */
int les_timers(double* t, int reset);
/* Enable (1) or disable (0) per-kernel-family HIP-event timing; while on,
 * les_timers also returns, for kernel family f (0 row r2c, 1 row c2r,
 * 2 column, 3 row resample, 4 column resample, 5 point-wise, 6 tridiagonal,
 * 7 Lagrangian, 8 MLP): t[5 + 3f] summed ms, t[6 + 3f] launches and
 * t[7 + 3f] algorithmic bytes (inputs read once + outputs written once).
 * t must hold 32 doubles.  Timing synchronises after every timed launch. */
int les_set_timing(int on);
/* Per-kernel view of the same timing mode (no counterpart in the reference,
 * whose profiling.txt has the five phases only): *count receives the number
 * of kernels timed since the last les_timers reset; for 0 <= i < *count,
 * name receives the kernel's name (template precisions and variant, e.g.
 * "k_cols_c16<float> [GRAD]"; memset and copy groups are named after the
 * runtime call), v[0] its summed ms, v[1] launches, v[2] algorithmic bytes,
 * v[3] its family.  i < 0 with name and v NULL only sets *count. */
int les_kernel_stat(int i, int* count, char* name, int name_cap, double* v);
/* Schedule of les_step (no counterpart in the reference, whose step is one
 * sequence): 1 (default) runs the SGS stage and the spectra of the stress
 * divergence on a second stream beside the padded convection, and in mixed
 * precision the wall law beside the horizontal gradients; 0 issues the whole
 * step on one stream.  Results are the same either way.  The per-kernel
 * timing mode always uses one stream. */
int les_set_overlap(int on);

/* --- operator-level entry points (the compute_* module procedures) -------- */
/* obukhov + wallstress (src/compute_wall_law.f90:107-254). */
int les_wall_law(void);
/* compute_vel_grad (src/LES_engine.f90:449-498). */
int les_compute_vel_grad(void);
/* grad_velocity_filt(f, dfdx, dfdy) (src/compute_derivatives.f90:94-137). */
int les_grad_velocity_filt(int f, int dfdx, int dfdy);
/* ddz_uv(dfdz, f) / ddz_w(dfdz, f) (src/compute_derivatives.f90:375-441). */
int les_ddz_uv(int dfdz, int f);
int les_ddz_w(int dfdz, int f);
/* sgs_stress(jt, jtt) (src/compute_sgs.f90:105-247), includes
 * compute_DynSmag_coeff (src/compute_DynSmag.f90:61-245; updates when jt is a
 * multiple of cs_count) and compute_LagScaleDep_coeff
 * (src/compute_LagScaleDep.f90:204-298; on jtt).  les_step passes the number
 * of steps since les_init as jt. */
int les_sgs_stress(int jt, int jtt);
/* compute_dealias_convec(cx, cy, cz) (src/compute_convective_padd.f90:115-262). */
int les_compute_dealias_convec(int cx, int cy, int cz);
/* div_stress(rhs, tx, ty, tz, stag[, bottom_wk_press, top_wk_press])
 * (src/compute_derivatives.f90:464-577). */
int les_div_stress(int rhs, int tx, int ty, int tz, int stag, int save_wk);
/* compute_RHS_adv (src/LES_engine.f90:509-572). */
int les_compute_rhs_adv(void);
/* press_stag_array(tadv1) (src/compute_pressure.f90:116-242). */
int les_press_stag_array(double tadv1);
/* test_filter(f, G) on levels 1..nz+1, alpha = 2 or 4 (src/testfilter.f90:144-184). */
int les_test_filter(int f, int alpha);
/* Advection step u += dt (tadv1 rhs + tadv2 rhs_f) on levels 1..nz and
 * w = 0 on the ghost levels (src/LES_engine.f90:253-283). */
int les_ab2(double tadv1, double tadv2);
/* Projection u -= dt tadv1 grad p, rhs -= grad p, dpdz, w(1) = w(nz+1) = 0
 * and the bottom/top set_bc (src/LES_engine.f90:299-352;
 * src/compute_pressure.f90:353-369). */
int les_projection(double tadv1);

/* execute_decomp_2d_fft_2d (src/decomp_2d_fft_2d.f90:498-577) on device
 * buffers of the library's compute precision: r2c scales the input by
 * `scale`; nplanes planes of (nx, ny) real <-> (nx/2+1, ny) complex. */
int les_fft2d_r2c(const void* d_in, void* d_out, int nplanes, double scale);
int les_fft2d_c2r(const void* d_in, void* d_out, int nplanes);

/* --- running statistics and diagnostics (post-step, outside the step) ----- */
#define LES_NSTAT 26
#define LES_NPROF 21
/* io_stat_init (src/IO_statistics.f90:40-244): allocates the running sums on
 * the first call, then zeroes them and the step count. */
int les_stat_init(void);
/* One call of io_stat_averages' accumulation (src/IO_statistics.f90:266-293,
 * 300-324, 363-440): FP64 running sums over levels 1..nz of, in the order of
 * RAV_<jt>_3D.dat, u v w p u2 v2 w2 u3 v3 w3 u4 v4 w4 uv uw vw txx tyy tzz
 * txy txz tyz dudz dvdz nu_t sqrt(cs_opt2) (uw, vw with u, v averaged to the
 * w node, zero at level 1), and of u* (RAV_<jt>_2D.dat). */
int les_stat_accumulate(void);
/* Host copies: sums3d = LES_NSTAT x (nx, ny, nz) float64 in the file's order
 * (may be NULL), ustar2d = (nx, ny) (may be NULL), *iterations = the number
 * of accumulated steps (iterations_av, RAV_<jt>_info.txt). */
int les_stat_download(double* sums3d, double* ustar2d, int* iterations);
/* Zeroes the sums and the step count, as after a RAV write
 * (src/IO_statistics.f90:286,351,501-509). */
int les_stat_reset(void);
/* Plane-mean profiles of the time averages (sum / steps, SGS stresses
 * negated), get_data_struct_RAV + get_profile_struct + compute_Cij
 * (matlab_src/get_data_struct_RAV.m:21-35; matlab_src/get_profile_struct.m:9-59;
 * matlab_src/compute_Cij.m:1-36).  prof = LES_NPROF rows of nz levels:
 * u v w p var_u var_v var_w txx tyy tzz R13 txz St13 R23 tyz St23 tke cs nut
 * C13 C23 (R_i3 = <u_i w> - u_i(w node) w; St_i3 = R_i3 + tau_i3 + C_i3;
 * tke = (var_u + var_v + var_w on the uvp node) / 2).  Error 1 before any
 * accumulation. */
int les_stat_profiles(double* prof);
/* Allocation-free variant of les_stat_profiles.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and summation order as les_stat_profiles,
 * hence the same profiles bit for bit, but the partial sums and the profiles
 * live in a fixed device array of the library (3.19 MB on every GPU it runs
 * on, held until the process ends) instead of les_stat_init's buffers; the
 * running sums themselves still come from les_stat_init.  nz at most 1024,
 * else status 10.  Calls on one GPU share that array: calls that overlap
 * (two host threads, or two processes' streams) race on it. */
int les_stat_profiles_noalloc(double* prof);
/* The geostrophic wind (ug, vg), nondimensional, of the Coriolis forcing
 * from the next operator or step on: the reference's module variables that
 * compute_geostrophic_correct rotates every step from geostr_start on when
 * geostr_rot is set (src/compute_geostrophic.f90:24-45; called at
 * src/LES_engine.f90:198-201).  les_init takes them from les_params;
 * les_step(n) keeps them for its n steps, so a rotating caller steps one
 * at a time (as the Fortran host does). */
int les_set_geostrophic(double ug, double vg);
/* TKE budget (src/tke_budget1D.f90:64-440): LES_NTKE per-level sums of
 * plane averages on the uvp nodes of levels 1..nz, in the order of the
 * reference's ta1_profiles/<name>.out files:
 *   u v w dudx dudy dudz dvdx dvdy dvdz dwdx dwdy dwdz uu vv ww uv uw vw
 *   txx tyy tzz txy txz tyz cs uuw vvw www uww wtzz utxz vtyz dxx dyy dzz
 *   dxy dxz dyz dudx2 dudy2 dudz2 dvdx2 dvdy2 dvdz2 dwdx2 dwdy2 dwdz2
 *   dudxdwdx dudydwdy dudzdwdz p dpdx dpdy dpdz pw pu pdudx pdvdy pdwdz
 *   pdwdx pdudz
 * (w-node fields interpolated to the uvp node; the p family is of the
 * cleaned pressure p - |u|^2 / 2, through grad_velocity_filt).
 * les_tke_init allocates the sums and the cleaned-pressure workspaces
 * (tke_budget_init, :64-195) and zeroes the sums. */
#define LES_NTKE 61
int les_tke_init(void);
/* One call of compute_tke_budget (:209-390): the plane averages of the
 * current state added to the sums.  Error 1 before les_tke_init. */
int les_tke_accumulate(void);
/* Host copy of the sums: LES_NTKE rows of nz levels (row-major by
 * quantity).  The reference writes sums / tke_period every tke_period
 * steps (avg_var, :415-436). */
int les_tke_download(double* sums);
/* Zeroes the sums, as after a record (:437). */
int les_tke_reset(void);
/* compute_running_diagnistics (src/running_diagnostics.f90:80-305) of the
 * current state: out[0] CFL, out[1] MKE, out[2] plane-mean u*, out[3] rms of
 * du/dx + dv/dy + dw/dz over levels 1..nz. */
int les_running_diagnostics(double* out);

/* --- learned SGS surrogate (no counterpart in the reference) -------------- */
/* A per-point MLP 12 -> 64 -> 64 -> 64 -> 2 (GELU, BF16 MFMA, FP32
 * accumulation) predicting the Lagrangian dynamic model's Cs^2 and beta from
 * strain/rotation invariants at delta, 2 delta, 4 delta and log(z/delta).
 * With precision = 2 the update steps of the dynamic procedure use it above
 * the near-wall band (z / delta >= learned_zmin, the training range); the band
 * keeps the plain procedure on every update, and every judge_every steps the
 * plain procedure runs on the whole column and judges the surrogate.
 *
 * Training samples at the current state (call right after a step with a
 * plain dynamic update, mixed-precision modes): every stride-th point of the
 * levels above the near-wall band, at most max_samples, raw features to device d_x
 * (12 float32 per sample) and targets (Cs^2, beta) to device d_y. */
int les_sgs_samples(void* d_x, void* d_y, long max_samples, long* n_written, int stride);
/* Number of float32 values of the compact weight layout: W1[64][12] b1[64]
 * W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2], then the feature
 * means[12] and scales[12], target means[2] and scales[2], and the
 * out-of-range limit (in scale units). */
int les_mlp_layout(long* nweights);
int les_mlp_set_weights(const float* w, long nweights);
int les_mlp_get_weights(float* w);
/* Raw samples -> normalised training buffers (device): d_x bf16 [n][16],
 * d_y float32 [n][16] (targets in columns 0, 1). */
int les_mlp_prepare(const void* d_raw, const void* d_tgt, long n, void* d_x, void* d_y);
/* Gradient of 0.5 * mean |y - t|^2 over a batch (n multiple of 16) in the
 * compact layout without the normalisers (MLP weights and biases only). */
int les_mlp_grad(const void* d_x, const void* d_y, long n, float* grad, double* loss);
/* One Adam step (beta1 0.9, beta2 0.999) on a batch; FP32 master weights,
 * BF16 copies for the MFMA kernels. */
int les_mlp_train_step(const void* d_x, const void* d_y, long n, double lr, double* loss);
/* Fused inference on normalised features (device bf16 [n][16]) ->
 * normalised outputs (device float32 [n][2]). */
int les_mlp_forward(const void* d_x, long n, void* d_out);
/* Judge statistics of the last judge step over the levels above the
 * near-wall band: [0] rel. L2 of Cs^2 (surrogate vs plain procedure), [1]
 * rel. L2 of beta (both compared up to 8), [2] number of band levels,
 * [3] the step. */
int les_mlp_judge_stats(double* out);

#ifdef __cplusplus
}
#endif
#endif

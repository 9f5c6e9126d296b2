/*
 * libcaps_amd: the CaPS plane-sphere round-trip-matrix library, MI355X
 * (gfx950) edition.
 *
 * C ABI of libcaps_amd.so.  Every entry point keeps the name, argument
 * meaning, ownership rule and error behaviour of the CaPS v0.5 function it
 * replaces (cited per declaration, reference file:line), so that existing
 * callers (the block CLI caps_logdetD, the MPI driver caps, the AuAl
 * example, cass, capc) link against this library unchanged.  The numerical
 * work behind caps_logdetD, caps_M_init and kernel_logdet runs in FP64 HIP
 * kernels on the GPU; the library fails loudly (abort with a message) when
 * no gfx950 device is usable.
 *
 * Additions that have no counterpart in the reference are marked "extension"
 * and carry the prefix caps_ or capsx_.
 */
#ifndef LIBCAPS_AMD_H
#define LIBCAPS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#include <stdio.h>
#include <stdint.h>

/* Polarization of a k-integral.  Reference: src/include/libcaps.h:17. */
typedef enum { TE, TM } polarization_t;

/* Sign of a log-represented number.  Reference: src/include/constants.h:15. */
typedef signed char sign_t;

/* Determinant algorithm.  Reference: src/include/matrix.h:14.  Here
 * DETALG_HODLR (the default) and DETALG_CHOLESKY select the dense
 * determinant with the HODLR-mode shortcuts of kernel_logdet, DETALG_LU and
 * DETALG_QR the dense LU with the LU-mode shortcuts (Mercator series).  All
 * of them factor with the GPU's blocked LU with partial pivoting. */
typedef enum { DETALG_HODLR, DETALG_LU, DETALG_QR, DETALG_CHOLESKY } detalg_t;

/* Defaults.  Reference: src/include/libcaps.h:29-33. */
#define CAPS_MINIMUM_LDIM 20
#define CAPS_FACTOR_LDIM 5
#define CAPS_EPSREL 1e-8

/* Physical constants.  Reference: src/include/constants.h:45-48. */
#define CAPS_hbar    1.0545718e-34
#define CAPS_hbar_eV 6.582119514e-16
#define CAPS_kB      1.38064852e-23
#define CAPS_c       299792458.

/* The CaPS object.  Reference: src/include/libcaps.h:41-78 (same fields in
 * the same order; one private pointer appended for the device context, so
 * the offsets of all public fields are those of the reference).  Do not
 * modify the fields yourself. */
typedef struct caps
{
    double L;
    double R;
    double calL;
    double LbyR;
    double y;
    double (*epsilonm1_plate)(double xi_, void *userdata);
    void *userdata_plate;
    double (*epsilonm1_sphere)(double xi_, void *userdata);
    void *userdata_sphere;
    int ldim;
    double epsrel;
    detalg_t detalg;
    void *priv;      /* extension: device context (owned by the library) */
} caps_t;

/* Opaque per-block state (the reference's integration_t).  */
typedef struct caps_block_state integration_t;
typedef struct caps_block_state_plasma integration_plasma_t;

/* Round-trip matrix object of one (ξ, m) block.  Reference:
 * src/include/libcaps.h:101-109 (same fields and order).  al/bl hold
 * log|a_l|, log|b_l| for l = lmin .. lmin+ldim-1 once caps_M_init returns. */
typedef struct
{
    caps_t *caps;
    int m, lmin, ldim;
    integration_t *integration;
    integration_plasma_t *integration_plasma;
    double xi_;
    double *al, *bl;
} caps_M_t;

/* ---- object, setters, info: src/include/libcaps.h:113-140 ------------- */

/* Reference: src/libcaps.c:183-214.  NULL if R <= 0 or L <= 0. */
caps_t *caps_init(double R, double L);
/* Reference: src/libcaps.c:223-226. */
void caps_free(caps_t *self);
/* Reference: src/libcaps.c:239-262. */
void caps_build(FILE *stream, const char *prefix);
/* Reference: src/libcaps.c:273-294. */
void caps_info(caps_t *self, FILE *stream, const char *prefix);

/* Reference: src/libcaps.c:306-313, 322-325.  Return 1 on success, 0 on
 * invalid input (epsrel <= 0). */
int    caps_set_epsrel(caps_t *self, double epsrel);
double caps_get_epsrel(caps_t *self);
/* Reference: src/libcaps.c:436-443, 453-456 (0 if ldim < 1). */
int caps_set_ldim(caps_t *self, int ldim);
int caps_get_ldim(caps_t *self);
/* Reference: src/libcaps.c:401-411, 420-423. */
int      caps_set_detalg(caps_t *self, detalg_t detalg);
detalg_t caps_get_detalg(caps_t *self);

/* ε(iξ)-1 callbacks, evaluated on the host once per ξ and body.
 * Reference: src/libcaps.c:337-381; userdata is borrowed. */
void caps_set_epsilonm1(caps_t *self, double (*epsilonm1)(double xi_, void *userdata), void *userdata);
void caps_set_epsilonm1_plate(caps_t *self, double (*epsilonm1)(double xi_, void *userdata), void *userdata);
void caps_set_epsilonm1_sphere(caps_t *self, double (*epsilonm1)(double xi_, void *userdata), void *userdata);
/* Reference: src/libcaps.c:104-120. */
double caps_epsilonm1_plate(caps_t *self, double xi_);
double caps_epsilonm1_sphere(caps_t *self, double xi_);
/* Reference: src/libcaps.c:129-132 (+∞) and 150-156 (Drude, userdata =
 * {ωp, γ} in rad/s). */
double caps_epsilonm1_perf(double xi_, void *userdata);
double caps_epsilonm1_drude(double xi_, void *userdata);

/* Extension: tabulated dielectric functions (the file format of the
 * reference's material_init, src/material.c:63-180: "# omegap_low = <eV>",
 * "# gamma_low", "# omegap_high", "# gamma_high" headers and data lines
 * "xi[rad/s] epsilon(i xi)" with xi ascending).  ε(iξ) - 1 is interpolated
 * linearly, with Drude extrapolation below and above the table
 * (src/material.c:255-293); evaluated on the host once per ξ and body.
 * Use as caps_set_epsilonm1(caps, capsx_material_epsilonm1, material). */
typedef struct capsx_material capsx_material_t;
capsx_material_t *capsx_material_init(const char *filename, double calL);   /* NULL on error */
void capsx_material_free(capsx_material_t *material);
double capsx_material_epsilonm1(double xi, void *material);
void capsx_material_get_extrapolation(const capsx_material_t *material, double *omegap_low, double *gamma_low,
                                      double *omegap_high, double *gamma_high);

/* ---- scalar physics: src/include/libcaps.h:116,142-157 ---------------- */

/* log Λ_{l1 l2}^{(m)}.  Reference: src/libcaps.c:45-48. */
double caps_lnLambda(int l1, int l2, int m);
/* ℓ-window [lmin, lmax = lmin+ldim).  Reference: src/libcaps.c:63-88. */
int caps_estimate_lminmax(caps_t *self, int m, size_t *lmin_p, size_t *lmax_p);
/* log|a_l|, log|b_l| (GPU).  Reference: src/libcaps.c:485-520, 557-602. */
void caps_mie(caps_t *self, double xi_, int l, double *lna, double *lnb);
void caps_mie_perf(caps_t *self, double xi_, int l, double *lna, double *lnb);
/* Fresnel coefficients.  Reference: src/libcaps.c:616-645. */
void caps_fresnel(caps_t *self, double xi_, double k_, double *r_TE, double *r_TM);

/* ---- round-trip matrix and determinant: src/include/libcaps.h:145-150 -- */

/* Builds the whole block on the GPU (Mie, K, I, matrix) and keeps a host
 * copy.  Aborts if xi_ <= 0.  Reference: src/libcaps.c:673-697. */
caps_M_t *caps_M_init(caps_t *self, int m, double xi_);
/* Element M^{(m)}_{l1 l2}(p1,p2), p = 'E' or 'M', lmin <= l1,l2 < lmin+ldim.
 * Reference: src/libcaps.c:761-831. */
double caps_M_elem(caps_M_t *self, int l1, int l2, char p1, char p2);
/* Reference: src/libcaps.c:717-743 (interleaved 2x2 index map). */
double caps_kernel_M(int i, int j, void *args_);
/* Reference: src/libcaps.c:840-846. */
void caps_M_free(caps_M_t *self);

/* log det(1 - M^{(m)}(ξ)), GPU.  Aborts if xi_ <= 0.  Reference:
 * src/libcaps.c:950-962. */
double caps_logdetD(caps_t *self, double xi_, int m);

/* log det(1 - A) of an N x N kernel given as a host callback; the matrix
 * is filled on the host and factored on the GPU.  sym_spd is accepted for
 * compatibility.  Reference: src/matrix.c:57-149. */
double kernel_logdet(int dim, double (*kernel)(int, int, void *), void *args, int sym_spd, detalg_t detalg);

/* ξ = 0 Drude term.  Reference: src/libcaps.c:1043-1065. */
double caps_ht_drude(caps_t *caps);

/* ξ = 0 closed-form kernels (args: caps_M_t with caps, m, lmin).
 * Reference: src/libcaps.c:860-873, 925-941. */
double caps_kernel_M0_EE(int i, int j, void *args);
double caps_kernel_M0_MM(int i, int j, void *args);
/* ξ = 0 plasma-model MM kernel (args: caps_M_t with caps, m, lmin and
 * integration_plasma).  Reference: src/libcaps.c:882-906. */
double caps_kernel_M0_MM_plasma(int i, int j, void *args);
/* ξ = 0 plasma integrals: ω = ωp (L+R)/c, I(l1+l2) = (1/ν!) ∫_0^∞ -r_TE(t)
 * e^{-t} t^ν dt and the ratios I_{l-1/2}(α)/I_{l+1/2}(α), α = ω/(1 + L/R),
 * cached per object (epsrel is accepted for compatibility: the integral is
 * evaluated to full precision).  Reference: src/integration.c:897-1031
 * (src/include/integration.h:24-26). */
integration_plasma_t *caps_integrate_plasma_init(caps_t *caps, double omegap, double epsrel);
double caps_integrate_plasma(integration_plasma_t *self, int l1, int l2, int m, double *ratio1, double *ratio2);
void caps_integrate_plasma_free(integration_plasma_t *self);
/* The k-integrals of one block (ξ, m): caps_integrate_init computes the
 * block's whole K_ν and I_{l1 l2} tables on the GPU (ν from 2m, or 2 for
 * m = 0, to 2(lmin + ldim) + 1; l1, l2 over the block's ℓ-window and one
 * more on either side); K, I, A, B, C, D return log|value| and set *sign.
 * I is 0 (sign 0) for l1 < m or l2 < m; arguments outside the tables abort
 * with a message.  epsrel is accepted for compatibility (the device
 * quadrature has a fixed accuracy); NULL for xi_ <= 0, m < 0 or
 * epsrel <= 0.  Reference: src/include/integration.h:13-22,
 * src/integration.c:452-493, 627-888. */
integration_t *caps_integrate_init(caps_t *self, double xi_, int m, double epsrel);
void caps_integrate_free(integration_t *integration);
double caps_integrate_I(integration_t *self, int l1, int l2, polarization_t p, sign_t *sign);
double caps_integrate_K(integration_t *self, int nu, polarization_t p, sign_t *sign);
double caps_integrate_A(integration_t *self, int l1, int l2, polarization_t p, sign_t *sign);
double caps_integrate_B(integration_t *self, int l1, int l2, polarization_t p, sign_t *sign);
double caps_integrate_C(integration_t *self, int l1, int l2, polarization_t p, sign_t *sign);
double caps_integrate_D(integration_t *self, int l1, int l2, polarization_t p, sign_t *sign);
/* ξ = 0 blocks EE, MM and the plasma-model MM (omegap in rad/s), each
 * ldim x ldim, filled and factored on the GPU; NULL outputs are skipped.
 * Reference: src/libcaps.c:984-1017. */
void caps_logdetD0(caps_t *self, int m, double omegap, double *EE, double *MM, double *MM_plasma);
/* ξ = 0 perfect-reflector free energy.  Reference: src/libcaps.c:1089-1116. */
double caps_ht_perf(caps_t *caps, double eps);
/* ξ = 0 plasma-model free energy (omegap in rad/s), in units of k_B T.
 * Reference: src/libcaps.c:1127-1145. */
double caps_ht_plasma(caps_t *caps, double omegap, double eps);

/* ---- extensions ------------------------------------------------------- */

/* Extension: n blocks (xi_[i], m[i]) of one caps object in one batched GPU
 * launch sequence; results in out[i], identical to caps_logdetD(self,
 * xi_[i], m[i]).  Returns 0 on success, nonzero on invalid input (a
 * message goes to stderr).  Replaces the reference's one-block-per-MPI-task
 * loop (src/caps.c:163-232, 869-957). */
int caps_logdetD_batch(caps_t *self, const double *xi_, const int *m, int n, double *out);

/* Extension: n blocks, block i being (objs[i], xi_[i], m[i]), in one batched
 * launch sequence; the objects may differ in R/L, ldim and material, so the
 * blocks have uneven dimensions 2 ldim.  out[i] equals caps_logdetD(objs[i],
 * xi_[i], m[i]).  All objects must use the same determinant mode (HODLR
 * shortcut rules, or LU/QR shortcut rules); the batch runs on objs[0]'s
 * device context.  Returns 0 on success, nonzero on invalid input.  Serves
 * drivers that sweep several geometries at once (the reference's workers
 * build one caps object per geometry, src/caps.c:869-957, at 913). */
int capsx_logdetD_multi(caps_t *const *objs, const double *xi_, const int *m, int n, double *out);

/* Extension: LU schedule of the determinant stage.  levels 1 = 64-column
 * steps, 2 = 128-column steps of two panels with one K = 128 trailing
 * update, 3 = 256-column steps of four panels with one K = 256 update;
 * streams = groups of blocks factored on separate HIP streams; xcd =
 * trailing-update tile order (1: all tiles of a block on one XCD); tile =
 * trailing-update workgroup tile (128, 256 or 64 rows).  By default the library picks,
 * per batch, the nearest shape (largest dimension, number of blocks) of the
 * autotuned table lu_tuning.txt next to the library (or the file named by
 * CAPS_AMD_TUNING), written by caps_amd/autotune.py from capsx_lu_time.
 * No counterpart in the reference (which calls LAPACK dgetrf_,
 * src/matrix.c:555-577). */
/* Extension: capture each batch shape of this object into a HIP graph once
 * and replay it for later batches of the same shape (on != 0); 0 (default)
 * launches eagerly.  Descriptor uploads and the status reset stay outside
 * the graph.  Stage timings of replayed batches read 0 (the total is
 * measured).  No counterpart in the reference. */
void capsx_set_graph(caps_t *self, int on);
/* Fix the schedule for all later batches; levels = 0 returns to the table.
 * Returns 0 on success, 1 on invalid values. */
int capsx_lu_config(int levels, int streams, int xcd, int tile);
/* Schedule used for n blocks of largest dimension maxN: {levels, streams,
 * xcd, tile}. */
void capsx_lu_config_get(int maxN, int n, int cfg[4]);
/* Replace the table by the file at path; returns the entries read. */
int capsx_lu_tuning_load(const char *path);
/* Panels taller than rows are factored column by column by grid-wide
 * kernels (default 6144, above which one workgroup would hold more than 12
 * rows per lane); rows <= 0 only queries.  Returns the previous value. */
int capsx_lu_grid_panel_rows(int rows);
/* Kernel variants of the LU for all later batches (values < 0 keep the
 * current one; prev, if not NULL, receives the values before the call):
 * gemm_kernel = the 128 x 128 trailing update (2, the default: operand
 * stages moved straight into LDS, 8 deep, three workgroups per CU; 1: 16
 * deep, two per CU; 0: stages through registers), panel_set (1, the
 * default: leaves of 8 columns; 0: leaves of 4), narrow_tiles (1, the
 * default: 128 x 64 and 64 x 128 tiles for regions one panel wide or high;
 * 0: 128 x 128 for all).  All give bit-identical factors.  Same as the
 * CAPS_AMD_GEMM_KERNEL, CAPS_AMD_PANEL_SET and CAPS_AMD_NARROW_TILES
 * variables read at start-up.  Returns 1 for invalid values.  No
 * counterpart in the reference. */
int capsx_lu_kernels(int gemm_kernel, int panel_set, int narrow_tiles, int prev[3]);
/* Milliseconds of one batched LU of n synthetic dim x dim matrices with the
 * given schedule (values < 0 keep the configured ones), averaged over reps.
 * Returns 0 on success. */
int capsx_lu_time(int dim, int n, int reps, int levels, int streams, int xcd, int tile, double *ms);
/* Extension without a counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): the batched LU log|det A_b| of n column-major dim x dim matrices
 * already in device memory (dA, leading dimension lda, matrix b at
 * dA + b*lda*dim; overwritten), written to the device array d_logdet,
 * enqueued on `stream` (a hipStream_t; NULL: the null stream) without
 * waiting.  It allocates nothing: descriptors and the LU's side buffers
 * (about 16*lda*64 doubles per matrix) live in a fixed 64 MiB device array
 * per GPU; batches needing more are refused (3) before any launch.  Calls
 * that overlap on one GPU share the array and race.  Returns 0; 1 invalid
 * arguments, 2 no scratch address, 3 too large, 4 launch error. */
int capsx_lu_logdet_batch_noalloc(double *dA, int dim, int lda, int n, double *d_logdet, void *stream);

/* Extension: per-stage timing of the last batch on this caps object, in
 * milliseconds: [0] Mie, [1] K, [2] I, [3] assembly, [4] LU, [5] total
 * (HIP events on the library's stream). */
void caps_last_timing(caps_t *self, double ms[6]);

/* Extension: free energy of the plane-sphere system at temperature T (K),
 * with the sums of the reference driver (src/caps.c:235-396, 722-850).
 * Every frequency's m-sum stops at the first m with logdetD/logdetD(m=0) <
 * cutoff (that m included, m = 0 halved).
 *   T > 0: Matsubara sum, stopped at |v_n/v_0| < epsrel, n = 0 halved;
 *          drude_xi0 nonzero adds the analytic ξ = 0 Drude term
 *          (caps_ht_drude), as the reference driver does for finite ωp and
 *          γ > 0 (required: other ξ = 0 terms abort).  Returns T_scaled/π
 *          Σ'_n v_n, T_scaled = 2π k_B (L+R) T/(ħc).
 *   T = 0: ∫_0^∞ F_ξ(x/α) dx/(απ), α = 2L/(L+R), by the Fourier-Chebyshev
 *          rule on [0, ∞) (the reference's --fcqs option, src/fcqs.c:107)
 *          with levels 5, 10, ..., 2560 refined until two successive
 *          levels agree to epsrel; the integrand is 0 above the PFA
 *          frequency cutoff for R/L > 10 (src/caps.c:256-263).
 * mode 0 = plain (one block at a time, in the driver's order), mode 1 =
 * batched (all m of a frequency, or of a quadrature level, in one launch
 * sequence, verified against the same cutoff rule), mode 2 = batched with the
 * blocks planned by the surrogate (capsx_set_surrogate).  nblocks
 * (may be NULL) receives the number of determinants computed, nused (may be
 * NULL) the number that enter the sums.  Returns E·(L+R)/(ħc), the value
 * the reference driver prints. */
double caps_free_energy(caps_t *self, double T, double cutoff, double epsrel, int mode,
                        int drude_xi0, long *nblocks, long *nused);

/* Extension: the free energy with the reference driver's further options
 * (src/caps.c:270-396, 635-639, 752-850).  xi0 = ξ = 0 term: 1 Drude
 * (caps_ht_drude), 2 perfect reflector and 3 plasma model with omegap
 * (rad/s): the Drude term plus Σ'_m of the MM blocks of caps_logdetD0 up to
 * the m-cutoff.  psd_order = 0: Matsubara sum; > 0: Padé spectrum
 * decomposition of that order, F = T_scaled/π [v_0/2 + Σ_j η_j
 * F_ξ(ξ_j T_scaled/(2π))] (capsx_psd); < 0: the driver's order
 * ceil((1 - 1.5 log10 epsrel)/sqrt(T_eff)), T_eff = 4π k_B T L/(ħc).  mode
 * as for caps_free_energy.  T = 0: as caps_free_energy. */
double capsx_free_energy(caps_t *self, double T, double cutoff, double epsrel, int mode, int xi0,
                         double omegap, int psd_order, long *nblocks, long *nused);
/* Extension: the resume cache of the free-energy drivers, read from an
 * earlier run's output as the reference driver's --resume option does
 * (src/caps.c:72-101): every line with "# xi*(L+R)/c=" and "logdetD="
 * gives one (ξ(L+R)/c, F_ξ) pair, appended to this object's cache (path
 * NULL or "" clears it).  caps_free_energy and capsx_free_energy then take
 * F_ξ of a matching frequency (equal, or |1 - ξ_cached/ξ| < 1e-11,
 * src/caps.c:316-327) from the cache and compute no blocks for it.
 * Returns the number of pairs read; -1 the file cannot be opened, -2 a line
 * has the wrong format (the reference's two errors, src/caps.c:78, 90, 96). */
int capsx_resume_load(caps_t *self, const char *path);
/* Extension: F_ξ of frequency xi_ from the resume cache (the matching rule
 * above).  Returns 1 and sets *value if found, else 0. */
int capsx_resume_lookup(caps_t *self, double xi_, double *value);
/* Extension: the free-energy drivers append a line
 * "# xi*(L+R)/c=%.16g, logdetD=%.16g, t=%g" for every F_ξ they compute
 * (the reference's output line, src/caps.c:265, 807, 827; cached ones are
 * not repeated) to path, opened for appending: the file capsx_resume_load
 * reads back.  NULL or "" closes it.  Returns 0, -1 if it cannot be opened. */
int capsx_set_output(caps_t *self, const char *path);
/* Extension: checkpointing of the T = 0 free energy for runs longer than one
 * job (the reference's way is to kill the MPI job and rerun with --resume,
 * src/caps.c:74-105, losing only the frequencies in flight).  The driver
 * batches at most `nodes` quadrature nodes at a time (0: all new nodes of a
 * refinement level, the default), writes each F_ξ to the output file as its
 * batch completes, and with seconds > 0 stops before its next batch once
 * that many seconds have passed since the call began, returning NaN (every
 * call runs at least one batch); a later call after capsx_resume_load of that
 * file continues.  Returns 0, -1 for a
 * negative argument. */
int capsx_set_checkpoint(caps_t *self, int nodes, double seconds);
/* Extension: poles ξ_j (ascending) and residues η_j, j < N, of the Padé
 * spectrum decomposition of the Bose function (Hu, Xu, Yan, J. Chem. Phys.
 * 133, 101106 (2010)): 1/(1 - e^{-x}) ≈ 1/x + 1/2 + Σ_j 2η_j x/(x² + ξ_j²).
 * Returns 0 on success.  Reference: src/psd.c:80-136 (LAPACK dstemr there;
 * bisection on Sturm sequences here). */
int capsx_psd(int N, double *xi, double *eta);

/* ---- matrix files ------------------------------------------------------
 * CAPS_DUMP: with a dense determinant algorithm (detalg LU, QR, CHOLESKY),
 * caps_logdetD, caps_logdetD0 and kernel_logdet write each matrix M they
 * factor (not the trace-shortcut ones) to the file the environment variable
 * names, as the reference's kernel_logdet does (src/matrix.c:84-109); the
 * batched extensions do not.  The byte layout is the reference's. */
/* Extension: write a dim x dim matrix (element (i, j) at M[i*dim + j]) as
 * .npy: '<f8', 'fortran_order': True, the elements in that order (the
 * reference's matrix_save_to_file, src/matrix.c:211-254).  Returns 0 on
 * success. */
int capsx_npy_save(const char *path, const double *M, int dim);
/* Extension: read such a file (the reference's matrix_load_from_file,
 * src/matrix.c:270-335): a malloc'd dim*dim array in file order, NULL for a
 * file of another layout.  Release with capsx_npy_free. */
double *capsx_npy_load(const char *path, int *dim);
void capsx_npy_free(double *M);

/* ---- planner surrogate (no counterpart in the reference) ---------------
 * An MLP 5 -> 64 -> 64 -> 1 (tanh, FP32) predicting log|logdetD| from
 * (log R/L, log ξ(L+R)/c, m/ldim, log ωp/eV, log γ/eV), trained and
 * evaluated by the library's HIP kernels.  Parameter layout W1[64][5] b1[64]
 * W2[64][64] b2[64] W3[64] b3. */
struct capsx_mlp;
int capsx_mlp_nparams(void);
/* New network, Glorot-uniform weights from SplitMix64(seed), zero biases. */
struct capsx_mlp *capsx_mlp_new(unsigned long long seed);
void capsx_mlp_free(struct capsx_mlp *h);
void capsx_mlp_get_params(struct capsx_mlp *h, float *out);
void capsx_mlp_set_params(struct capsx_mlp *h, const float *in);
/* Standardisation x' = (x - mean)/std per feature, y' = (y - ymean)/ystd. */
void capsx_mlp_set_norm(struct capsx_mlp *h, const float *xmean, const float *xstd, float ymean, float ystd);
void capsx_mlp_get_norm(struct capsx_mlp *h, float *xmean, float *xstd, float *ymean, float *ystd);
/* Predictions in raw units for raw features X[n][5]. */
void capsx_mlp_predict(struct capsx_mlp *h, const float *X, int n, float *Y);
/* Gradient of the standardised MSE over n standardised samples. */
void capsx_mlp_grad(struct capsx_mlp *h, const float *Xn, const float *Tn, int n, float *grad);
/* Adam on raw (X, Y): minibatches of `batch`, order shuffled per epoch,
 * step size cosine-annealed from lr to lr/20; sets the standardisation from
 * the data; returns the final standardised MSE (loss_curve: one value per
 * epoch, may be NULL). */
float capsx_mlp_train(struct capsx_mlp *h, const float *X, const float *Y, int n, int epochs, int batch,
                      float lr, unsigned long long seed, float *loss_curve);
/* Attach a network (borrowed) for caps_free_energy mode 2, with the
 * material features it is evaluated at. */
void capsx_set_surrogate(caps_t *self, struct capsx_mlp *mlp, float f_omegap, float f_gamma);
/* Launch sequences used by the free-energy driver since the last reset
 * (resets when reset != 0). */
long capsx_plan_batches(caps_t *self, int reset);
/* Surrogate accuracy on the blocks computed for its plans since the last
 * reset: out = {blocks compared, RMS and largest |log|logdetD| error|,
 * frequencies compared, mean and largest |predicted - actual m-cutoff|}. */
void capsx_surrogate_report(caps_t *self, double out[6], int reset);

/* ---- further extensions: batched points, diagnostics, test hooks ----------
 * None has a counterpart in the reference's public API; the point functions
 * evaluate the reference's per-point routines on many points in one launch. */

/* n blocks as capsx_logdetD_multi, with per block tr M (trace, may be NULL)
 * and flags (may be NULL): bits 0-1 how the value was obtained (0 LU, 1 trace
 * shortcut |tr M| < 1e-8, 2 LU-mode Mercator shortcut, 3 non-finite trace;
 * the rules of src/matrix.c:57-149), bit 2 set when the LU ran in delta form
 * (1 - M column diagonally dominant: no pivot search), bit 3 set when the
 * HODLR-mode rule |tr M| > |log det| replaced the LU value by -tr M. */
int capsx_logdetD_multi_info(caps_t *const *objs, const double *xi_, const int *m, int n, double *out,
                             double *trace, int *flags);
/* log|a_l|, log|b_l| of the sphere at n points (xi_[i], l[i]), as caps_mie
 * (src/libcaps.c:557-602) per point.  Returns 0 (1: n < 0). */
int capsx_mie_points(caps_t *self, int n, const double *xi_, const int *l, double *lna, double *lnb);
/* Fresnel r_TE, r_TM of the plate at n points (xi_[i], k_[i]), as
 * caps_fresnel (src/libcaps.c:616-645) per point.  Returns 0 (1: n < 0). */
int capsx_fresnel_points(caps_t *self, int n, const double *xi_, const double *k_, double *r_TE, double *r_TM);
/* log I_{n+1/2}(x), log K_{n+1/2}(x) (order[i] = n >= 0; NaN for order < 0)
 * and I_nu(x)/I_{nu+1}(x) (nu[i] >= 0; NaN for nu < 0) at n points: the device
 * functions of the Mie coefficients (src/bessel.c:635-660, 781-833). */
int capsx_bessel_points(int n, const int *order, const double *nu, const double *x, double *logI, double *logK,
                        double *ratio);
/* Number of quadrature nodes of the K integrals of block (xi_, m). */
int capsx_block_nodes(caps_t *self, double xi_, int m);
/* Test hook: the K table (log|K_nu^p|, nu = mu..nu_max, 2*nK values) and the I
 * table (log|I|, sign; 2*nl*nl values, l from lo) of one block as computed
 * on the device.  Returns 0, 1 if nK > cap_nK or nl > cap_nl. */
int capsx_block_tables(caps_t *self, double xi_, int m, int *mu, int *nK, double *lnK, int *lo, int *nl,
                       double *logI, signed char *sgnI, int cap_nK, int cap_nl);
/* Test hook: copy the N x N matrix M of a caps_M_t (column-major); returns N. */
int capsx_M_copy(caps_M_t *self, double *out);
/* Test hook: log|det A| of n column-major dim x dim matrices (lda = dim) on
 * the batched pivoted LU.  Returns 0 on success. */
int capsx_lu_logdet_batch(const double *A, int dim, int n, double *out);
/* Test hook: log|det(1 - M)| of n column-major dim x dim matrices M through
 * the device's column dominance test (delta[b] = 1: delta form, no pivot
 * search; 0: pivoted LU).  No trace shortcut.  Returns 0 on success. */
int capsx_logdet_1mM_batch(const double *M, int dim, int n, double *out, int *delta);
/* Test hook: 1 = delta form for column-dominant blocks (default), 0 = the
 * pivoted LU for every block.  Returns the previous setting. */
int capsx_lu_delta_mode(int mode);
/* Autotuning: capsx_lu_time with the matrix form: 0 = 1 - 0.3 X/sqrt(dim)
 * (pivoted LU), 1 = 1 - M, M = 0.3 X/dim (column dominant: delta form). */
int capsx_lu_time_form(int dim, int n, int reps, int levels, int streams, int xcd, int tile, int form,
                       double *ms);
/* Test hook: D = A(16x4) B(4x16) through one FP64 MFMA instruction (the
 * register layout check).  Returns 0 on success. */
int capsx_mfma_probe(const double *A, const double *B, double *D);
/* Diagnostics: HIP events around every trailing-update launch (on != 0). */
void capsx_profile(caps_t *self, int on);
/* Accumulated statistics since the last reset: out[0..5] stage milliseconds
 * (Mie, K, I, assembly, LU, total; the Mie kernel runs beside the K
 * quadrature, whose time is what it adds after the Mie kernel ends),
 * out[6] blocks, out[7] blocks factored,
 * out[8] algorithmic LU flops, out[9] batches, out[10] I integrals; with
 * capsx_profile on, out[11..14] the milliseconds, algorithmic flops,
 * launches and algorithmic bytes of the trailing updates run by
 * k_lu_gemm_dma, out[15] the milliseconds of the other trailing-update
 * launches (k_lu_gemm tiles for regions one panel wide or high). */
void capsx_stats(caps_t *self, double out[16]);
void capsx_stats_reset(caps_t *self);

#ifdef __cplusplus
}
#endif

#endif

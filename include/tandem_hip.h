/*
 * tandem_hip.h -- C ABI of the MI355X FP64 compact-scheme time stepper.
 *
 * The reference's hot path is a set of Fortran module procedures working on
 * module-global arrays of one MPI rank (src/Subroutines3D.f90, the stage body
 * inline in src/Main3D.f90:552-965).  This ABI is the drop-in for that path:
 * a context owns device mirrors of those arrays for every sub-block of every
 * case, and each entry point below replaces one reference routine (or one
 * block of inline statements) and acts on all sub-blocks at once, with the
 * MPI exchanges performed as device-side copies.
 *
 * Conventions
 *   - every function returns 0 on success, a nonzero TP_E_* code otherwise;
 *     tp_last_error() gives the message (per thread);
 *   - arrays use the reference's layout: a(0:lmx, var) in Fortran order with
 *     l = (k*(let+1)+j)*(lxi+1)+i, faces face(0:F-1, c, 0:1);
 *   - directions nn = 1 (xi), 2 (eta), 3 (zeta) and columns nz, m are 1-based
 *     exactly as in the reference;
 *   - calls are asynchronous on the context's HIP stream; tp_download and
 *     tp_sync synchronise.
 *   - the library never falls back to a CPU path: without a usable GPU,
 *     tp_create fails.
 */
#ifndef TANDEM_HIP_H
#define TANDEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tp_context tp_context;

enum {
  TP_OK = 0, TP_E_ARG = 1, TP_E_STATE = 2, TP_E_HIP = 3, TP_E_SIZE = 4, TP_E_LAYOUT = 5
};

/* integer case parameters, iparams[] (inputo.dat lines 5-10, 20; src/Main3D.f90:39-56).
 * TP_I_INFLOW switches on an inflow disturbance through the sponge target
 * (the reference's spongego gust path, src/SingleTurbulence3D.f90:332-372, which
 * it ships disabled): a velocity table over the inflow-sponge points sampled
 * at TP_I_NTZ+1 times over one period, filled from
 *   1: TP_I_NMODES seeded random Fourier modes of turbulence (tp_inflow_modes);
 *   2: the single mode of a sinusoidal transverse gust (tp_gust_mode);
 *   3: the reference's synthetic eddies (src/SingleTurbulence3D.f90:262-291),
 *      given by tp_set_inflow_eddies, the table scaled by TP_D_CFIT.
 * TP_I_FORCING = 1 adds the reference's localised spanwise-modal body force
 * (forceup/forcego, src/rpt.f90:728-769; switched off in its decks). */
enum { TP_I_NKRK = 0, TP_I_NVISCOUS, TP_I_NSMF, TP_I_LES, TP_I_NEXTRABC, TP_I_NEXTGCIC,
       TP_I_INFLOW, TP_I_NMODES, TP_I_SEED, TP_I_NTZ, TP_I_FORCING, TP_I_NPARAM = 16 };
/* double case parameters, dparams[] (inputo.dat, inputp.dat; fltk/fltkbc in units of pi).
 * Inflow: TP_D_TGUSTO start time (tgusto), TP_D_SLIT length of the periodic
 * turbulence box along x (the table's period is slit/u_inf, as slit/uoo(1) in
 * the reference), TP_D_TU turbulence intensity (rms of each velocity
 * component over |u_inf|), TP_D_LTURB integral length scale, TP_D_GUSTW half
 * width in y of the band of inflow-sponge points that receive the turbulence.
 * Gust (TP_I_INFLOW = 2): TP_D_TU amplitude over |u_inf|, TP_D_GUSTK reduced
 * frequency omega c / (2 |u_inf|) with chord c = 1; TP_D_SLIT is derived.
 * Eddies (TP_I_INFLOW = 3): TP_D_CFIT scale of the table (cfit), TP_D_SLIT
 * the x-period of the eddy train.  Forcing: TP_D_AMFOR amplitude, TP_D_XFOR,
 * TP_D_YFOR centre, TP_D_RFOR radius, TP_D_TSFOR, TP_D_TEFOR time window
 * (amfor, xfor, yfor, rfor, tsfor, tefor of src/MainVar3D.f90:167). */
enum {
  TP_D_REOO = 0, TP_D_TEMPOO, TP_D_AMACH1, TP_D_AMACH2, TP_D_AMACH3, TP_D_WTEMP, TP_D_CFL,
  TP_D_TMAX, TP_D_TIMF, TP_D_FLTK, TP_D_FLTKBC, TP_D_DTO, TP_D_AOA0, TP_D_AOA1, TP_D_TPS,
  TP_D_TP, TP_D_SMAGO1, TP_D_SMAGO2, TP_D_DOMLEN, TP_D_SZTH1, TP_D_SZTH2, TP_D_SZXT,
  TP_D_SZCO, TP_D_SPAN, TP_D_WLEW, TP_D_WLEA, TP_D_TGUSTO, TP_D_SLIT, TP_D_TU, TP_D_LTURB,
  TP_D_GUSTW, TP_D_GUSTK, TP_D_CFIT, TP_D_AMFOR, TP_D_XFOR, TP_D_YFOR, TP_D_RFOR, TP_D_TSFOR,
  TP_D_TEFOR, TP_D_NPARAM = 48
};
/* arrays reachable by tp_upload / tp_download (names of src/MainVar3D.f90:78-102) */
enum {
  TP_QA = 0, TP_QO, TP_DE, TP_RR, TP_SS, TP_XIM, TP_ETM, TP_ZEM, TP_P, TP_YACO, TP_TV,
  TP_DRVA1, TP_DRVA2, TP_DRVA3, TP_DRVB1, TP_DRVB2, TP_DRVB3, TP_SEND1, TP_SEND2, TP_SEND3,
  TP_RECV1, TP_RECV2, TP_RECV3, TP_CM1, TP_CM2, TP_CM3, TP_XU, TP_XL, TP_YU, TP_YL, TP_RPEX,
  TP_NARRAY
};
/* scalars of a case (tp_get_scalars / tp_set_scalars), src/MainVar3D.f90:139-141 */
enum {
  TP_S_N = 0, TP_S_NDT, TP_S_DT, TP_S_DTS, TP_S_DTE, TP_S_TIMO, TP_S_UMF1, TP_S_UMF2, TP_S_UMF3,
  TP_S_DUDT1, TP_S_DUDT2, TP_S_DUDT3, TP_S_DTK, TP_S_DTKO, TP_S_AOA, TP_S_NSCALAR = 16
};
/* kernel classes reported by tp_profile_read (TP_K_FLUX: the pass forming the
 * viscous stresses and all fluxes; TP_K_VISC: none since the stresses are
 * formed inside that pass, kept so that the numbering stays) */
enum {
  TP_K_DERIV = 0, TP_K_FILTER, TP_K_EXCH, TP_K_POINT, TP_K_FLUX, TP_K_VISC, TP_K_GCBC, TP_K_RK,
  TP_K_MISC, TP_K_NCLASS
};

/* ---- context ---- */
int32_t tp_create(tp_context **ctx);            /* on the current HIP device */
int32_t tp_destroy(tp_context *ctx);
const char *tp_last_error(void);

/* A case is one reference run: its sub-blocks (MPI ranks of the reference),
 * their boundary codes and neighbours (src/Main3D.f90:92-178,
 * src/SingleTurbulence3D.f90:121-164), its corner junctions
 * (src/SingleTurbulence3D.f90:378-443) and its input-deck parameters.
 *   dims[b*3+d]             lxi, let, lze of sub-block b
 *   nbc/ncd/nsz[b*6+d*2+ip] boundary code, neighbour sub-block (case-local
 *                           index, -1 none), sponge flag of face ip of direction d
 *   jct[j*12+mbr*3+{0,1,2}] (sub-block, ip, jp) of corner member mbr of junction j
 *   jvel[j*2+{0,1}]         members whose velocities are averaged
 * Several cases in one context are advanced together (ensembles).  A case
 * added after tp_finalize waits, pending, until tp_admit puts it on the
 * device; until then it takes no device memory and is not advanced. */
int32_t tp_add_case(tp_context *ctx, int32_t nblocks, const int32_t *dims, const int32_t *nbc,
                    const int32_t *ncd, const int32_t *nsz, int32_t njct, const int32_t *jct,
                    const int32_t *jvel, const int32_t *iparams, const double *dparams,
                    int32_t *case_id);
/* the checks tp_add_case makes of a case, without a context (and so
 * without a GPU): boundary codes, neighbours and matching faces, line
 * lengths the closures need (src/Main3D.f90:92-178), junction members, and
 * block sizes -- a block holds fewer than 2^29 points (the line solves
 * address a field with 32-bit byte offsets).  No counterpart in the
 * reference, which reads its deck and aborts on a bad one. */
int32_t tp_check_case(int32_t nblocks, const int32_t *dims, const int32_t *nbc, const int32_t *ncd,
                      int32_t njct, const int32_t *jct, const int32_t *jvel, const int32_t *iparams);
/* allocate device state of the cases added so far, compute the scheme
 * coefficients and LU factors (src/Main3D.f90:293-332), build the launch
 * tables.  A context may be finalized with no case (cases then come through
 * tp_admit). */
int32_t tp_finalize(tp_context *ctx);

/* grid of one sub-block, ss(0:lmx,1:3) = x, y, z (src/rpt.f90:673-696); the
 * host keeps a copy (a pending case's grid goes to the device on tp_admit) */
int32_t tp_set_grid(tp_context *ctx, int32_t case_id, int32_t block, const double *xyz);
/* metrics, Jacobian, face normals, wall-slot flags and sponge coefficients
 * from the grid (src/Main3D.f90:365-480, src/SingleTurbulence3D.f90:209-249),
 * for every case on the device whose geometry is not set yet; the fields of
 * the other cases are not touched */
int32_t tp_setup_geometry(tp_context *ctx);
/* uniform initial state (src/SingleTurbulence3D.f90:179-205) and n = 0, t = 0 */
int32_t tp_initialo(tp_context *ctx, int32_t case_id);

int32_t tp_upload(tp_context *ctx, int32_t case_id, int32_t block, int32_t array_id,
                  const double *src, int64_t count);
int32_t tp_download(tp_context *ctx, int32_t case_id, int32_t block, int32_t array_id,
                    double *dst, int64_t count);
int64_t tp_array_count(tp_context *ctx, int32_t case_id, int32_t block, int32_t array_id);
int32_t tp_get_scalars(tp_context *ctx, int32_t case_id, double *s16);
int32_t tp_set_scalars(tp_context *ctx, int32_t case_id, const double *s16);

/* ---- per-call drop-ins, all sub-blocks of all cases ---- */
/* mpigo(nt,nrt,n45,itag): src/Subroutines3D.f90:16-80 (closure sums to the neighbours' recv) */
int32_t tp_mpigo(tp_context *ctx, int32_t nt, int32_t nrt, int32_t n45);
/* deriv(nn,nz,m): src/Subroutines3D.f90:84-141 (rr(:,nn) = d rr(:,nz)/d index_nn; m = 1..5
 * keeps the end values in drva(:,m,0:1) as the reference does, m = 0 skips that) */
int32_t tp_deriv(tp_context *ctx, int32_t nn, int32_t nz, int32_t m);
/* filte(nn,nz): src/Subroutines3D.f90:145-206 (rr(:,nz) += filter increment) */
int32_t tp_filte(tp_context *ctx, int32_t nn, int32_t nz);

/* deriv(nn,nz,m) on one field held by the caller in device memory:
 * dst = d src / d index_nn over the block of (lxi+1) x (let+1) x (lze+1)
 * points in the reference's layout (xi fastest), the pentadiagonal compact
 * scheme of src/Subroutines3D.f90:84-141 with the closures of the line ends
 * nbc2[0], nbc2[1]: 10, 20, 25 or 30 (one-sided rows) or both 45 (periodic,
 * the closure sums of mpigo, src/Subroutines3D.f90:16-80, with n45 = 0,
 * formed from the other end of each line).  src and dst must not overlap.
 * Lines of at most 512 points (longer ones: a context and tp_deriv).  The
 * solve is enqueued on stream (a hipStream_t; NULL: the null stream); the
 * library allocates its tables in device memory for the call and returns
 * once the solve has run. */
int32_t tp_field_deriv(int32_t lxi, int32_t let, int32_t lze, const int32_t *nbc2, int32_t nn,
                       const double *src, double *dst, void *stream);
/* tp_field_deriv without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc); no counterpart in the reference.  The tables go
 * to a work area the library holds on each GPU (1 MiB, kept until the
 * process ends): calls needing more (over about 4 M lines) fail with
 * TP_E_SIZE.  The solve is only enqueued on stream, and the next call on
 * the same GPU reuses the area: calls that overlap on one GPU (other
 * streams, other host threads) race; calls in turn on one stream do not.
 * Same results as tp_field_deriv, bit for bit. */
int32_t tp_field_deriv_noalloc(int32_t lxi, int32_t let, int32_t lze, const int32_t *nbc2, int32_t nn,
                               const double *src, double *dst, void *stream);

/* ---- coarse entry points ---- */
/* filtering of qa and qo = qa: src/Main3D.f90:554-561 */
int32_t tp_filter_step(tp_context *ctx);
/* one Runge-Kutta stage nk: src/Main3D.f90:567-958 */
int32_t tp_stage(tp_context *ctx, int32_t nk);

/* ---- the sections of a stage as drop-ins ----
 * tp_stage(ctx, nk) is these nine calls in this order; called one by one they
 * compute the same state bit for bit (the stresses then pass through tv at
 * every point, which the fused stage keeps in registers).  Each acts on all
 * sub-blocks of all cases, like the inline blocks of src/Main3D.f90 it
 * replaces on every MPI rank. */
/* stage start: frame velocity (movef, src/Main3D.f90:571-572), primitives
 * de(:,1:5), p, ss(:,1) = viscosity (:576-583), and at nk = 1 the time step
 * size (:587-622) */
int32_t tp_primitives(tp_context *ctx, int32_t nk);
/* viscous stresses and heat fluxes into tv(:,1:9) = txx tyy tzz txy tyz tzx
 * hxx hyy hzz at every point (src/Main3D.f90:626-693; gradients of de(:,2:5)
 * from the compact derivative); nothing for inviscid cases */
int32_t tp_viscous(tp_context *ctx);
/* at nk = 1 the surface sample (src/Main3D.f90:695-706), then the
 * metric-weighted Euler and viscous fluxes (from tv) and their compact
 * derivatives into de(:,1:5), the end values kept in drva (:708-768) */
int32_t tp_flux_divergence(tp_context *ctx, int32_t nk);
/* sponge relaxation towards the rest state or the inflow target
 * (spongego, src/SingleTurbulence3D.f90:332-374) */
int32_t tp_sponge(tp_context *ctx);
/* characteristic boundary and interface conditions on de: preparation,
 * interface exchange and application (src/Main3D.f90:778-869) */
int32_t tp_gcbc(tp_context *ctx);
/* frame velocity for the update and qa = qo - dtk yaco de
 * (src/Main3D.f90:873-881) */
int32_t tp_rk_update(tp_context *ctx, int32_t nk);
/* wall temperature and velocity (src/Main3D.f90:884-895) */
int32_t tp_wall(tp_context *ctx);
/* corner junctions (junction, src/SingleTurbulence3D.f90:378-443) */
int32_t tp_junction(tp_context *ctx);
/* interface averaging of qa (src/Main3D.f90:897-952) */
int32_t tp_interface_avg(tp_context *ctx);
/* nsteps whole time steps (filter + nkrk stages + n, t advance): src/Main3D.f90:552-965.
 * qa, qo, p, de and the scalars end as the reference leaves them.  ss(:,1),
 * the viscosity (stage scratch), does not: the flux pass forms it without
 * storing it, and the step stores it only on the steps whose time-step limit
 * reads it (n = 1, 11, 21, ...). */
int32_t tp_step(tp_context *ctx, int32_t nsteps);
/* the same, replaying one step captured in a HIP graph */
int32_t tp_step_graph(tp_context *ctx, int32_t nsteps);
int32_t tp_sync(tp_context *ctx);

/* ---- options ----
 * TP_OPT_EXACT_LINES = 1: line solves run the reference's sequential LU
 *   sweeps one lane per line (bit-identical to a sequential CPU sweep);
 *   0 (default): chunked sweeps -- each line cut into P chunks swept by P
 *   waves, the chunks joined by composing their 2x2 carry maps -- equal to
 *   the sequential sweep up to rounding of the chunk carries (~1e-16
 *   relative).  May be changed between steps.
 * TP_OPT_CHUNKS = P (1 to 16) fixes the chunks per line of the
 *   chunked sweeps for every block and direction whose lines L allow chunks
 *   of ceil(L/P) <= 32 points; 0 (default) lets the library choose per block
 *   and direction.  The chunk count of a block depends only on its own line
 *   lengths (and on the autotuner's choice for its line-length class), so a
 *   case computes bit-identical results alone or batched with others when
 *   the chunk counts are fixed or left at their defaults.
 * TP_OPT_STRESSES = 1: every stage leaves the viscous stresses and heat
 *   fluxes (the reference's txx, tyy, tzz, txy, tyz, tzx, hxx, hyy, hzz) in
 *   TP_TV at every point; 0 (default): the stresses stay inside the flux
 *   pass and TP_TV holds them only at the wall rows of blocks with a wall
 *   surface, after the first stage of a step (what the surface forces read).
 *   Either way the stage's results are the same.  May be changed between
 *   steps. */
enum { TP_OPT_EXACT_LINES = 0, TP_OPT_CHUNKS = 1, TP_OPT_STRESSES = 2 };
int32_t tp_set_option(tp_context *ctx, int32_t option, int64_t value);
int32_t tp_get_option(tp_context *ctx, int32_t option, int64_t *value);

/* Time every usable line-solver configuration (sequential kernel and P =
 * 1..16 chunks) on the context's own blocks, reps launches each, and keep
 * the fastest per direction, kind and line-length class (blocks whose lines
 * have the same default chunk count form a class; the blocks of one chunk
 * count share a launch), five fields per launch as a time step runs them.
 * Overwrites the stage scratch (the flux and gradient work array), so call
 * it between steps.  chosen[d*2+kind] (kind 0 deriv, 1 filter) receives the
 * chunk count used for the most points (0: sequential kernel), us[d*2+kind]
 * the time of the direction's launches; either may be NULL.  No counterpart
 * in the reference: it replaces the fixed 1-D loop structure of deriv/filte
 * with a measured launch shape. */
int32_t tp_autotune(tp_context *ctx, int32_t reps, int32_t *chosen, double *us);
/* chunks per line used for the most points, [d*2+kind] (0: sequential kernel) */
int32_t tp_col_config(tp_context *ctx, int32_t *chunks6);

/* ---- turbulent inflow ----
 * Seeded random Fourier modes of homogeneous isotropic turbulence with a von
 * Karman spectrum: u'(x) = sum_n a_n cos(k_n . x + phi_n), a_n perpendicular
 * to k_n (divergence free), <u'.u'> = sum_n |a_n|^2 / 2 = 3 (tu |uoo|)^2
 * (rms tu|uoo| in each component), k_x a multiple of 2 pi/slit
 * and k_z of 2 pi/span (periodic in x and in the span; span <= 0: no z
 * constraint).  Magnitudes log-uniform in [k_e/5, 40 k_e] with k_e =
 * sqrt(pi) Gamma(5/6) / (Gamma(1/3) lturb), directions uniform on the sphere,
 * phases uniform; splitmix64 stream from seed.  out8[n*8 + {0..7}] = k_n (3),
 * a_n (3), phi_n, 0.  Needs no device.  No counterpart in the reference
 * (its eddy generator is switched off there); used for the inflow table. */
int32_t tp_inflow_modes(int32_t nmodes, uint64_t seed, double tu, double lturb, const double *uoo3,
                        double slit, double span, double *out8);

/* Sinusoidal transverse gust as one Fourier mode in the layout of
 * tp_inflow_modes: wavevector along u_inf with |k| = 2 kr (reduced frequency
 * kr = omega c / (2 |u_inf|), chord 1), amplitude tu |u_inf| normal to u_inf
 * in the x-y plane, phase 0; *slit receives the table's x-period, so that
 * slit / u_x is the gust period.  Needs no device.  No counterpart in the
 * reference. */
int32_t tp_gust_mode(double tu, double kr, const double *uoo3, double *out8, double *slit);

/* ---- the reference's synthetic eddies (TP_I_INFLOW = 3) ----
 * The eddies of the reference's inflow (inputext, src/SingleTurbulence3D.f90:
 * 86-115): from its random numbers (randnum.dat, the first 8 nits values in
 * file order: ran(:,1), ait(:,1), ran(:,2), ait(:,2), ran(:,3), ait(:,3),
 * yit(:), zit(:)) and deck values deck[18] = span, slit, gaus, cutlb,
 * domlen, szth1, tl0(0), tl1(0), tlw(0), tl0(1), tl1(1), tlw(1), tam(1:3,0),
 * tam(1:3,1) (inputp.dat): out10[n*10 + {0..9}] = centre xit, yit, zit,
 * sit(1:3) (1/size^2), ait(1:3) (amplitudes), type mxc.  Needs no device. */
int32_t tp_eddy_params(int32_t nits, const double *randnum, const double *deck, double *out10);
/* the eddies a case with TP_I_INFLOW = 3 builds its inflow table from
 * (src/SingleTurbulence3D.f90:262-298: each eddy and its images one slit
 * upstream and one span to either side, convected with u_inf, the table
 * scaled by TP_D_CFIT); before the case's geometry is set up */
int32_t tp_set_inflow_eddies(tp_context *ctx, int32_t case_id, int32_t nits, const double *eddies10);

/* ---- the reference's file formats ----
 * The reference writes one restart and one raw grid file per parent block
 * (a block of its bkx x bky x bkz layout, cut into sub-blocks for MPI) and
 * one PLOT3D file for all parent blocks.  tp_set_parents tells which parent
 * each sub-block belongs to: parent[b], the sub-block's first point in it
 * start3[b*3+d], and the parents' last indices pdims3[mb*3+d] (lxi, let, lze
 * of the parent).  Without it every sub-block is its own parent. */
int32_t tp_set_parents(tp_context *ctx, int32_t case_id, int32_t nparent, const int32_t *parent,
                       const int32_t *start3, const int32_t *pdims3);
/* the parent layout in use (outputs may be NULL; parent and start3 hold one
 * entry per sub-block, pdims3 one per parent) */
int32_t tp_get_parents(tp_context *ctx, int32_t case_id, int32_t *nparent, int32_t *parent,
                       int32_t *start3, int32_t *pdims3);
/* the grid of a sub-block as tp_set_grid gave it (count 3 * points) */
int32_t tp_get_grid(tp_context *ctx, int32_t case_id, int32_t block, double *xyz, int64_t count);
/* lxi, let, lze of every sub-block, and the case's iparams / dparams as added */
int32_t tp_case_dims(tp_context *ctx, int32_t case_id, int32_t *dims3);
int32_t tp_case_params(tp_context *ctx, int32_t case_id, int32_t *iparams, double *dparams);
/* <rundir>/rsta/restartNNN.dat per parent block NNN (src/rpt.f90:555-623,
 * wrRsta): n and ndt as INTEGER4 at bytes 0 and 8, dt, dts, dte, timo as
 * REAL8 at 16..40, then qa(0:lmx,1:5) of the parent block (REAL8, Fortran
 * order) from byte 48.  Directories are created when missing. */
int32_t tp_write_restart(tp_context *ctx, int32_t case_id, const char *rundir);
/* read them back into qa and n, ndt, dt, dts, dte, timo (rdRsta,
 * src/rpt.f90:628-661); the case must be on the device */
int32_t tp_read_restart(tp_context *ctx, int32_t case_id, const char *rundir);
/* <rundir>/misc/gridNNN.dat per parent block: the grid ss(0:lmx,1:3) as set
 * by tp_set_grid, REAL8, no header (written by src/Grid.f90:452-459, read by
 * rdGrid, src/rpt.f90:666-696) */
int32_t tp_write_grid(tp_context *ctx, int32_t case_id, const char *rundir);
/* read them and set every sub-block's grid (tp_set_grid) */
int32_t tp_read_grid(tp_context *ctx, int32_t case_id, const char *rundir);
/* PLOT3D multiblock files in <rundir>/out (src/rpt.f90:115-330): what & 1
 * grid.xyz (INTEGER4 block count and IMax, JMax, KMax per block, then REAL4
 * x, y, z per block); what & 2 solT<timo as f8.4, blanks as 0>.q (same
 * header, then per block REAL4 Mach, AoA, Re, time and REAL4 rho, u + umf,
 * v + umf, w + umf, p -- p being the pressure of the last stage, as the
 * reference writes it) */
int32_t tp_write_plot3d(tp_context *ctx, int32_t case_id, const char *rundir, int32_t what);

/* ---- surface statistics on the device ----
 * The wall of a case is the eta face of each sub-block with the wall code
 * (20 + 5 nviscous), as src/rpt.f90:774-864 (wallArea, walldir) sets it up:
 * cell areas |dr/dxi x dr/dzeta| halved at the parent block's xi and zeta
 * ends (faces that are not BC 40 splits), unit normals along the eta metric
 * into the fluid, points ordered xi fastest.  At the first stage of every
 * step whose n is a multiple of the interval, a sample takes
 *   - the force coefficients of src/rpt.f90:1017-1056 (clpost): pressure
 *     p n area and wall shear rho (tau . n) area summed over the wall, times
 *     2 / (M^2 span), turned into lift and drag with the angle of attack as
 *     src/Main3D.f90:700-704 prints them (Cl, Cd);
 *   - p at every wall point.
 * Time means are trapezoidal integrals over the sampled times divided by the
 * window, the running form of the time averages of src/rptPost.f90:870-925
 * (p3dStats); rms is the square root of mean(p^2) - mean(p)^2.  Sampling
 * runs inside the captured step, so graph replays sample too. */
/* start (or restart, clearing the accumulators) sampling every `every`
 * steps from the next step on; every = 0 stops sampling */
int32_t tp_stats_start(tp_context *ctx, int32_t case_id, int32_t every);
/* out8 = samples, first and last sample time, last Cl and Cd, their time
 * means, window length (means equal the last values for one sample) */
int32_t tp_stats_forces(tp_context *ctx, int32_t case_id, double *out8);
/* wall points of a sub-block (0: it has no wall face; -1: bad arguments) */
int64_t tp_stats_wall_count(tp_context *ctx, int32_t case_id, int32_t block);
/* per wall point: time mean and rms of p, cell area, unit normal [3][count];
 * any output may be NULL; count must equal tp_stats_wall_count.  Cp is
 * (mean p - 1/gamma) 2 / M^2, src/rptPost.f90:693-694. */
int32_t tp_stats_wall(tp_context *ctx, int32_t case_id, int32_t block, double *mean_p, double *rms_p,
                      double *area, double *wnor3, int64_t count);

/* ---- cases joining and leaving a running context ----
 * No counterpart in the reference (one case per MPI job there): these serve
 * ensembles and the request loop below.  Cases on the device keep their state
 * and scalars while others join or leave; a join or leave drops a captured
 * step graph (tp_step_graph captures again). */
/* device bytes case_id will hold once admitted: block arrays, solver tables,
 * sponge and inflow arrays (needs the case's grids) -- the figure an HBM
 * budget admits by */
int32_t tp_case_bytes(tp_context *ctx, int32_t case_id, int64_t *bytes);
/* put a pending case on the device: allocate, set up its geometry (metrics,
 * sponge, inflow table) from the grids given by tp_set_grid, and include it
 * in every later launch.  Its state is zero: follow with tp_initialo or
 * tp_upload.  On failure the case stays pending and the context unchanged. */
int32_t tp_admit(tp_context *ctx, int32_t case_id);
/* the same for n pending cases at once (one rebuild of the launch tables and
 * one geometry pass for all of them); status[i] is 0 for an admitted case,
 * else the error that left case_ids[i] pending.  Returns 0 when every case
 * was admitted, else the first error; a failure after the cases were built
 * (linking, geometry) leaves all of them pending. */
int32_t tp_admit_cases(tp_context *ctx, int32_t n, const int32_t *case_ids, int32_t *status);
/* drop a case (pending or on the device) and free its device memory */
int32_t tp_remove_case(tp_context *ctx, int32_t case_id);
/* the same for n cases at once (one rebuild of the launch tables); nothing
 * is removed if any id is unknown */
int32_t tp_remove_cases(tp_context *ctx, int32_t n, const int32_t *case_ids);
/* move n cases that are set up on the device (geometry done) from their
 * contexts srcs[i] into dst, on the same device, without copying or
 * recomputing anything: each keeps its device memory, state and scalars and
 * gets the handle new_ids[i] in dst (one rebuild of dst's launch tables).
 * A host can thus set cases up in contexts of their own on other host
 * threads and streams while dst steps, and join them between steps. */
int32_t tp_adopt_cases(tp_context *dst, int32_t n, tp_context *const *srcs, const int32_t *case_ids,
                       int32_t *new_ids);
/* out[0] state (0 pending, 1 on the device, 2 with geometry), out[1] points,
 * out[2] sub-blocks, out[3] device bytes held */
int32_t tp_case_info(tp_context *ctx, int32_t case_id, int64_t *out4);

/* ---- request loop (no counterpart in the reference) ----
 * A server owns a context.  Requests -- whole cases with their grids and a
 * step count -- queue on the host with an arrival time.  Two staging
 * threads take arrived requests in arrival order and set each up in a
 * context of its own (device memory, metrics, sponge, inflow table, the
 * uniform initial state)
 * while their device bytes fit the HBM budget beside the cases on the device
 * and fewer than tp_server_set_staging requests wait set up; this runs on
 * its own streams, beside the stepping.  Each iteration
 *   1. admits set-up requests in arrival order (a request whose set-up
 *      finished first may go before an earlier one still being set up)
 *      while a slot is free and the cases on the device hold fewer points
 *      than the fill target (tp_server_set_fill), joining them to the cases
 *      on the device
 *      (tp_adopt_cases: no copy, no recomputation); an idle device first
 *      waits until the requests that have arrived are set up, or until those
 *      set up fill the slots, the fill target or the staging window, so that
 *      a burst runs as one batch;
 *   2. advances every case on the device by the same number of steps (at
 *      most max_steps, at most what any of them still needs) in shared
 *      launches (captured step graph when use_graph);
 *   3. copies the final qa of every case that made its steps to the host and
 *      takes it off the device.
 * Cases join and leave while others are resident; each case's result does
 * not depend on the others (see tp_set_option on line-solver chunks for
 * bit-identity).  Times are seconds on the server's clock. */
typedef struct tp_server tp_server;
/* budget_bytes <= 0: 90 % of the device memory free at creation */
int32_t tp_server_create(int64_t budget_bytes, int32_t max_active, int32_t use_graph, tp_server **srv);
int32_t tp_server_destroy(tp_server *srv);
tp_context *tp_server_context(tp_server *srv);   /* for tp_set_option / tp_autotune */
/* requests set up ahead of their admission (default: max_active): each
 * request is set up in a context of its own (allocation, metrics, sponge, inflow table,
 * initial state) on a host thread and stream of its own while the cases on
 * the device step, and joins them between steps (tp_adopt_cases).  Staged
 * cases hold their device bytes within the budget. */
int32_t tp_server_set_staging(tp_server *srv, int32_t ahead);
/* admit no further request while the cases on the device hold at least
 * `points` grid points (one case is always admitted; 0: no limit, the
 * default).  Enough points to fill the GPU keep its throughput while large
 * cases finish one after another instead of sharing it. */
int32_t tp_server_set_fill(tp_server *srv, int64_t points);
/* message of the last failed server call on the calling thread */
const char *tp_server_error(tp_server *srv);
double tp_server_clock(tp_server *srv);
/* queue a case (arguments of tp_add_case, then xyz[b] = grid of sub-block b);
 * t_arrival < 0: the time of the call.  Fails with TP_E_SIZE if the case
 * alone exceeds the budget.  May be called from another host thread while
 * tp_server_run or tp_server_iterate runs (a client submitting into a
 * running loop); the other server calls belong to one thread. */
int32_t tp_server_submit(tp_server *srv, int32_t nblocks, const int32_t *dims, const int32_t *nbc,
                         const int32_t *ncd, const int32_t *nsz, int32_t njct, const int32_t *jct,
                         const int32_t *jvel, const int32_t *iparams, const double *dparams,
                         const double *const *xyz, int32_t nsteps, double t_arrival, int32_t *req);
/* one iteration; status6 = on the device, queued (arrived), not yet arrived,
 * finished, admitted in this iteration, steps made */
int32_t tp_server_iterate(tp_server *srv, int32_t max_steps, int32_t *status6);
/* iterate until every request submitted so far has finished, idling until
 * arrivals */
int32_t tp_server_run(tp_server *srv, int32_t max_steps);
/* info6 = state (0 queued, 1 on the device, 2 finished, 3 failed), arrival,
 * admission and finish times, steps made, device bytes; qa (may be NULL):
 * the final qa of every sub-block in order, count = tp_server_result_count.
 * The server keeps each final state in page-locked host memory (the
 * retirement copy then runs at full link speed) until tp_server_destroy. */
int32_t tp_server_result(tp_server *srv, int32_t req, double *qa, int64_t count, double *info6);
int64_t tp_server_result_count(tp_server *srv, int32_t req);
/* out16 = finished, failed, p50, p99, mean and max time to solution (arrival
 * to final state on the host), wall time from the first arrival to the last
 * finish, cases per hour over it, peak admitted bytes, peak cases on the
 * device, admissions that joined running cases, time spent stepping, in
 * admissions, in retirements, in step-graph captures, number of captures */
int32_t tp_server_stats(tp_server *srv, double *out16);
/* out6 = time spent setting requests up (staging thread), requests set up,
 * set-ups that failed, most requests set up at once, iterations stepped
 * eagerly because a set-up held the step-graph capture off, staging window */
int32_t tp_server_stats2(tp_server *srv, double *out6);

/* ---- measurement ---- */
/* record HIP events around every launch (on the context's stream) */
int32_t tp_profile(tp_context *ctx, int32_t on);
/* the same for the kernel classes in mask only (bit k: class TP_K_* k; 0:
 * off): each event pair costs the stream a few microseconds, so timing one
 * class disturbs the step less than timing all of them */
int32_t tp_profile_classes(tp_context *ctx, int32_t mask);
/* accumulated device time (ms) and launch count of a kernel class since the
 * last reset; algorithmic HBM bytes those launches move (bytes per point of
 * the class times points); reset != 0 clears the counters afterwards */
int32_t tp_profile_read(tp_context *ctx, int32_t kclass, double *ms, int64_t *launches,
                        double *bytes, int32_t reset);
/* out[0] points, out[1] sub-blocks, out[2] cases on the device, out[3] device bytes held */
int32_t tp_info(tp_context *ctx, int64_t *out4);

#ifdef __cplusplus
}
#endif
#endif

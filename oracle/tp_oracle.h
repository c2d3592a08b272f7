/*
 * tp_oracle.h -- CPU restatement of the reference solver's time-marching hot
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker for the HIP
 * path and the CPU baseline timed beside it.  No product code may link,
 * load or call it.
 *
 * Parity status of each function is recorded in oracle/tp_oracle.c (file
 * header) next to the reference file:line each one restates.
 *
 * Layout conventions follow the reference (all arrays FP64, Fortran order):
 *   a(0:lmx, var)          -> a[var*N + l],  l = (k*(let+1)+j)*(lxi+1)+i
 *   face(0:F-1, c, 0:1)    -> face[(ip*C + c)*F + jk]
 * Blocks are the reference's MPI sub-blocks; one process owns all of them
 * and the MPI exchanges become copies between block buffers.
 */
#ifndef TP_ORACLE_H
#define TP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* integer parameters (index into iparams[]) */
enum {
  ORC_I_NKRK = 0, ORC_I_NVISCOUS, ORC_I_NSMF, ORC_I_LES, ORC_I_NEXTRABC,
  ORC_I_NEXTGCIC, ORC_I_INFLOW, ORC_I_NMODES, ORC_I_SEED, ORC_I_NTZ, ORC_I_FORCING, ORC_I_NPARAM = 16
};
/* double parameters (index into dparams[]) */
enum {
  ORC_D_REOO = 0, ORC_D_TEMPOO, ORC_D_AMACH1, ORC_D_AMACH2, ORC_D_AMACH3,
  ORC_D_WTEMP, ORC_D_CFL, ORC_D_TMAX, ORC_D_TIMF, ORC_D_FLTK, ORC_D_FLTKBC,
  ORC_D_DTO, ORC_D_AOA0, ORC_D_AOA1, ORC_D_TPS, ORC_D_TP, ORC_D_SMAGO1,
  ORC_D_SMAGO2, ORC_D_DOMLEN, ORC_D_SZTH1, ORC_D_SZTH2, ORC_D_SZXT,
  ORC_D_SZCO, ORC_D_SPAN, ORC_D_WLEW, ORC_D_WLEA, ORC_D_TGUSTO, ORC_D_SLIT,
  ORC_D_TU, ORC_D_LTURB, ORC_D_GUSTW, ORC_D_GUSTK, ORC_D_CFIT, ORC_D_AMFOR, ORC_D_XFOR, ORC_D_YFOR,
  ORC_D_RFOR, ORC_D_TSFOR, ORC_D_TEFOR, ORC_D_NPARAM = 48
};

/* per-block arrays reachable through orc_array() */
enum {
  ORC_A_QA = 0, ORC_A_QO, ORC_A_DE, ORC_A_RR, ORC_A_SS, ORC_A_XIM, ORC_A_ETM,
  ORC_A_ZEM, ORC_A_P, ORC_A_YACO, ORC_A_TV,  /* tv = txx tyy tzz txy tyz tzx hxx hyy hzz */
  ORC_A_DRVA1, ORC_A_DRVA2, ORC_A_DRVA3, ORC_A_DRVB1, ORC_A_DRVB2, ORC_A_DRVB3,
  ORC_A_SEND1, ORC_A_SEND2, ORC_A_SEND3, ORC_A_RECV1, ORC_A_RECV2, ORC_A_RECV3,
  ORC_A_CM1, ORC_A_CM2, ORC_A_CM3, ORC_A_XU, ORC_A_XL, ORC_A_YU, ORC_A_YL,
  ORC_A_RPEX, ORC_A_SBCC, ORC_A_ASZ, ORC_A_BSZ, ORC_A_VITO, ORC_A_ATZ, ORC_A_NARRAY
};

/* scalar state reachable through orc_scalars() */
enum {
  ORC_S_N = 0, ORC_S_NDT, ORC_S_DT, ORC_S_DTS, ORC_S_DTE, ORC_S_TIMO,
  ORC_S_UMF1, ORC_S_UMF2, ORC_S_UMF3, ORC_S_DUDT1, ORC_S_DUDT2, ORC_S_DUDT3,
  ORC_S_DTK, ORC_S_DTKO, ORC_S_AOA, ORC_S_NSCALAR = 16
};

typedef struct orc orc;

/*
 * dims[b*3+d]   : lxi, let, lze of block b (last index, points-1)
 * nbc[b*6+nn*2+ip], ncd[...], nsz[...] : BC code, neighbour block, sponge flag
 *                 for face ip (0 start, 1 end) of direction nn (0 xi,1 eta,2 zeta)
 * jct[j*12 + mbr*3 + {0,1,2}] : (block, ip, jp) of member mbr (0..3) of junction j
 * jvel[j*2 + {0,1}]           : members whose velocity is averaged
 */
/* NULL for a case the oracle cannot hold (lines shorter than their faces
 * need or longer than 8192 points, no memory); orc_last_error() says why */
orc *orc_create(int nb, const int32_t *dims, const int32_t *nbc,
                const int32_t *ncd, const int32_t *nsz, int njct,
                const int32_t *jct, const int32_t *jvel,
                const int32_t *iparams, const double *dparams);
void orc_destroy(orc *c);
const char *orc_last_error(void);
/* inflow: the Fourier modes (8 doubles each: k, a, phase, 0) the velocity
 * table is built from; call before orc_metrics */
void orc_set_inflow_modes(orc *c, int nmodes, const double *modes8);
/* inflow 3: the synthetic eddies the table is built from, 10 doubles each:
 * centre xit, yit, zit, sit(1:3), ait(1:3), mxc (src/SingleTurbulence3D.f90:
 * 86-115); call before orc_metrics */
void orc_set_inflow_eddies(orc *c, int nits, const double *eddies10);
/* forcing points of block b and their coefficients xafor(:,1:3) (forceup);
 * returns the count; either pointer may be NULL */
int orc_force_list(const orc *c, int b, int32_t *l, double *xafor3);
/* time interpolation of the inflow table at (timo, dtk): table rows il[5],
 * node times tl[5] and weights w[5]; returns 0 before tgusto */
int orc_inflow_weights(const orc *c, double timo, double dtk, int *il5, double *tl5, double *w5);
/* sponge points of block b (set by orc_metrics): their point indices l[q]
 * and inflow indices g[q] (-1: not an inflow point); returns the count;
 * either pointer may be NULL */
int orc_sponge_list(const orc *c, int b, int32_t *l, int32_t *g);
int orc_npoints(const orc *c, int b);
int orc_count(const orc *c, int b, int which); /* element count of an array */
double *orc_array(orc *c, int b, int which);
double *orc_scalars(orc *c);
/* surface statistics (src/rpt.f90:774-864, 1017-1056; src/rptPost.f90:870-925):
 * sample every `every` steps at the first stage (0: off), from the next step */
void orc_stats_start(orc *c, int every);
/* out8 = samples, first and last sample time, last Cl, Cd, their time means, window */
void orc_stats_forces(const orc *c, double *out8);
/* per wall point of block b (xi fastest): time mean and rms of p, cell area,
 * unit normal [3][n]; returns n (0: no wall face); any pointer may be NULL */
int orc_stats_wall(const orc *c, int b, double *mean_p, double *rms_p, double *area, double *wnor3);
void orc_set_threads(int n);

/* setup: coefficient tables (module-global in the reference) */
void orc_coefficients(orc *c, double *out_pbci, double *out_pbco,
                      double *out_pbcot, double *out_albef, double *out_fabc,
                      double *out_sbci);

/* grid -> metrics, face normals, GCBC weights, sponge (reads ss of every block) */
void orc_metrics(orc *c);
void orc_initialo(orc *c);

/* per-call drop-ins over all blocks */
/* these three return 0, or 1 for a selector out of range (orc_last_error) */
int orc_mpigo(orc *c, int nt, int nrt, int n45);
int orc_deriv(orc *c, int nn, int nz, int m);
int orc_filte(orc *c, int nn, int nz);

/* step pieces and whole steps */
void orc_filter_step(orc *c);
void orc_stage(orc *c, int nk);
/* the sections of a stage, in the order orc_stage runs them:
 *   BEGIN     frame velocity, primitives, time step (src/Main3D.f90:571-622)
 *   VISCOUS   stresses and heat fluxes into tv (:626-693; nothing when inviscid)
 *   FLUX      surface sample, flux divergence into de (:695-768)
 *   SPONGE    sponge and inflow target (src/SingleTurbulence3D.f90:332-374)
 *   GCBC      characteristic boundary/interface conditions (:778-869)
 *   RK        RK update of qa (:873-881)
 *   WALL      wall temperature and velocity (:884-895)
 *   JUNCTION  corner junctions (src/SingleTurbulence3D.f90:378-443)
 *   AVG       interface averaging (:897-952)
 * returns 1 for a stage or section out of range */
enum { ORC_P_BEGIN = 0, ORC_P_VISCOUS, ORC_P_FLUX, ORC_P_SPONGE, ORC_P_GCBC, ORC_P_RK,
       ORC_P_WALL, ORC_P_JUNCTION, ORC_P_AVG, ORC_P_NPART };
int orc_stage_part(orc *c, int nk, int part);
void orc_step(orc *c, int nsteps);

/* x**1.5 of the viscosity law, correctly rounded */
double orc_pow15(double x);
/* sin and cos, correctly rounded (the per-stage scalars) */
void orc_sincos(double x, double *sn, double *cs);

/* standalone numerics (setup routines) */
void orc_penta(int n, const double *albes, const double *albee, double alpha,
               double beta, double *xu, double *xl);
void orc_fcint(double fltk, double fltr, double *out5);
void orc_fcbcm(double fltk, double fltkbc, double *albef15, double *fam3,
               double *fbm3, double *fcm3);
void orc_sbcco(double fltk, double *pbci, double *pbco, double *pbcot);
void orc_mtrxi(int n, const double *ax, double *sx, double *last_ra0);
void orc_xtq2r(const double *q5, double p, const double *cm3,
               const double *umf3, double *xt25);
void orc_xtr2q(const double *q5, double p, const double *cm3,
               const double *umf3, double *xt25);
void orc_movef(int nsmf, double timf, double timo, double dtko, double dtk,
               const double *uoo3, double tp, double tps, double aoa0,
               double aoa1, double amachoo, double *umf3, double *dudtmf3,
               double *aoa);

#ifdef __cplusplus
}
#endif
#endif

/*
 * tp_oracle.c -- CPU restatement of the reference's time-marching hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is
 * compared against and the CPU baseline timed beside it; product code never
 * links or calls it.  It restates, operation by operation and in the same
 * evaluation order, the reference Fortran (all citations relative to the
 * reference tree):
 *
 *   function            restates                                   pinned by
 *   ------------------  -----------------------------------------  -----------------------------
 *   orc_penta           src/Subroutineso.f90:15-67                 golden vectors of the compiled reference
 *   orc_fcint           src/Subroutineso.f90:91-103                golden vectors + interior transfer function
 *   orc_fcbcm           src/Subroutineso.f90:71-87                 golden vectors
 *   orc_sbcco/mtrxi     src/Subroutineso.f90:107-188               golden vectors
 *   orc_movef (pitch)   src/Subroutineso.f90:192-259               golden vectors
 *   index map (indx3)   src/Subroutineso.f90:364-375               golden vectors
 *   GCBC weights        src/Main3D.f90:441-480                     golden vectors (mtrxi pivot state, sbci)
 *   mpigo               src/Subroutines3D.f90:16-80                closure identity: split line == unsplit line
 *   deriv               src/Subroutines3D.f90:84-141               dense solve, polynomial exactness,
 *                                                                  modified wavenumber
 *   filte               src/Subroutines3D.f90:145-206              constants invariant for every closure,
 *                                                                  interior transfer function
 *   eleme/xtq2r/xtr2q   src/Subroutines3D.f90:235-324              X_R * X_L = I, X_L A X_R diagonal
 *   metrics, normals    src/Main3D.f90:365-439                     exact metrics of affine grids, yaco = 1/J
 *   sponge mask         src/SingleTurbulence3D.f90:209-249         zone edges from the domain geometry (the
 *                                                                  single-precision 0.7e0 literal included),
 *                                                                  1 - sigma/szco = product of cos^2(pi r/2),
 *                                                                  lambda's ends; mutations caught
 *   extrabcc            src/SingleTurbulence3D.f90:447-464         slot points = the wall points inside slots
 *                                                                  of width 0.4 wlea centred a quarter pitch
 *                                                                  into each whole pitch of the span, along
 *                                                                  the first 2 wlea of chord; mutations caught
 *   stage sections (orc_stage_part; tests/test_oracle_stage_pins.py, every pin checked to fail on a
 *   planted mistake in tests/test_oracle_mutations.py):
 *     BEGIN             src/Main3D.f90:571-622                     CFL and viscous limits of a Cartesian grid
 *                                                                  in closed form; sin^2 ramp symmetry
 *     VISCOUS + FLUX    src/Main3D.f90:626-768                     analytic divergence of the Euler and
 *                                                                  Navier-Stokes fluxes of linear fields
 *                                                                  (moving frame too), quadratic shear,
 *                                                                  heat flux of a linear temperature; the
 *                                                                  LES subgrid stress by the model's form
 *                                                                  (its constants through the GPU parity)
 *     SPONGE            src/SingleTurbulence3D.f90:332-374         isentropic target with total enthalpy
 *                                                                  1/(gamma-1); relaxation linear in the
 *                                                                  departure; forcing (rpt.f90:758-769) in
 *                                                                  its window, closed-form coefficients
 *     GCBC              src/Main3D.f90:778-869                     incoming amplitudes zero after a far-field
 *                                                                  face; R+ = R- after a reflecting wall; the
 *                                                                  two copies of an interface point carry the
 *                                                                  same amplitudes after a whole stage, and so
 *                                                                  do partner wall slots (extrabcs) for the
 *                                                                  waves entering one and leaving the other
 *     RK                src/Main3D.f90:873-881                     4th-order Taylor polynomial on y' = ly
 *     WALL              src/Main3D.f90:884-895                     no-slip, gamma p / rho = T_wall, slots kept
 *     JUNCTION          src/SingleTurbulence3D.f90:378-443         copies agree: mean rho, E of four, momentum
 *                                                                  of the two wall blocks
 *     AVG               src/Main3D.f90:897-952                     interfaces continuous, two copies averaged
 *   inflow table        src/SingleTurbulence3D.f90:253-292         tests/test_inflow_cpu.py: periodic over one
 *     (Fourier modes)                                              period; one gust mode gives the closed-form
 *                                                                  travelling sinusoid A cos(theta - 2 pi it/ntz)
 *   inflow table        src/SingleTurbulence3D.f90:262-298         tests/test_eddy_forcing_cpu.py: one isotropic
 *     (eddies)                                                     eddy: u = f(r) a x d, perpendicular to d and
 *                                                                  a, convected with u_inf, scaled by cfit
 *   orc_inflow_weights  src/SingleTurbulence3D.f90:336-356         tests/test_inflow_cpu.py: exact for quartics
 *                                                                  in the node times, wrapped nodes included
 *   forceup             src/rpt.f90:728-753                        tests/test_eddy_forcing_cpu.py: points within
 *                                                                  rfor, Gaussian weight 1e-4 at rfor, the
 *                                                                  reference's reuse of ra1 as wavenumber
 *   initialo            src/SingleTurbulence3D.f90:179-205         uniform-flow test (trivial state)
 *   orc_sincos          sin, cos of src/Subroutineso.f90:192-259,  tests/test_rounding_cpu.py: correct rounding
 *                       Main3D.f90:610, rpt.f90:761-763            against 300-bit mpmath values
 *   orc_pow15 (x**1.5)  src/Main3D.f90:582                         tests/test_rounding_cpu.py: correct rounding
 *                                                                  decided in exact integer arithmetic; libm
 *                                                                  pow within one unit in the last place
 *
 *   orc_stage (order)   src/Main3D.f90:567-958                     after one whole stage: the interior residual of
 *                                                                  a viscous linear field (stresses before fluxes),
 *                                                                  zero incoming amplitudes at open faces (faces
 *                                                                  after fluxes), state = qo - dtk yaco de (update
 *                                                                  after the residual), wall conditions and
 *                                                                  continuous interfaces (wall, averaging after
 *                                                                  the update), junction corners with the wall
 *                                                                  pair's momentum (wall, junction, averaging
 *                                                                  in that order), interface amplitudes with
 *                                                                  the interface in the sponge (sponge before
 *                                                                  the characteristic section); mutations caught
 * The reference program itself cannot be built here (it needs an MPI
 * Fortran module).
 *
 * Layout: Fortran order, see tp_oracle.h.  All arithmetic is FP64 and is
 * compiled with -ffp-contract=off so that no fused multiply-add changes the
 * reference's rounding.
 */
#include "tp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: src/MainVar3D.f90:18-68 ---- */
#define LMD 11
#define LMF 8
#define LMP 11
#define MBCI 3
#define MAXLINE 8192

static const double zero = 0.0, half = 0.5, one = 1.0, two = 2.0, three = 3.0;
#define PI 3.14159265358979323846
static double gam = 1.4, gamm1, ham, hamm1, prndtl = 0.71, gamm1prndtli;
static double tprndtl = 0.99, tgamm1prndtli, two_three, sml = 1.0e-6;

static const double alpha = 0.5862704032801503, beta = 0.09549533555017055;
static const double aa = 0.6431406736919156, ab = 0.2586011023495066,
                    ac = 0.007140953479797375;
static const double beta20 = 0.03250008295108466, alpha21 = 0.3998040493524358,
                    alpha23 = 0.7719261277615860, beta24 = 0.1626635931256900;
static const double a20 = -0.1219006056449124, a21 = -0.6301651351188667,
                    a23 = 0.6521195063966084, a24 = 0.3938843551210350,
                    a25 = 0.01904944407973912, a26 = -0.001027260523947668;
static const double alpha10 = 0.08360703307833438, alpha12 = 2.058102869495757,
                    beta13 = 0.9704052014790193;
static const double a10 = -0.3177447290722621, a12 = -0.02807631929593225,
                    a13 = 1.593461635747659, a14 = 0.2533027046976367,
                    a15 = -0.03619652460174756, a16 = 0.004080281419108407;
static const double alpha01 = 5.912678614078549, beta02 = 3.775623951744012;
static const double a01 = -3.456878182643609, a02 = 5.839043358834730,
                    a03 = 1.015886726041007, a04 = -0.2246526470654333,
                    a05 = 0.08564940889936562, a06 = -0.01836710059356763;

static void init_constants(void) {
  gamm1 = gam - one;
  ham = one / gam;
  hamm1 = one / gamm1;
  gamm1prndtli = one / (gamm1 * prndtl);
  tgamm1prndtli = 1.0 / (gamm1 * tprndtl);
  two_three = two / three;
}

/* ---- data structures ---- */
typedef struct {
  int lxi, let, lze, N, lim;
  int ijk[3][3];  /* ijk[nn][0..2] == ijk(1..3, nn+1) */
  int nbsize[3];
  int nbc[3][2], ncd[3][2], ndf[3][2], nsz[3][2];
  double *qa, *qo, *de, *rr, *ss, *xim, *etm, *zem, *p, *yaco, *tv;
  double *drva[3], *drvb[3], *send[3], *recv[3], *cm[3];
  double *xu, *xl, *yu, *yl;
  int nwall;        /* wall face points (rpex entries) */
  double *rpex;
  int nsbcc;        /* GCBC weight entries */
  double *sbcc;
  int lsz;          /* sponge points - 1 */
  int *lcsz;
  double *asz, *bsz;
  int ltz;          /* inflow points - 1 */
  int *lctz;        /* their sponge-list positions */
  double *atz, *vito;   /* ramp; table vito(0:ltz, 0:ntz, 3) */
  double xbc_ra0;   /* state of the global ra0 seen by extrabcc */
  int nsurf, sside; /* wall surface points (eta face), wall side (0 j = 0, 1 j = let) */
  double *sarea, *snor, *sst;   /* area(0:lcwall), wnor(0:lcwall,3), stats p_prev, S1, S2 */
  int lfor;         /* forcing points - 1 */
  int *lcfor;       /* their point indices */
  double *xafor;    /* xafor(0:lfor, 3) */
} oblk;

struct orc {
  int nb;
  oblk *b;
  int njct;
  int *jct, *jvel;
  int nkrk, nviscous, nsmf, les, nextrabc, nextgcic;
  int inflow, nmodes, ntz;
  double tgusto, slit, gustw, *modes;
  int nits;            /* synthetic eddies (inflow 3): xit, yit, zit, sit(3), ait(3), mxc each */
  double *eddies, cfit;
  int forcing;         /* body force, src/rpt.f90:728-769 */
  double amfor, xfor, yfor, rfor, tsfor, tefor;
  double reoo, tempoo, amach1, amach2, amach3, wtemp, cfl, tmax, timf, fltk,
      fltkbc, dto, aoa0, aoa1, tps, tp, smago1, smago2, domlen, szth1, szth2,
      szxt, szco, span, wlew, wlea;
  double rhooo, poo, aoo, amachoo, srefoo, srefp1dre, uoo[3];
  /* coefficients (module globals of the reference) */
  double albed[2][3][5], albef[2][3][5], abc[3][6];
  double pbci[2][2][LMP + 1], pbco[2][2][LMP + 1], pbcot[2][2];
  double fam[3], fbm[3], fcm[3], alphf, betf, fa, fb, fc;
  double sbci[MBCI], cbca_ra0;
  /* scalars */
  double sc[ORC_S_NSCALAR];
  /* surface statistics: every, t0, t_prev, samples, Cl, Cd, trapezoid sums of Cl, Cd */
  double st[8];
};

#define IDX3(B, i, j, k, nn)                                                   \
  ((nn) == 0 ? ((k) * ((B)->let + 1) + (j)) * ((B)->lxi + 1) + (i)            \
   : (nn) == 1 ? ((j) * ((B)->let + 1) + (i)) * ((B)->lxi + 1) + (k)          \
               : ((i) * ((B)->let + 1) + (k)) * ((B)->lxi + 1) + (j))

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
static int nthreads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* setup numerics                                                          */
/* ---------------------------------------------------------------------- */

/* src/Subroutineso.f90:15-67.  xu[(c)*n + r] for c = 0..2, xl[(c)*n + r] c = 0..1;
 * albes/albee are albed(-2:2,0:2) as [row][col+2]. */
void orc_penta(int n, const double *albes, const double *albee, double alph,
               double bet, double *xu, double *xl) {
#define XU(r, c) xu[(c) * n + (r)]
#define XL(r, c) xl[(c) * n + (r)]
#define AS(c, r) albes[(r) * 5 + (c) + 2]
#define AE(c, r) albee[(r) * 5 + (c) + 2]
  int is = 0, ie = n - 1, i;
  for (i = is; i <= ie; i++) {
    XL(i, 0) = one; XL(i, 1) = one;
    XU(i, 0) = one; XU(i, 1) = one; XU(i, 2) = one;
  }
  i = is;
  XU(i, 0) = one;
  XU(i, 1) = AS(1, 0);
  XU(i, 2) = AS(2, 0);
  i = is + 1;
  XL(i, 1) = AS(-1, 1) * XU(i - 1, 0);
  XU(i, 0) = one / (one - XU(i - 1, 1) * XL(i, 1));
  XU(i, 1) = AS(1, 1) - XU(i - 1, 2) * XL(i, 1);
  XU(i, 2) = AS(2, 1);
  i = is + 2;
  XL(i, 0) = AS(-2, 2) * XU(i - 2, 0);
  XL(i, 1) = (AS(-1, 2) - XU(i - 2, 1) * XL(i, 0)) * XU(i - 1, 0);
  XU(i, 0) = one / (one - XU(i - 2, 2) * XL(i, 0) - XU(i - 1, 1) * XL(i, 1));
  XU(i, 1) = AS(1, 2) - XU(i - 1, 2) * XL(i, 1);
  XU(i, 2) = AS(2, 2);
  for (i = is + 3; i <= ie - 3; i++) {
    XL(i, 0) = bet * XU(i - 2, 0);
    XL(i, 1) = (alph - XU(i - 2, 1) * XL(i, 0)) * XU(i - 1, 0);
    XU(i, 0) = one / (one - XU(i - 2, 2) * XL(i, 0) - XU(i - 1, 1) * XL(i, 1));
    XU(i, 1) = alph - XU(i - 1, 2) * XL(i, 1);
    XU(i, 2) = bet;
  }
  i = ie - 2;
  XL(i, 0) = AE(2, 2) * XU(i - 2, 0);
  XL(i, 1) = (AE(1, 2) - XU(i - 2, 1) * XL(i, 0)) * XU(i - 1, 0);
  XU(i, 0) = one / (one - XU(i - 2, 2) * XL(i, 0) - XU(i - 1, 1) * XL(i, 1));
  XU(i, 1) = AE(-1, 2) - XU(i - 1, 2) * XL(i, 1);
  XU(i, 2) = AE(-2, 2);
  i = ie - 1;
  XL(i, 0) = AE(2, 1) * XU(i - 2, 0);
  XL(i, 1) = (AE(1, 1) - XU(i - 2, 1) * XL(i, 0)) * XU(i - 1, 0);
  XU(i, 0) = one / (one - XU(i - 2, 2) * XL(i, 0) - XU(i - 1, 1) * XL(i, 1));
  XU(i, 1) = AE(-1, 1) - XU(i - 1, 2) * XL(i, 1);
  i = ie;
  XL(i, 0) = AE(2, 0) * XU(i - 2, 0);
  XL(i, 1) = (AE(1, 0) - XU(i - 2, 1) * XL(i, 0)) * XU(i - 1, 0);
  XU(i, 0) = one / (one - XU(i - 2, 2) * XL(i, 0) - XU(i - 1, 1) * XL(i, 1));
  for (i = is; i <= ie; i++) {
    XU(i, 1) = XU(i, 1) * XU(i, 0);
    XU(i, 2) = XU(i, 2) * XU(i, 0);
  }
#undef XU
#undef XL
#undef AS
#undef AE
}

/* src/Subroutineso.f90:91-103; out5 = alphz, betz, za, zb, zc */
void orc_fcint(double fltk, double fltr, double *o) {
  double c1 = cos(fltk), c2 = cos(2 * fltk), c3 = cos(3 * fltk);
  double fctr = 1 / (30 + 5 * (7 - 16 * fltr) * c1 + 2 * (1 + 8 * fltr) * c2 - 3 * c3);
  double ch = cos(half * fltk), ch2 = ch * ch;
  o[0] = fctr * (20 * (2 * fltr - 1) - 30 * c1 + 12 * (2 * fltr - 1) * c2 - 2 * c3);
  o[1] = half * fctr * (2 * (13 - 8 * fltr) + (33 - 48 * fltr) * c1 + 6 * c2 - c3);
  o[2] = 60 * (1 - fltr) * (ch2 * ch2) * fctr;
  o[3] = -2 * o[2] / 5;
  o[4] = o[2] / 15;
}

/* src/Subroutineso.f90:71-87; albef15 = albef(-2:2,0:2) as [row][col+2] */
void orc_fcbcm(double fltk, double fltkbc, double *albef, double *fam,
               double *fbm, double *fcm) {
  double z[5];
  double res = (fltkbc - fltk) / 9, ra2 = fltk + res, ra1 = fltk + 4 * res,
         ra0 = fltk + 9 * res, fctr;
  orc_fcint(ra0, half, z);
  fctr = one / (one + z[0] / 2 + z[1] / 4);
  albef[0] = zero; albef[1] = zero; albef[2] = one;
  albef[3] = z[0] * fctr; albef[4] = z[1] * fctr;
  fam[0] = z[2] * fctr; fbm[0] = z[3] * fctr; fcm[0] = z[4] * fctr;
  orc_fcint(ra1, half, z);
  albef[5] = zero; albef[6] = z[0] + z[1] / 2; albef[7] = one;
  albef[8] = z[0]; albef[9] = z[1];
  fam[1] = z[2]; fbm[1] = z[3]; fcm[1] = z[4];
  orc_fcint(ra2, half, z);
  albef[10] = z[1]; albef[11] = z[0]; albef[12] = one;
  albef[13] = z[0]; albef[14] = z[1];
  fam[2] = z[2]; fbm[2] = z[3]; fcm[2] = z[4];
}

/* src/Subroutineso.f90:163-188 (Gauss-Jordan inverse with row pivoting).
 * ax, sx row-major n x n.  last_ra0 receives the final pivot reciprocal
 * (the global ra0 the reference leaves behind). */
void orc_mtrxi(int n, const double *ax, double *sx, double *last_ra0) {
  double *rx = (double *)malloc(sizeof(double) * n * n);
  double *temp = (double *)malloc(sizeof(double) * n);
  int *ipvt = (int *)malloc(sizeof(int) * n);
  double ra0 = 0, ra1;
  int i, j, r, m;
  memcpy(rx, ax, sizeof(double) * n * n);
  for (i = 0; i < n; i++) ipvt[i] = i;
  for (i = 0; i < n; i++) {
    double best = -1;
    m = i;
    for (r = i; r < n; r++) {
      double v = fabs(rx[r * n + i]);
      if (v > best) { best = v; m = r; }
    }
    if (m != i) {
      int t = ipvt[m]; ipvt[m] = ipvt[i]; ipvt[i] = t;
      for (j = 0; j < n; j++) {
        double tv = rx[m * n + j]; rx[m * n + j] = rx[i * n + j]; rx[i * n + j] = tv;
      }
    }
    ra0 = one / rx[i * n + i];
    for (r = 0; r < n; r++) temp[r] = rx[r * n + i];
    for (j = 0; j < n; j++) {
      ra1 = ra0 * rx[i * n + j];
      for (r = 0; r < n; r++) rx[r * n + j] = rx[r * n + j] - ra1 * temp[r];
      rx[i * n + j] = ra1;
    }
    for (r = 0; r < n; r++) rx[r * n + i] = -ra0 * temp[r];
    rx[i * n + i] = ra0;
  }
  for (r = 0; r < n; r++)
    for (j = 0; j < n; j++) sx[r * n + ipvt[j]] = rx[r * n + j];
  if (last_ra0) *last_ra0 = ra0;
  free(rx); free(temp); free(ipvt);
}

static void matmul(int n, const double *a, const double *b, double *c) {
  for (int r = 0; r < n; r++)
    for (int col = 0; col < n; col++) {
      double s = 0;
      for (int k = 0; k < n; k++) s = s + a[r * n + k] * b[k * n + col];
      c[r * n + col] = s;
    }
}

/* src/Subroutineso.f90:107-159.  pbci/pbco [nt][row][0..LMP], pbcot[nt][row]
 * (pbcot is Main3D.f90:310-312). */
void orc_sbcco(double fltk, double *pbci, double *pbco, double *pbcot) {
  double abc[3][6] = {{a01, a02, a03, a04, a05, a06},
                      {a10, a12, a13, a14, a15, a16},
                      {a20, a21, a23, a24, a25, a26}};
  double albed0[3][5] = {{zero, zero, one, alpha01, beta02},
                         {zero, alpha10, one, alpha12, beta13},
                         {beta20, alpha21, one, alpha23, beta24}};
  memset(pbci, 0, sizeof(double) * 2 * 2 * (LMP + 1));
  memset(pbco, 0, sizeof(double) * 2 * 2 * (LMP + 1));
  for (int nt = 0; nt <= 1; nt++) {
    int lp = 2 * nt - 1;
    int ll = nt == 0 ? LMD : LMF;
    int n = 2 * (ll + 1);
    double *ax = (double *)calloc(n * n, sizeof(double));
    double *bx = (double *)calloc(n * n, sizeof(double));
    double *rx = (double *)calloc(n * n, sizeof(double));
    double *sx = (double *)calloc(n * n, sizeof(double));
    double *tx = (double *)calloc(n * n, sizeof(double));
#define AX(r, c) ax[((r) - 1) * n + (c) - 1]
#define BX(r, c) bx[((r) - 1) * n + (c) - 1]
#define RX(r, c) rx[((r) - 1) * n + (c) - 1]
    if (nt == 0) {
      double s0 = 0, s1 = 0, s2 = 0;
      for (int c = 0; c < 6; c++) s0 = s0 + abc[0][c];
      for (int c = 0; c < 6; c++) s1 = s1 + abc[1][c];
      for (int c = 0; c < 6; c++) s2 = s2 + abc[2][c];
      for (int c = 0; c < 3; c++) AX(1, 1 + c) = albed0[0][2 + c];
      { double v[7] = {-s0, a01, a02, a03, a04, a05, a06};
        for (int c = 0; c < 7; c++) BX(1, 1 + c) = v[c]; }
      for (int c = 0; c < 4; c++) AX(2, 1 + c) = albed0[1][1 + c];
      { double v[7] = {a10, -s1, a12, a13, a14, a15, a16};
        for (int c = 0; c < 7; c++) BX(2, 1 + c) = v[c]; }
      for (int c = 0; c < 5; c++) AX(3, 1 + c) = albed0[2][c];
      { double v[7] = {a20, a21, -s2, a23, a24, a25, a26};
        for (int c = 0; c < 7; c++) BX(3, 1 + c) = v[c]; }
      for (int i = 4; i <= n - 3; i++) {
        double va[5] = {beta, alpha, one, alpha, beta};
        double vb[7] = {-ac, -ab, -aa, zero, aa, ab, ac};
        for (int c = 0; c < 5; c++) AX(i, i - 2 + c) = va[c];
        for (int c = 0; c < 7; c++) BX(i, i - 3 + c) = vb[c];
      }
    } else {
      double z[5], zv[5], al, be, fctr, ra0, ra1, ra2;
      orc_fcint(fltk, half, z);
      al = z[0]; be = z[1];
      zv[0] = z[3] + 5 * z[4]; zv[1] = z[2] - 10 * z[4]; zv[2] = z[2] - 5 * z[4];
      zv[3] = z[3] + z[4]; zv[4] = z[4];
      {
        double al2 = al * al, al3 = al2 * al, be2 = be * be, be3 = be2 * be;
        fctr = 1 / (1 + 5 * al3 + al2 * (8 + 22 * be) + be * (5 + 4 * be + 60 * be2) +
                    5 * al * (1 + be * (3 + 10 * be)));
        ra0 = (al * (1 + al) * (1 + 4 * al) + 2 * al * (7 + 3 * al) * be +
               24 * (1 - al) * be2 - 80 * be3) * fctr;
        ra1 = (al3 + be + 14 * al2 * be + 60 * be3 + al * be * (3 + 46 * be)) * fctr;
        AX(1, 1) = one; AX(1, 2) = ra0; AX(1, 3) = ra1;
        fctr = 1 / (1 + al2 * (9 - 5 * be) + al * (5 + (35 - 29 * be) * be) +
                    be * (5 + 6 * (9 - 10 * be) * be));
        ra0 = (9 * al3 + 10 * be2 * (8 * be - 1) + 5 * al2 * (1 + 8 * be) +
               al * (1 + be * (4 + 81 * be))) * fctr;
        ra1 = (al * (1 + al * (5 + 9 * al)) + al * (5 + 36 * al) * be +
               (55 * al - 1) * be2 + 10 * be3) * fctr;
        ra2 = be * (1 + al * (5 + 9 * al) + 5 * be + 35 * al * be + 50 * be2) * fctr;
        AX(2, 1) = ra0; AX(2, 2) = one; AX(2, 3) = ra1; AX(2, 4) = ra2;
      }
      {
        double va[5] = {be, al, one, al, be};
        double s = 0;
        for (int c = 0; c < 5; c++) s = s + zv[c];
        double vb[6] = {zv[0], zv[1], -s, zv[2], zv[3], zv[4]};
        for (int c = 0; c < 5; c++) AX(3, 1 + c) = va[c];
        for (int c = 0; c < 6; c++) BX(3, 1 + c) = vb[c];
      }
      for (int i = 4; i <= n - 3; i++) {
        double va[5] = {be, al, one, al, be};
        double vb[7] = {z[4], z[3], z[2], -2 * (z[2] + z[3] + z[4]), z[2], z[3], z[4]};
        for (int c = 0; c < 5; c++) AX(i, i - 2 + c) = va[c];
        for (int c = 0; c < 7; c++) BX(i, i - 3 + c) = vb[c];
      }
    }
    for (int c = 1; c <= n; c++) {
      AX(n - 2, n + 1 - c) = AX(3, c); BX(n - 2, n + 1 - c) = lp * BX(3, c);
      AX(n - 1, n + 1 - c) = AX(2, c); BX(n - 1, n + 1 - c) = lp * BX(2, c);
      AX(n, n + 1 - c) = AX(1, c);     BX(n, n + 1 - c) = lp * BX(1, c);
    }
    orc_mtrxi(n, ax, sx, NULL);
    memcpy(rx, ax, sizeof(double) * n * n);
    {
      int i = n / 2 - 1;
      RX(i, i + 2) = 0; RX(i + 1, i + 2) = 0; RX(i + 1, i + 3) = 0;
      i = n / 2 + 2;
      RX(i, i - 2) = 0; RX(i - 1, i - 2) = 0; RX(i - 1, i - 3) = 0;
    }
    matmul(n, sx, bx, tx);
    matmul(n, rx, tx, ax);
    for (int row = 0; row <= 1; row++) {
      int i = n / 2 + 1 + row;
      for (int c = 0; c <= ll; c++) {
        pbco[(nt * 2 + row) * (LMP + 1) + (ll - c)] = AX(i, 1 + c);
        pbci[(nt * 2 + row) * (LMP + 1) + c] = AX(i, ll + 2 + c);
      }
    }
#undef AX
#undef BX
#undef RX
    free(ax); free(bx); free(rx); free(sx); free(tx);
  }
  for (int nt = 0; nt <= 1; nt++)
    for (int j = 0; j <= 1; j++) {
      int ii = LMD + nt * (LMF - LMD);
      double s = 0;
      for (int i = 0; i <= ii; i++) s = s + pbco[(nt * 2 + j) * (LMP + 1) + i];
      pbcot[nt * 2 + j] = s;
    }
}

/* ---- sin and cos rounded correctly ----
 * The per-stage scalars (moving frame, time-step ramp, body force) follow
 * the reference's sin and cos, whose last bit its math library leaves open;
 * the oracle fixes the correctly rounded values.  Method: x - k pi/2 with
 * pi/2 split into three doubles (Cody and Waite) in double-double arithmetic
 * (Dekker's two-sum and an fma two-product), then the Taylor series of sin
 * and cos to the 1/29! term, summed by Horner's rule in double-double with
 * the inverse factorials as double-double constants; the leading double of
 * the normalised sum is the result.  |x| >= 2^20 (and NaN) is left to libm. */
typedef struct { double h, l; } ddbl;
static ddbl dd_fast(double a, double b) { double s = a + b; ddbl r = {s, b - (s - a)}; return r; }
static ddbl dd_sum(double a, double b) {
  double s = a + b, bb = s - a;
  ddbl r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static ddbl dd_add(ddbl a, ddbl b) { ddbl s = dd_sum(a.h, b.h); return dd_fast(s.h, s.l + (a.l + b.l)); }
static ddbl dd_mul(ddbl a, ddbl b) {
  double p = a.h * b.h, e = fma(a.h, b.h, -p);
  return dd_fast(p, e + (a.h * b.l + a.l * b.h));
}
/* 1/n!, n = 0..29, as double-double */
static const double INV_FACT[30][2] = {
    {1.0, 0.0}, {1.0, 0.0}, {0.5, 0.0}, {0.16666666666666666, 9.25185853854297e-18},
    {0.041666666666666664, 2.3129646346357427e-18}, {0.008333333333333333, 1.1564823173178714e-19},
    {0.001388888888888889, -5.300543954373577e-20}, {0.0001984126984126984, 1.7209558293420705e-22},
    {2.48015873015873e-05, 2.1511947866775882e-23}, {2.7557319223985893e-06, -1.858393274046472e-22},
    {2.755731922398589e-07, 2.3767714622250297e-23}, {2.505210838544172e-08, -1.448814070935912e-24},
    {2.08767569878681e-09, -1.20734505911326e-25}, {1.6059043836821613e-10, 1.2585294588752098e-26},
    {1.1470745597729725e-11, 2.0655512752830745e-28}, {7.647163731819816e-13, 7.03872877733453e-30},
    {4.779477332387385e-14, 4.399205485834081e-31}, {2.8114572543455206e-15, 1.6508842730861433e-31},
    {1.5619206968586225e-16, 1.1910679660273754e-32}, {8.22063524662433e-18, 2.2141894119604265e-34},
    {4.110317623312165e-19, 1.4412973378659527e-36}, {1.9572941063391263e-20, -1.3643503830087908e-36},
    {8.896791392450574e-22, -7.911402614872376e-38}, {3.868170170630684e-23, -8.843177655482344e-40},
    {1.6117375710961184e-24, -3.6846573564509766e-41}, {6.446950284384474e-26, -1.9330404233703465e-42},
    {2.4795962632247976e-27, -1.2953730964765229e-43}, {9.183689863795546e-29, 1.4303150396787322e-45},
    {3.279889237069838e-30, 1.5117542744029879e-46}, {1.1309962886447716e-31, 1.0498015412959506e-47}};
static ddbl inv_fact(int n, int neg) {
  ddbl r = {neg ? -INV_FACT[n][0] : INV_FACT[n][0], neg ? -INV_FACT[n][1] : INV_FACT[n][1]};
  return r;
}
void orc_sincos(double x, double *sn, double *cs) {
  if (!(fabs(x) < 0x1p20)) { *sn = sin(x); *cs = cos(x); return; }
  if (x == 0.0) { *sn = x; *cs = 1.0; return; }
  const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  const double k = nearbyint(x * 0.6366197723675814);
  const double p = k * P1, pe = fma(k, P1, -p);
  ddbl r = dd_sum(x, -p);
  ddbl t = {-pe, 0.0};
  r = dd_add(r, t);
  const double q = k * P2, qe = fma(k, P2, -q);
  t.h = -q; t.l = -qe;
  r = dd_add(r, t);
  t.h = -(k * P3); t.l = 0.0;
  r = dd_add(r, t);
  const ddbl z = dd_mul(r, r);
  ddbl ps = inv_fact(29, 1), pc = inv_fact(28, 0);   /* n = 14: -z^14/29!, +z^14/28! */
  for (int n = 13; n >= 0; n--) {
    ps = dd_add(dd_mul(ps, z), inv_fact(2 * n + 1, n & 1));
    pc = dd_add(dd_mul(pc, z), inv_fact(2 * n, n & 1));
  }
  const double S = dd_mul(r, ps).h, C = pc.h;
  switch (((long long)k) & 3) {
    case 0: *sn = S; *cs = C; break;
    case 1: *sn = C; *cs = -S; break;
    case 2: *sn = -S; *cs = -C; break;
    default: *sn = -C; *cs = S; break;
  }
}
static double crsin(double x) { double s, c; orc_sincos(x, &s, &c); return s; }
static double crcos(double x) { double s, c; orc_sincos(x, &s, &c); return c; }

/* src/Subroutineso.f90:192-259 */
void orc_movef(int nsmf, double timf, double timo, double dtko, double dtk,
               const double *uoo, double tp, double tps, double aoa0,
               double aoa1, double amachoo, double *umf, double *dudtmf,
               double *aoa) {
  if (nsmf == 0) {
    double ra0 = PI / timf;
    double ra1 = ra0 * fmin(timo, timf), ra2 = ra0 * fmin(timo + dtko, timf);
    double fctr = one - crcos(ra1), dfdt = ra0 * crsin(ra2);
    double progmf = half * (fctr + dtk * dfdt);
    for (int d = 0; d < 3; d++) umf[d] = progmf * uoo[d];
    fctr = crsin(ra1);
    dfdt = ra0 * crcos(ra2);
    progmf = half * ra0 * (fctr + dtk * dfdt);
    for (int d = 0; d < 3; d++) dudtmf[d] = progmf * uoo[d];
    if (tp > 0 && timo + dtk > tps) {
      double daoa = (aoa1 - aoa0) * PI / 180, raoa0 = aoa0 * PI / 180;
      double ra3 = timo - tps, dtsdt = PI / tp;
      double ts = fmin(dtsdt * ra3, PI), tsdt = fmin(dtsdt * (ra3 + dtko), PI);
      double aoat = raoa0 + half * daoa * (1 - crcos(ts));
      double aoatdt = raoa0 + half * daoa * (1 - crcos(tsdt));
      double f = half * daoa * dtsdt;
      double ut[3] = {amachoo * crcos(aoat), amachoo * crsin(aoat), zero};
      double du[3] = {-(amachoo * f * crsin(aoatdt) * crsin(tsdt)),
                      amachoo * f * crcos(aoatdt) * crsin(tsdt), zero};
      for (int d = 0; d < 3; d++) umf[d] = ut[d] + dtk * du[d];
      double dd[3] = {-(amachoo * f * crsin(aoat) * crsin(ts)),
                      amachoo * f * crcos(aoat) * crsin(ts), zero};
      double st = crsin(tsdt);
      double d2[3] = {
          -(amachoo * f * (crcos(aoatdt) * f * (st * st) + dtsdt * crsin(aoatdt) * crcos(tsdt))),
          -(amachoo * f * (crsin(aoatdt) * f * (st * st) - dtsdt * crcos(aoatdt) * crcos(tsdt))),
          zero};
      for (int d = 0; d < 3; d++) dudtmf[d] = dd[d] + dtk * d2[d];
      if (aoa) *aoa = atan(umf[1] / umf[0]) * 180 / PI;
    }
  } else {
    for (int d = 0; d < 3; d++) { umf[d] = uoo[d]; dudtmf[d] = zero; }
  }
}

/* ---- characteristic algebra: src/Subroutines3D.f90:235-324 ---- */
typedef struct { double rhoi, ao, aoi, ve[3], hv2, vn, vs; } elem_t;

static inline void eleme(const double *q, double p, const double *cm,
                         const double *umf, elem_t *e) {
  e->rhoi = one / q[0];
  e->ao = sqrt(gam * e->rhoi * p);
  e->aoi = one / e->ao;
  e->ve[0] = e->rhoi * q[1]; e->ve[1] = e->rhoi * q[2]; e->ve[2] = e->rhoi * q[3];
  e->hv2 = half * (e->ve[0] * e->ve[0] + e->ve[1] * e->ve[1] + e->ve[2] * e->ve[2]);
  e->vn = cm[0] * e->ve[0] + cm[1] * e->ve[1] + cm[2] * e->ve[2];
  e->vs = cm[0] * umf[0] + cm[1] * umf[1] + cm[2] * umf[2];
}

static inline void xtq2r(const elem_t *e, const double *cm, double xt[5][5]) {
  double ho = gamm1 * e->aoi * e->aoi, bo = 1 - ho * e->hv2, co = e->aoi * e->vn;
  double dm[3] = {e->aoi * cm[0], e->aoi * cm[1], e->aoi * cm[2]};
  double rv[3] = {ho * e->ve[0], ho * e->ve[1], ho * e->ve[2]};
  const double *ve = e->ve;
  xt[0][0] = bo * cm[0] + dm[1] * ve[2] - dm[2] * ve[1];
  xt[0][1] = cm[0] * rv[0];
  xt[0][2] = cm[0] * rv[1] + dm[2];
  xt[0][3] = cm[0] * rv[2] - dm[1];
  xt[0][4] = -ho * cm[0];
  xt[1][0] = bo * cm[1] + dm[2] * ve[0] - dm[0] * ve[2];
  xt[1][1] = cm[1] * rv[0] - dm[2];
  xt[1][2] = cm[1] * rv[1];
  xt[1][3] = cm[1] * rv[2] + dm[0];
  xt[1][4] = -ho * cm[1];
  xt[2][0] = bo * cm[2] + dm[0] * ve[1] - dm[1] * ve[0];
  xt[2][1] = cm[2] * rv[0] + dm[1];
  xt[2][2] = cm[2] * rv[1] - dm[0];
  xt[2][3] = cm[2] * rv[2];
  xt[2][4] = -ho * cm[2];
  xt[3][0] = one - bo - co;
  xt[3][1] = dm[0] - rv[0];
  xt[3][2] = dm[1] - rv[1];
  xt[3][3] = dm[2] - rv[2];
  xt[3][4] = ho;
  xt[4][0] = one - bo + co;
  xt[4][1] = -dm[0] - rv[0];
  xt[4][2] = -dm[1] - rv[1];
  xt[4][3] = -dm[2] - rv[2];
  xt[4][4] = ho;
}

static inline void xtr2q(const elem_t *e, const double *cm, double xt[5][5]) {
  double bo = e->hv2 + hamm1 * e->ao * e->ao, co = e->ao * e->vn;
  double dm[3] = {e->ao * cm[0], e->ao * cm[1], e->ao * cm[2]};
  const double *ve = e->ve;
  xt[0][0] = cm[0]; xt[0][1] = cm[1]; xt[0][2] = cm[2]; xt[0][3] = half; xt[0][4] = half;
  xt[1][0] = cm[0] * ve[0];
  xt[1][1] = cm[1] * ve[0] - dm[2];
  xt[1][2] = cm[2] * ve[0] + dm[1];
  xt[1][3] = half * (ve[0] + dm[0]);
  xt[1][4] = xt[1][3] - dm[0];
  xt[2][0] = cm[0] * ve[1] + dm[2];
  xt[2][1] = cm[1] * ve[1];
  xt[2][2] = cm[2] * ve[1] - dm[0];
  xt[2][3] = half * (ve[1] + dm[1]);
  xt[2][4] = xt[2][3] - dm[1];
  xt[3][0] = cm[0] * ve[2] - dm[1];
  xt[3][1] = cm[1] * ve[2] + dm[0];
  xt[3][2] = cm[2] * ve[2];
  xt[3][3] = half * (ve[2] + dm[2]);
  xt[3][4] = xt[3][3] - dm[2];
  xt[4][0] = e->hv2 * cm[0] + dm[2] * ve[1] - dm[1] * ve[2];
  xt[4][1] = e->hv2 * cm[1] + dm[0] * ve[2] - dm[2] * ve[0];
  xt[4][2] = e->hv2 * cm[2] + dm[1] * ve[0] - dm[0] * ve[1];
  xt[4][3] = half * (bo + co);
  xt[4][4] = xt[4][3] - co;
}

void orc_xtq2r(const double *q5, double p, const double *cm3, const double *umf3,
               double *xt25) {
  elem_t e; double xt[5][5];
  init_constants();
  eleme(q5, p, cm3, umf3, &e);
  xtq2r(&e, cm3, xt);
  memcpy(xt25, xt, sizeof(xt));
}
void orc_xtr2q(const double *q5, double p, const double *cm3, const double *umf3,
               double *xt25) {
  elem_t e; double xt[5][5];
  init_constants();
  eleme(q5, p, cm3, umf3, &e);
  xtr2q(&e, cm3, xt);
  memcpy(xt25, xt, sizeof(xt));
}

/* ---------------------------------------------------------------------- */
/* case construction                                                        */
/* ---------------------------------------------------------------------- */

/* Errors: the functions that take caller input return NULL or a nonzero code
 * and leave a message here instead of ending the process, so that a test
 * that feeds the oracle a bad case fails alone (orc_last_error). */
static char g_err[256];
static int g_oom;
const char *orc_last_error(void) { return g_err; }
static void set_err(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

static double *dalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) { g_oom = 1; set_err("oracle: out of memory"); }
  return p;
}

static void setup_coefficients(orc *c) {
  /* src/Main3D.f90:71-73, 295-312 */
  double abc[3][6] = {{a01, a02, a03, a04, a05, a06},
                      {a10, a12, a13, a14, a15, a16},
                      {a20, a21, a23, a24, a25, a26}};
  memcpy(c->abc, abc, sizeof(abc));
  double ab0[3][5] = {{zero, zero, one, alpha01, beta02},
                      {zero, alpha10, one, alpha12, beta13},
                      {beta20, alpha21, one, alpha23, beta24}};
  double ab1[3][5] = {{zero, zero, one, alpha, beta},
                      {zero, alpha, one, alpha, beta},
                      {beta, alpha, one, alpha, beta}};
  memcpy(c->albed[0], ab0, sizeof(ab0));
  memcpy(c->albed[1], ab1, sizeof(ab1));
  orc_fcbcm(c->fltk, c->fltkbc, &c->albef[0][0][0], c->fam, c->fbm, c->fcm);
  double z[5];
  orc_fcint(c->fltk, half, z);
  c->alphf = z[0]; c->betf = z[1]; c->fa = z[2]; c->fb = z[3]; c->fc = z[4];
  double af1[3][5] = {{zero, zero, one, c->alphf, c->betf},
                      {zero, c->alphf, one, c->alphf, c->betf},
                      {c->betf, c->alphf, one, c->alphf, c->betf}};
  memcpy(c->albef[1], af1, sizeof(af1));
  orc_sbcco(c->fltk, &c->pbci[0][0][0], &c->pbco[0][0][0], &c->pbcot[0][0]);
  /* GCBC weights, src/Main3D.f90:443-459 (mbci = 3) */
  double cbca[MBCI * MBCI] = {0}, cbcs[MBCI * MBCI], rbci[MBCI];
  cbca[0 * 3 + 0] = c->albed[0][0][3]; cbca[0 * 3 + 1] = c->albed[0][0][4];
  cbca[1 * 3 + 0] = c->albed[0][1][2]; cbca[1 * 3 + 1] = c->albed[0][1][3];
  cbca[1 * 3 + 2] = c->albed[0][1][4];
  cbca[2 * 3 + 0] = c->albed[0][2][1]; cbca[2 * 3 + 1] = c->albed[0][2][2];
  cbca[2 * 3 + 2] = c->albed[0][2][3];
  rbci[0] = one; rbci[1] = c->albed[0][1][1]; rbci[2] = c->albed[0][2][0];
  orc_mtrxi(MBCI, cbca, cbcs, &c->cbca_ra0);
  for (int i = 0; i < MBCI; i++) {
    double s = 0;
    for (int k = 0; k < MBCI; k++) s = s + cbcs[i * 3 + k] * rbci[k];
    c->sbci[i] = -s;
  }
  double fctr = PI / (MBCI + 1), res = zero;
  for (int i = 0; i < MBCI; i++) {
    res = res + one;
    c->sbci[i] = half * c->sbci[i] * (one + cos(res * fctr));
  }
}

orc *orc_create(int nb, const int32_t *dims, const int32_t *nbc,
                const int32_t *ncd, const int32_t *nsz, int njct,
                const int32_t *jct, const int32_t *jvel,
                const int32_t *ip, const double *dp) {
  init_constants();
  g_oom = 0;
  g_err[0] = 0;
  if (nb < 1) { set_err("oracle: a case needs at least one block"); return NULL; }
  for (int b = 0; b < nb; b++)
    for (int nn = 0; nn < 3; nn++)
      if (dims[b * 3 + nn] < 1 || dims[b * 3 + nn] + 1 > MAXLINE) {
        snprintf(g_err, sizeof(g_err), "oracle: block %d direction %d has %d points (1 < points <= %d)",
                 b, nn + 1, dims[b * 3 + nn] + 1, MAXLINE);
        return NULL;
      }
  orc *c = (orc *)calloc(1, sizeof(orc));
  c->nb = nb;
  c->b = (oblk *)calloc(nb, sizeof(oblk));
  c->njct = njct;
  c->jct = (int *)calloc(njct * 12 + 1, sizeof(int));
  c->jvel = (int *)calloc(njct * 2 + 1, sizeof(int));
  if (njct) {
    memcpy(c->jct, jct, sizeof(int) * 12 * njct);
    memcpy(c->jvel, jvel, sizeof(int) * 2 * njct);
  }
  c->nkrk = ip[ORC_I_NKRK]; c->nviscous = ip[ORC_I_NVISCOUS];
  c->nsmf = ip[ORC_I_NSMF]; c->les = ip[ORC_I_LES];
  c->nextrabc = ip[ORC_I_NEXTRABC]; c->nextgcic = ip[ORC_I_NEXTGCIC];
  c->inflow = ip[ORC_I_INFLOW]; c->ntz = ip[ORC_I_NTZ];
  c->tgusto = dp[ORC_D_TGUSTO]; c->slit = dp[ORC_D_SLIT]; c->gustw = dp[ORC_D_GUSTW];
  c->cfit = dp[ORC_D_CFIT];
  c->forcing = ip[ORC_I_FORCING];
  c->amfor = dp[ORC_D_AMFOR]; c->xfor = dp[ORC_D_XFOR]; c->yfor = dp[ORC_D_YFOR];
  c->rfor = dp[ORC_D_RFOR]; c->tsfor = dp[ORC_D_TSFOR]; c->tefor = dp[ORC_D_TEFOR];
  c->reoo = dp[ORC_D_REOO]; c->tempoo = dp[ORC_D_TEMPOO];
  c->amach1 = dp[ORC_D_AMACH1]; c->amach2 = dp[ORC_D_AMACH2]; c->amach3 = dp[ORC_D_AMACH3];
  c->wtemp = dp[ORC_D_WTEMP]; c->cfl = dp[ORC_D_CFL]; c->tmax = dp[ORC_D_TMAX];
  c->timf = dp[ORC_D_TIMF]; c->dto = dp[ORC_D_DTO];
  c->aoa0 = dp[ORC_D_AOA0]; c->aoa1 = dp[ORC_D_AOA1]; c->tps = dp[ORC_D_TPS]; c->tp = dp[ORC_D_TP];
  c->smago1 = dp[ORC_D_SMAGO1]; c->smago2 = dp[ORC_D_SMAGO2];
  c->domlen = dp[ORC_D_DOMLEN]; c->szth1 = dp[ORC_D_SZTH1]; c->szth2 = dp[ORC_D_SZTH2];
  c->szxt = dp[ORC_D_SZXT]; c->szco = dp[ORC_D_SZCO]; c->span = dp[ORC_D_SPAN];
  c->wlew = dp[ORC_D_WLEW]; c->wlea = dp[ORC_D_WLEA];
  /* src/Main3D.f90:60-63 */
  c->fltk = PI * dp[ORC_D_FLTK]; c->fltkbc = PI * dp[ORC_D_FLTKBC];
  c->rhooo = one; c->poo = one / gam; c->aoo = sqrt(gam * c->poo / c->rhooo);
  c->amachoo = sqrt(c->amach1 * c->amach1 + c->amach2 * c->amach2 + c->amach3 * c->amach3);
  c->srefoo = 111 / c->tempoo; c->srefp1dre = (c->srefoo + one) / c->reoo;
  c->uoo[0] = c->amach1 * c->aoo; c->uoo[1] = c->amach2 * c->aoo; c->uoo[2] = c->amach3 * c->aoo;
  setup_coefficients(c);
  c->sc[ORC_S_AOA] = c->aoa0;

  for (int b = 0; b < nb; b++) {
    oblk *B = &c->b[b];
    B->lxi = dims[b * 3 + 0]; B->let = dims[b * 3 + 1]; B->lze = dims[b * 3 + 2];
    B->N = (B->lxi + 1) * (B->let + 1) * (B->lze + 1);
    B->lim = (B->lxi + 1) + (B->let + 1) + (B->lze + 1) - 1;
    int L[3] = {B->lxi, B->let, B->lze};
    /* src/Main3D.f90:203-207 */
    for (int nn = 0; nn < 3; nn++) {
      B->ijk[nn][0] = L[nn]; B->ijk[nn][1] = L[(nn + 1) % 3]; B->ijk[nn][2] = L[(nn + 2) % 3];
      B->nbsize[nn] = (B->ijk[nn][1] + 1) * (B->ijk[nn][2] + 1);
      for (int s = 0; s < 2; s++) {
        B->nbc[nn][s] = nbc[b * 6 + nn * 2 + s];
        B->ncd[nn][s] = ncd[b * 6 + nn * 2 + s];
        B->nsz[nn][s] = nsz[b * 6 + nn * 2 + s];
        int np = B->nbc[nn][s];
        /* src/Main3D.f90:322-327 */
        B->ndf[nn][s] = (np == 35 || np == 40 || np == 45) ? 1 : 0;
      }
    }
    size_t N = B->N;
    B->qa = dalloc(5 * N); B->qo = dalloc(5 * N); B->de = dalloc(5 * N);
    B->rr = dalloc(3 * N); B->ss = dalloc(3 * N);
    B->xim = dalloc(3 * N); B->etm = dalloc(3 * N); B->zem = dalloc(3 * N);
    B->p = dalloc(N); B->yaco = dalloc(N); B->tv = dalloc(9 * N);
    for (int nn = 0; nn < 3; nn++) {
      size_t F = B->nbsize[nn];
      B->drva[nn] = dalloc(10 * F); B->drvb[nn] = dalloc(10 * F);
      B->send[nn] = dalloc(6 * F); B->recv[nn] = dalloc(6 * F);
      B->cm[nn] = dalloc(6 * F);
    }
    int nl = B->lim + 1;
    B->xu = dalloc(3 * nl); B->xl = dalloc(2 * nl);
    B->yu = dalloc(3 * nl); B->yl = dalloc(2 * nl);
    /* src/Main3D.f90:316-332 */
    for (int nn = 0; nn < 3; nn++) {
      int is = nn == 0 ? 0 : nn == 1 ? B->lxi + 1 : B->lxi + B->let + 2;
      int n = L[nn] + 1;
      double xu[3 * MAXLINE], xl[2 * MAXLINE];
      if (g_oom) { orc_destroy(c); return NULL; }
      int ns = B->ndf[nn][0], ne = B->ndf[nn][1];
      orc_penta(n, &c->albed[ns][0][0], &c->albed[ne][0][0], alpha, beta, xu, xl);
      for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) B->xu[k * nl + is + i] = xu[k * n + i];
        for (int k = 0; k < 2; k++) B->xl[k * nl + is + i] = xl[k * n + i];
      }
      orc_penta(n, &c->albef[ns][0][0], &c->albef[ne][0][0], c->alphf, c->betf, xu, xl);
      for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) B->yu[k * nl + is + i] = xu[k * n + i];
        for (int k = 0; k < 2; k++) B->yl[k * nl + is + i] = xl[k * n + i];
      }
    }
    /* closure faces read lmd+1 points inward plus the face point (mpigo,
       src/Subroutines3D.f90:39; deriv :107-108): lines need >= lmd+2 points */
    for (int nn = 0; nn < 3; nn++) {
      int need = 7;
      for (int s = 0; s < 2; s++) {
        int np = B->nbc[nn][s];
        if (np == 40 && need < LMD + 1) need = LMD + 1;
        if ((np == 35 || np == 45) && need < LMD + 2) need = LMD + 2;
      }
      if (L[nn] + 1 < need) {
        snprintf(g_err, sizeof(g_err), "oracle: block %d direction %d has %d points; its faces need >= %d",
                 b, nn + 1, L[nn] + 1, need);
        orc_destroy(c);
        return NULL;
      }
    }
    B->xbc_ra0 = c->cbca_ra0;
    B->lsz = -1;
  }
  return c;
}

void orc_destroy(orc *c) {
  if (!c) return;
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    free(B->qa); free(B->qo); free(B->de); free(B->rr); free(B->ss);
    free(B->xim); free(B->etm); free(B->zem); free(B->p); free(B->yaco); free(B->tv);
    for (int nn = 0; nn < 3; nn++) {
      free(B->drva[nn]); free(B->drvb[nn]); free(B->send[nn]); free(B->recv[nn]);
      free(B->cm[nn]);
    }
    free(B->xu); free(B->xl); free(B->yu); free(B->yl);
    free(B->rpex); free(B->sbcc); free(B->lcsz); free(B->asz); free(B->bsz);
    free(B->lctz); free(B->atz); free(B->vito);
    free(B->sarea); free(B->snor); free(B->sst);
    free(B->lcfor); free(B->xafor);
  }
  free(c->modes);
  free(c->eddies);
  free(c->b); free(c->jct); free(c->jvel); free(c);
}

int orc_npoints(const orc *c, int b) { return c->b[b].N; }

int orc_count(const orc *c, int b, int w) {
  const oblk *B = &c->b[b];
  int N = B->N, nl = B->lim + 1;
  switch (w) {
    case ORC_A_QA: case ORC_A_QO: case ORC_A_DE: return 5 * N;
    case ORC_A_RR: case ORC_A_SS: case ORC_A_XIM: case ORC_A_ETM: case ORC_A_ZEM: return 3 * N;
    case ORC_A_P: case ORC_A_YACO: return N;
    case ORC_A_TV: return 9 * N;
    case ORC_A_DRVA1: case ORC_A_DRVB1: return 10 * B->nbsize[0];
    case ORC_A_DRVA2: case ORC_A_DRVB2: return 10 * B->nbsize[1];
    case ORC_A_DRVA3: case ORC_A_DRVB3: return 10 * B->nbsize[2];
    case ORC_A_SEND1: case ORC_A_RECV1: case ORC_A_CM1: return 6 * B->nbsize[0];
    case ORC_A_SEND2: case ORC_A_RECV2: case ORC_A_CM2: return 6 * B->nbsize[1];
    case ORC_A_SEND3: case ORC_A_RECV3: case ORC_A_CM3: return 6 * B->nbsize[2];
    case ORC_A_XU: case ORC_A_YU: return 3 * nl;
    case ORC_A_XL: case ORC_A_YL: return 2 * nl;
    case ORC_A_RPEX: return B->nwall;
    case ORC_A_SBCC: return B->nsbcc;
    case ORC_A_ASZ: case ORC_A_BSZ: return B->lsz + 1;
    case ORC_A_VITO: return B->vito ? 3 * (B->ltz + 1) * (c->ntz + 1) : 0;
    case ORC_A_ATZ: return B->atz ? B->ltz + 1 : 0;
  }
  return 0;
}

double *orc_array(orc *c, int b, int w) {
  oblk *B = &c->b[b];
  switch (w) {
    case ORC_A_QA: return B->qa; case ORC_A_QO: return B->qo; case ORC_A_DE: return B->de;
    case ORC_A_RR: return B->rr; case ORC_A_SS: return B->ss;
    case ORC_A_XIM: return B->xim; case ORC_A_ETM: return B->etm; case ORC_A_ZEM: return B->zem;
    case ORC_A_P: return B->p; case ORC_A_YACO: return B->yaco; case ORC_A_TV: return B->tv;
    case ORC_A_DRVA1: return B->drva[0]; case ORC_A_DRVA2: return B->drva[1];
    case ORC_A_DRVA3: return B->drva[2];
    case ORC_A_DRVB1: return B->drvb[0]; case ORC_A_DRVB2: return B->drvb[1];
    case ORC_A_DRVB3: return B->drvb[2];
    case ORC_A_SEND1: return B->send[0]; case ORC_A_SEND2: return B->send[1];
    case ORC_A_SEND3: return B->send[2];
    case ORC_A_RECV1: return B->recv[0]; case ORC_A_RECV2: return B->recv[1];
    case ORC_A_RECV3: return B->recv[2];
    case ORC_A_CM1: return B->cm[0]; case ORC_A_CM2: return B->cm[1]; case ORC_A_CM3: return B->cm[2];
    case ORC_A_XU: return B->xu; case ORC_A_XL: return B->xl;
    case ORC_A_YU: return B->yu; case ORC_A_YL: return B->yl;
    case ORC_A_RPEX: return B->rpex; case ORC_A_SBCC: return B->sbcc;
    case ORC_A_ASZ: return B->asz; case ORC_A_BSZ: return B->bsz;
    case ORC_A_VITO: return B->vito; case ORC_A_ATZ: return B->atz;
  }
  return NULL;
}

double *orc_scalars(orc *c) { return c->sc; }

int orc_sponge_list(const orc *c, int b, int32_t *l, int32_t *g) {
  const oblk *B = &c->b[b];
  const int n = B->lsz + 1;
  if (l)
    for (int q = 0; q < n; q++) l[q] = B->lcsz[q];
  if (g) {
    for (int q = 0; q < n; q++) g[q] = -1;
    for (int k = 0; k <= B->ltz; k++) g[B->lctz[k]] = k;
  }
  return n;
}

int orc_force_list(const orc *c, int b, int32_t *l, double *xafor3) {
  const oblk *B = &c->b[b];
  const int n = B->lfor + 1;
  for (int q = 0; q < n; q++) {
    if (l) l[q] = B->lcfor[q];
    if (xafor3)
      for (int m = 0; m < 3; m++) xafor3[m * n + q] = B->xafor[m * n + q];
  }
  return n;
}

void orc_set_inflow_eddies(orc *c, int nits, const double *eddies10) {
  free(c->eddies);
  c->nits = nits;
  c->eddies = dalloc((size_t)10 * nits);
  memcpy(c->eddies, eddies10, sizeof(double) * 10 * nits);
}

void orc_set_inflow_modes(orc *c, int nmodes, const double *modes8) {
  free(c->modes);
  c->nmodes = nmodes;
  c->modes = dalloc((size_t)8 * nmodes);
  memcpy(c->modes, modes8, sizeof(double) * 8 * nmodes);
}

/* src/SingleTurbulence3D.f90:336-356: five-point Lagrange interpolation in the
 * periodic table time tt(ii) = ii*fctr, ii = 0..ntz (fctr = slit/(ntz*uoo(1)),
 * period ra1 = slit/uoo(1)), at ra3 = (timo - tgusto + dtk) mod ra1 */
int orc_inflow_weights(const orc *c, double timo, double dtk, int *il5, double *tl5, double *w5) {
  if (!(timo > c->tgusto)) return 0;
  double fctr = c->slit / (c->ntz * c->uoo[0]);
  double ra0 = timo - c->tgusto + dtk, ra1 = c->slit / c->uoo[0], ra2 = ra0 / ra1;
  double ra3 = ra0 - ra1 * (double)(int)ra2;
  int is = 0, ie = c->ntz;
  int ii = 0; /* minloc(abs(tt(:)-ra3)): first minimum */
  double best = fabs(0 * fctr - ra3);
  for (int i = 1; i <= ie; i++) {
    double d = fabs(i * fctr - ra3);
    if (d < best) { best = d; ii = i; }
  }
  int *il = il5 + 2;       /* il[-2..2] */
  double *tl = tl5 + 2;
  for (int jj = -2; jj <= 2; jj++) {
    int v = ii + jj; v = v < is ? is : v; v = v > ie ? ie : v;
    il[jj] = v; tl[jj] = v * fctr;
  }
  if (ii - is == 0) { il[-2] = ie - 2; il[-1] = ie - 1; tl[-2] = il[-2] * fctr - ra1; tl[-1] = il[-1] * fctr - ra1; }
  if (ii - is == 1) { il[-2] = ie - 1; tl[-2] = il[-2] * fctr - ra1; }
  if (ie - ii == 1) { il[2] = is + 1; tl[2] = il[2] * fctr + ra1; }
  if (ie - ii == 0) { il[1] = is + 1; il[2] = is + 2; tl[1] = il[1] * fctr + ra1; tl[2] = il[2] * fctr + ra1; }
  double al[5];
  for (int q = -2; q <= 2; q++) al[q + 2] = ra3 - tl[q];
  for (int jj = -2; jj <= 2; jj++) {
    double ao = one, bo = one;
    for (int q = -2; q <= 2; q++)
      if (q != jj) { ao = ao * al[q + 2]; bo = bo * (tl[jj] - tl[q]); }
    w5[jj + 2] = ao / bo;
  }
  return 1;
}

void orc_coefficients(orc *c, double *pbci, double *pbco, double *pbcot,
                      double *albef, double *fabc, double *sbci) {
  if (pbci) memcpy(pbci, c->pbci, sizeof(c->pbci));
  if (pbco) memcpy(pbco, c->pbco, sizeof(c->pbco));
  if (pbcot) memcpy(pbcot, c->pbcot, sizeof(c->pbcot));
  if (albef) memcpy(albef, c->albef, sizeof(c->albef));
  if (fabc) {
    fabc[0] = c->alphf; fabc[1] = c->betf; fabc[2] = c->fa; fabc[3] = c->fb; fabc[4] = c->fc;
    for (int i = 0; i < 3; i++) { fabc[5 + i] = c->fam[i]; fabc[8 + i] = c->fbm[i]; fabc[11 + i] = c->fcm[i]; }
  }
  if (sbci) memcpy(sbci, c->sbci, sizeof(c->sbci));
}

/* ---------------------------------------------------------------------- */
/* mpigo: src/Subroutines3D.f90:16-80, over all blocks                      */
/* ---------------------------------------------------------------------- */
int orc_mpigo(orc *c, int nt, int nrt, int n45) {
  if (nt < 0 || nt > 1 || nrt < 0 || nrt > 1 || n45 < 0 || n45 > 1) {
    snprintf(g_err, sizeof(g_err), "oracle: mpigo(%d, %d, %d) out of range", nt, nrt, n45);
    return 1;
  }
  int mp = nt == 0 ? LMD : LMF;
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    for (int nn = 0; nn < 3; nn++) {
      int nz = (1 - nrt) * nn; /* 0-based column of rr */
      const double *rz = B->rr + (size_t)nz * B->N;
      int F = B->nbsize[nn];
      for (int ipp = 0; ipp <= 1; ipp++) {
        if (B->ndf[nn][ipp] != 1) continue;
        int is = ipp * B->ijk[nn][0], ie = 1 - 2 * ipp, ii = 0;
        double ra0 = 0;
        switch (B->nbc[nn][ipp]) {
          case 35: ra0 = zero; ii = 1; break;
          case 40: ra0 = zero; ii = 0; break;
          case 45: ra0 = n45; ii = 1; break;
        }
        double *snd = B->send[nn];
        const double *po0 = c->pbco[nt][0], *po1 = c->pbco[nt][1];
        double pt0 = c->pbcot[nt][0], pt1 = c->pbcot[nt][1];
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (int k = 0; k <= B->ijk[nn][2]; k++) {
          int kp = k * (B->ijk[nn][1] + 1);
          double sap[LMP + 1];
          for (int j = 0; j <= B->ijk[nn][1]; j++) {
            int jk = kp + j;
            double res = ra0 * rz[IDX3(B, is, j, k, nn)];
            for (int i = 0; i <= mp; i++) sap[i] = rz[IDX3(B, is + ie * (i + ii), j, k, nn)];
            double s0 = 0, s1 = 0;
            for (int i = 0; i <= mp; i++) s0 = s0 + po0[i] * sap[i];
            for (int i = 0; i <= mp; i++) s1 = s1 + po1[i] * sap[i];
            snd[(ipp * 3 + 0) * F + jk] = s0 - res * pt0;
            snd[(ipp * 3 + 1) * F + jk] = s1 - res * pt1;
            snd[(ipp * 3 + 2) * F + jk] = sap[0] - res;
          }
        }
      }
    }
  }
  /* exchange: my recv(:,:,ip) <- neighbour's send(:,:,1-ip) */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    for (int nn = 0; nn < 3; nn++)
      for (int ipp = 0; ipp <= 1; ipp++) {
        if (B->ndf[nn][ipp] != 1) continue;
        oblk *Nb = &c->b[B->ncd[nn][ipp]];
        int F = B->nbsize[nn];
        memcpy(B->recv[nn] + (size_t)ipp * 3 * F, Nb->send[nn] + (size_t)(1 - ipp) * 3 * F,
               sizeof(double) * 3 * F);
      }
  }
  if (n45 == 1) {
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      for (int nn = 0; nn < 3; nn++) {
        int nz = (1 - nrt) * nn;
        const double *rz = B->rr + (size_t)nz * B->N;
        int F = B->nbsize[nn];
        for (int ipp = 0; ipp <= 1; ipp++) {
          if (B->nbc[nn][ipp] != 45) continue;
          int is = ipp * B->ijk[nn][0];
          double *rcv = B->recv[nn];
          for (int k = 0; k <= B->ijk[nn][2]; k++)
            for (int j = 0; j <= B->ijk[nn][1]; j++) {
              int jk = k * (B->ijk[nn][1] + 1) + j;
              double v = rz[IDX3(B, is, j, k, nn)];
              rcv[(ipp * 3 + 0) * F + jk] = rcv[(ipp * 3 + 0) * F + jk] + v * c->pbcot[nt][0];
              rcv[(ipp * 3 + 1) * F + jk] = rcv[(ipp * 3 + 1) * F + jk] + v * c->pbcot[nt][1];
              rcv[(ipp * 3 + 2) * F + jk] = rcv[(ipp * 3 + 2) * F + jk] + v;
            }
        }
      }
    }
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* deriv: src/Subroutines3D.f90:84-141                                      */
/* ---------------------------------------------------------------------- */
static void deriv_block(orc *c, oblk *B, int nn, int nz, int m) {
  int ns = B->ndf[nn][0], ne = B->ndf[nn][1];
  int nl = B->lim + 1;
  int is = 0, ie = B->ijk[nn][0];
  int off = nn == 0 ? 0 : nn == 1 ? B->lxi + 1 : B->lxi + B->let + 2;
  const double *xl1 = B->xl + off, *xl2 = B->xl + nl + off;
  const double *xu1 = B->xu + off, *xu2 = B->xu + nl + off, *xu3 = B->xu + 2 * nl + off;
  const double *in = B->rr + (size_t)(nz - 1) * B->N;
  double *out = B->rr + (size_t)nn * B->N;
  const double *rcv = B->recv[nn];
  double *drv = B->drva[nn];
  int F = B->nbsize[nn];
  const double(*abc)[6] = c->abc;
  const double *pb0 = c->pbci[0][0], *pb1 = c->pbci[0][1];
#pragma omp parallel for num_threads(nthreads()) schedule(static)
  for (int k = 0; k <= B->ijk[nn][2]; k++) {
    int kp = k * (B->ijk[nn][1] + 1);
    double sa[MAXLINE], sb[MAXLINE];
    int li[MAXLINE];
    for (int j = 0; j <= B->ijk[nn][1]; j++) {
      int jk = kp + j;
      for (int i = is; i <= ie; i++) { int l = IDX3(B, i, j, k, nn); li[i] = l; sa[i] = in[l]; }
      if (ns == 0) {
        static const int o0[6] = {1, 2, 3, 4, 5, 6}, o1[6] = {0, 2, 3, 4, 5, 6},
                         o2[6] = {0, 1, 3, 4, 5, 6};
        double s;
        s = 0; for (int q = 0; q < 6; q++) s = s + abc[0][q] * (sa[is + o0[q]] - sa[is]); sb[is] = s;
        s = 0; for (int q = 0; q < 6; q++) s = s + abc[1][q] * (sa[is + o1[q]] - sa[is + 1]); sb[is + 1] = s;
        s = 0; for (int q = 0; q < 6; q++) s = s + abc[2][q] * (sa[is + o2[q]] - sa[is + 2]); sb[is + 2] = s;
      } else {
        double s0 = 0, s1 = 0;
        for (int q = 0; q <= LMD; q++) s0 = s0 + pb0[q] * sa[is + q];
        for (int q = 0; q <= LMD; q++) s1 = s1 + pb1[q] * sa[is + q];
        sb[is] = s0 + rcv[(0 * 3 + 0) * F + jk];
        sb[is + 1] = s1 + rcv[(0 * 3 + 1) * F + jk];
        sb[is + 2] = aa * (sa[is + 3] - sa[is + 1]) + ab * (sa[is + 4] - sa[is]) +
                     ac * (sa[is + 5] - rcv[(0 * 3 + 2) * F + jk]);
      }
      for (int i = is + 3; i <= ie - 3; i++)
        sb[i] = aa * (sa[i + 1] - sa[i - 1]) + ab * (sa[i + 2] - sa[i - 2]) + ac * (sa[i + 3] - sa[i - 3]);
      if (ne == 0) {
        static const int o0[6] = {1, 2, 3, 4, 5, 6}, o1[6] = {0, 2, 3, 4, 5, 6},
                         o2[6] = {0, 1, 3, 4, 5, 6};
        double s;
        s = 0; for (int q = 0; q < 6; q++) s = s + abc[0][q] * (sa[ie] - sa[ie - o0[q]]); sb[ie] = s;
        s = 0; for (int q = 0; q < 6; q++) s = s + abc[1][q] * (sa[ie - 1] - sa[ie - o1[q]]); sb[ie - 1] = s;
        s = 0; for (int q = 0; q < 6; q++) s = s + abc[2][q] * (sa[ie - 2] - sa[ie - o2[q]]); sb[ie - 2] = s;
      } else {
        double s0 = 0, s1 = 0;
        for (int q = 0; q <= LMD; q++) s0 = s0 + pb0[q] * sa[ie - q];
        for (int q = 0; q <= LMD; q++) s1 = s1 + pb1[q] * sa[ie - q];
        sb[ie] = -s0 - rcv[(1 * 3 + 0) * F + jk];
        sb[ie - 1] = -s1 - rcv[(1 * 3 + 1) * F + jk];
        sb[ie - 2] = aa * (sa[ie - 1] - sa[ie - 3]) + ab * (sa[ie] - sa[ie - 4]) +
                     ac * (rcv[(1 * 3 + 2) * F + jk] - sa[ie - 5]);
      }
      sa[is] = sb[is];
      sa[is + 1] = sb[is + 1] - xl2[is + 1] * sa[is];
      for (int i = is + 2; i <= ie; i++) sa[i] = sb[i] - xl1[i] * sa[i - 2] - xl2[i] * sa[i - 1];
      sb[ie] = xu1[ie] * sa[ie];
      sb[ie - 1] = xu1[ie - 1] * sa[ie - 1] - xu2[ie - 1] * sb[ie];
      for (int i = ie - 2; i >= is; i--) sb[i] = xu1[i] * sa[i] - xu2[i] * sb[i + 1] - xu3[i] * sb[i + 2];
      for (int i = is; i <= ie; i++) out[li[i]] = sb[i];
      if (m > 0) {   /* m = 0: no boundary capture (internal use only) */
        drv[(0 * 5 + m - 1) * F + jk] = sb[is];
        drv[(1 * 5 + m - 1) * F + jk] = sb[ie];
      }
    }
  }
}

int orc_deriv(orc *c, int nn, int nz, int m) {
  if (nn < 1 || nn > 3 || nz < 1 || nz > 3 || m < 0 || m > 5) {
    snprintf(g_err, sizeof(g_err), "oracle: deriv(%d, %d, %d) out of range", nn, nz, m);
    return 1;
  }
  for (int b = 0; b < c->nb; b++) deriv_block(c, &c->b[b], nn - 1, nz, m);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* filte: src/Subroutines3D.f90:145-206                                     */
/* ---------------------------------------------------------------------- */
static void filte_block(orc *c, oblk *B, int nn, int nz) {
  const double fex[3] = {45 / (1 * 30.0), -9 / (1 * 30.0), 1 / (1 * 30.0)};
  const double l2 = 8 * log(two);
  const double ffx[3] = {4 / l2, 6 / l2, 7 / l2};
  int ns = B->ndf[nn][0], ne = B->ndf[nn][1];
  int nl = B->lim + 1;
  int is = 0, ie = B->ijk[nn][0];
  int off = nn == 0 ? 0 : nn == 1 ? B->lxi + 1 : B->lxi + B->let + 2;
  const double *yl1 = B->yl + off, *yl2 = B->yl + nl + off;
  const double *yu1 = B->yu + off, *yu2 = B->yu + nl + off, *yu3 = B->yu + 2 * nl + off;
  double *io = B->rr + (size_t)(nz - 1) * B->N;
  const double *rcv = B->recv[nn];
  int F = B->nbsize[nn];
  const double *fam = c->fam, *fbm = c->fbm, *fcm = c->fcm;
  double fa = c->fa, fb = c->fb, fc = c->fc;
  const double *pb0 = c->pbci[1][0], *pb1 = c->pbci[1][1];
#pragma omp parallel for num_threads(nthreads()) schedule(static)
  for (int k = 0; k <= B->ijk[nn][2]; k++) {
    int kp = k * (B->ijk[nn][1] + 1);
    double sa[MAXLINE], sb[MAXLINE];
    int li[MAXLINE];
    for (int j = 0; j <= B->ijk[nn][1]; j++) {
      int jk = kp + j;
      double ra0, ra1, ra2, res, rv[3];
      for (int i = is; i <= ie; i++) { int l = IDX3(B, i, j, k, nn); li[i] = l; sa[i] = io[l]; }
      if (ns == 0) {
        ra0 = two * sa[is]; ra1 = two * sa[is + 1]; ra2 = two * sa[is + 2];
        res = 0;
        for (int q = 0; q < 3; q++) res = res + fex[q] * (sa[is + 1 + q] - sa[is]);
        for (int q = 0; q < 3; q++) rv[q] = sa[is] - res * ffx[q];
        sb[is] = fam[0] * (rv[0] + sa[is + 1] - ra0) + fbm[0] * (rv[1] + sa[is + 2] - ra0) +
                 fcm[0] * (rv[2] + sa[is + 3] - ra0);
        sb[is + 1] = fam[1] * (sa[is] + sa[is + 2] - ra1) + fbm[1] * (rv[0] + sa[is + 3] - ra1) +
                     fcm[1] * (rv[1] + sa[is + 4] - ra1);
        sb[is + 2] = fam[2] * (sa[is + 1] + sa[is + 3] - ra2) + fbm[2] * (sa[is] + sa[is + 4] - ra2) +
                     fcm[2] * (rv[0] + sa[is + 5] - ra2);
      } else {
        ra2 = two * sa[is + 2];
        double s0 = 0, s1 = 0;
        for (int q = 0; q <= LMF; q++) s0 = s0 + pb0[q] * sa[is + q];
        for (int q = 0; q <= LMF; q++) s1 = s1 + pb1[q] * sa[is + q];
        sb[is] = s0 + rcv[(0 * 3 + 0) * F + jk];
        sb[is + 1] = s1 + rcv[(0 * 3 + 1) * F + jk];
        sb[is + 2] = fa * (sa[is + 1] + sa[is + 3] - ra2) + fb * (sa[is] + sa[is + 4] - ra2) +
                     fc * (rcv[(0 * 3 + 2) * F + jk] + sa[is + 5] - ra2);
      }
      for (int i = is + 3; i <= ie - 3; i++) {
        res = two * sa[i];
        sb[i] = fa * (sa[i - 1] + sa[i + 1] - res) + fb * (sa[i - 2] + sa[i + 2] - res) +
                fc * (sa[i - 3] + sa[i + 3] - res);
      }
      if (ne == 0) {
        ra0 = two * sa[ie]; ra1 = two * sa[ie - 1]; ra2 = two * sa[ie - 2];
        res = 0;
        for (int q = 0; q < 3; q++) res = res + fex[q] * (sa[ie] - sa[ie - 1 - q]);
        for (int q = 0; q < 3; q++) rv[q] = sa[ie] + res * ffx[q];
        sb[ie] = fam[0] * (sa[ie - 1] + rv[0] - ra0) + fbm[0] * (sa[ie - 2] + rv[1] - ra0) +
                 fcm[0] * (sa[ie - 3] + rv[2] - ra0);
        sb[ie - 1] = fam[1] * (sa[ie - 2] + sa[ie] - ra1) + fbm[1] * (sa[ie - 3] + rv[0] - ra1) +
                     fcm[1] * (sa[ie - 4] + rv[1] - ra1);
        sb[ie - 2] = fam[2] * (sa[ie - 3] + sa[ie - 1] - ra2) + fbm[2] * (sa[ie - 4] + sa[ie] - ra2) +
                     fcm[2] * (sa[ie - 5] + rv[0] - ra2);
      } else {
        ra2 = two * sa[ie - 2];
        double s0 = 0, s1 = 0;
        for (int q = 0; q <= LMF; q++) s0 = s0 + pb0[q] * sa[ie - q];
        for (int q = 0; q <= LMF; q++) s1 = s1 + pb1[q] * sa[ie - q];
        sb[ie] = s0 + rcv[(1 * 3 + 0) * F + jk];
        sb[ie - 1] = s1 + rcv[(1 * 3 + 1) * F + jk];
        sb[ie - 2] = fa * (sa[ie - 3] + sa[ie - 1] - ra2) + fb * (sa[ie - 4] + sa[ie] - ra2) +
                     fc * (sa[ie - 5] + rcv[(1 * 3 + 2) * F + jk] - ra2);
      }
      sa[is] = sb[is];
      sa[is + 1] = sb[is + 1] - yl2[is + 1] * sa[is];
      for (int i = is + 2; i <= ie; i++) sa[i] = sb[i] - yl1[i] * sa[i - 2] - yl2[i] * sa[i - 1];
      sb[ie] = yu1[ie] * sa[ie];
      sb[ie - 1] = yu1[ie - 1] * sa[ie - 1] - yu2[ie - 1] * sb[ie];
      for (int i = ie - 2; i >= is; i--) sb[i] = yu1[i] * sa[i] - yu2[i] * sb[i + 1] - yu3[i] * sb[i + 2];
      for (int i = is; i <= ie; i++) io[li[i]] = io[li[i]] + sb[i];
    }
  }
}

int orc_filte(orc *c, int nn, int nz) {
  if (nn < 1 || nn > 3 || nz < 1 || nz > 3) {
    snprintf(g_err, sizeof(g_err), "oracle: filte(%d, %d) out of range", nn, nz);
    return 1;
  }
  for (int b = 0; b < c->nb; b++) filte_block(c, &c->b[b], nn - 1, nz);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* setup on the grid: metrics, normals, GCBC weights, sponge                */
/* ---------------------------------------------------------------------- */

/* src/SingleTurbulence3D.f90:447-464, with the caller's global ra0 as state */
static double extrabcc(orc *c, oblk *B, int l) {
  double flag = zero;
  const double *ss = B->ss;
  if (ss[l] < B->xbc_ra0) {
    double slitl = c->wlea, slitw = 0.4 * slitl;
    B->xbc_ra0 = -half + c->wlea + slitl;
    double ra1 = -half * c->span + 0.25 * c->wlew - half * slitw, ra2 = ra1 + slitw;
    int ie = (int)((c->span + sml) / c->wlew) - 1;
    for (int ii = 0; ii <= ie; ii++) {
      double ra3 = ii * c->wlew;
      double z = ss[2 * B->N + l];
      double res = (z - (ra1 + ra3)) * (z - (ra2 + ra3));
      if (res < zero && c->nextrabc == 1) flag = one;
    }
  }
  return flag;
}

void orc_metrics(orc *c) {
  /* src/Main3D.f90:365-375: derivatives of x, y, z */
  for (int m = 1; m <= 3; m++) {
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      memcpy(B->rr, B->ss + (size_t)(m - 1) * B->N, sizeof(double) * B->N);
    }
    orc_mpigo(c, 0, 1, 1);
    orc_deriv(c, 3, 1, m);
    orc_deriv(c, 2, 1, m);
    orc_deriv(c, 1, 1, m);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      double *dst = m == 1 ? B->qo : m == 2 ? B->qa : B->de;
      memcpy(dst, B->rr, sizeof(double) * 3 * B->N);
    }
  }
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    double *qo = B->qo, *qa = B->qa, *de = B->de;
    double *xim = B->xim, *etm = B->etm, *zem = B->zem;
    /* src/Main3D.f90:378-386, 411-413 */
    for (int l = 0; l < N; l++) {
      double qo1 = qo[l], qo2 = qo[N + l], qo3 = qo[2 * N + l];
      double qa1 = qa[l], qa2 = qa[N + l], qa3 = qa[2 * N + l];
      double de1 = de[l], de2 = de[N + l], de3 = de[2 * N + l];
      xim[l] = qa2 * de3 - de2 * qa3;
      xim[N + l] = de2 * qo3 - qo2 * de3;
      xim[2 * N + l] = qo2 * qa3 - qa2 * qo3;
      etm[l] = qa3 * de1 - de3 * qa1;
      etm[N + l] = de3 * qo1 - qo3 * de1;
      etm[2 * N + l] = qo3 * qa1 - qa3 * qo1;
      zem[l] = qa1 * de2 - de1 * qa2;
      zem[N + l] = de1 * qo2 - qo1 * de2;
      zem[2 * N + l] = qo1 * qa2 - qa1 * qo2;
      B->yaco[l] = three / (qo1 * xim[l] + qo2 * etm[l] + qo3 * zem[l] + qa1 * xim[N + l] +
                            qa2 * etm[N + l] + qa3 * zem[N + l] + de1 * xim[2 * N + l] +
                            de2 * etm[2 * N + l] + de3 * zem[2 * N + l]);
    }
    /* face unit normals, src/Main3D.f90:429-439 */
    for (int nn = 0; nn < 3; nn++) {
      const double *mt = nn == 0 ? xim : nn == 1 ? etm : zem;
      int F = B->nbsize[nn];
      for (int ipp = 0; ipp <= 1; ipp++) {
        int i = ipp * B->ijk[nn][0];
        for (int k = 0; k <= B->ijk[nn][2]; k++)
          for (int j = 0; j <= B->ijk[nn][1]; j++) {
            int jk = k * (B->ijk[nn][1] + 1) + j, l = IDX3(B, i, j, k, nn);
            double rv[3] = {B->yaco[l] * mt[l], B->yaco[l] * mt[N + l], B->yaco[l] * mt[2 * N + l]};
            double fctr = one / sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
            for (int d = 0; d < 3; d++) B->cm[nn][(ipp * 3 + d) * F + jk] = fctr * rv[d];
          }
      }
    }
    /* wall surface, src/rpt.f90:774-864 (wallArea, walldir): the eta face of
     * the block that is a wall; points in the reference's lwall order (xi
     * fastest), cell areas from the coordinate derivatives still held in
     * qo, qa, de, halved at the parent block's xi and zeta ends (the faces
     * that are not splits of the parent, BC 40), unit normals from etm
     * pointing into the fluid */
    {
      const int bcw = 20 + 5 * c->nviscous;
      int j = -1;
      double coef = one;
      if (B->nbc[1][0] == bcw) { j = 0; coef = one; B->sside = 0; }
      else if (B->nbc[1][1] == bcw) { j = B->let; coef = -one; B->sside = 1; }
      if (j >= 0) {
        int nwl = (B->lxi + 1) * (B->lze + 1), ll = -1;
        free(B->sarea); free(B->snor); free(B->sst);
        B->nsurf = nwl;
        B->sarea = dalloc(nwl); B->snor = dalloc(3 * (size_t)nwl); B->sst = dalloc(3 * (size_t)nwl);
        for (int k = 0; k <= B->lze; k++)
          for (int i = 0; i <= B->lxi; i++) {
            int l = IDX3(B, j, k, i, 1);
            double g11 = qo[l] * qo[l] + qa[l] * qa[l] + de[l] * de[l];
            double g33 = qo[2 * N + l] * qo[2 * N + l] + qa[2 * N + l] * qa[2 * N + l] + de[2 * N + l] * de[2 * N + l];
            double g13 = qo[l] * qo[2 * N + l] + qa[l] * qa[2 * N + l] + de[l] * de[2 * N + l];
            ll = ll + 1;
            double ar = sqrt(g11 * g33 - g13 * g13);
            if (i == 0 && B->nbc[0][0] != 40) ar = ar * half;
            else if (i == B->lxi && B->nbc[0][1] != 40) ar = ar * half;
            if (k == 0 && B->nbc[2][0] != 40) ar = ar * half;
            else if (k == B->lze && B->nbc[2][1] != 40) ar = ar * half;
            B->sarea[ll] = ar;
            double tmp = coef / sqrt(etm[l] * etm[l] + etm[N + l] * etm[N + l] + etm[2 * N + l] * etm[2 * N + l]);
            for (int m = 0; m < 3; m++) B->snor[(size_t)m * nwl + ll] = etm[(size_t)m * N + l] * tmp;
          }
      }
    }
    /* GCBC weights and wall flags, src/Main3D.f90:460-480 */
    int nw = 0, ns = 0;
    for (int nn = 0; nn < 3; nn++)
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp];
        if (np == 10 || np == 20 || np == 25 || np == 30) {
          int F = B->nbsize[nn];
          ns += F * (MBCI + 1);
          if (np == 20 || np == 25) nw += F;
        }
      }
    free(B->rpex); free(B->sbcc);
    B->nwall = nw; B->nsbcc = ns;
    B->rpex = dalloc(nw); B->sbcc = dalloc(ns);
    double *cnt = B->rr, *wgt = dalloc(ns);
    int *lst = (int *)calloc(ns + 1, sizeof(int));
    for (int l = 0; l < N; l++) cnt[l] = zero;
    int lp = -1, ll = -1;
    for (int nn = 0; nn < 3; nn++)
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp], i = ipp * B->ijk[nn][0], iq = 1 - 2 * ipp;
        if (!(np == 10 || np == 20 || np == 25 || np == 30)) continue;
        for (int k = 0; k <= B->ijk[nn][2]; k++)
          for (int j = 0; j <= B->ijk[nn][1]; j++) {
            int l = IDX3(B, i, j, k, nn);
            if (np == 20 || np == 25) { lp++; B->rpex[lp] = extrabcc(c, B, l); }
            ll++;
            double res = one / B->yaco[l];
            cnt[l] = cnt[l] + one; wgt[ll] = res; lst[ll] = l;
            for (int ii = 1; ii <= MBCI; ii++) {
              int l2 = IDX3(B, i + iq * ii, j, k, nn);
              ll++;
              cnt[l2] = cnt[l2] + one; wgt[ll] = res * c->sbci[ii - 1]; lst[ll] = l2;
            }
          }
      }
    for (ll = 0; ll < ns; ll++) B->sbcc[ll] = wgt[ll] / cnt[lst[ll]];
    free(wgt); free(lst);
  }
  /* sponge, src/SingleTurbulence3D.f90:209-249 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    double szr[2][3], szp[2][3];
    double p07 = (double)0.7f; /* default-real literal 0.7e0 in the reference */
    szr[0][0] = 1 / c->szth1; szp[0][0] = -p07 * c->domlen + c->szth1;
    szr[1][0] = 1 / (c->szth1 + c->szxt); szp[1][0] = c->domlen - c->szth1;
    szr[0][1] = 1 / c->szth2; szp[0][1] = -p07 * c->domlen + c->szth2;
    szr[1][1] = 1 / c->szth2; szp[1][1] = c->domlen + c->szxt - c->szth2;
    szr[0][2] = 0; szp[0][2] = 0; szr[1][2] = 0; szp[1][2] = 0;
    double ra1 = 0.25, ra2 = PI / (two * c->domlen);
    int *lst = (int *)malloc(sizeof(int) * (N + 1));
    double *sg = dalloc(N), *lm = dalloc(N);
    int ll = -1;
    for (int l = 0; l < N; l++) {
      double r[3];
      for (int d = 0; d < 3; d++) {
        double x = B->ss[(size_t)d * N + l];
        r[d] = B->nsz[d][0] * szr[0][d] * fmax(szp[0][d] - x, zero) +
               B->nsz[d][1] * szr[1][d] * fmax(x - szp[1][d], zero);
      }
      sg[l] = c->szco * (one - 0.125 * (one + cos(PI * r[0])) * (one + cos(PI * r[1])) *
                                   (one + cos(PI * r[2])));
      lm[l] = ra1 + half * (one - ra1) * (one + cos(ra2 * (fmin(B->ss[l], c->domlen) + c->domlen)));
      if (sg[l] > zero) { ll++; lst[ll] = l; }
    }
    free(B->lcsz); free(B->asz); free(B->bsz);
    B->lsz = ll;
    B->lcsz = (int *)malloc(sizeof(int) * (ll + 1 > 0 ? ll + 1 : 1));
    B->asz = dalloc(ll + 1); B->bsz = dalloc(ll + 1);
    for (int q = 0; q <= ll; q++) {
      int l = lst[q];
      double res = one / B->yaco[l];
      B->lcsz[q] = l;
      B->asz[q] = res * sg[l] * lm[l];
      B->bsz[q] = res * hamm1 * sg[l];
    }
    free(lst); free(sg); free(lm);
    /* inflow points and their velocity table, src/SingleTurbulence3D.f90:253-292
     * (inflow 3: the synthetic eddies; 1, 2: the given Fourier modes) */
    free(B->lctz); free(B->atz); free(B->vito);
    B->lctz = NULL; B->atz = NULL; B->vito = NULL; B->ltz = -1;
    if ((c->inflow == 3 && c->eddies) || (c->inflow && c->inflow != 3 && c->modes)) {
      double ra0 = PI / c->szth1;
      int *gl = (int *)malloc(sizeof(int) * (B->lsz + 2));
      int lt = -1;
      for (int q = 0; q <= B->lsz; q++) {
        int l = B->lcsz[q];
        if (B->ss[l] < szp[0][0] && fabs(B->ss[N + l]) < c->gustw) gl[++lt] = q;
      }
      B->ltz = lt;
      B->lctz = gl;
      B->atz = dalloc(lt + 1);
      int nt1 = c->ntz + 1;
      B->vito = dalloc((size_t)3 * (lt + 1) * nt1);
      double fctr = c->slit / (c->ntz * c->uoo[0]);
      for (int g = 0; g <= lt; g++) {
        int l = B->lcsz[gl[g]];
        double sn = sin(fmin(ra0 * (B->ss[l] + c->domlen), PI / 2));
        B->atz[g] = sn * sn;
      }
      if (c->inflow == 3) {
        /* each eddy nn and its images one slit upstream and one span to
         * either side, convected with uoo to the table times tt(ii):
         * velocity rr x d at the point, d its offset from the eddy centre,
         * src/SingleTurbulence3D.f90:268-291; then scaled by cfit (:298) */
        const int ng = lt + 1;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (int e = 0; e < ng * nt1; e++) {
          int g = e % ng, ii = e / ng;
          int l = B->lcsz[gl[g]];
          double tt = ii * fctr;
          const double px = B->ss[l], py = B->ss[N + l], pz = B->ss[2 * (size_t)N + l];
          double v0 = zero, v1 = zero, v2 = zero;
          for (int nn = 0; nn < c->nits; nn++) {
            const double *E = c->eddies + 10 * (size_t)nn;
            const double mxc = E[9];
            double ve[3];
            for (int m = 0; m < 3; m++) ve[m] = (-9 + mxc) * E[3 + m] * E[3 + m];
            const double ra1 = -36 - 60 * mxc, ra2 = two_three * mxc;
            for (int i = 0; i <= 1; i++)
              for (int k = -1; k <= 1; k++) {
                const double dm0 = E[0] - i * c->slit, dm1 = E[1], dm2 = E[2] - k * c->span;
                const double rv0 = dm0 + tt * c->uoo[0], rv1 = dm1 + tt * c->uoo[1], rv2 = dm2 + tt * c->uoo[2];
                const double d1 = px - rv0, d2 = py - rv1, d3 = pz - rv2;
                const double d4 = d1 * d1 + d2 * d2 + d3 * d3;
                double d5 = d4 * d4;
                double r1 = ve[0] * d5, r2 = ve[1] * d5, r3 = ve[2] * d5;
                d5 = ra1 * d4;
                r1 = E[6] * d5 * exp(r1) * (1 + ra2 * r1);
                r2 = E[7] * d5 * exp(r2) * (1 + ra2 * r2);
                r3 = E[8] * d5 * exp(r3) * (1 + ra2 * r3);
                v0 = v0 + r2 * d3 - r3 * d2;
                v1 = v1 + r3 * d1 - r1 * d3;
                v2 = v2 + r1 * d2 - r2 * d1;
              }
          }
          B->vito[((size_t)0 * nt1 + ii) * ng + g] = c->cfit * v0;
          B->vito[((size_t)1 * nt1 + ii) * ng + g] = c->cfit * v1;
          B->vito[((size_t)2 * nt1 + ii) * ng + g] = c->cfit * v2;
        }
      } else {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
      for (int e = 0; e < (lt + 1) * nt1; e++) {
        int g = e % (lt + 1), ii = e / (lt + 1);
        int l = B->lcsz[gl[g]];
        double tt = ii * fctr;
        double x0 = B->ss[l] - tt * c->uoo[0], x1 = B->ss[N + l] - tt * c->uoo[1],
               x2 = B->ss[2 * (size_t)N + l] - tt * c->uoo[2];
        double s0 = zero, s1 = zero, s2 = zero;
        for (int n = 0; n < c->nmodes; n++) {
          const double *m = c->modes + 8 * (size_t)n;
          double cs = cos(m[0] * x0 + m[1] * x1 + m[2] * x2 + m[6]);
          s0 = s0 + m[3] * cs; s1 = s1 + m[4] * cs; s2 = s2 + m[5] * cs;
        }
        B->vito[((size_t)0 * nt1 + ii) * (lt + 1) + g] = s0;
        B->vito[((size_t)1 * nt1 + ii) * (lt + 1) + g] = s1;
        B->vito[((size_t)2 * nt1 + ii) * (lt + 1) + g] = s2;
      }
      }
    }
    /* forcing points and coefficients (forceup, src/rpt.f90:728-753): the
     * points within rfor of (xfor, yfor), three spanwise modes each with a
     * Gaussian weight ~ exp(-ra0 r^2), ra0 = -ln(1e-4) / rfor^2.  As in the
     * reference, ra1 holds 2 pi / span only for the first point: the loop
     * overwrites it with each point's weight, which the next point's modes
     * then take as their wavenumber. */
    free(B->lcfor); free(B->xafor);
    B->lcfor = NULL; B->xafor = NULL; B->lfor = -1;
    if (c->forcing == 1) {
      double ra0 = -log(0.0001) / (c->rfor * c->rfor), ra1 = 2 * PI / c->span, ra2 = c->rfor * c->rfor;
      int ll = -1;
      int *lst = (int *)malloc(sizeof(int) * (N + 1));
      double *r2v = dalloc(N);
      for (int l = 0; l < N; l++) {
        r2v[l] = (B->ss[l] - c->xfor) * (B->ss[l] - c->xfor) + (B->ss[N + l] - c->yfor) * (B->ss[N + l] - c->yfor);
        if (r2v[l] - ra2 < 0) lst[++ll] = l;
      }
      B->lfor = ll;
      if (ll >= 0) {
        B->lcfor = lst;
        B->xafor = dalloc((size_t)3 * (ll + 1));
        for (int q = 0; q <= ll; q++) {
          int l = lst[q];
          double z = B->ss[2 * (size_t)N + l];
          double b1 = cos(ra1 * z), b2 = cos(3 * ra1 * z), b3 = cos(4 * ra1 * z);
          ra1 = half * exp(-ra0 * r2v[l]) / B->yaco[l];
          B->xafor[q] = ra1 * b1;
          B->xafor[(ll + 1) + q] = ra1 * b2;
          B->xafor[2 * (size_t)(ll + 1) + q] = ra1 * b3;
        }
      } else {
        free(lst);
      }
      free(r2v);
    }
  }
}

/* src/SingleTurbulence3D.f90:203 */
void orc_initialo(orc *c) {
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    for (int l = 0; l < N; l++) {
      B->qa[l] = c->rhooo;
      B->qa[N + l] = zero; B->qa[2 * N + l] = zero; B->qa[3 * N + l] = zero;
      B->qa[4 * N + l] = hamm1 * c->poo;
    }
  }
  c->sc[ORC_S_N] = 0; c->sc[ORC_S_NDT] = 0; c->sc[ORC_S_DT] = 0;
  c->sc[ORC_S_DTS] = 0; c->sc[ORC_S_DTE] = 0; c->sc[ORC_S_TIMO] = 0;
}

/* ---------------------------------------------------------------------- */
/* the time step: src/Main3D.f90:552-965                                     */
/* ---------------------------------------------------------------------- */

static void movef_c(orc *c, double dtko, double dtk) {
  double aoa = c->sc[ORC_S_AOA];
  orc_movef(c->nsmf, c->timf, c->sc[ORC_S_TIMO], dtko, dtk, c->uoo, c->tp, c->tps,
            c->aoa0, c->aoa1, c->amachoo, &c->sc[ORC_S_UMF1], &c->sc[ORC_S_DUDT1], &aoa);
  c->sc[ORC_S_AOA] = aoa;
  c->sc[ORC_S_DTK] = dtk; c->sc[ORC_S_DTKO] = dtko;
}

/* src/Main3D.f90:554-561 */
void orc_filter_step(orc *c) {
  for (int m = 1; m <= 5; m++) {
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      memcpy(B->rr, B->qa + (size_t)(m - 1) * B->N, sizeof(double) * B->N);
    }
    orc_mpigo(c, 1, 1, 0);
    orc_filte(c, 1, 1);
    orc_filte(c, 2, 1);
    orc_filte(c, 3, 1);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      memcpy(B->qa + (size_t)(m - 1) * B->N, B->rr, sizeof(double) * B->N);
    }
  }
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    memcpy(B->qo, B->qa, sizeof(double) * 5 * B->N);
  }
}

/* gradient of one field and its metric combination, src/Main3D.f90:629-651 */
static void viscous_grad(orc *c, int m, int src_de, double **dst) {
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    memcpy(B->rr, B->de + (size_t)src_de * B->N, sizeof(double) * B->N);
  }
  orc_mpigo(c, 0, 1, 0);
  orc_deriv(c, 3, 1, m);
  orc_deriv(c, 2, 1, m);
  orc_deriv(c, 1, 1, m);
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    double *o1 = dst[b * 3 + 0], *o2 = dst[b * 3 + 1], *o3 = dst[b * 3 + 2];
    const double *r1 = B->rr, *r2 = B->rr + N, *r3 = B->rr + 2 * N;
    const double *xi = B->xim, *et = B->etm, *ze = B->zem;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int l = 0; l < N; l++) {
      double v1 = xi[l] * r1[l] + et[l] * r2[l] + ze[l] * r3[l];
      double v2 = xi[N + l] * r1[l] + et[N + l] * r2[l] + ze[N + l] * r3[l];
      double v3 = xi[2 * N + l] * r1[l] + et[2 * N + l] * r2[l] + ze[2 * N + l] * r3[l];
      o1[l] = v1; o2[l] = v2; o3[l] = v3;
    }
  }
}

/* One sample of the surface statistics, when sampling is on and n is a
 * multiple of its interval: the force coefficients of src/rpt.f90:1017-1056
 * (clpost: pressure and wall-shear force over the wall cells, per block
 * (clp + clv) dynp with dynp = 2/(M^2 span), summed over the blocks as the
 * reference's MPI_REDUCE does; wall shear tw = rho (tau . n) in the
 * reference's form), turned into lift and drag with the angle of attack as
 * src/Main3D.f90:700-704 prints them; and trapezoidal time integrals, over
 * the sampled times, of Cl, Cd and of p and p^2 at every wall point -- the
 * running form of the time averages of src/rptPost.f90:870-925 (p3dStats). */
static void wall_sample(orc *c) {
  double *st = c->st, *sc = c->sc;
  if (!(st[0] > 0) || fmod(sc[ORC_S_N], st[0]) != 0.0) return;
  const double t = sc[ORC_S_TIMO];
  const double dynp = two / (c->amachoo * c->amachoo * c->span);
  double cl[2] = {zero, zero};
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    if (B->nsurf <= 0) continue;
    const int N = B->N, nwl = B->nsurf, j = B->sside ? B->let : 0;
    for (int dir = 0; dir < 2; dir++) {
      double clp = zero, clv = zero;
      int ll = -1;
      for (int k = 0; k <= B->lze; k++)
        for (int i = 0; i <= B->lxi; i++) {
          int l = IDX3(B, j, k, i, 1);
          ll = ll + 1;
          clp = clp + (B->p[l] * B->snor[(size_t)dir * nwl + ll] * B->sarea[ll]);
        }
      if (c->nviscous == 1) {
        ll = -1;
        for (int k = 0; k <= B->lze; k++)
          for (int i = 0; i <= B->lxi; i++) {
            int l = IDX3(B, j, k, i, 1);
            ll = ll + 1;
            const double *tv = B->tv;
            double n1 = B->snor[ll], n2 = B->snor[(size_t)nwl + ll], n3 = B->snor[2 * (size_t)nwl + ll];
            double tw = dir == 0 ? B->qa[l] * (tv[l] * n1 + tv[3 * N + l] * n2 + tv[5 * N + l] * n3)
                                 : B->qa[l] * (tv[3 * N + l] * n1 + tv[N + l] * n2 + tv[4 * N + l] * n3);
            clv = clv + tw * B->sarea[ll];
          }
      }
      cl[dir] = cl[dir] + (clp + clv) * dynp;
    }
    /* p statistics per wall point */
    int ll = -1;
    for (int k = 0; k <= B->lze; k++)
      for (int i = 0; i <= B->lxi; i++) {
        int l = IDX3(B, j, k, i, 1);
        ll = ll + 1;
        double p = B->p[l], *ps = B->sst;
        if (st[3] > 0) {
          double dt = t - st[2];
          ps[(size_t)nwl + ll] = ps[(size_t)nwl + ll] + half * dt * (ps[ll] + p);
          ps[2 * (size_t)nwl + ll] = ps[2 * (size_t)nwl + ll] + half * dt * (ps[ll] * ps[ll] + p * p);
        }
        ps[ll] = p;
      }
  }
  double raoa = sc[ORC_S_AOA] * PI / 180;
  double clift = cl[1] * cos(raoa) - cl[0] * sin(raoa);
  double cdrag = cl[1] * sin(raoa) + cl[0] * cos(raoa);
  if (st[3] > 0) {
    double dt = t - st[2];
    st[6] = st[6] + half * dt * (st[4] + clift);
    st[7] = st[7] + half * dt * (st[5] + cdrag);
  } else {
    st[1] = t;
  }
  st[4] = clift; st[5] = cdrag; st[2] = t; st[3] = st[3] + 1;
}

void orc_stats_start(orc *c, int every) {
  for (int q = 0; q < 8; q++) c->st[q] = zero;
  c->st[0] = every;
  for (int b = 0; b < c->nb; b++)
    if (c->b[b].nsurf > 0) memset(c->b[b].sst, 0, sizeof(double) * 3 * (size_t)c->b[b].nsurf);
}

void orc_stats_forces(const orc *c, double *out8) {
  const double *st = c->st;
  double T = st[2] - st[1];
  out8[0] = st[3]; out8[1] = st[1]; out8[2] = st[2]; out8[3] = st[4]; out8[4] = st[5];
  out8[5] = T > 0 ? st[6] / T : st[4];
  out8[6] = T > 0 ? st[7] / T : st[5];
  out8[7] = T;
}

int orc_stats_wall(const orc *c, int b, double *mean_p, double *rms_p, double *area, double *wnor3) {
  const oblk *B = &c->b[b];
  const int nwl = B->nsurf;
  if (nwl <= 0) return 0;
  double T = c->st[2] - c->st[1];
  for (int ll = 0; ll < nwl; ll++) {
    double m = T > 0 ? B->sst[(size_t)nwl + ll] / T : B->sst[ll];
    double v = T > 0 ? B->sst[2 * (size_t)nwl + ll] / T - m * m : zero;
    if (mean_p) mean_p[ll] = m;
    if (rms_p) rms_p[ll] = sqrt(v > zero ? v : zero);
    if (area) area[ll] = B->sarea[ll];
    if (wnor3)
      for (int q = 0; q < 3; q++) wnor3[(size_t)q * nwl + ll] = B->snor[(size_t)q * nwl + ll];
  }
  return nwl;
}

/* stage start (frame velocity for the stage, src/Main3D.f90:571-572), primitives
 * (:576-583) and the time step size at the first stage (:587-622) */
/* x**1.5 (the viscosity law, src/Main3D.f90:582) rounded correctly: the
 * reference leaves the last bit to its compiler's pow, which differs between
 * libraries, so the oracle fixes the correctly rounded value.  y = x sqrt(x)
 * is within two units in the last place of x^1.5; the residual x^3 - y^2 is
 * formed from error-free products (fma), its leading difference exact by
 * Sterbenz's lemma, and one Newton step y + (x^3 - y^2) / 2y lands on the
 * nearest double unless x^1.5 lies within ~2^-50 units of a rounding midpoint.
 * A perfect square x gives sqrt(x) exactly, and x sqrt(x) is then already
 * rounded once.  Outside [2^-300, 2^300] (and for x <= 0, NaN) it is pow's. */
double orc_pow15(double x) {
  if (!(x > 0x1p-300 && x < 0x1p300)) return pow(x, 1.5);
  double s = sqrt(x), y = x * s;
  if (fma(s, s, -x) == 0.0) return y;
  double P = y * y, p = fma(y, y, -P);  /* y^2 = P + p */
  double Q = x * x, q = fma(x, x, -Q);  /* x^2 = Q + q */
  double S = x * Q, e = fma(x, Q, -S);  /* x Q = S + e */
  double r = (S - P) + ((e - p) + x * q);
  return y + r / (y + y);
}

static void stage_begin(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* src/Main3D.f90:571-572 */
  {
    double dtko = dt * (double)(nk - 2 > 0 ? (nk - 2 < 1 ? nk - 2 : 1) : 0) / (nkrk - nk + 3);
    double dtk = dt * (double)(nk - 1 < 1 ? nk - 1 : 1) / (nkrk - nk + 2);
    movef_c(c, dtko, dtk);
  }
  /* primitives, src/Main3D.f90:576-583 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    double *qa = B->qa, *de = B->de;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int l = 0; l < N; l++) {
      double d1 = one / qa[l];
      double d2 = qa[N + l] * d1, d3 = qa[2 * N + l] * d1, d4 = qa[3 * N + l] * d1;
      double p = gamm1 * (qa[4 * N + l] - half * (qa[N + l] * d2 + qa[2 * N + l] * d3 + qa[3 * N + l] * d4));
      double d5 = gam * p * d1;
      de[l] = d1; de[N + l] = d2; de[2 * N + l] = d3; de[3 * N + l] = d4; de[4 * N + l] = d5;
      B->p[l] = p;
      B->ss[l] = c->srefp1dre * orc_pow15(d5) / (d5 + c->srefoo);
    }
  }
  /* time step size, src/Main3D.f90:587-610 */
  if (nk == 1) {
    int n = (int)sc[ORC_S_N];
    if (n % 10 == 1) {
      sc[ORC_S_NDT] = n;
      sc[ORC_S_DTS] = sc[ORC_S_DTE];
      if (c->dto < zero) {
        double fi = -INFINITY, fv = -INFINITY;
        for (int b = 0; b < c->nb; b++) {
          oblk *B = &c->b[b];
          int N = B->N;
          const double *de = B->de, *xi = B->xim, *et = B->etm, *ze = B->zem;
#pragma omp parallel for num_threads(nthreads()) reduction(max : fi, fv) schedule(static)
          for (int l = 0; l < N; l++) {
            double r1 = xi[l] * xi[l] + xi[N + l] * xi[N + l] + xi[2 * N + l] * xi[2 * N + l] +
                        et[l] * et[l] + et[N + l] * et[N + l] + et[2 * N + l] * et[2 * N + l] +
                        ze[l] * ze[l] + ze[N + l] * ze[N + l] + ze[2 * N + l] * ze[2 * N + l];
            double u = de[N + l] + umf[0], v = de[2 * N + l] + umf[1], w = de[3 * N + l] + umf[2];
            double r2 = fabs(xi[l] * u + xi[N + l] * v + xi[2 * N + l] * w) +
                        fabs(et[l] * u + et[N + l] * v + et[2 * N + l] * w) +
                        fabs(ze[l] * u + ze[N + l] * v + ze[2 * N + l] * w);
            double s2 = fabs(B->yaco[l]);
            double a = (sqrt(de[4 * N + l] * r1) + r2) * s2;
            if (a > fi) fi = a;
            double vv = de[l] * B->ss[l] * r1 * s2 * s2;
            if (vv > fv) fv = vv;
          }
        }
        double ra0 = c->cfl / fi, ra1 = ra0;
        if (c->nviscous == 1) ra1 = half / fv;
        sc[ORC_S_DTE] = fmin(ra0, ra1);
      } else {
        sc[ORC_S_DTE] = c->dto;
      }
    }
    double s = crsin(0.05 * PI * (double)(n - (int)sc[ORC_S_NDT]));
    sc[ORC_S_DT] = sc[ORC_S_DTS] + (sc[ORC_S_DTE] - sc[ORC_S_DTS]) * (s * s);
    dt = sc[ORC_S_DT];
  }
}

/* viscous stresses and heat fluxes into tv, src/Main3D.f90:626-693 */
static void stage_viscous(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* viscous stresses and heat fluxes, src/Main3D.f90:626-693 */
  if (c->nviscous == 1) {
    double **dst = (double **)malloc(sizeof(double *) * 3 * c->nb);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      memcpy(B->de, B->ss, sizeof(double) * B->N); /* de(:,1)=ss(:,1) */
    }
    /* tv order: 0 txx 1 tyy 2 tzz 3 txy 4 tyz 5 tzx 6 hxx 7 hyy 8 hzz */
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b]; int N = B->N;
      dst[b * 3 + 0] = B->tv + 0 * N; dst[b * 3 + 1] = B->tv + 8 * N; dst[b * 3 + 2] = B->tv + 5 * N;
    }
    viscous_grad(c, 2, 1, dst);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b]; int N = B->N;
      dst[b * 3 + 0] = B->tv + 3 * N; dst[b * 3 + 1] = B->tv + 1 * N; dst[b * 3 + 2] = B->tv + 6 * N;
    }
    viscous_grad(c, 3, 2, dst);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b]; int N = B->N;
      dst[b * 3 + 0] = B->tv + 7 * N; dst[b * 3 + 1] = B->tv + 4 * N; dst[b * 3 + 2] = B->tv + 2 * N;
    }
    viscous_grad(c, 4, 3, dst);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b]; int N = B->N;
      dst[b * 3 + 0] = B->ss; dst[b * 3 + 1] = B->ss + N; dst[b * 3 + 2] = B->ss + 2 * N;
    }
    viscous_grad(c, 5, 4, dst);
    free(dst);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      int N = B->N;
      double *de = B->de, *tv = B->tv, *ss = B->ss, *qa = B->qa, *ya = B->yaco;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
      for (int l = 0; l < N; l++) {
        double txx = tv[l], tyy = tv[N + l], tzz = tv[2 * N + l], txy = tv[3 * N + l],
               tyz = tv[4 * N + l], tzx = tv[5 * N + l], hxx = tv[6 * N + l],
               hyy = tv[7 * N + l], hzz = tv[8 * N + l];
        double r1 = de[l] * ya[l];
        double r2 = gamm1prndtli * r1, d5;
        if (c->les == 1) {
          double d1 = txx * txx + tyy * tyy + tzz * tzz + (hzz + txy) * (hzz + txy) +
                      (hyy + tzx) * (hyy + tzx) + (hxx + tyz) * (hxx + tyz);
          double varr = (double)(float)pow(-1 / ya[l], 1.5);
          double r3 = qa[l] * (c->smago1 * c->smago1) * varr * sqrt(2 * d1);
          r1 = r1 + r3 * ya[l];
          r2 = r2 + tgamm1prndtli * r3 * ya[l];
          r3 = two_three * (qa[l] * c->smago2 * varr * d1);
          d5 = two_three * (txx + tyy + tzz);
          double ntxx = r1 * (2 * txx - d5) - r3, ntyy = r1 * (2 * tyy - d5) - r3,
                 ntzz = r1 * (2 * tzz - d5) - r3;
          double ntxy = r1 * (txy + hzz), ntyz = r1 * (tyz + hxx), ntzx = r1 * (tzx + hyy);
          txx = ntxx; tyy = ntyy; tzz = ntzz; txy = ntxy; tyz = ntyz; tzx = ntzx;
          de[l] = d1;
        } else {
          d5 = two_three * (txx + tyy + tzz);
          double ntxx = r1 * (2 * txx - d5), ntyy = r1 * (2 * tyy - d5), ntzz = r1 * (2 * tzz - d5);
          double ntxy = r1 * (txy + hzz), ntyz = r1 * (tyz + hxx), ntzx = r1 * (tzx + hyy);
          txx = ntxx; tyy = ntyy; tzz = ntzz; txy = ntxy; tyz = ntyz; tzx = ntzx;
        }
        de[4 * N + l] = d5;
        double u = de[N + l], v = de[2 * N + l], w = de[3 * N + l];
        hxx = r2 * ss[l] + u * txx + v * txy + w * tzx;
        hyy = r2 * ss[N + l] + u * txy + v * tyy + w * tyz;
        hzz = r2 * ss[2 * N + l] + u * tzx + v * tyz + w * tzz;
        tv[l] = txx; tv[N + l] = tyy; tv[2 * N + l] = tzz; tv[3 * N + l] = txy;
        tv[4 * N + l] = tyz; tv[5 * N + l] = tzx; tv[6 * N + l] = hxx; tv[7 * N + l] = hyy;
        tv[8 * N + l] = hzz;
        B->rr[l] = r1; B->rr[N + l] = r2;
      }
    }
  }
}

/* surface sample at the first stage (:695-706), contravariant velocities and the
 * divergence of the metric-weighted fluxes into de, src/Main3D.f90:708-768 */
static void stage_flux(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* surface forces and pressure statistics at the first stage, src/Main3D.f90:695-706 */
  if (nk == 1) wall_sample(c);
  /* fluxes, src/Main3D.f90:710-768 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    double *de = B->de, *ss = B->ss, *xi = B->xim, *et = B->etm, *ze = B->zem;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int l = 0; l < N; l++) {
      double r1 = de[N + l] + umf[0], r2 = de[2 * N + l] + umf[1], r3 = de[3 * N + l] + umf[2];
      ss[l] = xi[l] * r1 + xi[N + l] * r2 + xi[2 * N + l] * r3;
      ss[N + l] = et[l] * r1 + et[N + l] * r2 + et[2 * N + l] * r3;
      ss[2 * N + l] = ze[l] * r1 + ze[N + l] * r2 + ze[2 * N + l] * r3;
    }
  }
  for (int m = 1; m <= 5; m++) {
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      int N = B->N;
      double *qa = B->qa, *de = B->de, *ss = B->ss, *xi = B->xim, *et = B->etm, *ze = B->zem,
             *p = B->p, *tv = B->tv, *rr = B->rr;
      int visc = c->nviscous == 1;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
      for (int l = 0; l < N; l++) {
        double r1, r2, r3;
        double x1 = xi[l], x2 = xi[N + l], x3 = xi[2 * N + l];
        double e1 = et[l], e2 = et[N + l], e3 = et[2 * N + l];
        double z1 = ze[l], z2 = ze[N + l], z3 = ze[2 * N + l];
        double s1 = ss[l], s2 = ss[N + l], s3 = ss[2 * N + l], pl = p[l];
        if (m == 1) {
          r1 = qa[l] * s1; r2 = qa[l] * s2; r3 = qa[l] * s3;
        } else if (m <= 4) {
          double q = qa[(m - 1) * N + l];
          double mx = m == 2 ? x1 : m == 3 ? x2 : x3;
          double me = m == 2 ? e1 : m == 3 ? e2 : e3;
          double mz = m == 2 ? z1 : m == 3 ? z2 : z3;
          r1 = q * s1 + mx * pl; r2 = q * s2 + me * pl; r3 = q * s3 + mz * pl;
          if (visc) {
            double ta, tb, tc;
            if (m == 2) { ta = tv[l]; tb = tv[3 * N + l]; tc = tv[5 * N + l]; }
            else if (m == 3) { ta = tv[3 * N + l]; tb = tv[N + l]; tc = tv[4 * N + l]; }
            else { ta = tv[5 * N + l]; tb = tv[4 * N + l]; tc = tv[2 * N + l]; }
            r1 = r1 - x1 * ta - x2 * tb - x3 * tc;
            r2 = r2 - e1 * ta - e2 * tb - e3 * tc;
            r3 = r3 - z1 * ta - z2 * tb - z3 * tc;
          }
        } else {
          double d5 = qa[4 * N + l] + pl;
          de[4 * N + l] = d5;
          r1 = d5 * s1 - pl * (umf[0] * x1 + umf[1] * x2 + umf[2] * x3);
          r2 = d5 * s2 - pl * (umf[0] * e1 + umf[1] * e2 + umf[2] * e3);
          r3 = d5 * s3 - pl * (umf[0] * z1 + umf[1] * z2 + umf[2] * z3);
          if (visc) {
            double ha = tv[6 * N + l], hb = tv[7 * N + l], hc = tv[8 * N + l];
            r1 = r1 - x1 * ha - x2 * hb - x3 * hc;
            r2 = r2 - e1 * ha - e2 * hb - e3 * hc;
            r3 = r3 - z1 * ha - z2 * hb - z3 * hc;
          }
        }
        rr[l] = r1; rr[N + l] = r2; rr[2 * N + l] = r3;
      }
    }
    orc_mpigo(c, 0, 0, 0);
    orc_deriv(c, 1, 1, m);
    orc_deriv(c, 2, 2, m);
    orc_deriv(c, 3, 3, m);
    for (int b = 0; b < c->nb; b++) {
      oblk *B = &c->b[b];
      int N = B->N;
      double *de = B->de + (size_t)(m - 1) * N, *rr = B->rr;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
      for (int l = 0; l < N; l++) de[l] = rr[l] + rr[N + l] + rr[2 * N + l];
    }
  }
}

/* sponge relaxation with the inflow target, src/SingleTurbulence3D.f90:332-374 */
static void stage_sponge(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* sponge, src/SingleTurbulence3D.f90:332-372 */
  int lag_on = 0, il5[5];
  double tl5[5], w5[5];
  if (c->inflow) lag_on = orc_inflow_weights(c, c->sc[ORC_S_TIMO], c->sc[ORC_S_DTK], il5, tl5, w5);
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    for (int q = 0; q <= B->lsz; q++) {
      int l = B->lcsz[q];
      B->rr[l] = one;
      B->ss[l] = zero; B->ss[N + l] = zero; B->ss[2 * N + l] = zero;
    }
    if (B->vito) {
      int nt1 = c->ntz + 1, ng = B->ltz + 1;
      double fctr = half * gamm1;
      for (int g = 0; g < ng; g++) {
        double vit[3] = {zero, zero, zero};
        if (lag_on)
          for (int jj = 0; jj < 5; jj++)
            for (int d = 0; d < 3; d++)
              vit[d] = vit[d] + w5[jj] * B->vito[((size_t)d * nt1 + il5[jj]) * ng + g];
        int l = B->lcsz[B->lctz[g]];
        B->rr[l] = pow(one - fctr * (vit[0] * vit[0] + vit[1] * vit[1] + vit[2] * vit[2]), hamm1);
        for (int d = 0; d < 3; d++) B->ss[(size_t)d * N + l] = B->atz[g] * B->rr[l] * vit[d];
      }
    }
    for (int q = 0; q <= B->lsz; q++) {
      int l = B->lcsz[q];
      double a = B->asz[q];
      B->de[l] = B->de[l] + a * (B->qa[l] - B->rr[l]);
      for (int d = 0; d < 3; d++)
        B->de[(1 + d) * N + l] = B->de[(1 + d) * N + l] + a * (B->qa[(1 + d) * N + l] - B->ss[d * N + l]);
      B->de[4 * N + l] = B->de[4 * N + l] + B->bsz[q] * (B->p[l] - ham * pow(B->rr[l], gam));
    }
  }
  /* body force, forcego (src/rpt.f90:758-769; src/Main3D.f90:776 calls it
   * right after the sponge) */
  if (c->forcing == 1) {
    const double t = sc[ORC_S_TIMO] + sc[ORC_S_DTK];
    if (t > c->tsfor && t < c->tefor) {
      const double ra0 = t - c->tsfor;
      const double ra1 = c->amfor * crcos(ra0 * 48.76 * c->amachoo) / 3;
      const double ra2 = c->amfor * crcos(ra0 * 53.6 * c->amachoo) / 3;
      const double ra3 = c->amfor * crcos(ra0 * 53.6 * c->amachoo) / 3;
      for (int b = 0; b < c->nb; b++) {
        oblk *B = &c->b[b];
        const int N = B->N, nf = B->lfor + 1;
        for (int q = 0; q < nf; q++) {
          const int l = B->lcfor[q];
          const double f = ra1 * B->xafor[q] + ra2 * B->xafor[nf + q] + ra3 * B->xafor[2 * (size_t)nf + q];
          B->de[N + l] = B->de[N + l] + B->qa[l] * f;
          B->de[2 * N + l] = B->de[2 * N + l] - B->qa[l] * f;
        }
      }
    }
  }
}

/* characteristic boundary and interface conditions, src/Main3D.f90:778-869 */
static void stage_gcbc(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* GCBC/GCIC preparation, src/Main3D.f90:780-795 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    for (int nn = 0; nn < 3; nn++) {
      int F = B->nbsize[nn];
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp], i = ipp * B->ijk[nn][0];
        if (!(np == 10 || np == 20 || np == 25 || np == 30)) continue;
        double ra0 = (double)((20 - np) * (25 - np) * (30 - np) / 3000), ra1 = one - ra0;
        double *drva = B->drva[nn], *cm = B->cm[nn];
        for (int k = 0; k <= B->ijk[nn][2]; k++)
          for (int j = 0; j <= B->ijk[nn][1]; j++) {
            int jk = k * (B->ijk[nn][1] + 1) + j, l = IDX3(B, i, j, k, nn);
            double q[5], cmv[3], cha[5], xt[5][5];
            elem_t e;
            for (int v = 0; v < 5; v++) q[v] = B->qa[(size_t)v * N + l];
            for (int d = 0; d < 3; d++) cmv[d] = cm[(ipp * 3 + d) * F + jk];
            eleme(q, B->p[l], cmv, umf, &e);
            xtq2r(&e, cmv, xt);
            for (int v = 0; v < 5; v++)
              cha[v] = ra0 * drva[(ipp * 5 + v) * F + jk] + ra1 * B->de[(size_t)v * N + l];
            for (int r = 0; r < 5; r++) {
              double s = 0;
              for (int v = 0; v < 5; v++) s = s + xt[r][v] * (B->yaco[l] * cha[v]);
              drva[(ipp * 5 + r) * F + jk] = s;
            }
          }
      }
    }
  }
  /* GCIC exchange, src/Main3D.f90:799-813 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    for (int nn = 0; nn < 3; nn++)
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp];
        if (np == 30 || (c->nextgcic == 1 && (np == 20 || np == 25))) {
          oblk *Nb = &c->b[B->ncd[nn][ipp]];
          int F = B->nbsize[nn];
          memcpy(B->drvb[nn] + (size_t)ipp * 5 * F, Nb->drva[nn] + (size_t)(1 - ipp) * 5 * F,
                 sizeof(double) * 5 * F);
        }
      }
  }
  /* GCBC/GCIC application, src/Main3D.f90:817-869 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    int lp = -1, ll = -1;
    for (int nn = 0; nn < 3; nn++) {
      int F = B->nbsize[nn];
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp], i = ipp * B->ijk[nn][0], iq = 1 - 2 * ipp;
        double ra0 = iq;
        if (!(np == 10 || np == 20 || np == 25 || np == 30)) continue;
        double *drva = B->drva[nn], *drvb = B->drvb[nn], *cm = B->cm[nn];
        for (int k = 0; k <= B->ijk[nn][2]; k++)
          for (int j = 0; j <= B->ijk[nn][1]; j++) {
            int jk = k * (B->ijk[nn][1] + 1) + j, l = IDX3(B, i, j, k, nn);
            double q[5], cmv[3], cha[5], dha[5], xt[5][5];
            elem_t e;
            for (int v = 0; v < 5; v++) q[v] = B->qa[(size_t)v * N + l];
            for (int d = 0; d < 3; d++) cmv[d] = cm[(ipp * 3 + d) * F + jk];
            eleme(q, B->p[l], cmv, umf, &e);
            for (int v = 0; v < 5; v++) {
              cha[v] = drva[(ipp * 5 + v) * F + jk];
              dha[v] = drvb[(ipp * 5 + v) * F + jk];
            }
            double vv = e.vn + e.vs;
            if (np == 10) {
              if (ra0 * (vv + e.ao) > zero) cha[3] = zero;
              if (ra0 * (vv - e.ao) > zero) cha[4] = zero;
            } else if (np == 20 || np == 25) {
              lp++;
              if (B->rpex[lp] == zero) {
                const double *dd = &c->sc[ORC_S_DUDT1];
                double s = cmv[0] * dd[0] + cmv[1] * dd[1] + cmv[2] * dd[2];
                cha[3 + ipp] = cha[4 - ipp] + two * ra0 * e.aoi * q[0] * (s + 50.0 * vv);
              } else {
                /* extrabcs, src/SingleTurbulence3D.f90:468-474 */
                if (ra0 * vv > zero) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }
                if (ra0 * (vv + e.ao) > zero) cha[3] = dha[3];
                if (ra0 * (vv - e.ao) > zero) cha[4] = dha[4];
              }
            } else {
              if (ra0 * vv > zero) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }
              if (ra0 * (vv + e.ao) > zero) cha[3] = dha[3];
              if (ra0 * (vv - e.ao) > zero) cha[4] = dha[4];
            }
            xtr2q(&e, cmv, xt);
            for (int r = 0; r < 5; r++) {
              double s = 0;
              for (int v = 0; v < 5; v++) s = s + xt[r][v] * (cha[v] - drva[(ipp * 5 + v) * F + jk]);
              dha[r] = s;
            }
            for (int ii = 0; ii <= MBCI; ii++) {
              int l2 = IDX3(B, i + iq * ii, j, k, nn);
              ll++;
              for (int v = 0; v < 5; v++)
                B->de[(size_t)v * N + l2] = B->de[(size_t)v * N + l2] + B->sbcc[ll] * dha[v];
            }
          }
      }
    }
  }
}

/* frame velocity for the update (:873-874) and the RK update, src/Main3D.f90:876-881 */
static void stage_rk(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* RK update, src/Main3D.f90:873-881 */
  {
    double dtko = dt * (double)(nk - 1 < 1 ? nk - 1 : 1) / (nkrk - nk + 2);
    double dtk = dt / (nkrk - nk + 1);
    movef_c(c, dtko, dtk);
  }
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    double dtk = sc[ORC_S_DTK];
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int l = 0; l < N; l++) {
      double r1 = dtk * B->yaco[l];
      B->rr[l] = r1;
      for (int v = 0; v < 5; v++)
        B->qa[(size_t)v * N + l] = B->qo[(size_t)v * N + l] - r1 * B->de[(size_t)v * N + l];
    }
  }
}

/* wall temperature and velocity, src/Main3D.f90:884-895 */
static void stage_wall(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* wall temperature and velocity, src/Main3D.f90:886-895 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N, lp = -1;
    double ra0 = ham * hamm1 * c->wtemp;
    for (int nn = 0; nn < 3; nn++)
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp], i = ipp * B->ijk[nn][0];
        if (!(np == 20 || np == 25)) continue;
        double ra1 = (double)((np - 20) / 5);
        for (int k = 0; k <= B->ijk[nn][2]; k++)
          for (int j = 0; j <= B->ijk[nn][1]; j++) {
            int l = IDX3(B, i, j, k, nn);
            lp++;
            double ra2 = one - B->rpex[lp], ra3 = one - ra1 * ra2;
            double *qa = B->qa;
            double fctr = (one - ra3) * qa[l];
            for (int d = 0; d < 3; d++) qa[(1 + d) * N + l] = ra3 * qa[(1 + d) * N + l] - fctr * umf[d];
            qa[4 * N + l] = B->rpex[lp] * qa[4 * N + l] +
                            ra2 * (ra0 * qa[l] + half * (qa[N + l] * qa[N + l] + qa[2 * N + l] * qa[2 * N + l] +
                                                       qa[3 * N + l] * qa[3 * N + l]) / qa[l]);
          }
      }
  }
}

/* corner junctions, src/SingleTurbulence3D.f90:378-443 */
static void stage_junction(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* junction, src/SingleTurbulence3D.f90:378-443 */
  for (int jn = 0; jn < c->njct; jn++) {
    const int *J = c->jct + jn * 12;
    oblk *M[4];
    int lze = c->b[J[0]].lze;
    for (int q = 0; q < 4; q++) M[q] = &c->b[J[q * 3]];
    int va = c->jvel[jn * 2 + 0], vb = c->jvel[jn * 2 + 1];
    for (int k = 0; k <= lze; k++) {
      double v[4][5];
      int lq[4];
      for (int q = 0; q < 4; q++) {
        oblk *B = M[q];
        int i = J[q * 3 + 1] * B->lxi, j = J[q * 3 + 2] * B->let;
        lq[q] = IDX3(B, i, j, k, 0);
        for (int m = 0; m < 5; m++) v[q][m] = B->qa[(size_t)m * B->N + lq[q]];
      }
      double r[5];
      r[0] = 0.25 * (v[0][0] + v[1][0] + v[2][0] + v[3][0]);
      for (int m = 1; m <= 3; m++) r[m] = half * (v[va][m] + v[vb][m]);
      r[4] = 0.25 * (v[0][4] + v[1][4] + v[2][4] + v[3][4]);
      for (int q = 0; q < 4; q++)
        for (int m = 0; m < 5; m++) M[q]->qa[(size_t)m * M[q]->N + lq[q]] = r[m];
    }
  }
}

/* interface averaging, src/Main3D.f90:897-952 */
static void stage_avg(orc *c, int nk) {
  double *sc = c->sc;
  const int nkrk = c->nkrk;
  double dt = sc[ORC_S_DT];
  const double *umf = &sc[ORC_S_UMF1];
  (void)sc; (void)nkrk; (void)dt; (void)umf; (void)nk;
  /* interface averaging, src/Main3D.f90:901-952 */
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N;
    for (int nn = 0; nn < 3; nn++) {
      int F = B->nbsize[nn];
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp], i = ipp * B->ijk[nn][0];
        if (np == 30 || np == 35 || np == 45 || (c->nextgcic == 1 && (np == 20 || np == 25))) {
          for (int k = 0; k <= B->ijk[nn][2]; k++)
            for (int j = 0; j <= B->ijk[nn][1]; j++) {
              int jk = k * (B->ijk[nn][1] + 1) + j, l = IDX3(B, i, j, k, nn);
              for (int v = 0; v < 5; v++) B->drva[nn][(ipp * 5 + v) * F + jk] = B->qa[(size_t)v * N + l];
              B->rr[l] = one;
            }
        }
      }
    }
  }
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    for (int nn = 0; nn < 3; nn++)
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp];
        if (np == 30 || np == 35 || np == 45 || (c->nextgcic == 1 && (np == 20 || np == 25))) {
          oblk *Nb = &c->b[B->ncd[nn][ipp]];
          int F = B->nbsize[nn];
          memcpy(B->drvb[nn] + (size_t)ipp * 5 * F, Nb->drva[nn] + (size_t)(1 - ipp) * 5 * F,
                 sizeof(double) * 5 * F);
        }
      }
  }
  for (int b = 0; b < c->nb; b++) {
    oblk *B = &c->b[b];
    int N = B->N, lp = -1;
    double *rr = B->rr, *qa = B->qa;
    for (int nn = 0; nn < 3; nn++) {
      int F = B->nbsize[nn];
      for (int ipp = 0; ipp <= 1; ipp++) {
        int np = B->nbc[nn][ipp], i = ipp * B->ijk[nn][0];
        const double *drvb = B->drvb[nn];
        if (np == 30 || np == 35 || np == 45) {
          for (int k = 0; k <= B->ijk[nn][2]; k++)
            for (int j = 0; j <= B->ijk[nn][1]; j++) {
              int jk = k * (B->ijk[nn][1] + 1) + j, l = IDX3(B, i, j, k, nn);
              rr[l] = rr[l] + one;
              double r2 = one / rr[l];
              rr[N + l] = r2;
              for (int v = 0; v < 5; v++)
                qa[(size_t)v * N + l] = r2 * ((rr[l] - one) * qa[(size_t)v * N + l] + drvb[(ipp * 5 + v) * F + jk]);
            }
        }
        if (c->nextgcic == 1 && (np == 20 || np == 25)) {
          for (int k = 0; k <= B->ijk[nn][2]; k++)
            for (int j = 0; j <= B->ijk[nn][1]; j++) {
              int jk = k * (B->ijk[nn][1] + 1) + j, l = IDX3(B, i, j, k, nn);
              lp++;
              double ra0 = B->rpex[lp], ra1 = one - ra0;
              rr[l] = rr[l] + ra0;
              double r2 = one / rr[l];
              rr[N + l] = r2;
              for (int v = 0; v < 5; v++)
                qa[(size_t)v * N + l] = ra0 * r2 * ((rr[l] - one) * qa[(size_t)v * N + l] + drvb[(ipp * 5 + v) * F + jk]) +
                                        ra1 * qa[(size_t)v * N + l];
            }
        }
      }
    }
  }
}

/* The stage's sections, indexed by ORC_P_*. */
static void (*const stage_parts[ORC_P_NPART])(orc *, int) = {
    stage_begin, stage_viscous, stage_flux, stage_sponge, stage_gcbc,
    stage_rk, stage_wall, stage_junction, stage_avg};

/* One stage, src/Main3D.f90:567-958: the sections in the reference's order. */
static const int stage_order[ORC_P_NPART] = {
    ORC_P_BEGIN, ORC_P_VISCOUS, ORC_P_FLUX, ORC_P_SPONGE, ORC_P_GCBC,
    ORC_P_RK, ORC_P_WALL, ORC_P_JUNCTION, ORC_P_AVG};

int orc_stage_part(orc *c, int nk, int part) {
  if (nk < 1 || nk > c->nkrk || part < 0 || part >= ORC_P_NPART) {
    snprintf(g_err, sizeof(g_err), "oracle: stage part (%d, %d) out of range", nk, part);
    return 1;
  }
  stage_parts[part](c, nk);
  return 0;
}

void orc_stage(orc *c, int nk) {
  for (int q = 0; q < ORC_P_NPART; q++) stage_parts[stage_order[q]](c, nk);
}

void orc_step(orc *c, int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    orc_filter_step(c);
    for (int nk = 1; nk <= c->nkrk; nk++) orc_stage(c, nk);
    c->sc[ORC_S_N] = c->sc[ORC_S_N] + 1;
    c->sc[ORC_S_TIMO] = c->sc[ORC_S_TIMO] + c->sc[ORC_S_DT];
  }
}

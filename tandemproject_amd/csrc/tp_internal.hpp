// tp_internal.hpp -- device data model of the MI355X time stepper.
//
// One process owns every sub-block of every case.  Each sub-block keeps its
// FP64 fields in HBM in the reference's layout (a(0:lmx, var), xi fastest,
// l = (k*(let+1)+j)*(lxi+1)+i; face buffers face(0:F-1, c, 0:1)), so a field
// column is one contiguous plane-major array and every face exchange is a
// device-side copy between block buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace tp {

constexpr int LMD = 11;   // deriv closure length - 1 (src/MainVar3D.f90:19)
constexpr int LMF = 8;    // filter closure length - 1
constexpr int LMP = 11;
constexpr int MBCI = 3;   // GCBC spreading depth
constexpr int MAX_CASE = 4096;
constexpr int RCV_SLOTS = 5;  // fields whose closure sums one exchange launch can hold

// ---- per-case scalars on the device (src/MainVar3D.f90:139-142) ----
enum Scalar {
  S_N = 0, S_NDT, S_DT, S_DTS, S_DTE, S_TIMO, S_UMF1, S_UMF2, S_UMF3,
  S_DUDT1, S_DUDT2, S_DUDT3, S_DTK, S_DTKO, S_AOA, S_NSCALAR = 16,
  // internal: weights and table rows of the inflow's time interpolation for the current stage
  S_LAGW = 16, S_LAGI = 21, S_LAGON = 26,
  // surface statistics: sampling interval (0 off), first and previous sample
  // time, samples taken, last Cl, Cd, trapezoid integrals of Cl, Cd
  S_ST_EVERY = 27, S_ST_T0 = 28, S_ST_TP = 29, S_ST_NS = 30,
  S_CL = 32, S_CD = 33, S_CLS = 34, S_CDS = 35,
  S_STRIDE = 48
};

// ---- per-case constants (scheme coefficients and physics) ----
struct CaseCoef {
  double abc[3][6];
  double pbci[2][2][LMP + 1];
  double pbco[2][2][LMP + 1];
  double pbcot[2][2];
  double fa, fb, fc, fam[3], fbm[3], fcm[3];
  // the filter's left-hand side: rows 0..2 of a one-sided end (flhs[0]) and
  // of an interface or periodic end (flhs[1]), entries at columns i-2..i+2;
  // interior rows (fbe, fal, 1, fal, fbe)
  double flhs[2][3][5], fal, fbe;
  double sbci[MBCI];
  double fex[3], ffx[3];
  double srefoo, srefp1dre, cfl, dto, wtemp, smago1, smago2;
  double timf, tp, tps, aoa0, aoa1, amachoo, uoo[3];
  int nkrk, nviscous, nsmf, les, nextgcic, pad;
  // turbulent inflow through the sponge target (SingleTurbulence3D.f90:332-372)
  int inflow, ntz;
  double tgusto, period, fctr;   // table period slit/uoo(1) and spacing slit/(ntz uoo(1))
  double span;                   // spanwise extent (force coefficients' reference area per chord)
  // body force (forcego, src/rpt.f90:758-769)
  int forcing, pad2;
  double amfor, tsfor, tefor;
};

// ---- device view of one sub-block ----
struct Block {
  int nx, ny, nz, N;          // points per direction and total
  int L[3];                   // line length per direction (nx, ny, nz)
  int F[3];                   // face size per direction (nbsize)
  int nbc[3][2], ncd[3][2], ndf[3][2];
  int cs;                     // case index
  int lu_off[3];              // offset of direction d in LU tables
  int nlu;                    // LU table stride (lim + 1)
  int wall_off[3][2];         // rpex offset of each wall face (-1 none)
  int nsponge;
  double *qa, *qo, *de, *rr, *ss, *xim, *etm, *zem, *p, *yaco, *tv;
  double *fx;                 // [15][N] scratch: gradients of u,v,w,T, then the 15 flux components
  float *varr;                // [N] LES cases: (-1/yaco)^1.5 in single precision (the
                              // reference's real(4) work array, src/Main3D.f90:664), formed
                              // once from the metrics on the host; null otherwise
  double *drva[3], *drvb[3], *cm[3];
  double *gq;                 // [2][7][F0]: the stage-start q (5), p and yaco at each xi face
                              // point, gathered by the characteristic preparation for its
                              // application (contiguous reads instead of one row each)
  double *recv[3];            // [RCV_SLOTS][6][F]: closure sums per field of a merged launch
  double *lud, *luf;          // [5][nlu]: l1, l2, u1, u2, u3 for deriv / filter
  double *rpex;
  int *lcsz;
  double *asz, *bsz;
  int ngust;                  // inflow points (a subset of the sponge points)
  int *gix;                   // per sponge point: its inflow index, -1 if none
  double *atz;                // per inflow point: ramp sin^2 of the reference
  double *vito;               // [3][ntz+1][ngust] inflow velocity table
  int nforce;                 // forcing points (forceup, src/rpt.f90:728-753)
  int *lcfor;                 // their point indices
  double *xafor;              // [3][nforce] modal coefficients
  // wall surface (the eta face that is a wall): points ll = k*nx + i
  int nsurf, sside;           // wall points (0: none), side (0: j = 0, 1: j = ny-1)
  double *sarea;              // [nsurf] cell areas (src/rpt.f90:774-830, wallArea)
  double *snor;               // [3][nsurf] unit normals into the fluid (walldir)
  double *sst;                // [3][nsurf] previous sampled p, trapezoid integrals of p and p^2
  const double *col[3][2];    // chunked-solver carry maps per direction, kind (deriv, filter)
};

// array ids exposed through the C ABI (order shared with include/tandem_hip.h)
enum ArrayId {
  A_QA = 0, A_QO, A_DE, A_RR, A_SS, A_XIM, A_ETM, A_ZEM, A_P, A_YACO, A_TV,
  A_DRVA1, A_DRVA2, A_DRVA3, A_DRVB1, A_DRVB2, A_DRVB3, A_SEND1, A_SEND2, A_SEND3,
  A_RECV1, A_RECV2, A_RECV3, A_CM1, A_CM2, A_CM3, A_XU, A_XL, A_YU, A_YL, A_RPEX,
  A_NARRAY
};

// field selector for generic line kernels: (array, column)
enum FieldArr { F_RR = 0, F_DE, F_QA, F_SS, F_QO, F_FX };

__host__ __device__ inline double *field_ptr(const Block &b, int arr, int col) {
  double *base = arr == F_RR ? b.rr : arr == F_DE ? b.de : arr == F_QA ? b.qa
               : arr == F_SS ? b.ss : arr == F_QO ? b.qo : b.fx;
  return base + (size_t)col * b.N;
}

// linear index of point (i along direction d, face indices j, k), Subroutineso.f90:364-375
__host__ __device__ inline int indx3(const Block &b, int i, int j, int k, int d) {
  return d == 0 ? (k * b.ny + j) * b.nx + i
       : d == 1 ? (j * b.ny + i) * b.nx + k
                : (i * b.ny + k) * b.nx + j;
}

// work item of a launch: one workgroup's share
struct Work {
  int blk;    // block index
  int a;      // first element (point / line / face point)
  int n;      // count
  int d;      // direction (line and face launches), side in bit 8
};

// Streaming access with the nontemporal hint, for values a pass reads or
// writes once in arrays far larger than the caches (the next reader finds
// them in HBM in any case): the pointwise passes' inputs and outputs and the
// xi line solves' outputs (whole rows).  Not for the line solves' window
// loads, whose halo rows the neighbouring chunks read again from the cache,
// nor for the eta/zeta solves' row stores, whose edge cache lines the
// neighbouring work item completes.
__device__ inline double ld_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ inline void st_stream(double *p, double v) { __builtin_nontemporal_store(v, p); }

// Work item of workgroup bid out of n.  Workgroups go to the GPU's 8 XCDs in
// turn (bid mod 8), each XCD with its own L2; this hands each XCD a
// contiguous run of work items instead (a bijection on [0, n)), so that
// neighbouring items -- adjacent lines or face points, whose edge cache lines
// they share -- meet in the same L2.  Not for kernels whose per-workgroup
// outputs are summed by work-item ranges (those index by blockIdx.x).
__device__ inline int xcd_item(int bid, int n) {
  const int x = bid & 7, k = bid >> 3, per = n >> 3, rem = n & 7;
  return x * per + min(x, rem) + k;
}

// Closure sums that face (d, ip) of block b receives from its neighbour s
// (mpigo, Subroutines3D.f90:16-80): over the neighbour's MP+1 points next to
// the shared face (NT 0: the derivative's closure, MP = LMD; 1: the filter's,
// MP = LMF), r0 and r1 the sums with the closure rows pbco, r2 the first of
// those points, each less the periodic jump terms (n45 = 1 and a periodic
// face: the neighbour's face point, and b's own face point added back).
// src_s: the neighbour's field; own: b's field; jk: the face index.
template <int NT>
__device__ inline void closure_sums(const Block &b, const Block &s, const CaseCoef &c, const double *src_s,
                                    const double *own, int d, int ip, int jk, int n45, double &r0, double &r1,
                                    double &r2) {
  constexpr int MP = NT == 0 ? LMD : LMF;
  const int sp = 1 - ip;
  const int snb = s.nbc[d][sp];
  const int ii = snb == 40 ? 0 : 1;
  const double ra0 = snb == 45 ? (double)n45 : 0.0;
  const int fd = d == 0 ? s.ny : d == 1 ? s.nz : s.nx;   // ijk(2,nn)+1
  const int j = jk % fd, k = jk / fd;
  const int is = sp * (s.L[d] - 1), step = 1 - 2 * sp;
  // every load first, then the sums in the reference's order
  double v[MP + 1];
  // the face point itself enters only through the periodic jump (ra0 != 0)
  const double e0 = ra0 != 0.0 ? src_s[indx3(s, is, j, k, d)] : 0.0;
#pragma unroll
  for (int i = 0; i <= MP; i++) v[i] = src_s[indx3(s, is + step * (i + ii), j, k, d)];
  double ow = 0.0;
  const bool self45 = n45 == 1 && b.nbc[d][ip] == 45;
  if (self45) ow = own[indx3(b, ip * (b.L[d] - 1), j, k, d)];
  const double res = ra0 * e0;
  double t0 = 0, t1 = 0;
#pragma unroll
  for (int i = 0; i <= MP; i++) {
    t0 = t0 + c.pbco[NT][0][i] * v[i];
    t1 = t1 + c.pbco[NT][1][i] * v[i];
  }
  r0 = t0 - res * c.pbcot[NT][0];
  r1 = t1 - res * c.pbcot[NT][1];
  r2 = v[0] - res;
  if (self45) {
    r0 = r0 + ow * c.pbcot[NT][0];
    r1 = r1 + ow * c.pbcot[NT][1];
    r2 = r2 + ow;
  }
}

// Face launches: the face point jk of this thread (256 threads per work item).
// A xi or zeta face takes 256 consecutive face indices: on a zeta face they
// are consecutive in memory, on a xi face each is its own row.  An eta face
// (face index xi*nz + zeta, memory xi-fastest) is taken in tiles of 32 xi x 8
// zeta, work item a = first xi, n = first zeta: a wave's 32-lane rows read
// consecutive doubles of the volume arrays, and the tile's 8 rows share the
// cache lines of the face arrays.
constexpr int FACE_TX = 32, FACE_TZ = 8;
__device__ inline bool face_point(const Block &b, const Work &w, int d, int &jk) {
  const int t = threadIdx.x;
  if (d == 1) {
    const int xi = w.a + (t & (FACE_TX - 1)), ze = w.n + t / FACE_TX;
    jk = xi * b.nz + ze;
    return xi < b.nx && ze < b.nz;
  }
  jk = w.a + t;
  return t < w.n;
}

// kernel classes timed by the profiler hooks
enum KClass {
  K_DERIV = 0, K_FILTER, K_EXCH, K_POINT, K_FLUX, K_VISC, K_GCBC, K_RK, K_MISC, K_NCLASS
};


// ---- kernel argument blocks (filled by the host, passed by value) ----
struct LineArgs {
  const Block *blocks;
  const Work *work;
  const CaseCoef *coef;
  int in_arr, in_col;     // source field
  int out_arr, out_col;   // destination (deriv)
  int m;                  // drva capture column 1..5, 0 = none
  int acc;                // deriv: 1 = add into destination
  int nlp;                // LDS row stride (lines per tile + 1)
  int in_step, out_step;  // column steps per field: field f = blockIdx.y uses
                          // in_col + f in_step, out_col + f out_step, recv slot f
                          // and drva column m + f
  int copy_arr, copy_col; // filter: also write the result to column copy_col +
                          // f out_step of copy_arr (copy_arr < 0: no copy)
  int direct, n45;        // chunked solves: form the closure sums of a block's
                          // faces onto itself here (closure_sums, periodic jump
                          // n45) instead of reading the exchange's recv slots
};

struct ExchArgs {
  const Block *blocks;
  const Work *work;        // face work items (face_point), d = dir | side << 8
  const CaseCoef *coef;
  int nt, nrt, n45;
  int arr, col0;           // source field: column col0 + f fstep + (1-nrt)*dir,
  int fstep;               // for field f = blockIdx.y (recv slot f)
};

struct StageArgs {
  const Block *blocks;
  const Work *work;
  const CaseCoef *coef;
  double *sc;        // [ncase][S_STRIDE]
  double *part;      // reduction partials [nwork][2]
  double *wpart;     // wall-force partials [nwallwork][4]
  int nk, m, d, opt; // stage number, field, direction, option (k_flux_all: write tv everywhere;
                     // k_wall, k_avg_apply: form the primitives of the points they change)
  int arr, col;      // generic field selector
  int o0, o1, o2;    // tv slots for gradient combination (-1..-3 = ss columns)
  int qin;           // 1: the stage's state is in qo, not qa, until its update (the
                     // first stage of a whole time step: the filter sweep leaves the
                     // filtered state in qo only, and the update writes qa from qo)
};


// launchers (tp_line.hip, tp_col.hip, tp_stage.hip)
void launch_inflow_table(const double *modes, const double *rot, int nmodes, const double *pos, int ng, int ntz,
                         double fctr, const double *uoo, double *vito, hipStream_t st);
void launch_eddy_table(const double *eddies, int nits, const double *pos, int ng, int ntz, double fctr,
                       const double *uoo, double slit, double span, double cfit, double *vito, hipStream_t st);
// chunked line solves: P waves per 64 lines, chunk capacity C points per lane
int col_capacity(int chunk);
size_t col_lds_bytes(int d, int P);
// e0, e1 (optional): timing events carried in the launch's own dispatch
// (hipExtLaunchKernel), recorded as the kernel starts and ends, instead of
// event packets of their own between the launches
void launch_col(int d, int kind, int C, int P, int nwg, int nf, hipStream_t st, const LineArgs &a,
                hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_deriv(int d, int nwg, int nf, size_t lds, hipStream_t st, const LineArgs &a,
                  hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_filter(int d, int nwg, int nf, size_t lds, hipStream_t st, const LineArgs &a,
                   hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// a launch with the timing events e0, e1 in its dispatch when either is set
#define TP_LAUNCH_EV(K, G, B, LDS, ST, E0, E1, ...)                                        \
  do {                                                                                  \
    if ((E0) || (E1)) hipExtLaunchKernelGGL(K, G, B, (uint32_t)(LDS), ST, E0, E1, 0, __VA_ARGS__); \
    else hipLaunchKernelGGL(K, G, B, LDS, ST, __VA_ARGS__);                             \
  } while (0)
void launch_exchange(int nwg, int nf, hipStream_t st, const ExchArgs &a);
void launch_point(int which, int nwg, hipStream_t st, const StageArgs &a);
void launch_face(int which, int nwg, hipStream_t st, const StageArgs &a);
void launch_sc_stage(const CaseCoef *coef, double *sc, int ncase, int nk, hipStream_t st);
void launch_sc_dt(const CaseCoef *coef, double *sc, const double *part, const int *wr, int ncase,
                  int reduced, hipStream_t st);
// next: also the next stage's start scalars (inside a whole time step)
void launch_sc_rk(const CaseCoef *coef, double *sc, int ncase, int nk, int next, hipStream_t st);
void launch_sc_end(double *sc, int ncase, const int *active, hipStream_t st);
void launch_junction(const Block *blocks, const CaseCoef *coef, const int *jct, const int *jvel, int njct, int maxk,
                     int prim, hipStream_t st);
void launch_surf_reduce(const CaseCoef *coef, double *sc, const double *wpart, const int *wwr, int ncase,
                        hipStream_t st);

// host: record the message of a failure (tp_last_error) and return its code
int set_error(int code, const std::string &msg);

enum PointKernel { PK_PRIM = 0, PK_DTRED, PK_FORCE = 5, PK_SPONGE = 6, PK_RK, PK_RK_PRIM, PK_COPYF = 9, PK_INIT, PK_METRICS,
                   PK_FLUXALL = 14, PK_STRESS = 15, PK_FLUX_TV = 16 };
enum FaceKernel { FK_GCBC_PREP = 0, FK_COPY, FK_GCBC_APPLY, FK_WALL, FK_AVG_CAP, FK_AVG_APPLY, FK_NORMALS,
                  FK_SURF_SETUP, FK_SURF_SAMPLE, FK_GCBC_APPLY_XI };

}  // namespace tp

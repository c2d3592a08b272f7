// tp_stage.hip -- pointwise, boundary, face-exchange and scalar kernels of one
// Runge-Kutta stage.
//
// Reference: src/Main3D.f90:571-952 (stage body), src/Subroutines3D.f90:235-324
// (characteristic transforms), src/SingleTurbulence3D.f90:332-443 (sponge,
// junction), src/Subroutineso.f90:192-259 (moving frame).  Every statement keeps
// the reference's evaluation order; the library is compiled without FP
// contraction.
//
// All per-case scalars (n, dt, time, frame velocity) live on the device and are
// advanced by single-thread kernels, so a whole step is enqueued without a host
// synchronisation and can be captured in a HIP graph.
#include "tp_internal.hpp"

namespace tp {

constexpr double GAM = 1.4;
constexpr double PI_D = 3.14159265358979323846;

struct Phys {
  double gamm1, ham, hamm1, gamm1prndtli, tgamm1prndtli, two_three;
};
__device__ inline Phys phys() {
  Phys p;
  p.gamm1 = GAM - 1.0;
  p.ham = 1.0 / GAM;
  p.hamm1 = 1.0 / p.gamm1;
  p.gamm1prndtli = 1.0 / (p.gamm1 * 0.71);
  p.tgamm1prndtli = 1.0 / (p.gamm1 * 0.99);
  p.two_three = 2.0 / 3.0;
  return p;
}



#define WORK_POINTS                                                            \
  const Work w = a.work[blockIdx.x];                                           \
  const Block &b = a.blocks[w.blk];                                            \
  const int N = b.N;                                                           \
  (void)N;

// ---- sin and cos rounded correctly, for the per-stage scalars (moving
// frame, time-step ramp, body force): the reference leaves their last bit to
// its math library, and a one-unit difference in the frame acceleration
// grows over a run.  x - k pi/2 with pi/2 split into three doubles (Cody and
// Waite), in double-double arithmetic (two-sum, fma two-product); Taylor
// series of sin and cos to the 1/29! term by Horner's rule in double-double;
// the normalised sum's leading double is the result.  |x| >= 2^20: libm. ----
struct DD {
  double h, l;
};
__device__ inline DD dd_fast(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ inline DD dd_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ inline DD dd_add(DD a, DD b) {
  const DD s = dd_sum(a.h, b.h);
  return dd_fast(s.h, s.l + (a.l + b.l));
}
__device__ inline DD dd_mul(DD a, DD b) {
  const double p = a.h * b.h, e = fma(a.h, b.h, -p);
  return dd_fast(p, e + (a.h * b.l + a.l * b.h));
}
// 1/n! as double-double, n = 0..29
__constant__ double kInvFact[30][2] = {
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
__device__ inline DD inv_fact(int n, bool neg) {
  return neg ? DD{-kInvFact[n][0], -kInvFact[n][1]} : DD{kInvFact[n][0], kInvFact[n][1]};
}
__device__ void cr_sincos(double x, double *sn, double *cs) {
  if (!(fabs(x) < 0x1p20)) { *sn = sin(x); *cs = cos(x); return; }
  if (x == 0.0) { *sn = x; *cs = 1.0; return; }
  const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  const double k = rint(x * 0.6366197723675814);
  const double p = k * P1, pe = fma(k, P1, -p);
  DD r = dd_sum(x, -p);
  r = dd_add(r, DD{-pe, 0.0});
  const double q = k * P2, qe = fma(k, P2, -q);
  r = dd_add(r, DD{-q, -qe});
  r = dd_add(r, DD{-(k * P3), 0.0});
  const DD z = dd_mul(r, r);
  DD ps = inv_fact(29, true), pc = inv_fact(28, false);
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
__device__ inline double cr_sin(double x) {
  double s, c;
  cr_sincos(x, &s, &c);
  return s;
}
__device__ inline double cr_cos(double x) {
  double s, c;
  cr_sincos(x, &s, &c);
  return c;
}

// ---- moving frame, Subroutineso.f90:192-259 ----
__device__ void movef_dev(const CaseCoef &c, double *sc, double dtko, double dtk) {
  double *umf = sc + S_UMF1, *dudt = sc + S_DUDT1;
  const double timo = sc[S_TIMO];
  if (c.nsmf == 0) {
    // each angle's sine and cosine from one evaluation (cr_sin and cr_cos of
    // an angle are the two results of cr_sincos)
    double ra0 = PI_D / c.timf;
    double ra1 = ra0 * fmin(timo, c.timf), ra2 = ra0 * fmin(timo + dtko, c.timf);
    double s1, c1, s2, c2;
    cr_sincos(ra1, &s1, &c1);
    cr_sincos(ra2, &s2, &c2);
    double fctr = 1.0 - c1, dfdt = ra0 * s2;
    double progmf = 0.5 * (fctr + dtk * dfdt);
    for (int d = 0; d < 3; d++) umf[d] = progmf * c.uoo[d];
    fctr = s1;
    dfdt = ra0 * c2;
    progmf = 0.5 * ra0 * (fctr + dtk * dfdt);
    for (int d = 0; d < 3; d++) dudt[d] = progmf * c.uoo[d];
    if (c.tp > 0 && timo + dtk > c.tps) {
      double daoa = (c.aoa1 - c.aoa0) * PI_D / 180, raoa0 = c.aoa0 * PI_D / 180;
      double ra3 = timo - c.tps, dtsdt = PI_D / c.tp;
      double ts = fmin(dtsdt * ra3, PI_D), tsdt = fmin(dtsdt * (ra3 + dtko), PI_D);
      double sts, cts, stsdt, ctsdt;
      cr_sincos(ts, &sts, &cts);
      cr_sincos(tsdt, &stsdt, &ctsdt);
      double aoat = raoa0 + 0.5 * daoa * (1 - cts);
      double aoatdt = raoa0 + 0.5 * daoa * (1 - ctsdt);
      double sa, ca, sad, cad;
      cr_sincos(aoat, &sa, &ca);
      cr_sincos(aoatdt, &sad, &cad);
      double f = 0.5 * daoa * dtsdt, am = c.amachoo;
      double ut0 = am * ca, ut1 = am * sa;
      double du0 = -(am * f * sad * stsdt), du1 = am * f * cad * stsdt;
      umf[0] = ut0 + dtk * du0; umf[1] = ut1 + dtk * du1; umf[2] = 0.0 + dtk * 0.0;
      double dd0 = -(am * f * sa * sts), dd1 = am * f * ca * sts;
      double st = stsdt;
      double d20 = -(am * f * (cad * f * (st * st) + dtsdt * sad * ctsdt));
      double d21 = -(am * f * (sad * f * (st * st) - dtsdt * cad * ctsdt));
      dudt[0] = dd0 + dtk * d20; dudt[1] = dd1 + dtk * d21; dudt[2] = 0.0 + dtk * 0.0;
      sc[S_AOA] = atan(umf[1] / umf[0]) * 180 / PI_D;
    }
  } else {
    for (int d = 0; d < 3; d++) { umf[d] = c.uoo[d]; dudt[d] = 0.0; }
  }
  sc[S_DTK] = dtk;
  sc[S_DTKO] = dtko;
}

// stage start: Main3D.f90:571-572
// ---- time interpolation of the inflow table, SingleTurbulence3D.f90:336-356 ----
// Five-point Lagrange interpolation in the periodic table time tt(i) = i*fctr,
// i = 0..ntz, at ra3 = (timo - tgusto + dtk) mod period; nodes next to the
// table ends wrap around by one period.  The weights and table rows are the
// same for every inflow point of a case: stored in its scalar slots.
__device__ void inflow_weights(const CaseCoef &c, double *s) {
  s[S_LAGON] = 0.0;
  const double timo = s[S_TIMO];
  if (!(timo > c.tgusto)) return;
  const double ra0 = timo - c.tgusto + s[S_DTK], ra1 = c.period, ra2 = ra0 / ra1;
  const double ra3 = ra0 - ra1 * (double)(int)ra2;
  const int is = 0, ie = c.ntz;
  int ii = 0;
  double best = fabs(0 * c.fctr - ra3);
  for (int i = 1; i <= ie; i++) {
    const double d = fabs(i * c.fctr - ra3);
    if (d < best) { best = d; ii = i; }
  }
  int il[5];
  double tl[5];
  for (int jj = -2; jj <= 2; jj++) {
    il[jj + 2] = min(max(ii + jj, is), ie);
    tl[jj + 2] = il[jj + 2] * c.fctr;
  }
  if (ii - is == 0) { il[0] = ie - 2; il[1] = ie - 1; tl[0] = il[0] * c.fctr - ra1; tl[1] = il[1] * c.fctr - ra1; }
  if (ii - is == 1) { il[0] = ie - 1; tl[0] = il[0] * c.fctr - ra1; }
  if (ie - ii == 1) { il[4] = is + 1; tl[4] = il[4] * c.fctr + ra1; }
  if (ie - ii == 0) { il[3] = is + 1; il[4] = is + 2; tl[3] = il[3] * c.fctr + ra1; tl[4] = il[4] * c.fctr + ra1; }
  double al[5];
  for (int q = 0; q < 5; q++) al[q] = ra3 - tl[q];
  for (int jj = 0; jj < 5; jj++) {
    double ao = 1.0, bo = 1.0;
    for (int q = 0; q < 5; q++) {
      if (q != jj) {
        ao = ao * al[q];
        bo = bo * (tl[jj] - tl[q]);
      }
    }
    s[S_LAGW + jj] = ao / bo;
    s[S_LAGI + jj] = (double)il[jj];
  }
  s[S_LAGON] = 1.0;
}

__global__ void k_sc_stage(const CaseCoef *coef, double *sc, int ncase, int nk) {
  int cs = blockIdx.x * blockDim.x + threadIdx.x;
  if (cs >= ncase) return;
  const CaseCoef &c = coef[cs];
  double *s = sc + cs * S_STRIDE;
  int nkrk = c.nkrk;
  double dt = s[S_DT];
  int f1 = nk - 2 > 0 ? (nk - 2 < 1 ? nk - 2 : 1) : 0;
  int f2 = nk - 1 < 1 ? nk - 1 : 1;
  double dtko = dt * (double)f1 / (nkrk - nk + 3);
  double dtk = dt * (double)f2 / (nkrk - nk + 2);
  movef_dev(c, s, dtko, dtk);
  if (c.inflow) inflow_weights(c, s);
}

// time step size: Main3D.f90:587-610 (maxima reduced from the partials of the
// blocks of each case; wr[cs], wr[cs+1] bound its work items)
__global__ void k_sc_dt(const CaseCoef *coef, double *sc, const double *part, const int *wr,
                        int ncase, int reduced) {
  int cs = blockIdx.x;
  if (cs >= ncase) return;
  __shared__ double mi[256], mv[256];
  double fi = -INFINITY, fv = -INFINITY;
  double *s = sc + cs * S_STRIDE;
  int n = (int)s[S_N];
  bool need = reduced && (n % 10 == 1) && coef[cs].dto < 0.0;
  if (need) {
    for (int q = wr[cs] + threadIdx.x; q < wr[cs + 1]; q += blockDim.x) {
      fi = fmax(fi, part[2 * q]);
      fv = fmax(fv, part[2 * q + 1]);
    }
  }
  mi[threadIdx.x] = fi; mv[threadIdx.x] = fv;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      mi[threadIdx.x] = fmax(mi[threadIdx.x], mi[threadIdx.x + h]);
      mv[threadIdx.x] = fmax(mv[threadIdx.x], mv[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const CaseCoef &c = coef[cs];
    if (n % 10 == 1) {
      s[S_NDT] = n;
      s[S_DTS] = s[S_DTE];
      if (c.dto < 0.0) {
        double ra0 = c.cfl / mi[0], ra1 = ra0;
        if (c.nviscous == 1) ra1 = 0.5 / mv[0];
        s[S_DTE] = fmin(ra0, ra1);
      } else {
        s[S_DTE] = c.dto;
      }
    }
    double sn = cr_sin(0.05 * PI_D * (double)(n - (int)s[S_NDT]));
    s[S_DT] = s[S_DTS] + (s[S_DTE] - s[S_DTS]) * (sn * sn);
  }
}

// before the RK update: Main3D.f90:873-874.  Its frame arguments (dtko, dtk)
// are those of the next stage's start (k_sc_stage(nk + 1): the same time,
// the same step), so with next set (inside a whole time step) it also forms
// the next stage's inflow weights and that stage needs no k_sc_stage; this
// stage's sponge, their only reader, has run.
__global__ void k_sc_rk(const CaseCoef *coef, double *sc, int ncase, int nk, int next) {
  int cs = blockIdx.x * blockDim.x + threadIdx.x;
  if (cs >= ncase) return;
  const CaseCoef &c = coef[cs];
  double *s = sc + cs * S_STRIDE;
  double dt = s[S_DT];
  int f2 = nk - 1 < 1 ? nk - 1 : 1;
  double dtko = dt * (double)f2 / (c.nkrk - nk + 2);
  double dtk = dt / (c.nkrk - nk + 1);
  movef_dev(c, s, dtko, dtk);
  if (next && c.inflow) inflow_weights(c, s);
}

// end of step: Main3D.f90:964-965
__global__ void k_sc_end(double *sc, int ncase, const int *active) {
  int cs = blockIdx.x * blockDim.x + threadIdx.x;
  if (cs >= ncase || (active && !active[cs])) return;
  double *s = sc + cs * S_STRIDE;
  s[S_N] = s[S_N] + 1;
  s[S_TIMO] = s[S_TIMO] + s[S_DT];
}

// ---- primitives, Main3D.f90:576-583 ----
// 1/rho, u, v, w, gamma p/rho, p and the viscosity of the conserved state q
struct Prim {
  double d1, d2, d3, d4, d5, p, mu;
};
// x^1.5 of the viscosity law (Main3D.f90:582), rounded correctly, so that the
// result does not depend on a math library's last bit: y = x sqrt(x) is
// within two units in the last place; the residual x^3 - y^2 comes from
// error-free fma products (its leading difference is exact, the two terms
// being within a factor two), and one Newton step y + (x^3 - y^2) / 2y gives
// the nearest double.  A perfect square x has an exact sqrt, and x sqrt(x)
// is then rounded once already.  Outside [2^-300, 2^300]: pow's value.
__device__ inline double pow15(double x) {
  if (!(x > 0x1p-300 && x < 0x1p300)) return pow(x, 1.5);
  const double s = sqrt(x), y = x * s;
  if (fma(s, s, -x) == 0.0) return y;
  const double P = y * y, p = fma(y, y, -P);
  const double Q = x * x, q = fma(x, x, -Q);
  const double S = x * Q, e = fma(x, Q, -S);
  const double r = (S - P) + ((e - p) + x * q);
  return y + r / (y + y);
}

__device__ inline Prim primitives(const CaseCoef &c, const Phys &ph, double q1, double q2, double q3, double q4,
                                  double q5) {
  Prim r;
  r.d1 = 1.0 / q1;
  r.d2 = q2 * r.d1; r.d3 = q3 * r.d1; r.d4 = q4 * r.d1;
  r.p = ph.gamm1 * (q5 - 0.5 * (q2 * r.d2 + q3 * r.d3 + q4 * r.d4));
  r.d5 = GAM * r.p * r.d1;
  r.mu = c.srefp1dre * pow15(r.d5) / (r.d5 + c.srefoo);
  return r;
}

// the conserved state a stage reads before its update (qo on the first stage
// of a whole time step, whose filter sweep leaves it there only)
__device__ inline const double *stage_q(const StageArgs &a, const Block &b) { return a.qin ? b.qo : b.qa; }

// the primitives of the point's conserved state q that a later stage of the
// same time step reads: u, v, w, T into de(:,2:5) and p (plain stores: the
// face kernels' points).  1/rho (de(:,1)) and the viscosity (ss(:,1)) are
// read only by the time-step limit, which follows the primitives pass of a
// step's first stage; the flux pass forms the viscosity again from q.
__device__ inline void put_primitives(const Block &b, const CaseCoef &c, const Phys &ph, int l) {
  const int N = b.N;
  const Prim r = primitives(c, ph, b.qa[l], b.qa[N + l], b.qa[2 * N + l], b.qa[3 * N + l], b.qa[4 * N + l]);
  b.de[N + l] = r.d2; b.de[2 * N + l] = r.d3; b.de[3 * N + l] = r.d4; b.de[4 * N + l] = r.d5;
  b.p[l] = r.p;
}

// opt (inside a whole time step): 1/rho (de(:,1)) and the viscosity
// (ss(:,1)) only on the steps whose time-step limit reads them (the test of
// k_dt_reduce); nothing else in the step reads them
__global__ void __launch_bounds__(256) k_primitives(StageArgs a) {
  WORK_POINTS
  const CaseCoef &c = a.coef[b.cs];
  const Phys ph = phys();
  const bool all = !a.opt || (int)a.sc[b.cs * S_STRIDE + S_N] % 10 == 1;
  const double *qa = stage_q(a, b);
  for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) {
    const Prim r = primitives(c, ph, ld_stream(&qa[l]), ld_stream(&qa[N + l]), ld_stream(&qa[2 * N + l]),
                              ld_stream(&qa[3 * N + l]), ld_stream(&qa[4 * N + l]));
    st_stream(&b.de[N + l], r.d2);
    st_stream(&b.de[2 * N + l], r.d3);
    st_stream(&b.de[3 * N + l], r.d4);
    st_stream(&b.de[4 * N + l], r.d5);
    st_stream(&b.p[l], r.p);
    if (all) {
      st_stream(&b.de[l], r.d1);
      st_stream(&b.ss[l], r.mu);
    }
  }
}

// ---- time-step maxima, Main3D.f90:590-603 (partials per workgroup) ----
// The limit is formed on every tenth step (n = 1, 11, ...); k_sc_dt reads
// the partials only then, so on the other steps a workgroup returns at once
// (uniformly: a work item lies in one block of one case).
__global__ void __launch_bounds__(256) k_dt_reduce(StageArgs a) {
  WORK_POINTS
  const double *sc = a.sc + b.cs * S_STRIDE;
  if ((int)sc[S_N] % 10 != 1) return;
  __shared__ double mi[256], mv[256];
  double fi = -INFINITY, fv = -INFINITY;
  {
    const double u0 = sc[S_UMF1], u1 = sc[S_UMF2], u2 = sc[S_UMF3];
    const double *xi = b.xim, *et = b.etm, *ze = b.zem, *de = b.de;
    for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) {
      double x1 = xi[l], x2 = xi[N + l], x3 = xi[2 * N + l];
      double e1 = et[l], e2 = et[N + l], e3 = et[2 * N + l];
      double z1 = ze[l], z2 = ze[N + l], z3 = ze[2 * N + l];
      double r1 = x1 * x1 + x2 * x2 + x3 * x3 + e1 * e1 + e2 * e2 + e3 * e3 + z1 * z1 + z2 * z2 + z3 * z3;
      double u = de[N + l] + u0, v = de[2 * N + l] + u1, ww = de[3 * N + l] + u2;
      double r2 = fabs(x1 * u + x2 * v + x3 * ww) + fabs(e1 * u + e2 * v + e3 * ww) + fabs(z1 * u + z2 * v + z3 * ww);
      double s2 = fabs(b.yaco[l]);
      fi = fmax(fi, (sqrt(de[4 * N + l] * r1) + r2) * s2);
      fv = fmax(fv, de[l] * b.ss[l] * r1 * s2 * s2);
    }
  }
  mi[threadIdx.x] = fi; mv[threadIdx.x] = fv;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      mi[threadIdx.x] = fmax(mi[threadIdx.x], mi[threadIdx.x + h]);
      mv[threadIdx.x] = fmax(mv[threadIdx.x], mv[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { a.part[2 * blockIdx.x] = mi[0]; a.part[2 * blockIdx.x + 1] = mv[0]; }
}

// ---- viscous stresses and heat fluxes (Main3D.f90:627-692), contravariant
// velocities and the metric-weighted Euler + viscous fluxes of all five
// fields (Main3D.f90:710-766), one pass ----
// On entry fx(:,3f+d) holds d(field f)/d(index d) for f = u, v, w, T
// (de(:,2:5)) and d = xi, eta, zeta.  Each gradient is combined with the
// metric cofactors exactly as the reference does into its work arrays
// (Main3D.f90:631-651: u -> txx, hzz, tzx; v -> txy, tyy, hxx; w -> hyy, tyz,
// tzz; T -> ss(:,1:3)) and the stresses and heat fluxes follow in the
// reference's operation order (Main3D.f90:653-692); the viscosity is ss(:,1)
// (the reference copies it to de(:,1) first, :627, because its T gradient
// overwrites ss).  The fluxes then overwrite the gradients of the same point:
// fx(:,3(m-1)+d) = flux of field m along direction d, with contravariant
// U_d = m_d . (u + umf) (:710-715), per field as :718-766.
// The reference keeps the stresses in module arrays (tv here) and passes
// de(:,1) and de(:,5) through as temporaries that the flux derivatives
// overwrite; here the stresses stay in registers.  tv is written where it is
// read: at the wall rows of blocks with a wall surface on the first stage
// (the surface forces), everywhere when a.opt is set.  u, v, w, p and the
// viscosity are formed again from q with k_primitives' operations (equal to
// what it left in de, p and ss; q is read here anyway).
__device__ inline void flux_put(double *fx, int N, int l, int m, double r1, double r2, double r3) {
  st_stream(&fx[(size_t)(3 * (m - 1)) * N + l], r1);
  st_stream(&fx[(size_t)(3 * (m - 1) + 1) * N + l], r2);
  st_stream(&fx[(size_t)(3 * (m - 1) + 2) * N + l], r3);
}

// MODE: FUSED forms stresses and fluxes from the gradients in one pass (the
// stage's path); STRESS forms only the stresses and heat fluxes and writes
// them to tv at every point (the viscous drop-in, Main3D.f90:626-693);
// FROM_TV forms the fluxes from the stresses in tv (the flux drop-in,
// :708-768).  STRESS followed by FROM_TV gives the fluxes of FUSED bit for bit.
enum FluxMode { FM_FUSED = 0, FM_STRESS = 1, FM_FROM_TV = 2 };

template <int MODE>
__global__ void __launch_bounds__(256) k_flux_all(StageArgs a) {
  WORK_POINTS
  const CaseCoef &c = a.coef[b.cs];
  const Phys ph = phys();
  const double *sc = a.sc + b.cs * S_STRIDE;
  const double u0 = sc[S_UMF1], u1 = sc[S_UMF2], u2 = sc[S_UMF3];
  const int visc = c.nviscous == 1;
  // rows of this block whose stresses the surface forces read
  const int wall_row = a.nk == 1 && b.nsurf > 0 ? (b.sside ? b.ny - 1 : 0) : -1;
  const double *qa = stage_q(a, b), *xi = b.xim, *et = b.etm, *ze = b.zem;
  const double *ya = b.yaco;
  double *fx = b.fx, *tv = b.tv;
  for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) {
    const double x1 = ld_stream(&xi[l]), x2 = ld_stream(&xi[N + l]), x3 = ld_stream(&xi[2 * N + l]);
    const double e1 = ld_stream(&et[l]), e2 = ld_stream(&et[N + l]), e3 = ld_stream(&et[2 * N + l]);
    const double z1 = ld_stream(&ze[l]), z2 = ld_stream(&ze[N + l]), z3 = ld_stream(&ze[2 * N + l]);
    const double qv[5] = {ld_stream(&qa[l]), ld_stream(&qa[N + l]), ld_stream(&qa[2 * N + l]),
                          ld_stream(&qa[3 * N + l]), ld_stream(&qa[4 * N + l])};
    // the 12 gradients loaded with the rest, before the primitives are
    // formed, so that all of the point's loads are in flight at once
    double gr[4][3];
    double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (visc && MODE != FM_FROM_TV) {
#pragma unroll
      for (int f = 0; f < 4; f++)
#pragma unroll
        for (int m = 0; m < 3; m++) gr[f][m] = ld_stream(&fx[(size_t)(3 * f + m) * N + l]);
    }
    if (visc && MODE == FM_FROM_TV) {
#pragma unroll
      for (int k = 0; k < 9; k++) t[k] = ld_stream(&tv[(size_t)k * N + l]);
    }
    const Prim pr = primitives(c, ph, qv[0], qv[1], qv[2], qv[3], qv[4]);
    const double u = pr.d2, v = pr.d3, ww = pr.d4;
    if (visc && MODE != FM_FROM_TV) {
      double cb[4][3];
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const double g1 = gr[f][0], g2 = gr[f][1], g3 = gr[f][2];
        cb[f][0] = x1 * g1 + e1 * g2 + z1 * g3;
        cb[f][1] = x2 * g1 + e2 * g2 + z2 * g3;
        cb[f][2] = x3 * g1 + e3 * g2 + z3 * g3;
      }
      double txx = cb[0][0], hzz = cb[0][1], tzx = cb[0][2];
      double txy = cb[1][0], tyy = cb[1][1], hxx = cb[1][2];
      double hyy = cb[2][0], tyz = cb[2][1], tzz = cb[2][2];
      const double sx = cb[3][0], sy = cb[3][1], sz = cb[3][2];
      double r1 = pr.mu * ya[l];
      double r2 = ph.gamm1prndtli * r1, d5;
      if (c.les == 1) {
        double d1 = txx * txx + tyy * tyy + tzz * tzz + (hzz + txy) * (hzz + txy) + (hyy + tzx) * (hyy + tzx) +
                    (hxx + tyz) * (hxx + tyz);
        // the single-precision work array (-1/yaco)^1.5, formed at set-up with
        // the host's pow: yaco does not change, and a one-ulp difference of the
        // device pow() could flip its rounding to single precision
        double varr = (double)b.varr[l];
        double r3 = qv[0] * (c.smago1 * c.smago1) * varr * sqrt(2 * d1);
        r1 = r1 + r3 * ya[l];
        r2 = r2 + ph.tgamm1prndtli * r3 * ya[l];
        r3 = ph.two_three * (qv[0] * c.smago2 * varr * d1);
        d5 = ph.two_three * (txx + tyy + tzz);
        double ntxx = r1 * (2 * txx - d5) - r3, ntyy = r1 * (2 * tyy - d5) - r3, ntzz = r1 * (2 * tzz - d5) - r3;
        double ntxy = r1 * (txy + hzz), ntyz = r1 * (tyz + hxx), ntzx = r1 * (tzx + hyy);
        txx = ntxx; tyy = ntyy; tzz = ntzz; txy = ntxy; tyz = ntyz; tzx = ntzx;
      } else {
        d5 = ph.two_three * (txx + tyy + tzz);
        double ntxx = r1 * (2 * txx - d5), ntyy = r1 * (2 * tyy - d5), ntzz = r1 * (2 * tzz - d5);
        double ntxy = r1 * (txy + hzz), ntyz = r1 * (tyz + hxx), ntzx = r1 * (tzx + hyy);
        txx = ntxx; tyy = ntyy; tzz = ntzz; txy = ntxy; tyz = ntyz; tzx = ntzx;
      }
      hxx = r2 * sx + u * txx + v * txy + ww * tzx;
      hyy = r2 * sy + u * txy + v * tyy + ww * tyz;
      hzz = r2 * sz + u * tzx + v * tyz + ww * tzz;
      t[0] = txx; t[1] = tyy; t[2] = tzz; t[3] = txy; t[4] = tyz; t[5] = tzx; t[6] = hxx; t[7] = hyy; t[8] = hzz;
      if (MODE == FM_STRESS || a.opt || (wall_row >= 0 && (l / b.nx) % b.ny == wall_row)) {
#pragma unroll
        for (int k = 0; k < 9; k++) tv[(size_t)k * N + l] = t[k];
      }
    }
    if (MODE == FM_STRESS) continue;
    const double q1 = u + u0, q2 = v + u1, q3 = ww + u2;
    const double s1 = x1 * q1 + x2 * q2 + x3 * q3;
    const double s2 = e1 * q1 + e2 * q2 + e3 * q3;
    const double s3 = z1 * q1 + z2 * q2 + z3 * q3;
    const double pl = pr.p;
    {
      const double q = qv[0];
      flux_put(fx, N, l, 1, q * s1, q * s2, q * s3);
    }
#pragma unroll
    for (int m = 2; m <= 4; m++) {
      const double q = qv[m - 1];
      const double mx = m == 2 ? x1 : m == 3 ? x2 : x3;
      const double me = m == 2 ? e1 : m == 3 ? e2 : e3;
      const double mz = m == 2 ? z1 : m == 3 ? z2 : z3;
      double r1 = q * s1 + mx * pl, r2 = q * s2 + me * pl, r3 = q * s3 + mz * pl;
      if (visc) {
        // (txx, txy, tzx), (txy, tyy, tyz), (tzx, tyz, tzz): slots 0..5 = txx tyy tzz txy tyz tzx
        const double ta = m == 2 ? t[0] : m == 3 ? t[3] : t[5];
        const double tb = m == 2 ? t[3] : m == 3 ? t[1] : t[4];
        const double tc = m == 2 ? t[5] : m == 3 ? t[4] : t[2];
        r1 = r1 - x1 * ta - x2 * tb - x3 * tc;
        r2 = r2 - e1 * ta - e2 * tb - e3 * tc;
        r3 = r3 - z1 * ta - z2 * tb - z3 * tc;
      }
      flux_put(fx, N, l, m, r1, r2, r3);
    }
    {
      const double d5 = qv[4] + pl;
      double r1 = d5 * s1 - pl * (u0 * x1 + u1 * x2 + u2 * x3);
      double r2 = d5 * s2 - pl * (u0 * e1 + u1 * e2 + u2 * e3);
      double r3 = d5 * s3 - pl * (u0 * z1 + u1 * z2 + u2 * z3);
      if (visc) {
        r1 = r1 - x1 * t[6] - x2 * t[7] - x3 * t[8];
        r2 = r2 - e1 * t[6] - e2 * t[7] - e3 * t[8];
        r3 = r3 - z1 * t[6] - z2 * t[7] - z3 * t[8];
      }
      flux_put(fx, N, l, 5, r1, r2, r3);
    }
  }
}

// ---- sponge, SingleTurbulence3D.f90:360-372 (eddy inflow switched off in the
// reference: its target is the uniform state rho = 1, rho u = 0) ----
__global__ void __launch_bounds__(256) k_sponge(StageArgs a) {
  WORK_POINTS
  const Phys ph = phys();
  const CaseCoef &c = a.coef[b.cs];
  const double *sc = a.sc + b.cs * S_STRIDE;
  const bool gust = c.inflow && sc[S_LAGON] != 0.0;
  const int ng = b.ngust, nt1 = c.ntz + 1;
  for (int q = w.a + threadIdx.x; q < w.a + w.n; q += blockDim.x) {
    const int l = b.lcsz[q];
    // rest-state target: rho = 1, no momentum, gamma p = rho^gamma = 1
    // (pow(1, gamma) is 1 exactly, so pt = ham * 1 is the reference's value
    // without the pow, which the sponge points' lanes would all evaluate)
    double rt = 1.0, m0 = 0.0, m1 = 0.0, m2 = 0.0, pt = ph.ham;
    const int ll = c.inflow ? b.gix[q] : -1;
    if (ll >= 0) {
      // target of an inflow point: isentropic density and momentum of the
      // interpolated velocity (SingleTurbulence3D.f90:361-366)
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
      if (gust) {
        for (int jj = 0; jj < 5; jj++) {
          const double wj = sc[S_LAGW + jj];
          const int it = (int)sc[S_LAGI + jj];
          v0 = v0 + wj * b.vito[((size_t)0 * nt1 + it) * ng + ll];
          v1 = v1 + wj * b.vito[((size_t)1 * nt1 + it) * ng + ll];
          v2 = v2 + wj * b.vito[((size_t)2 * nt1 + it) * ng + ll];
        }
      }
      const double fctr = 0.5 * ph.gamm1;
      rt = pow(1.0 - fctr * (v0 * v0 + v1 * v1 + v2 * v2), ph.hamm1);
      const double at = b.atz[ll] * rt;
      m0 = at * v0; m1 = at * v1; m2 = at * v2;
      pt = ph.ham * pow(rt, GAM);
    }
    const double as = b.asz[q];
    const double *qs = stage_q(a, b);
    b.de[l] = b.de[l] + as * (qs[l] - rt);
    b.de[N + l] = b.de[N + l] + as * (qs[N + l] - m0);
    b.de[2 * N + l] = b.de[2 * N + l] + as * (qs[2 * N + l] - m1);
    b.de[3 * N + l] = b.de[3 * N + l] + as * (qs[3 * N + l] - m2);
    b.de[4 * N + l] = b.de[4 * N + l] + b.bsz[q] * (b.p[l] - pt);
  }
}

// ---- body force, forcego (src/rpt.f90:758-769): between tsfor and tefor,
// three spanwise modes with temporal cosines added to the x-momentum
// residual and taken from the y-momentum residual at the forcing points ----
__global__ void __launch_bounds__(256) k_force(StageArgs a) {
  WORK_POINTS
  const CaseCoef &c = a.coef[b.cs];
  const double *sc = a.sc + b.cs * S_STRIDE;
  const double t = sc[S_TIMO] + sc[S_DTK];
  if (!(c.forcing == 1 && t > c.tsfor && t < c.tefor)) return;
  const double ra0 = t - c.tsfor;
  const double ra1 = c.amfor * cr_cos(ra0 * 48.76 * c.amachoo) / 3;
  const double ra2 = c.amfor * cr_cos(ra0 * 53.6 * c.amachoo) / 3;
  const double ra3 = c.amfor * cr_cos(ra0 * 53.6 * c.amachoo) / 3;
  const int nf = b.nforce;
  for (int q = w.a + threadIdx.x; q < w.a + w.n; q += blockDim.x) {
    const int l = b.lcfor[q];
    const double f = ra1 * b.xafor[q] + ra2 * b.xafor[nf + q] + ra3 * b.xafor[2 * nf + q];
    const double rho = stage_q(a, b)[l];
    b.de[N + l] = b.de[N + l] + rho * f;
    b.de[2 * N + l] = b.de[2 * N + l] - rho * f;
  }
}

// ---- inflow velocity table from the reference's synthetic eddies
// (src/SingleTurbulence3D.f90:268-291): each eddy nn and its images one slit
// upstream and one span to either side, convected with uoo to the table
// time tt(ii), induce rr x d at the point (d its offset from the centre,
// rr_m = ait_m (ra1 r^2) exp(ve_m r^4) (1 + ra2 ve_m r^4)); the sum, in the
// reference's order, is scaled by cfit.  One thread per (point, time). ----
__global__ void __launch_bounds__(256) k_eddy_table(const double *eddies, int nits, const double *pos, int ng,
                                                    int ntz, double fctr, double u0, double u1, double u2, double slit,
                                                    double span, double cfit, double *vito) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int nt1 = ntz + 1;
  if (e >= ng * nt1) return;
  const int g = e % ng, ii = e / ng;
  const double px = pos[g], py = pos[ng + g], pz = pos[2 * (size_t)ng + g];
  const double tt = ii * fctr;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  for (int nn = 0; nn < nits; nn++) {
    const double *E = eddies + 10 * (size_t)nn;
    const double mxc = E[9];
    const double ve0 = (-9 + mxc) * E[3] * E[3], ve1 = (-9 + mxc) * E[4] * E[4], ve2 = (-9 + mxc) * E[5] * E[5];
    const double ra1 = -36 - 60 * mxc, ra2 = (2.0 / 3.0) * mxc;
    for (int i = 0; i <= 1; i++)
      for (int k = -1; k <= 1; k++) {
        const double rv0 = (E[0] - i * slit) + tt * u0, rv1 = E[1] + tt * u1, rv2 = (E[2] - k * span) + tt * u2;
        const double d1 = px - rv0, d2 = py - rv1, d3 = pz - rv2;
        const double d4 = d1 * d1 + d2 * d2 + d3 * d3;
        double d5 = d4 * d4;
        double r1 = ve0 * d5, r2 = ve1 * d5, r3 = ve2 * d5;
        d5 = ra1 * d4;
        r1 = E[6] * d5 * exp(r1) * (1 + ra2 * r1);
        r2 = E[7] * d5 * exp(r2) * (1 + ra2 * r2);
        r3 = E[8] * d5 * exp(r3) * (1 + ra2 * r3);
        v0 = v0 + r2 * d3 - r3 * d2;
        v1 = v1 + r3 * d1 - r1 * d3;
        v2 = v2 + r1 * d2 - r2 * d1;
      }
  }
  vito[((size_t)0 * nt1 + ii) * ng + g] = cfit * v0;
  vito[((size_t)1 * nt1 + ii) * ng + g] = cfit * v1;
  vito[((size_t)2 * nt1 + ii) * ng + g] = cfit * v2;
}

void launch_eddy_table(const double *eddies, int nits, const double *pos, int ng, int ntz, double fctr,
                       const double *uoo, double slit, double span, double cfit, double *vito, hipStream_t st) {
  const int tot = ng * (ntz + 1);
  hipLaunchKernelGGL(k_eddy_table, dim3((tot + 255) / 256), dim3(256), 0, st, eddies, nits, pos, ng, ntz, fctr,
                     uoo[0], uoo[1], uoo[2], slit, span, cfit, vito);
}

// ---- inflow velocity table: frozen turbulence convected with uoo, sampled
// at the inflow points at the table times tt(i) = i*fctr (the reference fills
// vito(ll, i, :) the same way from its eddies, SingleTurbulence3D.f90:262-292) ----
// the table at IT_BLK consecutive times per thread: mode n's phase at time
// index it is a_n - it b_n (a_n = k.x + phi, b_n = fctr k.u), so one sincos
// per mode gives the first time and a rotation by -b_n (cos b_n, sin b_n in
// rot) each next one
constexpr int IT_BLK = 8;
__global__ void __launch_bounds__(256) k_inflow_table(const double *modes, const double *rot, int nmodes,
                                                      const double *pos, int ng, int ntz, double fctr, double u0,
                                                      double u1, double u2, double *vito) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int nt1 = ntz + 1, nblk = (nt1 + IT_BLK - 1) / IT_BLK;
  if (e >= ng * nblk) return;
  const int ll = e % ng, it0 = (e / ng) * IT_BLK;
  const double px = pos[ll], py = pos[ng + ll], pz = pos[2 * (size_t)ng + ll];
  double s0[IT_BLK], s1[IT_BLK], s2[IT_BLK];
#pragma unroll
  for (int k = 0; k < IT_BLK; k++) s0[k] = s1[k] = s2[k] = 0.0;
  for (int n = 0; n < nmodes; n++) {
    const double *m = modes + 8 * (size_t)n;
    const double a = m[0] * px + m[1] * py + m[2] * pz + m[6];
    const double b = fctr * (m[0] * u0 + m[1] * u1 + m[2] * u2);
    const double cb = rot[2 * n], sb = rot[2 * n + 1];
    double sn, cs;
    sincos(a - it0 * b, &sn, &cs);
#pragma unroll
    for (int k = 0; k < IT_BLK; k++) {
      s0[k] = s0[k] + m[3] * cs;
      s1[k] = s1[k] + m[4] * cs;
      s2[k] = s2[k] + m[5] * cs;
      const double c2 = cs * cb + sn * sb;
      sn = sn * cb - cs * sb;
      cs = c2;
    }
  }
#pragma unroll
  for (int k = 0; k < IT_BLK; k++) {
    const int it = it0 + k;
    if (it > ntz) break;
    vito[((size_t)0 * nt1 + it) * ng + ll] = s0[k];
    vito[((size_t)1 * nt1 + it) * ng + ll] = s1[k];
    vito[((size_t)2 * nt1 + it) * ng + ll] = s2[k];
  }
}

void launch_inflow_table(const double *modes, const double *rot, int nmodes, const double *pos, int ng, int ntz,
                         double fctr, const double *uoo, double *vito, hipStream_t st) {
  const int tot = ng * ((ntz + IT_BLK) / IT_BLK);
  hipLaunchKernelGGL(k_inflow_table, dim3((tot + 255) / 256), dim3(256), 0, st, modes, rot, nmodes, pos, ng, ntz,
                     fctr, uoo[0], uoo[1], uoo[2], vito);
}

// ---- characteristic algebra, Subroutines3D.f90:235-324 ----
struct Elem { double rhoi, ao, aoi, ve[3], hv2, vn, vs; };

__device__ inline void eleme(const double *q, double p, const double *cm, const double *umf, Elem &e) {
  e.rhoi = 1.0 / q[0];
  e.ao = sqrt(GAM * e.rhoi * p);
  e.aoi = 1.0 / e.ao;
  e.ve[0] = e.rhoi * q[1]; e.ve[1] = e.rhoi * q[2]; e.ve[2] = e.rhoi * q[3];
  e.hv2 = 0.5 * (e.ve[0] * e.ve[0] + e.ve[1] * e.ve[1] + e.ve[2] * e.ve[2]);
  e.vn = cm[0] * e.ve[0] + cm[1] * e.ve[1] + cm[2] * e.ve[2];
  e.vs = cm[0] * umf[0] + cm[1] * umf[1] + cm[2] * umf[2];
}

// X_L . v (Q -> R) without forming the matrix
__device__ inline void xtq2r_apply(const Elem &e, const double *cm, const double *v, double *o) {
  const double gamm1 = GAM - 1.0;
  double ho = gamm1 * e.aoi * e.aoi, bo = 1 - ho * e.hv2, co = e.aoi * e.vn;
  double dm0 = e.aoi * cm[0], dm1 = e.aoi * cm[1], dm2 = e.aoi * cm[2];
  double rv0 = ho * e.ve[0], rv1 = ho * e.ve[1], rv2 = ho * e.ve[2];
  const double *ve = e.ve;
  double xt[5][5];
  xt[0][0] = bo * cm[0] + dm1 * ve[2] - dm2 * ve[1];
  xt[0][1] = cm[0] * rv0; xt[0][2] = cm[0] * rv1 + dm2; xt[0][3] = cm[0] * rv2 - dm1; xt[0][4] = -ho * cm[0];
  xt[1][0] = bo * cm[1] + dm2 * ve[0] - dm0 * ve[2];
  xt[1][1] = cm[1] * rv0 - dm2; xt[1][2] = cm[1] * rv1; xt[1][3] = cm[1] * rv2 + dm0; xt[1][4] = -ho * cm[1];
  xt[2][0] = bo * cm[2] + dm0 * ve[1] - dm1 * ve[0];
  xt[2][1] = cm[2] * rv0 + dm1; xt[2][2] = cm[2] * rv1 - dm0; xt[2][3] = cm[2] * rv2; xt[2][4] = -ho * cm[2];
  xt[3][0] = 1.0 - bo - co; xt[3][1] = dm0 - rv0; xt[3][2] = dm1 - rv1; xt[3][3] = dm2 - rv2; xt[3][4] = ho;
  xt[4][0] = 1.0 - bo + co; xt[4][1] = -dm0 - rv0; xt[4][2] = -dm1 - rv1; xt[4][3] = -dm2 - rv2; xt[4][4] = ho;
  for (int r = 0; r < 5; r++) {
    double s = 0;
    for (int k = 0; k < 5; k++) s = s + xt[r][k] * v[k];
    o[r] = s;
  }
}

__device__ inline void xtr2q_apply(const Elem &e, const double *cm, const double *v, double *o) {
  const double hamm1 = 1.0 / (GAM - 1.0);
  double bo = e.hv2 + hamm1 * e.ao * e.ao, co = e.ao * e.vn;
  double dm0 = e.ao * cm[0], dm1 = e.ao * cm[1], dm2 = e.ao * cm[2];
  const double *ve = e.ve;
  double xt[5][5];
  xt[0][0] = cm[0]; xt[0][1] = cm[1]; xt[0][2] = cm[2]; xt[0][3] = 0.5; xt[0][4] = 0.5;
  xt[1][0] = cm[0] * ve[0]; xt[1][1] = cm[1] * ve[0] - dm2; xt[1][2] = cm[2] * ve[0] + dm1;
  xt[1][3] = 0.5 * (ve[0] + dm0); xt[1][4] = xt[1][3] - dm0;
  xt[2][0] = cm[0] * ve[1] + dm2; xt[2][1] = cm[1] * ve[1]; xt[2][2] = cm[2] * ve[1] - dm0;
  xt[2][3] = 0.5 * (ve[1] + dm1); xt[2][4] = xt[2][3] - dm1;
  xt[3][0] = cm[0] * ve[2] - dm1; xt[3][1] = cm[1] * ve[2] + dm0; xt[3][2] = cm[2] * ve[2];
  xt[3][3] = 0.5 * (ve[2] + dm2); xt[3][4] = xt[3][3] - dm2;
  xt[4][0] = e.hv2 * cm[0] + dm2 * ve[1] - dm1 * ve[2];
  xt[4][1] = e.hv2 * cm[1] + dm0 * ve[2] - dm2 * ve[0];
  xt[4][2] = e.hv2 * cm[2] + dm1 * ve[0] - dm0 * ve[1];
  xt[4][3] = 0.5 * (bo + co); xt[4][4] = xt[4][3] - co;
  for (int r = 0; r < 5; r++) {
    double s = 0;
    for (int k = 0; k < 5; k++) s = s + xt[r][k] * v[k];
    o[r] = s;
  }
}

__device__ inline bool gcbc_code(int np) { return np == 10 || np == 20 || np == 25 || np == 30; }

__device__ inline void face_jk(const Block &b, int d, int jk, int &j, int &k) {
  int fd = d == 0 ? b.ny : d == 1 ? b.nz : b.nx;
  j = jk % fd; k = jk / fd;
}

// ---- GCBC/GCIC preparation, Main3D.f90:780-795 ----
__global__ void __launch_bounds__(256) k_gcbc_prep(StageArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const int N = b.N, d = w.d & 0xff, ip = w.d >> 8, F = b.F[d];
  int jk;
  if (!face_point(b, w, d, jk)) return;
  const int np = b.nbc[d][ip];
  const double *sc = a.sc + b.cs * S_STRIDE;
  int j, k;
  face_jk(b, d, jk, j, k);
  int l = indx3(b, ip * (b.L[d] - 1), j, k, d);
  double ra0 = (double)((20 - np) * (25 - np) * (30 - np) / 3000), ra1 = 1.0 - ra0;
  double q[5], cm[3], cha[5], v[5], o[5];
  const double *qs = stage_q(a, b);
  for (int c = 0; c < 5; c++) q[c] = qs[(size_t)c * N + l];
  for (int c = 0; c < 3; c++) cm[c] = b.cm[d][(ip * 3 + c) * F + jk];
  Elem e;
  eleme(q, b.p[l], cm, sc + S_UMF1, e);
  double *drva = b.drva[d];
  for (int c = 0; c < 5; c++) cha[c] = ra0 * drva[(ip * 5 + c) * F + jk] + ra1 * b.de[(size_t)c * N + l];
  double ya = b.yaco[l];
  for (int c = 0; c < 5; c++) v[c] = ya * cha[c];
  xtq2r_apply(e, cm, v, o);
  for (int c = 0; c < 5; c++) drva[(ip * 5 + c) * F + jk] = o[c];
  if (d == 0) {
    // the xi faces' state for k_gcbc_apply<true>, one face array each
    for (int c = 0; c < 5; c++) b.gq[(ip * 7 + c) * F + jk] = q[c];
    b.gq[(ip * 7 + 5) * F + jk] = b.p[l];
    b.gq[(ip * 7 + 6) * F + jk] = ya;
  }
}

// ---- face copy from the neighbour: drvb(:,:,ip) <- drva(:,:,1-ip) of ncd
// (GCIC exchange Main3D.f90:799-813 and interface averaging :915-925) ----
__global__ void __launch_bounds__(256) k_face_copy(StageArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const int d = w.d & 0xff, ip = w.d >> 8, F = b.F[d];
  int jk;
  if (!face_point(b, w, d, jk)) return;
  const Block &s = a.blocks[b.ncd[d][ip]];
  for (int c = 0; c < 5; c++) b.drvb[d][(ip * 5 + c) * F + jk] = s.drva[d][((1 - ip) * 5 + c) * F + jk];
}

// number of characteristic faces whose band (face point + MBCI inward) holds
// index (i,j,k): the counter the reference accumulates in rr(:,1), Main3D.f90:460-473
__device__ inline double gcbc_count(const Block &b, int x, int y, int z) {
  int idx[3] = {x, y, z};
  double cnt = 0.0;
  for (int d = 0; d < 3; d++) {
    if (gcbc_code(b.nbc[d][0]) && idx[d] <= MBCI) cnt = cnt + 1.0;
    if (gcbc_code(b.nbc[d][1]) && idx[d] >= b.L[d] - 1 - MBCI) cnt = cnt + 1.0;
  }
  return cnt;
}

__device__ inline void point_ijk(const Block &b, int l, int &x, int &y, int &z) {
  x = l % b.nx; y = (l / b.nx) % b.ny; z = l / (b.nx * b.ny);
}

// ---- GCBC/GCIC application along direction d (both ends), Main3D.f90:817-869 ----
// XI (the xi faces): a face point's band (it and MBCI points inward) is a
// contiguous piece of one row, and the rows of neighbouring face points are a
// row pitch apart.  A thread per band would touch its own cache lines with
// every access, so the workgroup first forms each face point's correction
// (dq, 1/yaco) into LDS, then updates the bands together, consecutive lanes
// on consecutive band points.  The other directions' bands run across
// neighbouring face points' threads, coalesced as they are.
template <bool XI>
__global__ void __launch_bounds__(256) k_gcbc_apply(StageArgs a) {
  __shared__ double sdq[XI ? 256 : 1][6];
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const int N = b.N, d = w.d & 0xff, ip = w.d >> 8, F = b.F[d];
  const int i = ip * (b.L[d] - 1), iq = 1 - 2 * ip;
  int jk;
  const bool on = face_point(b, w, d, jk);
  if (on) {
    const int np = b.nbc[d][ip];
    const double *sc = a.sc + b.cs * S_STRIDE;
    int j, k;
    face_jk(b, d, jk, j, k);
    const double ra0 = iq;
    int l = indx3(b, i, j, k, d);
    double q[5], cm[3], cha[5], dha[5], dv[5], dq[5], pl, yl;
    if (XI) {
      // gathered by k_gcbc_prep: the values at l, from one face array each
      for (int v = 0; v < 5; v++) q[v] = b.gq[(ip * 7 + v) * F + jk];
      pl = b.gq[(ip * 7 + 5) * F + jk];
      yl = b.gq[(ip * 7 + 6) * F + jk];
    } else {
      const double *qs = stage_q(a, b);
      for (int v = 0; v < 5; v++) q[v] = qs[(size_t)v * N + l];
      pl = b.p[l];
      yl = b.yaco[l];
    }
    for (int v = 0; v < 3; v++) cm[v] = b.cm[d][(ip * 3 + v) * F + jk];
    Elem e;
    eleme(q, pl, cm, sc + S_UMF1, e);
    const double *drva = b.drva[d], *drvb = b.drvb[d];
    for (int v = 0; v < 5; v++) { cha[v] = drva[(ip * 5 + v) * F + jk]; dha[v] = drvb[(ip * 5 + v) * F + jk]; }
    double vv = e.vn + e.vs;
    if (np == 10) {
      if (ra0 * (vv + e.ao) > 0.0) cha[3] = 0.0;
      if (ra0 * (vv - e.ao) > 0.0) cha[4] = 0.0;
    } else if (np == 20 || np == 25) {
      double rp = b.rpex[b.wall_off[d][ip] + jk];
      if (rp == 0.0) {
        const double *dd = sc + S_DUDT1;
        double s = cm[0] * dd[0] + cm[1] * dd[1] + cm[2] * dd[2];
        cha[3 + ip] = cha[4 - ip] + 2.0 * ra0 * e.aoi * q[0] * (s + 50.0 * vv);
      } else {
        if (ra0 * vv > 0.0) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }
        if (ra0 * (vv + e.ao) > 0.0) cha[3] = dha[3];
        if (ra0 * (vv - e.ao) > 0.0) cha[4] = dha[4];
      }
    } else {
      if (ra0 * vv > 0.0) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }
      if (ra0 * (vv + e.ao) > 0.0) cha[3] = dha[3];
      if (ra0 * (vv - e.ao) > 0.0) cha[4] = dha[4];
    }
    for (int v = 0; v < 5; v++) dv[v] = cha[v] - drva[(ip * 5 + v) * F + jk];
    xtr2q_apply(e, cm, dv, dq);
    const double res = 1.0 / yl;
    if (XI) {
      for (int v = 0; v < 5; v++) sdq[threadIdx.x][v] = dq[v];
      sdq[threadIdx.x][5] = res;
    } else {
      // the band's de values are all read before any is written back, so
      // the loads are issued together (no point of the band is another
      // thread's)
      double band[MBCI + 1][5];
      int lb[MBCI + 1];
#pragma unroll
      for (int ii = 0; ii <= MBCI; ii++) {
        lb[ii] = indx3(b, i + iq * ii, j, k, d);
#pragma unroll
        for (int v = 0; v < 5; v++) band[ii][v] = b.de[(size_t)v * N + lb[ii]];
      }
#pragma unroll
      for (int ii = 0; ii <= MBCI; ii++) {
        int x, y, z;
        point_ijk(b, lb[ii], x, y, z);
        double wgt = ii == 0 ? res : res * c.sbci[ii - 1];
        double sbcc = wgt / gcbc_count(b, x, y, z);
#pragma unroll
        for (int v = 0; v < 5; v++) b.de[(size_t)v * N + lb[ii]] = band[ii][v] + sbcc * dq[v];
      }
    }
  }
  if (!XI) return;
  __syncthreads();
  // xi faces: element e = (v, face point t, band point ii), ii fastest; the
  // work item's face points are jk = w.a + t, t < w.n (at most 256)
  constexpr int NB = MBCI + 1;
  for (int e = threadIdx.x; e < 5 * NB * w.n; e += blockDim.x) {
    const int ii = e % NB, r = e / NB, t = r % w.n, v = r / w.n;
    int j, k;
    face_jk(b, 0, w.a + t, j, k);
    const int x = i + iq * ii;
    const int l = indx3(b, x, j, k, 0);
    const double res = sdq[t][5];
    const double wgt = ii == 0 ? res : res * c.sbci[ii - 1];
    const double sbcc = wgt / gcbc_count(b, x, j, k);
    b.de[(size_t)v * N + l] = b.de[(size_t)v * N + l] + sbcc * sdq[t][v];
  }
}

// ---- RK update, Main3D.f90:876-881 ----
// PRIM: also the next stage's primitives of the updated state (k_primitives'
// arithmetic on the same values) that the stage reads, u, v, w, T into
// de(:,2:5) and p, once the residual is read; the wall, junction and
// averaging passes that follow form them again at the points they change
template <bool PRIM>
__global__ void __launch_bounds__(256) k_rk(StageArgs a) {
  WORK_POINTS
  const double dtk = a.sc[b.cs * S_STRIDE + S_DTK];
  const CaseCoef &c = a.coef[b.cs];
  const Phys ph = phys();
  for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) {
    double r1 = dtk * ld_stream(&b.yaco[l]);
    double q[5];
    for (int v = 0; v < 5; v++) {
      q[v] = ld_stream(&b.qo[(size_t)v * N + l]) - r1 * ld_stream(&b.de[(size_t)v * N + l]);
      st_stream(&b.qa[(size_t)v * N + l], q[v]);
    }
    if (PRIM) {
      const Prim r = primitives(c, ph, q[0], q[1], q[2], q[3], q[4]);
      st_stream(&b.de[N + l], r.d2);
      st_stream(&b.de[2 * N + l], r.d3);
      st_stream(&b.de[3 * N + l], r.d4);
      st_stream(&b.de[4 * N + l], r.d5);
      st_stream(&b.p[l], r.p);
    }
  }
}

// ---- wall temperature and velocity, Main3D.f90:886-895 ----
__global__ void __launch_bounds__(256) k_wall(StageArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const Phys ph = phys();
  const int N = b.N, d = w.d & 0xff, ip = w.d >> 8;
  int jk;
  if (!face_point(b, w, d, jk)) return;
  const int np = b.nbc[d][ip];
  const double *umf = a.sc + b.cs * S_STRIDE + S_UMF1;
  int j, k;
  face_jk(b, d, jk, j, k);
  int l = indx3(b, ip * (b.L[d] - 1), j, k, d);
  double ra0 = ph.ham * ph.hamm1 * c.wtemp;
  double ra1 = (double)((np - 20) / 5);
  double rp = b.rpex[b.wall_off[d][ip] + jk];
  double ra2 = 1.0 - rp, ra3 = 1.0 - ra1 * ra2;
  double *qa = b.qa;
  double fctr = (1.0 - ra3) * qa[l];
  for (int v = 0; v < 3; v++) qa[(1 + v) * N + l] = ra3 * qa[(1 + v) * N + l] - fctr * umf[v];
  double q1 = qa[l], q2 = qa[N + l], q3 = qa[2 * N + l], q4 = qa[3 * N + l];
  qa[4 * N + l] = rp * qa[4 * N + l] + ra2 * (ra0 * q1 + 0.5 * (q2 * q2 + q3 * q3 + q4 * q4) / q1);
  if (a.opt) put_primitives(b, c, ph, l);   // the next stage's primitives, formed in the update
}

// ---- junction averaging, SingleTurbulence3D.f90:378-443 ----
// jct[j*12 + mbr*3 + {block, ip, jp}], jvel[j*2 + {0,1}]; one thread per (junction, k)
// prim: also the primitives of the new corner states (the next stage's,
// formed in the update)
__global__ void k_junction(const Block *blocks, const CaseCoef *coef, const int *jct, const int *jvel, int njct,
                           int maxk, int prim) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int jn = t / maxk, k = t % maxk;
  if (jn >= njct) return;
  const int *J = jct + jn * 12;
  if (k > blocks[J[0]].nz - 1) return;
  double v[4][5];
  int lq[4];
  for (int q = 0; q < 4; q++) {
    const Block &b = blocks[J[q * 3]];
    int i = J[q * 3 + 1] * (b.nx - 1), j = J[q * 3 + 2] * (b.ny - 1);
    lq[q] = indx3(b, i, j, k, 0);
    for (int m = 0; m < 5; m++) v[q][m] = b.qa[(size_t)m * b.N + lq[q]];
  }
  int va = jvel[jn * 2], vb = jvel[jn * 2 + 1];
  double r[5];
  r[0] = 0.25 * (v[0][0] + v[1][0] + v[2][0] + v[3][0]);
  for (int m = 1; m <= 3; m++) r[m] = 0.5 * (v[va][m] + v[vb][m]);
  r[4] = 0.25 * (v[0][4] + v[1][4] + v[2][4] + v[3][4]);
  for (int q = 0; q < 4; q++) {
    const Block &b = blocks[J[q * 3]];
    for (int m = 0; m < 5; m++) b.qa[(size_t)m * b.N + lq[q]] = r[m];
    if (prim) put_primitives(b, coef[b.cs], phys(), lq[q]);
  }
}

// ---- interface averaging, Main3D.f90:901-952 ----
__device__ inline bool avg_code(int np) { return np == 30 || np == 35 || np == 45; }

// capture face values into drva (first loop of the reference)
__global__ void __launch_bounds__(256) k_avg_capture(StageArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const int N = b.N, d = w.d & 0xff, ip = w.d >> 8, F = b.F[d];
  int jk;
  if (!face_point(b, w, d, jk)) return;
  int j, k;
  face_jk(b, d, jk, j, k);
  int l = indx3(b, ip * (b.L[d] - 1), j, k, d);
  for (int v = 0; v < 5; v++) b.drva[d][(ip * 5 + v) * F + jk] = b.qa[(size_t)v * N + l];
}

// running-average weight of face (d, ip) at (x,y,z): the reference's rr(l,1)
// counter after this face's increment (faces processed in the order
// xi start, xi end, eta start, ..., each adding one (or rpex for GCIC walls))
__device__ inline double avg_count(const Block &b, const CaseCoef &c, int x, int y, int z, int d,
                                   double self_inc) {
  int idx[3] = {x, y, z};
  double cnt = 1.0;
  for (int dd = 0; dd < d; dd++)
    for (int s = 0; s < 2; s++) {
      int np = b.nbc[dd][s];
      bool on = s == 0 ? idx[dd] == 0 : idx[dd] == b.L[dd] - 1;
      if (!on) continue;
      if (avg_code(np)) cnt = cnt + 1.0;
      else if (c.nextgcic == 1 && (np == 20 || np == 25)) {
        int fjk;
        int fj = dd == 0 ? y : dd == 1 ? z : x, fk = dd == 0 ? z : dd == 1 ? x : y;
        int fd = dd == 0 ? b.ny : dd == 1 ? b.nz : b.nx;
        fjk = fk * fd + fj;
        cnt = cnt + b.rpex[b.wall_off[dd][s] + fjk];
      }
    }
  return cnt + self_inc;
}

__global__ void __launch_bounds__(256) k_avg_apply(StageArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const int N = b.N, d = w.d & 0xff, ip = w.d >> 8, F = b.F[d];
  int jk;
  if (!face_point(b, w, d, jk)) return;
  const int np = b.nbc[d][ip];
  int j, k;
  face_jk(b, d, jk, j, k);
  int l = indx3(b, ip * (b.L[d] - 1), j, k, d);
  int x, y, z;
  point_ijk(b, l, x, y, z);
  // a face of the same direction at the other end cannot hold this point
  // (lines have >= 7 points), so only earlier directions contribute
  double *qa = b.qa;
  const double *drvb = b.drvb[d];
  // the point's state: on xi faces (the first direction) the value captured
  // for this face, from one face array each instead of a row per point (no
  // earlier direction has changed it); later directions read qa, which an
  // earlier face of the point may have averaged
  double q[5];
  for (int v = 0; v < 5; v++) q[v] = d == 0 ? b.drva[0][(ip * 5 + v) * F + jk] : qa[(size_t)v * N + l];
  if (avg_code(np)) {
    double r1 = avg_count(b, c, x, y, z, d, 1.0);
    double r2 = 1.0 / r1;
    for (int v = 0; v < 5; v++)
      qa[(size_t)v * N + l] = r2 * ((r1 - 1.0) * q[v] + drvb[(ip * 5 + v) * F + jk]);
  } else {
    double rp = b.rpex[b.wall_off[d][ip] + jk];
    double ra1 = 1.0 - rp;
    double r1 = avg_count(b, c, x, y, z, d, rp);
    double r2 = 1.0 / r1;
    for (int v = 0; v < 5; v++)
      qa[(size_t)v * N + l] = rp * r2 * ((r1 - 1.0) * q[v] + drvb[(ip * 5 + v) * F + jk]) + ra1 * q[v];
  }
  if (a.opt) put_primitives(b, c, phys(), l);   // the next stage's primitives, formed in the update
}

// ---- copies ----
__global__ void __launch_bounds__(256) k_copy_field(StageArgs a) {  // rr(:,1) = field(arr, col)
  WORK_POINTS
  const double *s = field_ptr(b, a.arr, a.col);
  for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) b.rr[l] = s[l];
}

// ---- initial state, SingleTurbulence3D.f90:203 ----
__global__ void __launch_bounds__(256) k_initialo(StageArgs a) {
  WORK_POINTS
  const Phys ph = phys();
  double e = ph.hamm1 * (1.0 / GAM);
  for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) {
    b.qa[l] = 1.0; b.qa[N + l] = 0.0; b.qa[2 * N + l] = 0.0; b.qa[3 * N + l] = 0.0; b.qa[4 * N + l] = e;
  }
}

// ---- metrics from coordinate derivatives, Main3D.f90:378-386, 411-413 ----
// qo/qa/de columns 0..2 hold d(x,y,z)/d(xi,eta,zeta)
__global__ void __launch_bounds__(256) k_metrics(StageArgs a) {
  WORK_POINTS
  for (int l = w.a + threadIdx.x; l < w.a + w.n; l += blockDim.x) {
    double qo1 = b.qo[l], qo2 = b.qo[N + l], qo3 = b.qo[2 * N + l];
    double qa1 = b.qa[l], qa2 = b.qa[N + l], qa3 = b.qa[2 * N + l];
    double de1 = b.de[l], de2 = b.de[N + l], de3 = b.de[2 * N + l];
    double x1 = qa2 * de3 - de2 * qa3, x2 = de2 * qo3 - qo2 * de3, x3 = qo2 * qa3 - qa2 * qo3;
    double e1 = qa3 * de1 - de3 * qa1, e2 = de3 * qo1 - qo3 * de1, e3 = qo3 * qa1 - qa3 * qo1;
    double z1 = qa1 * de2 - de1 * qa2, z2 = de1 * qo2 - qo1 * de2, z3 = qo1 * qa2 - qa1 * qo2;
    b.xim[l] = x1; b.xim[N + l] = x2; b.xim[2 * N + l] = x3;
    b.etm[l] = e1; b.etm[N + l] = e2; b.etm[2 * N + l] = e3;
    b.zem[l] = z1; b.zem[N + l] = z2; b.zem[2 * N + l] = z3;
    b.yaco[l] = 3.0 / (qo1 * x1 + qo2 * e1 + qo3 * z1 + qa1 * x2 + qa2 * e2 + qa3 * z2 + de1 * x3 + de2 * e3 + de3 * z3);
  }
}

// ---- face unit normals, Main3D.f90:429-439 ----
__global__ void __launch_bounds__(256) k_normals(StageArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const int N = b.N, d = w.d & 0xff, ip = w.d >> 8, F = b.F[d];
  int jk;
  if (!face_point(b, w, d, jk)) return;
  int j, k;
  face_jk(b, d, jk, j, k);
  int l = indx3(b, ip * (b.L[d] - 1), j, k, d);
  const double *mt = d == 0 ? b.xim : d == 1 ? b.etm : b.zem;
  double rv0 = b.yaco[l] * mt[l], rv1 = b.yaco[l] * mt[N + l], rv2 = b.yaco[l] * mt[2 * N + l];
  double f = 1.0 / sqrt(rv0 * rv0 + rv1 * rv1 + rv2 * rv2);
  b.cm[d][(ip * 3 + 0) * F + jk] = f * rv0;
  b.cm[d][(ip * 3 + 1) * F + jk] = f * rv1;
  b.cm[d][(ip * 3 + 2) * F + jk] = f * rv2;
}

// ---- host launchers ----
// ---- wall surface, src/rpt.f90:774-864 (wallArea, walldir) ----
// One thread per point ll = k*nx + i of the eta face that is a wall (the
// reference's lwall order, xi fastest).  The cell area comes from the
// coordinate derivatives the metrics set-up leaves in qo, qa, de (x, y, z
// along xi and zeta), halved at the parent block's xi and zeta ends: the
// faces that are not splits of the parent (BC 40).  The unit normal is the
// eta metric, pointing into the fluid.
__global__ void __launch_bounds__(256) k_surf_setup(StageArgs a) {
  const Work w = a.work[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= w.n) return;
  const Block &b = a.blocks[w.blk];
  const int N = b.N, nw = b.nsurf, ll = w.a + t;
  const int i = ll % b.nx, k = ll / b.nx, j = b.sside ? b.ny - 1 : 0;
  const int l = (k * b.ny + j) * b.nx + i;
  const double *qo = b.qo, *qa = b.qa, *de = b.de, *et = b.etm;
  const double g11 = qo[l] * qo[l] + qa[l] * qa[l] + de[l] * de[l];
  const double g33 = qo[2 * N + l] * qo[2 * N + l] + qa[2 * N + l] * qa[2 * N + l] + de[2 * N + l] * de[2 * N + l];
  const double g13 = qo[l] * qo[2 * N + l] + qa[l] * qa[2 * N + l] + de[l] * de[2 * N + l];
  double ar = sqrt(g11 * g33 - g13 * g13);
  if (i == 0 && b.nbc[0][0] != 40) ar = ar * 0.5;
  else if (i == b.nx - 1 && b.nbc[0][1] != 40) ar = ar * 0.5;
  if (k == 0 && b.nbc[2][0] != 40) ar = ar * 0.5;
  else if (k == b.nz - 1 && b.nbc[2][1] != 40) ar = ar * 0.5;
  b.sarea[ll] = ar;
  const double coef = b.sside ? -1.0 : 1.0;
  const double tmp = coef / sqrt(et[l] * et[l] + et[N + l] * et[N + l] + et[2 * N + l] * et[2 * N + l]);
  b.snor[ll] = et[l] * tmp;
  b.snor[nw + ll] = et[N + l] * tmp;
  b.snor[2 * nw + ll] = et[2 * N + l] * tmp;
}

// ---- one sample of the surface statistics at the first stage ----
// When the case samples at this step (n a multiple of its interval): the
// pressure and wall-shear force of each wall cell, p n area and
// rho (tau . n) area as src/rpt.f90:1017-1056 (clpost) forms them, summed per
// workgroup in a fixed order into four partials; and the trapezoidal time
// integrals of p and p^2 at each wall point over the sampled times (the
// running form of the time averages of src/rptPost.f90:870-925).
__global__ void __launch_bounds__(256) k_surf_sample(StageArgs a) {
  const Work w = a.work[blockIdx.x];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const double *s = a.sc + b.cs * S_STRIDE;
  const double every = s[S_ST_EVERY];
  const bool on = every > 0 && fmod(s[S_N], every) == 0.0;
  __shared__ double red[4][256];
  double f0 = 0, f1 = 0, f2 = 0, f3 = 0;
  const int t = threadIdx.x;
  if (on && t < w.n) {
    const int N = b.N, nw = b.nsurf, ll = w.a + t;
    const int i = ll % b.nx, k = ll / b.nx, j = b.sside ? b.ny - 1 : 0;
    const int l = (k * b.ny + j) * b.nx + i;
    const double p = b.p[l], ar = b.sarea[ll];
    const double n1 = b.snor[ll], n2 = b.snor[nw + ll], n3 = b.snor[2 * nw + ll];
    f0 = p * n1 * ar;
    f1 = p * n2 * ar;
    if (c.nviscous == 1) {
      const double *tv = b.tv, rho = stage_q(a, b)[l];
      const double tw1 = rho * (tv[l] * n1 + tv[3 * N + l] * n2 + tv[5 * N + l] * n3);
      const double tw2 = rho * (tv[3 * N + l] * n1 + tv[N + l] * n2 + tv[4 * N + l] * n3);
      f2 = tw1 * ar;
      f3 = tw2 * ar;
    }
    double *ps = b.sst;
    if (s[S_ST_NS] > 0) {
      const double dt = s[S_TIMO] - s[S_ST_TP];
      ps[nw + ll] = ps[nw + ll] + 0.5 * dt * (ps[ll] + p);
      ps[2 * nw + ll] = ps[2 * nw + ll] + 0.5 * dt * (ps[ll] * ps[ll] + p * p);
    }
    ps[ll] = p;
  }
  red[0][t] = f0; red[1][t] = f1; red[2][t] = f2; red[3][t] = f3;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 4; q++) red[q][t] = red[q][t] + red[q][t + h];
    __syncthreads();
  }
  if (t < 4) a.wpart[4 * blockIdx.x + t] = red[t][0];
}

// per case: the force coefficients from the partials of its wall work items,
// lift and drag with the angle of attack (src/Main3D.f90:700-704), and their
// trapezoidal time integrals
__global__ void k_surf_reduce(const CaseCoef *coef, double *sc, const double *wpart, const int *wwr,
                              int ncase) {
  const int cs = blockIdx.x;
  if (cs >= ncase) return;
  double *s = sc + cs * S_STRIDE;
  const double every = s[S_ST_EVERY];
  if (!(every > 0 && fmod(s[S_N], every) == 0.0)) return;
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  double v[4] = {0, 0, 0, 0};
  for (int q = wwr[cs] + t; q < wwr[cs + 1]; q += blockDim.x)
    for (int r = 0; r < 4; r++) v[r] = v[r] + wpart[4 * q + r];
  for (int r = 0; r < 4; r++) red[r][t] = v[r];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int r = 0; r < 4; r++) red[r][t] = red[r][t] + red[r][t + h];
    __syncthreads();
  }
  if (t == 0) {
    const CaseCoef &c = coef[cs];
    const double dynp = 2.0 / (c.amachoo * c.amachoo * c.span);
    const double cl1 = (red[0][0] + red[2][0]) * dynp, cl2 = (red[1][0] + red[3][0]) * dynp;
    const double raoa = s[S_AOA] * PI_D / 180;
    const double clift = cl2 * cos(raoa) - cl1 * sin(raoa);
    const double cdrag = cl2 * sin(raoa) + cl1 * cos(raoa);
    const double tm = s[S_TIMO];
    if (s[S_ST_NS] > 0) {
      const double dt = tm - s[S_ST_TP];
      s[S_CLS] = s[S_CLS] + 0.5 * dt * (s[S_CL] + clift);
      s[S_CDS] = s[S_CDS] + 0.5 * dt * (s[S_CD] + cdrag);
    } else {
      s[S_ST_T0] = tm;
    }
    s[S_CL] = clift;
    s[S_CD] = cdrag;
    s[S_ST_TP] = tm;
    s[S_ST_NS] = s[S_ST_NS] + 1;
  }
}

#define LAUNCH(K, NWG, BS, ST, A) hipLaunchKernelGGL(K, dim3(NWG), dim3(BS), 0, ST, A)
void launch_point(int which, int nwg, hipStream_t st, const StageArgs &a) {
  switch (which) {
    case 0: LAUNCH(k_primitives, nwg, 256, st, a); break;
    case 1: LAUNCH(k_dt_reduce, nwg, 256, st, a); break;
    case 5: LAUNCH(k_force, nwg, 256, st, a); break;
    case 6: LAUNCH(k_sponge, nwg, 256, st, a); break;
    case 7: LAUNCH(k_rk<false>, nwg, 256, st, a); break;
    case 8: LAUNCH(k_rk<true>, nwg, 256, st, a); break;
    case 9: LAUNCH(k_copy_field, nwg, 256, st, a); break;
    case 10: LAUNCH(k_initialo, nwg, 256, st, a); break;
    case 11: LAUNCH(k_metrics, nwg, 256, st, a); break;
    case 14: LAUNCH(k_flux_all<FM_FUSED>, nwg, 256, st, a); break;
    case 15: LAUNCH(k_flux_all<FM_STRESS>, nwg, 256, st, a); break;
    case 16: LAUNCH(k_flux_all<FM_FROM_TV>, nwg, 256, st, a); break;
  }
}
void launch_face(int which, int nwg, hipStream_t st, const StageArgs &a) {
  switch (which) {
    case 0: LAUNCH(k_gcbc_prep, nwg, 256, st, a); break;
    case 1: LAUNCH(k_face_copy, nwg, 256, st, a); break;
    case 2: LAUNCH(k_gcbc_apply<false>, nwg, 256, st, a); break;
    case 3: LAUNCH(k_wall, nwg, 256, st, a); break;
    case 4: LAUNCH(k_avg_capture, nwg, 256, st, a); break;
    case 5: LAUNCH(k_avg_apply, nwg, 256, st, a); break;
    case 6: LAUNCH(k_normals, nwg, 256, st, a); break;
    case 7: LAUNCH(k_surf_setup, nwg, 256, st, a); break;
    case 8: LAUNCH(k_surf_sample, nwg, 256, st, a); break;
    case 9: LAUNCH(k_gcbc_apply<true>, nwg, 256, st, a); break;
  }
}
void launch_surf_reduce(const CaseCoef *coef, double *sc, const double *wpart, const int *wwr, int ncase,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_surf_reduce, dim3(ncase), dim3(256), 0, st, coef, sc, wpart, wwr, ncase);
}
void launch_sc_stage(const CaseCoef *coef, double *sc, int ncase, int nk, hipStream_t st) {
  hipLaunchKernelGGL(k_sc_stage, dim3((ncase + 63) / 64), dim3(64), 0, st, coef, sc, ncase, nk);
}
void launch_sc_dt(const CaseCoef *coef, double *sc, const double *part, const int *wr, int ncase, int reduced,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_sc_dt, dim3(ncase), dim3(256), 0, st, coef, sc, part, wr, ncase, reduced);
}
void launch_sc_rk(const CaseCoef *coef, double *sc, int ncase, int nk, int next, hipStream_t st) {
  hipLaunchKernelGGL(k_sc_rk, dim3((ncase + 63) / 64), dim3(64), 0, st, coef, sc, ncase, nk, next);
}
void launch_sc_end(double *sc, int ncase, const int *active, hipStream_t st) {
  hipLaunchKernelGGL(k_sc_end, dim3((ncase + 63) / 64), dim3(64), 0, st, sc, ncase, active);
}
void launch_junction(const Block *blocks, const CaseCoef *coef, const int *jct, const int *jvel, int njct, int maxk,
                     int prim, hipStream_t st) {
  int tot = njct * maxk;
  hipLaunchKernelGGL(k_junction, dim3((tot + 63) / 64), dim3(64), 0, st, blocks, coef, jct, jvel, njct, maxk, prim);
}

}  // namespace tp

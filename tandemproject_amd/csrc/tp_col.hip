// tp_col.hip -- compact line solves with one lane per line and one wave per
// chunk of the line.
//
// Arithmetic of src/Subroutines3D.f90:84-141 (deriv) and :145-206 (filte):
// pentadiagonal RHS, forward substitution with the factors xl (penta,
// src/Subroutineso.f90), backward substitution with xu, in the reference's
// operation order.
//
// Layout of the work: a workgroup owns 64 adjacent lines (one per lane) of one
// sub-block; its P waves own consecutive chunks of ceil(L/P) points of those
// lines.  For eta and zeta lines the 64 lanes read and write 64 consecutive
// doubles of each row (the lines are adjacent in xi), so no staging is needed;
// xi lines are contiguous rows: their windows come through a small per-wave
// LDS transpose in pieces of 8 points, and their results are gathered in LDS
// and stored by the whole workgroup as contiguous runs of rows.  The factors
// of a row are the same for every lane and are read from LDS as broadcasts.
//
// Each wave keeps its chunk in registers and runs:
//   forward sweep from zero carries -> the pair (y[e-1], y[e]) that is an
//     affine function of the true incoming pair (its 2x2 matrix depends on
//     the factors only: tabulated per block, direction, kind and P);
//   composition of those maps over the preceding chunks -> the true carries;
//   forward sweep from the true carries (the reference's recurrence), and the
//   same three steps backwards.
// The carries come out of the composition rounded differently from the
// sequential sweep (~1 ulp).  Inside a chunk that difference decays by ~0.74
// per point, but the two roundings re-synchronise only by chance, and the
// last rows of a one-sided closure amplify whatever ulp difference is left
// (|l2| ~ 72 in the last row, |u2|, |u3| ~ 6 in the first rows).  So near such
// an end the chunks take their carries in a chain instead: each from the
// sweep its neighbour just finished, over at least COL_DECAY points, which
// in practice brings the end rows back to the sequential sweep's bits.
#include "tp_internal.hpp"

namespace tp {

template <int D>
__device__ inline void col_line(const Block &b, int q, int &base, int &jk) {
  if (D == 0) {            // q = zeta*ny + eta
    base = q * b.nx; jk = q;
  } else if (D == 1) {     // q = zeta*nx + xi, face index xi*nz + zeta
    const int xi = q % b.nx, ze = q / b.nx;
    base = ze * b.ny * b.nx + xi; jk = xi * b.nz + ze;
  } else {                 // q = eta*nx + xi
    base = q; jk = q;
  }
}

// element at a 32-bit byte offset from a wave-uniform base: lets the loads and
// stores use the scalar-base + 32-bit vector-offset addressing form
__device__ inline const double &at(const double *base, unsigned off) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + off);
}
__device__ inline double &at(double *base, unsigned off) {
  return *reinterpret_cast<double *>(reinterpret_cast<char *>(base) + off);
}
// a xi output value, stored with the streaming hint
__device__ inline void st(double *base, unsigned off, double v) { st_stream(&at(base, off), v); }

// strided view of one line in global memory (boundary rows)
struct GLine {
  const double *p;
  int s;
  __device__ double operator()(int i) const { return p[(size_t)i * s]; }
};

// closure sums r0, r1, r2 of one face: read from the exchange's recv slots
// where they are used, or held as values
struct RcvFace {
  const double *rcv;
  int F, jk, ip;
  __device__ double operator()(int q) const { return rcv[(ip * 3 + q) * F + jk]; }
};
struct ValFace {
  double r[3];
  // a select, not an index: q may be known only at run time, and an indexed
  // register array would go to scratch memory
  __device__ double operator()(int q) const { return q == 0 ? r[0] : q == 1 ? r[1] : r[2]; }
};

// boundary rows of the derivative RHS (rows 0..2 and L-3..L-1); fs, fe: the
// closure sums received at the start and end faces
template <class SA, class FA>
__device__ inline double deriv_bnd_row(const Block &b, const CaseCoef &c, int D, int i, int L, const SA &sa,
                                       const FA &fs, const FA &fe) {
  const double aa = 0.6431406736919156, ab = 0.2586011023495066, ac = 0.007140953479797375;
  const int ie = L - 1;
  if (i <= 2) {
    if (b.ndf[D][0] == 0) {
      const double s0 = sa(0), s1 = sa(1), s2 = sa(2), s3 = sa(3), s4 = sa(4), s5 = sa(5), s6 = sa(6);
      double r = 0;
      if (i == 0) {
        r = r + c.abc[0][0] * (s1 - s0); r = r + c.abc[0][1] * (s2 - s0); r = r + c.abc[0][2] * (s3 - s0);
        r = r + c.abc[0][3] * (s4 - s0); r = r + c.abc[0][4] * (s5 - s0); r = r + c.abc[0][5] * (s6 - s0);
      } else if (i == 1) {
        r = r + c.abc[1][0] * (s0 - s1); r = r + c.abc[1][1] * (s2 - s1); r = r + c.abc[1][2] * (s3 - s1);
        r = r + c.abc[1][3] * (s4 - s1); r = r + c.abc[1][4] * (s5 - s1); r = r + c.abc[1][5] * (s6 - s1);
      } else {
        r = r + c.abc[2][0] * (s0 - s2); r = r + c.abc[2][1] * (s1 - s2); r = r + c.abc[2][2] * (s3 - s2);
        r = r + c.abc[2][3] * (s4 - s2); r = r + c.abc[2][4] * (s5 - s2); r = r + c.abc[2][5] * (s6 - s2);
      }
      return r;
    }
    if (i <= 1) {
      double t = 0;
      for (int q = 0; q <= LMD; q++) t = t + c.pbci[0][i][q] * sa(q);
      return t + fs(i);
    }
    return aa * (sa(3) - sa(1)) + ab * (sa(4) - sa(0)) + ac * (sa(5) - fs(2));
  }
  const int row = ie - i;  // 0: last row
  if (b.ndf[D][1] == 0) {
    const double e0 = sa(ie), e1 = sa(ie - 1), e2 = sa(ie - 2), e3 = sa(ie - 3), e4 = sa(ie - 4),
                 e5 = sa(ie - 5), e6 = sa(ie - 6);
    double r = 0;
    if (row == 0) {
      r = r + c.abc[0][0] * (e0 - e1); r = r + c.abc[0][1] * (e0 - e2); r = r + c.abc[0][2] * (e0 - e3);
      r = r + c.abc[0][3] * (e0 - e4); r = r + c.abc[0][4] * (e0 - e5); r = r + c.abc[0][5] * (e0 - e6);
    } else if (row == 1) {
      r = r + c.abc[1][0] * (e1 - e0); r = r + c.abc[1][1] * (e1 - e2); r = r + c.abc[1][2] * (e1 - e3);
      r = r + c.abc[1][3] * (e1 - e4); r = r + c.abc[1][4] * (e1 - e5); r = r + c.abc[1][5] * (e1 - e6);
    } else {
      r = r + c.abc[2][0] * (e2 - e0); r = r + c.abc[2][1] * (e2 - e1); r = r + c.abc[2][2] * (e2 - e3);
      r = r + c.abc[2][3] * (e2 - e4); r = r + c.abc[2][4] * (e2 - e5); r = r + c.abc[2][5] * (e2 - e6);
    }
    return r;
  }
  if (row <= 1) {
    double t = 0;
    for (int q = 0; q <= LMD; q++) t = t + c.pbci[0][row][q] * sa(ie - q);
    return -t - fe(row);
  }
  return aa * (sa(ie - 1) - sa(ie - 3)) + ab * (sa(ie) - sa(ie - 4)) + ac * (fe(2) - sa(ie - 5));
}

// boundary rows of the filter RHS
template <class SA, class FA>
__device__ inline double filter_bnd_row(const Block &b, const CaseCoef &c, int D, int i, int L, const SA &sa,
                                        const FA &fs, const FA &fe) {
  const int ie = L - 1;
  const double fa = c.fa, fb = c.fb, fc = c.fc;
  if (i <= 2) {
    if (b.ndf[D][0] == 0) {
      const double s0 = sa(0), s1 = sa(1), s2 = sa(2), s3 = sa(3), s4 = sa(4), s5 = sa(5);
      double res = 0;
      res = res + c.fex[0] * (s1 - s0); res = res + c.fex[1] * (s2 - s0); res = res + c.fex[2] * (s3 - s0);
      const double rv0 = s0 - res * c.ffx[0], rv1 = s0 - res * c.ffx[1], rv2 = s0 - res * c.ffx[2];
      if (i == 0) {
        const double ra0 = 2.0 * s0;
        return c.fam[0] * (rv0 + s1 - ra0) + c.fbm[0] * (rv1 + s2 - ra0) + c.fcm[0] * (rv2 + s3 - ra0);
      } else if (i == 1) {
        const double ra1 = 2.0 * s1;
        return c.fam[1] * (s0 + s2 - ra1) + c.fbm[1] * (rv0 + s3 - ra1) + c.fcm[1] * (rv1 + s4 - ra1);
      }
      const double ra2 = 2.0 * s2;
      return c.fam[2] * (s1 + s3 - ra2) + c.fbm[2] * (s0 + s4 - ra2) + c.fcm[2] * (rv0 + s5 - ra2);
    }
    if (i <= 1) {
      double t = 0;
      for (int q = 0; q <= LMF; q++) t = t + c.pbci[1][i][q] * sa(q);
      return t + fs(i);
    }
    const double ra2 = 2.0 * sa(2);
    return fa * (sa(1) + sa(3) - ra2) + fb * (sa(0) + sa(4) - ra2) + fc * (fs(2) + sa(5) - ra2);
  }
  const int row = ie - i;
  if (b.ndf[D][1] == 0) {
    const double e0 = sa(ie), e1 = sa(ie - 1), e2 = sa(ie - 2), e3 = sa(ie - 3), e4 = sa(ie - 4), e5 = sa(ie - 5);
    double res = 0;
    res = res + c.fex[0] * (e0 - e1); res = res + c.fex[1] * (e0 - e2); res = res + c.fex[2] * (e0 - e3);
    const double rv0 = e0 + res * c.ffx[0], rv1 = e0 + res * c.ffx[1], rv2 = e0 + res * c.ffx[2];
    if (row == 0) {
      const double ra0 = 2.0 * e0;
      return c.fam[0] * (e1 + rv0 - ra0) + c.fbm[0] * (e2 + rv1 - ra0) + c.fcm[0] * (e3 + rv2 - ra0);
    } else if (row == 1) {
      const double ra1 = 2.0 * e1;
      return c.fam[1] * (e2 + e0 - ra1) + c.fbm[1] * (e3 + rv0 - ra1) + c.fcm[1] * (e4 + rv1 - ra1);
    }
    const double ra2 = 2.0 * e2;
    return c.fam[2] * (e3 + e1 - ra2) + c.fbm[2] * (e4 + e0 - ra2) + c.fcm[2] * (e5 + rv0 - ra2);
  }
  if (row <= 1) {
    double t = 0;
    for (int q = 0; q <= LMF; q++) t = t + c.pbci[1][row][q] * sa(ie - q);
    return t + fe(row);
  }
  const double ra2 = 2.0 * sa(ie - 2);
  return fa * (sa(ie - 3) + sa(ie - 1) - ra2) + fb * (sa(ie - 4) + sa(ie) - ra2) +
         fc * (sa(ie - 5) + fe(2) - ra2);
}


constexpr int COL_TR = 65;   // row pitch (doubles) of the xi transpose buffer
constexpr int COL_DECAY = 48;  // points a carry travels sequentially before an amplifying closure

// dynamic LDS of a k_col launch, in doubles: carry pairs [2][P][64] (double2),
// chunk factors [P][192], and for xi the transposes [P][8][COL_TR]
__host__ __device__ constexpr int col_lds_doubles(int d, int P) {
  return 2 * P * 64 * 2 + P * 192 + (d == 0 ? P * 8 * COL_TR : 0);
}

// a zero the optimiser cannot see through: offsetting the factor pointer by
// it at each sweep keeps the factor loads inside the sweep instead of being
// hoisted out of the chain-round loops and held in registers throughout
__device__ inline int opaque_zero() {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}

// Orders the LDS accesses of one wave around a transpose step.  The LDS
// instructions of a wave execute in issue order, so no memory fence (which
// would also wait for the wave's outstanding global loads) is needed; this
// only keeps the compiler from moving memory operations across.
__device__ inline void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// value of a double held in lane l (wave-uniform result)
__device__ inline double lane_val(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Factors of the chunk's rows, copied to the wave's LDS slice:
// F[k] = l1, F[32+k] = l2, F[64+k] = u1, F[96+k] = u2, F[128+k] = u3 of row
// k.  Every lane reads the same address (an LDS broadcast).
struct ChunkLU {
  const double *F;
};

// Forward substitution over the chunk rows from the carry (ym2, ym1), in
// place when store is set.  Rows k >= n are capacity padding: their factors
// (l1 = -1, l2 = 0) and zero right-hand side make each of them swap the
// carried pair exactly (y = 0 + ym2 - 0), so after them the pair is the one
// of row n-1, swapped once more if their count is odd.  The opaque offset
// keeps each sweep's factor loads inside the sweep.
template <int C>
__device__ inline void fwd_sweep(double (&r)[C], int n, const ChunkLU &f, double &ym2, double &ym1,
                                 bool store) {
  const double *F = f.F + opaque_zero();
#pragma unroll
  for (int k = 0; k < C; k++) {
    const double f1 = F[k];
    const double f2 = F[32 + k];
    const double y = r[k] - f1 * ym2 - f2 * ym1;
    if (store) r[k] = y;
    ym2 = ym1;
    ym1 = y;
  }
  if ((C - n) & 1) {
    const double t = ym2;
    ym2 = ym1;
    ym1 = t;
  }
}

// backward substitution from the carry (xp1, xp2) = (x[e+1], x[e+2]): the
// padding rows k >= n come first, each swapping the pair (u1 = u2 = 0,
// u3 = -1), so the carry enters them pre-swapped when their count is odd
template <int C>
__device__ inline void bwd_sweep(double (&r)[C], int n, const ChunkLU &f, double &xp1, double &xp2,
                                 bool store) {
  const double *F = f.F + opaque_zero();
  if ((C - n) & 1) {
    const double t = xp1;
    xp1 = xp2;
    xp2 = t;
  }
#pragma unroll
  for (int k = C - 1; k >= 0; k--) {
    const double g1 = F[64 + k];
    const double g2 = F[96 + k];
    const double g3 = F[128 + k];
    const double x = g1 * r[k] - g2 * xp1 - g3 * xp2;
    if (store) r[k] = x;
    xp2 = xp1;
    xp1 = x;
  }
}

// DIRECT: the launch forms the closure sums itself (a.direct) instead of
// reading the exchange's recv slots
template <int D, int KIND, int C, bool DIRECT>
__global__ void __launch_bounds__(1024) k_col(LineArgs a) {
  extern __shared__ double lds[];
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const int L = b.L[D];
  const int P = blockDim.x >> 6;
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int Cl = (L + P - 1) / P;
  const int Pe = (L + Cl - 1) / Cl;               // chunks holding points
  const int s0 = p * Cl;
  const int n = p < Pe ? min(Cl, L - s0) : 0;     // points of this wave's chunk
  const int nl = w.n;
  const bool act = lane < nl;
  const int stride = D == 0 ? 1 : D == 1 ? b.nx : b.nx * b.ny;
  int base, jk;
  col_line<D>(b, w.a + (act ? lane : 0), base, jk);
  const int fi = blockIdx.y;   // field of a merged launch
  const double *src = field_ptr(b, a.in_arr, a.in_col + fi * a.in_step);
  double *dst = field_ptr(b, a.out_arr, a.out_col + fi * a.out_step);
  const double *rcv = b.recv[D] + (size_t)fi * 6 * b.F[D];
  const int mcap = a.m > 0 ? a.m + fi : 0;
  double2 *pairA = reinterpret_cast<double2 *>(lds);      // [P][64] carry-map inputs
  double2 *pairB = pairA + P * 64;                        // [P][64] pairs handed along a chain
  double *fl = reinterpret_cast<double *>(pairB + P * 64) + p * 192;   // [P][192] chunk factors
  double *tr = reinterpret_cast<double *>(pairB + P * 64) + P * 192 + p * (8 * COL_TR);
  const double *T = b.col[D][KIND];
  const double mapf = T[lane], mapb = T[64 + lane];   // carry maps of all chunks
  ChunkLU f;
  {
    const double *F = T + 128 + 192 * min(p, Pe - 1);
    fl[lane] = F[lane]; fl[64 + lane] = F[64 + lane]; fl[128 + lane] = F[128 + lane];
    f.F = fl;
    wave_sync();   // the wave's own LDS writes precede its reads in issue order
  }

  // Chains at one-sided closures.  End: the last Rhi chunks take their
  // forward carry from the chunk before them, one after the other.  Start:
  // the first Rlo chunks take their backward carry from the chunk after
  // them, which needs exact forward values there, so chunks 2..Rlo take
  // their forward carry from their predecessor too (chunk 1's carry, the
  // first chunk's zero-carry result, is exact already).
  const int Rc = (COL_DECAY + Cl - 1) / Cl;
  const int Rhi = b.ndf[D][1] == 0 ? min(Rc, Pe - 1) : 0;
  const int Rlo = b.ndf[D][0] == 0 ? min(Rc, Pe - 1) : 0;
  int tf = 0, Tf = 0, tb = 0, Tb = 0;   // chain round of this wave, rounds in all
  {
    int t = 0;
    for (int q = 0; q < Pe; q++) {
      const bool ch = (q >= 2 && q <= Rlo) || (Rhi > 0 && q >= Pe - Rhi);
      t = ch ? t + 1 : 0;
      if (q == p) tf = t;
      Tf = max(Tf, t);
    }
    t = 0;
    for (int q = Pe - 1; q >= 0; q--) {
      t = q < Rlo ? t + 1 : 0;
      if (q == p) tb = t;
      Tb = max(Tb, t);
    }
  }

  // DIRECT: the closure sums of the faces this wave's boundary rows need,
  // formed from the neighbour's field (its lines on the far side of the
  // face: the other end of the same lines for a periodic block, the partner
  // block's lines -- whose work items run beside these -- otherwise), as the
  // exchange forms them.  First, so that only the three sums per face stay
  // live while the window is loaded.
  ValFace fsum[2] = {{{0.0, 0.0, 0.0}}, {{0.0, 0.0, 0.0}}};
  if (DIRECT) {
#pragma unroll
    for (int ip = 0; ip < 2; ip++) {
      const bool need = ip == 0 ? n > 0 && s0 <= 2 : n > 0 && s0 + n - 1 >= L - 3;
      if (!need || !b.ndf[D][ip]) continue;
      const Block &nb = a.blocks[b.ncd[D][ip]];
      closure_sums<KIND>(b, nb, c, field_ptr(nb, a.in_arr, a.in_col + fi * a.in_step), src, D, ip, jk, a.n45,
                         fsum[ip].r[0], fsum[ip].r[1], fsum[ip].r[2]);
    }
  }

  // ---- window rows s0-3 .. s0+n+2, clamped to the line; the rows of the
  // capacity padding past them are not read (zero) ----
  double v[C + 6];
  const int nw = n + 6;   // window rows in use
  if (D == 0) {
    // all global loads of the window first (8 lines x 8 consecutive points per
    // instruction), then the transposes through LDS piece by piece
    constexpr int NPC = (C + 6 + 7) / 8;
    double tmp[NPC][8];
#pragma unroll
    for (int t = 0; t < NPC; t++) {
      const int row = min(max(s0 - 3 + 8 * t + (lane & 7), 0), L - 1);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int rl = min((lane >> 3) + 8 * j, nl - 1);
        tmp[t][j] = 8 * t < nw ? at(src, 8u * (unsigned)((w.a + rl) * b.nx + row)) : 0.0;
      }
    }
#pragma unroll
    for (int t = 0; t < NPC; t++) {
      if (8 * t < nw) {
#pragma unroll
        for (int j = 0; j < 8; j++) tr[(lane & 7) * COL_TR + (lane >> 3) + 8 * j] = tmp[t][j];
        wave_sync();
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
          if (8 * t + rr < C + 6) v[8 * t + rr] = tr[rr * COL_TR + lane];
        wave_sync();
      } else {
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
          if (8 * t + rr < C + 6) v[8 * t + rr] = 0.0;
      }
    }
  } else {
    const unsigned b8 = 8u * (unsigned)base, st8 = 8u * (unsigned)stride;
#pragma unroll
    for (int k = 0; k < C + 6; k++) {
      const int row = min(max(s0 - 3 + k, 0), L - 1);
      v[k] = k < nw ? at(src, b8 + (unsigned)row * st8) : 0.0;
    }
  }

  // ---- right-hand side ----
  // rows 0..2 and L-3..L-1 use the closure formulas, evaluated once, outside
  // the unrolled loop, by the (at most two) chunks that hold them
  // the closure rows at the line ends, with the closure sums formed above
  // (DIRECT) or read from the exchange's recv slots
  double rb0 = 0.0, rb1 = 0.0, rb2 = 0.0, re2 = 0.0, re1 = 0.0, re0 = 0.0;
  {
    const GLine g{src + base, stride};
    auto rows = [&](const auto &fs, const auto &fe) {
      if (n > 0 && s0 <= 2) {
        if (KIND == 0) {
          rb0 = deriv_bnd_row(b, c, D, 0, L, g, fs, fe);
          rb1 = deriv_bnd_row(b, c, D, 1, L, g, fs, fe);
          rb2 = deriv_bnd_row(b, c, D, 2, L, g, fs, fe);
        } else {
          rb0 = filter_bnd_row(b, c, D, 0, L, g, fs, fe);
          rb1 = filter_bnd_row(b, c, D, 1, L, g, fs, fe);
          rb2 = filter_bnd_row(b, c, D, 2, L, g, fs, fe);
        }
      }
      if (n > 0 && s0 + n - 1 >= L - 3) {
        if (KIND == 0) {
          re2 = deriv_bnd_row(b, c, D, L - 3, L, g, fs, fe);
          re1 = deriv_bnd_row(b, c, D, L - 2, L, g, fs, fe);
          re0 = deriv_bnd_row(b, c, D, L - 1, L, g, fs, fe);
        } else {
          re2 = filter_bnd_row(b, c, D, L - 3, L, g, fs, fe);
          re1 = filter_bnd_row(b, c, D, L - 2, L, g, fs, fe);
          re0 = filter_bnd_row(b, c, D, L - 1, L, g, fs, fe);
        }
      }
    };
    if (DIRECT) rows(fsum[0], fsum[1]);
    else rows(RcvFace{rcv, b.F[D], jk, 0}, RcvFace{rcv, b.F[D], jk, 1});
  }
  // derivative: interior stencil on every row, then the closure rows patched
  // in by the (at most two) chunks at each end that hold them; then the
  // padding rows zeroed
  double r[C];
  if (KIND == 1) {
    // filter: the solve yields the filtered field itself, A out = R + A in
    // (A the filter's left-hand side: interior rows (fbe, fal, 1, fal, fbe),
    // closure rows of each end from flhs, mirrored at the far end), rather
    // than the increment A^-1 R added to a second read of the input.  The
    // closure rows' A in is formed first, from the window (its rows found by
    // selects, not by indexing the register array), then every row's
    // R + A in with the window dying row by row.
    if (n > 0 && s0 <= 2) {
      const double *hd = &c.flhs[b.ndf[D][0]][0][0];
      double ah[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3 && k < C; k++) {
        const int i = s0 + k;   // row i sits at k; entries at columns i-2 .. i+2
        if (i <= 2) {
          const double *e = hd + 5 * i;
          const double s = e[0] * v[k + 1] + e[1] * v[k + 2] + v[k + 3] + e[3] * v[k + 4] + e[4] * v[k + 5];
          ah[0] = i == 0 ? s : ah[0]; ah[1] = i == 1 ? s : ah[1]; ah[2] = i == 2 ? s : ah[2];
        }
      }
      rb0 = rb0 + ah[0]; rb1 = rb1 + ah[1]; rb2 = rb2 + ah[2];
    }
    if (n > 0 && s0 + n - 1 >= L - 3) {
      const double *tl = &c.flhs[b.ndf[D][1]][0][0];
      double at0 = 0.0, at1 = 0.0, at2 = 0.0;
#pragma unroll
      for (int k = 0; k < C; k++) {
        const int i = s0 + k;   // mirrored: entry 4-q at column i+q-2
        if (i >= L - 3 && i <= L - 1) {
          const double *e = tl + 5 * (L - 1 - i);
          const double s = e[4] * v[k + 1] + e[3] * v[k + 2] + v[k + 3] + e[1] * v[k + 4] + e[0] * v[k + 5];
          at0 = i == L - 1 ? s : at0; at1 = i == L - 2 ? s : at1; at2 = i == L - 3 ? s : at2;
        }
      }
      re0 = re0 + at0; re1 = re1 + at1; re2 = re2 + at2;
    }
#pragma unroll
    for (int k = 0; k < C; k++) {
      const double res = 2.0 * v[k + 3];
      r[k] = c.fa * (v[k + 2] + v[k + 4] - res) + c.fb * (v[k + 1] + v[k + 5] - res) +
             c.fc * (v[k] + v[k + 6] - res) + (v[k + 3] + c.fal * (v[k + 2] + v[k + 4]) + c.fbe * (v[k + 1] + v[k + 5]));
    }
  } else {
#pragma unroll
    for (int k = 0; k < C; k++) {
      const double aa = 0.6431406736919156, ab = 0.2586011023495066, ac = 0.007140953479797375;
      r[k] = aa * (v[k + 4] - v[k + 2]) + ab * (v[k + 5] - v[k + 1]) + ac * (v[k + 6] - v[k]);
    }
  }
  // the closure rows patched in by the (at most two) chunks at each end that
  // hold them
  {
    if (n > 0 && s0 <= 2) {
#pragma unroll
      for (int k = 0; k < 3 && k < C; k++) {
        const int i = s0 + k;
        r[k] = i == 0 ? rb0 : i == 1 ? rb1 : i == 2 ? rb2 : r[k];
      }
    }
    if (n > 0 && s0 + n - 1 >= L - 3) {
#pragma unroll
      for (int k = 0; k < C; k++) {
        const int i = s0 + k;
        r[k] = i == L - 1 ? re0 : i == L - 2 ? re1 : i == L - 3 ? re2 : r[k];
      }
    }
  }
  if (n < C) {
#pragma unroll
    for (int k = 0; k < C; k++) r[k] = k < n ? r[k] : 0.0;   // padding rows carry a zero right-hand side
  }

  // ---- forward substitution ----
  double ym2 = 0.0, ym1 = 0.0;
  fwd_sweep<C>(r, n, f, ym2, ym1, false);
  pairA[p * 64 + lane] = make_double2(ym2, ym1);
  __syncthreads();
  double c0 = 0.0, c1 = 0.0;
  for (int q = 0; q < min(p, Pe); q++) {
    const double m0 = lane_val(mapf, 4 * q), m1 = lane_val(mapf, 4 * q + 1);
    const double m2 = lane_val(mapf, 4 * q + 2), m3 = lane_val(mapf, 4 * q + 3);
    const double2 e = pairA[q * 64 + lane];
    const double t0 = m0 * c0 + m1 * c1 + e.x;
    const double t1 = m2 * c0 + m3 * c1 + e.y;
    c0 = t0; c1 = t1;
  }
  for (int t = 0; t <= Tf; t++) {
    if (t > 0) __syncthreads();
    if (tf == t) {
      if (t > 0) {
        const double2 e = pairB[(p - 1) * 64 + lane];
        c0 = e.x; c1 = e.y;
      }
      ym2 = c0; ym1 = c1;
      fwd_sweep<C>(r, n, f, ym2, ym1, true);
      pairB[p * 64 + lane] = make_double2(ym2, ym1);
    }
  }
  __syncthreads();   // every wave has read pairA and pairB before they are reused

  // ---- backward substitution ----
  double xp1 = 0.0, xp2 = 0.0;
  bwd_sweep<C>(r, n, f, xp1, xp2, false);
  pairA[p * 64 + lane] = make_double2(xp1, xp2);
  __syncthreads();
  double d0 = 0.0, d1 = 0.0;
  for (int q = Pe - 1; q > p; q--) {
    const double m0 = lane_val(mapb, 4 * q), m1 = lane_val(mapb, 4 * q + 1);
    const double m2 = lane_val(mapb, 4 * q + 2), m3 = lane_val(mapb, 4 * q + 3);
    const double2 e = pairA[q * 64 + lane];
    const double t0 = m0 * d0 + m1 * d1 + e.x;
    const double t1 = m2 * d0 + m3 * d1 + e.y;
    d0 = t0; d1 = t1;
  }
  for (int t = 0; t <= Tb; t++) {
    if (t > 0) __syncthreads();
    if (tb == t) {
      if (t > 0) {
        const double2 e = pairB[(p + 1) * 64 + lane];
        d0 = e.x; d1 = e.y;
      }
      xp1 = d0; xp2 = d1;
      bwd_sweep<C>(r, n, f, xp1, xp2, true);
      pairB[p * 64 + lane] = make_double2(xp1, xp2);
    }
  }

  // ---- boundary derivatives kept for the characteristic conditions ----
  if (KIND == 0 && mcap > 0 && act && n > 0) {
    const int F = b.F[D];
    if (p == 0) b.drva[D][(0 * 5 + mcap - 1) * F + jk] = r[0];
    if (p == Pe - 1) {
      double xe = r[0];
#pragma unroll
      for (int k = 1; k < C; k++) xe = k == n - 1 ? r[k] : xe;
      b.drva[D][(1 * 5 + mcap - 1) * F + jk] = xe;
    }
  }

  // ---- store: the filtered field, or the derivative (which may accumulate) ----
  const bool add = KIND == 0 && a.acc;
  const double *ob = dst;
  double *cpy = KIND == 1 && a.copy_arr >= 0 ? field_ptr(b, a.copy_arr, a.copy_col + fi * a.out_step) : nullptr;
  // xi: the work item's lines are consecutive rows of the block, one
  // contiguous span of nl * L doubles.  Once every wave is done with the
  // carry, factor and transpose areas, the whole LDS holds R of these rows at
  // a time (R the largest power of two that fits, at least 8), and the
  // workgroup stores them as one run of whole cache lines: a wave's own
  // pieces (8 rows x 8 points) would straddle 64-byte write granules.
  const int pitch = L + 1;
  int R = 64;
  while (R >= 8 && R * pitch > col_lds_doubles(0, P)) R >>= 1;
  if (D == 0 && R >= 8) {
    double *tile = lds;
    __syncthreads();
    for (int h = 0; h * R < nl; h++) {
      if (act && lane >= h * R && lane < (h + 1) * R) {
        double *trow = tile + (lane - h * R) * pitch + s0;
#pragma unroll
        for (int k = 0; k < C; k++)
          if (k < n) trow[k] = r[k];
      }
      __syncthreads();
      const int rows = min(R, nl - h * R);
      const int tot = rows * L;
      const unsigned base0 = (unsigned)(w.a + h * R) * (unsigned)b.nx;
      // element idx of the run sits at (row, col) = (idx / L, idx % L) of the
      // tile; each thread steps by blockDim.x = dq L + dr elements, so
      // (row, col) advance without a division, four elements at a time
      const int dq = (int)blockDim.x / L, dr = (int)blockDim.x - dq * L;
      int row = (int)threadIdx.x / L, col = (int)threadIdx.x - row * L;
      for (int i0 = threadIdx.x; i0 < tot; i0 += 4 * (int)blockDim.x) {
        double x4[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          x4[u] = i0 + u * (int)blockDim.x < tot ? tile[row * pitch + col] : 0.0;
          col += dr;
          row += dq;
          if (col >= L) { col -= L; row++; }
        }
        double o4[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int idx = i0 + u * (int)blockDim.x;
          o4[u] = add && idx < tot ? ld_stream(&at(ob, 8u * (base0 + (unsigned)idx))) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int idx = i0 + u * (int)blockDim.x;
          if (idx < tot) {
            const unsigned o = 8u * (base0 + (unsigned)idx);
            const double y = add ? o4[u] + x4[u] : x4[u];
            st(dst, o, y);
            if (KIND == 1 && cpy) st(cpy, o, y);
          }
        }
      }
      __syncthreads();   // the tile is read before the next rows overwrite it
    }
  } else if (D == 0) {
#pragma unroll
    for (int t = 0; t < (C + 7) / 8; t++) {
      if (8 * t >= n) break;   // pieces past the chunk hold no rows of it
#pragma unroll
      for (int rr = 0; rr < 8; rr++)
        if (8 * t + rr < C) tr[rr * COL_TR + lane] = r[8 * t + rr];
      wave_sync();
      const int k = 8 * t + (lane & 7);
      const int row = min(s0 + k, L - 1);
      double xv[8], ov[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int rl = (lane >> 3) + 8 * j;
        xv[j] = tr[(lane & 7) * COL_TR + rl];
        ov[j] = add ? at(ob, 8u * (unsigned)((w.a + min(rl, nl - 1)) * b.nx + row)) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int rl = (lane >> 3) + 8 * j;
        if (k < n && rl < nl) {
          const unsigned o = 8u * (unsigned)((w.a + rl) * b.nx + s0 + k);
          const double y = add ? ov[j] + xv[j] : xv[j];
          st(dst, o, y);
          if (KIND == 1 && cpy) st(cpy, o, y);
        }
      }
      wave_sync();
    }
  } else {
    const unsigned b8 = 8u * (unsigned)base, st8 = 8u * (unsigned)stride;
    double ov[C];
#pragma unroll
    for (int k = 0; k < C; k++)
      ov[k] = add && k < n ? at(ob, b8 + (unsigned)min(s0 + k, L - 1) * st8) : 0.0;
    if (act) {
#pragma unroll
      for (int k = 0; k < C; k++)
        if (k < n) {
          const unsigned o = b8 + (unsigned)(s0 + k) * st8;
          const double y = add ? ov[k] + r[k] : r[k];
          // the default policy: a row's 512 bytes start anywhere in a cache
          // line, and the neighbouring work item writes the rest of its edge
          // lines; the streaming hint evicts them half-written
          at(dst, o) = y;
          if (KIND == 1 && cpy) at(cpy, o) = y;
        }
    }
  }
}

size_t col_lds_bytes(int d, int P) { return (size_t)col_lds_doubles(d, P) * sizeof(double); }

template <int D, int KIND>
static void launch_col_c(int C, int P, int nwg, int nf, hipStream_t st, const LineArgs &a, hipEvent_t e0,
                         hipEvent_t e1) {
  const size_t lds = col_lds_bytes(D, P);
  switch (C) {
#define CASE_C(CV)                                                                                     \
  case CV:                                                                                           \
    if (KIND == 0 && a.direct)                                                                       \
      TP_LAUNCH_EV((k_col<D, KIND, CV, KIND == 0>), dim3(nwg, nf), dim3(64 * P), lds, st, e0, e1, a); \
    else                                                                                             \
      TP_LAUNCH_EV((k_col<D, KIND, CV, false>), dim3(nwg, nf), dim3(64 * P), lds, st, e0, e1, a);    \
    break;
    CASE_C(8) CASE_C(12) CASE_C(16) CASE_C(20) CASE_C(24) CASE_C(28) CASE_C(32)
#undef CASE_C
  }
}

// register capacity of a chunk of `chunk` points: the next multiple of 4 from
// 8 (rows past the chunk are padding the sweeps run over)
int col_capacity(int chunk) {
  if (chunk > 32) return 0;
  return chunk <= 8 ? 8 : (chunk + 3) / 4 * 4;
}

void launch_col(int d, int kind, int C, int P, int nwg, int nf, hipStream_t st, const LineArgs &a, hipEvent_t e0,
                hipEvent_t e1) {
  if (kind == 0) {
    if (d == 0) launch_col_c<0, 0>(C, P, nwg, nf, st, a, e0, e1);
    else if (d == 1) launch_col_c<1, 0>(C, P, nwg, nf, st, a, e0, e1);
    else launch_col_c<2, 0>(C, P, nwg, nf, st, a, e0, e1);
  } else {
    if (d == 0) launch_col_c<0, 1>(C, P, nwg, nf, st, a, e0, e1);
    else if (d == 1) launch_col_c<1, 1>(C, P, nwg, nf, st, a, e0, e1);
    else launch_col_c<2, 1>(C, P, nwg, nf, st, a, e0, e1);
  }
}

}  // namespace tp

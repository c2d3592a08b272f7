// tp_capi.hip -- host side of the MI355X time stepper: context, device memory,
// scheme coefficients, launch tables, stage sequencing, and the C ABI of
// include/tandem_hip.h.
//
// The scheme coefficients are computed here from their definitions
// (penta: banded LU without pivoting of the pentadiagonal LHS; fcint/fcbcm:
// cut-off-wavenumber filter coefficients; sbcco: quasi-disjoint closure rows
// from the inverse of the coupled two-sided system), following
// src/Subroutineso.f90:15-188 and src/Main3D.f90:293-332 so that the device
// consumes the same tables the reference's setup produces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tandem_hip.h"
#include "tp_internal.hpp"

using namespace tp;

// ----------------------------------------------------------------------------
// errors
// ----------------------------------------------------------------------------
static thread_local std::string g_err;
static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      (void)hipGetLastError(); /* reported here: not again by the next call */            \
      return fail(TP_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));               \
    }                                                                                      \
  } while (0)

namespace tp {
// the message and code of a failure in another host file of the library
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace tp

// ----------------------------------------------------------------------------
// scheme coefficients (host, once per case)
// ----------------------------------------------------------------------------
namespace {

constexpr double kPi = 3.14159265358979323846;

// Kim's interior derivative scheme and its one-sided closures, src/MainVar3D.f90:34-68
constexpr double kAlpha = 0.5862704032801503, kBeta = 0.09549533555017055;
const double kBnd[3][5] = {{0, 0, 1, 5.912678614078549, 3.775623951744012},
                           {0, 0.08360703307833438, 1, 2.058102869495757, 0.9704052014790193},
                           {0.03250008295108466, 0.3998040493524358, 1, 0.7719261277615860,
                            0.1626635931256900}};
const double kAbc[3][6] = {{-3.456878182643609, 5.839043358834730, 1.015886726041007,
                            -0.2246526470654333, 0.08564940889936562, -0.01836710059356763},
                           {-0.3177447290722621, -0.02807631929593225, 1.593461635747659,
                            0.2533027046976367, -0.03619652460174756, 0.004080281419108407},
                           {-0.1219006056449124, -0.6301651351188667, 0.6521195063966084,
                            0.3938843551210350, 0.01904944407973912, -0.001027260523947668}};
constexpr double kAA = 0.6431406736919156, kAB = 0.2586011023495066, kAC = 0.007140953479797375;

struct FiltCoef { double alpha, beta, a, b, c; };

// filter coefficients from the cut-off wavenumber fk and fltr = 1/2
FiltCoef filter_coef(double fk, double fr) {
  const double c1 = std::cos(fk), c2 = std::cos(2 * fk), c3 = std::cos(3 * fk);
  const double f = 1 / (30 + 5 * (7 - 16 * fr) * c1 + 2 * (1 + 8 * fr) * c2 - 3 * c3);
  const double h = std::cos(0.5 * fk), h2 = h * h;
  FiltCoef o;
  o.alpha = f * (20 * (2 * fr - 1) - 30 * c1 + 12 * (2 * fr - 1) * c2 - 2 * c3);
  o.beta = 0.5 * f * (2 * (13 - 8 * fr) + (33 - 48 * fr) * c1 + 6 * c2 - c3);
  o.a = 60 * (1 - fr) * (h2 * h2) * f;
  o.b = -2 * o.a / 5;
  o.c = o.a / 15;
  return o;
}

// LHS rows 0..2 of the filter next to a boundary (cut-off ramped to fltkbc)
void filter_boundary(double fk, double fkbc, double rows[3][5], double fam[3], double fbm[3],
                     double fcm[3]) {
  const double step = (fkbc - fk) / 9;
  const double cut[3] = {fk + 9 * step, fk + 4 * step, fk + step};
  for (int r = 0; r < 3; r++) {
    FiltCoef z = filter_coef(cut[r], 0.5);
    if (r == 0) {
      const double s = 1 / (1 + z.alpha / 2 + z.beta / 4);
      const double v[5] = {0, 0, 1, z.alpha * s, z.beta * s};
      std::copy(v, v + 5, rows[0]);
      fam[0] = z.a * s; fbm[0] = z.b * s; fcm[0] = z.c * s;
    } else if (r == 1) {
      const double v[5] = {0, z.alpha + z.beta / 2, 1, z.alpha, z.beta};
      std::copy(v, v + 5, rows[1]);
      fam[1] = z.a; fbm[1] = z.b; fcm[1] = z.c;
    } else {
      const double v[5] = {z.beta, z.alpha, 1, z.alpha, z.beta};
      std::copy(v, v + 5, rows[2]);
      fam[2] = z.a; fbm[2] = z.b; fcm[2] = z.c;
    }
  }
}

// banded LU (no pivoting) of the pentadiagonal matrix whose first three rows
// are `head`, last three rows the mirror of `tail` and interior (beta, alpha,
// 1, alpha, beta); l1/l2 multipliers, u1 = 1/pivot, u2/u3 scaled upper band
void band_lu(int n, const double head[3][5], const double tail[3][5], double al, double be, double *l1,
             double *l2, double *u1, double *u2, double *u3) {
  for (int i = 0; i < n; i++) { l1[i] = l2[i] = 1; u1[i] = u2[i] = u3[i] = 1; }
  auto pivot = [&](int i, double lo2, double lo1, double up1, double up2, bool set_up2) {
    // row i: lo2 at i-2, lo1 at i-1, 1 at i, up1 at i+1, up2 at i+2
    if (i >= 2) {
      l1[i] = lo2 * u1[i - 2];
      l2[i] = (lo1 - u2[i - 2] * l1[i]) * u1[i - 1];
      u1[i] = 1 / (1 - u3[i - 2] * l1[i] - u2[i - 1] * l2[i]);
    } else if (i == 1) {
      l2[i] = lo1 * u1[i - 1];
      u1[i] = 1 / (1 - u2[i - 1] * l2[i]);
    } else {
      u1[i] = 1;
    }
    if (i >= 1) u2[i] = up1 - u3[i - 1] * l2[i];
    else u2[i] = up1;
    if (set_up2) u3[i] = up2;
  };
  pivot(0, 0, 0, head[0][3], head[0][4], true);
  pivot(1, 0, head[1][1], head[1][3], head[1][4], true);
  pivot(2, head[2][0], head[2][1], head[2][3], head[2][4], true);
  for (int i = 3; i <= n - 4; i++) pivot(i, be, al, al, be, true);
  // mirrored tail rows: row n-1-r has entries tail[r][4-c] at column n-1-r+(c-2)
  pivot(n - 3, tail[2][4], tail[2][3], tail[2][1], tail[2][0], true);
  pivot(n - 2, tail[1][4], tail[1][3], tail[1][1], 0, false);
  // last row: only the pivot; the upper entries stay unset as in the reference
  {
    const int i = n - 1;
    l1[i] = tail[0][4] * u1[i - 2];
    l2[i] = (tail[0][3] - u2[i - 2] * l1[i]) * u1[i - 1];
    u1[i] = 1 / (1 - u3[i - 2] * l1[i] - u2[i - 1] * l2[i]);
  }
  for (int i = 0; i < n; i++) { u2[i] = u2[i] * u1[i]; u3[i] = u3[i] * u1[i]; }
}

// Gauss-Jordan inverse with row pivoting (row-major n x n); returns the
// reciprocal of the last pivot
double gj_inverse(int n, const std::vector<double> &a, std::vector<double> &inv) {
  std::vector<double> w(a);
  std::vector<int> perm(n);
  std::vector<double> col(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  double rp = 0;
  for (int i = 0; i < n; i++) {
    int m = i;
    double best = -1;
    for (int r = i; r < n; r++) {
      double v = std::fabs(w[r * n + i]);
      if (v > best) { best = v; m = r; }
    }
    if (m != i) {
      std::swap(perm[m], perm[i]);
      for (int j = 0; j < n; j++) std::swap(w[m * n + j], w[i * n + j]);
    }
    rp = 1 / w[i * n + i];
    for (int r = 0; r < n; r++) col[r] = w[r * n + i];
    for (int j = 0; j < n; j++) {
      const double f = rp * w[i * n + j];
      for (int r = 0; r < n; r++) w[r * n + j] = w[r * n + j] - f * col[r];
      w[i * n + j] = f;
    }
    for (int r = 0; r < n; r++) w[r * n + i] = -rp * col[r];
    w[i * n + i] = rp;
  }
  inv.assign(n * n, 0.0);
  for (int r = 0; r < n; r++)
    for (int j = 0; j < n; j++) inv[r * n + perm[j]] = w[r * n + j];
  return rp;
}

std::vector<double> matmul(int n, const std::vector<double> &a, const std::vector<double> &b) {
  std::vector<double> c(n * n);
  for (int r = 0; r < n; r++)
    for (int j = 0; j < n; j++) {
      double s = 0;
      for (int k = 0; k < n; k++) s = s + a[r * n + k] * b[k * n + j];
      c[r * n + j] = s;
    }
  return c;
}

// quasi-disjoint closure rows for a line split into two halves: build the
// coupled 2(ll+1) system A f' = B f, decouple the two halves in A (R), and
// read the first two rows of the right half of R A^-1 B
void closure_rows(double fk, CaseCoef &c) {
  std::memset(c.pbci, 0, sizeof(c.pbci));
  std::memset(c.pbco, 0, sizeof(c.pbco));
  for (int nt = 0; nt <= 1; nt++) {
    const int ll = nt == 0 ? LMD : LMF, n = 2 * (ll + 1), sgn = 2 * nt - 1;
    std::vector<double> A(n * n, 0.0), B(n * n, 0.0);
    auto at = [&](std::vector<double> &M, int r, int col) -> double & { return M[(r - 1) * n + col - 1]; };
    if (nt == 0) {
      double s[3];
      for (int r = 0; r < 3; r++) { s[r] = 0; for (int q = 0; q < 6; q++) s[r] = s[r] + kAbc[r][q]; }
      for (int q = 0; q < 3; q++) at(A, 1, 1 + q) = kBnd[0][2 + q];
      at(B, 1, 1) = -s[0];
      for (int q = 0; q < 6; q++) at(B, 1, 2 + q) = kAbc[0][q];
      for (int q = 0; q < 4; q++) at(A, 2, 1 + q) = kBnd[1][1 + q];
      at(B, 2, 1) = kAbc[1][0]; at(B, 2, 2) = -s[1];
      for (int q = 1; q < 6; q++) at(B, 2, 2 + q) = kAbc[1][q];
      for (int q = 0; q < 5; q++) at(A, 3, 1 + q) = kBnd[2][q];
      at(B, 3, 1) = kAbc[2][0]; at(B, 3, 2) = kAbc[2][1]; at(B, 3, 3) = -s[2];
      for (int q = 2; q < 6; q++) at(B, 3, 2 + q) = kAbc[2][q];
      for (int i = 4; i <= n - 3; i++) {
        const double va[5] = {kBeta, kAlpha, 1, kAlpha, kBeta};
        const double vb[7] = {-kAC, -kAB, -kAA, 0, kAA, kAB, kAC};
        for (int q = 0; q < 5; q++) at(A, i, i - 2 + q) = va[q];
        for (int q = 0; q < 7; q++) at(B, i, i - 3 + q) = vb[q];
      }
    } else {
      FiltCoef z = filter_coef(fk, 0.5);
      const double al = z.alpha, be = z.beta, al2 = al * al, al3 = al2 * al, be2 = be * be, be3 = be2 * be;
      const double zv[5] = {z.b + 5 * z.c, z.a - 10 * z.c, z.a - 5 * z.c, z.b + z.c, z.c};
      double f = 1 / (1 + 5 * al3 + al2 * (8 + 22 * be) + be * (5 + 4 * be + 60 * be2) +
                      5 * al * (1 + be * (3 + 10 * be)));
      at(A, 1, 1) = 1;
      at(A, 1, 2) = (al * (1 + al) * (1 + 4 * al) + 2 * al * (7 + 3 * al) * be + 24 * (1 - al) * be2 - 80 * be3) * f;
      at(A, 1, 3) = (al3 + be + 14 * al2 * be + 60 * be3 + al * be * (3 + 46 * be)) * f;
      f = 1 / (1 + al2 * (9 - 5 * be) + al * (5 + (35 - 29 * be) * be) + be * (5 + 6 * (9 - 10 * be) * be));
      at(A, 2, 1) = (9 * al3 + 10 * be2 * (8 * be - 1) + 5 * al2 * (1 + 8 * be) + al * (1 + be * (4 + 81 * be))) * f;
      at(A, 2, 2) = 1;
      at(A, 2, 3) = (al * (1 + al * (5 + 9 * al)) + al * (5 + 36 * al) * be + (55 * al - 1) * be2 + 10 * be3) * f;
      at(A, 2, 4) = be * (1 + al * (5 + 9 * al) + 5 * be + 35 * al * be + 50 * be2) * f;
      const double r3[5] = {be, al, 1, al, be};
      for (int q = 0; q < 5; q++) at(A, 3, 1 + q) = r3[q];
      double zs = 0;
      for (int q = 0; q < 5; q++) zs = zs + zv[q];
      const double b3[6] = {zv[0], zv[1], -zs, zv[2], zv[3], zv[4]};
      for (int q = 0; q < 6; q++) at(B, 3, 1 + q) = b3[q];
      for (int i = 4; i <= n - 3; i++) {
        const double vb[7] = {z.c, z.b, z.a, -2 * (z.a + z.b + z.c), z.a, z.b, z.c};
        for (int q = 0; q < 5; q++) at(A, i, i - 2 + q) = r3[q];
        for (int q = 0; q < 7; q++) at(B, i, i - 3 + q) = vb[q];
      }
    }
    for (int col = 1; col <= n; col++) {
      for (int r = 1; r <= 3; r++) {
        at(A, n + 1 - r, n + 1 - col) = at(A, r, col);
        at(B, n + 1 - r, n + 1 - col) = sgn * at(B, r, col);
      }
    }
    std::vector<double> Ai;
    gj_inverse(n, A, Ai);
    std::vector<double> R(A);
    int i = n / 2 - 1;
    at(R, i, i + 2) = 0; at(R, i + 1, i + 2) = 0; at(R, i + 1, i + 3) = 0;
    i = n / 2 + 2;
    at(R, i, i - 2) = 0; at(R, i - 1, i - 2) = 0; at(R, i - 1, i - 3) = 0;
    std::vector<double> M = matmul(n, R, matmul(n, Ai, B));
    for (int row = 0; row <= 1; row++) {
      const int r = n / 2 + 1 + row;
      for (int q = 0; q <= ll; q++) {
        c.pbco[nt][row][ll - q] = M[(r - 1) * n + q];
        c.pbci[nt][row][q] = M[(r - 1) * n + ll + 1 + q];
      }
    }
  }
  for (int nt = 0; nt <= 1; nt++)
    for (int row = 0; row <= 1; row++) {
      double s = 0;
      for (int q = 0; q <= (nt == 0 ? LMD : LMF); q++) s = s + c.pbco[nt][row][q];
      c.pbcot[nt][row] = s;
    }
}

// Tables of the chunked line solver for one line family cut into P chunks of
// Cl = ceil(L/P) <= 32 points, laid out for one 64-lane load per register:
//   [0, 64)    forward carry maps: lane 4q+t holds entry t of chunk q's 2x2
//              response of its last pair (y[e-1], y[e]) to the incoming
//              pair (y[s-2], y[s-1]);
//   [64, 128)  backward carry maps: the response of (x[s], x[s+1]) to
//              (x[e+1], x[e+2]);
//   then per chunk q three registers: {l1 | l2}, {u1 | u2}, {u3 | 0}, row k
//   of the chunk in lane k (first half) or 32+k (second half).
// Factor entries the sweeps never use (l1, l2 of rows 0, 1; u2, u3 of rows
// L-1, L-2) are zero.  Rows past the chunk's end (up to the 32 slots) hold
// l1 = -1, l2 = 0 and u1 = u2 = 0, u3 = -1: with a zero right-hand side such
// a row only swaps the two carried values, exactly, so the sweeps run over
// the whole capacity without selects (see fwd_sweep / bwd_sweep).
std::vector<double> col_tables(int L, int P, const double *l1, const double *l2, const double *u1,
                               const double *u2, const double *u3) {
  const int Cl = (L + P - 1) / P, Pe = (L + Cl - 1) / Cl;
  std::vector<double> T(128 + 192 * (size_t)Pe, 0.0);
  auto L1 = [&](int i) { return i < 2 ? 0.0 : l1[i]; };
  auto L2 = [&](int i) { return i < 1 ? 0.0 : l2[i]; };
  auto U2 = [&](int i) { return i >= L - 1 ? 0.0 : u2[i]; };
  auto U3 = [&](int i) { return i >= L - 2 ? 0.0 : u3[i]; };
  for (int q = 0; q < Pe; q++) {
    const int s0 = q * Cl, e = std::min(L, s0 + Cl);
    double *F = &T[128 + 192 * (size_t)q];
    for (int i = s0; i < e; i++) {
      const int k = i - s0;
      F[k] = L1(i); F[32 + k] = L2(i);
      F[64 + k] = u1[i]; F[96 + k] = U2(i);
      F[128 + k] = U3(i);
    }
    for (int k = e - s0; k < 32; k++) {
      F[k] = -1.0; F[32 + k] = 0.0;
      F[64 + k] = 0.0; F[96 + k] = 0.0;
      F[128 + k] = -1.0;
    }
    for (int which = 0; which < 2; which++) {
      double ym2 = which == 0 ? 1 : 0, ym1 = which == 0 ? 0 : 1;
      for (int i = s0; i < e; i++) {
        const double y = -L1(i) * ym2 - L2(i) * ym1;
        ym2 = ym1; ym1 = y;
      }
      T[4 * q + which] = ym2;
      T[4 * q + 2 + which] = ym1;
      double xp1 = which == 0 ? 1 : 0, xp2 = which == 0 ? 0 : 1;
      for (int i = e - 1; i >= s0; i--) {
        const double x = -U2(i) * xp1 - U3(i) * xp2;
        xp2 = xp1; xp1 = x;
      }
      T[64 + 4 * q + which] = xp1;
      T[64 + 4 * q + 2 + which] = xp2;
    }
  }
  return T;
}

// chunk counts the chunked solver can run with (P waves of 64 lanes); the
// defaults (no autotuning) come from the powers of two among them
constexpr int kColP[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
constexpr int kNColP = 16;
constexpr int kColDefault[] = {1, 2, 4, 8, 16};

// chunk length of P chunks over L points, 0 if the chunked solver cannot run it
int col_chunk(int L, int P) {
  const int Cl = (L + P - 1) / P;
  return col_capacity(Cl) ? Cl : 0;
}

// doubles of the chunked-solver table of P chunks over L points (col_tables),
// padded to whole 128-byte lines
size_t col_table_doubles(int L, int P) {
  const int Cl = (L + P - 1) / P, Pe = (L + Cl - 1) / Cl;
  return (128 + 192 * (size_t)Pe + 15) / 16 * 16;
}

// doubles of all chunked-solver tables of a block with lines L[3]: every
// chunk count its lines allow, both kinds
size_t col_block_doubles(const int *L) {
  size_t n = 0;
  for (int d = 0; d < 3; d++)
    for (int P : kColP)
      if (col_chunk(L[d], P)) n += 2 * col_table_doubles(L[d], P);
  return n;
}

// default chunk count of lines of L points: the fewest chunks (a power of
// two) whose length stays within 24 points (registers for 4+ waves per SIMD),
// else within 32; 0 (the sequential kernel) if none fits
int col_default(int L) {
  for (int lim : {24, 32})
    for (int P : kColDefault)
      if ((L + P - 1) / P <= lim) return P;
  return 0;
}

// index of a chunk count in {0 (sequential), kColP...}: the line-length class
// of blocks whose default is that count
int col_class(int P) {
  for (int i = 0; i < kNColP; i++)
    if (kColP[i] == P) return i + 1;
  return 0;
}
constexpr int kNClass = kNColP + 1;

// GCBC spreading weights over the MBCI points inside a characteristic face
double gcbc_weights(double sbci[MBCI]) {
  std::vector<double> m = {kBnd[0][3], kBnd[0][4], 0, kBnd[1][2], kBnd[1][3], kBnd[1][4],
                           kBnd[2][1], kBnd[2][2], kBnd[2][3]};
  const double rhs[3] = {1, kBnd[1][1], kBnd[2][0]};
  std::vector<double> inv;
  double last = gj_inverse(3, m, inv);
  for (int i = 0; i < MBCI; i++) {
    double s = 0;
    for (int k = 0; k < 3; k++) s = s + inv[i * 3 + k] * rhs[k];
    sbci[i] = -s;
  }
  const double f = kPi / (MBCI + 1);
  double r = 0;
  for (int i = 0; i < MBCI; i++) {
    r = r + 1;
    sbci[i] = 0.5 * sbci[i] * (1 + std::cos(r * f));
  }
  return last;
}

}  // namespace

// ----------------------------------------------------------------------------
// context
// ----------------------------------------------------------------------------
// device allocations with one owner: a case (its blocks, tables and sponge
// lists) or the context (launch tables, descriptors, scalars)
struct Mem {
  std::vector<void *> p;
  size_t bytes = 0;
  void release() {
    for (void *q : p) (void)hipFree(q);
    p.clear();
    bytes = 0;
  }
};

// chunked-solver tables of one block: direction x kind x chunk-count index
using ColTab = std::array<const double *, 3 * 2 * kNColP>;

struct HostCase {
  int id = 0;                         // stable handle returned by tp_add_case
  int nb = 0, njct = 0;
  int block0 = 0, slot = -1;          // first global block and scalar slot while linked
  bool built = false, geom = false;   // device memory allocated / metrics and sponge set
  std::vector<int> dims, nbc, ncd, nsz, jct, jvel;
  int ip[TP_I_NPARAM] = {0};
  double dp[TP_D_NPARAM] = {0};
  CaseCoef coef;
  double cbca_ra0 = 0;
  double albed[2][3][5], albef[2][3][5], alphf = 0, betf = 0;
  Mem mem;
  std::vector<Block> blk;                   // host views of the case's blocks
  std::vector<ColTab> coltab;               // per block
  std::vector<std::vector<double>> grid;    // host copy of ss per block
  double sc[S_STRIDE];                      // scalars kept across a relink
  bool sc_saved = false;
  // parent blocks of the reference's file formats (tp_set_parents): parent of
  // each sub-block, its first point in the parent, the parents' last indices
  int npar = 0;
  std::vector<int> parent, pstart, pdims;
  // synthetic eddies of inflow 3 (tp_set_inflow_eddies), 10 doubles each
  std::vector<double> eddies;
};

struct DevList {
  Work *d = nullptr;
  int n = 0;
};

// launch tables over a set of blocks
// one launch of the line solves of a direction and kind: the blocks whose
// lines use P chunks (P = 0: the sequential kernel) of capacity C points
struct ColGroup {
  int P = 0, C = 0;
  bool closure = false;   // a block of the group has a closure face in this direction
  long long npts = 0;
  DevList work;
};

struct ListSet {
  DevList points, sponge, force, lines[3], exch, gprep, gapply[3], gcopy, wall, acap, aapply[3], normals;
  DevList surf;                     // wall-surface points, 256 per item
  std::vector<ColGroup> cg[3][2];   // line-solve launches per direction and kind
  std::vector<int> wr;   // first point-work item of each linked case, then the end
  std::vector<int> wwr;  // first wall-surface item of each linked case, then the end
};

struct ProfRec {
  hipEvent_t a, b;
  double bytes;
};

struct tp_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool finalized = false;
  int next_id = 0, nkrk = 0;
  std::vector<HostCase> cases;        // pending and built cases, in order of addition
  // linked view of the built cases, rebuilt whenever a case joins or leaves
  std::vector<Block> blocks;          // host copies of the device views
  std::vector<std::pair<int, int>> bsrc;   // (case index, block of the case) per global block
  std::vector<int> linked;            // case ids in scalar-slot order
  Mem link;
  Block *d_blocks = nullptr;
  CaseCoef *d_coef = nullptr;
  double *d_sc = nullptr, *d_part = nullptr, *d_wpart = nullptr;
  int *d_wr = nullptr, *d_wwr = nullptr, *d_jct = nullptr, *d_jvel = nullptr;
  int njct = 0, maxk = 0;
  ListSet L;
  int nlp[3] = {0, 0, 0};
  size_t lds[3] = {0, 0, 0};
  bool exact_lines = false;        // sequential one-lane-per-line sweeps
  bool keep_stress = false;        // tv written at every point of every stage
  bool qin = false;                // kernels enqueued now read the stage's state from qo
                                   // (StageArgs::qin): set by enqueue_step only
  int col_fixed = 0;               // requested chunks per line (0: chosen per direction)
  // autotuned chunks per line for each direction, kind and line-length class (-1: none)
  int tuned[3][2][kNClass];
  Mem lmem;                        // line-solve launch tables (rebuilt when the chunk choice changes)
  long long npts = 0;
  bool any_visc = false;
  // measurement
  bool prof = false;
  int prof_mask = 0;   // kernel classes timed while prof is on (bit k: class k)
  std::vector<ProfRec> pend[K_NCLASS];
  std::vector<hipEvent_t> pool;
  double ms[K_NCLASS] = {0}, bytes[K_NCLASS] = {0};
  int64_t cnt[K_NCLASS] = {0};
  // graph of one step
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  tp_context() {
    for (auto &a : tuned)
      for (auto &b : a)
        for (int &v : b) v = -1;
  }
};

namespace {

// bytes an allocation of n bytes is accounted with (dalloc)
size_t acct(size_t n) { return n == 0 ? 256 : n; }

int dalloc(Mem &m, void **p, size_t bytes) {
  bytes = acct(bytes);
  HIPCHK(hipMalloc(p, bytes));
  m.p.push_back(*p);
  m.bytes += bytes;
  return 0;
}

template <class T>
int upload_vec(Mem &m, T **dptr, const std::vector<T> &v) {
  int rc = dalloc(m, (void **)dptr, sizeof(T) * std::max<size_t>(v.size(), 1));
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*dptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return 0;
}

int upload_list(Mem &m, DevList &L, const std::vector<Work> &v) {
  L.n = (int)v.size();
  return upload_vec(m, &L.d, v);
}

hipEvent_t take_event(tp_context *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Times the launches enqueued in its scope (profiling on, class selected).
// Plain: event packets recorded on the stream before and after them.  ext:
// the two events are handed to the first and the last launch (start() and
// stop()), which carry them in their own dispatches: no packets between the
// launches, so timing costs the stream next to nothing.
struct Timed {
  tp_context *c;
  int cls;
  double bytes;
  bool ext;
  hipEvent_t a = nullptr, b = nullptr;
  bool used = false;
  Timed(tp_context *cc, int k, double by, bool ext_ = false) : c(cc), cls(k), bytes(by), ext(ext_) {
    if (c->prof && (c->prof_mask >> cls & 1)) {
      a = take_event(c);
      if (a && ext) {
        b = take_event(c);
        if (!b) { c->pool.push_back(a); a = nullptr; }
      } else if (a) {
        (void)hipEventRecord(a, c->stream);
      }
    }
  }
  // the events of launch i of n (ext)
  hipEvent_t start(int i) { return a && i == 0 ? a : nullptr; }
  hipEvent_t stop(int i, int n) {
    if (!a) return nullptr;
    used = true;
    return i == n - 1 ? b : nullptr;
  }
  ~Timed() {
    if (!a) return;
    if (ext) {
      if (used) {
        c->pend[cls].push_back({a, b, bytes});
      } else {
        c->pool.push_back(a);
        c->pool.push_back(b);
      }
      return;
    }
    hipEvent_t e = take_event(c);
    if (e) {
      (void)hipEventRecord(e, c->stream);
      c->pend[cls].push_back({a, e, bytes});
    }
  }
};

// ---------------- launch helpers ----------------
void run_points(tp_context *c, int which, StageArgs a, int cls, double bytes_per_point) {
  if (!c->L.points.n) return;
  Timed t(c, cls, bytes_per_point * (double)c->npts);
  a.work = c->L.points.d;
  launch_point(which, c->L.points.n, c->stream, a);
}

void run_faces(tp_context *c, const DevList &L, int which, StageArgs a, int cls) {
  if (!L.n) return;
  Timed t(c, cls, 0.0);
  a.work = L.d;
  launch_face(which, L.n, c->stream, a);
}

StageArgs base_args(tp_context *c) {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.blocks = c->d_blocks;
  a.coef = c->d_coef;
  a.sc = c->d_sc;
  a.part = c->d_part;
  a.wpart = c->d_wpart;
  a.qin = c->qin ? 1 : 0;
  return a;
}

// closure sums of nf fields (columns col0 + f fstep, f < nf) in one launch,
// into recv slots 0..nf-1
void run_exchange(tp_context *c, int nt, int nrt, int n45, int arr, int col0, int nf = 1, int fstep = 0) {
  const DevList &L = c->L.exch;
  if (!L.n) return;
  Timed t(c, K_EXCH, 0.0);
  ExchArgs a;
  a.blocks = c->d_blocks; a.work = L.d; a.coef = c->d_coef;
  a.nt = nt; a.nrt = nrt; a.n45 = n45; a.arr = arr; a.col0 = col0; a.fstep = fstep;
  launch_exchange(L.n, nf, c->stream, a);
}

// line solves of direction d over nf fields in one launch: field f reads
// column in_col + f in_step, writes out_col + f out_step, uses recv slot f
// (filled by an exchange over the same fields) and captures drva column m + f
void run_line(tp_context *c, int kind, int d, int in_arr, int in_col, int out_arr, int out_col, int m,
              int acc, int nf = 1, int in_step = 0, int out_step = 0, int copy_arr = -1, int copy_col = 0,
              bool direct = false) {
  if (!c->L.lines[d].n) return;
  // algorithmic bytes: one field in, one out; an accumulated output is also
  // read, a copy of the output (the last filter's qo) also written
  Timed t(c, kind == 0 ? K_DERIV : K_FILTER, ((acc ? 24.0 : 16.0) + (copy_arr >= 0 ? 8.0 : 0.0)) * (double)c->npts * nf,
          true);
  LineArgs a;
  a.blocks = c->d_blocks; a.work = c->L.lines[d].d; a.coef = c->d_coef;
  a.in_arr = in_arr; a.in_col = in_col; a.out_arr = out_arr; a.out_col = out_col;
  a.m = m; a.acc = acc; a.nlp = c->nlp[d];
  a.in_step = in_step; a.out_step = out_step;
  a.copy_arr = copy_arr; a.copy_col = copy_col;
  a.direct = direct ? 1 : 0;
  a.n45 = 0;
  if (c->exact_lines) {
    hipEvent_t e0 = t.start(0), e1 = t.stop(0, 1);
    if (kind == 0) launch_deriv(d, c->L.lines[d].n, nf, c->lds[d], c->stream, a, e0, e1);
    else launch_filter(d, c->L.lines[d].n, nf, c->lds[d], c->stream, a, e0, e1);
    return;
  }
  int ng = 0, i = 0;
  for (const ColGroup &g : c->L.cg[d][kind]) ng += g.work.n ? 1 : 0;
  for (const ColGroup &g : c->L.cg[d][kind]) {
    if (!g.work.n) continue;
    a.work = g.work.d;
    a.direct = direct && g.closure ? 1 : 0;
    hipEvent_t e0 = t.start(i), e1 = t.stop(i, ng);
    if (g.P) launch_col(d, kind, g.C, g.P, g.work.n, nf, c->stream, a, e0, e1);
    else if (kind == 0) launch_deriv(d, g.work.n, nf, c->lds[d], c->stream, a, e0, e1);
    else launch_filter(d, g.work.n, nf, c->lds[d], c->stream, a, e0, e1);
    i++;
  }
}

// drop a captured step: it holds the launch configuration it was built with
void drop_graph(tp_context *c) {
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
}

// chunks per line of block g's line solves in direction d, kind: the tuned
// count of its line-length class if that fits its lines, else the requested
// count if it fits, else the default (0: the sequential kernel)
int block_chunks(const tp_context *c, int g, int d, int kind) {
  const int L = c->blocks[g].L[d];
  const int def = col_default(L);
  auto fits = [&](int P) { return P == 0 || col_chunk(L, P) != 0; };
  const int t = c->tuned[d][kind][col_class(def)];
  if (t >= 0 && fits(t)) return t;
  if (c->col_fixed && fits(c->col_fixed)) return c->col_fixed;
  return def;
}

// point every block's chunked-solver table pointers at its chosen chunk
// count (host copies of the blocks only)
void choose_lines(tp_context *c) {
  for (size_t g = 0; g < c->blocks.size(); g++) {
    const auto &src = c->bsrc[g];
    for (int d = 0; d < 3; d++)
      for (int kind = 0; kind < 2; kind++) {
        const int ci = col_class(block_chunks(c, (int)g, d, kind)) - 1;
        c->blocks[g].col[d][kind] =
            ci < 0 ? nullptr : c->cases[src.first].coltab[src.second][(d * 2 + kind) * kNColP + ci];
      }
  }
}

// line-solve launches over the global blocks gs: one per distinct chunk
// count and chunk capacity of each direction and kind
int make_groups(tp_context *c, const std::vector<int> &gs, Mem &m, ListSet &L) {
  for (int d = 0; d < 3; d++)
    for (int kind = 0; kind < 2; kind++) {
      std::vector<ColGroup> groups;
      std::vector<std::vector<Work>> works;
      for (int g : gs) {
        const Block &b = c->blocks[g];
        const int P = block_chunks(c, g, d, kind);
        const int C = P ? col_capacity((b.L[d] + P - 1) / P) : 0;
        // a chunked launch reads the block's carry-map and factor table
        if (P && (!C || !b.col[d][kind]))
          return fail(TP_E_STATE, "block " + std::to_string(g) + " direction " + std::to_string(d + 1) +
                                      ": no chunked-solver table for " + std::to_string(P) + " chunks");
        size_t k = 0;
        while (k < groups.size() && (groups[k].P != P || groups[k].C != C)) k++;
        if (k == groups.size()) {
          groups.emplace_back();
          groups.back().P = P;
          groups.back().C = C;
          works.emplace_back();
        }
        for (int ip = 0; ip < 2; ip++)
          if (b.ndf[d][ip]) groups[k].closure = true;
        const int nlines = b.N / b.L[d];
        const int step = P ? 64 : c->nlp[d] - 1;
        for (int q = 0; q < nlines; q += step) works[k].push_back({g, q, std::min(step, nlines - q), d});
        groups[k].npts += b.N;
      }
      // a block and its neighbour across the end face of a closure (the two
      // halves of a split line) interleave their work items, so that the line
      // solves that form their closure sums from each other's lines run side
      // by side (with xcd_item, on one XCD) and find those lines in its L2
      for (auto &wl : works) {
        std::vector<Work> out;
        out.reserve(wl.size());
        std::vector<char> done(wl.size(), 0);
        for (size_t i = 0; i < wl.size(); i++) {
          if (done[i]) continue;
          const int ga = wl[i].blk;
          size_t ea = i;
          while (ea < wl.size() && wl[ea].blk == ga) ea++;   // the items of ga: [i, ea)
          const Block &ba = c->blocks[ga];
          const int gb = ba.ndf[d][1] ? ba.ncd[d][1] : -1;
          size_t sb = wl.size();
          if (gb >= 0 && gb != ga)
            for (size_t q = ea; q < wl.size(); q++)
              if (wl[q].blk == gb && !done[q]) { sb = q; break; }
          size_t eb = sb;
          while (eb < wl.size() && wl[eb].blk == gb) eb++;
          if (sb < wl.size() && eb - sb == ea - i) {
            for (size_t q = 0; q < ea - i; q++) {
              out.push_back(wl[i + q]); done[i + q] = 1;
              out.push_back(wl[sb + q]); done[sb + q] = 1;
            }
          } else {
            for (size_t q = i; q < ea; q++) { out.push_back(wl[q]); done[q] = 1; }
          }
          i = ea - 1;
        }
        wl.swap(out);
      }
      for (size_t k = 0; k < groups.size(); k++)
        if (int rc = upload_list(m, groups[k].work, works[k])) return rc;
      L.cg[d][kind] = std::move(groups);
    }
  return 0;
}

// rebuild the line-solve launches of the linked blocks after the chunk
// choice changed, and refresh the device block views
int regroup(tp_context *c) {
  drop_graph(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  choose_lines(c);
  c->lmem.release();
  std::vector<int> gs(c->blocks.size());
  for (size_t g = 0; g < gs.size(); g++) gs[g] = (int)g;
  if (int rc = make_groups(c, gs, c->lmem, c->L)) return rc;
  if (!c->d_blocks) return 0;
  HIPCHK(hipMemcpy(c->d_blocks, c->blocks.data(), sizeof(Block) * c->blocks.size(), hipMemcpyHostToDevice));
  return 0;
}

// the chunked line solves of this kind can form their closure sums
// themselves (every launch group chunked; the sequential solves read the
// exchange's recv slots), so that no exchange launch is needed: a periodic
// block reads the other end of its own lines, which the workgroup's other
// waves load at the same time; a split line reads its partner block's lines,
// whose work items run beside its own.  Only for solves whose input no earlier
// solve of the same exchange changed (the derivatives; not the in-place
// filters).
bool direct_ok(const tp_context *c, int kind) {
  if (c->exact_lines) return false;
  for (int d = 0; d < 3; d++)
    for (const ColGroup &g : c->L.cg[d][kind])
      if (g.work.n && !g.P) return false;
  return true;
}

// lean (inside a whole time step): the last filter leaves the result in qo
// only, and the step's first stage reads its state from there until its
// update writes qa from qo
void enqueue_filter_step(tp_context *c, bool lean = false) {
  // src/Main3D.f90:555-560: per field m, mpigo then the xi, eta, zeta filters
  // in place; the five fields are independent, so each of these steps runs
  // over all five in one launch; the last also leaves the result in qo
  // (qo = qa, :561)
  // (the filters work in place, so the eta and zeta filters take the closure
  // sums of the unfiltered field: one exchange before the three, as mpigo)
  run_exchange(c, 1, 1, 0, F_QA, 0, 5, 1);
  run_line(c, 1, 0, F_QA, 0, F_QA, 0, 0, 0, 5, 1, 1);
  run_line(c, 1, 1, F_QA, 0, F_QA, 0, 0, 0, 5, 1, 1);
  if (lean) run_line(c, 1, 2, F_QA, 0, F_QO, 0, 0, 0, 5, 1, 1);
  else run_line(c, 1, 2, F_QA, 0, F_QA, 0, 0, 0, 5, 1, 1, F_QO, 0);
}

// ---- the sections of a stage (src/Main3D.f90:567-958), shared by the fused
// stage and the per-section drop-ins ----

// stage start: frame velocity and inflow weights (:571-572), primitives
// (:576-583), time step at the first stage (:587-622)
// prim_ready: the previous stage's update already left this stage's
// primitives (enqueue_rk with prim)
// lean (inside a whole time step): the primitives pass leaves 1/rho and the
// viscosity only on the steps whose time-step limit reads them
void enqueue_begin(tp_context *c, int nk, bool prim_ready = false, bool lean = false) {
  const int ncase = (int)c->linked.size();
  // (with prim_ready the previous stage's k_sc_rk left this stage's start
  // scalars too)
  if (!prim_ready) launch_sc_stage(c->d_coef, c->d_sc, ncase, nk, c->stream);
  StageArgs a = base_args(c);
  a.nk = nk;
  a.opt = lean ? 1 : 0;
  if (!prim_ready) run_points(c, PK_PRIM, a, K_POINT, lean ? 80.0 : 96.0);
  a.opt = 0;
  if (nk == 1) {
    run_points(c, PK_DTRED, a, K_POINT, 0.0);
    launch_sc_dt(c->d_coef, c->d_sc, c->d_part, c->d_wr, ncase, 1, c->stream);
  }
}

// src/Main3D.f90:627-651: gradients of u, v, w, T (de(:,2:5)) along the
// three index directions into fx(:,3f+d) -- the four fields in one launch per
// direction
void enqueue_gradients(tp_context *c) {
  if (!c->any_visc) return;
  const bool dir = direct_ok(c, 0);
  if (!dir) run_exchange(c, 0, 1, 0, F_DE, 1, 4, 1);
  run_line(c, 0, 2, F_DE, 1, F_FX, 2, 0, 0, 4, 1, 3, -1, 0, dir);
  run_line(c, 0, 1, F_DE, 1, F_FX, 1, 0, 0, 4, 1, 3, -1, 0, dir);
  run_line(c, 0, 0, F_DE, 1, F_FX, 0, 0, 0, 4, 1, 3, -1, 0, dir);
}

// surface forces and pressure statistics, src/Main3D.f90:695-706 (the
// kernels return at once for cases that do not sample this step)
void enqueue_surface(tp_context *c, int nk) {
  if (nk != 1 || !c->L.surf.n) return;
  StageArgs a = base_args(c);
  a.nk = nk;
  run_faces(c, c->L.surf, FK_SURF_SAMPLE, a, K_MISC);
  launch_surf_reduce(c->d_coef, c->d_sc, c->d_wpart, c->d_wwr, (int)c->linked.size(), c->stream);
}

// de(:,m) = d/dxi F + d/deta G + d/dzeta H of the flux components
// fx(:,3(m-1)+d), the five fields m in one launch per direction, summed in
// the reference's order (src/Main3D.f90:721-768)
void enqueue_flux_lines(tp_context *c) {
  const bool dirf = direct_ok(c, 0);
  if (!dirf) run_exchange(c, 0, 0, 0, F_FX, 0, 5, 3);
  run_line(c, 0, 0, F_FX, 0, F_DE, 0, 1, 0, 5, 3, 1, -1, 0, dirf);
  run_line(c, 0, 1, F_FX, 1, F_DE, 0, 1, 1, 5, 3, 1, -1, 0, dirf);
  run_line(c, 0, 2, F_FX, 2, F_DE, 0, 1, 1, 5, 3, 1, -1, 0, dirf);
}

// sponge with the inflow target, src/SingleTurbulence3D.f90:332-374
void enqueue_sponge(tp_context *c) {
  StageArgs s = base_args(c);
  if (c->L.sponge.n) {
    Timed t(c, K_MISC, 0.0);
    s.work = c->L.sponge.d;
    launch_point(PK_SPONGE, c->L.sponge.n, c->stream, s);
  }
  // the body force, forcego (src/rpt.f90:758-769), which src/Main3D.f90:776
  // applies right after the sponge
  if (c->L.force.n) {
    Timed t(c, K_MISC, 0.0);
    s.work = c->L.force.d;
    launch_point(PK_FORCE, c->L.force.n, c->stream, s);
  }
}

// characteristic boundary and interface conditions, src/Main3D.f90:778-869
void enqueue_gcbc(tp_context *c) {
  const StageArgs a = base_args(c);
  run_faces(c, c->L.gprep, FK_GCBC_PREP, a, K_GCBC);
  run_faces(c, c->L.gcopy, FK_COPY, a, K_GCBC);
  for (int d = 0; d < 3; d++) run_faces(c, c->L.gapply[d], d == 0 ? FK_GCBC_APPLY_XI : FK_GCBC_APPLY, a, K_GCBC);
}

// frame velocity for the update and the RK update, src/Main3D.f90:873-881
// prim: the update also leaves the primitives the next stage reads
// (de(:,2:5) and p of the new state, k_primitives' arithmetic), and the wall, junction
// and averaging passes after it form them again where they change the state,
// so that the next stage starts without the primitives pass and its read of
// qa (the residual in de is read first)
void enqueue_rk(tp_context *c, int nk, bool prim = false) {
  launch_sc_rk(c->d_coef, c->d_sc, (int)c->linked.size(), nk, prim ? 1 : 0, c->stream);
  StageArgs a = base_args(c);
  a.nk = nk;
  if (prim) run_points(c, PK_RK_PRIM, a, K_RK, 168.0);
  else run_points(c, PK_RK, a, K_RK, 128.0);
}

// wall temperature and velocity, src/Main3D.f90:884-895
void enqueue_wall(tp_context *c, bool prim = false) {
  StageArgs a = base_args(c);
  a.opt = prim ? 1 : 0;
  run_faces(c, c->L.wall, FK_WALL, a, K_GCBC);
}

// corner junctions, src/SingleTurbulence3D.f90:378-443
void enqueue_junction(tp_context *c, bool prim = false) {
  if (c->njct)
    launch_junction(c->d_blocks, c->d_coef, c->d_jct, c->d_jvel, c->njct, c->maxk, prim ? 1 : 0, c->stream);
}

// interface averaging, src/Main3D.f90:897-952
void enqueue_avg(tp_context *c, bool prim = false) {
  StageArgs a = base_args(c);
  run_faces(c, c->L.acap, FK_AVG_CAP, a, K_GCBC);
  run_faces(c, c->L.acap, FK_COPY, a, K_GCBC);
  a.opt = prim ? 1 : 0;
  for (int d = 0; d < 3; d++) run_faces(c, c->L.aapply[d], FK_AVG_APPLY, a, K_GCBC);
}

// one stage, src/Main3D.f90:567-958: the stresses and heat fluxes stay in
// registers of the pass that forms the fluxes (tv is written only where it
// is read, or everywhere with TP_OPT_STRESSES)
// prim_ready: the previous stage left this one's primitives; prim_next: leave
// the next stage's (inside a time step only: after a single stage the state
// is the reference's, with the residual in de)
void enqueue_stage(tp_context *c, int nk, bool prim_ready = false, bool prim_next = false, bool lean = false) {
  enqueue_begin(c, nk, prim_ready, lean);
  enqueue_gradients(c);
  {
    StageArgs f = base_args(c);
    f.nk = nk;
    f.opt = c->keep_stress ? 1 : 0;
    run_points(c, PK_FLUXALL, f, K_FLUX, 8.0 * (c->any_visc ? 27 + 15 : 14 + 15));
  }
  enqueue_surface(c, nk);
  enqueue_flux_lines(c);
  enqueue_sponge(c);
  enqueue_gcbc(c);
  c->qin = false;   // the update reads qo and writes qa in any case
  enqueue_rk(c, nk, prim_next);
  enqueue_wall(c, prim_next);
  enqueue_junction(c, prim_next);
  enqueue_avg(c, prim_next);
}

// a time step: the filter step, then the stages, each but the last leaving
// the next one's primitives
void enqueue_step(tp_context *c) {
  if (c->linked.empty()) return;
  enqueue_filter_step(c, true);
  c->qin = true;    // until the first stage's update
  for (int nk = 1; nk <= c->nkrk; nk++) enqueue_stage(c, nk, nk > 1, nk < c->nkrk, true);
  c->qin = false;
  launch_sc_end(c->d_sc, (int)c->linked.size(), nullptr, c->stream);
}

int check_ready(tp_context *c) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (!c->finalized) return fail(TP_E_STATE, "context not finalized (call tp_finalize)");
  return 0;
}

double *array_ptr(const Block &b, int id, int64_t *count) {
  const int64_t N = b.N;
  switch (id) {
    case TP_QA: *count = 5 * N; return b.qa;
    case TP_QO: *count = 5 * N; return b.qo;
    case TP_DE: *count = 5 * N; return b.de;
    case TP_RR: *count = 3 * N; return b.rr;
    case TP_SS: *count = 3 * N; return b.ss;
    case TP_XIM: *count = 3 * N; return b.xim;
    case TP_ETM: *count = 3 * N; return b.etm;
    case TP_ZEM: *count = 3 * N; return b.zem;
    case TP_P: *count = N; return b.p;
    case TP_YACO: *count = N; return b.yaco;
    case TP_TV: *count = 9 * N; return b.tv;
    case TP_DRVA1: case TP_DRVA2: case TP_DRVA3: *count = 10LL * b.F[id - TP_DRVA1]; return b.drva[id - TP_DRVA1];
    case TP_DRVB1: case TP_DRVB2: case TP_DRVB3: *count = 10LL * b.F[id - TP_DRVB1]; return b.drvb[id - TP_DRVB1];
    case TP_RECV1: case TP_RECV2: case TP_RECV3: *count = 6LL * b.F[id - TP_RECV1]; return b.recv[id - TP_RECV1];
    case TP_CM1: case TP_CM2: case TP_CM3: *count = 6LL * b.F[id - TP_CM1]; return b.cm[id - TP_CM1];
    case TP_XL: *count = 2LL * b.nlu; return b.lud;
    case TP_XU: *count = 3LL * b.nlu; return b.lud + 2 * b.nlu;
    case TP_YL: *count = 2LL * b.nlu; return b.luf;
    case TP_YU: *count = 3LL * b.nlu; return b.luf + 2 * b.nlu;
    case TP_RPEX: {
      int n = 0;
      for (int d = 0; d < 3; d++)
        for (int s = 0; s < 2; s++)
          if (b.nbc[d][s] == 20 || b.nbc[d][s] == 25) n += b.F[d];
      *count = n;
      return b.rpex;
    }
  }
  *count = 0;
  return nullptr;
}

// index in c->cases of the case with handle id (-1: none)
int case_index(const tp_context *c, int id) {
  for (size_t i = 0; i < c->cases.size(); i++)
    if (c->cases[i].id == id) return (int)i;
  return -1;
}

// the linked case with handle id (nullptr: unknown, or not on the device)
HostCase *linked_case(tp_context *c, int id) {
  const int i = case_index(c, id);
  return i >= 0 && c->cases[i].built && c->cases[i].slot >= 0 ? &c->cases[i] : nullptr;
}

int block_index(tp_context *c, int cs, int b) {
  const HostCase *h = linked_case(c, cs);
  if (!h || b < 0 || b >= h->nb) return -1;
  return h->block0 + b;
}

}  // namespace

// ---- seeded random Fourier modes of the inflow turbulence ----
namespace {
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};
}  // namespace

// ---- a case's device memory, and the linked view of all built cases ----
namespace {

// sub-array offsets of one block's arena (256-byte aligned)
struct Arena {
  size_t qa, qo, de, rr, ss, xi, et, ze, p, ya, tv, fx, f[3][4], gq, lud, luf, rp, sa, sn, ss2, va, total;
};

// les: the case needs the single-precision Smagorinsky volume term per point
Arena arena_layout(const Block &b, int nwall, bool les) {
  Arena A;
  size_t off = 0;
  auto take = [&](size_t n) { size_t o = off; off += ((n * sizeof(double) + 255) / 256) * 256; return o; };
  const size_t N = b.N;
  A.qa = take(5 * N); A.qo = take(5 * N); A.de = take(5 * N); A.rr = take(3 * N); A.ss = take(3 * N);
  A.xi = take(3 * N); A.et = take(3 * N); A.ze = take(3 * N); A.p = take(N); A.ya = take(N);
  A.tv = take(9 * N); A.fx = take(15 * N);
  for (int d = 0; d < 3; d++) {
    A.f[d][0] = take(10 * (size_t)b.F[d]); A.f[d][1] = take(10 * (size_t)b.F[d]);
    A.f[d][2] = take(6 * RCV_SLOTS * (size_t)b.F[d]); A.f[d][3] = take(6 * (size_t)b.F[d]);
  }
  A.gq = take(14 * (size_t)b.F[0]);
  A.lud = take(5 * (size_t)b.nlu); A.luf = take(5 * (size_t)b.nlu); A.rp = take(std::max(nwall, 1));
  A.sa = take(b.nsurf); A.sn = take(3 * (size_t)b.nsurf); A.ss2 = take(3 * (size_t)b.nsurf);
  A.va = les ? take((N + 1) / 2) : 0;   // N floats
  A.total = off;
  return A;
}

// sizes, boundary codes, LU offsets and wall slots of block lb of a case;
// ncd holds the neighbour's block index within the case
Block block_shape(const HostCase &h, int lb, int *nwall) {
  Block b;
  std::memset(&b, 0, sizeof(b));
  b.nx = h.dims[3 * lb] + 1; b.ny = h.dims[3 * lb + 1] + 1; b.nz = h.dims[3 * lb + 2] + 1;
  b.N = b.nx * b.ny * b.nz;
  b.L[0] = b.nx; b.L[1] = b.ny; b.L[2] = b.nz;
  b.F[0] = b.ny * b.nz; b.F[1] = b.nz * b.nx; b.F[2] = b.nx * b.ny;
  b.cs = -1;
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      b.nbc[d][s] = h.nbc[6 * lb + 2 * d + s];
      b.ncd[d][s] = h.ncd[6 * lb + 2 * d + s];
      const int np = b.nbc[d][s];
      b.ndf[d][s] = (np == 35 || np == 40 || np == 45) ? 1 : 0;
    }
  b.nlu = b.nx + b.ny + b.nz;
  b.lu_off[0] = 0; b.lu_off[1] = b.nx; b.lu_off[2] = b.nx + b.ny;
  int nw = 0;
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      if (b.nbc[d][s] == 20 || b.nbc[d][s] == 25) { b.wall_off[d][s] = nw; nw += b.F[d]; }
      else b.wall_off[d][s] = -1;
    }
  *nwall = nw;
  // the wall surface of src/rpt.f90:774-830: the eta face with the wall code
  const int bcw = 20 + 5 * h.ip[TP_I_NVISCOUS];
  b.nsurf = 0;
  b.sside = 0;
  if (b.nbc[1][0] == bcw) { b.nsurf = b.nx * b.nz; b.sside = 0; }
  else if (b.nbc[1][1] == bcw) { b.nsurf = b.nx * b.nz; b.sside = 1; }
  return b;
}

// run fn(lo, hi, t) over [0, n) in contiguous ranges on nt host threads
// (range t covers [t n/nt, (t+1) n/nt))
template <class F>
void host_parallel(int n, int nt, F fn) {
  if (nt <= 1 || n < 65536) {
    fn(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
    th.emplace_back([=]() { fn(lo, hi, t); });
  }
  for (auto &t : th) t.join();
}

int host_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}

// sponge points of block lb from its coordinates, SingleTurbulence3D.f90:209-249:
// the point list with the strength sg and the streamwise factor lm of each
// (the coefficients are (1/yaco) sg lm and (1/yaco) sg/(gam-1))
void sponge_select(const HostCase &h, int lb, const Block &b, const double *x, const double *y,
                   const double *z, std::vector<int> &lst, std::vector<double> &sgv, std::vector<double> &lmv) {
  const double domlen = h.dp[TP_D_DOMLEN], szth1 = h.dp[TP_D_SZTH1], szth2 = h.dp[TP_D_SZTH2];
  const double szxt = h.dp[TP_D_SZXT], szco = h.dp[TP_D_SZCO];
  const double p07 = (double)0.7f;   // single-precision literal 0.7e0 of the reference
  const double szr[2][3] = {{1 / szth1, 1 / szth2, 0}, {1 / (szth1 + szxt), 1 / szth2, 0}};
  const double szp[2][3] = {{-p07 * domlen + szth1, -p07 * domlen + szth2, 0},
                            {domlen - szth1, domlen + szxt - szth2, 0}};
  int nsz[3][2];
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) nsz[d][s] = h.nsz[6 * lb + 2 * d + s];
  const double ra1 = 0.25, ra2 = kPi / (2 * domlen);
  const int nt = host_threads();
  std::vector<std::vector<int>> tl(nt);
  std::vector<std::vector<double>> ts(nt), tm(nt);
  host_parallel(b.N, nt, [&](int lo, int hi, int t) {
    for (int l = lo; l < hi; l++) {
      const double cx[3] = {x[l], y[l], z[l]};
      double r[3];
      for (int d = 0; d < 3; d++)
        r[d] = nsz[d][0] * szr[0][d] * std::fmax(szp[0][d] - cx[d], 0.0) +
               nsz[d][1] * szr[1][d] * std::fmax(cx[d] - szp[1][d], 0.0);
      // 1 + cos(pi r) is exactly 2 at r = 0 (outside the band of that
      // direction): most points need no cosine
      const auto f = [](double rr) { return rr == 0.0 ? 2.0 : 1 + std::cos(kPi * rr); };
      const double sg = szco * (1 - 0.125 * f(r[0]) * f(r[1]) * f(r[2]));
      if (sg > 0) {
        const double lm = ra1 + 0.5 * (1 - ra1) * (1 + std::cos(ra2 * (std::fmin(x[l], domlen) + domlen)));
        tl[t].push_back(l);
        ts[t].push_back(sg);
        tm[t].push_back(lm);
      }
    }
  });
  lst.clear(); sgv.clear(); lmv.clear();
  for (int t = 0; t < nt; t++) {
    lst.insert(lst.end(), tl[t].begin(), tl[t].end());
    sgv.insert(sgv.end(), ts[t].begin(), ts[t].end());
    lmv.insert(lmv.end(), tm[t].begin(), tm[t].end());
  }
}

// inflow points: sponge points upstream of the inflow sponge edge inside the
// band |y| < gustw, with the ramp sin^2 of SingleTurbulence3D.f90:253-268
void inflow_select(const HostCase &h, const double *x, const double *y, const std::vector<int> &lst,
                   std::vector<int> &gix, std::vector<int> &pts, std::vector<double> &atz) {
  const double domlen = h.dp[TP_D_DOMLEN], szth1 = h.dp[TP_D_SZTH1];
  const double edge = -(double)0.7f * domlen + szth1;
  const double gw = h.dp[TP_D_GUSTW], ra0 = kPi / szth1;
  gix.assign(lst.size(), -1);
  pts.clear(); atz.clear();
  for (size_t q = 0; q < lst.size(); q++) {
    const int l = lst[q];
    if (x[l] < edge && std::fabs(y[l]) < gw) {
      gix[q] = (int)pts.size();
      pts.push_back(l);
      const double sn = std::sin(std::fmin(ra0 * (x[l] + domlen), kPi / 2));
      atz.push_back(sn * sn);
    }
  }
}

int inflow_nmodes(const HostCase &h) { return h.coef.inflow == 2 ? 1 : h.ip[TP_I_NMODES]; }

// forcing points of a block and their squared distance from the forcing
// centre (forceup, src/rpt.f90:731-740): the points within rfor of (xfor, yfor)
void force_select(const HostCase &h, const Block &b, const double *x, const double *y, std::vector<int> &lst,
                  std::vector<double> &r2) {
  lst.clear(); r2.clear();
  if (h.ip[TP_I_FORCING] != 1) return;
  const double xf = h.dp[TP_D_XFOR], yf = h.dp[TP_D_YFOR], rf = h.dp[TP_D_RFOR], ra2 = rf * rf;
  for (int l = 0; l < b.N; l++) {
    const double r = (x[l] - xf) * (x[l] - xf) + (y[l] - yf) * (y[l] - yf);
    if (r - ra2 < 0) { lst.push_back(l); r2.push_back(r); }
  }
}

// device bytes a case takes once admitted: block arenas, chunked-solver
// tables, and its sponge and inflow arrays (from the host grid)
size_t case_bytes(const HostCase &h) {
  size_t total = 0;
  for (int lb = 0; lb < h.nb; lb++) {
    int nwall = 0;
    const Block b = block_shape(h, lb, &nwall);
    total += acct(arena_layout(b, nwall, h.ip[TP_I_LES] == 1).total);
    total += acct(sizeof(double) * col_block_doubles(b.L));
    if (lb >= (int)h.grid.size() || h.grid[lb].empty()) continue;
    const double *x = h.grid[lb].data(), *y = x + b.N, *z = y + b.N;
    std::vector<int> lst, gix, pts;
    std::vector<double> sg, lm, atz;
    sponge_select(h, lb, b, x, y, z, lst, sg, lm);
    if (!lst.empty()) total += acct(4 * lst.size()) + 2 * acct(8 * lst.size());
    {
      std::vector<int> fl;
      std::vector<double> fr;
      force_select(h, b, x, y, fl, fr);
      if (!fl.empty()) total += acct(4 * fl.size()) + acct(24 * fl.size());
    }
    if (h.coef.inflow) {
      inflow_select(h, x, y, lst, gix, pts, atz);
      total += acct(4 * std::max<size_t>(gix.size(), 1));
      if (!pts.empty())
        total += acct(8 * pts.size()) + acct(sizeof(double) * 3 * (h.coef.ntz + 1) * pts.size());
    }
  }
  return total;
}

// device memory of a case: block arenas (zeroed), LU factors of the
// derivative and filter matrices (src/Main3D.f90:316-332), chunked-solver
// tables for every chunk count the block's lines allow, and the grid if set
int build_case(HostCase &h) {
  h.blk.assign(h.nb, Block{});
  h.coltab.assign(h.nb, ColTab{});
  h.grid.resize(h.nb);
  int rc;
  for (int lb = 0; lb < h.nb; lb++) {
    int nwall = 0;
    Block b = block_shape(h, lb, &nwall);
    const bool les = h.ip[TP_I_LES] == 1;
    const Arena A = arena_layout(b, nwall, les);
    char *base = nullptr;
    if ((rc = dalloc(h.mem, (void **)&base, A.total))) return rc;
    HIPCHK(hipMemset(base, 0, A.total));
    auto P = [&](size_t o) { return (double *)(base + o); };
    b.qa = P(A.qa); b.qo = P(A.qo); b.de = P(A.de); b.rr = P(A.rr); b.ss = P(A.ss);
    b.xim = P(A.xi); b.etm = P(A.et); b.zem = P(A.ze); b.p = P(A.p); b.yaco = P(A.ya); b.tv = P(A.tv);
    b.fx = P(A.fx);
    for (int d = 0; d < 3; d++) {
      b.drva[d] = P(A.f[d][0]); b.drvb[d] = P(A.f[d][1]); b.recv[d] = P(A.f[d][2]); b.cm[d] = P(A.f[d][3]);
    }
    b.gq = P(A.gq);
    b.lud = P(A.lud); b.luf = P(A.luf); b.rpex = P(A.rp);
    b.sarea = P(A.sa); b.snor = P(A.sn); b.sst = P(A.ss2);
    b.varr = les ? reinterpret_cast<float *>(base + A.va) : nullptr;
    std::vector<double> lud(5 * b.nlu), luf(5 * b.nlu);
    const int nl = b.nlu;
    for (int d = 0; d < 3; d++) {
      const int n = b.L[d], o = b.lu_off[d];
      const int ns = b.ndf[d][0], ne = b.ndf[d][1];
      band_lu(n, h.albed[ns], h.albed[ne], kAlpha, kBeta, &lud[0 * nl + o], &lud[1 * nl + o], &lud[2 * nl + o],
              &lud[3 * nl + o], &lud[4 * nl + o]);
      band_lu(n, h.albef[ns], h.albef[ne], h.alphf, h.betf, &luf[0 * nl + o], &luf[1 * nl + o], &luf[2 * nl + o],
              &luf[3 * nl + o], &luf[4 * nl + o]);
    }
    HIPCHK(hipMemcpy(b.lud, lud.data(), sizeof(double) * lud.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.luf, luf.data(), sizeof(double) * luf.size(), hipMemcpyHostToDevice));
    {
      // every table of the block in one allocation, each at a 128-byte offset
      std::vector<double> all;
      all.reserve(col_block_doubles(b.L));
      std::vector<std::pair<int, size_t>> slot;
      for (int d = 0; d < 3; d++) {
        const int L = b.L[d], o = b.lu_off[d];
        for (int pi = 0; pi < kNColP; pi++) {
          if (!col_chunk(L, kColP[pi])) continue;
          for (int kind = 0; kind < 2; kind++) {
            const std::vector<double> &lu = kind == 0 ? lud : luf;
            std::vector<double> T = col_tables(L, kColP[pi], &lu[0 * nl + o], &lu[1 * nl + o], &lu[2 * nl + o],
                                               &lu[3 * nl + o], &lu[4 * nl + o]);
            slot.push_back({(d * 2 + kind) * kNColP + pi, all.size()});
            T.resize(col_table_doubles(L, kColP[pi]), 0.0);
            all.insert(all.end(), T.begin(), T.end());
          }
        }
      }
      double *dptr = nullptr;
      if ((rc = upload_vec(h.mem, &dptr, all))) return rc;
      for (const auto &e : slot) h.coltab[lb][e.first] = dptr + e.second;
    }
    if (!h.grid[lb].empty())
      HIPCHK(hipMemcpy(b.ss, h.grid[lb].data(), sizeof(double) * 3 * b.N, hipMemcpyHostToDevice));
    h.blk[lb] = b;
  }
  h.built = true;
  h.geom = false;
  h.sc_saved = false;
  return 0;
}

// launch tables over the global blocks gs (sorted by case)
int make_lists(tp_context *c, const std::vector<int> &gs, Mem &m, ListSet &L) {
  std::vector<Work> pts, sp, fo, lw[3], ex, gp, ga[3], gc, wl, ac, aa[3], nm, sf;
  L.wr.assign(c->linked.size() + 1, 0);
  L.wwr.assign(c->linked.size() + 1, 0);
  int slot = -1;
  for (int g : gs) {
    const Block &b = c->blocks[g];
    while (slot < b.cs) {
      ++slot;
      L.wr[slot] = (int)pts.size();
      L.wwr[slot] = (int)sf.size();
    }
    for (int a0 = 0; a0 < b.nsurf; a0 += 256) sf.push_back({g, a0, std::min(256, b.nsurf - a0), 1});
    const int CH = 256;
    for (int a0 = 0; a0 < b.N; a0 += CH) pts.push_back({g, a0, std::min(CH, b.N - a0), 0});
    for (int a0 = 0; a0 < b.nsponge; a0 += 256) sp.push_back({g, a0, std::min(256, b.nsponge - a0), 0});
    for (int a0 = 0; a0 < b.nforce; a0 += 256) fo.push_back({g, a0, std::min(256, b.nforce - a0), 0});
    for (int d = 0; d < 3; d++) {
      const int nlines = b.N / b.L[d], nl = c->nlp[d] - 1;
      for (int q = 0; q < nlines; q += nl) lw[d].push_back({g, q, std::min(nl, nlines - q), d});
    }
    const HostCase &h = c->cases[c->bsrc[g].first];
    const bool gcic = h.ip[TP_I_NEXTGCIC] == 1;
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        const int np = b.nbc[d][s], F = b.F[d];
        // eta faces in tiles of FACE_TX xi x FACE_TZ zeta (a = first xi,
        // n = first zeta), the others in runs of 256 face indices
        auto add = [&](std::vector<Work> &v) {
          if (d == 1) {
            for (int x0 = 0; x0 < b.nx; x0 += FACE_TX)
              for (int z0 = 0; z0 < b.nz; z0 += FACE_TZ) v.push_back({g, x0, z0, d | (s << 8)});
          } else {
            for (int a0 = 0; a0 < F; a0 += 256) v.push_back({g, a0, std::min(256, F - a0), d | (s << 8)});
          }
        };
        const bool gc_face = np == 10 || np == 20 || np == 25 || np == 30;
        const bool wall = np == 20 || np == 25;
        if (b.ndf[d][s]) add(ex);
        if (gc_face) { add(gp); add(ga[d]); }
        if (np == 30 || (gcic && wall)) add(gc);
        if (wall) add(wl);
        if (np == 30 || np == 35 || np == 45 || (gcic && wall)) { add(ac); add(aa[d]); }
        add(nm);
      }
  }
  while (slot < (int)c->linked.size()) {
    ++slot;
    L.wr[slot] = (int)pts.size();
    L.wwr[slot] = (int)sf.size();
  }
  int rc;
  if ((rc = upload_list(m, L.surf, sf))) return rc;
  if ((rc = upload_list(m, L.points, pts))) return rc;
  if ((rc = upload_list(m, L.sponge, sp))) return rc;
  if ((rc = upload_list(m, L.force, fo))) return rc;
  if ((rc = upload_list(m, L.exch, ex))) return rc;
  if ((rc = upload_list(m, L.gprep, gp))) return rc;
  if ((rc = upload_list(m, L.gcopy, gc))) return rc;
  if ((rc = upload_list(m, L.wall, wl))) return rc;
  if ((rc = upload_list(m, L.acap, ac))) return rc;
  if ((rc = upload_list(m, L.normals, nm))) return rc;
  for (int d = 0; d < 3; d++) {
    if ((rc = upload_list(m, L.lines[d], lw[d]))) return rc;
    if ((rc = upload_list(m, L.gapply[d], ga[d]))) return rc;
    if ((rc = upload_list(m, L.aapply[d], aa[d]))) return rc;
  }
  return 0;
}

// rebuild the linked view after a case joined or left: global block
// numbering, scalar slots (each case keeps its scalars), chunk counts,
// descriptors, coefficients and launch tables
int link(tp_context *c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->d_sc && !c->linked.empty()) {
    std::vector<double> sc(c->linked.size() * S_STRIDE);
    HIPCHK(hipMemcpy(sc.data(), c->d_sc, sizeof(double) * sc.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < c->linked.size(); i++) {
      const int ci = case_index(c, c->linked[i]);
      if (ci < 0 || !c->cases[ci].built) continue;
      std::memcpy(c->cases[ci].sc, &sc[i * S_STRIDE], sizeof(double) * S_STRIDE);
      c->cases[ci].sc_saved = true;
    }
  }
  drop_graph(c);
  c->link.release();
  c->lmem.release();
  c->d_blocks = nullptr; c->d_coef = nullptr; c->d_sc = nullptr; c->d_part = nullptr; c->d_wpart = nullptr;
  c->d_wr = nullptr; c->d_wwr = nullptr; c->d_jct = nullptr; c->d_jvel = nullptr;
  c->L = ListSet();
  c->blocks.clear(); c->bsrc.clear(); c->linked.clear();
  c->npts = 0; c->any_visc = false; c->njct = 0; c->maxk = 0;
  std::vector<CaseCoef> coefs;
  std::vector<double> sc;
  for (size_t i = 0; i < c->cases.size(); i++) {
    HostCase &h = c->cases[i];
    h.slot = -1;
    if (!h.built) continue;
    h.block0 = (int)c->blocks.size();
    h.slot = (int)c->linked.size();
    c->linked.push_back(h.id);
    coefs.push_back(h.coef);
    if (h.sc_saved) {
      sc.insert(sc.end(), h.sc, h.sc + S_STRIDE);
    } else {
      sc.insert(sc.end(), S_STRIDE, 0.0);
      sc[h.slot * S_STRIDE + S_AOA] = h.dp[TP_D_AOA0];
    }
    if (h.ip[TP_I_NVISCOUS] == 1) c->any_visc = true;
    for (int lb = 0; lb < h.nb; lb++) {
      Block b = h.blk[lb];
      b.cs = h.slot;
      for (int d = 0; d < 3; d++)
        for (int s = 0; s < 2; s++) {
          const int nbr = h.ncd[6 * lb + 2 * d + s];
          b.ncd[d][s] = nbr >= 0 ? h.block0 + nbr : -1;
        }
      c->blocks.push_back(b);
      c->bsrc.push_back({(int)i, lb});
      c->npts += b.N;
    }
  }
  const int nb = (int)c->blocks.size();
  int rc;
  // sequential kernels: NL lines of one block per 64-lane workgroup, LDS budget ~40 KB
  for (int d = 0; d < 3; d++) {
    int Lmax = 1;
    for (auto &b : c->blocks) Lmax = std::max(Lmax, b.L[d]);
    int nl = std::max(1, std::min(64, (int)(40960 / (8 * Lmax)) - 1));
    if (nl >= 16) nl = (nl / 16) * 16;
    c->nlp[d] = nl + 1;
    c->lds[d] = (size_t)Lmax * (nl + 1) * sizeof(double);
    if (c->lds[d] > 64 * 1024) return fail(TP_E_SIZE, "line too long for the LDS tile");
  }
  // chunked solver: chunks per line per block, one launch per chunk count
  choose_lines(c);
  std::vector<int> gs(nb);
  for (int g = 0; g < nb; g++) gs[g] = g;
  if ((rc = make_lists(c, gs, c->link, c->L))) return rc;
  if ((rc = make_groups(c, gs, c->lmem, c->L))) return rc;
  if ((rc = upload_vec(c->link, &c->d_blocks, c->blocks))) return rc;
  if ((rc = upload_vec(c->link, &c->d_coef, coefs))) return rc;
  if ((rc = upload_vec(c->link, &c->d_sc, sc))) return rc;
  std::vector<double> part(2 * std::max(c->L.points.n, 1), 0.0);
  if ((rc = upload_vec(c->link, &c->d_part, part))) return rc;
  if ((rc = upload_vec(c->link, &c->d_wr, c->L.wr))) return rc;
  if ((rc = upload_vec(c->link, &c->d_wwr, c->L.wwr))) return rc;
  std::vector<double> wpart(4 * std::max(c->L.surf.n, 1), 0.0);
  if ((rc = upload_vec(c->link, &c->d_wpart, wpart))) return rc;
  // junctions
  std::vector<int> jct, jvel;
  for (const HostCase &h : c->cases) {
    if (!h.built) continue;
    for (int j = 0; j < h.njct; j++) {
      for (int q = 0; q < 4; q++) {
        const int lb = h.jct[12 * j + 3 * q];
        jct.push_back(h.block0 + lb);
        jct.push_back(h.jct[12 * j + 3 * q + 1]);
        jct.push_back(h.jct[12 * j + 3 * q + 2]);
        c->maxk = std::max(c->maxk, c->blocks[h.block0 + lb].nz);
      }
      jvel.push_back(h.jvel[2 * j]);
      jvel.push_back(h.jvel[2 * j + 1]);
      c->njct++;
    }
  }
  if ((rc = upload_vec(c->link, &c->d_jct, jct))) return rc;
  if ((rc = upload_vec(c->link, &c->d_jvel, jvel))) return rc;
  return 0;
}

// the single Fourier mode of a sinusoidal transverse gust: wavevector along
// the free stream with |k| = 2 kr (reduced frequency kr = omega c / (2 |u|),
// chord 1), velocity amplitude tu |u| normal to the free stream in the x-y
// plane; slit is the table's x-period, so that slit / u_x is the gust period
void gust_mode(double tu, double kr, const double *u, double *m8, double *slit) {
  const double um = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double uxy = std::sqrt(u[0] * u[0] + u[1] * u[1]);
  const double ca = u[0] / uxy, sa = u[1] / uxy;
  m8[0] = 2 * kr * u[0] / um; m8[1] = 2 * kr * u[1] / um; m8[2] = 2 * kr * u[2] / um;
  m8[3] = -tu * um * sa; m8[4] = tu * um * ca; m8[5] = 0;
  m8[6] = 0; m8[7] = 0;
  *slit = kPi * u[0] / (kr * um);
}

}  // namespace

// ----------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------
extern "C" {

int32_t tp_inflow_modes(int32_t nmodes, uint64_t seed, double tu, double lturb, const double *uoo3,
                        double slit, double span, double *out8) {
  if (nmodes < 1 || !out8 || !uoo3 || !(lturb > 0) || !(slit > 0) || !(tu >= 0))
    return fail(TP_E_ARG, "tp_inflow_modes: bad arguments");
  const double ke = std::sqrt(kPi) * std::tgamma(5.0 / 6.0) / (std::tgamma(1.0 / 3.0) * lturb);
  const double klo = ke / 5, khi = 40 * ke, lr = std::log(khi / klo);
  const double dkx = 2 * kPi / slit, dkz = span > 0 ? 2 * kPi / span : 0.0;
  SplitMix rng{seed};
  std::vector<double> w(nmodes);
  double esum = 0;
  for (int n = 0; n < nmodes; n++) {
    double k[3], kk, sg[3];
    double kmag = 0;
    for (;;) {
      kmag = klo * std::exp(lr * rng.uniform());           // log-uniform magnitude
      const double ct = 2 * rng.uniform() - 1, st = std::sqrt(std::fmax(0.0, 1 - ct * ct));
      const double ph = 2 * kPi * rng.uniform();
      k[0] = kmag * st * std::cos(ph); k[1] = kmag * ct; k[2] = kmag * st * std::sin(ph);
      k[0] = std::nearbyint(k[0] / dkx) * dkx;              // periodic along x (period slit)
      if (dkz > 0) k[2] = std::nearbyint(k[2] / dkz) * dkz; // and along the span
      kk = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
      if (kk <= 0) continue;
      // direction: a random vector with its component along k removed
      double r[3] = {2 * rng.uniform() - 1, 2 * rng.uniform() - 1, 2 * rng.uniform() - 1};
      const double rk = (r[0] * k[0] + r[1] * k[1] + r[2] * k[2]) / (kk * kk);
      for (int d = 0; d < 3; d++) sg[d] = r[d] - rk * k[d];
      const double sn = std::sqrt(sg[0] * sg[0] + sg[1] * sg[1] + sg[2] * sg[2]);
      if (sn < 1e-3) continue;
      for (int d = 0; d < 3; d++) sg[d] /= sn;
      break;
    }
    // von Karman energy spectrum at |k| times the sampling density of kmag
    const double x = kk / ke;
    const double e = x * x * x * x / std::pow(1 + x * x, 17.0 / 6.0);
    w[n] = e * kmag;
    esum += w[n];
    double *o = out8 + 8 * (size_t)n;
    o[0] = k[0]; o[1] = k[1]; o[2] = k[2];
    o[3] = sg[0]; o[4] = sg[1]; o[5] = sg[2];
    o[6] = 2 * kPi * rng.uniform();
    o[7] = 0;
  }
  // amplitudes: <u'.u'> = sum_n |a_n|^2 / 2 = 3 (tu |uoo|)^2, i.e. rms tu |uoo|
  // in each component
  const double u2 = tu * tu * (uoo3[0] * uoo3[0] + uoo3[1] * uoo3[1] + uoo3[2] * uoo3[2]);
  for (int n = 0; n < nmodes; n++) {
    const double amp = std::sqrt(6 * u2 * w[n] / esum);
    double *o = out8 + 8 * (size_t)n;
    for (int d = 0; d < 3; d++) o[3 + d] *= amp;
  }
  return 0;
}


int32_t tp_eddy_params(int32_t nits, const double *randnum, const double *deck, double *out10) {
  if (nits < 1 || !randnum || !deck || !out10) return fail(TP_E_ARG, "tp_eddy_params: bad arguments");
  // deck: span, slit, gaus, cutlb, domlen, szth1, then tl0, tl1, tlw of the
  // two eddy types, then tam(1:3) of each
  const double span = deck[0], slit = deck[1], gaus = deck[2], cutlb = deck[3], domlen = deck[4],
               szth1 = deck[5];
  const double tl0[2] = {deck[6], deck[9]}, tl1[2] = {deck[7], deck[10]}, tlw[2] = {deck[8], deck[11]};
  const double tam[2][3] = {{deck[12], deck[13], deck[14]}, {deck[15], deck[16], deck[17]}};
  const double tlb = std::max(std::max(tl0[0], tl0[1]), std::max(tl1[0], tl1[1]));
  // randnum.dat in file order: ran(:,1), ait(:,1), ran(:,2), ait(:,2),
  // ran(:,3), ait(:,3), yit(:), zit(:), src/SingleTurbulence3D.f90:88-93
  const size_t n = (size_t)nits;
  std::vector<double> ran(3 * n), ait(3 * n), sit(3 * n), xit(n), yit(n), zit(n);
  std::vector<int> mxc(n);
  for (int m = 0; m < 3; m++)
    for (size_t q = 0; q < n; q++) {
      ran[m * n + q] = randnum[(2 * m) * n + q];
      ait[m * n + q] = randnum[(2 * m + 1) * n + q];
    }
  for (size_t q = 0; q < n; q++) { yit[q] = randnum[6 * n + q]; zit[q] = randnum[7 * n + q]; }
  // eddy type, sizes and amplitudes (:95-102)
  for (size_t q = 0; q < n; q++) mxc[q] = ran[q] <= gaus ? 0 : 1;
  double res = std::pow(2.0 * (2.0 - cutlb) * tlb * span * slit / nits, 1.0 / 3.0);
  for (int m = 0; m < 3; m++)
    for (size_t q = 0; q < n; q++) {
      const int t = mxc[q];
      double &r = ran[m * n + q];
      r = tl0[t] + (tl1[t] - tl0[t]) * std::pow(r, tlw[t]);
      sit[m * n + q] = 1.0 / std::pow(r, 2.0);
      ait[m * n + q] = res * tam[t][m] * (2.0 * ait[m * n + q] - 1.0) * std::pow(sit[m * n + q], 1.5);
    }
  // centres: a train along x, each eddy spaced by the sizes of its
  // neighbours, over one slit upstream of the inflow sponge edge (:103-107)
  double sr = 0;
  for (size_t q = 0; q < n; q++) sr = sr + ran[q];
  const double denxit = slit / (2.0 * sr);
  xit[0] = -domlen + 0.5 * (szth1 + slit) - denxit * ran[0];
  for (size_t q = 1; q < n; q++) xit[q] = xit[q - 1] - denxit * (ran[q - 1] + ran[q]);
  for (size_t q = 0; q < n; q++) {
    res = 3.0 * span - 2.0 * std::max(std::max(ran[q], ran[n + q]), ran[2 * n + q]);
    yit[q] = (1.0 - cutlb) * ran[n + q] * (2.0 * yit[q] - 1.0);
    zit[q] = std::min(span, res) * (zit[q] - 0.5);
  }
  for (size_t q = 0; q < n; q++) {
    double *o = out10 + 10 * q;
    o[0] = xit[q]; o[1] = yit[q]; o[2] = zit[q];
    for (int m = 0; m < 3; m++) { o[3 + m] = sit[m * n + q]; o[6 + m] = ait[m * n + q]; }
    o[9] = mxc[q];
  }
  return 0;
}

int32_t tp_set_inflow_eddies(tp_context *c, int32_t cs, int32_t nits, const double *eddies10) {
  if (!c) return fail(TP_E_ARG, "null context");
  const int ci = case_index(c, cs);
  if (ci < 0 || nits < 1 || !eddies10) return fail(TP_E_ARG, "tp_set_inflow_eddies: bad arguments");
  HostCase &h = c->cases[ci];
  if (h.coef.inflow != 3) return fail(TP_E_STATE, "tp_set_inflow_eddies: the case's TP_I_INFLOW is not 3");
  if (h.geom) return fail(TP_E_STATE, "tp_set_inflow_eddies: the case's geometry is set up already");
  h.eddies.assign(eddies10, eddies10 + 10 * (size_t)nits);
  return 0;
}

const char *tp_last_error(void) { return g_err.c_str(); }

namespace {
int scratch_address(char **out);

// the HIP runtime the process has loaded must be no older than the one the
// library is built against: a process can hold another runtime first (an
// application's bundled copy answers to the same soname), and this library
// is not run on an older one
int runtime_check() {
  int v = 0;
  if (hipRuntimeGetVersion(&v) != hipSuccess) {
    (void)hipGetLastError();
    return fail(TP_E_HIP, "hipRuntimeGetVersion failed");
  }
  if (v / 100000 < HIP_VERSION / 100000)
    return fail(TP_E_HIP, "the process has loaded HIP runtime " + std::to_string(v / 10000000) + "." +
                              std::to_string(v / 100000 % 100) + ", older than the " +
                              std::to_string(HIP_VERSION_MAJOR) + "." + std::to_string(HIP_VERSION_MINOR) +
                              " this library is built for: load the library before any other HIP user "
                              "(in Python, tandemproject_amd.api.load() before `import torch`)");
  return 0;
}
}  // namespace

int32_t tp_create(tp_context **out) {
  if (!out) return fail(TP_E_ARG, "null output pointer");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(TP_E_HIP, "no HIP device available: the MI355X time stepper has no CPU path");
  if (int rc = runtime_check()) return rc;
  tp_context *c = new tp_context();
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return fail(TP_E_HIP, "hipGetDevice failed"); }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(TP_E_HIP, "hipStreamCreate failed");
  }
  // the work area of tp_field_deriv_noalloc on this device, looked up (and
  // placed by the runtime) now rather than at that entry point's first call
  char *ws = nullptr;
  if (int rc = scratch_address(&ws)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int32_t tp_destroy(tp_context *c) {
  if (!c) return 0;
  (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (int k = 0; k < K_NCLASS; k++)
    for (auto &r : c->pend[k]) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  for (auto &h : c->cases) h.mem.release();
  c->link.release();
  c->lmem.release();
  (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int32_t tp_check_case(int32_t nb, const int32_t *dims, const int32_t *nbc, const int32_t *ncd, int32_t njct,
                      const int32_t *jct, const int32_t *jvel, const int32_t *iparams) {
  if (nb <= 0 || !dims || !nbc || !ncd || !iparams) return fail(TP_E_ARG, "tp_check_case: missing arrays");
  if (njct < 0 || (njct > 0 && (!jct || !jvel))) return fail(TP_E_ARG, "junction arrays missing");
  for (int j = 0; j < njct; j++)
    for (int q = 0; q < 4; q++) {
      const int lb = jct[12 * j + 3 * q];
      if (lb < 0 || lb >= nb) return fail(TP_E_LAYOUT, "junction member outside its case");
      if (dims[3 * lb + 2] != dims[3 * jct[12 * j] + 2])
        return fail(TP_E_LAYOUT, "junction members differ in zeta size");
    }
  // block size: the line solves address a field with 32-bit byte offsets
  // from its base, so a block holds fewer than 2^29 points
  for (int b = 0; b < nb; b++) {
    int64_t np = 1;
    for (int d = 0; d < 3; d++) {
      if (dims[3 * b + d] < 1) return fail(TP_E_LAYOUT, "block " + std::to_string(b) + ": fewer than 2 points in a direction");
      np *= (int64_t)dims[3 * b + d] + 1;
    }
    if (8 * np >= (int64_t)1 << 32)
      return fail(TP_E_SIZE, "block " + std::to_string(b) + " has " + std::to_string(np) +
                                 " points; a block holds fewer than 2^29 (536870912)");
  }
  // validate the topology: line lengths the stencils and closures need, and
  // neighbours that exist and match
  for (int b = 0; b < nb; b++) {
    for (int d = 0; d < 3; d++) {
      const int L = dims[3 * b + d] + 1;
      int need = 7;
      // both ends characteristic: their bands of MBCI + 1 points, updated by
      // one launch, must not share a point
      const auto gc = [](int np) { return np == 10 || np == 20 || np == 25 || np == 30; };
      if (gc(nbc[6 * b + 2 * d]) && gc(nbc[6 * b + 2 * d + 1])) need = 2 * (MBCI + 1);
      for (int s = 0; s < 2; s++) {
        const int np = nbc[6 * b + 2 * d + s];
        if (np != 10 && np != 20 && np != 25 && np != 30 && np != 35 && np != 40 && np != 45)
          return fail(TP_E_LAYOUT, "block " + std::to_string(b) + ": unknown boundary code " + std::to_string(np));
        if (np == 40) need = std::max(need, LMD + 1);
        if (np == 35 || np == 45) need = std::max(need, LMD + 2);
        const int nbr = ncd[6 * b + 2 * d + s];
        const bool exchanges = np == 30 || np == 35 || np == 40 || np == 45 ||
                               ((np == 20 || np == 25) && iparams[TP_I_NEXTGCIC] == 1);
        if (exchanges) {
          if (nbr < 0 || nbr >= nb)
            return fail(TP_E_LAYOUT, "block " + std::to_string(b) + " face " + std::to_string(2 * d + s) +
                                         " needs a neighbour");
          for (int e = 0; e < 3; e++)
            if (e != d && dims[3 * nbr + e] != dims[3 * b + e])
              return fail(TP_E_LAYOUT, "block " + std::to_string(b) + " and neighbour " + std::to_string(nbr) +
                                           " faces differ in size");
        }
      }
      if (L < need)
        return fail(TP_E_LAYOUT, "block " + std::to_string(b) + " direction " + std::to_string(d + 1) + " has " +
                                     std::to_string(L) + " points; its faces need >= " + std::to_string(need));
    }
  }
  if (iparams[TP_I_NKRK] < 1) return fail(TP_E_ARG, "nkrk must be >= 1");
  return 0;
}

int32_t tp_add_case(tp_context *c, int32_t nb, const int32_t *dims, const int32_t *nbc, const int32_t *ncd,
                    const int32_t *nsz, int32_t njct, const int32_t *jct, const int32_t *jvel,
                    const int32_t *iparams, const double *dparams, int32_t *case_id) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (nb <= 0 || !dims || !nbc || !ncd || !nsz || !iparams || !dparams)
    return fail(TP_E_ARG, "tp_add_case: missing arrays");
  if ((int)c->cases.size() >= MAX_CASE) return fail(TP_E_SIZE, "too many cases");
  if (int rc = tp_check_case(nb, dims, nbc, ncd, njct, jct, jvel, iparams)) return rc;
  HostCase h;
  h.nb = nb;
  h.njct = njct;
  h.dims.assign(dims, dims + 3 * nb);
  h.nbc.assign(nbc, nbc + 6 * nb);
  h.ncd.assign(ncd, ncd + 6 * nb);
  h.nsz.assign(nsz, nsz + 6 * nb);
  if (njct > 0) {
    if (!jct || !jvel) return fail(TP_E_ARG, "junction arrays missing");
    h.jct.assign(jct, jct + 12 * njct);
    h.jvel.assign(jvel, jvel + 2 * njct);
  }
  std::memcpy(h.ip, iparams, sizeof(h.ip));
  std::memcpy(h.dp, dparams, sizeof(h.dp));
  if (c->nkrk && c->nkrk != iparams[TP_I_NKRK])
    return fail(TP_E_ARG, "all cases of a context must use the same number of RK stages");
  if (iparams[TP_I_NKRK] < 1) return fail(TP_E_ARG, "nkrk must be >= 1");
  // coefficients, src/Main3D.f90:60-63, 71-73, 295-312
  CaseCoef &k = h.coef;
  std::memset(&k, 0, sizeof(k));
  std::memcpy(k.abc, kAbc, sizeof(kAbc));
  const double fltk = kPi * dparams[TP_D_FLTK], fltkbc = kPi * dparams[TP_D_FLTKBC];
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 5; q++) {
      h.albed[0][r][q] = kBnd[r][q];
      const double inner[3][5] = {{0, 0, 1, kAlpha, kBeta}, {0, kAlpha, 1, kAlpha, kBeta},
                                  {kBeta, kAlpha, 1, kAlpha, kBeta}};
      h.albed[1][r][q] = inner[r][q];
    }
  filter_boundary(fltk, fltkbc, h.albef[0], k.fam, k.fbm, k.fcm);
  FiltCoef fi = filter_coef(fltk, 0.5);
  k.fa = fi.a; k.fb = fi.b; k.fc = fi.c;
  h.alphf = fi.alpha; h.betf = fi.beta;
  {
    const double inner[3][5] = {{0, 0, 1, fi.alpha, fi.beta}, {0, fi.alpha, 1, fi.alpha, fi.beta},
                                {fi.beta, fi.alpha, 1, fi.alpha, fi.beta}};
    std::memcpy(h.albef[1], inner, sizeof(inner));
  }
  std::memcpy(k.flhs, h.albef, sizeof(k.flhs));
  k.fal = fi.alpha; k.fbe = fi.beta;
  closure_rows(fltk, k);
  h.cbca_ra0 = gcbc_weights(k.sbci);
  k.fex[0] = 45 / 30.0; k.fex[1] = -9 / 30.0; k.fex[2] = 1 / 30.0;
  const double l2 = 8 * std::log(2.0);
  k.ffx[0] = 4 / l2; k.ffx[1] = 6 / l2; k.ffx[2] = 7 / l2;
  const double gam = 1.4, poo = 1 / gam, aoo = std::sqrt(gam * poo / 1.0);
  k.srefoo = 111 / dparams[TP_D_TEMPOO];
  k.srefp1dre = (k.srefoo + 1) / dparams[TP_D_REOO];
  k.cfl = dparams[TP_D_CFL]; k.dto = dparams[TP_D_DTO]; k.wtemp = dparams[TP_D_WTEMP];
  k.smago1 = dparams[TP_D_SMAGO1]; k.smago2 = dparams[TP_D_SMAGO2];
  k.timf = dparams[TP_D_TIMF]; k.tp = dparams[TP_D_TP]; k.tps = dparams[TP_D_TPS];
  k.aoa0 = dparams[TP_D_AOA0]; k.aoa1 = dparams[TP_D_AOA1];
  const double m1 = dparams[TP_D_AMACH1], m2 = dparams[TP_D_AMACH2], m3 = dparams[TP_D_AMACH3];
  k.amachoo = std::sqrt(m1 * m1 + m2 * m2 + m3 * m3);
  k.uoo[0] = m1 * aoo; k.uoo[1] = m2 * aoo; k.uoo[2] = m3 * aoo;
  k.nkrk = iparams[TP_I_NKRK]; k.nviscous = iparams[TP_I_NVISCOUS]; k.nsmf = iparams[TP_I_NSMF];
  k.les = iparams[TP_I_LES]; k.nextgcic = iparams[TP_I_NEXTGCIC];
  k.span = dparams[TP_D_SPAN];
  k.inflow = iparams[TP_I_INFLOW];
  if (k.inflow < 0 || k.inflow > 3) return fail(TP_E_ARG, "TP_I_INFLOW must be 0, 1, 2 or 3");
  k.forcing = iparams[TP_I_FORCING];
  if (k.forcing != 0 && k.forcing != 1) return fail(TP_E_ARG, "TP_I_FORCING must be 0 or 1");
  if (k.forcing == 1 && !(dparams[TP_D_RFOR] > 0 && dparams[TP_D_SPAN] > 0))
    return fail(TP_E_ARG, "forcing needs TP_D_RFOR > 0 and TP_D_SPAN > 0");
  k.amfor = dparams[TP_D_AMFOR]; k.tsfor = dparams[TP_D_TSFOR]; k.tefor = dparams[TP_D_TEFOR];
  if (k.inflow) {
    k.ntz = iparams[TP_I_NTZ];
    if (k.ntz < 4) return fail(TP_E_ARG, "inflow table needs TP_I_NTZ >= 4 time intervals");
    if (!(k.uoo[0] > 0)) return fail(TP_E_ARG, "inflow needs u_inf with a positive x component");
    double slit = dparams[TP_D_SLIT];
    if (k.inflow == 1) {
      if (iparams[TP_I_NMODES] < 1) return fail(TP_E_ARG, "inflow needs TP_I_NMODES >= 1");
      if (!(slit > 0) || !(dparams[TP_D_LTURB] > 0))
        return fail(TP_E_ARG, "turbulent inflow needs TP_D_SLIT > 0 and TP_D_LTURB > 0");
    } else if (k.inflow == 3) {
      if (!(slit > 0)) return fail(TP_E_ARG, "eddy inflow needs TP_D_SLIT > 0");
    } else {
      if (!(dparams[TP_D_GUSTK] > 0)) return fail(TP_E_ARG, "gust inflow needs TP_D_GUSTK > 0");
      double m8[8];
      gust_mode(dparams[TP_D_TU], dparams[TP_D_GUSTK], k.uoo, m8, &slit);
    }
    k.tgusto = dparams[TP_D_TGUSTO];
    k.period = slit / k.uoo[0];
    k.fctr = slit / (k.ntz * k.uoo[0]);
  }
  h.id = c->next_id++;
  c->nkrk = iparams[TP_I_NKRK];
  if (case_id) *case_id = h.id;
  c->cases.push_back(std::move(h));
  return 0;
}

int32_t tp_finalize(tp_context *c) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (c->finalized) return fail(TP_E_STATE, "already finalized");
  HIPCHK(hipSetDevice(c->device));
  int rc;
  for (auto &h : c->cases)
    if ((rc = build_case(h))) return rc;
  if ((rc = link(c))) return rc;
  c->finalized = true;
  return 0;
}

int32_t tp_set_grid(tp_context *c, int32_t cs, int32_t lb, const double *xyz) {
  if (!c) return fail(TP_E_ARG, "null context");
  const int ci = case_index(c, cs);
  if (ci < 0 || !xyz || lb < 0 || lb >= c->cases[ci].nb) return fail(TP_E_ARG, "tp_set_grid: bad case/block");
  HostCase &h = c->cases[ci];
  const size_t N = (size_t)(h.dims[3 * lb] + 1) * (h.dims[3 * lb + 1] + 1) * (h.dims[3 * lb + 2] + 1);
  h.grid.resize(h.nb);
  h.grid[lb].assign(xyz, xyz + 3 * N);
  h.geom = false;
  if (h.built) {
    HIPCHK(hipMemcpyAsync(h.blk[lb].ss, xyz, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int32_t tp_setup_geometry(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  // the built cases whose metrics are not set yet, and their global blocks
  std::vector<int> todo, gs;
  for (size_t i = 0; i < c->cases.size(); i++) {
    const HostCase &h = c->cases[i];
    if (!h.built || h.geom) continue;
    for (int lb = 0; lb < h.nb; lb++)
      if ((int)h.grid.size() <= lb || h.grid[lb].empty())
        return fail(TP_E_STATE, "grid of case " + std::to_string(h.id) + " block " + std::to_string(lb) +
                                    " not set");
    if (h.coef.inflow == 3 && h.eddies.empty())
      return fail(TP_E_STATE, "case " + std::to_string(h.id) + ": eddy inflow needs tp_set_inflow_eddies");
    todo.push_back((int)i);
    for (int lb = 0; lb < h.nb; lb++) gs.push_back(h.block0 + lb);
  }
  if (todo.empty()) return 0;
  // coordinate derivatives, src/Main3D.f90:365-375 (periodic jump correction
  // on), metrics and face normals, over these cases' blocks only: the other
  // cases' fields are left as they are
  {
    Mem tmp;
    ListSet sub;
    int rc = make_lists(c, gs, tmp, sub);
    if (!rc) rc = make_groups(c, gs, tmp, sub);
    if (rc) { tmp.release(); return rc; }
    std::swap(c->L, sub);
    const int dst[3] = {F_QO, F_QA, F_DE};
    for (int m = 0; m < 3; m++) {
      run_exchange(c, 0, 1, 1, F_SS, m);
      run_line(c, 0, 2, F_SS, m, dst[m], 2, 0, 0);
      run_line(c, 0, 1, F_SS, m, dst[m], 1, 0, 0);
      run_line(c, 0, 0, F_SS, m, dst[m], 0, 0, 0);
    }
    run_points(c, PK_METRICS, base_args(c), K_MISC, 0.0);
    run_faces(c, c->L.normals, FK_NORMALS, base_args(c), K_MISC);
    run_faces(c, c->L.surf, FK_SURF_SETUP, base_args(c), K_MISC);
    std::swap(c->L, sub);
    const hipError_t e = hipStreamSynchronize(c->stream);
    tmp.release();
    if (e != hipSuccess) return fail(TP_E_HIP, std::string("metrics: ") + hipGetErrorString(e));
  }
  // wall-slot flags (SingleTurbulence3D.f90:447-464), sponge coefficients
  // (SingleTurbulence3D.f90:209-249) and the inflow table, on the host from
  // the grid and Jacobian
  for (int ci : todo) {
    HostCase &h = c->cases[ci];
    for (int lb = 0; lb < h.nb; lb++) {
      Block &b = h.blk[lb];
      const double *x = h.grid[lb].data(), *y = x + b.N, *z = y + b.N;
      std::vector<double> yaco(b.N);
      HIPCHK(hipMemcpy(yaco.data(), b.yaco, sizeof(double) * b.N, hipMemcpyDeviceToHost));
      if (b.varr) {
        // Smagorinsky volume term (-1/yaco)^1.5 in single precision, as the
        // reference keeps it (real(4) varr, src/MainVar3D.f90:105; Main3D.f90:664)
        std::vector<float> va(b.N);
        for (int l = 0; l < b.N; l++) va[l] = (float)std::pow(-1 / yaco[l], 1.5);
        HIPCHK(hipMemcpy(b.varr, va.data(), sizeof(float) * b.N, hipMemcpyHostToDevice));
      }
      // wall flags in the reference's face order
      {
        int64_t cnt;
        array_ptr(b, TP_RPEX, &cnt);
        std::vector<double> rp(std::max<int64_t>(cnt, 1), 0.0);
        double state = h.cbca_ra0;
        const double wlea = h.dp[TP_D_WLEA], wlew = h.dp[TP_D_WLEW], span = h.dp[TP_D_SPAN];
        int lp = 0;
        for (int d = 0; d < 3; d++)
          for (int s = 0; s < 2; s++) {
            const int np = b.nbc[d][s];
            if (!(np == 10 || np == 20 || np == 25 || np == 30)) continue;
            const int fd = d == 0 ? b.ny : d == 1 ? b.nz : b.nx, fo = b.F[d] / fd;
            for (int kk = 0; kk < fo; kk++)
              for (int jj = 0; jj < fd; jj++) {
                if (!(np == 20 || np == 25)) continue;
                const int l = indx3(b, s * (b.L[d] - 1), jj, kk, d);
                double flag = 0;
                if (x[l] < state) {
                  const double slitw = 0.4 * wlea;
                  state = -0.5 + wlea + wlea;
                  const double r1 = -0.5 * span + 0.25 * wlew - 0.5 * slitw, r2 = r1 + slitw;
                  const int ie = (int)((span + 1.0e-6) / wlew) - 1;
                  for (int ii = 0; ii <= ie; ii++) {
                    const double r3 = ii * wlew, res = (z[l] - (r1 + r3)) * (z[l] - (r2 + r3));
                    if (res < 0 && h.ip[TP_I_NEXTRABC] == 1) flag = 1;
                  }
                }
                rp[lp++] = flag;
              }
          }
        HIPCHK(hipMemcpy(b.rpex, rp.data(), sizeof(double) * rp.size(), hipMemcpyHostToDevice));
      }
      // sponge
      std::vector<int> lst;
      std::vector<double> sgv, lmv, as, bs;
      sponge_select(h, lb, b, x, y, z, lst, sgv, lmv);
      for (size_t q = 0; q < lst.size(); q++) {
        const double res = 1 / yaco[lst[q]], sg = sgv[q], lm = lmv[q];
        as.push_back(res * sg * lm);
        bs.push_back(res * (1 / (1.4 - 1)) * sg);
      }
      int rc;
      if (!lst.empty()) {
        if ((rc = upload_vec(h.mem, &b.lcsz, lst))) return rc;
        if ((rc = upload_vec(h.mem, &b.asz, as))) return rc;
        if ((rc = upload_vec(h.mem, &b.bsz, bs))) return rc;
      }
      b.nsponge = (int)lst.size();
      // forcing coefficients (forceup, src/rpt.f90:741-751): three spanwise
      // modes with a Gaussian weight; as in the reference, ra1 holds 2 pi /
      // span only for the first point, the loop overwriting it with each
      // point's weight, which the next point's modes take as wavenumber
      {
        std::vector<int> fl;
        std::vector<double> fr;
        force_select(h, b, x, y, fl, fr);
        b.nforce = (int)fl.size();
        if (b.nforce) {
          const double rf = h.dp[TP_D_RFOR];
          const double ra0 = -std::log(0.0001) / (rf * rf);
          double ra1 = 2 * kPi / h.dp[TP_D_SPAN];
          const size_t nf = fl.size();
          std::vector<double> xa(3 * nf);
          for (size_t q = 0; q < nf; q++) {
            const int l = fl[q];
            const double zz = z[l];
            const double b1 = std::cos(ra1 * zz), b2 = std::cos(3 * ra1 * zz), b3 = std::cos(4 * ra1 * zz);
            ra1 = 0.5 * std::exp(-ra0 * fr[q]) / yaco[l];
            xa[q] = ra1 * b1; xa[nf + q] = ra1 * b2; xa[2 * nf + q] = ra1 * b3;
          }
          if ((rc = upload_vec(h.mem, &b.lcfor, fl))) return rc;
          if ((rc = upload_vec(h.mem, &b.xafor, xa))) return rc;
        }
      }
      if (h.coef.inflow) {
        std::vector<int> gix, pts;
        std::vector<double> atz, pos;
        inflow_select(h, x, y, lst, gix, pts, atz);
        b.ngust = (int)pts.size();
        if ((rc = upload_vec(h.mem, &b.gix, gix))) return rc;
        if (b.ngust > 0) {
          pos.resize(3 * (size_t)b.ngust);
          for (int q = 0; q < b.ngust; q++) {
            pos[q] = x[pts[q]]; pos[b.ngust + q] = y[pts[q]]; pos[2 * (size_t)b.ngust + q] = z[pts[q]];
          }
          if ((rc = upload_vec(h.mem, &b.atz, atz))) return rc;
          if (h.coef.inflow == 3) {
            // the reference's synthetic eddies, src/SingleTurbulence3D.f90:262-298
            if ((rc = dalloc(h.mem, (void **)&b.vito, sizeof(double) * 3 * (h.coef.ntz + 1) * (size_t)b.ngust)))
              return rc;
            Mem tmp;
            double *ded = nullptr, *dpos = nullptr;
            if ((rc = upload_vec(tmp, &ded, h.eddies)) || (rc = upload_vec(tmp, &dpos, pos))) {
              tmp.release();
              return rc;
            }
            launch_eddy_table(ded, (int)(h.eddies.size() / 10), dpos, b.ngust, h.coef.ntz, h.coef.fctr, h.coef.uoo,
                              h.dp[TP_D_SLIT], h.dp[TP_D_SPAN], h.dp[TP_D_CFIT], b.vito, c->stream);
            const hipError_t e = hipStreamSynchronize(c->stream);
            tmp.release();
            if (e != hipSuccess) return fail(TP_E_HIP, std::string("eddy table: ") + hipGetErrorString(e));
            HIPCHK(hipGetLastError());
            continue;   // what follows builds the Fourier-mode table of inflow 1 and 2
          }
          const int nm = inflow_nmodes(h);
          std::vector<double> modes(8 * (size_t)nm);
          if (h.coef.inflow == 2) {
            double slit;
            gust_mode(h.dp[TP_D_TU], h.dp[TP_D_GUSTK], h.coef.uoo, modes.data(), &slit);
          } else if ((rc = tp_inflow_modes(nm, (uint64_t)(uint32_t)h.ip[TP_I_SEED], h.dp[TP_D_TU],
                                           h.dp[TP_D_LTURB], h.coef.uoo, h.dp[TP_D_SLIT], h.dp[TP_D_SPAN],
                                           modes.data()))) {
            return rc;
          }
          if ((rc = dalloc(h.mem, (void **)&b.vito, sizeof(double) * 3 * (h.coef.ntz + 1) * (size_t)b.ngust)))
            return rc;
          // per mode the cosine and sine of its phase step between table
          // times, fctr k.u: the table kernel advances each mode by rotation
          std::vector<double> rot(2 * (size_t)nm);
          for (int n = 0; n < nm; n++) {
            const double *m = &modes[8 * (size_t)n];
            const double st =
                h.coef.fctr * (m[0] * h.coef.uoo[0] + m[1] * h.coef.uoo[1] + m[2] * h.coef.uoo[2]);
            rot[2 * (size_t)n] = std::cos(st);
            rot[2 * (size_t)n + 1] = std::sin(st);
          }
          Mem tmp;
          double *dmodes = nullptr, *dpos = nullptr, *drot = nullptr;
          if ((rc = upload_vec(tmp, &dmodes, modes)) || (rc = upload_vec(tmp, &dpos, pos)) ||
              (rc = upload_vec(tmp, &drot, rot))) {
            tmp.release();
            return rc;
          }
          launch_inflow_table(dmodes, drot, nm, dpos, b.ngust, h.coef.ntz, h.coef.fctr, h.coef.uoo, b.vito,
                              c->stream);
          const hipError_t e = hipStreamSynchronize(c->stream);
          tmp.release();
          if (e != hipSuccess) return fail(TP_E_HIP, std::string("inflow table: ") + hipGetErrorString(e));
          HIPCHK(hipGetLastError());
        }
      }
    }
    h.geom = true;
  }
  return link(c);
}

int32_t tp_initialo(tp_context *c, int32_t cs) {
  if (int rc = check_ready(c)) return rc;
  const HostCase *h = linked_case(c, cs);
  if (!h) return fail(TP_E_ARG, "bad case");
  StageArgs a = base_args(c);
  a.work = c->L.points.d + c->L.wr[h->slot];
  const int n = c->L.wr[h->slot + 1] - c->L.wr[h->slot];
  if (n) launch_point(PK_INIT, n, c->stream, a);
  double s[S_NSCALAR] = {0};
  s[S_AOA] = h->dp[TP_D_AOA0];
  HIPCHK(hipMemcpyAsync(c->d_sc + h->slot * S_STRIDE, s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int64_t tp_array_count(tp_context *c, int32_t cs, int32_t lb, int32_t id) {
  if (check_ready(c)) return -1;
  const int g = block_index(c, cs, lb);
  if (g < 0) return -1;
  int64_t n = 0;
  array_ptr(c->blocks[g], id, &n);
  return n;
}

int32_t tp_upload(tp_context *c, int32_t cs, int32_t lb, int32_t id, const double *src, int64_t count) {
  if (int rc = check_ready(c)) return rc;
  const int g = block_index(c, cs, lb);
  if (g < 0 || !src) return fail(TP_E_ARG, "tp_upload: bad case/block");
  int64_t n = 0;
  double *p = array_ptr(c->blocks[g], id, &n);
  if (!p) return fail(TP_E_ARG, "tp_upload: unknown array id " + std::to_string(id));
  if (count != n) return fail(TP_E_SIZE, "tp_upload: expected " + std::to_string(n) + " values");
  HIPCHK(hipMemcpyAsync(p, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int32_t tp_download(tp_context *c, int32_t cs, int32_t lb, int32_t id, double *dst, int64_t count) {
  if (int rc = check_ready(c)) return rc;
  const int g = block_index(c, cs, lb);
  if (g < 0 || !dst) return fail(TP_E_ARG, "tp_download: bad case/block");
  int64_t n = 0;
  double *p = array_ptr(c->blocks[g], id, &n);
  if (!p) return fail(TP_E_ARG, "tp_download: unknown array id " + std::to_string(id));
  if (count != n) return fail(TP_E_SIZE, "tp_download: expected " + std::to_string(n) + " values");
  HIPCHK(hipMemcpyAsync(dst, p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int32_t tp_get_scalars(tp_context *c, int32_t cs, double *s16) {
  if (int rc = check_ready(c)) return rc;
  const HostCase *h = linked_case(c, cs);
  if (!h || !s16) return fail(TP_E_ARG, "bad case");
  HIPCHK(hipMemcpyAsync(s16, c->d_sc + h->slot * S_STRIDE, sizeof(double) * S_NSCALAR, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int32_t tp_set_scalars(tp_context *c, int32_t cs, const double *s16) {
  if (int rc = check_ready(c)) return rc;
  const HostCase *h = linked_case(c, cs);
  if (!h || !s16) return fail(TP_E_ARG, "bad case");
  HIPCHK(hipMemcpyAsync(c->d_sc + h->slot * S_STRIDE, s16, sizeof(double) * S_NSCALAR, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int32_t tp_mpigo(tp_context *c, int32_t nt, int32_t nrt, int32_t n45) {
  if (int rc = check_ready(c)) return rc;
  if (nt < 0 || nt > 1 || nrt < 0 || nrt > 1 || n45 < 0 || n45 > 1) return fail(TP_E_ARG, "tp_mpigo: bad selector");
  run_exchange(c, nt, nrt, n45, F_RR, 0);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t tp_deriv(tp_context *c, int32_t nn, int32_t nz, int32_t m) {
  if (int rc = check_ready(c)) return rc;
  if (nn < 1 || nn > 3 || nz < 1 || nz > 3 || m < 0 || m > 5) return fail(TP_E_ARG, "tp_deriv: bad selector");
  run_line(c, 0, nn - 1, F_RR, nz - 1, F_RR, nn - 1, m, 0);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t tp_filte(tp_context *c, int32_t nn, int32_t nz) {
  if (int rc = check_ready(c)) return rc;
  if (nn < 1 || nn > 3 || nz < 1 || nz > 3) return fail(TP_E_ARG, "tp_filte: bad selector");
  run_line(c, 1, nn - 1, F_RR, nz - 1, F_RR, nz - 1, 0, 0);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t tp_filter_step(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  enqueue_filter_step(c);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t tp_stage(tp_context *c, int32_t nk) {
  if (int rc = check_ready(c)) return rc;
  if (c->linked.empty()) return 0;
  if (nk < 1 || nk > c->nkrk) return fail(TP_E_ARG, "tp_stage: bad stage");
  enqueue_stage(c, nk);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- the sections of a stage as drop-ins (all sub-blocks of all cases) ----
namespace {
int stage_ok(tp_context *c, int nk) {
  if (int rc = check_ready(c)) return rc;
  if (nk < 1 || nk > std::max(c->nkrk, 1)) return fail(TP_E_ARG, "bad stage " + std::to_string(nk));
  return 0;
}
int launched() {
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace

int32_t tp_primitives(tp_context *c, int32_t nk) {
  if (int rc = stage_ok(c, nk)) return rc;
  if (!c->linked.empty()) enqueue_begin(c, nk);
  return launched();
}

int32_t tp_viscous(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  if (c->linked.empty() || !c->any_visc) return 0;
  enqueue_gradients(c);
  run_points(c, PK_STRESS, base_args(c), K_FLUX, 8.0 * (27 + 9));
  return launched();
}

int32_t tp_flux_divergence(tp_context *c, int32_t nk) {
  if (int rc = stage_ok(c, nk)) return rc;
  if (c->linked.empty()) return 0;
  enqueue_surface(c, nk);
  StageArgs f = base_args(c);
  f.nk = nk;
  run_points(c, PK_FLUX_TV, f, K_FLUX, 8.0 * (c->any_visc ? 27 + 15 : 14 + 15));
  enqueue_flux_lines(c);
  return launched();
}

int32_t tp_sponge(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  if (!c->linked.empty()) enqueue_sponge(c);
  return launched();
}

int32_t tp_gcbc(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  if (!c->linked.empty()) enqueue_gcbc(c);
  return launched();
}

int32_t tp_rk_update(tp_context *c, int32_t nk) {
  if (int rc = stage_ok(c, nk)) return rc;
  if (!c->linked.empty()) enqueue_rk(c, nk);
  return launched();
}

int32_t tp_wall(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  if (!c->linked.empty()) enqueue_wall(c);
  return launched();
}

int32_t tp_junction(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  if (!c->linked.empty()) enqueue_junction(c);
  return launched();
}

int32_t tp_interface_avg(tp_context *c) {
  if (int rc = check_ready(c)) return rc;
  if (!c->linked.empty()) enqueue_avg(c);
  return launched();
}

int32_t tp_step(tp_context *c, int32_t nsteps) {
  if (int rc = check_ready(c)) return rc;
  for (int s = 0; s < nsteps; s++) enqueue_step(c);
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t tp_step_graph(tp_context *c, int32_t nsteps) {
  if (int rc = check_ready(c)) return rc;
  if (!c->gexec) {
    const bool prof = c->prof;
    c->prof = false;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
    c->prof = prof;
    if (e != hipSuccess) {
      (void)hipGetLastError();
      c->graph = nullptr;
      return fail(TP_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  }
  for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  return 0;
}

int32_t tp_sync(tp_context *c) {
  if (!c) return fail(TP_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipGetLastError());
  return 0;
}

int32_t tp_set_option(tp_context *c, int32_t option, int64_t value) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (option == TP_OPT_EXACT_LINES) {
    c->exact_lines = value != 0;
    drop_graph(c);   // a captured step keeps its kernels: capture again
    return 0;
  }
  if (option == TP_OPT_STRESSES) {
    c->keep_stress = value != 0;
    drop_graph(c);
    return 0;
  }
  if (option == TP_OPT_CHUNKS) {
    const int P = (int)value;
    if (P < 0 || P > 16)
      return fail(TP_E_ARG, "chunks per line must be 0 to 16, not " + std::to_string(value));
    c->col_fixed = P;
    for (auto &a : c->tuned)
      for (auto &b : a)
        for (int &v : b) v = -1;
    if (!c->finalized) return 0;
    return regroup(c);
  }
  return fail(TP_E_ARG, "unknown option " + std::to_string(option));
}

int32_t tp_get_option(tp_context *c, int32_t option, int64_t *value) {
  if (!c || !value) return fail(TP_E_ARG, "null argument");
  switch (option) {
    case TP_OPT_EXACT_LINES: *value = c->exact_lines ? 1 : 0; return 0;
    case TP_OPT_CHUNKS: *value = c->col_fixed; return 0;
    case TP_OPT_STRESSES: *value = c->keep_stress ? 1 : 0; return 0;
  }
  return fail(TP_E_ARG, "unknown option " + std::to_string(option));
}

int32_t tp_autotune(tp_context *c, int32_t reps, int32_t *chosen, double *us) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (!c->finalized) return fail(TP_E_STATE, "tp_autotune needs tp_finalize");
  if (reps < 1) reps = 1;
  const bool prof = c->prof, exact = c->exact_lines;
  c->prof = false;
  c->exact_lines = false;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = 0;
  // time the line solves of one direction and kind over all blocks
  auto timed = [&](int d, int kind, float *ms) -> int {
    // fx is scratch between steps: five fields at a time, as the stage runs
    // them: deriv fx(:,1:5) -> fx(:,6:10), filter fx(:,11:15) in place
    auto once = [&]() {
      if (kind == 0) run_line(c, 0, d, F_FX, 0, F_FX, 5, 0, 0, 5, 1, 1);
      else run_line(c, 1, d, F_FX, 10, F_FX, 10, 0, 0, 5, 1, 1);
    };
    once(); once();
    if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(TP_E_HIP, "event record");
    for (int r = 0; r < reps; r++) once();
    if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipGetLastError() != hipSuccess)
      return fail(TP_E_HIP, "autotune launch failed");
    if (hipEventElapsedTime(ms, e0, e1) != hipSuccess) return fail(TP_E_HIP, "autotune timing failed");
    return 0;
  };
  for (int d = 0; d < 3 && !rc; d++) {
    for (int kind = 0; kind < 2 && !rc; kind++) {
      // the line-length classes present, each tuned in turn with the others fixed
      bool present[kNClass] = {false};
      for (const Block &b : c->blocks) present[col_class(col_default(b.L[d]))] = true;
      float bestms = 1e30f;
      for (int cls = 0; cls < kNClass && !rc; cls++) {
        if (!present[cls]) continue;
        int best = -2;
        bestms = 1e30f;
        for (int pi = -1; pi < kNColP && !rc; pi++) {
          const int P = pi < 0 ? 0 : kColP[pi];   // 0: the sequential kernel
          bool ok = true;
          for (const Block &b : c->blocks)
            if (col_class(col_default(b.L[d])) == cls && P && !col_chunk(b.L[d], P)) ok = false;
          if (!ok) continue;
          c->tuned[d][kind][cls] = P;
          if ((rc = regroup(c))) break;
          float ms = 0;
          if ((rc = timed(d, kind, &ms))) break;
          if (ms < bestms) { bestms = ms; best = P; }
        }
        if (rc) break;
        c->tuned[d][kind][cls] = best >= 0 ? best : -1;
      }
      if (rc) break;
      if ((rc = regroup(c))) break;
      float ms = 0;
      if ((rc = timed(d, kind, &ms))) break;
      // report the chunk count that covers the most points
      long long top = -1;
      int P = 0;
      for (const ColGroup &g : c->L.cg[d][kind])
        if (g.npts > top) { top = g.npts; P = g.P; }
      if (chosen) chosen[d * 2 + kind] = P;
      if (us) us[d * 2 + kind] = 1e3 * ms / reps;
    }
  }
  if (!rc) rc = regroup(c);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->prof = prof;
  c->exact_lines = exact;
  return rc;
}

int32_t tp_col_config(tp_context *c, int32_t *chunks6) {
  if (!c || !chunks6) return fail(TP_E_ARG, "null argument");
  for (int i = 0; i < 6; i++) {
    long long top = -1;
    int P = 0;
    if (!c->exact_lines)
      for (const ColGroup &g : c->L.cg[i / 2][i % 2])
        if (g.npts > top) { top = g.npts; P = g.P; }
    chunks6[i] = P;
  }
  return 0;
}

int32_t tp_profile(tp_context *c, int32_t on) {
  if (!c) return fail(TP_E_ARG, "null context");
  return tp_profile_classes(c, on ? (1 << K_NCLASS) - 1 : 0);
}

int32_t tp_profile_classes(tp_context *c, int32_t mask) {
  if (!c) return fail(TP_E_ARG, "null context");
  c->prof_mask = mask & ((1 << K_NCLASS) - 1);
  c->prof = c->prof_mask != 0;
  return 0;
}

int32_t tp_profile_read(tp_context *c, int32_t k, double *ms, int64_t *launches, double *bytes, int32_t reset) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (k < 0 || k >= K_NCLASS) return fail(TP_E_ARG, "bad kernel class");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < K_NCLASS; q++) {
    for (auto &r : c->pend[q]) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      c->ms[q] += t;
      c->bytes[q] += r.bytes;
      c->cnt[q] += 1;
      c->pool.push_back(r.a);
      c->pool.push_back(r.b);
    }
    c->pend[q].clear();
  }
  if (ms) *ms = c->ms[k];
  if (launches) *launches = c->cnt[k];
  if (bytes) *bytes = c->bytes[k];
  if (reset)
    for (int q = 0; q < K_NCLASS; q++) { c->ms[q] = 0; c->cnt[q] = 0; c->bytes[q] = 0; }
  return 0;
}

int32_t tp_info(tp_context *c, int64_t *out) {
  if (!c || !out) return fail(TP_E_ARG, "null argument");
  size_t bytes = c->link.bytes;
  for (const auto &h : c->cases) bytes += h.mem.bytes;
  out[0] = c->npts;
  out[1] = (int64_t)c->blocks.size();
  out[2] = (int64_t)c->linked.size();
  out[3] = (int64_t)bytes;
  return 0;
}

int32_t tp_case_bytes(tp_context *c, int32_t cs, int64_t *bytes) {
  if (!c || !bytes) return fail(TP_E_ARG, "null argument");
  const int ci = case_index(c, cs);
  if (ci < 0) return fail(TP_E_ARG, "bad case");
  const HostCase &h = c->cases[ci];
  for (int lb = 0; lb < h.nb; lb++)
    if ((int)h.grid.size() <= lb || h.grid[lb].empty())
      return fail(TP_E_STATE, "tp_case_bytes needs the grid of every block (tp_set_grid)");
  *bytes = (int64_t)case_bytes(h);
  return 0;
}

int32_t tp_admit_cases(tp_context *c, int32_t n, const int32_t *ids, int32_t *status) {
  if (int rc = check_ready(c)) return rc;
  if (n < 0 || (n > 0 && (!ids || !status))) return fail(TP_E_ARG, "tp_admit_cases: bad arguments");
  int first = 0;
  std::vector<int> built;   // case ids built here
  for (int i = 0; i < n; i++) {
    const int ci = case_index(c, ids[i]);
    int rc = 0;
    if (ci < 0) rc = fail(TP_E_ARG, "bad case " + std::to_string(ids[i]));
    else if (c->cases[ci].built) rc = fail(TP_E_STATE, "case " + std::to_string(ids[i]) + " already admitted");
    else {
      for (int lb = 0; lb < c->cases[ci].nb && !rc; lb++)
        if ((int)c->cases[ci].grid.size() <= lb || c->cases[ci].grid[lb].empty())
          rc = fail(TP_E_STATE, "admitting case " + std::to_string(ids[i]) +
                                    " needs the grid of every block (tp_set_grid)");
      if (!rc) rc = build_case(c->cases[ci]);
      if (rc) {
        HostCase &h = c->cases[ci];
        h.mem.release();
        h.built = false;
        h.geom = false;
      } else {
        built.push_back(ids[i]);
      }
    }
    status[i] = rc;
    if (rc && !first) first = rc;
  }
  if (built.empty()) return first;
  int rc = link(c);
  if (!rc) rc = tp_setup_geometry(c);
  if (rc) {
    // leave the context as it was: these cases stay pending
    const std::string msg = g_err;
    (void)hipStreamSynchronize(c->stream);
    for (int id : built) {
      HostCase &h = c->cases[case_index(c, id)];
      h.mem.release();
      h.built = false;
      h.geom = false;
    }
    for (int i = 0; i < n; i++)
      if (std::find(built.begin(), built.end(), ids[i]) != built.end()) status[i] = rc;
    link(c);
    g_err = msg;
    return rc;
  }
  return first;   // tp_last_error holds the last failing case's message
}

int32_t tp_admit(tp_context *c, int32_t cs) {
  int32_t st = 0;
  const int rc = tp_admit_cases(c, 1, &cs, &st);
  return rc ? rc : st;
}

int32_t tp_remove_cases(tp_context *c, int32_t n, const int32_t *ids) {
  if (!c) return fail(TP_E_ARG, "null context");
  if (n < 0 || (n > 0 && !ids)) return fail(TP_E_ARG, "tp_remove_cases: bad arguments");
  for (int i = 0; i < n; i++)
    if (case_index(c, ids[i]) < 0) return fail(TP_E_ARG, "bad case " + std::to_string(ids[i]));
  HIPCHK(hipStreamSynchronize(c->stream));
  bool any_built = false;
  for (int i = 0; i < n; i++) {
    const int ci = case_index(c, ids[i]);
    if (ci < 0) continue;   // named twice
    any_built = any_built || c->cases[ci].built;
    c->cases[ci].mem.release();
    c->cases.erase(c->cases.begin() + ci);
  }
  if (c->cases.empty()) c->nkrk = 0;
  if (any_built && c->finalized) return link(c);
  return 0;
}

int32_t tp_remove_case(tp_context *c, int32_t cs) { return tp_remove_cases(c, 1, &cs); }

int32_t tp_adopt_cases(tp_context *dst, int32_t n, tp_context *const *srcs, const int32_t *case_ids,
                       int32_t *new_ids) {
  if (int rc = check_ready(dst)) return rc;
  if (n < 0 || (n > 0 && (!srcs || !case_ids || !new_ids))) return fail(TP_E_ARG, "tp_adopt_cases: bad arguments");
  // every case must be set up on the device of dst, in a context other than dst
  for (int i = 0; i < n; i++) {
    tp_context *src = srcs[i];
    if (!src || src == dst) return fail(TP_E_ARG, "tp_adopt_cases: a source context is missing or is dst");
    if (src->device != dst->device) return fail(TP_E_ARG, "tp_adopt_cases: contexts on different devices");
    const int ci = case_index(src, case_ids[i]);
    if (ci < 0) return fail(TP_E_ARG, "tp_adopt_cases: bad case " + std::to_string(case_ids[i]));
    const HostCase &h = src->cases[ci];
    if (!h.built || !h.geom || h.slot < 0)
      return fail(TP_E_STATE, "tp_adopt_cases: case " + std::to_string(case_ids[i]) + " is not set up on the device");
    if (dst->nkrk && h.ip[TP_I_NKRK] != dst->nkrk)
      return fail(TP_E_ARG, "all cases of a context must use the same number of RK stages");
    for (int j = 0; j < i; j++)
      if (srcs[j] == src && case_ids[j] == case_ids[i]) return fail(TP_E_ARG, "tp_adopt_cases: a case named twice");
  }
  HIPCHK(hipStreamSynchronize(dst->stream));
  std::vector<tp_context *> touched;
  for (int i = 0; i < n; i++) {
    tp_context *src = srcs[i];
    HIPCHK(hipStreamSynchronize(src->stream));
    const int ci = case_index(src, case_ids[i]);
    HostCase h = std::move(src->cases[ci]);
    // the scalars of the case travel with it (n, t, dt, frame velocity, ...)
    HIPCHK(hipMemcpy(h.sc, src->d_sc + (size_t)h.slot * S_STRIDE, sizeof(double) * S_STRIDE, hipMemcpyDeviceToHost));
    h.sc_saved = true;
    src->cases.erase(src->cases.begin() + ci);
    if (src->cases.empty()) src->nkrk = 0;
    if (std::find(touched.begin(), touched.end(), src) == touched.end()) touched.push_back(src);
    h.id = dst->next_id++;
    h.slot = -1;
    new_ids[i] = h.id;
    dst->nkrk = h.ip[TP_I_NKRK];
    dst->cases.push_back(std::move(h));
  }
  for (tp_context *src : touched)
    if (int rc = link(src)) return rc;
  return link(dst);
}

int32_t tp_case_info(tp_context *c, int32_t cs, int64_t *out4) {
  if (!c || !out4) return fail(TP_E_ARG, "null argument");
  const int ci = case_index(c, cs);
  if (ci < 0) return fail(TP_E_ARG, "bad case");
  const HostCase &h = c->cases[ci];
  int64_t np = 0;
  for (int lb = 0; lb < h.nb; lb++)
    np += (int64_t)(h.dims[3 * lb] + 1) * (h.dims[3 * lb + 1] + 1) * (h.dims[3 * lb + 2] + 1);
  out4[0] = h.geom ? 2 : h.built ? 1 : 0;
  out4[1] = np;
  out4[2] = h.nb;
  out4[3] = (int64_t)h.mem.bytes;
  return 0;
}

int32_t tp_stats_start(tp_context *c, int32_t cs, int32_t every) {
  if (int rc = check_ready(c)) return rc;
  HostCase *h = linked_case(c, cs);
  if (!h || every < 0) return fail(TP_E_ARG, "tp_stats_start: bad case or interval");
  double s[S_STRIDE];
  HIPCHK(hipMemcpyAsync(s, c->d_sc + h->slot * S_STRIDE, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  s[S_ST_EVERY] = every;
  s[S_ST_T0] = s[S_ST_TP] = s[S_ST_NS] = 0.0;
  s[S_CL] = s[S_CD] = s[S_CLS] = s[S_CDS] = 0.0;
  HIPCHK(hipMemcpyAsync(c->d_sc + h->slot * S_STRIDE, s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  for (const Block &b : h->blk)
    if (b.nsurf > 0) HIPCHK(hipMemsetAsync(b.sst, 0, sizeof(double) * 3 * (size_t)b.nsurf, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int32_t tp_stats_forces(tp_context *c, int32_t cs, double *out8) {
  if (int rc = check_ready(c)) return rc;
  const HostCase *h = linked_case(c, cs);
  if (!h || !out8) return fail(TP_E_ARG, "tp_stats_forces: bad case");
  double s[S_STRIDE];
  HIPCHK(hipMemcpyAsync(s, c->d_sc + h->slot * S_STRIDE, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double T = s[S_ST_TP] - s[S_ST_T0];
  out8[0] = s[S_ST_NS]; out8[1] = s[S_ST_T0]; out8[2] = s[S_ST_TP];
  out8[3] = s[S_CL]; out8[4] = s[S_CD];
  out8[5] = T > 0 ? s[S_CLS] / T : s[S_CL];
  out8[6] = T > 0 ? s[S_CDS] / T : s[S_CD];
  out8[7] = T;
  return 0;
}

int64_t tp_stats_wall_count(tp_context *c, int32_t cs, int32_t lb) {
  if (check_ready(c)) return -1;
  const int g = block_index(c, cs, lb);
  return g < 0 ? -1 : c->blocks[g].nsurf;
}

int32_t tp_stats_wall(tp_context *c, int32_t cs, int32_t lb, double *mean_p, double *rms_p, double *area,
                      double *wnor3, int64_t count) {
  if (int rc = check_ready(c)) return rc;
  const int g = block_index(c, cs, lb);
  if (g < 0) return fail(TP_E_ARG, "tp_stats_wall: bad case/block");
  const Block &b = c->blocks[g];
  const int nw = b.nsurf;
  if (count != nw) return fail(TP_E_SIZE, "tp_stats_wall: the block has " + std::to_string(nw) + " wall points");
  if (nw == 0) return 0;
  const HostCase *h = linked_case(c, cs);
  double s[S_STRIDE];
  std::vector<double> st(3 * (size_t)nw), ar(nw), nr(3 * (size_t)nw);
  HIPCHK(hipMemcpyAsync(s, c->d_sc + h->slot * S_STRIDE, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(st.data(), b.sst, sizeof(double) * st.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(ar.data(), b.sarea, sizeof(double) * ar.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(nr.data(), b.snor, sizeof(double) * nr.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double T = s[S_ST_TP] - s[S_ST_T0];
  for (int ll = 0; ll < nw; ll++) {
    const double m = T > 0 ? st[nw + ll] / T : st[ll];
    const double v = T > 0 ? st[2 * (size_t)nw + ll] / T - m * m : 0.0;
    if (mean_p) mean_p[ll] = m;
    if (rms_p) rms_p[ll] = std::sqrt(v > 0.0 ? v : 0.0);
    if (area) area[ll] = ar[ll];
  }
  if (wnor3) std::memcpy(wnor3, nr.data(), sizeof(double) * nr.size());
  return 0;
}

int32_t tp_set_parents(tp_context *c, int32_t cs, int32_t nparent, const int32_t *parent, const int32_t *start3,
                       const int32_t *pdims3) {
  if (!c || !parent || !start3 || !pdims3 || nparent < 1) return fail(TP_E_ARG, "tp_set_parents: bad arguments");
  const int ci = case_index(c, cs);
  if (ci < 0) return fail(TP_E_ARG, "bad case");
  HostCase &h = c->cases[ci];
  for (int b = 0; b < h.nb; b++) {
    const int mb = parent[b];
    if (mb < 0 || mb >= nparent) return fail(TP_E_LAYOUT, "sub-block " + std::to_string(b) + ": bad parent");
    for (int d = 0; d < 3; d++)
      if (start3[3 * b + d] < 0 || start3[3 * b + d] + h.dims[3 * b + d] > pdims3[3 * mb + d])
        return fail(TP_E_LAYOUT, "sub-block " + std::to_string(b) + " does not fit in its parent block");
  }
  h.npar = nparent;
  h.parent.assign(parent, parent + h.nb);
  h.pstart.assign(start3, start3 + 3 * h.nb);
  h.pdims.assign(pdims3, pdims3 + 3 * nparent);
  return 0;
}

int32_t tp_get_parents(tp_context *c, int32_t cs, int32_t *nparent, int32_t *parent, int32_t *start3,
                       int32_t *pdims3) {
  if (!c) return fail(TP_E_ARG, "null context");
  const int ci = case_index(c, cs);
  if (ci < 0) return fail(TP_E_ARG, "bad case");
  const HostCase &h = c->cases[ci];
  const bool set = h.npar > 0;
  if (nparent) *nparent = set ? h.npar : h.nb;
  for (int b = 0; b < h.nb; b++) {
    if (parent) parent[b] = set ? h.parent[b] : b;
    for (int d = 0; d < 3; d++) {
      if (start3) start3[3 * b + d] = set ? h.pstart[3 * b + d] : 0;
      if (pdims3 && !set) pdims3[3 * b + d] = h.dims[3 * b + d];
    }
  }
  if (pdims3 && set)
    for (int q = 0; q < 3 * h.npar; q++) pdims3[q] = h.pdims[q];
  return 0;
}

int32_t tp_get_grid(tp_context *c, int32_t cs, int32_t lb, double *xyz, int64_t count) {
  if (!c || !xyz) return fail(TP_E_ARG, "null argument");
  const int ci = case_index(c, cs);
  if (ci < 0 || lb < 0 || lb >= c->cases[ci].nb) return fail(TP_E_ARG, "tp_get_grid: bad case/block");
  const HostCase &h = c->cases[ci];
  if ((int)h.grid.size() <= lb || h.grid[lb].empty()) return fail(TP_E_STATE, "grid not set");
  if (count != (int64_t)h.grid[lb].size())
    return fail(TP_E_SIZE, "tp_get_grid: expected " + std::to_string(h.grid[lb].size()) + " values");
  std::memcpy(xyz, h.grid[lb].data(), sizeof(double) * h.grid[lb].size());
  return 0;
}

int32_t tp_case_dims(tp_context *c, int32_t cs, int32_t *dims3) {
  if (!c || !dims3) return fail(TP_E_ARG, "null argument");
  const int ci = case_index(c, cs);
  if (ci < 0) return fail(TP_E_ARG, "bad case");
  std::memcpy(dims3, c->cases[ci].dims.data(), sizeof(int32_t) * c->cases[ci].dims.size());
  return 0;
}

int32_t tp_case_params(tp_context *c, int32_t cs, int32_t *iparams, double *dparams) {
  if (!c) return fail(TP_E_ARG, "null context");
  const int ci = case_index(c, cs);
  if (ci < 0) return fail(TP_E_ARG, "bad case");
  if (iparams) std::memcpy(iparams, c->cases[ci].ip, sizeof(c->cases[ci].ip));
  if (dparams) std::memcpy(dparams, c->cases[ci].dp, sizeof(c->cases[ci].dp));
  return 0;
}

// ---- the derivative of one field held by the caller ----
namespace {

// device bytes of the work area of tp_field_deriv: block view, coefficients,
// chunked-solver table, work list (each 256-byte aligned)
struct FieldWs {
  size_t blk, coef, tab, work, total;
};
FieldWs field_ws(int L, int P, int nwork) {
  FieldWs w;
  size_t off = 0;
  auto take = [&](size_t n) { size_t o = off; off += (n + 255) / 256 * 256; return o; };
  w.blk = take(sizeof(Block));
  w.coef = take(sizeof(CaseCoef));
  w.tab = take(sizeof(double) * col_table_doubles(L, P));
  w.work = take(sizeof(Work) * (size_t)nwork);
  w.total = off;
  return w;
}

// everything tp_field_deriv launches with, built on the host: the block
// view of the caller's arrays, the derivative's coefficients, the table of
// P chunks and the work list; the launch needs the work area's address
struct FieldCall {
  int d = 0, L = 0, P = 0, C = 0, nwork = 0;
  bool periodic = false;
  Block blk;
  CaseCoef coef;
  std::vector<double> tab;
  std::vector<Work> work;
  FieldWs ws;
};

int field_prepare(int32_t lxi, int32_t let, int32_t lze, const int32_t *nbc2, int32_t nn, const double *src,
                  double *dst, FieldCall &f) {
  if (nn < 1 || nn > 3 || !nbc2 || !src || !dst) return fail(TP_E_ARG, "tp_field_deriv: bad arguments");
  const int d = nn - 1;
  const int dims[3] = {lxi, let, lze};
  int64_t np = 1;
  for (int e = 0; e < 3; e++) {
    if (dims[e] < 1) return fail(TP_E_LAYOUT, "tp_field_deriv: fewer than 2 points in a direction");
    np *= (int64_t)dims[e] + 1;
  }
  if (8 * np >= (int64_t)1 << 32) return fail(TP_E_SIZE, "tp_field_deriv: a block holds fewer than 2^29 points");
  for (int s = 0; s < 2; s++) {
    const int c = nbc2[s];
    if (c != 10 && c != 20 && c != 25 && c != 30 && c != 45)
      return fail(TP_E_ARG, "tp_field_deriv: end codes are 10, 20, 25, 30 (one-sided) or 45 (periodic)");
  }
  if ((nbc2[0] == 45) != (nbc2[1] == 45)) return fail(TP_E_ARG, "tp_field_deriv: a periodic line has both ends 45");
  f.d = d;
  f.L = dims[d] + 1;
  f.periodic = nbc2[0] == 45;
  if (f.L < (f.periodic ? LMD + 2 : 7))
    return fail(TP_E_LAYOUT, "tp_field_deriv: lines of " + std::to_string(f.L) + " points are too short");
  f.P = col_default(f.L);
  if (!f.P) return fail(TP_E_SIZE, "tp_field_deriv: lines longer than 512 points need a context (tp_deriv)");
  f.C = col_capacity((f.L + f.P - 1) / f.P);
  std::memset(&f.blk, 0, sizeof(Block));
  Block &b = f.blk;
  b.nx = lxi + 1; b.ny = let + 1; b.nz = lze + 1;
  b.N = (int)np;
  b.L[0] = b.nx; b.L[1] = b.ny; b.L[2] = b.nz;
  b.F[0] = b.ny * b.nz; b.F[1] = b.nz * b.nx; b.F[2] = b.nx * b.ny;
  for (int e = 0; e < 3; e++)
    for (int s = 0; s < 2; s++) { b.nbc[e][s] = 10; b.ncd[e][s] = -1; }
  for (int s = 0; s < 2; s++) {
    b.nbc[d][s] = nbc2[s];
    b.ndf[d][s] = f.periodic ? 1 : 0;
    b.ncd[d][s] = f.periodic ? 0 : -1;
  }
  b.cs = 0;
  b.rr = const_cast<double *>(src);   // read as rr(:,1) ...
  b.de = dst;                         // ... written as de(:,1)
  // coefficients: the derivative's closure rows, src/Main3D.f90:71-73, 309-312
  std::memset(&f.coef, 0, sizeof(CaseCoef));
  std::memcpy(f.coef.abc, kAbc, sizeof(kAbc));
  closure_rows(0.85 * kPi, f.coef);
  // LU factors of this direction (src/Main3D.f90:316-332) and their chunk table
  double albed[2][3][5];
  const double inner[3][5] = {{0, 0, 1, kAlpha, kBeta}, {0, kAlpha, 1, kAlpha, kBeta}, {kBeta, kAlpha, 1, kAlpha, kBeta}};
  std::memcpy(albed[0], kBnd, sizeof(kBnd));
  std::memcpy(albed[1], inner, sizeof(inner));
  std::vector<double> lu(5 * (size_t)f.L);
  double *l1 = &lu[0], *l2 = l1 + f.L, *u1 = l2 + f.L, *u2 = u1 + f.L, *u3 = u2 + f.L;
  const int end = f.periodic ? 1 : 0;
  band_lu(f.L, albed[end], albed[end], kAlpha, kBeta, l1, l2, u1, u2, u3);
  f.tab = col_tables(f.L, f.P, l1, l2, u1, u2, u3);
  const int64_t nlines = np / f.L;
  f.work.clear();
  for (int64_t q = 0; q < nlines; q += 64) f.work.push_back({0, (int)q, (int)std::min<int64_t>(64, nlines - q), d});
  f.nwork = (int)f.work.size();
  f.ws = field_ws(f.L, f.P, f.nwork);
  return 0;
}

// copy the call's tables into the work area ws and launch the solve on st
int field_launch(FieldCall &f, char *ws, hipStream_t st) {
  f.blk.col[f.d][0] = reinterpret_cast<const double *>(ws + f.ws.tab);
  HIPCHK(hipMemcpyAsync(ws + f.ws.blk, &f.blk, sizeof(Block), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ws + f.ws.coef, &f.coef, sizeof(CaseCoef), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ws + f.ws.tab, f.tab.data(), sizeof(double) * f.tab.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ws + f.ws.work, f.work.data(), sizeof(Work) * f.work.size(), hipMemcpyHostToDevice, st));
  LineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.blocks = reinterpret_cast<const Block *>(ws + f.ws.blk);
  a.work = reinterpret_cast<const Work *>(ws + f.ws.work);
  a.coef = reinterpret_cast<const CaseCoef *>(ws + f.ws.coef);
  a.in_arr = F_RR; a.in_col = 0; a.out_arr = F_DE; a.out_col = 0;
  a.m = 0; a.acc = 0; a.nlp = 65; a.copy_arr = -1;
  a.direct = f.periodic ? 1 : 0;   // periodic: closure sums from the other end of the same lines
  a.n45 = 0;
  launch_col(f.d, 0, f.C, f.P, f.nwork, 1, st, a);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

int32_t tp_field_deriv(int32_t lxi, int32_t let, int32_t lze, const int32_t *nbc2, int32_t nn, const double *src,
                       double *dst, void *stream) {
  FieldCall f;
  if (int rc = field_prepare(lxi, let, lze, nbc2, nn, src, dst, f)) return rc;
  if (int rc = runtime_check()) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  char *ws = nullptr;
  HIPCHK(hipMalloc(&ws, f.ws.total));
  int rc = field_launch(f, ws, st);
  // the work area lives until the solve has run
  const hipError_t e = hipStreamSynchronize(st);
  (void)hipFree(ws);
  if (rc) return rc;
  if (e != hipSuccess) return fail(TP_E_HIP, std::string("tp_field_deriv: ") + hipGetErrorString(e));
  return 0;
}

// ---- the same without allocating: the work area in a device array ----
// One array per GPU the library runs on, held until the process ends.  Its
// name is only a host-side stand-in: host code uses the address that
// hipGetSymbolAddress gives for the current device, looked up on the first
// call on each device.  Calls that overlap on one GPU (other streams, other
// host threads) share it and race; calls in turn on one stream are ordered.
// Nothing in it needs to start at zero: every call writes the parts it reads.
constexpr size_t kScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch[kMaxDevices] = {nullptr};

int scratch_address(char **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(TP_E_HIP, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TP_E_HIP, "hipGetSymbolAddress gave no address for the work area");
    g_scratch[dev] = static_cast<char *>(p);
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace

int32_t tp_field_deriv_noalloc(int32_t lxi, int32_t let, int32_t lze, const int32_t *nbc2, int32_t nn,
                               const double *src, double *dst, void *stream) {
  FieldCall f;
  if (int rc = field_prepare(lxi, let, lze, nbc2, nn, src, dst, f)) return rc;
  if (int rc = runtime_check()) return rc;
  if (f.ws.total > kScratchBytes)
    return fail(TP_E_SIZE, "tp_field_deriv_noalloc: the call needs " + std::to_string(f.ws.total) +
                               " bytes of work area, the library holds " + std::to_string(kScratchBytes) +
                               " (at most about 4 M lines); use tp_field_deriv");
  char *ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return field_launch(f, ws, static_cast<hipStream_t>(stream));
}

int32_t tp_gust_mode(double tu, double kr, const double *uoo3, double *out8, double *slit) {
  if (!uoo3 || !out8 || !slit || !(kr > 0) || !(tu >= 0)) return fail(TP_E_ARG, "tp_gust_mode: bad arguments");
  if (!(uoo3[0] > 0)) return fail(TP_E_ARG, "tp_gust_mode: u_inf needs a positive x component");
  gust_mode(tu, kr, uoo3, out8, slit);
  return 0;
}

}  // extern "C"

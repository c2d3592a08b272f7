// tp_line.hip -- batched pentadiagonal line solves (compact derivative and
// compact filter) and the device-side closure exchange.
//
// Reference: src/Subroutines3D.f90:16-80 (mpigo), :84-141 (deriv),
// :145-206 (filte).  The arithmetic of every line follows the reference's
// operation order (RHS stencil, forward substitution with xl, backward
// substitution with xu) and the library is compiled without FP contraction,
// so a line solved here rounds like the reference's.
//
// Mapping to CDNA4: one 64-lane wave owns a tile of NL lines of one block.
// The tile is staged in LDS with coalesced HBM loads (along the line for xi,
// across consecutive xi lines for eta and zeta), each lane then runs the
// sequential recurrences of one line entirely in LDS (RHS fused into the
// forward sweep through a 7-value register window), and the tile is written
// back with coalesced stores.  Lines of every block of every case are one
// launch: the work table maps workgroups to (block, first line).
#include "tp_internal.hpp"

namespace tp {



// address of point p of tile line q in direction D, and its face index jk
template <int D>
__device__ inline void line_addr(const Block &b, int q, int &base, int &stride, int &jk) {
  if (D == 0) {            // rows: q == jk == zeta*ny + eta
    base = q * b.nx; stride = 1; jk = q;
  } else if (D == 1) {     // q = zeta*nx + xi, jk = xi*nz + zeta
    int xi = q % b.nx, ze = q / b.nx;
    base = ze * b.ny * b.nx + xi; stride = b.nx; jk = xi * b.nz + ze;
  } else {                 // q = eta*nx + xi == jk
    base = q; stride = b.nx * b.ny; jk = q;
  }
}

template <int D>
__device__ inline void tile_load(const Block &b, const double *__restrict__ src, int q0, int nl,
                                 int L, int nlp, double *lds) {
  if (D == 0) {
    for (int r = 0; r < nl; r++) {
      const double *row = src + (size_t)(q0 + r) * b.nx;
      for (int p = threadIdx.x; p < L; p += blockDim.x) lds[p * nlp + r] = row[p];
    }
  } else {
    int tot = nl * L;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      int r = e % nl, p = e / nl;
      int base, stride, jk;
      line_addr<D>(b, q0 + r, base, stride, jk);
      lds[p * nlp + r] = src[base + (size_t)p * stride];
    }
  }
}

// ADD: dst = (addsrc, or dst itself) + the tile
template <int D, int ADD>
__device__ inline void tile_store(const Block &b, double *__restrict__ dst, int q0, int nl, int L,
                                  int nlp, const double *lds, double *copy = nullptr,
                                  const double *addsrc = nullptr) {
  if (D == 0) {
    for (int r = 0; r < nl; r++) {
      double *row = dst + (size_t)(q0 + r) * b.nx;
      const double *arow = addsrc ? addsrc + (size_t)(q0 + r) * b.nx : row;
      for (int p = threadIdx.x; p < L; p += blockDim.x) {
        double v = lds[p * nlp + r];
        v = ADD ? arow[p] + v : v;
        row[p] = v;
        if (copy) copy[(size_t)(q0 + r) * b.nx + p] = v;
      }
    }
  } else {
    int tot = nl * L;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      int r = e % nl, p = e / nl;
      int base, stride, jk;
      line_addr<D>(b, q0 + r, base, stride, jk);
      const size_t o = base + (size_t)p * stride;
      double v = lds[p * nlp + r];
      v = ADD ? (addsrc ? addsrc[o] : dst[o]) + v : v;
      dst[o] = v;
      if (copy) copy[o] = v;
    }
  }
}

// ---- compact first derivative along direction D, Subroutines3D.f90:84-141 ----
template <int D>
__global__ void __launch_bounds__(64) k_deriv(LineArgs a) {
  extern __shared__ double lds[];
  const Work w = a.work[blockIdx.x];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const int L = b.L[D], nl = w.n, q0 = w.a, nlp = a.nlp;
  const int f = blockIdx.y;
  const double *src = field_ptr(b, a.in_arr, a.in_col + f * a.in_step);
  double *dst = field_ptr(b, a.out_arr, a.out_col + f * a.out_step);
  const int m = a.m > 0 ? a.m + f : 0;
  tile_load<D>(b, src, q0, nl, L, nlp, lds);
  __syncthreads();
  const int r = threadIdx.x;
  if (r < nl) {
    double *s = lds + r;
#define S(i) s[(i) * nlp]
    int base, stride, jk;
    line_addr<D>(b, q0 + r, base, stride, jk);
    const int F = b.F[D];
    const double *rcv = b.recv[D] + (size_t)f * 6 * F;
    const double aa = 0.6431406736919156, ab = 0.2586011023495066, ac = 0.007140953479797375;
    const int ie = L - 1;
    double rs0, rs1, rs2, re0, re1, re2;
    if (b.ndf[D][0] == 0) {
      double s0 = S(0), s1 = S(1), s2 = S(2), s3 = S(3), s4 = S(4), s5 = S(5), s6 = S(6);
      rs0 = 0; rs0 = rs0 + c.abc[0][0] * (s1 - s0); rs0 = rs0 + c.abc[0][1] * (s2 - s0);
      rs0 = rs0 + c.abc[0][2] * (s3 - s0); rs0 = rs0 + c.abc[0][3] * (s4 - s0);
      rs0 = rs0 + c.abc[0][4] * (s5 - s0); rs0 = rs0 + c.abc[0][5] * (s6 - s0);
      rs1 = 0; rs1 = rs1 + c.abc[1][0] * (s0 - s1); rs1 = rs1 + c.abc[1][1] * (s2 - s1);
      rs1 = rs1 + c.abc[1][2] * (s3 - s1); rs1 = rs1 + c.abc[1][3] * (s4 - s1);
      rs1 = rs1 + c.abc[1][4] * (s5 - s1); rs1 = rs1 + c.abc[1][5] * (s6 - s1);
      rs2 = 0; rs2 = rs2 + c.abc[2][0] * (s0 - s2); rs2 = rs2 + c.abc[2][1] * (s1 - s2);
      rs2 = rs2 + c.abc[2][2] * (s3 - s2); rs2 = rs2 + c.abc[2][3] * (s4 - s2);
      rs2 = rs2 + c.abc[2][4] * (s5 - s2); rs2 = rs2 + c.abc[2][5] * (s6 - s2);
    } else {
      double t0 = 0, t1 = 0;
      for (int q = 0; q <= LMD; q++) { double v = S(q); t0 = t0 + c.pbci[0][0][q] * v; t1 = t1 + c.pbci[0][1][q] * v; }
      rs0 = t0 + rcv[(0 * 3 + 0) * F + jk];
      rs1 = t1 + rcv[(0 * 3 + 1) * F + jk];
      rs2 = aa * (S(3) - S(1)) + ab * (S(4) - S(0)) + ac * (S(5) - rcv[(0 * 3 + 2) * F + jk]);
    }
    if (b.ndf[D][1] == 0) {
      double e0 = S(ie), e1 = S(ie - 1), e2 = S(ie - 2), e3 = S(ie - 3), e4 = S(ie - 4), e5 = S(ie - 5),
             e6 = S(ie - 6);
      re0 = 0; re0 = re0 + c.abc[0][0] * (e0 - e1); re0 = re0 + c.abc[0][1] * (e0 - e2);
      re0 = re0 + c.abc[0][2] * (e0 - e3); re0 = re0 + c.abc[0][3] * (e0 - e4);
      re0 = re0 + c.abc[0][4] * (e0 - e5); re0 = re0 + c.abc[0][5] * (e0 - e6);
      re1 = 0; re1 = re1 + c.abc[1][0] * (e1 - e0); re1 = re1 + c.abc[1][1] * (e1 - e2);
      re1 = re1 + c.abc[1][2] * (e1 - e3); re1 = re1 + c.abc[1][3] * (e1 - e4);
      re1 = re1 + c.abc[1][4] * (e1 - e5); re1 = re1 + c.abc[1][5] * (e1 - e6);
      re2 = 0; re2 = re2 + c.abc[2][0] * (e2 - e0); re2 = re2 + c.abc[2][1] * (e2 - e1);
      re2 = re2 + c.abc[2][2] * (e2 - e3); re2 = re2 + c.abc[2][3] * (e2 - e4);
      re2 = re2 + c.abc[2][4] * (e2 - e5); re2 = re2 + c.abc[2][5] * (e2 - e6);
    } else {
      double t0 = 0, t1 = 0;
      for (int q = 0; q <= LMD; q++) { double v = S(ie - q); t0 = t0 + c.pbci[0][0][q] * v; t1 = t1 + c.pbci[0][1][q] * v; }
      re0 = -t0 - rcv[(1 * 3 + 0) * F + jk];
      re1 = -t1 - rcv[(1 * 3 + 1) * F + jk];
      re2 = aa * (S(ie - 1) - S(ie - 3)) + ab * (S(ie) - S(ie - 4)) + ac * (rcv[(1 * 3 + 2) * F + jk] - S(ie - 5));
    }
    const double *l1 = b.lud + 0 * b.nlu + b.lu_off[D], *l2 = b.lud + 1 * b.nlu + b.lu_off[D];
    const double *u1 = b.lud + 2 * b.nlu + b.lu_off[D], *u2 = b.lud + 3 * b.nlu + b.lu_off[D];
    const double *u3 = b.lud + 4 * b.nlu + b.lu_off[D];
    // forward substitution with the RHS generated on the fly
    double wm3 = S(0), wm2 = S(1), wm1 = S(2), w0 = S(3), wp1 = S(4), wp2 = S(5);
    double y0 = rs0;
    double y1 = rs1 - l2[1] * y0;
    double y2 = rs2 - l1[2] * y0 - l2[2] * y1;
    S(0) = y0; S(1) = y1; S(2) = y2;
    double ym2 = y1, ym1 = y2;
    for (int i = 3; i <= ie - 3; i++) {
      double wp3 = S(i + 3);
      double sb = aa * (wp1 - wm1) + ab * (wp2 - wm2) + ac * (wp3 - wm3);
      double y = sb - l1[i] * ym2 - l2[i] * ym1;
      S(i) = y;
      ym2 = ym1; ym1 = y;
      wm3 = wm2; wm2 = wm1; wm1 = w0; w0 = wp1; wp1 = wp2; wp2 = wp3;
    }
    {
      double y = re2 - l1[ie - 2] * ym2 - l2[ie - 2] * ym1; S(ie - 2) = y; ym2 = ym1; ym1 = y;
      y = re1 - l1[ie - 1] * ym2 - l2[ie - 1] * ym1; S(ie - 1) = y; ym2 = ym1; ym1 = y;
      y = re0 - l1[ie] * ym2 - l2[ie] * ym1; S(ie) = y;
    }
    // backward substitution
    double xp2 = u1[ie] * S(ie);
    S(ie) = xp2;
    double xp1 = u1[ie - 1] * S(ie - 1) - u2[ie - 1] * xp2;
    S(ie - 1) = xp1;
    for (int i = ie - 2; i >= 0; i--) {
      double x = u1[i] * S(i) - u2[i] * xp1 - u3[i] * xp2;
      S(i) = x;
      xp2 = xp1; xp1 = x;
    }
    if (m > 0) {
      b.drva[D][(0 * 5 + m - 1) * F + jk] = S(0);
      b.drva[D][(1 * 5 + m - 1) * F + jk] = S(ie);
    }
#undef S
  }
  __syncthreads();
  if (a.acc) tile_store<D, 1>(b, dst, q0, nl, L, nlp, lds);
  else tile_store<D, 0>(b, dst, q0, nl, L, nlp, lds);
}

// ---- compact filter along direction D, Subroutines3D.f90:145-206: in place,
// or into another field (out_arr, out_col) when the caller asks ----
template <int D>
__global__ void __launch_bounds__(64) k_filter(LineArgs a) {
  extern __shared__ double lds[];
  const Work w = a.work[blockIdx.x];
  const Block &b = a.blocks[w.blk];
  const CaseCoef &c = a.coef[b.cs];
  const int L = b.L[D], nl = w.n, q0 = w.a, nlp = a.nlp;
  const int f = blockIdx.y;
  double *io = field_ptr(b, a.in_arr, a.in_col + f * a.in_step);
  tile_load<D>(b, io, q0, nl, L, nlp, lds);
  __syncthreads();
  const int r = threadIdx.x;
  if (r < nl) {
    double *s = lds + r;
#define S(i) s[(i) * nlp]
    int base, stride, jk;
    line_addr<D>(b, q0 + r, base, stride, jk);
    const int F = b.F[D];
    const double *rcv = b.recv[D] + (size_t)f * 6 * F;
    const double fa = c.fa, fb = c.fb, fc = c.fc;
    const int ie = L - 1;
    double rs0, rs1, rs2, re0, re1, re2;
    if (b.ndf[D][0] == 0) {
      double s0 = S(0), s1 = S(1), s2 = S(2), s3 = S(3), s4 = S(4), s5 = S(5);
      double ra0 = 2.0 * s0, ra1 = 2.0 * s1, ra2 = 2.0 * s2;
      double res = 0;
      res = res + c.fex[0] * (s1 - s0); res = res + c.fex[1] * (s2 - s0); res = res + c.fex[2] * (s3 - s0);
      double rv0 = s0 - res * c.ffx[0], rv1 = s0 - res * c.ffx[1], rv2 = s0 - res * c.ffx[2];
      rs0 = c.fam[0] * (rv0 + s1 - ra0) + c.fbm[0] * (rv1 + s2 - ra0) + c.fcm[0] * (rv2 + s3 - ra0);
      rs1 = c.fam[1] * (s0 + s2 - ra1) + c.fbm[1] * (rv0 + s3 - ra1) + c.fcm[1] * (rv1 + s4 - ra1);
      rs2 = c.fam[2] * (s1 + s3 - ra2) + c.fbm[2] * (s0 + s4 - ra2) + c.fcm[2] * (rv0 + s5 - ra2);
    } else {
      double ra2 = 2.0 * S(2);
      double t0 = 0, t1 = 0;
      for (int q = 0; q <= LMF; q++) { double v = S(q); t0 = t0 + c.pbci[1][0][q] * v; t1 = t1 + c.pbci[1][1][q] * v; }
      rs0 = t0 + rcv[(0 * 3 + 0) * F + jk];
      rs1 = t1 + rcv[(0 * 3 + 1) * F + jk];
      rs2 = fa * (S(1) + S(3) - ra2) + fb * (S(0) + S(4) - ra2) + fc * (rcv[(0 * 3 + 2) * F + jk] + S(5) - ra2);
    }
    if (b.ndf[D][1] == 0) {
      double e0 = S(ie), e1 = S(ie - 1), e2 = S(ie - 2), e3 = S(ie - 3), e4 = S(ie - 4), e5 = S(ie - 5);
      double ra0 = 2.0 * e0, ra1 = 2.0 * e1, ra2 = 2.0 * e2;
      double res = 0;
      res = res + c.fex[0] * (e0 - e1); res = res + c.fex[1] * (e0 - e2); res = res + c.fex[2] * (e0 - e3);
      double rv0 = e0 + res * c.ffx[0], rv1 = e0 + res * c.ffx[1], rv2 = e0 + res * c.ffx[2];
      re0 = c.fam[0] * (e1 + rv0 - ra0) + c.fbm[0] * (e2 + rv1 - ra0) + c.fcm[0] * (e3 + rv2 - ra0);
      re1 = c.fam[1] * (e2 + e0 - ra1) + c.fbm[1] * (e3 + rv0 - ra1) + c.fcm[1] * (e4 + rv1 - ra1);
      re2 = c.fam[2] * (e3 + e1 - ra2) + c.fbm[2] * (e4 + e0 - ra2) + c.fcm[2] * (e5 + rv0 - ra2);
    } else {
      double ra2 = 2.0 * S(ie - 2);
      double t0 = 0, t1 = 0;
      for (int q = 0; q <= LMF; q++) { double v = S(ie - q); t0 = t0 + c.pbci[1][0][q] * v; t1 = t1 + c.pbci[1][1][q] * v; }
      re0 = t0 + rcv[(1 * 3 + 0) * F + jk];
      re1 = t1 + rcv[(1 * 3 + 1) * F + jk];
      re2 = fa * (S(ie - 3) + S(ie - 1) - ra2) + fb * (S(ie - 4) + S(ie) - ra2) +
            fc * (S(ie - 5) + rcv[(1 * 3 + 2) * F + jk] - ra2);
    }
    const double *l1 = b.luf + 0 * b.nlu + b.lu_off[D], *l2 = b.luf + 1 * b.nlu + b.lu_off[D];
    const double *u1 = b.luf + 2 * b.nlu + b.lu_off[D], *u2 = b.luf + 3 * b.nlu + b.lu_off[D];
    const double *u3 = b.luf + 4 * b.nlu + b.lu_off[D];
    double wm3 = S(0), wm2 = S(1), wm1 = S(2), w0 = S(3), wp1 = S(4), wp2 = S(5);
    double y0 = rs0;
    double y1 = rs1 - l2[1] * y0;
    double y2 = rs2 - l1[2] * y0 - l2[2] * y1;
    S(0) = y0; S(1) = y1; S(2) = y2;
    double ym2 = y1, ym1 = y2;
    for (int i = 3; i <= ie - 3; i++) {
      double wp3 = S(i + 3);
      double res = 2.0 * w0;
      double sb = fa * (wm1 + wp1 - res) + fb * (wm2 + wp2 - res) + fc * (wm3 + wp3 - res);
      double y = sb - l1[i] * ym2 - l2[i] * ym1;
      S(i) = y;
      ym2 = ym1; ym1 = y;
      wm3 = wm2; wm2 = wm1; wm1 = w0; w0 = wp1; wp1 = wp2; wp2 = wp3;
    }
    {
      double y = re2 - l1[ie - 2] * ym2 - l2[ie - 2] * ym1; S(ie - 2) = y; ym2 = ym1; ym1 = y;
      y = re1 - l1[ie - 1] * ym2 - l2[ie - 1] * ym1; S(ie - 1) = y; ym2 = ym1; ym1 = y;
      y = re0 - l1[ie] * ym2 - l2[ie] * ym1; S(ie) = y;
    }
    double xp2 = u1[ie] * S(ie);
    S(ie) = xp2;
    double xp1 = u1[ie - 1] * S(ie - 1) - u2[ie - 1] * xp2;
    S(ie - 1) = xp1;
    for (int i = ie - 2; i >= 0; i--) {
      double x = u1[i] * S(i) - u2[i] * xp1 - u3[i] * xp2;
      S(i) = x;
      xp2 = xp1; xp1 = x;
    }
#undef S
  }
  __syncthreads();
  double *out = field_ptr(b, a.out_arr, a.out_col + f * a.out_step);
  tile_store<D, 1>(b, out, q0, nl, L, nlp, lds,
                   a.copy_arr >= 0 ? field_ptr(b, a.copy_arr, a.copy_col + f * a.out_step) : nullptr,
                   out != io ? io : nullptr);
}

// ---- closure sums received from the neighbour, Subroutines3D.f90:16-80 ----
// One thread per receiving face point; the sender's interior points are read
// straight from its field (no send buffer, no host round trip).  NT: 0 the
// derivative's closure (LMD + 1 points), 1 the filter's (LMF + 1).
template <int NT>
__global__ void __launch_bounds__(256) k_exchange(ExchArgs a) {
  const Work w = a.work[xcd_item(blockIdx.x, gridDim.x)];
  const int d = w.d & 0xff, ip = w.d >> 8;
  const Block &b = a.blocks[w.blk];
  int jk;
  if (!face_point(b, w, d, jk)) return;
  const Block &s = a.blocks[b.ncd[d][ip]];
  const int f = blockIdx.y;
  const int col = a.col0 + f * a.fstep + (1 - a.nrt) * d;
  double r0, r1, r2;
  closure_sums<NT>(b, s, a.coef[b.cs], field_ptr(s, a.arr, col), field_ptr(b, a.arr, col), d, ip, jk, a.n45, r0,
                   r1, r2);
  const int F = b.F[d];
  double *rcv = b.recv[d] + (size_t)f * 6 * F;
  rcv[(ip * 3 + 0) * F + jk] = r0;
  rcv[(ip * 3 + 1) * F + jk] = r1;
  rcv[(ip * 3 + 2) * F + jk] = r2;
}

// ---- host launchers ----
void launch_deriv(int d, int nwg, int nf, size_t lds, hipStream_t st, const LineArgs &a, hipEvent_t e0,
                  hipEvent_t e1) {
  if (d == 0) TP_LAUNCH_EV(k_deriv<0>, dim3(nwg, nf), dim3(64), lds, st, e0, e1, a);
  else if (d == 1) TP_LAUNCH_EV(k_deriv<1>, dim3(nwg, nf), dim3(64), lds, st, e0, e1, a);
  else TP_LAUNCH_EV(k_deriv<2>, dim3(nwg, nf), dim3(64), lds, st, e0, e1, a);
}

void launch_filter(int d, int nwg, int nf, size_t lds, hipStream_t st, const LineArgs &a, hipEvent_t e0,
                   hipEvent_t e1) {
  if (d == 0) TP_LAUNCH_EV(k_filter<0>, dim3(nwg, nf), dim3(64), lds, st, e0, e1, a);
  else if (d == 1) TP_LAUNCH_EV(k_filter<1>, dim3(nwg, nf), dim3(64), lds, st, e0, e1, a);
  else TP_LAUNCH_EV(k_filter<2>, dim3(nwg, nf), dim3(64), lds, st, e0, e1, a);
}

void launch_exchange(int nwg, int nf, hipStream_t st, const ExchArgs &a) {
  if (a.nt == 0) hipLaunchKernelGGL(k_exchange<0>, dim3(nwg, nf), dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_exchange<1>, dim3(nwg, nf), dim3(256), 0, st, a);
}

}  // namespace tp

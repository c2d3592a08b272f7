// tp_io.cpp -- the reference's file formats from device-resident state:
// restart files (src/rpt.f90:555-661, wrRsta/rdRsta), raw grid files
// (src/rpt.f90:666-696 rdGrid; written by src/Grid.f90:452-459), and the
// PLOT3D multiblock grid and solution files (src/rpt.f90:115-330,
// wrP3dS/wrP3dG).
//
// The reference writes one restart or raw grid file per parent block (its
// MPI-IO file views place every sub-block's part) and one PLOT3D file for all
// parent blocks.  Here the sub-blocks of a case are gathered into their
// parent blocks on the host (tp_get_parents: each sub-block is its own parent
// unless tp_set_parents says otherwise) and written with the same byte
// layout: native little-endian, Fortran order, file names relative to the
// run directory as the reference forms them.  Host code over the C ABI.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tandem_hip.h"
#include "tp_internal.hpp"

namespace {

using tp::set_error;

// the sub-blocks of a case and their parent blocks
struct Parents {
  int nb = 0, np = 0;
  std::vector<int32_t> dims, par, start, pdims;
  size_t points(int mb) const {
    return (size_t)(pdims[3 * mb] + 1) * (pdims[3 * mb + 1] + 1) * (pdims[3 * mb + 2] + 1);
  }
  size_t bpoints(int b) const {
    return (size_t)(dims[3 * b] + 1) * (dims[3 * b + 1] + 1) * (dims[3 * b + 2] + 1);
  }
};

int parents(tp_context *ctx, int32_t cs, Parents &P) {
  int64_t info[4];
  if (int rc = tp_case_info(ctx, cs, info)) return rc;
  P.nb = (int)info[2];
  P.dims.resize(3 * (size_t)P.nb);
  if (int rc = tp_case_dims(ctx, cs, P.dims.data())) return rc;
  int32_t np = 0;
  if (int rc = tp_get_parents(ctx, cs, &np, nullptr, nullptr, nullptr)) return rc;
  P.np = np;
  P.par.resize(P.nb);
  P.start.resize(3 * (size_t)P.nb);
  P.pdims.resize(3 * (size_t)np);
  return tp_get_parents(ctx, cs, &np, P.par.data(), P.start.data(), P.pdims.data());
}

// copy field arrays (nv columns) between a sub-block and its parent block,
// both a(0:lmx, 1:nv) in Fortran order
void scatter_gather(const Parents &P, int b, int nv, double *parent, double *sub, bool to_parent) {
  const int mb = P.par[b];
  const int I = P.pdims[3 * mb] + 1, J = P.pdims[3 * mb + 1] + 1, K = P.pdims[3 * mb + 2] + 1;
  const int nx = P.dims[3 * b] + 1, ny = P.dims[3 * b + 1] + 1, nz = P.dims[3 * b + 2] + 1;
  const int si = P.start[3 * b], sj = P.start[3 * b + 1], sk = P.start[3 * b + 2];
  const size_t N = (size_t)nx * ny * nz, NP = (size_t)I * J * K;
  for (int v = 0; v < nv; v++)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        double *pr = parent + v * NP + ((size_t)(k + sk) * J + (j + sj)) * I + si;
        double *sr = sub + v * N + ((size_t)k * ny + j) * nx;
        if (to_parent) std::memcpy(pr, sr, sizeof(double) * nx);
        else std::memcpy(sr, pr, sizeof(double) * nx);
      }
}

std::string zone3(int mb) {
  char b[8];
  std::snprintf(b, sizeof(b), "%03d", mb % 1000);
  return b;
}

int ensure_dir(const std::string &d) {
  struct stat st;
  if (stat(d.c_str(), &st) == 0)
    return S_ISDIR(st.st_mode) ? 0 : set_error(TP_E_ARG, d + " is not a directory");
  if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return set_error(TP_E_ARG, "cannot create " + d);
  return 0;
}

// Fortran "(f8.4)" of t with blanks turned into '0' (src/rpt.f90:175-180)
std::string ctime_f84(double t) {
  char b[32];
  std::snprintf(b, sizeof(b), "%8.4f", t);
  std::string s(b);
  if (s.size() > 8) s = std::string(8, '*');
  for (char &ch : s)
    if (ch == ' ') ch = '0';
  return s;
}

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

int write_at(FILE *f, long off, const void *p, size_t n) {
  if (std::fseek(f, off, SEEK_SET) != 0 || std::fwrite(p, 1, n, f) != n)
    return set_error(TP_E_ARG, "write failed");
  return 0;
}

int read_at(FILE *f, long off, void *p, size_t n) {
  if (std::fseek(f, off, SEEK_SET) != 0 || std::fread(p, 1, n, f) != n)
    return set_error(TP_E_SIZE, "file shorter than its layout");
  return 0;
}

// gather field `id` (nv columns) of every sub-block into its parent block;
// id = -1: the grid as set by tp_set_grid (the device's ss is stage scratch,
// as the reference's is once it has read its grid)
int gather(tp_context *ctx, int32_t cs, const Parents &P, int id, int nv, std::vector<std::vector<double>> &out) {
  out.assign(P.np, {});
  for (int mb = 0; mb < P.np; mb++) out[mb].assign(nv * P.points(mb), 0.0);
  std::vector<double> sub;
  for (int b = 0; b < P.nb; b++) {
    sub.resize(nv * P.bpoints(b));
    if (id < 0) {
      if (int rc = tp_get_grid(ctx, cs, b, sub.data(), (int64_t)sub.size())) return rc;
    } else {
      const int64_t n = tp_array_count(ctx, cs, b, id);
      if (n < (int64_t)sub.size()) return set_error(TP_E_STATE, "case is not on the device");
      std::vector<double> all(n);
      if (int rc = tp_download(ctx, cs, b, id, all.data(), n)) return rc;
      std::memcpy(sub.data(), all.data(), sizeof(double) * sub.size());
    }
    scatter_gather(P, b, nv, out[P.par[b]].data(), sub.data(), true);
  }
  return 0;
}

}  // namespace

extern "C" {

int32_t tp_write_restart(tp_context *ctx, int32_t cs, const char *rundir) {
  if (!ctx || !rundir) return set_error(TP_E_ARG, "tp_write_restart: null argument");
  Parents P;
  if (int rc = parents(ctx, cs, P)) return rc;
  double s[TP_S_NSCALAR];
  if (int rc = tp_get_scalars(ctx, cs, s)) return rc;
  std::vector<std::vector<double>> q;
  if (int rc = gather(ctx, cs, P, TP_QA, 5, q)) return rc;
  const std::string dir = std::string(rundir) + "/rsta";
  if (int rc = ensure_dir(rundir)) return rc;
  if (int rc = ensure_dir(dir)) return rc;
  for (int mb = 0; mb < P.np; mb++) {
    const std::string name = dir + "/restart" + zone3(mb) + ".dat";
    File F;
    F.f = std::fopen(name.c_str(), "wb");
    if (!F.f) return set_error(TP_E_ARG, "cannot open " + name);
    // header in REAL8-sized slots (src/rpt.f90:587-604): n, ndt as INTEGER4,
    // then dt, dts, dte, timo; the unwritten half of the integer slots is zero
    unsigned char head[48];
    std::memset(head, 0, sizeof(head));
    const int32_t n = (int32_t)s[TP_S_N], ndt = (int32_t)s[TP_S_NDT];
    std::memcpy(head + 0, &n, 4);
    std::memcpy(head + 8, &ndt, 4);
    std::memcpy(head + 16, &s[TP_S_DT], 8);
    std::memcpy(head + 24, &s[TP_S_DTS], 8);
    std::memcpy(head + 32, &s[TP_S_DTE], 8);
    std::memcpy(head + 40, &s[TP_S_TIMO], 8);
    if (int rc = write_at(F.f, 0, head, sizeof(head))) return rc;
    if (int rc = write_at(F.f, 48, q[mb].data(), sizeof(double) * q[mb].size())) return rc;
  }
  return 0;
}

int32_t tp_read_restart(tp_context *ctx, int32_t cs, const char *rundir) {
  if (!ctx || !rundir) return set_error(TP_E_ARG, "tp_read_restart: null argument");
  Parents P;
  if (int rc = parents(ctx, cs, P)) return rc;
  std::vector<std::vector<double>> q(P.np);
  int32_t n = 0, ndt = 0;
  double dts[4] = {0, 0, 0, 0};
  for (int mb = 0; mb < P.np; mb++) {
    const std::string name = std::string(rundir) + "/rsta/restart" + zone3(mb) + ".dat";
    File F;
    F.f = std::fopen(name.c_str(), "rb");
    if (!F.f) return set_error(TP_E_ARG, "cannot open " + name);
    unsigned char head[48];
    if (int rc = read_at(F.f, 0, head, sizeof(head))) return rc;
    std::memcpy(&n, head + 0, 4);
    std::memcpy(&ndt, head + 8, 4);
    std::memcpy(dts, head + 16, 32);
    q[mb].resize(5 * P.points(mb));
    if (int rc = read_at(F.f, 48, q[mb].data(), sizeof(double) * q[mb].size())) return rc;
    std::fseek(F.f, 0, SEEK_END);
    if (std::ftell(F.f) != (long)(48 + sizeof(double) * q[mb].size()))
      return set_error(TP_E_SIZE, name + " does not hold the parent block's 5 fields");
  }
  std::vector<double> sub;
  for (int b = 0; b < P.nb; b++) {
    sub.assign(5 * P.bpoints(b), 0.0);
    scatter_gather(P, b, 5, q[P.par[b]].data(), sub.data(), false);
    if (int rc = tp_upload(ctx, cs, b, TP_QA, sub.data(), (int64_t)sub.size())) return rc;
  }
  double s[TP_S_NSCALAR];
  if (int rc = tp_get_scalars(ctx, cs, s)) return rc;
  s[TP_S_N] = n; s[TP_S_NDT] = ndt;
  s[TP_S_DT] = dts[0]; s[TP_S_DTS] = dts[1]; s[TP_S_DTE] = dts[2]; s[TP_S_TIMO] = dts[3];
  return tp_set_scalars(ctx, cs, s);
}

int32_t tp_write_grid(tp_context *ctx, int32_t cs, const char *rundir) {
  if (!ctx || !rundir) return set_error(TP_E_ARG, "tp_write_grid: null argument");
  Parents P;
  if (int rc = parents(ctx, cs, P)) return rc;
  std::vector<std::vector<double>> g;
  if (int rc = gather(ctx, cs, P, -1, 3, g)) return rc;
  const std::string dir = std::string(rundir) + "/misc";
  if (int rc = ensure_dir(rundir)) return rc;
  if (int rc = ensure_dir(dir)) return rc;
  for (int mb = 0; mb < P.np; mb++) {
    const std::string name = dir + "/grid" + zone3(mb) + ".dat";
    File F;
    F.f = std::fopen(name.c_str(), "wb");
    if (!F.f) return set_error(TP_E_ARG, "cannot open " + name);
    if (int rc = write_at(F.f, 0, g[mb].data(), sizeof(double) * g[mb].size())) return rc;
  }
  return 0;
}

int32_t tp_read_grid(tp_context *ctx, int32_t cs, const char *rundir) {
  if (!ctx || !rundir) return set_error(TP_E_ARG, "tp_read_grid: null argument");
  Parents P;
  if (int rc = parents(ctx, cs, P)) return rc;
  std::vector<std::vector<double>> g(P.np);
  for (int mb = 0; mb < P.np; mb++) {
    const std::string name = std::string(rundir) + "/misc/grid" + zone3(mb) + ".dat";
    File F;
    F.f = std::fopen(name.c_str(), "rb");
    if (!F.f) return set_error(TP_E_ARG, "cannot open " + name);
    g[mb].resize(3 * P.points(mb));
    if (int rc = read_at(F.f, 0, g[mb].data(), sizeof(double) * g[mb].size())) return rc;
  }
  std::vector<double> sub;
  for (int b = 0; b < P.nb; b++) {
    sub.assign(3 * P.bpoints(b), 0.0);
    scatter_gather(P, b, 3, g[P.par[b]].data(), sub.data(), false);
    if (int rc = tp_set_grid(ctx, cs, b, sub.data())) return rc;
  }
  return 0;
}

int32_t tp_write_plot3d(tp_context *ctx, int32_t cs, const char *rundir, int32_t what) {
  if (!ctx || !rundir || what < 1 || what > 3) return set_error(TP_E_ARG, "tp_write_plot3d: bad arguments");
  Parents P;
  if (int rc = parents(ctx, cs, P)) return rc;
  const std::string dir = std::string(rundir) + "/out";
  if (int rc = ensure_dir(rundir)) return rc;
  if (int rc = ensure_dir(dir)) return rc;
  // header: number of blocks, then IMax, JMax, KMax of each (INTEGER4)
  std::vector<int32_t> head(1 + 3 * (size_t)P.np);
  head[0] = P.np;
  for (int mb = 0; mb < P.np; mb++)
    for (int d = 0; d < 3; d++) head[1 + 3 * mb + d] = P.pdims[3 * mb + d] + 1;
  if (what & 1) {
    // src/rpt.f90:265-330: REAL4 x, y, z of each block after the header
    std::vector<std::vector<double>> g;
    if (int rc = gather(ctx, cs, P, -1, 3, g)) return rc;
    const std::string name = dir + "/grid.xyz";
    File F;
    F.f = std::fopen(name.c_str(), "wb");
    if (!F.f) return set_error(TP_E_ARG, "cannot open " + name);
    if (int rc = write_at(F.f, 0, head.data(), 4 * head.size())) return rc;
    long lh = (long)head.size();
    for (int mb = 0; mb < P.np; mb++) {
      std::vector<float> x4(g[mb].begin(), g[mb].end());
      if (int rc = write_at(F.f, 4 * lh, x4.data(), 4 * x4.size())) return rc;
      lh += (long)x4.size();
    }
  }
  if (what & 2) {
    // src/rpt.f90:130-258: per block REAL4 amachoo, aoa, reoo, timo, then
    // rho, u + umf, v + umf, w + umf and p (the pressure array of the last
    // stage, as the reference writes it)
    double s[TP_S_NSCALAR], dp[TP_D_NPARAM];
    if (int rc = tp_get_scalars(ctx, cs, s)) return rc;
    if (int rc = tp_case_params(ctx, cs, nullptr, dp)) return rc;
    const double m1 = dp[TP_D_AMACH1], m2 = dp[TP_D_AMACH2], m3 = dp[TP_D_AMACH3];
    const float hb[4] = {(float)std::sqrt(m1 * m1 + m2 * m2 + m3 * m3), (float)s[TP_S_AOA],
                         (float)dp[TP_D_REOO], (float)s[TP_S_TIMO]};
    std::vector<std::vector<double>> q, pr;
    if (int rc = gather(ctx, cs, P, TP_QA, 5, q)) return rc;
    if (int rc = gather(ctx, cs, P, TP_P, 1, pr)) return rc;
    const std::string name = dir + "/solT" + ctime_f84(s[TP_S_TIMO]) + ".q";
    File F;
    F.f = std::fopen(name.c_str(), "wb");
    if (!F.f) return set_error(TP_E_ARG, "cannot open " + name);
    if (int rc = write_at(F.f, 0, head.data(), 4 * head.size())) return rc;
    long lh = (long)head.size();
    for (int mb = 0; mb < P.np; mb++) {
      const size_t NP = P.points(mb);
      std::vector<float> q4(5 * NP);
      const double *qa = q[mb].data();
      for (size_t l = 0; l < NP; l++) {
        q4[l] = (float)qa[l];
        for (int i = 1; i <= 3; i++) q4[i * NP + l] = (float)((qa[i * NP + l] / qa[l]) + s[TP_S_UMF1 + i - 1]);
        q4[4 * NP + l] = (float)pr[mb][l];
      }
      if (int rc = write_at(F.f, 4 * lh, hb, sizeof(hb))) return rc;
      if (int rc = write_at(F.f, 4 * (lh + 4), q4.data(), 4 * q4.size())) return rc;
      lh += 4 + (long)q4.size();
    }
  }
  return 0;
}

}  // extern "C"

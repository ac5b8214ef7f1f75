// kitamr_oracle.cpp -- CPU FP64 restatement of KitAMR.jl's per-step phase-space update.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path; never linked by the product).
//
// Every function below follows the reference function named in its comment, in the
// reference's own loop order: slope!  -> src/Flux/Slope.jl:20-1070
//                               flux!   -> src/Flux/Flux.jl:10-502, CAIDVM.jl, DVM.jl
//                               iterate!-> src/Theory/Iterate.jl:96-162
//                               KitCore -> lib/KitCore/{2D,3D,2D2F,3D1F}.jl
// Moment sums use Neumaier-compensated accumulation, so the oracle's own rounding error
// is far below the 1e-12 parity budget (the reference uses Julia's pairwise `sum`).
// Compiled with -ffp-contract=off: no fused multiply-adds, every operation is rounded
// as written, in the reference's association order.
#include "kitamr_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double EPS = 1e-12;                            // src/Abstract/Types.jl:3
constexpr double MACH_EPS = 2.220446049250313e-16;       // Julia eps(Float64)
constexpr double PI = 3.141592653589793;                 // Julia π as Float64

inline double jl_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }
inline double jl_min(double a, double b) {
  if (a != a || b != b) return std::numeric_limits<double>::quiet_NaN();
  return b < a ? b : a;
}
// minmod, src/Flux/Slope.jl:20-24
inline double minmod(double sL, double sR) {
  double SL = std::fabs(sL), SR = std::fabs(sR);
  return 0.5 * (jl_sign(sL) + jl_sign(sR)) * jl_min(SL, SR);
}

struct Neumaier {
  double s = 0.0, c = 0.0;
  void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x; else c += (x - t) + s;
    s = t;
  }
  double get() const { return s + c; }
};

}  // namespace

struct orc_ctx {
  int dim = 2, ndf = 2, flux_kind = 0, marching = 0, st_check = 100;
  double K = 1.0, gamma = 5.0 / 3.0, Pr = 2.0 / 3.0, mu = 0.0, omega = 0.5;
  int64_t nc = 0, ni = 0, np = 0;
  std::vector<int64_t> image_src;
  std::vector<double> mid, ds;
  std::vector<int32_t> benc, lvl;
  std::vector<int64_t> voff;
  std::vector<int8_t> vlev;
  std::vector<double> vmid, vw;
  bool vs_uniform = false;
  std::vector<int32_t> nbs;
  std::vector<int64_t> nboff, nbc;
  std::vector<int64_t> ph, pt;
  std::vector<int32_t> pd, pbc;
  std::vector<double> prot, parea, pmid;
  std::vector<int32_t> bck;
  std::vector<double> bcp;
  int vs_maxlevel = 0;
  int32_t vs_tn[3] = {1, 1, 1};
  double vs_lo[3] = {0, 0, 0}, vs_hi[3] = {0, 0, 0};
  std::vector<std::vector<int64_t>> cell_pairs;   // per cell: pair*2 + side, in pair order
  // immersed boundary (kah_topology's ib_* arrays) and the SolidNeighbor distributions
  std::vector<int64_t> ib_solid, ib_snb_off, ib_snb, ib_pair, ib_img_off, ib_img, ib_cut_off;
  std::vector<double> ib_aux, ib_nrm, ib_bc, ib_cut_gas, ib_cut_sol;
  std::vector<int32_t> ib_cut_idx;
  std::vector<int64_t> ib_of_pair;                 // IB face of each pair, or -1
  std::vector<double> ib_df;                       // [face][k][i], the SolidNeighbor df
  std::vector<int64_t> ib_df_off;                  // first velocity row of each face (donor grid)
  std::vector<double> ib_ssdf;                     // the limited boundary slope, as ib_df
  std::vector<std::vector<int64_t>> ib_faces_of;   // IB faces of each donor, in face-id order
  bool ib_done = false;
  // state
  std::vector<double> w, prim, qf, sw, flux;
  std::vector<double> df, sdf, vflux;
  std::string err;

  // Error message of a failing call; called from inside parallel regions, so the write is
  // serialized and the first message is kept.
  int fail(int code, const std::string& m) {
#pragma omp critical(orc_err)
    {
      if (err.empty()) err = m;
    }
    return code;
  }
  int NW() const { return dim + 2; }
  int64_t src(int64_t rec) const { return rec < nc ? rec : image_src[rec - nc]; }
  int64_t off(int64_t c) const { return voff[c]; }
  int64_t goff(int64_t c) const { return vs_uniform ? 0 : voff[c]; }  // grid arrays
  int64_t nv(int64_t c) const { return voff[c + 1] - voff[c]; }
  const double* rmid(int64_t rec) const { return &mid[rec * dim]; }
  const double* rds(int64_t rec) const { return &ds[rec * dim]; }
  int32_t rbenc(int64_t rec) const { return benc[src(rec)]; }
  double& F(int64_t c, int64_t i, int k) { return df[off(c) * ndf + k * nv(c) + i]; }
  double& S(int64_t c, int64_t i, int k, int d) {
    return sdf[off(c) * ndf * dim + (int64_t(d) * ndf + k) * nv(c) + i];
  }
  double& VF(int64_t c, int64_t i, int k) { return vflux[off(c) * ndf + k * nv(c) + i]; }
  double U(int64_t c, int64_t i, int d) const { return vmid[goff(c) * dim + d * nv(c) + i]; }
  double Wt(int64_t c, int64_t i) const { return vw[goff(c) + i]; }
  int L(int64_t c, int64_t i) const { return vlev[goff(c) + i]; }
  int32_t state(int64_t c, int f) const { return nbs[c * 2 * dim + f]; }
  int64_t nb_n(int64_t c, int f) const { return nboff[c * 2 * dim + f + 1] - nboff[c * 2 * dim + f]; }
  int64_t nb(int64_t c, int f, int64_t j) const { return nbc[nboff[c * 2 * dim + f] + j]; }
};

namespace {

template <class T>
void copy_in(std::vector<T>& v, const T* p, int64_t n) {
  v.assign(p ? p : nullptr, p ? p + n : nullptr);
  if (!p) v.assign(size_t(n), T());
}

// ---------------------------------------------------------------- KitCore (point forms)
// get_prim_2D / get_prim_3D, lib/KitCore/2D.jl:9-16, 3D.jl:12-20
void get_prim(int dim, const double* w, double g, double* p) {
  if (dim == 2) {
    p[0] = w[0];
    p[1] = w[1] / w[0];
    p[2] = w[2] / w[0];
    p[3] = 0.5 * w[0] / (g - 1.0) / (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
  } else {
    p[0] = w[0];
    p[1] = w[1] / w[0];
    p[2] = w[2] / w[0];
    p[3] = w[3] / w[0];
    p[4] = 0.5 * w[0] / (g - 1.0) /
           (w[4] - 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) / w[0]);
  }
}
// get_conserved_2D / 3D, lib/KitCore/2D.jl:1-8, 3D.jl:1-11
void get_conserved(int dim, const double* p, double g, double* w) {
  if (dim == 2) {
    w[0] = p[0];
    w[1] = p[0] * p[1];
    w[2] = p[0] * p[2];
    w[3] = 0.5 * p[0] / p[3] / (g - 1.0) + 0.5 * p[0] * (p[1] * p[1] + p[2] * p[2]);
  } else {
    w[0] = p[0];
    w[1] = p[0] * p[1];
    w[2] = p[0] * p[2];
    w[3] = p[0] * p[3];
    w[4] = 0.5 * p[0] / p[4] / (g - 1.0) +
           0.5 * p[0] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  }
}
// get_τ, src/Gas/Model.jl:14
double get_tau(int dim, const double* prim, double mu, double omega) {
  return mu * 2.0 * std::pow(prim[dim + 1], 1 - omega) / prim[0];
}
// discrete_maxwell_2D2F (lib/KitCore/2D2F.jl:15-21), discrete_maxwell_3D1F (3D1F.jl:14-23)
inline void maxwell(int dim, const double* u, const double* p, double K, double* out) {
  if (dim == 2) {
    double c2 = (u[0] - p[1]) * (u[0] - p[1]) + (u[1] - p[2]) * (u[1] - p[2]);
    double h = p[0] * (p[3] / PI) * std::exp(-p[3] * c2);
    out[0] = h;
    out[1] = h * K / (2.0 * p[3]);
  } else {
    double c2 = (u[0] - p[1]) * (u[0] - p[1]) + (u[1] - p[2]) * (u[1] - p[2]) +
                (u[2] - p[3]) * (u[2] - p[3]);
    out[0] = p[0] * std::pow(p[4] / PI, 1.5) * std::exp(-p[4] * c2);
  }
}
// shakhov_part_2D2F (2D2F.jl:22-44), shakhov_part_3D1F (3D1F.jl:24-40); F = Maxwellian
inline void shakhov(int dim, const double* u, const double* F, const double* p,
                    const double* q, double Pr, double K, double* out) {
  if (dim == 2) {
    double cu = u[0] - p[1], cv = u[1] - p[2];
    double c2 = cu * cu + cv * cv;
    double cq = cu * q[0] + cv * q[1];
    out[0] = 0.8 * (1 - Pr) * (p[3] * p[3]) / p[0] * cq * (2 * p[3] * c2 + K - 5) * F[0];
    out[1] = 0.8 * (1 - Pr) * (p[3] * p[3]) / p[0] * cq * (2 * p[3] * c2 + K - 3) * F[1];
  } else {
    double cu = u[0] - p[1], cv = u[1] - p[2], cw = u[2] - p[3];
    double c2 = cu * cu + cv * cv + cw * cw;
    double cq = cu * q[0] + cv * q[1] + cw * q[2];
    out[0] = 0.8 * (1 - Pr) * (p[4] * p[4]) / p[0] * cq * (2 * p[4] * c2 - 5) * F[0];
  }
}

// micro_to_macro_2D2F (2D2F.jl:119-126) / micro_to_macro_3D1F (3D1F.jl:109-123) over an
// index subset: values h[j], b[j] with velocities of rows rows[j] of cell c.
struct Moments {
  Neumaier m[5];
};

// ---------------------------------------------------------------- slope!
// diff_vs!, src/Flux/Slope.jl:29-64
void diff_vs(orc_ctx& o, int64_t c, int64_t n, double dsL, std::vector<double>& sL) {
  const int DIM = o.dim, NDF = o.ndf;
  int64_t nc_ = o.nv(c), nn = o.nv(n);
  int64_t index = 0;
  double flag = 0.0;
  for (int64_t i = 0; i < nc_; ++i) {
    int li = o.L(c, i);
    int ln = o.L(n, index);
    if (li == ln) {
      for (int j = 0; j < NDF; ++j) sL[j * nc_ + i] += (o.F(c, i, j) - o.F(n, index, j)) / dsL;
      index += 1;
    } else if (li < ln) {
      while (flag != 1.0) {
        int lnn = o.L(n, index);
        double den = std::ldexp(1.0, DIM * (lnn - li));
        for (int j = 0; j < NDF; ++j)
          sL[j * nc_ + i] += (o.F(c, i, j) - o.F(n, index, j)) / den / dsL;
        flag += 1 / den;
        index += 1;
      }
      flag = 0.0;
    } else {
      for (int j = 0; j < NDF; ++j) sL[j * nc_ + i] += (o.F(c, i, j) - o.F(n, index, j)) / dsL;
      flag += 1 / std::ldexp(1.0, DIM * (li - ln));
      if (flag == 1.0) {
        index += 1;
        flag = 0.0;
      }
    }
    (void)nn;
  }
}
// diff_vs_transverse!, src/Flux/Slope.jl:278-332
void diff_vs_transverse(orc_ctx& o, int64_t c, int64_t n, double dsL, const double* dm,
                        std::vector<double>& sL) {
  const int DIM = o.dim, NDF = o.ndf;
  int64_t nc_ = o.nv(c);
  int64_t index = 0;
  double flag = 0.0;
  auto proj = [&](int64_t idx, int j) {
    double p = o.F(n, idx, j);
    for (int t = 0; t < DIM; ++t) p += dm[t] * o.S(n, idx, j, t);
    return p;
  };
  for (int64_t i = 0; i < nc_; ++i) {
    int li = o.L(c, i);
    int ln = o.L(n, index);
    if (li == ln) {
      for (int j = 0; j < NDF; ++j) sL[j * nc_ + i] += (o.F(c, i, j) - proj(index, j)) / dsL;
      index += 1;
    } else if (li < ln) {
      while (flag != 1.0) {
        int lnn = o.L(n, index);
        double den = std::ldexp(1.0, DIM * (lnn - li));
        for (int j = 0; j < NDF; ++j)
          sL[j * nc_ + i] += (o.F(c, i, j) - proj(index, j)) / den / dsL;
        flag += 1 / den;
        index += 1;
      }
      flag = 0.0;
    } else {
      for (int j = 0; j < NDF; ++j) sL[j * nc_ + i] += (o.F(c, i, j) - proj(index, j)) / dsL;
      flag += 1 / std::ldexp(1.0, DIM * (li - ln));
      if (flag == 1.0) {
        index += 1;
        flag = 0.0;
      }
    }
  }
}

struct SlopeWs {
  std::vector<double> sL, sR;
};

// update_slope_bound_vs!, src/Flux/Slope.jl:68-86
void slope_bound(orc_ctx& o, int64_t c, int face, double ds, int dir, SlopeWs& ws) {
  int64_t n = o.nv(c);
  ws.sL.assign(size_t(n * o.ndf), 0.0);
  int64_t N = o.nb_n(c, face);
  for (int64_t j = 0; j < N; ++j) diff_vs(o, c, o.nb(c, face, j), ds, ws.sL);
  for (int k = 0; k < o.ndf; ++k)
    for (int64_t i = 0; i < n; ++i) o.S(c, i, k, dir) = ws.sL[k * n + i] / double(N);
}
// update_slope_inner_vs!, src/Flux/Slope.jl:90-116
void slope_inner(orc_ctx& o, int64_t c, int dir, double dsL, double dsR, SlopeWs& ws) {
  int64_t n = o.nv(c);
  int fL = 2 * dir, fR = 2 * dir + 1;
  ws.sL.assign(size_t(n * o.ndf), 0.0);
  ws.sR.assign(size_t(n * o.ndf), 0.0);
  int64_t nL = o.nb_n(c, fL), nR = o.nb_n(c, fR);
  for (int64_t j = 0; j < nL; ++j) diff_vs(o, c, o.nb(c, fL, j), dsL, ws.sL);
  for (int64_t j = 0; j < nR; ++j) diff_vs(o, c, o.nb(c, fR, j), dsR, ws.sR);
  for (int k = 0; k < o.ndf; ++k)
    for (int64_t i = 0; i < n; ++i)
      o.S(c, i, k, dir) = minmod(ws.sL[k * n + i] / double(nL), ws.sR[k * n + i] / double(nR));
}
// _ps_has_transverse_offset, src/Flux/Slope.jl:177-199
bool has_transverse_offset(const orc_ctx& o, int64_t c, int face, int dir) {
  if (o.dim == 1) return false;
  int64_t n = o.nb_n(c, face);
  if (n == 0) return false;
  for (int t = 0; t < o.dim; ++t) {
    if (t == dir) continue;
    double avg = 0.0;
    for (int64_t j = 0; j < n; ++j) avg += o.rmid(o.nb(c, face, j))[t];
    avg /= double(n);
    if (avg != o.rmid(c)[t]) return true;
  }
  return false;
}
// update_slope_inner_vs_transverse!, src/Flux/Slope.jl:339-383
void slope_inner_transverse(orc_ctx& o, int64_t c, int dir, double dsL, double dsR,
                            SlopeWs& ws) {
  int64_t n = o.nv(c);
  int fL = 2 * dir, fR = 2 * dir + 1;
  ws.sL.assign(size_t(n * o.ndf), 0.0);
  ws.sR.assign(size_t(n * o.ndf), 0.0);
  int64_t nL = o.nb_n(c, fL), nR = o.nb_n(c, fR);
  bool projL = has_transverse_offset(o, c, fL, dir);
  bool projR = has_transverse_offset(o, c, fR, dir);
  double dm[3];
  for (int64_t j = 0; j < nL; ++j) {
    int64_t nb = o.nb(c, fL, j);
    if (projL) {
      for (int t = 0; t < o.dim; ++t) dm[t] = (t == dir) ? 0.0 : (o.rmid(c)[t] - o.rmid(nb)[t]);
      diff_vs_transverse(o, c, nb, dsL, dm, ws.sL);
    } else {
      diff_vs(o, c, nb, dsL, ws.sL);
    }
  }
  for (int64_t j = 0; j < nR; ++j) {
    int64_t nb = o.nb(c, fR, j);
    if (projR) {
      for (int t = 0; t < o.dim; ++t) dm[t] = (t == dir) ? 0.0 : (o.rmid(c)[t] - o.rmid(nb)[t]);
      diff_vs_transverse(o, c, nb, dsR, dm, ws.sR);
    } else {
      diff_vs(o, c, nb, dsR, ws.sR);
    }
  }
  for (int k = 0; k < o.ndf; ++k)
    for (int64_t i = 0; i < n; ++i)
      o.S(c, i, k, dir) = minmod(ws.sL[k * n + i] / double(nL), ws.sR[k * n + i] / double(nR));
}
// update_slope_bound_vs_transverse!, src/Flux/Slope.jl:428-451
void slope_bound_transverse(orc_ctx& o, int64_t c, int face, double ds, int dir,
                            SlopeWs& ws) {
  int64_t n = o.nv(c);
  ws.sL.assign(size_t(n * o.ndf), 0.0);
  int64_t N = o.nb_n(c, face);
  double dm[3];
  for (int64_t j = 0; j < N; ++j) {
    int64_t nb = o.nb(c, face, j);
    for (int t = 0; t < o.dim; ++t) dm[t] = (t == dir) ? 0.0 : (o.rmid(c)[t] - o.rmid(nb)[t]);
    diff_vs_transverse(o, c, nb, ds, dm, ws.sL);
  }
  for (int k = 0; k < o.ndf; ++k)
    for (int64_t i = 0; i < n; ++i) o.S(c, i, k, dir) = ws.sL[k * n + i] / double(N);
}

inline bool is_fine(const orc_ctx& o, int s) { return s == (1 << (o.dim - 1)); }

// update_slope!(Val(sL),Val(sR),...) dispatch, src/Flux/Slope.jl:458-771
int slope_dispatch(orc_ctx& o, int64_t c, int dir, SlopeWs& ws) {
  int fL = 2 * dir, fR = 2 * dir + 1;
  int sL = o.state(c, fL), sR = o.state(c, fR);
  double ds = o.rds(c)[dir];
  const double xm = o.rmid(c)[dir];
  bool FL = is_fine(o, sL), FR = is_fine(o, sR);
  if (sL == 1 && sR == 1) {
    bool Ls = o.rbenc(o.nb(c, fL, 0)) < 0, Rs = o.rbenc(o.nb(c, fR, 0)) < 0;
    if (Ls && Rs) {
      for (int k = 0; k < o.ndf; ++k)
        for (int64_t i = 0; i < o.nv(c); ++i) o.S(c, i, k, dir) = 0.0;
    } else if (Ls) {
      slope_bound(o, c, fR, xm - o.rmid(o.nb(c, fR, 0))[dir], dir, ws);
    } else if (Rs) {
      slope_bound(o, c, fL, xm - o.rmid(o.nb(c, fL, 0))[dir], dir, ws);
    } else {
      slope_inner(o, c, dir, ds, -ds, ws);
    }
  } else if (sL == 1 && FR) {
    slope_inner(o, c, dir, ds, -0.75 * ds, ws);
  } else if (FL && sR == 1) {
    slope_inner(o, c, dir, 0.75 * ds, -ds, ws);
  } else if (FL && FR) {
    slope_inner(o, c, dir, 0.75 * ds, -0.75 * ds, ws);
  } else if (sL == -1 && sR == -1) {
    slope_inner(o, c, dir, 1.5 * ds, -1.5 * ds, ws);
  } else if (sL == 1 && sR == -1) {
    slope_inner(o, c, dir, ds, -1.5 * ds, ws);
  } else if (sL == -1 && sR == 1) {
    slope_inner(o, c, dir, 1.5 * ds, -ds, ws);
  } else if (FL && sR == -1) {
    slope_inner(o, c, dir, 0.75 * ds, -1.5 * ds, ws);
  } else if (sL == -1 && FR) {
    slope_inner(o, c, dir, 1.5 * ds, -0.75 * ds, ws);
  } else if (sL == 0 && sR != 0) {
    slope_bound(o, c, fR, xm - o.rmid(o.nb(c, fR, 0))[dir], dir, ws);
  } else if (sR == 0 && sL != 0) {
    slope_bound(o, c, fL, xm - o.rmid(o.nb(c, fL, 0))[dir], dir, ws);
  } else {
    return o.fail(KAH_EINVAL, "update_slope!: no method for neighbour states (" +
                                  std::to_string(sL) + "," + std::to_string(sR) + ")");
  }
  return 0;
}

// update_slope_transverse_level! per cell, src/Flux/Slope.jl:849-945
int slope_transverse_cell(orc_ctx& o, int64_t c, SlopeWs& ws) {
  for (int dir = 0; dir < o.dim; ++dir) {
    int sL = o.state(c, 2 * dir), sR = o.state(c, 2 * dir + 1);
    if (sL == -1 || sR == -1) continue;
    int rc = slope_dispatch(o, c, dir, ws);
    if (rc) return rc;
  }
  if (o.dim == 1) return 0;
  for (int dir = 0; dir < o.dim; ++dir) {
    int fL = 2 * dir, fR = 2 * dir + 1;
    int sL = o.state(c, fL), sR = o.state(c, fR);
    const double xm = o.rmid(c)[dir];
    if (sL != 0 && sR != 0) {
      if (!(sL == -1 || sR == -1)) continue;
      int64_t L0 = o.nb(c, fL, 0), R0 = o.nb(c, fR, 0);
      double dsL = xm - o.rmid(L0)[dir];
      double dsR = xm - o.rmid(R0)[dir];
      slope_inner_transverse(o, c, dir, dsL, dsR, ws);
      if (o.rbenc(L0) < 0)
        slope_bound_transverse(o, c, fR, dsR, dir, ws);
      else if (o.rbenc(R0) < 0)
        slope_bound_transverse(o, c, fL, dsL, dir, ws);
    } else if (sR == 0 && sL == -1) {
      int64_t L0 = o.nb(c, fL, 0);
      if (o.rbenc(L0) < 0) continue;
      slope_bound_transverse(o, c, fL, xm - o.rmid(L0)[dir], dir, ws);
    } else if (sL == 0 && sR == -1) {
      int64_t R0 = o.nb(c, fR, 0);
      if (o.rbenc(R0) < 0) continue;
      slope_bound_transverse(o, c, fR, xm - o.rmid(R0)[dir], dir, ws);
    }
  }
  return 0;
}

// calc_w0 over all cells of a slope/df layout: micro_to_macro_* over a full grid.
void moments_full(const orc_ctx& o, int64_t c, const double* a, int64_t stride_k,
                  double* wout) {
  // a[k*stride_k + i] holds the k-th distribution of velocity cell i
  int64_t n = o.nv(c);
  Neumaier m[5];
  for (int64_t i = 0; i < n; ++i) {
    double wt = o.Wt(c, i);
    if (o.dim == 2) {
      double u = o.U(c, i, 0), v = o.U(c, i, 1);
      double h = a[i], b = a[stride_k + i];
      m[0].add(wt * h);
      m[1].add(wt * u * h);
      m[2].add(wt * v * h);
      m[3].add(wt * ((u * u + v * v) * h + b));
    } else {
      double u = o.U(c, i, 0), v = o.U(c, i, 1), ww = o.U(c, i, 2);
      double f = a[i];
      m[0].add(wt * f);
      m[1].add(wt * u * f);
      m[2].add(wt * v * f);
      m[3].add(wt * ww * f);
      m[4].add(wt * (u * u + v * v + ww * ww) * f);
    }
  }
  if (o.dim == 2) {
    wout[0] = m[0].get(); wout[1] = m[1].get(); wout[2] = m[2].get();
    wout[3] = 0.5 * m[3].get();
  } else {
    wout[0] = m[0].get(); wout[1] = m[1].get(); wout[2] = m[2].get(); wout[3] = m[3].get();
    wout[4] = 0.5 * m[4].get();
  }
}

// heat_flux_2D2F (2D2F.jl:68-88) / heat_flux_3D1F (3D1F.jl:41-72) over a full grid
void heat_flux_full(const orc_ctx& o, int64_t c, const double* a, int64_t stride_k,
                    const double* p, double* q) {
  int64_t n = o.nv(c);
  if (o.dim == 2) {
    Neumaier s1, s2, s3, s4;
    for (int64_t i = 0; i < n; ++i) {
      double wt = o.Wt(c, i);
      double cu = o.U(c, i, 0) - p[1], cv = o.U(c, i, 1) - p[2];
      double c2 = cu * cu + cv * cv;
      double h = a[i], b = a[stride_k + i];
      s1.add(wt * cu * c2 * h);
      s2.add(wt * cu * b);
      s3.add(wt * cv * c2 * h);
      s4.add(wt * cv * b);
    }
    q[0] = 0.5 * (s1.get() + s2.get());
    q[1] = 0.5 * (s3.get() + s4.get());
  } else {
    Neumaier s[3];
    for (int64_t i = 0; i < n; ++i) {
      double wt = o.Wt(c, i);
      double cu = o.U(c, i, 0) - p[1], cv = o.U(c, i, 1) - p[2], cw = o.U(c, i, 2) - p[3];
      double c2 = cu * cu + cv * cv + cw * cw;
      double f = a[i];
      s[0].add(wt * cu * c2 * f);
      s[1].add(wt * cv * c2 * f);
      s[2].add(wt * cw * c2 * f);
    }
    for (int d = 0; d < 3; ++d) q[d] = 0.5 * s[d].get();
  }
}

// x = J^-1 g for a symmetric positive definite J (M <= 5) by Cholesky, L L^T = J
void cholesky_solve(int M, const double (&J)[5][5], const double* g, double* x) {
  double L[5][5] = {};
  for (int j = 0; j < M; ++j) {
    double s = J[j][j];
    for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
    L[j][j] = std::sqrt(s);
    for (int i = j + 1; i < M; ++i) {
      double t = J[j][i];
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
      L[i][j] = t / L[j][j];
    }
  }
  double y[5];
  for (int i = 0; i < M; ++i) {
    double t = g[i];
    for (int k = 0; k < i; ++k) t -= L[i][k] * y[k];
    y[i] = t / L[i][i];
  }
  for (int i = M - 1; i >= 0; --i) {
    double t = y[i];
    for (int k = i + 1; k < M; ++k) t -= L[k][i] * x[k];
    x[i] = t / L[i][i];
  }
}

// conserved_I_porjection! (src/Theory/I-projection.jl:144-158) of one cell: the h (2D2F) or f
// (3D1F) row a[0..n) is multiplied by exp(lambda.psi) with lambda from solve_I_projection
// (:55-141), so that its moments (with b's internal energy, 2D2F) equal w.
//   psi = (1, u, |u|^2/2) (build_psi_matrix, :7-21); negative entries are first shaved
//   (:74-84); Newton on the dual G(lambda) = <psi a exp(lambda.psi)> - W, at most 10 steps,
//   stop at |G| < 1e-10*max(1,|W|) or after two steps that reduce |G| by less than 10 %;
//   Newton step by a Cholesky solve of the symmetric J (the reference calls
//   Symmetric(J,:U) \ G: LAPACK's Bunch-Kaufman, equal up to rounding); Armijo backtracking
//   with the same sufficient-decrease inequality Phi(l + a*dl) <= Phi(l) + 1e-4*a*G.dl, its
//   left side evaluated as the difference sum(c_k*expm1(a*dl.psi_k)) - a*dl.W (the reference
//   subtracts two O(1) sums, whose rounding decides the test near convergence).
void cip_projection(const orc_ctx& o, int64_t c, double* a, int64_t n, const double* b,
                    const double* w) {
  const int D = o.dim, M = D + 2;
  double W[5];
  for (int q = 0; q < M; ++q) W[q] = w[q];
  if (D == 2) {   // e_int = dot(weight, df[:,2])/2 ; w_trans[end] -= e_int
    Neumaier e;
    for (int64_t i = 0; i < n; ++i) e.add(o.Wt(c, i) * b[i]);
    W[M - 1] -= e.get() / 2;
  }
  auto psi = [&](int64_t i, double* ps) {
    ps[0] = 1.0;
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
      const double x = o.U(c, i, d);
      ps[1 + d] = x;
      s += x * x;
    }
    ps[D + 1] = s / 2;
  };
  // shave negative distribution
  double mn = a[0];
  for (int64_t i = 1; i < n; ++i) mn = std::min(mn, a[i]);
  const double f_min = 1.1 * mn;
  if (f_min < 0) {
    double fp_min = HUGE_VAL;
    for (int64_t i = 0; i < n; ++i)
      if (a[i] > 0) fp_min = std::min(fp_min, a[i]);
    const double d = fp_min - f_min;
    for (int64_t i = 0; i < n; ++i)
      if (a[i] < 0) a[i] = (a[i] - f_min) / d * fp_min;
  }
  double lam[5] = {0, 0, 0, 0, 0};
  double wn = 0.0;
  for (int q = 0; q < M; ++q) wn += W[q] * W[q];
  const double tol_eff = 1e-10 * std::max(1.0, std::sqrt(wn));
  double g_prev = HUGE_VAL;
  int stall = 0;
  std::vector<double> ck(static_cast<size_t>(n));
  for (int it = 0; it < 10; ++it) {
    Neumaier Gs[5], Js[15];
    for (int64_t i = 0; i < n; ++i) {
      double ps[5];
      psi(i, ps);
      double dot = 0.0;
      for (int q = 0; q < M; ++q) dot += lam[q] * ps[q];
      const double cc = o.Wt(c, i) * a[i] * std::exp(dot);
      ck[size_t(i)] = cc;
      int t = 0;
      for (int p = 0; p < M; ++p) {
        const double ci = cc * ps[p];
        Gs[p].add(ci);
        for (int r = p; r < M; ++r) Js[t++].add(ci * ps[r]);
      }
    }
    double G[5], J[5][5], gn = 0.0;
    for (int p = 0; p < M; ++p) {
      G[p] = Gs[p].get() - W[p];
      gn += G[p] * G[p];
    }
    gn = std::sqrt(gn);
    if (gn < tol_eff) break;
    if (gn > 0.9 * g_prev) {
      if (++stall >= 2) break;
    } else {
      stall = 0;
    }
    g_prev = gn;
    int t = 0;
    for (int p = 0; p < M; ++p)
      for (int r = p; r < M; ++r) J[p][r] = J[r][p] = Js[t++].get();
    double dl[5];
    cholesky_solve(M, J, G, dl);
    double slope = 0.0, dW = 0.0;
    for (int p = 0; p < M; ++p) {
      dl[p] = -dl[p];
      slope += G[p] * dl[p];
      dW += dl[p] * W[p];
    }
    double alpha = 1.0;
    for (int tr = 0; tr < 10; ++tr) {
      Neumaier s;
      for (int64_t i = 0; i < n; ++i) {
        double ps[5];
        psi(i, ps);
        double e = 0.0;
        for (int q = 0; q < M; ++q) e += dl[q] * ps[q];
        s.add(ck[size_t(i)] * std::expm1(alpha * e));
      }
      if (s.get() - alpha * dW <= 1e-4 * alpha * slope) break;
      alpha *= 0.5;
    }
    for (int p = 0; p < M; ++p) lam[p] += alpha * dl[p];
  }
  for (int64_t i = 0; i < n; ++i) {
    double ps[5];
    psi(i, ps);
    double dot = 0.0;
    for (int q = 0; q < M; ++q) dot += lam[q] * ps[q];
    a[i] *= std::exp(dot);
  }
}

// ---------------------------------------------------------------- flux!
struct SideMicro {   // one side of a face: the upwind subset of a velocity grid
  int64_t cell = -1;             // data cell
  std::vector<int64_t> rows;     // velocity-cell indices in the subset, in grid order
  std::vector<double> micro;     // [k*rows.size() + j]
};

void side_moments(const orc_ctx& o, const SideMicro& s, double* fw) {
  // micro_to_macro(micro, mid[subset], weight[subset]) -- 2D2F.jl:119-126, 3D1F.jl:109-123
  int64_t m = int64_t(s.rows.size());
  Neumaier a[5];
  for (int64_t j = 0; j < m; ++j) {
    int64_t i = s.rows[j];
    double wt = o.Wt(s.cell, i);
    if (o.dim == 2) {
      double u = o.U(s.cell, i, 0), v = o.U(s.cell, i, 1);
      double h = s.micro[j], b = s.micro[m + j];
      a[0].add(wt * h);
      a[1].add(wt * u * h);
      a[2].add(wt * v * h);
      a[3].add(wt * ((u * u + v * v) * h + b));
    } else {
      double u = o.U(s.cell, i, 0), v = o.U(s.cell, i, 1), ww = o.U(s.cell, i, 2);
      double f = s.micro[j];
      a[0].add(wt * f);
      a[1].add(wt * u * f);
      a[2].add(wt * v * f);
      a[3].add(wt * ww * f);
      a[4].add(wt * (u * u + v * v + ww * ww) * f);
    }
  }
  for (int q = 0; q < o.dim + 1; ++q) fw[q] = a[q].get();
  fw[o.dim + 1] = 0.5 * a[o.dim + 1].get();
}

// positivity_preserving_reconstruct, src/Flux/CAIDVM.jl:127-141 (limited), or the plain
// df + dot(dx, sdf) reconstruction (unlimited); multiplied by the normal velocity.
void reconstruct(orc_ctx& o, SideMicro& s, const double* xc, const double* dsc,
                 const double* fmid, int dir, double dt, bool limited) {
  int64_t m = int64_t(s.rows.size());
  s.micro.assign(size_t(m * o.ndf), 0.0);
  for (int64_t j = 0; j < m; ++j) {
    int64_t i = s.rows[j];
    double dx[3];
    for (int k = 0; k < o.dim; ++k) dx[k] = fmid[k] - o.U(s.cell, i, k) * dt - xc[k];
    double vn = o.U(s.cell, i, dir);
    for (int k = 0; k < o.ndf; ++k) {
      double f = o.F(s.cell, i, k);
      double s_dx = 0.0;
      for (int t = 0; t < o.dim; ++t) s_dx = (t == 0) ? dx[t] * o.S(s.cell, i, k, t)
                                                      : s_dx + dx[t] * o.S(s.cell, i, k, t);
      if (limited) {
        double s_abs = 0.0;
        for (int t = 0; t < o.dim; ++t)
          s_abs = (t == 0) ? dsc[t] * std::fabs(o.S(s.cell, i, k, t))
                           : s_abs + dsc[t] * std::fabs(o.S(s.cell, i, k, t));
        double lim = jl_min(std::fabs((f - MACH_EPS) / (0.5 * s_abs + EPS)), 1.0);
        s.micro[k * m + j] = (f + lim * s_dx) * vn;
      } else {
        s.micro[k * m + j] = (f + s_dx) * vn;
      }
    }
  }
}
// df .* vn; fsrc: the df of a SolidNeighbor ([k][i] on the cell's grid) instead of the cell's
void plain_micro(orc_ctx& o, SideMicro& s, int dir, const double* fsrc = nullptr) {
  int64_t m = int64_t(s.rows.size());
  int64_t n = o.nv(s.cell);
  s.micro.assign(size_t(m * o.ndf), 0.0);
  for (int64_t j = 0; j < m; ++j) {
    int64_t i = s.rows[j];
    double vn = o.U(s.cell, i, dir);
    for (int k = 0; k < o.ndf; ++k)
      s.micro[k * m + j] = (fsrc ? fsrc[k * n + i] : o.F(s.cell, i, k)) * vn;
  }
}

// The there side of an IB face under the DVM flux (calc_flux, src/Flux/DVM.jl:90-92): the
// SolidNeighbor's df plus ndx . sdf, its only slope being the limited boundary slope along
// the face direction, ndx = x_face - u dt - x_solid, times the normal velocity.
void ib_dvm_micro(orc_ctx& o, SideMicro& s, int64_t k, const double* xs, const double* fm,
                  int dir, double dt) {
  const int D = o.dim, NDF = o.ndf;
  const int64_t m = int64_t(s.rows.size());
  const int64_t n = o.nv(s.cell);
  const double* f = &o.ib_df[size_t(o.ib_df_off[k]) * NDF];
  const double* sl = &o.ib_ssdf[size_t(o.ib_df_off[k]) * NDF];
  s.micro.assign(size_t(m * NDF), 0.0);
  for (int64_t j = 0; j < m; ++j) {
    const int64_t i = s.rows[j];
    double ndx[3];
    for (int t = 0; t < D; ++t) ndx[t] = fm[t] - o.U(s.cell, i, t) * dt - xs[t];
    const double vn = o.U(s.cell, i, dir);
    for (int q = 0; q < NDF; ++q) {
      const double sv = sl[q * n + i];
      double dot = ndx[0] * (dir == 0 ? sv : 0.0);
      for (int t = 1; t < D; ++t) dot += ndx[t] * (dir == t ? sv : 0.0);
      s.micro[q * m + j] = (f[q * n + i] + dot) * vn;
    }
  }
}

// update_micro_flux!, src/Flux/Flux.jl:151-281 (there local PsData, both branches)
void update_micro_flux(orc_ctx& o, const SideMicro& hs, const SideMicro& ts, double area,
                       int64_t h, int64_t t, bool there_solid, const std::vector<char>& heavi,
                       bool do_h, bool do_t) {
  // do_h / do_t select which side's accumulators this call updates; the walk is the same.
  const int DIM = o.dim, NDF = o.ndf;
  if (there_solid) do_t = false;
  int64_t mh = int64_t(hs.rows.size()), mt = int64_t(ts.rows.size());
  auto HM = [&](int64_t idx, int k) { return hs.micro[k * mh + idx] * area; };
  auto TM = [&](int64_t idx, int k) { return ts.micro[k * mt + idx] * area; };
  int64_t index = 0, j = 0, index_n = 0;
  double flag = 0.0;
  int64_t n = o.nv(h);
  for (int64_t i = 0; i < n; ++i) {
    int li = o.L(h, i);
    if (heavi[i]) {
      for (int k = 0; k < NDF; ++k) if (do_h) o.VF(h, i, k) += HM(index, k);
      int lj = o.L(t, j);
      if (li == lj) {
        if (do_t)
          for (int k = 0; k < NDF; ++k) o.VF(t, j, k) -= HM(index, k);
        j += 1;
      } else if (li < lj) {
        while (flag != 1.0) {
          int ljj = o.L(t, j);
          if (do_t)
            for (int k = 0; k < NDF; ++k) o.VF(t, j, k) -= HM(index, k);
          flag += 1 / std::ldexp(1.0, DIM * (ljj - li));
          j += 1;
        }
        flag = 0.0;
      } else {
        double den = std::ldexp(1.0, DIM * (li - lj));
        if (do_t)
          for (int k = 0; k < NDF; ++k) o.VF(t, j, k) -= HM(index, k) / den;
        flag += 1 / den;
        if (flag == 1.0) {
          j += 1;
          flag = 0.0;
        }
      }
      index += 1;
    } else {
      int lj = o.L(t, j);
      if (li == lj) {
        for (int k = 0; k < NDF; ++k) {
          if (do_t) o.VF(t, j, k) -= TM(index_n, k);
          if (do_h) o.VF(h, i, k) += TM(index_n, k);
        }
        j += 1;
        index_n += 1;
      } else if (li < lj) {
        while (flag != 1.0) {
          int ljj = o.L(t, j);
          double den = std::ldexp(1.0, DIM * (ljj - li));
          for (int k = 0; k < NDF; ++k) {
            if (do_t) o.VF(t, j, k) -= TM(index_n, k);
            if (do_h) o.VF(h, i, k) += TM(index_n, k) / den;
          }
          flag += 1 / den;
          j += 1;
          index_n += 1;
        }
        flag = 0.0;
      } else {
        for (int k = 0; k < NDF; ++k) if (do_h) o.VF(h, i, k) += TM(index_n, k);
        flag += 1 / std::ldexp(1.0, DIM * (li - lj));
        if (flag == 1.0) {
          if (do_t)
            for (int k = 0; k < NDF; ++k) o.VF(t, j, k) -= TM(index_n, k);
          j += 1;
          index_n += 1;
          flag = 0.0;
        }
      }
    }
  }
}

// calc_ρw, src/Theory/Math.jl:251-282 (wall density for zero net mass flux)
double calc_rho_w(const orc_ctx& o, const SideMicro& here_rec, const SideMicro& there,
                  const double* bc, int dir) {
  int64_t mh = int64_t(here_rec.rows.size());
  Neumaier SF, SG;
  for (int64_t j = 0; j < mh; ++j) {
    int64_t i = here_rec.rows[j];
    SF.add(o.Wt(here_rec.cell, i) * o.U(here_rec.cell, i, dir) * here_rec.micro[j]);
  }
  for (int64_t i : there.rows) {
    int64_t c = there.cell;
    double c2 = 0.0;
    for (int d = 0; d < o.dim; ++d) {
      double cu = o.U(c, i, d) - bc[1 + d];
      c2 = (d == 0) ? cu * cu : c2 + cu * cu;
    }
    SG.add(o.Wt(c, i) * o.U(c, i, dir) * std::exp(-bc[o.dim + 1] * c2));
  }
  double pref = (o.dim == 2) ? bc[3] / PI : std::pow(bc[4] / PI, 1.5);
  return -SF.get() / (pref * SG.get());
}

// The work of one face: the two upwind micro-flux sets and the macroscopic increment,
// computed independently of every other face and then applied to each side's
// accumulators (update_flux!/update_domain_flux!) in face-list order.
struct FaceWork {
  SideMicro hs, ts;
  std::vector<char> heavi;
  double fw[5] = {0, 0, 0, 0, 0};
  bool has_fw = false, domain = false, there_solid = false;
  int64_t h = -1, t = -1;
  double area = 0.0;
};

// flux!(F, face, ka) for one DomainFace, src/Flux/Flux.jl:19-24 with CAIDVM.jl:4-94 /
// DVM.jl:1-77 (the update_domain_flux! part, Flux.jl:64-83, is apply_face)
int domain_face(orc_ctx& o, int64_t p, double dt, FaceWork& W) {
  const int DIM = o.dim, NDF = o.ndf, NW = o.NW();
  int64_t h = o.ph[p];
  int dir = o.pd[p];
  double rot = o.prot[p], area = o.parea[p];
  const double* fm = &o.pmid[p * DIM];
  int kind = o.bck[o.pbc[p]];
  std::vector<double> bc(&o.bcp[o.pbc[p] * NW], &o.bcp[o.pbc[p] * NW] + NW);
  int64_t n = o.nv(h);
  W.domain = true;
  W.h = h;
  W.area = area;
  std::vector<char>& heavi = W.heavi;
  heavi.assign(static_cast<size_t>(n), 0);
  SideMicro& hs = W.hs;
  SideMicro& ts = W.ts;
  hs.cell = ts.cell = h;
  for (int64_t i = 0; i < n; ++i) {
    heavi[i] = (rot * o.U(h, i, dir) <= 0.0);
    if (heavi[i]) hs.rows.push_back(i); else ts.rows.push_back(i);
  }
  const double* xh = o.rmid(h);
  if (kind == KAH_BC_UNIFORM_OUTFLOW) {
    plain_micro(o, hs, dir);
    plain_micro(o, ts, dir);
  } else if (kind == KAH_BC_MAXWELLIAN || kind == KAH_BC_SUPERSONIC_INFLOW) {
    if (kind == KAH_BC_MAXWELLIAN && o.flux_kind == KAH_FLUX_DVM) {
      // src/Flux/DVM.jl:12 reads here_weight/there_weight that are never defined
      return o.fail(KAH_EUNSUPPORTED,
                    "UndefVarError: `here_weight` not defined (DVM Maxwellian domain flux)");
    }
    reconstruct(o, hs, xh, o.rds(h), fm, dir, dt, false);  // (df + dot(dx,sdf)) * vn
    if (kind == KAH_BC_MAXWELLIAN) {
      // calc_ρw receives the reconstructed df (before * vn): divide the vn back out by
      // recomputing it directly
      SideMicro rec = hs;
      int64_t mh = int64_t(hs.rows.size());
      for (int64_t j = 0; j < mh; ++j) {
        int64_t i = hs.rows[j];
        double dx[3];
        for (int k = 0; k < DIM; ++k) dx[k] = fm[k] - o.U(h, i, k) * dt - xh[k];
        for (int k = 0; k < NDF; ++k) {
          double s_dx = 0.0;
          for (int t = 0; t < DIM; ++t) s_dx = (t == 0) ? dx[t] * o.S(h, i, k, t)
                                                        : s_dx + dx[t] * o.S(h, i, k, t);
          rec.micro[k * mh + j] = o.F(h, i, k) + s_dx;
        }
      }
      bc[0] = calc_rho_w(o, rec, ts, bc.data(), dir);
    }
    int64_t mt = int64_t(ts.rows.size());
    ts.micro.assign(size_t(mt * NDF), 0.0);
    for (int64_t j = 0; j < mt; ++j) {
      int64_t i = ts.rows[j];
      double u[3] = {0, 0, 0}, M[2];
      for (int d = 0; d < DIM; ++d) u[d] = o.U(h, i, d);
      maxwell(DIM, u, bc.data(), o.K, M);
      for (int k = 0; k < NDF; ++k) ts.micro[k * mt + j] = M[k] * u[dir];
    }
  } else if (kind == KAH_BC_INTERPOLATED_OUTFLOW) {
    if (DIM != 2) {
      return o.fail(KAH_EUNSUPPORTED,
                    "InterpolatedOutflow is defined for DIM=2 only (src/Flux/CAIDVM.jl:71)");
    }
    double xt[3];
    for (int d = 0; d < DIM; ++d) xt[d] = 2.0 * fm[d] - xh[d];
    reconstruct(o, hs, xh, o.rds(h), fm, dir, dt, false);
    int64_t mt = int64_t(ts.rows.size());
    ts.micro.assign(size_t(mt * NDF), 0.0);
    for (int64_t j = 0; j < mt; ++j) {
      int64_t i = ts.rows[j];
      double ndx[3];
      for (int k = 0; k < DIM; ++k) ndx[k] = fm[k] - o.U(h, i, k) * dt - xt[k];
      for (int k = 0; k < NDF; ++k) {
        double tdf = o.F(h, i, k) + (xt[dir] - xh[dir]) * o.S(h, i, k, dir);
        double s_dx = 0.0;
        for (int t = 0; t < DIM; ++t) s_dx = (t == 0) ? ndx[t] * o.S(h, i, k, t)
                                                      : s_dx + ndx[t] * o.S(h, i, k, t);
        ts.micro[k * mt + j] = (tdf + s_dx) * o.U(h, i, dir);
      }
    }
  } else {
    return o.fail(KAH_EINVAL, "unknown domain boundary condition");
  }
  if (o.flux_kind == KAH_FLUX_CAIDVM) {
    double a[5], b[5];
    side_moments(o, hs, a);
    side_moments(o, ts, b);
    for (int q = 0; q < NW; ++q) W.fw[q] = area * (a[q] + b[q]);
    W.has_fw = true;
  }
  return 0;
}

// flux!(F, face::FullFace / FluxData, ka): make_face_vs (Flux.jl:362-424), calc_flux
// (CAIDVM.jl:99-120 / DVM.jl:79-99), update_flux! (Flux.jl:94-111)
int inner_face(orc_ctx& o, int64_t p, double dt, FaceWork& W) {
  const int DIM = o.dim, NW = o.NW();
  int64_t h = o.ph[p], trec = o.pt[p];
  int64_t t = o.src(trec);
  int dir = o.pd[p];
  double rot = o.prot[p], area = o.parea[p];
  const double* fm = &o.pmid[p * DIM];
  bool there_solid = o.benc[t] < 0;
  int64_t nh = o.nv(h), nt = o.nv(t);
  W.h = h;
  W.t = t;
  W.area = area;
  W.there_solid = there_solid;
  std::vector<char>& heavi = W.heavi;
  heavi.assign(static_cast<size_t>(nh), 0);
  SideMicro& hs = W.hs;
  SideMicro& ts = W.ts;
  hs.cell = h;
  ts.cell = t;
  for (int64_t i = 0; i < nh; ++i) {
    heavi[i] = (rot * o.U(h, i, dir) <= 0.0);
    if (heavi[i]) hs.rows.push_back(i);
  }
  // IB face: the there side is the SolidNeighbor, which carries the donor's grid
  // (initialize_solid_neighbor!, src/Boundary/Immersed_boundary.jl:264-273)
  const int64_t ibk = (there_solid && !o.ib_of_pair.empty()) ? o.ib_of_pair[p] : -1;
  if (ibk >= 0) {
    t = h;
    nt = nh;
    W.t = h;   // the micro-flux walk is then the identity; nothing is added to the solid side
    ts.cell = h;
  }
  for (int64_t j = 0; j < nt; ++j)
    if (rot * o.U(ts.cell, j, dir) > 0.0) ts.rows.push_back(j);
  if (there_solid) {
    reconstruct(o, hs, o.rmid(h), o.rds(h), fm, dir, dt, false);
    if (ibk >= 0 && o.flux_kind == KAH_FLUX_CAIDVM)   // CAIDVM.jl:110-112
      plain_micro(o, ts, dir, &o.ib_df[size_t(o.ib_df_off[ibk]) * o.ndf]);
    else if (ibk >= 0)   // DVM.jl:90-92: the SolidNeighbor's df with its slope (ssdf along dir)
      ib_dvm_micro(o, ts, ibk, o.rmid(trec), fm, dir, dt);
    else if (o.flux_kind == KAH_FLUX_CAIDVM)
      plain_micro(o, ts, dir);
    else
      reconstruct(o, ts, o.rmid(trec), o.rds(trec), fm, dir, dt, false);
  } else {
    reconstruct(o, hs, o.rmid(h), o.rds(h), fm, dir, dt, true);
    reconstruct(o, ts, o.rmid(trec), o.rds(trec), fm, dir, dt, true);
  }
  if (o.flux_kind == KAH_FLUX_CAIDVM) {
    double a[5], b[5];
    side_moments(o, hs, a);
    side_moments(o, ts, b);
    for (int q = 0; q < NW; ++q) W.fw[q] = (a[q] + b[q]) * area;
    W.has_fw = true;
  }
  return 0;
}

// update_flux! / update_domain_flux! (src/Flux/Flux.jl:64-139) restricted to one side of the
// face: do_h updates the here cell, do_t the there cell.
void apply_face(orc_ctx& o, const FaceWork& W, bool do_h, bool do_t) {
  const int NW = o.NW(), NDF = o.ndf;
  if (W.domain) {
    if (!do_h) return;
    if (W.has_fw)
      for (int q = 0; q < NW; ++q) o.flux[W.h * NW + q] += W.fw[q];
    int64_t mh = int64_t(W.hs.rows.size()), mt = int64_t(W.ts.rows.size());
    for (int64_t j = 0; j < mh; ++j)
      for (int k = 0; k < NDF; ++k) o.VF(W.h, W.hs.rows[j], k) += W.area * W.hs.micro[k * mh + j];
    for (int64_t j = 0; j < mt; ++j)
      for (int k = 0; k < NDF; ++k) o.VF(W.h, W.ts.rows[j], k) += W.area * W.ts.micro[k * mt + j];
    return;
  }
  if (W.has_fw) {
    if (do_h)
      for (int q = 0; q < NW; ++q) o.flux[W.h * NW + q] += W.fw[q];
    if (do_t && !W.there_solid)
      for (int q = 0; q < NW; ++q) o.flux[W.t * NW + q] -= W.fw[q];
  }
  update_micro_flux(o, W.hs, W.ts, W.area, W.h, W.t, W.there_solid, W.heavi, do_h, do_t);
}


// ---------------------------------------------------------------- immersed boundary
// Julia's norm / dot of short vectors: plain sums of products, in index order
double vnorm(int dim, const double* x) {
  double a = x[0] * x[0];
  for (int t = 1; t < dim; ++t) a += x[t] * x[t];
  return std::sqrt(a);
}
double vdot(int dim, const double* x, const double* y) {
  double a = x[0] * y[0];
  for (int t = 1; t < dim; ++t) a += x[t] * y[t];
  return a;
}

// Direction-weighted extrapolation of the listed fluid cells' f to the point x, onto the
// velocity grid of cell tc: out[k*n + i] = sum_j wn_j(i) (f_j + sdf_j . (x - x_j)) over the
// source cells j, with wn_j = w_j / sum w, w_j = max(0, u.l_j/|u|)^2, l_j = (x - x_j)/|x - x_j|
// (1/N when sum w = 0), u the target's velocity midpoint. Each source is walked against the
// target grid as vs_extrapolate! does (same level: add; target coarser: the mean of the
// finer source cells; target finer: injection). update_solid_cell! / image_df
// (src/Boundary/Immersed_boundary.jl:24-47, 122-141, 378-395).
void extrapolate(orc_ctx& o, const int64_t* cells, int64_t N, const double* x, int64_t tc,
                 double* out) {
  const int D = o.dim, NDF = o.ndf;
  const int64_t n = o.nv(tc);
  std::vector<double> w(static_cast<size_t>(N * n)), ws(static_cast<size_t>(n), 0.0);
  for (int64_t i = 0; i < n; ++i) {
    double u[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < D; ++t) u[t] = o.U(tc, i, t);
    const double nu = vnorm(D, u);
    for (int64_t j = 0; j < N; ++j) {
      double l[3] = {0.0, 0.0, 0.0};
      const double* xj = o.rmid(cells[j]);
      for (int t = 0; t < D; ++t) l[t] = x[t] - xj[t];
      const double nl = vnorm(D, l);
      for (int t = 0; t < D; ++t) l[t] = l[t] / nl;
      const double r = vdot(D, u, l) / nu;
      const double c = (r != r) ? r : (r > 0.0 ? r : 0.0);   // max(0., r)
      w[j * n + i] = c * c;
      ws[i] = (j == 0) ? w[j * n + i] : ws[i] + w[j * n + i];
    }
  }
  for (int64_t x0 = 0; x0 < n * NDF; ++x0) out[x0] = 0.0;
  for (int64_t j = 0; j < N; ++j) {
    const int64_t sc = cells[j];
    double dx[3] = {0.0, 0.0, 0.0};
    const double* xj = o.rmid(sc);
    for (int t = 0; t < D; ++t) dx[t] = x[t] - xj[t];
    auto term = [&](int64_t jj, int k) {   // df + ddf of source velocity cell jj
      double ddf = o.S(sc, jj, k, 0) * dx[0];
      for (int t = 1; t < D; ++t) ddf += o.S(sc, jj, k, t) * dx[t];
      return o.F(sc, jj, k) + ddf;
    };
    int64_t jj = 0;
    double flag = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double wn = ws[i] == 0.0 ? 1.0 / double(N) : w[j * n + i] / ws[i];
      const int lt = o.L(tc, i);
      const int ls = o.L(sc, jj);
      if (lt == ls) {
        for (int k = 0; k < NDF; ++k) out[k * n + i] += term(jj, k) * wn;
        jj += 1;
      } else if (lt < ls) {
        while (flag != 1.0) {
          const double den = std::ldexp(1.0, D * (o.L(sc, jj) - lt));
          for (int k = 0; k < NDF; ++k) out[k * n + i] += term(jj, k) / den * wn;
          flag += 1 / den;
          jj += 1;
        }
        flag = 0.0;
      } else {
        for (int k = 0; k < NDF; ++k) out[k * n + i] += term(jj, k) * wn;
        flag += 1 / std::ldexp(1.0, D * (lt - ls));
        if (flag == 1.0) {
          jj += 1;
          flag = 0.0;
        }
      }
    }
  }
}

// boundary_slope!'s one-sided difference (src/Boundary/Immersed_boundary.jl:396-426): the
// donor's f minus the solid cell's, walked across the two grids, over dxs
void boundary_diff(orc_ctx& o, int64_t c, int64_t s, double dxs, double* sL) {
  const int D = o.dim, NDF = o.ndf;
  const int64_t n = o.nv(c);
  for (int64_t x0 = 0; x0 < n * NDF; ++x0) sL[x0] = 0.0;
  int64_t idx = 0;
  double flag = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const int l = o.L(c, i), ln = o.L(s, idx);
    if (l == ln) {
      for (int k = 0; k < NDF; ++k) sL[k * n + i] = (o.F(c, i, k) - o.F(s, idx, k)) / dxs;
      idx += 1;
    } else if (l < ln) {
      while (flag != 1.0) {
        const double den = std::ldexp(1.0, D * (o.L(s, idx) - l));
        for (int k = 0; k < NDF; ++k)
          sL[k * n + i] += (o.F(c, i, k) - o.F(s, idx, k)) / den / dxs;
        flag += 1 / den;
        idx += 1;
      }
      flag = 0.0;
    } else {
      for (int k = 0; k < NDF; ++k) sL[k * n + i] += (o.F(c, i, k) - o.F(s, idx, k)) / dxs;
      flag += 1 / std::ldexp(1.0, D * (l - ln));
      if (flag == 1.0) {
        idx += 1;
        flag = 0.0;
      }
    }
  }
}

// update_solid_cell! (src/Boundary/Immersed_boundary.jl:122-148): f of each solid cell from
// its fluid face neighbours, then w and prim
void ib_solid_cells(orc_ctx& o) {
  const int NDF = o.ndf, NW = o.NW();
  const int64_t ns = int64_t(o.ib_solid.size());
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t q = 0; q < ns; ++q) {
    const int64_t s = o.ib_solid[q];
    const int64_t* nb = &o.ib_snb[o.ib_snb_off[q]];
    const int64_t N = o.ib_snb_off[q + 1] - o.ib_snb_off[q];
    const int64_t n = o.nv(s);
    std::vector<double> f(static_cast<size_t>(n * NDF));
    extrapolate(o, nb, N, o.rmid(s), s, f.data());
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < NDF; ++k) o.F(s, i, k) = f[k * n + i];
    double wv[5];
    moments_full(o, s, &o.df[o.off(s) * NDF], n, wv);
    for (int m = 0; m < NW; ++m) o.w[s * NW + m] = wv[m];
    get_prim(o.dim, &o.w[s * NW], o.gamma, &o.prim[s * NW]);
  }
}

// update_solid_neighbor! (src/Boundary/Immersed_boundary.jl:436-483) with image_df (:367-395),
// boundary_slope! (:396-435), the positivity scaling of the boundary slope, the cut-cell
// corrections cvc_gas_correction! / cvc_Mu / cvc_density / cvc_correction! (:332-366) and
// the wall Maxwellian of zero net mass flux: the df of each SolidNeighbor.
void ib_solid_neighbors(orc_ctx& o) {
  const int D = o.dim, NDF = o.ndf;
  const int64_t nf = int64_t(o.ib_pair.size());
#pragma omp parallel for schedule(dynamic, 2)
  for (int64_t k = 0; k < nf; ++k) {
    const int64_t p = o.ib_pair[k];
    const int64_t c = o.ph[p], s = o.src(o.pt[p]);
    const int dir = o.pd[p];
    const int64_t n = o.nv(c);
    const double* ap = &o.ib_aux[k * D];
    const double* nr = &o.ib_nrm[k * D];
    const double* xc = o.rmid(c);
    const double* xs = o.rmid(s);
    double ip[3];
    for (int t = 0; t < D; ++t) ip[t] = ap[t] + xc[t] - xs[t];
    const int64_t* img = &o.ib_img[o.ib_img_off[k]];
    const int64_t N = o.ib_img_off[k + 1] - o.ib_img_off[k];
    const double dxs = xc[dir] - xs[dir];
    const double dxf = ip[dir] - xc[dir];
    const double dxL = ap[dir] - ip[dir];
    double* aux = &o.ib_df[size_t(o.ib_df_off[k]) * NDF];
    std::vector<double> vn(static_cast<size_t>(n)), M(static_cast<size_t>(n * NDF));
    std::vector<double> img_f(static_cast<size_t>(n * NDF)), sLv(static_cast<size_t>(n * NDF));
    std::vector<char> th(static_cast<size_t>(n));
    double pw[5];
    for (int m = 0; m < D + 2; ++m) pw[m] = o.ib_bc[k * (D + 2) + m];
    pw[0] = 1.0;
    extrapolate(o, img, N, ip, c, img_f.data());
    boundary_diff(o, c, s, dxs, sLv.data());
    for (int64_t i = 0; i < n; ++i) {
      double u[3];
      for (int t = 0; t < D; ++t) u[t] = o.U(c, i, t);
      for (int q = 0; q < NDF; ++q) {
        const double fc = o.F(c, i, q);
        const double im = img_f[q * n + i];
        double sl = minmod(sLv[q * n + i], (im - fc) / dxf);
        sl = jl_min(std::fabs((im - MACH_EPS) / (sl * dxL + MACH_EPS)), 1.0) * sl;
        aux[q * n + i] = im + sl * dxL;
        o.ib_ssdf[size_t(o.ib_df_off[k]) * NDF + q * n + i] = sl;
      }
      vn[i] = vdot(D, u, nr);
      th[i] = vn[i] >= 0.0;
      double m2[2];
      maxwell(D, u, pw, o.K, m2);
      for (int q = 0; q < NDF; ++q) M[q * n + i] = m2[q];
    }
    const int64_t c0 = o.ib_cut_off[k], c1 = o.ib_cut_off[k + 1];
    std::vector<char> cut(static_cast<size_t>(n), 0);
    for (int64_t j = c0; j < c1; ++j) cut[o.ib_cut_idx[j]] = 1;
    Neumaier sf, mur;
    for (int64_t i = 0; i < n; ++i) {
      if (cut[i]) continue;   // cvc.weight is zero on the cut cells
      const double wv = o.Wt(c, i) * vn[i];
      if (th[i]) mur.add(wv * M[i]);
      else sf.add(wv * aux[i]);
    }
    double SF = sf.get(), MuR = mur.get();
    std::vector<double> gas(static_cast<size_t>((c1 - c0) * NDF));
    for (int64_t j = c0; j < c1; ++j) {
      const int64_t i = o.ib_cut_idx[j];
      MuR += o.ib_cut_sol[j] * vn[i] * M[i];
      SF += o.ib_cut_gas[j] * vn[i] * aux[i];
      for (int q = 0; q < NDF; ++q) gas[(j - c0) * NDF + q] = aux[q * n + i];
    }
    const double rho = -SF / MuR;
    for (int64_t i = 0; i < n * NDF; ++i) M[i] *= rho;
    for (int64_t i = 0; i < n; ++i)
      if (th[i])
        for (int q = 0; q < NDF; ++q) aux[q * n + i] = M[q * n + i];
    for (int64_t j = c0; j < c1; ++j) {
      const int64_t i = o.ib_cut_idx[j];
      const double g = o.ib_cut_gas[j], sw = o.ib_cut_sol[j];
      for (int q = 0; q < NDF; ++q)
        aux[q * n + i] = (g * gas[(j - c0) * NDF + q] + sw * M[q * n + i]) / (g + sw);
    }
  }
}

// positivity_preserving_ib! (src/Boundary/Positivity.jl:1-43) for a donor cell, after its w
// took the macroscopic flux: the micro flux through each IB face of the reconstruction
// perturbance (SolidNeighbor.vs_data.flux = ssdf*(x_s - aux)[dir], with the boundary slope
// ssdf as the SolidNeighbor's only slope) over rot*u[dir] > 0, its moments we*dt/V, the
// limiter theta that keeps density and internal energy positive, then w += theta*we and the
// velocity flux += theta*micro.
void ib_positivity(orc_ctx& o, int64_t c, double* w, double* vfl, double dt, double volume) {
  const int D = o.dim, NDF = o.ndf, NW = o.NW();
  if (o.ib_faces_of.empty() || o.ib_faces_of[size_t(c)].empty()) return;
  const auto& faces = o.ib_faces_of[size_t(c)];
  const int64_t n = o.nv(c);
  std::vector<double> micros(faces.size() * size_t(NDF * n), 0.0);
  double we[5] = {0, 0, 0, 0, 0};
  for (size_t a = 0; a < faces.size(); ++a) {
    const int64_t k = faces[a];
    const int64_t p = o.ib_pair[k];
    const int dir = o.pd[p];
    const double rot = o.prot[p], area = o.parea[p];
    const double* fm = &o.pmid[p * D];
    const double* xs = o.rmid(o.src(o.pt[p]));
    const double apd = o.ib_aux[k * D + dir];
    const double* ss = &o.ib_ssdf[size_t(o.ib_df_off[k]) * NDF];
    double* mi = &micros[a * size_t(NDF * n)];
    Neumaier acc[5];
    for (int64_t i = 0; i < n; ++i) {
      double u[3] = {0, 0, 0};
      for (int t = 0; t < D; ++t) u[t] = o.U(c, i, t);
      if (!(rot * u[dir] > 0.0)) continue;
      double ndx[3];
      for (int t = 0; t < D; ++t) ndx[t] = fm[t] - u[t] * dt - xs[t];
      for (int q = 0; q < NDF; ++q) {
        const double sl = ss[q * n + i];
        const double pert = sl * (xs[dir] - apd);
        double dot = ndx[0] * (dir == 0 ? sl : 0.0);
        for (int t = 1; t < D; ++t) dot += ndx[t] * (dir == t ? sl : 0.0);
        mi[q * n + i] = (pert + dot) * u[dir] * area;
      }
      const double wt = o.Wt(c, i);
      if (D == 2) {
        const double h = mi[i], b = mi[n + i];
        acc[0].add(wt * h);
        acc[1].add(wt * u[0] * h);
        acc[2].add(wt * u[1] * h);
        acc[3].add(wt * ((u[0] * u[0] + u[1] * u[1]) * h + b));
      } else {
        const double f = mi[i];
        acc[0].add(wt * f);
        acc[1].add(wt * u[0] * f);
        acc[2].add(wt * u[1] * f);
        acc[3].add(wt * u[2] * f);
        acc[4].add(wt * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * f);
      }
    }
    for (int q = 0; q < NW; ++q)
      we[q] += (q == NW - 1) ? 0.5 * acc[q].get() : acc[q].get();
  }
  const double dtv = dt / volume;
  for (int q = 0; q < NW; ++q) we[q] *= dtv;
  const double delta = 1e-3;
  const double th_rho = we[0] > 0.0 ? 1.0
                                    : jl_min(1.0, (1.0 - delta) * w[0] / (std::fabs(we[0]) + MACH_EPS));
  double sb = w[1] * w[1], se = we[1] * we[1], sd = w[1] * we[1];
  for (int t = 2; t <= D; ++t) {
    sb += w[t] * w[t];
    se += we[t] * we[t];
    sd += w[t] * we[t];
  }
  const double eb = w[NW - 1] - sb / (2.0 * w[0]);
  const double ee = we[NW - 1] - sd / w[0];
  const double gam = se / (2.0 * w[0]);
  const double th_e = jl_min(1.0, 2.0 * (1.0 - delta) * eb /
                                      (std::sqrt(ee * ee + 4.0 * gam * (1.0 - delta) * eb) - ee +
                                       MACH_EPS));
  const double th = jl_min(th_rho, th_e);
  for (int q = 0; q < NW; ++q) w[q] += th * we[q];
  for (size_t a = 0; a < faces.size(); ++a)
    for (int64_t x = 0; x < NDF * n; ++x) vfl[x] += th * micros[a * size_t(NDF * n) + x];
}

}  // namespace

// ===================================================================== C API
extern "C" {

orc_ctx* orc_create(const kah_params* P, const kah_topology* T) {
  if (!P || !T) return nullptr;
  orc_ctx* o = new orc_ctx();
  o->dim = P->dim;
  o->ndf = P->ndf;
  o->flux_kind = P->flux_kind;
  o->marching = P->marching;
  o->st_check = P->st_check_interval > 0 ? P->st_check_interval : 100;
  o->K = P->K;
  o->gamma = P->gamma;
  o->Pr = P->Pr;
  o->mu = P->mu_ref;
  o->omega = P->omega;
  const int D = o->dim;
  o->nc = T->n_cells;
  o->ni = T->n_images;
  int64_t nr = o->nc + o->ni;
  copy_in(o->image_src, T->image_src, o->ni);
  copy_in(o->mid, T->cell_mid, nr * D);
  copy_in(o->ds, T->cell_ds, nr * D);
  copy_in(o->benc, T->bound_enc, o->nc);
  copy_in(o->lvl, T->ps_level, o->nc);
  copy_in(o->voff, T->vs_offset, o->nc + 1);
  o->np = o->voff[o->nc];
  o->vs_uniform = T->vs_uniform != 0;
  int64_t ng = o->vs_uniform ? (o->nc > 0 ? o->voff[1] : 0) : o->np;
  if (o->vs_uniform)
    for (int64_t c = 0; c <= o->nc; ++c)
      if (o->voff[c] != c * ng) { delete o; return nullptr; }
  copy_in(o->vlev, T->vs_level, ng);
  copy_in(o->vmid, T->vs_mid, ng * D);
  copy_in(o->vw, T->vs_weight, ng);
  copy_in(o->nbs, T->nb_state, o->nc * 2 * D);
  copy_in(o->nboff, T->nb_offset, o->nc * 2 * D + 1);
  copy_in(o->nbc, T->nb_cell, o->nboff[o->nc * 2 * D]);
  int64_t npair = T->n_pairs;
  copy_in(o->ph, T->pair_here, npair);
  copy_in(o->pt, T->pair_there, npair);
  copy_in(o->pd, T->pair_dir, npair);
  copy_in(o->prot, T->pair_rot, npair);
  copy_in(o->parea, T->pair_area, npair);
  copy_in(o->pmid, T->pair_mid, npair * D);
  copy_in(o->pbc, T->pair_bc, npair);
  copy_in(o->bck, T->bc_kind, T->n_bc);
  copy_in(o->bcp, T->bc_prim, int64_t(T->n_bc) * (D + 2));
  o->vs_maxlevel = T->vs_maxlevel;
  for (int d = 0; d < 3; ++d) {
    o->vs_tn[d] = T->vs_trees_num[d];
    o->vs_lo[d] = T->vs_lo[d];
    o->vs_hi[d] = T->vs_hi[d];
  }
  if (T->n_ib_faces > 0 || T->n_ib_solid > 0) {
    const int64_t ns = T->n_ib_solid, nf = T->n_ib_faces;
    copy_in(o->ib_solid, T->ib_solid, ns);
    copy_in(o->ib_snb_off, T->ib_solid_nb_off, ns + 1);
    copy_in(o->ib_snb, T->ib_solid_nb, o->ib_snb_off[ns]);
    copy_in(o->ib_pair, T->ib_face_pair, nf);
    copy_in(o->ib_aux, T->ib_aux_point, nf * D);
    copy_in(o->ib_nrm, T->ib_normal, nf * D);
    copy_in(o->ib_bc, T->ib_bc, nf * (D + 2));
    copy_in(o->ib_img_off, T->ib_image_off, nf + 1);
    copy_in(o->ib_img, T->ib_image_cell, o->ib_img_off[nf]);
    copy_in(o->ib_cut_off, T->ib_cut_off, nf + 1);
    copy_in(o->ib_cut_idx, T->ib_cut_index, o->ib_cut_off[nf]);
    copy_in(o->ib_cut_gas, T->ib_cut_gas, o->ib_cut_off[nf]);
    copy_in(o->ib_cut_sol, T->ib_cut_solid, o->ib_cut_off[nf]);
    o->ib_of_pair.assign(size_t(npair), -1);
    for (int64_t k = 0; k < nf; ++k) o->ib_of_pair[o->ib_pair[k]] = k;
    o->ib_df_off.assign(size_t(nf + 1), 0);
    for (int64_t k = 0; k < nf; ++k) {
      const int64_t p = o->ib_pair[k];
      if (p < 0 || p >= npair) { delete o; return nullptr; }
      o->ib_df_off[k + 1] = o->ib_df_off[k] + o->nv(o->ph[p]);
    }
    o->ib_df.assign(size_t(o->ib_df_off[nf] * o->ndf), 0.0);
    o->ib_ssdf.assign(o->ib_df.size(), 0.0);
    o->ib_faces_of.assign(size_t(o->nc), {});
    for (int64_t k = 0; k < nf; ++k) o->ib_faces_of[size_t(o->ph[o->ib_pair[k]])].push_back(k);
  }
  int NW = D + 2;
  o->w.assign(size_t(o->nc * NW), 0.0);
  o->prim.assign(size_t(o->nc * NW), 0.0);
  o->qf.assign(size_t(o->nc * D), 0.0);
  o->sw.assign(size_t(o->nc * NW * D), 0.0);
  o->flux.assign(size_t(o->nc * NW), 0.0);
  o->df.assign(size_t(o->np * o->ndf), 0.0);
  o->sdf.assign(size_t(o->np * o->ndf * D), 0.0);
  o->vflux.assign(size_t(o->np * o->ndf), 0.0);
  return o;
}

void orc_destroy(orc_ctx* o) { delete o; }

// ------------------------------------------------------------- velocity-space adaptation
namespace {

double jl_max(double a, double b) {   // Julia max: NaN wins
  if (a != a) return a;
  if (b != b) return b;
  return b > a ? b : a;
}

// VsNeighborIndex of one cell's grid (src/Velocity_space/Neighbor.jl:61-124): the keys
// (cellkey) of its velocity cells, which are in key order.
struct VsIndex {
  int D = 2, L = 0, mb = 0;
  double vmin[3], hf[3];
  int64_t tn[3], stride[3];
  std::vector<uint64_t> keys;
  const std::vector<int8_t>* lev = nullptr;
  int64_t g0 = 0, n = 0;
  void init(const orc_ctx& o) {
    D = o.dim;
    L = o.vs_maxlevel;
    mb = D * L;
    int64_t s = 1;
    for (int d = 0; d < D; ++d) {
      vmin[d] = o.vs_lo[d];
      // ds0 = (quad_hi - quad_lo) / vs_trees_num, h_fine = ds0 / 2.0^maxlevel
      hf[d] = ((o.vs_hi[d] - o.vs_lo[d]) / double(o.vs_tn[d])) / std::ldexp(1.0, L);
      tn[d] = o.vs_tn[d];
      stride[d] = s;
      s *= tn[d];
    }
  }
  uint64_t code_of(const int64_t* g) const {
    int64_t root = 0;
    uint64_t morton = 0;
    const int64_t mask = (int64_t(1) << L) - 1;
    for (int d = 0; d < D; ++d) {
      root += (g[d] >> L) * stride[d];
      const int64_t q = g[d] & mask;
      for (int b = 0; b < L; ++b) morton |= uint64_t((q >> b) & 1) << (b * D + d);
    }
    return (uint64_t(root) << mb) | morton;
  }
  void build(const orc_ctx& o, int64_t c) {
    g0 = o.goff(c);
    n = o.nv(c);
    keys.resize(size_t(n));
    for (int64_t i = 0; i < n; ++i) {
      const int lv = o.vlev[g0 + i];
      int64_t g[3];
      for (int d = 0; d < D; ++d) {
        const double cell = hf[d] * double(int64_t(1) << (L - lv));
        g[d] = int64_t(std::nearbyint((o.vmid[(g0 * D) + d * n + i] - 0.5 * cell - vmin[d]) / hf[d]));
      }
      keys[size_t(i)] = code_of(g);
    }
  }
  double size(const orc_ctx& o, int64_t j, int d) const {
    return hf[d] * double(int64_t(1) << (L - o.vlev[g0 + j]));
  }
  // leaf containing the finest-level cell `code`, -1 if none (_vs_locate_code)
  int64_t locate(const orc_ctx& o, uint64_t code) const {
    int64_t lo = 0, hi = n - 1, k = -1;
    while (lo <= hi) {
      const int64_t m = (lo + hi) / 2;
      if (keys[size_t(m)] <= code) { k = m; lo = m + 1; } else { hi = m - 1; }
    }
    if (k < 0) return -1;
    const uint64_t span = uint64_t(1) << (D * (L - o.vlev[g0 + k]));
    return (code >= keys[size_t(k)] && code < keys[size_t(k)] + span) ? k : -1;
  }
  // vs_face_neighbor (Neighbor.jl:186-201): the leaf across the face of leaf i in dimension
  // dim, direction dir; -1 at the velocity-domain boundary
  int64_t face(const orc_ctx& o, int64_t i, int dim, int dir) const {
    const double cell = size(o, i, dim);
    int64_t g[3];
    for (int d = 0; d < D; ++d) {
      double coord = o.vmid[(g0 * D) + d * n + i];
      if (d == dim) coord += double(dir) * (0.5 * cell + 0.5 * hf[d]);
      const double x = std::floor((coord - vmin[d]) / hf[d]);
      if (x < 0 || x >= double(tn[d] << L)) return -1;
      g[d] = int64_t(x);
    }
    return locate(o, code_of(g));
  }
};

// _lohner_ratio, src/Velocity_space/Criteria.jl:221-227
double lohner_ratio(double L, double C, double R, double dsL, double dsR, double scale) {
  const double eflt = 1e-2, eabs = 1e-3;
  const double num = std::fabs(dsR * L - (dsL + dsR) * C + dsL * R);
  const double den = dsR * std::fabs(L - C) + dsL * std::fabs(R - C) +
                     eflt * (dsR * std::fabs(L) + (dsL + dsR) * std::fabs(C) + dsL * std::fabs(R)) +
                     eabs * scale * (dsL + dsR);
  return den > 0 ? num / den : 0.0;
}

}  // namespace

int orc_set_sdf(orc_ctx* o, const double* sdf) {
  if (!o || !sdf) return KAH_EINVAL;
  o->sdf.assign(sdf, sdf + o->np * o->ndf * o->dim);
  return 0;
}

int orc_vs_resolution(orc_ctx* o, double* vr) {
  if (!o || !vr) return KAH_EINVAL;
  const int D = o->dim, NDF = o->ndf, NW = D + 2;
  // weight of a finest cell: prod(du)/prod(vs_trees_num)/2^(DIM*maxlevel)
  double du = o->vs_hi[0] - o->vs_lo[0];
  int64_t tn = o->vs_tn[0];
  for (int d = 1; d < D; ++d) {
    du = du * (o->vs_hi[d] - o->vs_lo[d]);
    tn *= o->vs_tn[d];
  }
  const double weight = du / double(tn) / double(int64_t(1) << (D * o->vs_maxlevel));
  double dres = 0.0, eres = 0.0;
  for (int64_t c = 0; c < o->nc; ++c) {
    if (o->benc[c] < 0) continue;
    const int64_t n = o->nv(c), g0 = o->goff(c);
    const double* f = &o->df[size_t(o->off(c) * NDF)];
    const double* U = &o->prim[size_t(c * NW) + 1];
    double dmax = -HUGE_VAL, emax = -HUGE_VAL;
    for (int64_t x = 0; x < n * NDF; ++x) dmax = jl_max(dmax, f[x]);
    for (int64_t i = 0; i < n; ++i) {
      double c2 = 0.0;
      for (int d = 0; d < D; ++d) {
        const double t = U[d] - o->vmid[(g0 * D) + d * n + i];
        c2 = d == 0 ? t * t : c2 + t * t;
      }
      const double e = NDF == 1 ? f[i] * c2 : f[i] * c2 + f[n + i];
      emax = jl_max(emax, e);
    }
    dres = jl_max(dmax * weight, dres);
    eres = jl_max(0.5 * emax * weight, eres);
  }
  vr[0] = dres;
  vr[1] = eres;
  return 0;
}

int orc_vs_flags(orc_ctx* o, const kah_vs_adapt* p, const double* vr, uint8_t* flags) {
  if (!o || !p || !vr || !flags) return KAH_EINVAL;
  const int D = o->dim, NDF = o->ndf, NW = D + 2;
  const bool lohner = p->mode == KAH_VS_LOHNER;
  const double CL = p->coeff_local, CG = p->coeff_global, tauL = p->coeff_lohner;
  const double tauC = 0.3 * tauL;   // VS_LOHNER_COARSEN_RATIO * ADAPT_COEFFI_VS_LOHNER
  const double P2 = double(1 << D);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t c = 0; c < o->nc; ++c) {
    VsIndex X;
    X.init(*o);
    X.build(*o, c);
    const int64_t n = o->nv(c), g0 = o->goff(c), off = o->off(c);
    const double* f = &o->df[size_t(off * NDF)];
    const double* s = &o->sdf[size_t(off * NDF * D)];
    const double* w = &o->w[size_t(c * NW)];
    const double* U = &o->prim[size_t(c * NW) + 1];
    const double* ds = o->rds(c);
    double s1 = 0.0, s2 = 0.0;   // vs_lohner_scales
    for (int64_t i = 0; i < n; ++i) {
      const double a1 = std::fabs(f[i]);
      if (a1 > s1) s1 = a1;
      if (NDF == 2) {
        const double a2 = std::fabs(f[n + i]);
        if (a2 > s2) s2 = a2;
      }
    }
    double SU = 0.0;
    for (int d = 0; d < D; ++d) SU = d == 0 ? U[d] * U[d] : SU + U[d] * U[d];
    const double den = w[NW - 1] - 0.5 * w[0] * SU;
    for (int64_t i = 0; i < n; ++i) {
      // _criterion_cell!: cdf = df + max_d |sdf*ds|
      double cdf[2];
      for (int k = 0; k < NDF; ++k) {
        double m = 0.0;
        for (int d = 0; d < D; ++d) {
          const double a = std::fabs(s[(d * NDF + k) * n + i] * ds[d]);
          if (a > m) m = a;
        }
        cdf[k] = f[k * n + i] + m;
      }
      const double wt = o->vw[g0 + i];
      double S = 0.0;
      for (int d = 0; d < D; ++d) {
        const double t = U[d] - o->vmid[(g0 * D) + d * n + i];
        S = d == 0 ? t * t : S + t * t;
      }
      const double c1w = cdf[0] * wt;
      double loc_r, loc_c, e_w;
      if (NDF == 2) {
        e_w = 0.5 * (S * cdf[0] + cdf[1]) * wt;
        loc_r = jl_max(std::fabs(e_w) / den, c1w / w[0]);
        loc_c = jl_max(e_w / den, c1w / w[0]);
      } else {
        loc_r = jl_max(std::fabs(0.5 * S * cdf[0] * wt) / den, c1w / w[0]);
        e_w = 0.5 * (S * cdf[0]) * wt;
        loc_c = jl_max(e_w / den, c1w / w[0]);
      }
      const bool local_refine = loc_r > CL;
      const bool local_coarsen = loc_c < CL / P2;
      const bool global_refine = (NDF == 2 ? c1w > CG * vr[0] : c1w > vr[0] * CG) || e_w > vr[1] * CG;
      const bool global_coarsen =
          (NDF == 2 ? c1w < CG * vr[0] / P2 : c1w < vr[0] * CG / P2) && e_w < vr[1] * CG / P2;
      double eta = 0.0;
      if (lohner) {
        const double h0 = f[i], b0 = NDF == 2 ? f[n + i] : 0.0;
        for (int d = 0; d < D; ++d) {
          const double hi = X.size(*o, i, d);
          const int64_t Ln = X.face(*o, i, d, -1), Rn = X.face(*o, i, d, +1);
          const double dsL = Ln < 0 ? hi : 0.5 * (hi + X.size(*o, Ln, d));
          const double dsR = Rn < 0 ? hi : 0.5 * (hi + X.size(*o, Rn, d));
          const double hL = Ln < 0 ? 0.0 : f[Ln], hR = Rn < 0 ? 0.0 : f[Rn];
          eta = jl_max(eta, lohner_ratio(hL, h0, hR, dsL, dsR, s1));
          if (NDF == 2) {
            const double bL = Ln < 0 ? 0.0 : f[n + Ln], bR = Rn < 0 ? 0.0 : f[n + Rn];
            eta = jl_max(eta, lohner_ratio(bL, b0, bR, dsL, dsR, s2));
          }
        }
      }
      const bool base = lohner ? (eta > tauL || local_refine) : (local_refine || global_refine);
      const bool refine = o->vlev[g0 + i] < o->vs_maxlevel && base;
      const bool coarsen = lohner ? (eta < tauC && local_coarsen) : (local_coarsen && global_coarsen);
      flags[off + i] = uint8_t((refine ? KAH_VS_REFINE : 0) | (coarsen ? KAH_VS_COARSEN : 0));
    }
  }
  return 0;
}
int orc_get_ib(orc_ctx* o, double* aux_df) {
  if (!o || !aux_df) return KAH_EINVAL;
  if (!o->ib_done) return KAH_ESTATE;
  std::copy(o->ib_df.begin(), o->ib_df.end(), aux_df);
  return 0;
}

const char* orc_last_error(const orc_ctx* o) { return o ? o->err.c_str() : "null context"; }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int orc_set_state(orc_ctx* o, const double* w, const double* prim, const double* df) {
  int NW = o->NW();
  if (w) std::memcpy(o->w.data(), w, sizeof(double) * o->nc * NW);
  if (prim) std::memcpy(o->prim.data(), prim, sizeof(double) * o->nc * NW);
  if (df) std::memcpy(o->df.data(), df, sizeof(double) * o->np * o->ndf);
  std::fill(o->flux.begin(), o->flux.end(), 0.0);
  std::fill(o->vflux.begin(), o->vflux.end(), 0.0);
  return 0;
}

int orc_get_state(orc_ctx* o, double* w, double* prim, double* qf, double* sw, double* df,
                  double* sdf) {
  int NW = o->NW();
  if (w) std::memcpy(w, o->w.data(), sizeof(double) * o->nc * NW);
  if (prim) std::memcpy(prim, o->prim.data(), sizeof(double) * o->nc * NW);
  if (qf) std::memcpy(qf, o->qf.data(), sizeof(double) * o->nc * o->dim);
  if (sw) std::memcpy(sw, o->sw.data(), sizeof(double) * o->nc * NW * o->dim);
  if (df) std::memcpy(df, o->df.data(), sizeof(double) * o->np * o->ndf);
  if (sdf) std::memcpy(sdf, o->sdf.data(), sizeof(double) * o->np * o->ndf * o->dim);
  return 0;
}

int orc_get_flux(orc_ctx* o, double* flux, double* vs_flux) {
  if (flux) std::memcpy(flux, o->flux.data(), sizeof(double) * o->nc * o->NW());
  if (vs_flux) std::memcpy(vs_flux, o->vflux.data(), sizeof(double) * o->np * o->ndf);
  return 0;
}

// slope!(p4est,ka), src/Flux/Slope.jl:1047-1070
int orc_slope(orc_ctx* o) {
  o->err.clear();
  int Lmin = 1 << 30, Lmax = -1;
  for (int64_t c = 0; c < o->nc; ++c) {
    if (o->benc[c] < 0) continue;
    Lmin = std::min(Lmin, int(o->lvl[c]));
    Lmax = std::max(Lmax, int(o->lvl[c]));
  }
  if (Lmax < 0) return 0;
  int rc_all = 0;
  for (int L = Lmin; L <= Lmax; ++L) {
#pragma omp parallel
    {
      SlopeWs ws;
#pragma omp for schedule(dynamic, 16)
      for (int64_t c = 0; c < o->nc; ++c) {
        if (o->benc[c] < 0 || o->lvl[c] != L) continue;
        int rc = 0;
        if (L == Lmin) {   // update_slope_level!, Slope.jl:977-1019
          for (int dir = 0; dir < o->dim && !rc; ++dir) rc = slope_dispatch(*o, c, dir, ws);
        } else {           // update_slope_transverse_level!, Slope.jl:849-945
          rc = slope_transverse_cell(*o, c, ws);
        }
        if (rc) {
#pragma omp critical
          rc_all = rc;
        }
      }
    }
    if (rc_all) return rc_all;
  }
  // update_macro_slope!, Slope.jl:1022-1036
  int NW = o->NW();
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t c = 0; c < o->nc; ++c) {
    if (o->benc[c] < 0) continue;
    int64_t n = o->nv(c);
    for (int d = 0; d < o->dim; ++d) {
      double wv[5];
      moments_full(*o, c, &o->sdf[o->off(c) * o->ndf * o->dim + int64_t(d) * o->ndf * n], n, wv);
      for (int q = 0; q < NW; ++q) o->sw[(c * NW + q) * o->dim + d] = wv[q];
    }
  }
  return 0;
}

// flux!(p4est,ka), src/Flux/Flux.jl:458-495. Faces are computed in
// parallel; each cell then applies its side of its faces in face-list order, which gives
// every accumulator the same sequence of additions as the reference's serial face loop.
int orc_flux(orc_ctx* o, double dt) {
  o->err.clear();
  const int64_t npair = int64_t(o->ph.size());
  // update_solid_cell! and update_solid_neighbor! (src/Flux/Flux.jl:461-482): the solid
  // cells, then the SolidNeighbor df that the IB faces read
  if (!o->ib_solid.empty()) ib_solid_cells(*o);
  if (!o->ib_pair.empty()) {
    ib_solid_neighbors(*o);
    o->ib_done = true;
  }
  if (o->cell_pairs.empty() && o->nc > 0) {
    o->cell_pairs.assign(size_t(o->nc), {});
    for (int64_t p = 0; p < npair; ++p) {
      o->cell_pairs[o->ph[p]].push_back(p * 2);
      if (o->pt[p] >= 0) {
        int64_t t = o->src(o->pt[p]);
        if (o->benc[t] >= 0 && t != o->ph[p]) o->cell_pairs[t].push_back(p * 2 + 1);
      }
    }
  }
  std::vector<size_t> cursor(size_t(o->nc), 0);
  const int64_t CH = 4096;
  int rc_all = 0;
  for (int64_t p0 = 0; p0 < npair; p0 += CH) {
    const int64_t p1 = std::min(npair, p0 + CH);
    std::vector<FaceWork> work(size_t(p1 - p0));
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = p0; p < p1; ++p) {
      int rc = (o->pt[p] < 0) ? domain_face(*o, p, dt, work[p - p0])
                              : inner_face(*o, p, dt, work[p - p0]);
      if (rc) {
#pragma omp critical
        rc_all = rc;
      }
    }
    if (rc_all) return rc_all;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < o->nc; ++c) {
      const auto& lst = o->cell_pairs[c];
      size_t& k = cursor[c];
      while (k < lst.size() && (lst[k] >> 1) < p1) {
        int64_t p = lst[k] >> 1;
        bool here = (lst[k] & 1) == 0;
        apply_face(*o, work[p - p0], here, !here);
        ++k;
      }
    }
  }
  return 0;
}

// iterate!(p4est,ka) with CAIDVM_Marching (src/Theory/Iterate.jl:96-130) or Euler
// (:131-162); residual_check! (src/Solver/Finalize.jl:5-11)
int orc_iterate(orc_ctx* o, double dt, int64_t step, double* sum_res, double* sum_avg) {
  const int D = o->dim, NDF = o->ndf, NW = o->NW();
  bool check = (step % o->st_check) == 0;
  std::vector<double> res(size_t(o->nc * NW), 0.0), avg(size_t(o->nc * NW), 0.0);
#pragma omp parallel
  {
    std::vector<double> fc, Fm;
#pragma omp for schedule(dynamic, 16)
    for (int64_t c = 0; c < o->nc; ++c) {
      if (o->benc[c] < 0) continue;
      int64_t n = o->nv(c);
      double* w = &o->w[c * NW];
      double* fl = &o->flux[c * NW];
      const double* dsc = o->rds(c);
      double area = dsc[0];
      for (int d = 1; d < D; ++d) area = area * dsc[d];
      double* f = &o->df[o->off(c) * NDF];
      double* vfl = &o->vflux[o->off(c) * NDF];
      double prim_new[5];
      if (o->marching == KAH_MARCH_CAIDVM) {
        for (int q = 0; q < NW; ++q) w[q] += fl[q] * dt / area;
        double prim_c[5];
        get_prim(D, w, o->gamma, prim_c);
        for (int64_t x = 0; x < n * NDF; ++x) f[x] += dt / area * vfl[x];
        double wf[5], prim[5];
        moments_full(*o, c, f, n, wf);
        get_prim(D, wf, o->gamma, prim);
        double tau = get_tau(D, prim_c, o->mu, o->omega);
        fc.assign(size_t(n * NDF), 0.0);
        for (int64_t i = 0; i < n; ++i) {
          double u[3] = {0, 0, 0}, Mc[2], M[2];
          for (int d = 0; d < D; ++d) u[d] = o->U(c, i, d);
          maxwell(D, u, prim_c, o->K, Mc);
          maxwell(D, u, prim, o->K, M);
          for (int k = 0; k < NDF; ++k) {
            fc[k * n + i] = Mc[k];
            f[k * n + i] += Mc[k] - M[k];
          }
        }
        double q[3];
        heat_flux_full(*o, c, f, n, prim_c, q);
        for (int d = 0; d < D; ++d) o->qf[c * D + d] = q[d];
        for (int64_t i = 0; i < n; ++i) {
          double u[3] = {0, 0, 0}, Fc[2], Sp[2];
          for (int d = 0; d < D; ++d) u[d] = o->U(c, i, d);
          for (int k = 0; k < NDF; ++k) Fc[k] = fc[k * n + i];
          shakhov(D, u, Fc, prim_c, q, o->Pr, o->K, Sp);
          for (int k = 0; k < NDF; ++k) fc[k * n + i] = Fc[k] + Sp[k];
        }
        for (int64_t x = 0; x < n * NDF; ++x) f[x] *= tau / (tau + dt);
        for (int64_t x = 0; x < n * NDF; ++x) f[x] += dt / (tau + dt) * fc[x];
        for (int q2 = 0; q2 < NW; ++q2) prim_new[q2] = prim_c[q2];
      } else if (o->marching == KAH_MARCH_CIP) {
        // iterate!(CIP_Marching), src/Theory/I-projection.jl:161-192, with
        // positivity_preserving_ib! for donor cells (a no-op elsewhere)
        for (int q2 = 0; q2 < NW; ++q2) w[q2] += fl[q2] * dt / area;
        ib_positivity(*o, c, w, vfl, dt, area);
        double prim_c[5];
        get_prim(D, w, o->gamma, prim_c);
        for (int64_t x = 0; x < n * NDF; ++x) f[x] += dt / area * vfl[x];
        double q[3];
        heat_flux_full(*o, c, f, n, prim_c, q);
        for (int d = 0; d < D; ++d) o->qf[c * D + d] = q[d];
        fc.assign(size_t(n * NDF), 0.0);
        for (int64_t i = 0; i < n; ++i) {
          double u[3] = {0, 0, 0}, Fc[2], Sp[2];
          for (int d = 0; d < D; ++d) u[d] = o->U(c, i, d);
          maxwell(D, u, prim_c, o->K, Fc);
          shakhov(D, u, Fc, prim_c, q, o->Pr, o->K, Sp);
          for (int k = 0; k < NDF; ++k) fc[k * n + i] = Fc[k] + Sp[k];
        }
        cip_projection(*o, c, f, n, NDF == 2 ? f + n : nullptr, w);
        double tau = get_tau(D, prim_c, o->mu, o->omega);
        for (int64_t x = 0; x < n * NDF; ++x) f[x] *= tau / (tau + dt);
        for (int64_t x = 0; x < n * NDF; ++x) f[x] += dt / (tau + dt) * fc[x];
        for (int q2 = 0; q2 < NW; ++q2) prim_new[q2] = prim_c[q2];
      } else {
        // Euler: `isa(flux_type, MicroFlux)` is always false, so w always takes the
        // macroscopic flux (Iterate.jl:144-149)
        for (int q2 = 0; q2 < NW; ++q2) w[q2] += fl[q2] * dt / area;
        for (int q2 = 0; q2 < NW; ++q2) fl[q2] = 0.0;
        double prim[5];
        get_prim(D, w, o->gamma, prim);
        double tau = get_tau(D, prim, o->mu, o->omega);
        double q[3];
        heat_flux_full(*o, c, f, n, prim, q);
        for (int d = 0; d < D; ++d) o->qf[c * D + d] = q[d];
        for (int64_t i = 0; i < n; ++i) {
          double u[3] = {0, 0, 0}, M[2], Sp[2];
          for (int d = 0; d < D; ++d) u[d] = o->U(c, i, d);
          maxwell(D, u, prim, o->K, M);
          shakhov(D, u, M, prim, q, o->Pr, o->K, Sp);
          for (int k = 0; k < NDF; ++k) {
            double Fk = M[k] + Sp[k];
            double& fx = f[k * n + i];
            fx = (fx + dt / area * vfl[k * n + i]) * tau / (tau + dt) + dt / (tau + dt) * Fk;
          }
        }
        for (int q2 = 0; q2 < NW; ++q2) prim_new[q2] = prim[q2];
      }
      if (check) {
        for (int q2 = 0; q2 < NW; ++q2) {
          double dlt = prim_new[q2] - o->prim[c * NW + q2];
          res[c * NW + q2] = dlt * dlt;
          avg[c * NW + q2] = std::fabs(prim_new[q2]);
        }
      }
      for (int q2 = 0; q2 < NW; ++q2) o->prim[c * NW + q2] = prim_new[q2];
      for (int q2 = 0; q2 < NW; ++q2) fl[q2] = 0.0;
      for (int64_t x = 0; x < n * NDF; ++x) vfl[x] = 0.0;
    }
  }
  // residual sums in cell order (Finalize.jl:5-11)
  for (int q2 = 0; q2 < NW; ++q2) {
    double sr = 0.0, sa = 0.0;
    if (check)
      for (int64_t c = 0; c < o->nc; ++c) {
        if (o->benc[c] < 0) continue;
        sr += res[c * NW + q2];
        sa += avg[c * NW + q2];
      }
    if (sum_res) sum_res[q2] = sr;
    if (sum_avg) sum_avg[q2] = sa;
  }
  return 0;
}

// ---------------------------------------------------------------- point kernels (tests)
void orc_discrete_maxwell(int dim, int ndf, int64_t n, const double* vs_mid,
                          const double* prim, double K, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double u[3] = {0, 0, 0}, M[2];
    for (int d = 0; d < dim; ++d) u[d] = vs_mid[d * n + i];
    maxwell(dim, u, prim, K, M);
    for (int k = 0; k < ndf; ++k) out[k * n + i] = M[k];
  }
}
void orc_micro_to_macro(int dim, int ndf, int64_t n, const double* vs_mid, const double* df,
                        const double* weight, double* w_out) {
  orc_ctx o;
  o.dim = dim;
  o.ndf = ndf;
  o.voff = {0, n};
  o.vmid.assign(vs_mid, vs_mid + n * dim);
  o.vw.assign(weight, weight + n);
  moments_full(o, 0, df, n, w_out);
}
void orc_heat_flux(int dim, int ndf, int64_t n, const double* vs_mid, const double* df,
                   const double* weight, const double* prim, double* q_out) {
  orc_ctx o;
  o.dim = dim;
  o.ndf = ndf;
  o.voff = {0, n};
  o.vmid.assign(vs_mid, vs_mid + n * dim);
  o.vw.assign(weight, weight + n);
  heat_flux_full(o, 0, df, n, prim, q_out);
}
void orc_i_projection(int dim, int ndf, int64_t n, const double* vs_mid, const double* weight,
                      double* df, const double* w) {
  orc_ctx o;
  o.dim = dim;
  o.ndf = ndf;
  o.voff = {0, n};
  o.vmid.assign(vs_mid, vs_mid + n * dim);
  o.vw.assign(weight, weight + n);
  cip_projection(o, 0, df, n, ndf == 2 ? df + n : nullptr, w);
}
void orc_shakhov_part(int dim, int ndf, int64_t n, const double* vs_mid, const double* F,
                      const double* prim, const double* qf, double Pr, double K,
                      double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double u[3] = {0, 0, 0}, Fi[2], S[2];
    for (int d = 0; d < dim; ++d) u[d] = vs_mid[d * n + i];
    for (int k = 0; k < ndf; ++k) Fi[k] = F[k * n + i];
    shakhov(dim, u, Fi, prim, qf, Pr, K, S);
    for (int k = 0; k < ndf; ++k) out[k * n + i] = S[k];
  }
}
void orc_get_prim(int dim, const double* w, double gamma, double* prim) {
  get_prim(dim, w, gamma, prim);
}
void orc_get_conserved(int dim, const double* prim, double gamma, double* w) {
  get_conserved(dim, prim, gamma, w);
}
double orc_get_tau(int dim, const double* prim, double mu, double omega) {
  return get_tau(dim, prim, mu, omega);
}
// ref_vhs_vis, src/Gas/Model.jl:5-9
double orc_ref_vhs_vis(double Kn, double omega, double alpha, double T_ref) {
  double mu0 = 5.0 * (alpha + 1.0) * (alpha + 2.0) * std::sqrt(PI) /
               (4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega)) * Kn;
  return mu0 * std::pow(T_ref, 0.5 - omega);
}

}  // extern "C"

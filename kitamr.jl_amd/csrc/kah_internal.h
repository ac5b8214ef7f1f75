// kah_internal.h -- device-side descriptors and point math shared by the kernels and the
// host code that derives the descriptors from a kah_topology.
//
// Dataflow of one step on the device (single GPU, all state resident in HBM):
//   K1  slope   (one launch per physical level, coarse to fine): per fluid cell, the
//       velocity-space slopes sdf in every direction, and the macroscopic "half fluxes":
//       the moments of the upwind micro flux that this cell's own reconstruction sends
//       through each of its faces.
//   K2  flux+relax (one launch): per fluid cell, the micro flux of every face gathered per
//       velocity cell (own upwind cells from its own reconstruction, incoming cells from
//       the neighbour's), the macroscopic flux from the two half fluxes of each face, and
//       the implicit relaxation update, writing f of the next step into the other buffer.
// A face's macroscopic flux is the sum of two half fluxes that each side reads from the
// same buffer, so the two cells see bit-identical, opposite face fluxes (mass is conserved
// to round-off), and each cell accumulates its face contributions in face-list order, the
// order of the reference's scatter.
#pragma once
#include <stdint.h>

namespace kah {

constexpr double EPS = 1e-12;                        // src/Abstract/Types.jl:3
constexpr double MACH_EPS = 2.220446049250313e-16;   // eps(Float64)
constexpr double PI = 3.141592653589793;

// slope modes per (cell, direction)
enum : int8_t { SL_ZERO = 0, SL_INNER = 1, SL_BOUND = 2, SL_SKIP = 3 };
// micro-flux modes of one face side
enum : int8_t {
  MI_LIMITED = 0,     // positivity-preserving reconstruction, CAIDVM.jl:127-141
  MI_UNLIMITED = 1,   // (f + dx.sdf) * vn
  MI_PLAIN = 2,       // f * vn (no reconstruction)
  MI_MAXWELL = 3,     // Maxwellian of a boundary state * vn (inflow)
  MI_WALL = 4,        // Maxwellian with the zero-mass-flux wall density * vn
  MI_INTERP = 5,      // InterpolatedOutflow ghost reconstruction * vn
  MI_SELF_PLAIN = 6,  // UniformOutflow: the cell's own f * vn on the incoming set
};

// velocity-space adaptation: criteria mode and flag bits (the values of KAH_VS_*)
enum { VS_LOHNER = 0, VS_CONTRIBUTION = 1, VS_REFINE = 1, VS_COARSEN = 2 };

struct SlopeNb {        // one neighbour in a slope stencil
  int32_t cell;         // data cell
  int32_t proj;         // 1: project with the neighbour's transverse slopes
  int32_t map;          // overlap map id of (owner grid, neighbour grid), MAP_SAME, or -1
  int32_t pad;
  double dm[3];         // transverse offsets (dm[dir] = 0)
  float dm_f[3];        // the same rounded to float (FP32 advection form)
  float pad_f;
};

// Overlap map of adapted velocity grids: one per distinct ordered pair (owner grid, neighbour
// grid) of different grids; for velocity cell i of the owner grid, entry
// map[map_base[id] + i] = (j << 8) | cnt: j is the first overlapping velocity cell of the
// neighbour grid, cnt = 0 when it is as coarse or coarser than i (one cell, factor 1), else
// the number of finer cells j, j+1, ... inside i. A descriptor's map field holds the id, or
// MAP_SAME when the neighbour has the owner's grid (cell i meets cell i), or -1 (no map: the
// key search is done in place). own/nb are the grids' first entries in the geometry table.
constexpr int32_t MAP_SAME = -2;
struct OvJob {
  int64_t own, nb;
  int32_t n_own, n_nb;
  int64_t base;
};

// Descriptor fields are 32-bit or wider: the kernels read them with scalar loads, which
// have no byte form on gfx9.
struct SlopeDesc {      // per (fluid cell, direction)
  int32_t mode;
  int32_t pad;
  int32_t nb_beg[2];    // side 0 = L (or the one side of a BOUND), side 1 = R
  int32_t nb_cnt[2];
  double ds[2];         // dsL, dsR
  double rds[2];        // 1/dsL, 1/dsR (correctly rounded)
  double inv_cnt[2];    // 1/nb_cnt (exact: the counts are powers of two)
  float rds_f[2];       // rds and inv_cnt rounded to float (FP32 advection form)
  float inv_cnt_f[2];
};

struct HalfDesc {       // one half flux computed by its owner cell in K1
  int32_t slot;         // output row in half_fw (pair*2 + side)
  int32_t data_cell;    // cell whose f/sdf feed the micro flux
  int32_t mode;         // MI_*
  int32_t sel;          // 0: rot*u[dir] <= 0 ; 1: rot*u[dir] > 0
  int32_t dir;
  int32_t bc;           // boundary condition (domain halves) or -1
  double rot;
  double xrec[3];       // cell-record midpoint used for the reconstruction
  double dsrec[3];      // cell-record size used by the limiter
  double fmid[3];       // face midpoint
  // FP32 advection form: rot, fmid - xrec (formed in double, then rounded) and dsrec as float
  float rot_f;
  float dx_f[3];
  float ds_f[3];
  float pad_f;
};

struct EntryDesc {      // one face side of a fluid cell, gathered in K2
  int32_t pair;         // half fluxes at half_fw[pair*2 + 0/1]
  int32_t oth_cell;     // data cell of the other side (-1 for domain faces)
  int32_t side;         // 0: the cell is "here" (adds), 1: "there" (subtracts)
  int32_t own_mode;     // MI_* of the cell's own upwind set
  int32_t oth_mode;     // MI_* of the incoming set
  int32_t dir;
  int32_t bc;
  int32_t map;          // overlap map id of (owner grid, other grid), MAP_SAME, or -1
  double rot;
  double area;          // signed face area, rot*A (or rot*A/2^(DIM-1))
  double fmid[3];
  double xown[3], dsown[3];
  double xoth[3], dsoth[3];
  // FP32 advection form: rot, area, fmid - x (formed in double, then rounded) and ds as float
  float rot_f, area_f;
  float dxo_f[3], dso_f[3];
  float dxt_f[3], dst_f[3];
};

struct IbFaceDesc {     // one IB face (donor, SolidNeighbor), rebuilt every step
  int32_t donor, solid; // cells of the face
  int32_t pseudo;       // cell index whose f rows hold the SolidNeighbor's df
  int32_t dir;
  int32_t img_beg, img_cnt;   // image stencil (fluid neighbours, then the donor)
  int32_t cut_beg, cut_cnt;   // cut velocity cells
  int32_t slot;         // half_fw row of the face's there half (pair*2 + 1)
  int32_t pad;
  int64_t row;          // first row of the face in the per-face velocity arrays (ib_ssdf)
  double apd;           // aux_point[dir]
  double rot;
  double ip[3];         // image point aux + x_donor - x_solid
  double nrm[3];        // wall normal
  double dxs, dxf, dxL; // boundary_slope! distances along dir
  double bc[5];         // wall prim, density 1
};

// the double or the float copy of a descriptor value, for arithmetic type T
template <class T>
__host__ __device__ inline T desc_val(double d, float f) {
  if constexpr (sizeof(T) == 8) return d;
  else return f;
}

// a / b given rb = RN(1/b): q = RN(a*rb) corrected by one FMA step on the exact residual,
// which is the correctly rounded quotient (Markstein's theorem; no overflow or underflow
// in the quotient). Used where b is uniform over a workgroup.
__host__ __device__ inline double div_by(double a, double b, double rb) {
  double q = a * rb;
  double e = fma(-q, b, a);
  return fma(e, rb, q);
}

// ---------------------------------------------------------------- point math (FP64)
// Every expression keeps the reference's association order; the library is compiled
// without floating-point contraction so each operation rounds as written.

__host__ __device__ inline double jl_sign(double x) {
  return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x);
}
template <class T>
__host__ __device__ inline T jl_min(T a, T b) {
  if (a != a) return a;
  if (b != b) return b;
  return b < a ? b : a;
}
// minmod, src/Flux/Slope.jl:20-24: 0.5*(sign(sL)+sign(sR))*min(|sL|,|sR|)
__host__ __device__ inline double minmod(double sL, double sR) {
  double SL = fabs(sL), SR = fabs(sR);
  return 0.5 * (jl_sign(sL) + jl_sign(sR)) * jl_min(SL, SR);
}
// The same value in select form for the kernels: for same-sign nonzero slopes the factor
// 0.5*(±1±1) is exactly ±1; otherwise it is 0 or ±1/2 with min = 0, and the result is
// 0*min (NaN for a NaN input or for opposite infinities, as in the formula). Only the sign
// of a zero result may differ, which no later use of a slope can observe (slopes only
// enter products and sums with nonzero terms).
template <class T>
__host__ __device__ inline T minmod_sel(T sL, T sR) {
  const T SL = fabs(sL), SR = fabs(sR);
  const T m = (SL != SL || SR != SR) ? SL + SR : (SR < SL ? SR : SL);
  const bool pos = sL > T(0) && sR > T(0), neg = sL < T(0) && sR < T(0);
  return pos ? m : (neg ? -m : T(0) * m);
}

// x = J^-1 g for a symmetric positive definite J (M <= 5, upper triangle read) by Cholesky
// L L^T = J: the Newton step of the I-projection (the reference's Symmetric(J,:U) \ g,
// src/Theory/I-projection.jl:124, which LAPACK solves by Bunch-Kaufman: equal to rounding).
__host__ __device__ inline void cholesky_solve(int M, const double (&J)[5][5], const double* g,
                                               double* x) {
  double L[5][5] = {};
  for (int j = 0; j < M; ++j) {
    double s = J[j][j];
    for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
    L[j][j] = sqrt(s);
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

// psi = (1, u, |u|^2/2) of build_psi_matrix (src/Theory/I-projection.jl:7-21)
template <int D>
__host__ __device__ inline void cip_psi(const double (&u)[D], double (&ps)[D + 2]) {
  ps[0] = 1.0;
  double s = 0.0;
  for (int d = 0; d < D; ++d) {
    ps[1 + d] = u[d];
    s += u[d] * u[d];
  }
  ps[D + 1] = s / 2;
}

// get_prim_2D / get_prim_3D, lib/KitCore/2D.jl:9-16, 3D.jl:12-20
template <int D>
__host__ __device__ inline void get_prim(const double* w, double g, double* p) {
  if constexpr (D == 2) {
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

// discrete_maxwell_2D2F / 3D1F point forms, lib/KitCore/2D2F.jl:15-21, 3D1F.jl:14-23
template <int D>
__host__ __device__ inline void maxwell(const double* u, const double* p, double K,
                                        double* out) {
  if constexpr (D == 2) {
    double a = u[0] - p[1], b = u[1] - p[2];
    double c2 = a * a + b * b;
    double h = p[0] * (p[3] / PI) * exp(-p[3] * c2);
    out[0] = h;
    out[1] = h * K / (2.0 * p[3]);
  } else {
    double a = u[0] - p[1], b = u[1] - p[2], c = u[2] - p[3];
    double c2 = a * a + b * b + c * c;
    out[0] = p[0] * pow(p[4] / PI, 1.5) * exp(-p[4] * c2);
  }
}

// shakhov_part_2D2F / 3D1F point forms, lib/KitCore/2D2F.jl:22-44, 3D1F.jl:24-40
template <int D>
__host__ __device__ inline void shakhov(const double* u, const double* F, const double* p,
                                        const double* q, double Pr, double K, double* out) {
  if constexpr (D == 2) {
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

// The same two functions with the factors that depend only on prim hoisted out of the
// per-velocity-cell work (identical roundings: the hoisted subexpressions are the leading
// factors of the reference's left-to-right products).
template <int D>
struct MaxPre {
  double p[D + 2];
  double a0;      // rho*(lambda/pi) (2D) or rho*(lambda/pi)^(3/2) (3D)
  double two_l;   // 2*lambda
  double r2l;     // RN(1/(2*lambda))
  double s0;      // 0.8*(1-Pr)*lambda^2/rho
};
template <int D>
__host__ __device__ inline MaxPre<D> max_pre(const double* p, double Pr) {
  MaxPre<D> m;
  for (int q = 0; q < D + 2; ++q) m.p[q] = p[q];
  const double lam = p[D + 1];
  if constexpr (D == 2) m.a0 = p[0] * (lam / PI);
  else m.a0 = p[0] * pow(lam / PI, 1.5);
  m.two_l = 2.0 * lam;
  m.r2l = 1.0 / m.two_l;
  m.s0 = 0.8 * (1 - Pr) * (lam * lam) / p[0];
  return m;
}
// Discrete Maxwellian and Shakhov part of one velocity cell from the pre-factors; T is the
// arithmetic type (double, or float in the FP32 advection form, where the FP64 parameters
// are rounded once and exp is the float one).
template <int D, class T = double>
__host__ __device__ inline void maxwell_pre(const double* u, const MaxPre<D>& m, T K, T* out) {
  // the exponent in FP64 from the grid's FP64 velocities (a float u would put its rounding,
  // times 2*lambda*|c|, into every value); exp and the products in T
  if constexpr (D == 2) {
    double a = u[0] - m.p[1], b = u[1] - m.p[2];
    double c2 = a * a + b * b;
    T h = T(m.a0) * exp(T(-m.p[3] * c2));
    out[0] = h;
    if constexpr (sizeof(T) == 8) out[1] = div_by(h * K, m.two_l, m.r2l);
    else out[1] = h * K * T(m.r2l);
  } else {
    double a = u[0] - m.p[1], b = u[1] - m.p[2], c = u[2] - m.p[3];
    double c2 = a * a + b * b + c * c;
    out[0] = T(m.a0) * exp(T(-m.p[4] * c2));
  }
}
template <int D, class T = double>
__host__ __device__ inline void shakhov_pre(const T* u, const T* F, const MaxPre<D>& m,
                                            const T* q, T K, T* out) {
  const double* p = m.p;
  const T s0 = T(m.s0);
  if constexpr (D == 2) {
    T cu = u[0] - T(p[1]), cv = u[1] - T(p[2]);
    T c2 = cu * cu + cv * cv;
    T cq = cu * q[0] + cv * q[1];
    out[0] = s0 * cq * (2 * T(p[3]) * c2 + K - 5) * F[0];
    out[1] = s0 * cq * (2 * T(p[3]) * c2 + K - 3) * F[1];
  } else {
    T cu = u[0] - T(p[1]), cv = u[1] - T(p[2]), cw = u[2] - T(p[3]);
    T c2 = cu * cu + cv * cv + cw * cw;
    T cq = cu * q[0] + cv * q[1] + cw * q[2];
    out[0] = s0 * cq * (2 * T(p[4]) * c2 - 5) * F[0];
  }
}

// get_τ, src/Gas/Model.jl:14
template <int D>
__host__ __device__ inline double get_tau(const double* prim, double mu, double omega) {
  return mu * 2.0 * pow(prim[D + 1], 1 - omega) / prim[0];
}

// Contribution of one velocity cell to micro_to_macro (2D2F.jl:119-126, 3D1F.jl:109-123);
// the 1/2 of the energy row is applied after the sum.
// Terms of type T (float in the FP32 advection form), returned as double for the FP64 sums.
template <int D, int NDF, class T = double>
__host__ __device__ inline void moment_terms(T wt, const T* u, const T* f, double* m) {
  if constexpr (D == 2) {
    m[0] = wt * f[0];
    m[1] = wt * u[0] * f[0];
    m[2] = wt * u[1] * f[0];
    m[3] = wt * ((u[0] * u[0] + u[1] * u[1]) * f[0] + f[1]);
  } else {
    m[0] = wt * f[0];
    m[1] = wt * u[0] * f[0];
    m[2] = wt * u[1] * f[0];
    m[3] = wt * u[2] * f[0];
    m[4] = wt * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * f[0];
  }
}

}  // namespace kah

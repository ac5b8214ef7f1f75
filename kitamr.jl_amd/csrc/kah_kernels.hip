// kah_kernels.hip -- gfx950 kernels of the per-step phase-space update.
//
// K1 k_slope       one workgroup per fluid cell of one physical level
// K2 k_flux_relax  one workgroup per fluid cell (MODE_FLUX / MODE_ITER / MODE_FUSED)
// K3 k_macro_slope on demand: PsData.sw = moments of sdf (src/Flux/Slope.jl:1022-1036)
//
// Threads of a workgroup stride over the cell's velocity cells (velocity index fastest in
// memory, so every load and store of f, sdf and the flux is coalesced); per-cell moment
// sums are reduced deterministically (fixed wave shuffle tree, then a fixed-order sum of
// the four wave partials), with no atomics anywhere.
#include <hip/hip_runtime.h>

#include "kah_internal.h"
#include "kah_kernels.h"

namespace kah {

constexpr int BS = 256;       // threads per workgroup
constexpr int NWAVE = BS / 64;

// ------------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sums cnt values held per thread across the workgroup; every thread receives the totals
// in out[]. red must hold NWAVE*cnt doubles of LDS.
template <int CNT>
__device__ inline void block_sum(const double* vals, int cnt, double* red, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < CNT; ++q) {
    if (q < cnt) {
      double s = wave_sum(vals[q]);
      if (lane == 0) red[wv * CNT + q] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < CNT; ++q) {
    if (q < cnt) {
      double s = red[q];
#pragma unroll
      for (int w = 1; w < NWAVE; ++w) s += red[w * CNT + q];
      out[q] = s;
    }
  }
  __syncthreads();
}

// XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs; give every XCD
// a contiguous run of cells so that face neighbours (close in Morton order) share its L2.
__device__ inline int xcd_remap(int i, int n) {
  const int x = i & 7, j = i >> 3;
  const int q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + j;
}

// ------------------------------------------------------------------ grid access
template <int D, int NDF, bool UNI>
struct Grid {
  const DevArgs& A;
  int64_t off;   // first velocity cell of the cell
  int n;
  __device__ Grid(const DevArgs& a, int c) : A(a) {
    if constexpr (UNI) {
      n = a.uni_n;
      off = int64_t(c) * n;
    } else {
      off = a.voff[c];
      n = int(a.voff[c + 1] - off);
    }
  }
  __device__ int64_t goff() const { return UNI ? 0 : off; }
  __device__ double u(int i, int d) const { return A.gu[goff() * D + int64_t(d) * n + i]; }
  __device__ double wt(int i) const { return A.gw[goff() + i]; }
  __device__ int lev(int i) const { return UNI ? 0 : int(A.glev[off + i]); }
  __device__ uint32_t key(int i) const { return UNI ? uint32_t(i) : A.gkey[off + i]; }
  __device__ double f(int i, int k) const { return A.df[off * NDF + int64_t(k) * n + i]; }
  __device__ double s(int i, int k, int d) const {
    return A.sdf[off * NDF * D + int64_t(d * NDF + k) * n + i];
  }
  __device__ double* sp(int i, int k, int d) const {
    return &A.sdf[off * NDF * D + int64_t(d * NDF + k) * n + i];
  }
};

// Calls fn(j, den) for each velocity cell j of grid G that overlaps cell i (level li, key
// ki) of another grid on the same root lattice: den = 1 when j is coarser or as fine as
// i, else 2^(D*(lj-li)). This is the pairing of the two-pointer walks of diff_vs! and
// update_micro_flux! (src/Flux/Slope.jl:29-64, src/Flux/Flux.jl:151-344), in the same order.
template <int D, int NDF, bool UNI, class Fn>
__device__ inline void for_overlap(const Grid<D, NDF, UNI>& G, int i, int li, uint32_t ki,
                                   int lmax, Fn&& fn) {
  if constexpr (UNI) {
    fn(i, 1.0);
  } else {
    int lo = 0, hi = G.n - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (G.key(mid) <= ki) lo = mid; else hi = mid - 1;
    }
    int j = lo;
    int lj = G.lev(j);
    if (lj <= li) {
      fn(j, 1.0);
      return;
    }
    const uint32_t end = ki + (1u << (D * (lmax - li)));
    for (; j < G.n && G.key(j) < end; ++j) fn(j, ldexp(1.0, D * (G.lev(j) - li)));
  }
}

// Micro flux of one velocity cell: value times the normal velocity.
// LIMITED is positivity_preserving_reconstruct (src/Flux/CAIDVM.jl:127-141); UNLIMITED the
// plain reconstruction used next to solid cells and at domain walls/inflow (CAIDVM.jl:13-14,
// :110-112); PLAIN f*vn (UniformOutflow, CAIDVM.jl:53-67).
template <int D, int NDF>
__device__ inline void micro_recon(int mode, const double* f, const double (*s)[NDF],
                                   const double* u, double dt, const double* xrec,
                                   const double* dsrec, const double* fmid, int dir,
                                   double* out) {
  const double vn = u[dir];
  if (mode == MI_PLAIN || mode == MI_SELF_PLAIN) {
#pragma unroll
    for (int k = 0; k < NDF; ++k) out[k] = f[k] * vn;
    return;
  }
  double dx[D];
#pragma unroll
  for (int t = 0; t < D; ++t) dx[t] = fmid[t] - u[t] * dt - xrec[t];
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    double s_dx = dx[0] * s[0][k];
#pragma unroll
    for (int t = 1; t < D; ++t) s_dx = s_dx + dx[t] * s[t][k];
    if (mode == MI_LIMITED) {
      double s_abs = dsrec[0] * fabs(s[0][k]);
#pragma unroll
      for (int t = 1; t < D; ++t) s_abs = s_abs + dsrec[t] * fabs(s[t][k]);
      double lim = jl_min(fabs((f[k] - MACH_EPS) / (0.5 * s_abs + EPS)), 1.0);
      out[k] = (f[k] + lim * s_dx) * vn;
    } else {
      out[k] = (f[k] + s_dx) * vn;
    }
  }
}

// InterpolatedOutflow incoming micro flux (CAIDVM.jl:71-94): the ghost state mirrored
// through the face, reconstructed at the face from the cell's own slopes.
template <int D, int NDF>
__device__ inline void micro_interp(const double* f, const double (*s)[NDF], const double* u,
                                    double dt, const double* xh, const double* fmid, int dir,
                                    double* out) {
  double xt[D];
#pragma unroll
  for (int t = 0; t < D; ++t) xt[t] = 2.0 * fmid[t] - xh[t];
  double ndx[D];
#pragma unroll
  for (int t = 0; t < D; ++t) ndx[t] = fmid[t] - u[t] * dt - xt[t];
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    double tdf = f[k] + (xt[dir] - xh[dir]) * s[dir][k];
    double s_dx = ndx[0] * s[0][k];
#pragma unroll
    for (int t = 1; t < D; ++t) s_dx = s_dx + ndx[t] * s[t][k];
    out[k] = (tdf + s_dx) * u[dir];
  }
}

// Maxwellian inflow state (bc prim, density rho for walls) times vn.
template <int D, int NDF>
__device__ inline void micro_maxwell(const double* bc, double rho, bool wall, const double* u,
                                     double K, int dir, double* out) {
  double p[D + 2];
#pragma unroll
  for (int m = 0; m < D + 2; ++m) p[m] = bc[m];
  if (wall) p[0] = rho;
  double M[2];
  maxwell<D>(u, p, K, M);
#pragma unroll
  for (int k = 0; k < NDF; ++k) out[k] = M[k] * u[dir];
}

// ------------------------------------------------------------------ K1: slopes + half fluxes
// update_slope!/update_slope_transverse_level! per cell (src/Flux/Slope.jl:458-945) with the
// diff_vs!/diff_vs_transverse! sums (:29-64, :278-332) as key-interval overlaps; then the
// moments of the cell's own upwind micro flux through each face it owns a half of.
template <int D, int NDF, bool UNI, int MAXH>
__global__ __launch_bounds__(BS) void k_slope(DevArgs A, const int32_t* __restrict__ cells,
                                              int ncells) {
  constexpr int NW = D + 2;
  const int c = cells[xcd_remap(blockIdx.x, ncells)];
  const Grid<D, NDF, UNI> G(A, c);
  const int hb = A.half_off[c];
  const int nh = A.half_off[c + 1] - hb;
  const double dt = A.dt;
  __shared__ double red[NWAVE * MAXH * NW];
  __shared__ double rho_sh[MAXH];

  double acc[MAXH][NW];
#pragma unroll
  for (int h = 0; h < MAXH; ++h)
#pragma unroll
    for (int m = 0; m < NW; ++m) acc[h][m] = 0.0;

  for (int i = threadIdx.x; i < G.n; i += BS) {
    double u[D], f[NDF], s[D][NDF];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
#pragma unroll
    for (int k = 0; k < NDF; ++k) f[k] = G.f(i, k);
    const int li = G.lev(i);
    const uint32_t ki = G.key(i);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const SlopeDesc sd = A.sl[c * D + d];
      if (sd.mode == SL_SKIP) {
#pragma unroll
        for (int k = 0; k < NDF; ++k) s[d][k] = G.s(i, k, d);
        continue;
      }
      if (sd.mode == SL_ZERO) {
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          s[d][k] = 0.0;
          *G.sp(i, k, d) = 0.0;
        }
        continue;
      }
      double a[2][NDF];
      const int nsides = (sd.mode == SL_INNER) ? 2 : 1;
      for (int side = 0; side < nsides; ++side) {
#pragma unroll
        for (int k = 0; k < NDF; ++k) a[side][k] = 0.0;
        const double dsx = sd.ds[side];
        for (int e = 0; e < sd.nb_cnt[side]; ++e) {
          const SlopeNb nb = A.slnb[sd.nb_beg[side] + e];
          const Grid<D, NDF, UNI> N(A, nb.cell);
          for_overlap<D, NDF, UNI>(N, i, li, ki, A.lmax, [&](int j, double den) {
#pragma unroll
            for (int k = 0; k < NDF; ++k) {
              double pj = N.f(j, k);
              if (nb.proj) {
#pragma unroll
                for (int t = 0; t < D; ++t) pj += nb.dm[t] * N.s(j, k, t);
              }
              double term = f[k] - pj;
              if (den != 1.0) term = term / den;
              a[side][k] += term / dsx;
            }
          });
        }
      }
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        double v = (sd.mode == SL_INNER)
                       ? minmod(a[0][k] / double(sd.nb_cnt[0]), a[1][k] / double(sd.nb_cnt[1]))
                       : a[0][k] / double(sd.nb_cnt[0]);
        s[d][k] = v;
        *G.sp(i, k, d) = v;
      }
    }
    // half fluxes owned by this cell
    const double wt = G.wt(i);
#pragma unroll
    for (int h = 0; h < MAXH; ++h) {
      if (h < nh) {
        const HalfDesc& hd = A.halves[hb + h];
        const double ru = hd.rot * u[hd.dir];
        const bool in = hd.sel ? (ru > 0.0) : (ru <= 0.0);
        double mic[NDF];
        double m[NW];
        if (hd.mode == MI_WALL) {
          // first sweep of a wall: SF over the cell's upwind set, SG over the incoming set
          // (calc_ρw, src/Theory/Math.jl:251-282)
          const double* bc = &A.bc_prim[hd.bc * NW];
          if (in) {
            double c2 = 0.0;
#pragma unroll
            for (int t = 0; t < D; ++t) {
              double cu = u[t] - bc[1 + t];
              c2 = (t == 0) ? cu * cu : c2 + cu * cu;
            }
            acc[h][1] += wt * u[hd.dir] * exp(-bc[D + 1] * c2);
          } else {
            double dx[D];
#pragma unroll
            for (int t = 0; t < D; ++t) dx[t] = hd.fmid[t] - u[t] * dt - hd.xrec[t];
            double hrec = dx[0] * s[0][0];
#pragma unroll
            for (int t = 1; t < D; ++t) hrec = hrec + dx[t] * s[t][0];
            hrec = f[0] + hrec;
            acc[h][0] += wt * u[hd.dir] * hrec;
          }
          continue;
        }
        if (!in) continue;
        if (hd.mode == MI_MAXWELL) {
          micro_maxwell<D, NDF>(&A.bc_prim[hd.bc * NW], 0.0, false, u, A.K, hd.dir, mic);
        } else if (hd.mode == MI_INTERP) {
          micro_interp<D, NDF>(f, s, u, dt, hd.xrec, hd.fmid, hd.dir, mic);
        } else if (hd.data_cell == c) {
          micro_recon<D, NDF>(hd.mode, f, s, u, dt, hd.xrec, hd.dsrec, hd.fmid, hd.dir, mic);
        } else {
          // the there side of a face whose other cell is solid: f of that cell * vn
          // (uniform grids only; checked on the host)
          const Grid<D, NDF, UNI> S(A, hd.data_cell);
#pragma unroll
          for (int k = 0; k < NDF; ++k) mic[k] = S.f(i, k) * u[hd.dir];
        }
        moment_terms<D, NDF>(wt, u, mic, m);
#pragma unroll
        for (int q = 0; q < NW; ++q) acc[h][q] += m[q];
      }
    }
  }

  // reduce and publish the half fluxes
  double tot[MAXH * NW];
  block_sum<MAXH * NW>(&acc[0][0], nh * NW, red, tot);
  bool any_wall = false;
  for (int h = 0; h < nh; ++h) {
    const HalfDesc& hd = A.halves[hb + h];
    if (hd.mode == MI_WALL) {
      any_wall = true;
      const double* bc = &A.bc_prim[hd.bc * NW];
      double pref = (D == 2) ? bc[3] / PI : pow(bc[4] / PI, 1.5);
      double rho = -tot[h * NW + 0] / (pref * tot[h * NW + 1]);
      if (threadIdx.x == 0) {
        rho_sh[h] = rho;
        A.rho_w[hd.slot >> 1] = rho;
      }
      continue;
    }
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < NW; ++q)
        A.half_fw[int64_t(hd.slot) * NW + q] = (q == NW - 1) ? 0.5 * tot[h * NW + q]
                                                             : tot[h * NW + q];
    }
  }
  if (!any_wall) return;
  __syncthreads();
  // second sweep for walls: moments of the Maxwellian with the wall density
#pragma unroll
  for (int h = 0; h < MAXH; ++h)
#pragma unroll
    for (int m = 0; m < NW; ++m) acc[h][m] = 0.0;
  for (int i = threadIdx.x; i < G.n; i += BS) {
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
    const double wt = G.wt(i);
#pragma unroll
    for (int h = 0; h < MAXH; ++h) {
      if (h < nh) {
        const HalfDesc& hd = A.halves[hb + h];
        if (hd.mode != MI_WALL) continue;
        const double ru = hd.rot * u[hd.dir];
        if (!(ru > 0.0)) continue;
        double mic[NDF], m[NW];
        micro_maxwell<D, NDF>(&A.bc_prim[hd.bc * NW], rho_sh[h], true, u, A.K, hd.dir, mic);
        moment_terms<D, NDF>(wt, u, mic, m);
#pragma unroll
        for (int q = 0; q < NW; ++q) acc[h][q] += m[q];
      }
    }
  }
  block_sum<MAXH * NW>(&acc[0][0], nh * NW, red, tot);
  if (threadIdx.x == 0) {
    for (int h = 0; h < nh; ++h) {
      const HalfDesc& hd = A.halves[hb + h];
      if (hd.mode != MI_WALL) continue;
      for (int q = 0; q < NW; ++q)
        A.half_fw[int64_t(hd.slot) * NW + q] = (q == NW - 1) ? 0.5 * tot[h * NW + q]
                                                             : tot[h * NW + q];
    }
  }
}

// ------------------------------------------------------------------ K2: flux + relaxation
// Micro flux into velocity cell i of cell c through all its faces, in face-list order
// (update_flux!/update_domain_flux!/update_micro_flux!, src/Flux/Flux.jl:64-281).
template <int D, int NDF, bool UNI>
__device__ inline void gather_flux(const DevArgs& A, const Grid<D, NDF, UNI>& G, int c,
                                   int eb, int ne, int i, const double* u, const double* f,
                                   const double (*s)[NDF], int li, uint32_t ki,
                                   double* fl) {
  constexpr int NW = D + 2;
  const double dt = A.dt;
  for (int e = 0; e < ne; ++e) {
    const EntryDesc& ed = A.ents[eb + e];
    const double ru = ed.rot * u[ed.dir];
    const bool own = ed.side ? (ru > 0.0) : (ru <= 0.0);
    const double area = ed.area;
    double mic[NDF];
    if (own) {
      micro_recon<D, NDF>(ed.own_mode, f, s, u, dt, ed.xown, ed.dsown, ed.fmid, ed.dir, mic);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        double t = area * mic[k];
        fl[k] += ed.side ? -t : t;
      }
      continue;
    }
    switch (ed.oth_mode) {
      case MI_MAXWELL:
      case MI_WALL: {
        micro_maxwell<D, NDF>(&A.bc_prim[ed.bc * NW], A.rho_w[ed.pair],
                              ed.oth_mode == MI_WALL, u, A.K, ed.dir, mic);
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] += area * mic[k];
        break;
      }
      case MI_SELF_PLAIN: {
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] += area * (f[k] * u[ed.dir]);
        break;
      }
      case MI_INTERP: {
        micro_interp<D, NDF>(f, s, u, dt, ed.xown, ed.fmid, ed.dir, mic);
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] += area * mic[k];
        break;
      }
      default: {
        const Grid<D, NDF, UNI> N(A, ed.oth_cell);
        for_overlap<D, NDF, UNI>(N, i, li, ki, A.lmax, [&](int j, double den) {
          double uj[D], fj[NDF], sj[D][NDF];
#pragma unroll
          for (int t = 0; t < D; ++t) uj[t] = UNI ? u[t] : N.u(j, t);
#pragma unroll
          for (int k = 0; k < NDF; ++k) {
            fj[k] = N.f(j, k);
#pragma unroll
            for (int t = 0; t < D; ++t) sj[t][k] = (ed.oth_mode == MI_PLAIN) ? 0.0 : N.s(j, k, t);
          }
          double mj[NDF];
          micro_recon<D, NDF>(ed.oth_mode, fj, sj, uj, dt, ed.xoth, ed.dsoth, ed.fmid, ed.dir,
                              mj);
#pragma unroll
          for (int k = 0; k < NDF; ++k) {
            double t = area * mj[k];
            if (den != 1.0) t = t / den;
            fl[k] += ed.side ? -t : t;
          }
        });
      }
    }
  }
}

// iterate!(CAIDVM_Marching) (src/Theory/Iterate.jl:96-130) or iterate!(Euler) (:131-162)
// for one cell, fed by the gathered flux (MODE_FUSED), by the flux accumulators
// (MODE_ITER), or only accumulating the flux (MODE_FLUX).
template <int D, int NDF, bool UNI, int MODE, int MARCH, bool LDS>
__global__ __launch_bounds__(BS) void k_flux_relax(DevArgs A, const int32_t* __restrict__ cells,
                                                   int ncells) {
  constexpr int NW = D + 2;
  constexpr int NQ = (D == 2) ? 4 : 3;   // heat-flux partial sums
  const int c = cells[xcd_remap(blockIdx.x, ncells)];
  const Grid<D, NDF, UNI> G(A, c);
  const int n = G.n;
  const int eb = A.ent_off[c];
  const int ne = A.ent_off[c + 1] - eb;
  const double dt = A.dt;
  __shared__ double red[NWAVE * 8];
  extern __shared__ double lds[];
  double* sc = LDS ? lds : &A.scratch[G.off * (NDF + 1)];

  const double* dsc = &A.rec_ds[int64_t(c) * D];
  double V = dsc[0];
#pragma unroll
  for (int d = 1; d < D; ++d) V = V * dsc[d];

  // macroscopic face flux, in face order: flux += fw*area / flux -= fw*area
  double psf[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) psf[q] = A.psflux[int64_t(c) * NW + q];
  if (MODE != MODE_ITER && A.flux_kind == 0) {
    for (int e = 0; e < ne; ++e) {
      const EntryDesc& ed = A.ents[eb + e];
      const double* h0 = &A.half_fw[int64_t(ed.pair) * 2 * NW];
      const double* h1 = h0 + NW;
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        double t;
        if (ed.oth_cell < 0) t = ed.area * (h0[q] + h1[q]);    // update_domain_flux!
        else t = (h0[q] + h1[q]) * ed.area;                     // update_macro_flux!
        psf[q] = ed.side ? psf[q] - t : psf[q] + t;
      }
    }
  }
  if (MODE == MODE_FLUX) {
    if (threadIdx.x == 0)
      for (int q = 0; q < NW; ++q) A.psflux[int64_t(c) * NW + q] = psf[q];
    for (int i = threadIdx.x; i < n; i += BS) {
      double u[D], f[NDF], s[D][NDF], fl[NDF];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        f[k] = G.f(i, k);
#pragma unroll
        for (int d = 0; d < D; ++d) s[d][k] = G.s(i, k, d);
        fl[k] = A.vflux[G.off * NDF + int64_t(k) * n + i];
      }
      gather_flux<D, NDF, UNI>(A, G, c, eb, ne, i, u, f, s, G.lev(i), G.key(i), fl);
#pragma unroll
      for (int k = 0; k < NDF; ++k) A.vflux[G.off * NDF + int64_t(k) * n + i] = fl[k];
    }
    return;
  }

  double wnew[NW], prim_c[NW];
  const double* wold = &A.w[int64_t(c) * NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) wnew[q] = wold[q] + psf[q] * dt / V;
  get_prim<D>(wnew, A.gamma, prim_c);
  const double dtV = dt / V;
  double* out = A.df_out;

  if (MARCH == MARCH_CAIDVM) {
    // pass 1: convection, f* = f + dt/V*flux, and the moments of f*
    double mw[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) mw[q] = 0.0;
    for (int i = threadIdx.x; i < n; i += BS) {
      double u[D], f[NDF], fl[NDF];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
#pragma unroll
      for (int k = 0; k < NDF; ++k) f[k] = G.f(i, k);
      if (MODE == MODE_FUSED) {
        double s[D][NDF];
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fl[k] = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) s[d][k] = G.s(i, k, d);
        }
        gather_flux<D, NDF, UNI>(A, G, c, eb, ne, i, u, f, s, G.lev(i), G.key(i), fl);
      } else {
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fl[k] = A.vflux[G.off * NDF + int64_t(k) * n + i];
          A.vflux[G.off * NDF + int64_t(k) * n + i] = 0.0;
        }
      }
      double fs[NDF], m[NW];
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        fs[k] = f[k] + dtV * fl[k];
        sc[k * n + i] = fs[k];
      }
      moment_terms<D, NDF>(G.wt(i), u, fs, m);
#pragma unroll
      for (int q = 0; q < NW; ++q) mw[q] += m[q];
    }
    double wf[NW], prim[NW];
    block_sum<NW>(mw, NW, red, wf);
    wf[NW - 1] = 0.5 * wf[NW - 1];
    get_prim<D>(wf, A.gamma, prim);
    const double tau = get_tau<D>(prim_c, A.mu, A.omega);

    // pass 2: conservation correction f += M(prim_c) - M(prim); heat flux of the result
    double qa[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) qa[q] = 0.0;
    for (int i = threadIdx.x; i < n; i += BS) {
      double u[D], Mc[2], M[2];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
      maxwell<D>(u, prim_c, A.K, Mc);
      maxwell<D>(u, prim, A.K, M);
      double fs[NDF];
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        fs[k] = sc[k * n + i] + (Mc[k] - M[k]);
        sc[k * n + i] = fs[k];
      }
      sc[NDF * n + i] = Mc[0];
      const double wt = G.wt(i);
      if constexpr (D == 2) {
        double cu = u[0] - prim_c[1], cv = u[1] - prim_c[2];
        double c2 = cu * cu + cv * cv;
        qa[0] += wt * cu * c2 * fs[0];
        qa[1] += wt * cu * fs[1];
        qa[2] += wt * cv * c2 * fs[0];
        qa[3] += wt * cv * fs[1];
      } else {
        double cu = u[0] - prim_c[1], cv = u[1] - prim_c[2], cw = u[2] - prim_c[3];
        double c2 = cu * cu + cv * cv + cw * cw;
        qa[0] += wt * cu * c2 * fs[0];
        qa[1] += wt * cv * c2 * fs[0];
        qa[2] += wt * cw * c2 * fs[0];
      }
    }
    double qs[NQ], q[D];
    block_sum<NQ>(qa, NQ, red, qs);
    if constexpr (D == 2) {
      q[0] = 0.5 * (qs[0] + qs[1]);
      q[1] = 0.5 * (qs[2] + qs[3]);
    } else {
      q[0] = 0.5 * qs[0];
      q[1] = 0.5 * qs[1];
      q[2] = 0.5 * qs[2];
    }

    // pass 3: g+ = M(prim_c) + Shakhov part; f = f*tau/(tau+dt) + dt/(tau+dt)*g+
    const double a1 = tau / (tau + dt), a2 = dt / (tau + dt);
    for (int i = threadIdx.x; i < n; i += BS) {
      double u[D], Fc[2], Sp[2];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
      Fc[0] = sc[NDF * n + i];
      if constexpr (NDF == 2) Fc[1] = Fc[0] * A.K / (2.0 * prim_c[D + 1]);
      shakhov<D>(u, Fc, prim_c, q, A.Pr, A.K, Sp);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        double g = Fc[k] + Sp[k];
        double fn = sc[k * n + i] * a1;
        fn = fn + a2 * g;
        out[G.off * NDF + int64_t(k) * n + i] = fn;
      }
    }
    if (threadIdx.x == 0) {
      double* pr = &A.prim[int64_t(c) * NW];
      double* rs = &A.res[int64_t(c) * 2 * NW];
#pragma unroll
      for (int m = 0; m < NW; ++m) {
        double dl = prim_c[m] - pr[m];
        rs[m] = dl * dl;
        rs[NW + m] = fabs(prim_c[m]);
        pr[m] = prim_c[m];
        A.w[int64_t(c) * NW + m] = wnew[m];
        A.psflux[int64_t(c) * NW + m] = 0.0;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) A.qf[int64_t(c) * D + d] = q[d];
    }
  } else {
    // Euler: prim from the updated w, heat flux of f^n, then the implicit update
    const double tau = get_tau<D>(prim_c, A.mu, A.omega);
    double qa[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) qa[q] = 0.0;
    for (int i = threadIdx.x; i < n; i += BS) {
      double u[D], f[NDF], fl[NDF];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
#pragma unroll
      for (int k = 0; k < NDF; ++k) f[k] = G.f(i, k);
      if (MODE == MODE_FUSED) {
        double s[D][NDF];
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fl[k] = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) s[d][k] = G.s(i, k, d);
        }
        gather_flux<D, NDF, UNI>(A, G, c, eb, ne, i, u, f, s, G.lev(i), G.key(i), fl);
      } else {
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fl[k] = A.vflux[G.off * NDF + int64_t(k) * n + i];
          A.vflux[G.off * NDF + int64_t(k) * n + i] = 0.0;
        }
      }
#pragma unroll
      for (int k = 0; k < NDF; ++k) sc[k * n + i] = fl[k];
      const double wt = G.wt(i);
      if constexpr (D == 2) {
        double cu = u[0] - prim_c[1], cv = u[1] - prim_c[2];
        double c2 = cu * cu + cv * cv;
        qa[0] += wt * cu * c2 * f[0];
        qa[1] += wt * cu * f[1];
        qa[2] += wt * cv * c2 * f[0];
        qa[3] += wt * cv * f[1];
      } else {
        double cu = u[0] - prim_c[1], cv = u[1] - prim_c[2], cw = u[2] - prim_c[3];
        double c2 = cu * cu + cv * cv + cw * cw;
        qa[0] += wt * cu * c2 * f[0];
        qa[1] += wt * cv * c2 * f[0];
        qa[2] += wt * cw * c2 * f[0];
      }
    }
    double qs[NQ], q[D];
    block_sum<NQ>(qa, NQ, red, qs);
    if constexpr (D == 2) {
      q[0] = 0.5 * (qs[0] + qs[1]);
      q[1] = 0.5 * (qs[2] + qs[3]);
    } else {
      q[0] = 0.5 * qs[0];
      q[1] = 0.5 * qs[1];
      q[2] = 0.5 * qs[2];
    }
    const double tpd = tau + dt, a2 = dt / (tau + dt);
    for (int i = threadIdx.x; i < n; i += BS) {
      double u[D], M[2], Sp[2];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
      maxwell<D>(u, prim_c, A.K, M);
      shakhov<D>(u, M, prim_c, q, A.Pr, A.K, Sp);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        double F = M[k] + Sp[k];
        double fx = (G.f(i, k) + dtV * sc[k * n + i]) * tau / tpd + a2 * F;
        out[G.off * NDF + int64_t(k) * n + i] = fx;
      }
    }
    if (threadIdx.x == 0) {
      double* pr = &A.prim[int64_t(c) * NW];
      double* rs = &A.res[int64_t(c) * 2 * NW];
#pragma unroll
      for (int m = 0; m < NW; ++m) {
        double dl = prim_c[m] - pr[m];
        rs[m] = dl * dl;
        rs[NW + m] = fabs(prim_c[m]);
        pr[m] = prim_c[m];
        A.w[int64_t(c) * NW + m] = wnew[m];
        A.psflux[int64_t(c) * NW + m] = 0.0;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) A.qf[int64_t(c) * D + d] = q[d];
    }
  }
}

// ------------------------------------------------------------------ K3: macro slope
template <int D, int NDF, bool UNI>
__global__ __launch_bounds__(BS) void k_macro_slope(DevArgs A, const int32_t* __restrict__ cells,
                                                    int ncells, double* __restrict__ sw) {
  constexpr int NW = D + 2;
  const int c = cells[blockIdx.x];
  const Grid<D, NDF, UNI> G(A, c);
  __shared__ double red[NWAVE * NW];
  for (int d = 0; d < D; ++d) {
    double acc[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) acc[q] = 0.0;
    for (int i = threadIdx.x; i < G.n; i += BS) {
      double u[D], s[NDF], m[NW];
#pragma unroll
      for (int t = 0; t < D; ++t) u[t] = G.u(i, t);
#pragma unroll
      for (int k = 0; k < NDF; ++k) s[k] = G.s(i, k, d);
      moment_terms<D, NDF>(G.wt(i), u, s, m);
#pragma unroll
      for (int q = 0; q < NW; ++q) acc[q] += m[q];
    }
    double tot[NW];
    block_sum<NW>(acc, NW, red, tot);
    if (threadIdx.x == 0)
      for (int q = 0; q < NW; ++q)
        sw[(int64_t(c) * NW + q) * D + d] = (q == NW - 1) ? 0.5 * tot[q] : tot[q];
  }
}

// ------------------------------------------------------------------ launchers
template <int D, int NDF, bool UNI>
static hipError_t launch_slope_t(const DevArgs& A, const int32_t* cells, int n, int maxh,
                                 hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (maxh <= 8)
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, 8>), dim3(n), dim3(BS), 0, st, A, cells, n);
  else if (maxh <= 16)
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, 16>), dim3(n), dim3(BS), 0, st, A, cells, n);
  else
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, 24>), dim3(n), dim3(BS), 0, st, A, cells, n);
  return hipGetLastError();
}

template <int D, int NDF, bool UNI, int MODE, int MARCH>
static hipError_t launch_fr_t(const DevArgs& A, const int32_t* cells, int n, size_t lds,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (lds > 0)
    hipLaunchKernelGGL((k_flux_relax<D, NDF, UNI, MODE, MARCH, true>), dim3(n), dim3(BS), lds,
                       st, A, cells, n);
  else
    hipLaunchKernelGGL((k_flux_relax<D, NDF, UNI, MODE, MARCH, false>), dim3(n), dim3(BS), 0,
                       st, A, cells, n);
  return hipGetLastError();
}

template <int D, int NDF, bool UNI>
static hipError_t launch_fr_d(const DevArgs& A, const int32_t* cells, int n, int mode,
                              int march, size_t lds, hipStream_t st) {
  if (mode == MODE_FLUX) return launch_fr_t<D, NDF, UNI, MODE_FLUX, MARCH_CAIDVM>(A, cells, n, 0, st);
  if (mode == MODE_ITER) {
    if (march == MARCH_CAIDVM)
      return launch_fr_t<D, NDF, UNI, MODE_ITER, MARCH_CAIDVM>(A, cells, n, lds, st);
    return launch_fr_t<D, NDF, UNI, MODE_ITER, MARCH_EULER>(A, cells, n, lds, st);
  }
  if (march == MARCH_CAIDVM)
    return launch_fr_t<D, NDF, UNI, MODE_FUSED, MARCH_CAIDVM>(A, cells, n, lds, st);
  return launch_fr_t<D, NDF, UNI, MODE_FUSED, MARCH_EULER>(A, cells, n, lds, st);
}

hipError_t launch_slope(const DevArgs& A, int dim, bool uni, const int32_t* cells, int n,
                        int maxh, hipStream_t st) {
  if (dim == 2) {
    return uni ? launch_slope_t<2, 2, true>(A, cells, n, maxh, st)
               : launch_slope_t<2, 2, false>(A, cells, n, maxh, st);
  }
  return uni ? launch_slope_t<3, 1, true>(A, cells, n, maxh, st)
             : launch_slope_t<3, 1, false>(A, cells, n, maxh, st);
}

hipError_t launch_flux_relax(const DevArgs& A, int dim, bool uni, const int32_t* cells, int n,
                             int mode, int march, size_t lds, hipStream_t st) {
  if (dim == 2) {
    return uni ? launch_fr_d<2, 2, true>(A, cells, n, mode, march, lds, st)
               : launch_fr_d<2, 2, false>(A, cells, n, mode, march, lds, st);
  }
  return uni ? launch_fr_d<3, 1, true>(A, cells, n, mode, march, lds, st)
             : launch_fr_d<3, 1, false>(A, cells, n, mode, march, lds, st);
}

hipError_t launch_macro_slope(const DevArgs& A, int dim, bool uni, const int32_t* cells, int n,
                              double* sw, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (dim == 2) {
    if (uni) hipLaunchKernelGGL((k_macro_slope<2, 2, true>), dim3(n), dim3(BS), 0, st, A, cells, n, sw);
    else hipLaunchKernelGGL((k_macro_slope<2, 2, false>), dim3(n), dim3(BS), 0, st, A, cells, n, sw);
  } else {
    if (uni) hipLaunchKernelGGL((k_macro_slope<3, 1, true>), dim3(n), dim3(BS), 0, st, A, cells, n, sw);
    else hipLaunchKernelGGL((k_macro_slope<3, 1, false>), dim3(n), dim3(BS), 0, st, A, cells, n, sw);
  }
  return hipGetLastError();
}

}  // namespace kah

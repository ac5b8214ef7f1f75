// kah_device.h -- gfx950 device code of the per-step phase-space update (included by the
// per-dimension kernel translation units kah_slope_{2d,3d}.hip and kah_flux_{2d,3d}.hip).
//
// K1 k_slope_chunk one workgroup per (cell, chunk of velocity cells) of one projection
//                  depth: velocity slopes and per-chunk partial half fluxes
//    k_slope       one workgroup per cell, for cells with a Maxwellian wall face (the wall
//                  density needs the whole cell's sums before the half flux)
// K2 k_flux_relax  one workgroup per fluid cell (MODE_FLUX / MODE_ITER / MODE_FUSED)
// K3 k_macro_slope on demand: PsData.sw = moments of sdf (src/Flux/Slope.jl:1022-1036)
// and, off the per-step path: k_overlap_map (cross-grid overlap maps, once per topology),
// k_mirror (ghost-exchange buffers), k_restart (restart arrays), k_vs_res / k_vs_flags /
// k_vs_remap / k_vs_project (velocity-space adaptation).
//
// Velocity index fastest in memory, so every load and store of f, sdf and the flux is
// coalesced; per-cell moment sums are reduced deterministically (fixed-order trees through
// wave shuffles or LDS, then fixed-order sums of the partials), with no atomics anywhere.
// Private arrays are only ever indexed with compile-time indices (direction-dependent picks
// are selects), so nothing spills to scratch memory.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kah_internal.h"
#include "kah_kernels.h"

namespace kah {

// ------------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup sum of the first cnt of CNT per-thread partials. out[q] (LDS) receives the
// total of value q; red needs 2*NWV*red_stride(CNT) doubles of LDS. Three barriers, fixed
// order; the sum is compensated (block_totals).
template <int CNT, int NWV>
__device__ inline void block_reduce(const double (&vals)[CNT], int cnt, double* red,
                                    double* out);

// Compensated sums. two_sum(a, b) returns a + b rounded and the exact rounding error
// (Knuth's TwoSum, six additions, no branch); it is symmetric in a and b, so two lanes that add
// each other's values get bit-identical results.
__device__ inline double two_sum(double a, double b, double& err) {
  const double s = a + b;
  const double bb = s - a;
  err = (a - (s - bb)) + (b - bb);
  return s;
}

// Sum over the 64 lanes of a wave of (s, e) pairs (a value and its running error): an xor
// butterfly whose every level adds with two_sum and carries the errors, so every lane ends
// with the same compensated total s + e.
__device__ inline void wave_sum2(double& s, double& e) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double so = __shfl_xor(s, o, 64);
    const double eo = __shfl_xor(e, o, 64);
    double err;
    s = two_sum(s, so, err);
    e = (e + eo) + err;
  }
}

constexpr int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}
// per-wave stride of block_totals' LDS (in (s, e) pairs)
constexpr int red_stride(int cnt) { return pow2_at_least(cnt) <= 64 ? pow2_at_least(cnt) : cnt; }

// Compensated reduce-scatter of P (a power of two, at most 64) values over the 64 lanes of a
// wave. At the level with xor offset o (32, 16, ..., 64/P) every lane keeps one half of its
// current values (the upper half when its lane bit o is set) and adds the partner's copy of
// that half with two_sum; after log2(P) levels lane L holds, for value (L >> log2(64/P)), the
// sum over the lanes that share its high bits, and a plain two_sum butterfly over the low
// bits finishes it. P/2 + P/4 + ... + 1 + log2(64/P) exchanges instead of P*6.
// The levels are one template instance per value count M, each writing fresh arrays of M/2,
// so that every index is a compile-time constant: a loop over the levels, left to the
// unroller, stays rolled for 32 values and puts the arrays in scratch memory.
template <int M>
__device__ inline void scatter_levels(const double (&v)[M], const double (&e)[M], double& s0,
                                      double& e0, int lane, int o) {
  if constexpr (M == 1) {
    s0 = v[0];
    e0 = e[0];
  } else {
    const bool up = (lane & o) != 0;
    double nv[M / 2], ne[M / 2];
#pragma unroll
    for (int j = 0; j < M / 2; ++j) {
      const double lo = v[j], hi = v[M / 2 + j], loe = e[j], hie = e[M / 2 + j];
      const double keep = up ? hi : lo, ke = up ? hie : loe;
      const double send = up ? lo : hi, se = up ? loe : hie;
      const double recv = __shfl_xor(send, o, 64), re = __shfl_xor(se, o, 64);
      double err;
      nv[j] = two_sum(keep, recv, err);
      ne[j] = (ke + re) + err;
    }
    scatter_levels<M / 2>(nv, ne, s0, e0, lane, o >> 1);
  }
}

template <int P>
__device__ inline void wave_scatter2(double (&v)[P], double (&e)[P]) {
  static_assert(P >= 1 && P <= 64 && (P & (P - 1)) == 0, "power of two up to 64");
  const int lane = threadIdx.x & 63;
  double s0, e0;
  scatter_levels<P>(v, e, s0, e0, lane, 32);
#pragma unroll
  for (int o = 32 / P; o > 0; o >>= 1) {
    const double so = __shfl_xor(s0, o, 64), eo = __shfl_xor(e0, o, 64);
    double err;
    s0 = two_sum(s0, so, err);
    e0 = (e0 + eo) + err;
  }
  v[0] = s0;
  e[0] = e0;
}

// Compensated sum of the first cnt of CNT per-thread values over a workgroup of NWV waves:
// per wave a compensated reduce-scatter (or, for more than 64 values, one two_sum butterfly
// per value), then the waves' (s, e) pairs added in wave order with two_sum; thread q < cnt
// returns the total of value q. red needs 2*NWV*RED_STRIDE(CNT) doubles of LDS (the power of
// two at least CNT, or CNT above 64). Two barriers, fixed order.
template <int CNT, int NWV>
__device__ inline double block_totals(const double (&vals)[CNT], int cnt, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int P = pow2_at_least(CNT);
  if constexpr (P <= 64) {
    double v[P], e[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
      v[q] = q < CNT ? vals[q] : 0.0;
      e[q] = 0.0;
    }
    wave_scatter2<P>(v, e);
    constexpr int SH = (P == 1) ? 6 : (P == 2) ? 5 : (P == 4) ? 4 : (P == 8) ? 3 : (P == 16) ? 2 : (P == 32) ? 1 : 0;
    if ((lane & ((1 << SH) - 1)) == 0) {
      const int q = lane >> SH;
      red[(wv * P + q) * 2] = v[0];
      red[(wv * P + q) * 2 + 1] = e[0];
    }
  } else {
#pragma unroll
    for (int q = 0; q < CNT; ++q) {
      if (q < cnt) {
        double v = vals[q], e = 0.0;
        wave_sum2(v, e);
        if (lane == 0) {
          red[(wv * CNT + q) * 2] = v;
          red[(wv * CNT + q) * 2 + 1] = e;
        }
      }
    }
  }
  __syncthreads();
  constexpr int STRIDE = red_stride(CNT);
  double out = 0.0;
  if (int(threadIdx.x) < cnt) {
    const int q = threadIdx.x;
    double S = red[q * 2], E = red[q * 2 + 1];
#pragma unroll
    for (int w = 1; w < NWV; ++w) {
      double err;
      S = two_sum(S, red[(w * STRIDE + q) * 2], err);
      E = (E + red[(w * STRIDE + q) * 2 + 1]) + err;
    }
    out = S + E;
  }
  __syncthreads();
  return out;
}

template <int CNT, int NWV>
__device__ inline void block_reduce(const double (&vals)[CNT], int cnt, double* red,
                                    double* out) {
  const double t = block_totals<CNT, NWV>(vals, cnt, red);
  if (int(threadIdx.x) < cnt) out[threadIdx.x] = t;
  __syncthreads();
}

// XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs; give every XCD
// a contiguous run of cells so that face neighbours (close in Morton order) share its L2.
__device__ inline int xcd_remap(int i, int n) {
  const int x = i & 7, j = i >> 3;
  const int q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + j;
}

// a[d] for a runtime direction d, as selects over compile-time indices. With three
// candidates the optimizer would otherwise fold the select chain back into an indexed load
// from a stack copy of a (private memory); the empty asm makes each candidate an opaque
// register value.
template <int D, class T>
__device__ inline T pick(const T (&a)[D], int d) {
  T r = a[0];
#pragma unroll
  for (int t = 1; t < D; ++t) {
    T x = a[t];
    if constexpr (D > 2) __asm__("" : "+v"(x));
    r = (d == t) ? x : r;
  }
  return r;
}

// A value every lane holds alike (read from LDS after a barrier) as a wave-uniform scalar:
// kept in SGPRs instead of one VGPR pair per lane.
__device__ inline double uniform(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(b));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(b >> 32));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
template <int D>
__device__ inline MaxPre<D> uniform(const MaxPre<D>& m) {
  MaxPre<D> r;
#pragma unroll
  for (int q = 0; q < D + 2; ++q) r.p[q] = uniform(m.p[q]);
  r.a0 = uniform(m.a0);
  r.two_l = uniform(m.two_l);
  r.r2l = uniform(m.r2l);
  r.s0 = uniform(m.s0);
  return r;
}

// Arithmetic type of the per-velocity work for a storage type: FP64 for FP64 storage, FP32
// for the FP32 form (f stored and advected in FP32; moment sums, prim, tau, q and the
// Maxwellian parameters in FP64).
template <class FT>
struct Arith {
  using type = double;
};
template <>
struct Arith<float> {
  using type = float;
};

// ------------------------------------------------------------------ grid access
template <int D, int NDF, bool UNI, class FT>
struct Grid {
  const DevArgs& A;
  int64_t off;   // first velocity cell of the cell (state arrays)
  int64_t g;     // first entry of the cell's grid in the geometry table (adapted grids)
  int n;
  // per-cell base pointers (uniform over the workgroup): every access is base + a 32-bit
  // offset, which the loads take as a scalar base and one vector register
  const FT* fb;
  const FT* sb;
  FT* swb;
  FT* ofb;
  FT* ob;
  const double* gub;
  const double* gwb;
  __device__ Grid(const DevArgs& a, int c) : A(a) {
    if constexpr (UNI) {
      n = a.uni_n;
      off = int64_t(c) * n;
      g = 0;
    } else {
      off = a.voff[c];
      n = int(a.voff[c + 1] - off);
      g = a.gof[c];
    }
    fb = reinterpret_cast<const FT*>(a.df) + off * NDF;
    sb = reinterpret_cast<const FT*>(a.sdf_r) + off * NDF * D;
    swb = reinterpret_cast<FT*>(a.sdf) + off * NDF * D;
    ofb = reinterpret_cast<FT*>(a.oflux) + off * NDF * D;
    ob = reinterpret_cast<FT*>(a.df_out) + off * NDF;
    gub = a.gu + goff() * D;
    gwb = a.gw + goff();
  }
  __device__ int64_t goff() const { return UNI ? 0 : g; }
  // velocity cell handled at thread position p: the identity, or the wave-tile order of a
  // shared tensor-product grid (the 64 positions of a wave form one tile). Any bijection of
  // [0, n) gives the same per-cell results; the order only makes the upwind branches
  // uniform per wave.
  __device__ int vi(int p) const {
    if constexpr (UNI) {
      if (A.vt_mode != 0) {
        const int T = __builtin_amdgcn_readfirstlane(p >> 6);
        const int l = p & 63;
        const int o = (A.vt_mode == 2) ? (l >> 3) * A.vt_n0 + (l & 7)
                                       : ((l >> 4) * A.vt_n1 + ((l >> 2) & 3)) * A.vt_n0 + (l & 3);
        return A.vt_base[T] + o;
      }
    }
    return p;
  }
  // element x of a per-cell array as base + 32-bit byte offset
  template <class T>
  static __device__ T& at(T* b, int x) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(b) + uint32_t(x) * uint32_t(sizeof(T)));
  }
  template <class T>
  static __device__ const T& at(const T* b, int x) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(b) +
                                       uint32_t(x) * uint32_t(sizeof(T)));
  }
  __device__ double u(int i, int d) const { return at(gub, d * n + i); }
  __device__ double wt(int i) const { return at(gwb, i); }
  __device__ int lev(int i) const { return UNI ? 0 : int(A.glev[g + i]); }
  __device__ uint32_t key(int i) const { return UNI ? uint32_t(i) : A.gkey[g + i]; }
  // where key(i) is stored (adapted grids): the overlap walks load it only when a
  // neighbour has no overlap map
  __device__ const uint32_t* keyp(int i) const { return UNI ? nullptr : A.gkey + g + i; }
  // f, sdf and the next f are stored as FT (double, or float in the FP32 form); the
  // accessors widen to double, and the FP32 form's arithmetic narrows where it advects
  __device__ double f(int i, int k) const { return double(at(fb, k * n + i)); }
  __device__ double s(int i, int k, int d) const { return double(at(sb, (d * NDF + k) * n + i)); }
  // the outgoing micro flux of velocity cell i through the cell's face in direction d
  __device__ void set_of(int i, int k, int d, double v) const {
    at(ofb, (d * NDF + k) * n + i) = FT(v);
  }
  __device__ void set_s(int i, int k, int d, double v) const {
    at(swb, (d * NDF + k) * n + i) = FT(v);
  }
  __device__ void set_out(int i, int k, double v) const { at(ob, k * n + i) = FT(v); }
  // a value as the storage holds it
  static __device__ double stored(double v) { return double(FT(v)); }
};

// Calls fn(j, den, 1/den) for each velocity cell j of grid G that overlaps cell i (level li,
// key ki) of another grid on the same root lattice: den = 1 when j is coarser or as fine as
// i, else 2^(D*(lj-li)) (a power of two: 1/den is exact). This is the pairing of the two-pointer walks of diff_vs! and
// update_micro_flux! (src/Flux/Slope.jl:29-64, src/Flux/Flux.jl:151-344), in the same order.
template <int D, int NDF, bool UNI, class FT, class Fn>
__device__ inline void for_overlap(const Grid<D, NDF, UNI, FT>& G, int i, int li, uint32_t ki,
                                   int lmax, Fn&& fn) {
  if constexpr (UNI) {
    fn(i, 1.0, 1.0);
  } else {
    int lo = 0, hi = G.n - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (G.key(mid) <= ki) lo = mid; else hi = mid - 1;
    }
    int j = lo;
    int lj = G.lev(j);
    if (lj <= li) {
      fn(j, 1.0, 1.0);
      return;
    }
    const uint32_t end = ki + (1u << (D * (lmax - li)));
    for (; j < G.n && G.key(j) < end; ++j) {
      const int e = D * (G.lev(j) - li);
      fn(j, ldexp(1.0, e), ldexp(1.0, -e));
    }
  }
}

// for_overlap through the overlap map `slot` of (owner grid, neighbour grid): one coalesced
// load instead of a binary search over the neighbour's keys; the same cells and factors in
// the same order. A neighbour on the owner's grid (MAP_SAME) meets cell i itself; without a
// map (slot -1) the search is done in place.
template <int D, int NDF, bool UNI, class FT, class Fn>
__device__ inline void for_overlap_m(const DevArgs& A, const Grid<D, NDF, UNI, FT>& N, int slot,
                                     int i, int li, const uint32_t* kp, Fn&& fn) {
  if constexpr (UNI) {
    fn(i, 1.0, 1.0);
  } else {
    if (slot == MAP_SAME) {
      fn(i, 1.0, 1.0);
      return;
    }
    if (slot >= 0) {
      const uint32_t code = A.ovmap[A.map_base[slot] + i];
      const int j = int(code >> 8), cnt = int(code & 255u);
      if (cnt == 0) {
        fn(j, 1.0, 1.0);
        return;
      }
      for (int t = 0; t < cnt; ++t) {
        const int e = D * (N.lev(j + t) - li);
        fn(j + t, ldexp(1.0, e), ldexp(1.0, -e));
      }
      return;
    }
    for_overlap<D, NDF, UNI, FT>(N, i, li, *kp, A.lmax, fn);
  }
}

// Overlap map of one (owner grid, neighbour grid) job, one velocity cell of the owner per
// thread (the search of for_overlap, done once per topology).
template <int D, int NDF>
__global__ __launch_bounds__(256) void k_overlap_map(DevArgs A, const OvJob* __restrict__ jobs,
                                                     int njobs, uint32_t* __restrict__ map) {
  const OvJob jb = jobs[blockIdx.x];
  const int64_t go = jb.own, no = jb.nb;
  const int gn = jb.n_own, nn = jb.n_nb;
  const uint32_t* gk = A.gkey + go;
  const uint32_t* nk = A.gkey + no;
  const int8_t* gl = A.glev + go;
  const int8_t* nl = A.glev + no;
  for (int i = threadIdx.x; i < gn; i += 256) {
    const uint32_t ki = gk[i];
    const int li = gl[i];
    int lo = 0, hi = nn - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (nk[mid] <= ki) lo = mid; else hi = mid - 1;
    }
    uint32_t cnt = 0;
    if (nl[lo] > li) {
      const uint32_t end = ki + (1u << (D * (A.lmax - li)));
      for (int t = lo; t < nn && nk[t] < end; ++t) ++cnt;
    }
    map[jb.base + i] = (uint32_t(lo) << 8) | cnt;
  }
}

template <int D, int NDF>
inline hipError_t launch_overlap_map_t(const DevArgs& A, const OvJob* jobs, int njobs,
                                       uint32_t* map, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_overlap_map<D, NDF>), dim3(njobs), dim3(256), 0, st, A, jobs, njobs, map);
  return hipGetLastError();
}

// Micro flux of one velocity cell: value times the normal velocity.
// LIMITED is positivity_preserving_reconstruct (src/Flux/CAIDVM.jl:127-141); UNLIMITED the
// plain reconstruction used next to solid cells and at domain walls/inflow (CAIDVM.jl:13-14,
// :110-112); PLAIN f*vn (UniformOutflow, CAIDVM.jl:53-67).
// T is the arithmetic type: double, or float in the FP32 advection form. There the offset
// face midpoint - reconstruction point and the limiter's cell size come from the
// descriptor's float copies (dxf, dsf; the offset was formed in double before the rounding,
// being a difference of two coordinates of order one).
template <int D, int NDF, class T>
__device__ inline void micro_recon(int mode, const T (&f)[NDF], const T (&s)[D][NDF],
                                   const T (&u)[D], T dt, const double* xrec,
                                   const double* dsrec, const double* fmid, const float* dxf,
                                   const float* dsf, int dir, T (&out)[NDF]) {
  const T vn = pick<D>(u, dir);
  if (mode == MI_PLAIN || mode == MI_SELF_PLAIN) {
#pragma unroll
    for (int k = 0; k < NDF; ++k) out[k] = f[k] * vn;
    return;
  }
  T dx[D];
#pragma unroll
  for (int t = 0; t < D; ++t) {
    if constexpr (sizeof(T) == 8) dx[t] = fmid[t] - u[t] * dt - xrec[t];
    else dx[t] = dxf[t] - u[t] * dt;
  }
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    T s_dx = dx[0] * s[0][k];
#pragma unroll
    for (int t = 1; t < D; ++t) s_dx = s_dx + dx[t] * s[t][k];
    if (mode == MI_LIMITED) {
      T s_abs = desc_val<T>(dsrec[0], dsf[0]) * fabs(s[0][k]);
#pragma unroll
      for (int t = 1; t < D; ++t) s_abs = s_abs + desc_val<T>(dsrec[t], dsf[t]) * fabs(s[t][k]);
      const T a = f[k] - T(MACH_EPS), b = T(0.5) * s_abs + T(EPS);
      // b >= EPS > 0, so |a| >= b with b finite means |a/b| >= 1 and the limiter is 1
      // exactly; the quotient is only formed where it can be below 1 (or NaN)
      T lim = T(1);
      if (!(fabs(a) >= b && b < T(HUGE_VAL))) lim = jl_min(T(fabs(a / b)), T(1));
      out[k] = (f[k] + lim * s_dx) * vn;
    } else {
      out[k] = (f[k] + s_dx) * vn;
    }
  }
}

// InterpolatedOutflow incoming micro flux (CAIDVM.jl:71-94): the ghost state mirrored
// through the face, reconstructed at the face from the cell's own slopes.
template <int D, int NDF, class T>
__device__ inline void micro_interp(const T (&f)[NDF], const T (&s)[D][NDF], const T (&u)[D],
                                    T dt, const double* xh, const double* fmid, int dir,
                                    T (&out)[NDF]) {
  double xt[D], xhd[D];
#pragma unroll
  for (int t = 0; t < D; ++t) {
    xt[t] = 2.0 * fmid[t] - xh[t];
    xhd[t] = xh[t];
  }
  T ndx[D];
#pragma unroll
  for (int t = 0; t < D; ++t) {
    if constexpr (sizeof(T) == 8) ndx[t] = fmid[t] - u[t] * dt - xt[t];
    else ndx[t] = T(fmid[t] - xt[t]) - u[t] * dt;
  }
  const T shift = T(pick<D>(xt, dir) - pick<D>(xhd, dir));
  const T vn = pick<D>(u, dir);
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    T sk[D];
#pragma unroll
    for (int t = 0; t < D; ++t) sk[t] = s[t][k];
    T tdf = f[k] + shift * pick<D>(sk, dir);
    T s_dx = ndx[0] * s[0][k];
#pragma unroll
    for (int t = 1; t < D; ++t) s_dx = s_dx + ndx[t] * s[t][k];
    out[k] = (tdf + s_dx) * vn;
  }
}

// Maxwellian inflow state (bc prim, density rho for walls) times vn (evaluated in double).
template <int D, int NDF, class T>
__device__ inline void micro_maxwell(const double* bc, double rho, bool wall, const T (&u)[D],
                                     double K, int dir, T (&out)[NDF]) {
  double p[D + 2], ud[D];
#pragma unroll
  for (int m = 0; m < D + 2; ++m) p[m] = bc[m];
  if (wall) p[0] = rho;
#pragma unroll
  for (int t = 0; t < D; ++t) ud[t] = u[t];
  double M[2];
  maxwell<D>(ud, p, K, M);
  const T vn = pick<D>(u, dir);
#pragma unroll
  for (int k = 0; k < NDF; ++k) out[k] = T(M[k]) * vn;
}

// ------------------------------------------------------------------ K1: slopes + half fluxes
// Velocity-space slope of one direction (update_slope!/update_slope_transverse_level!,
// src/Flux/Slope.jl:458-945), the diff_vs!/diff_vs_transverse! sums (:29-64, :278-332)
// taken as key-interval overlaps.
template <class T>
__device__ inline T div_ds(T a, double ds, double rds, float rds_f) {
  if constexpr (sizeof(T) == 8) return div_by(a, ds, rds);
  else return a * rds_f;
}

template <int D, int NDF, bool UNI, class FT, class T>
__device__ inline void side_sum(const DevArgs& A, int c, int n_own, int beg, int cnt,
                                double dsx, double rdsx, float rdsx_f, int i, int li,
                                const uint32_t* kp, const T (&f)[NDF], T (&a)[NDF]) {
#pragma unroll
  for (int k = 0; k < NDF; ++k) a[k] = T(0);
  for (int e = 0; e < cnt; ++e) {
    const SlopeNb nb = A.slnb[beg + e];
    const Grid<D, NDF, UNI, FT> N(A, nb.cell);
    for_overlap_m<D, NDF, UNI, FT>(A, N, nb.map, i, li, kp,
                                   [&](int j, double den, double rden) {
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        T pj = T(N.f(j, k));
        if (nb.proj) {
#pragma unroll
          for (int t = 0; t < D; ++t) pj += desc_val<T>(nb.dm[t], nb.dm_f[t]) * T(N.s(j, k, t));
        }
        T term = f[k] - pj;
        if (den != 1.0) term = term * T(rden);
        a[k] += div_ds(term, dsx, rdsx, rdsx_f);
      }
    });
  }
}

template <int D, int NDF, bool UNI, class FT, class T>
__device__ inline void slope_dir(const DevArgs& A, const Grid<D, NDF, UNI, FT>& G, int c, int d,
                                 int i, int li, const uint32_t* kp, const T (&f)[NDF],
                                 T (&s)[NDF]) {
  const SlopeDesc sd = A.sl[c * D + d];
  if (sd.mode == SL_SKIP) {
#pragma unroll
    for (int k = 0; k < NDF; ++k) s[k] = T(G.s(i, k, d));
    return;
  }
  if (sd.mode == SL_ZERO) {
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      s[k] = T(0);
      G.set_s(i, k, d, 0.0);
    }
    return;
  }
  T aL[NDF];
  side_sum<D, NDF, UNI, FT, T>(A, c, G.n, sd.nb_beg[0], sd.nb_cnt[0], sd.ds[0], sd.rds[0],
                               sd.rds_f[0], i, li, kp, f, aL);
  const T ic0 = desc_val<T>(sd.inv_cnt[0], sd.inv_cnt_f[0]);
  if (sd.mode == SL_INNER) {
    T aR[NDF];
    side_sum<D, NDF, UNI, FT, T>(A, c, G.n, sd.nb_beg[1], sd.nb_cnt[1], sd.ds[1], sd.rds[1],
                                 sd.rds_f[1], i, li, kp, f, aR);
    const T ic1 = desc_val<T>(sd.inv_cnt[1], sd.inv_cnt_f[1]);
#pragma unroll
    for (int k = 0; k < NDF; ++k) s[k] = minmod_sel(aL[k] * ic0, aR[k] * ic1);
  } else {
#pragma unroll
    for (int k = 0; k < NDF; ++k) s[k] = aL[k] * ic0;
  }
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    s[k] = T(G.stored(s[k]));   // the flux uses the slope as stored
    G.set_s(i, k, d, s[k]);
  }
}

// Moments of one velocity cell's contribution to a half flux (the cell's own upwind micro
// flux through one face, micro_to_macro of src/Flux/CAIDVM.jl:119), or the two sums of the
// wall density (calc_ρw, src/Theory/Math.jl:251-282) in the first sweep of a wall.
template <int D, int NDF, bool UNI, class FT, class T>
__device__ inline void half_terms(const DevArgs& A, const Grid<D, NDF, UNI, FT>& G, int c,
                                  const HalfDesc& hd, int i, const T (&u)[D], const T (&f)[NDF],
                                  const T (&s)[D][NDF], T wt, double (&m)[D + 2]) {
  constexpr int NW = D + 2;
#pragma unroll
  for (int q = 0; q < NW; ++q) m[q] = 0.0;
  const T ud = pick<D>(u, hd.dir);
  const T ru = desc_val<T>(hd.rot, hd.rot_f) * ud;
  const bool in = hd.sel ? (ru > T(0)) : (ru <= T(0));
  if (hd.mode == MI_WALL) {
    const double* bc = &A.bc_prim[hd.bc * NW];
    if (in) {
      T c2 = T(0);
#pragma unroll
      for (int t = 0; t < D; ++t) {
        T cu = u[t] - T(bc[1 + t]);
        c2 = (t == 0) ? cu * cu : c2 + cu * cu;
      }
      m[1] = wt * ud * exp(-T(bc[D + 1]) * c2);
    } else {
      T dx[D];
#pragma unroll
      for (int t = 0; t < D; ++t) {
        if constexpr (sizeof(T) == 8) dx[t] = hd.fmid[t] - u[t] * A.dt - hd.xrec[t];
        else dx[t] = hd.dx_f[t] - u[t] * T(A.dt);
      }
      T hrec = dx[0] * s[0][0];
#pragma unroll
      for (int t = 1; t < D; ++t) hrec = hrec + dx[t] * s[t][0];
      hrec = f[0] + hrec;
      m[0] = wt * ud * hrec;
    }
    return;
  }
  if (!in) return;
  T mic[NDF];
  if (hd.mode == MI_MAXWELL) {
    micro_maxwell<D, NDF>(&A.bc_prim[hd.bc * NW], 0.0, false, u, A.K, hd.dir, mic);
  } else if (hd.mode == MI_INTERP) {
    micro_interp<D, NDF>(f, s, u, T(A.dt), hd.xrec, hd.fmid, hd.dir, mic);
  } else if (hd.data_cell == c) {
    micro_recon<D, NDF>(hd.mode, f, s, u, T(A.dt), hd.xrec, hd.dsrec, hd.fmid, hd.dx_f, hd.ds_f,
                        hd.dir, mic);
    // the cell's own outgoing flux through an inner face (a domain face's second half is the
    // ghost side: the cell's data flowing in)
    if (hd.bc < 0) {
#pragma unroll
      for (int k = 0; k < NDF; ++k) G.set_of(i, k, hd.dir, mic[k]);
    }
  } else {
    // the there side of a face whose other cell is solid: f of that cell * vn (shared
    // velocity grids only; checked on the host)
    const Grid<D, NDF, UNI, FT> S(A, hd.data_cell);
#pragma unroll
    for (int k = 0; k < NDF; ++k) mic[k] = T(S.f(i, k)) * ud;
  }
  moment_terms<D, NDF, T>(wt, u, mic, m);
}

// slope_dir of a regular cell (SL_INNER with one neighbour per side, no transverse
// projection): the neighbours' velocity cells that overlap i (cell i itself on the same grid);
// the same operations in the same order as slope_dir.
// SAME: both neighbours on the cell's velocity grid (or a shared grid): velocity cell i.
template <int D, int NDF, bool UNI, class FT, class T, bool SAME>
__device__ inline void slope_regular(const DevArgs& A, const Grid<D, NDF, UNI, FT>& G, int c,
                                     int d, int i, int li, const uint32_t* kp,
                                     const T (&f)[NDF], T (&s)[NDF]) {
  const SlopeDesc sd = A.sl[c * D + d];
  T a[2][NDF];
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const SlopeNb nb = A.slnb[sd.nb_beg[side]];
#pragma unroll
    for (int k = 0; k < NDF; ++k) a[side][k] = T(0);
    if constexpr (SAME) {
      const int64_t no = UNI ? int64_t(nb.cell) * G.n : A.voff[nb.cell];
      const FT* nf = reinterpret_cast<const FT*>(A.df) + no * NDF;
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        const T term = f[k] - T(double(Grid<D, NDF, UNI, FT>::at(nf, k * G.n + i)));
        a[side][k] += div_ds(term, sd.ds[side], sd.rds[side], sd.rds_f[side]);
      }
    } else {
      const Grid<D, NDF, UNI, FT> N(A, nb.cell);
      for_overlap_m<D, NDF, UNI, FT>(A, N, nb.map, i, li, kp, [&](int j, double den, double rden) {
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          T term = f[k] - T(N.f(j, k));
          if (den != 1.0) term = term * T(rden);
          a[side][k] += div_ds(term, sd.ds[side], sd.rds[side], sd.rds_f[side]);
        }
      });
    }
  }
  const T ic0 = desc_val<T>(sd.inv_cnt[0], sd.inv_cnt_f[0]);
  const T ic1 = desc_val<T>(sd.inv_cnt[1], sd.inv_cnt_f[1]);
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    s[k] = minmod_sel(a[0][k] * ic0, a[1][k] * ic1);
    s[k] = T(G.stored(s[k]));
    G.set_s(i, k, d, s[k]);
  }
}

// half_terms of a regular cell: every half flux it owns is its own limited or plain-slope
// upwind reconstruction.
template <int D, int NDF, bool UNI, class FT, class T>
__device__ inline void half_regular(const DevArgs& A, const Grid<D, NDF, UNI, FT>& G,
                                    const HalfDesc& hd, int i, const T (&u)[D],
                                    const T (&f)[NDF], const T (&s)[D][NDF], T wt,
                                    double (&m)[D + 2]) {
#pragma unroll
  for (int q = 0; q < D + 2; ++q) m[q] = 0.0;
  const T ud = pick<D>(u, hd.dir);
  const T ru = desc_val<T>(hd.rot, hd.rot_f) * ud;
  const bool in = hd.sel ? (ru > T(0)) : (ru <= T(0));
  if (!in) return;
  T mic[NDF];
  micro_recon<D, NDF>(hd.mode, f, s, u, T(A.dt), hd.xrec, hd.dsrec, hd.fmid, hd.dx_f, hd.ds_f,
                      hd.dir, mic);
#pragma unroll
  for (int k = 0; k < NDF; ++k) G.set_of(i, k, hd.dir, mic[k]);
  moment_terms<D, NDF, T>(wt, u, mic, m);
}

template <int D, int NDF, bool UNI, class FT, int MAXH, int NWV>
__global__ __launch_bounds__(NWV * 64) void k_slope(DevArgs A0, const int32_t* __restrict__ cells,
                                                    int ncells, KAH_RO_PARAMS) {
  constexpr int NW = D + 2;
  constexpr int BS = NWV * 64;
  constexpr int CNT = MAXH * NW;
  DevArgs A = A0;
  KAH_RO_BIND(A);
  const int c = cells[xcd_remap(blockIdx.x, ncells)];
  const Grid<D, NDF, UNI, FT> G(A, c);
  const int hb = A.half_off[c];
  const int nh = A.half_off[c + 1] - hb;
  __shared__ double red[2 * NWV * red_stride(CNT)];
  __shared__ double tot[CNT];
  __shared__ double rho_sh[MAXH];

  // FP64 arithmetic in both forms: the wall density zeroes the wall face's mass flux only
  // if its sums and the cell's own half fluxes round alike (in the FP32 form the slopes
  // computed here are still rounded to float when stored)
  using CT = double;
  double acc[CNT];
#pragma unroll
  for (int q = 0; q < CNT; ++q) acc[q] = 0.0;

  for (int i = threadIdx.x; i < G.n; i += BS) {
    CT u[D], f[NDF], s[D][NDF];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = CT(G.u(i, d));
#pragma unroll
    for (int k = 0; k < NDF; ++k) f[k] = CT(G.f(i, k));
    const int li = G.lev(i);
    const uint32_t* kp = G.keyp(i);
#pragma unroll
    for (int d = 0; d < D; ++d) slope_dir<D, NDF, UNI, FT, CT>(A, G, c, d, i, li, kp, f, s[d]);
    const CT wt = CT(G.wt(i));
#pragma unroll
    for (int h = 0; h < MAXH; ++h) {
      double m[NW];
      if (h < nh) {
        half_terms<D, NDF, UNI, FT, CT>(A, G, c, A.halves[hb + h], i, u, f, s, wt, m);
      } else {
#pragma unroll
        for (int q = 0; q < NW; ++q) m[q] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < NW; ++q) acc[h * NW + q] += m[q];
    }
  }

  block_reduce<CNT, NWV>(acc, nh * NW, red, tot);
  bool any_wall = false;
  for (int h = 0; h < nh; ++h) any_wall |= (A.halves[hb + h].mode == MI_WALL);
  if (int(threadIdx.x) < nh * NW) {
    const int h = threadIdx.x / NW, q = threadIdx.x % NW;
    const HalfDesc& hd = A.halves[hb + h];
    if (hd.mode != MI_WALL) A.half_fw[A.part_off[hd.slot] * NW + q] = tot[threadIdx.x];
  }
  if (!any_wall) return;
  // wall density rho_w = -SF/SG, then the moments of the wall Maxwellian (second sweep)
  if (int(threadIdx.x) < nh) {
    const HalfDesc& hd = A.halves[hb + threadIdx.x];
    if (hd.mode == MI_WALL) {
      const double* bc = &A.bc_prim[hd.bc * NW];
      const double pref = (D == 2) ? bc[3] / PI : pow(bc[4] / PI, 1.5);
      const double rho = -tot[threadIdx.x * NW + 0] / (pref * tot[threadIdx.x * NW + 1]);
      rho_sh[threadIdx.x] = rho;
      A.rho_w[hd.slot >> 1] = rho;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < CNT; ++q) acc[q] = 0.0;
  for (int i = threadIdx.x; i < G.n; i += BS) {
    CT u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = CT(G.u(i, d));
    const CT wt = CT(G.wt(i));
#pragma unroll
    for (int h = 0; h < MAXH; ++h) {
      if (h < nh) {
        const HalfDesc& hd = A.halves[hb + h];
        if (hd.mode != MI_WALL) continue;
        if (!(CT(hd.rot) * pick<D>(u, hd.dir) > CT(0))) continue;
        CT mic[NDF];
        double m[NW];
        micro_maxwell<D, NDF>(&A.bc_prim[hd.bc * NW], rho_sh[h], true, u, A.K, hd.dir, mic);
        moment_terms<D, NDF, CT>(wt, u, mic, m);
#pragma unroll
        for (int q = 0; q < NW; ++q) acc[h * NW + q] += m[q];
      }
    }
  }
  block_reduce<CNT, NWV>(acc, nh * NW, red, tot);
  if (int(threadIdx.x) < nh * NW) {
    const int h = threadIdx.x / NW, q = threadIdx.x % NW;
    const HalfDesc& hd = A.halves[hb + h];
    if (hd.mode == MI_WALL) A.half_fw[A.part_off[hd.slot] * NW + q] = tot[threadIdx.x];
  }
}

// VPT velocity cells per thread, one (cell, chunk) work item per workgroup: the same
// per-cell work as k_slope without the loop over the grid, for cells without wall faces.
// Each workgroup writes its partial sums of the cell's half fluxes to row
// part_off[slot]+chunk; k_flux_relax adds the rows in chunk order.
template <int D, int NDF, bool UNI, class FT, int MAXH, int VPT, int REG = 0, int WPI = 4>
__global__ __launch_bounds__(WPI * 64) void k_slope_chunk(DevArgs A0,
                                                          const int2* __restrict__ work,
                                                          int nwork, KAH_RO_PARAMS) {
  // WPI waves (3 or 4) per workgroup, 64*WPI*VPT velocity cells per chunk; the host picks
  // WPI per cell so that the chunks pad its grid least (a 568-cell grid fills one 3-wave
  // chunk of 576 where a 4-wave chunk would hold 768).
  static_assert(WPI == 3 || WPI == 4, "three or four waves per work item");
  constexpr int NW = D + 2;
  constexpr int CNT = MAXH * NW;
  constexpr int CH = 64 * WPI;   // threads per item
  DevArgs A = A0;
  KAH_RO_BIND(A);
  const int2 wk = work[xcd_remap(blockIdx.x, nwork)];
  const int c = wk.x, ch = wk.y;
  const int tid = threadIdx.x;
  const Grid<D, NDF, UNI, FT> G(A, c);
  const int hb = A.half_off[c];
  const int nh = A.half_off[c + 1] - hb;
  if constexpr (MAXH > 8) {
    // Cells with many half fluxes (hanging faces): the velocity cells' slopes stay in
    // registers and the half fluxes are summed one at a time (NW values per thread instead
    // of MAXH*NW), each with a per-wave reduce-scatter; one barrier, then the waves in order.
    using CT = typename Arith<FT>::type;
    constexpr int P = pow2_at_least(NW);
    constexpr int SH = (P == 4) ? 4 : 3;   // lanes per value after the reduce-scatter
    __shared__ double redh[2 * WPI * MAXH * P];
    CT u[VPT][D], f[VPT][NDF], s[VPT][D][NDF], wt[VPT];
    int iv[VPT];
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      const int p = (ch * VPT + r) * CH + tid;
      iv[r] = -1;
      if (p < G.n) {
        const int i = G.vi(p);
        iv[r] = i;
#pragma unroll
        for (int d = 0; d < D; ++d) u[r][d] = CT(G.u(i, d));
#pragma unroll
        for (int k = 0; k < NDF; ++k) f[r][k] = CT(G.f(i, k));
        const int li = G.lev(i);
        const uint32_t* kp = G.keyp(i);
#pragma unroll
        for (int d = 0; d < D; ++d)
          slope_dir<D, NDF, UNI, FT, CT>(A, G, c, d, i, li, kp, f[r], s[r][d]);
        wt[r] = CT(G.wt(i));
      }
    }
    const int lane = tid & 63, wv = tid >> 6;
    for (int h = 0; h < nh; ++h) {
      const HalfDesc hd = A.halves[hb + h];
      double v[P], e[P];
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = e[q] = 0.0;
#pragma unroll
      for (int r = 0; r < VPT; ++r) {
        if (iv[r] >= 0) {
          double m[NW];
          half_terms<D, NDF, UNI, FT, CT>(A, G, c, hd, iv[r], u[r], f[r], s[r], wt[r], m);
#pragma unroll
          for (int q = 0; q < NW; ++q) v[q] = (r == 0) ? m[q] : v[q] + m[q];
        }
      }
      wave_scatter2<P>(v, e);
      if ((lane & ((1 << SH) - 1)) == 0) {
        const int q = lane >> SH;
        redh[((wv * MAXH + h) * P + q) * 2] = v[0];
        redh[((wv * MAXH + h) * P + q) * 2 + 1] = e[0];
      }
    }
    __syncthreads();
    for (int t = tid; t < nh * NW; t += CH) {
      const int h = t / NW, q = t % NW;
      double S = redh[(h * P + q) * 2], E = redh[(h * P + q) * 2 + 1];
      for (int w = 1; w < WPI; ++w) {
        double err;
        S = two_sum(S, redh[((w * MAXH + h) * P + q) * 2], err);
        E = (E + redh[((w * MAXH + h) * P + q) * 2 + 1]) + err;
      }
      A.half_fw[(A.part_off[A.halves[hb + h].slot] + ch) * NW + q] = S + E;
    }
    return;
  }
  __shared__ double red[2 * WPI * red_stride(CNT)];

  // Per-thread partials (VPT <= 4 terms each) in FP64; the sum over the item's threads is
  // compensated (two_sum butterflies, then the waves in order).
  double acc[CNT];
#pragma unroll
  for (int q = 0; q < CNT; ++q) acc[q] = 0.0;
  // velocity cells (ch*VPT + r)*CH + tid, r < VPT
#pragma unroll
  for (int r = 0; r < VPT; ++r) {
    const int p = (ch * VPT + r) * CH + tid;
    if (p < G.n) {
      using CT = typename Arith<FT>::type;
      const int i = G.vi(p);
      CT u[D], f[NDF], s[D][NDF];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = CT(G.u(i, d));
#pragma unroll
      for (int k = 0; k < NDF; ++k) f[k] = CT(G.f(i, k));
      // REG 1: regular cell, neighbours on its grid; 2: regular cell, neighbours on any grid
      const int li = REG == 1 ? 0 : G.lev(i);
      const uint32_t* kp = REG == 1 ? nullptr : G.keyp(i);
      if constexpr (REG != 0) {
#pragma unroll
        for (int d = 0; d < D; ++d)
          slope_regular<D, NDF, UNI, FT, CT, REG == 1>(A, G, c, d, i, li, kp, f, s[d]);
      } else {
#pragma unroll
        for (int d = 0; d < D; ++d) slope_dir<D, NDF, UNI, FT, CT>(A, G, c, d, i, li, kp, f, s[d]);
      }
      const CT wt = CT(G.wt(i));
#pragma unroll
      for (int h = 0; h < MAXH; ++h) {
        if (h < nh) {
          double m[NW];
          if constexpr (REG) half_regular<D, NDF, UNI, FT, CT>(A, G, A.halves[hb + h], i, u, f, s, wt, m);
          else half_terms<D, NDF, UNI, FT, CT>(A, G, c, A.halves[hb + h], i, u, f, s, wt, m);
#pragma unroll
          for (int q = 0; q < NW; ++q) acc[h * NW + q] = (r == 0) ? m[q] : acc[h * NW + q] + m[q];
        }
      }
    }
  }
  const double tot = block_totals<CNT, WPI>(acc, nh * NW, red);
  if (tid < nh * NW) {
    const int h = tid / NW, q = tid % NW;
    A.half_fw[(A.part_off[A.halves[hb + h].slot] + ch) * NW + q] = tot;
  }
}

// ------------------------------------------------------------------ K2: flux + relaxation
// Micro flux into velocity cell i of cell c through all its faces, in face-list order
// (update_flux!/update_domain_flux!/update_micro_flux!, src/Flux/Flux.jl:64-281).
template <int D, int NDF, bool UNI, class FT, class T>
__device__ inline void gather_flux(const DevArgs& A, int c, int n_own, int eb, int ne, int i,
                                   const T (&u)[D], const T (&f)[NDF], const T (&s)[D][NDF],
                                   int li, const uint32_t* kp, T (&fl)[NDF]) {
  const T dt = T(A.dt);
  for (int e = 0; e < ne; ++e) {
    const EntryDesc& ed = A.ents[eb + e];
    const T ud = pick<D>(u, ed.dir);
    const T ru = desc_val<T>(ed.rot, ed.rot_f) * ud;
    const bool own = ed.side ? (ru > T(0)) : (ru <= T(0));
    const T area = desc_val<T>(ed.area, ed.area_f);
    T mic[NDF];
    if (own) {
      micro_recon<D, NDF>(ed.own_mode, f, s, u, dt, ed.xown, ed.dsown, ed.fmid, ed.dxo_f,
                          ed.dso_f, ed.dir, mic);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        T t = area * mic[k];
        fl[k] += ed.side ? -t : t;
      }
      continue;
    }
    switch (ed.oth_mode) {
      case MI_MAXWELL:
      case MI_WALL: {
        micro_maxwell<D, NDF>(&A.bc_prim[ed.bc * (D + 2)], A.rho_w_r[ed.pair],
                              ed.oth_mode == MI_WALL, u, A.K, ed.dir, mic);
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] += area * mic[k];
        break;
      }
      case MI_SELF_PLAIN: {
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] += area * (f[k] * ud);
        break;
      }
      case MI_INTERP: {
        micro_interp<D, NDF>(f, s, u, dt, ed.xown, ed.fmid, ed.dir, mic);
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] += area * mic[k];
        break;
      }
      default: {
        const Grid<D, NDF, UNI, FT> N(A, ed.oth_cell);
        const bool same = UNI || ed.map == MAP_SAME;   // j == i: the own velocities
        for_overlap_m<D, NDF, UNI, FT>(A, N, ed.map, i, li, kp,
                                       [&](int j, double den, double rden) {
          T uj[D], fj[NDF], sj[D][NDF];
#pragma unroll
          for (int t = 0; t < D; ++t) uj[t] = same ? u[t] : T(N.u(j, t));
#pragma unroll
          for (int k = 0; k < NDF; ++k) {
            fj[k] = T(N.f(j, k));
#pragma unroll
            for (int t = 0; t < D; ++t)
              sj[t][k] = (ed.oth_mode == MI_PLAIN) ? T(0) : T(N.s(j, k, t));
          }
          T mj[NDF];
          micro_recon<D, NDF>(ed.oth_mode, fj, sj, uj, dt, ed.xoth, ed.dsoth, ed.fmid, ed.dxt_f,
                              ed.dst_f, ed.dir, mj);
#pragma unroll
          for (int k = 0; k < NDF; ++k) {
            T t = area * mj[k];
            if (den != 1.0) t = t * T(rden);   // exact: den is a power of two
            fl[k] += ed.side ? -t : t;
          }
        });
      }
    }
  }
}

// gather_flux for a regular cell: its 2*DIM entries are faces to fluid cells on the same
// level (one face per side on both), reconstructed by the limited or plain-slope forms, whose
// micro flux through the face the slope kernel stored as the outgoing micro flux of the
// upwind cell per direction (its own reconstruction with the same face parameters and
// arithmetic as gather_flux's). Velocity cell i reads the cell's own outgoing flux, or the
// neighbour's at the velocity cells that overlap i (cell i itself on the same grid; across
// grids through the overlap map, with the 2^(-DIM*dl) factors): the same terms in the same
// order as gather_flux, so the same result, without slopes, neighbour velocities or
// reconstruction.
// SAME: every neighbour on the cell's velocity grid (or a shared grid): velocity cell i.
template <int D, int NDF, bool UNI, class FT, class T, bool SAME>
__device__ inline void gather_regular(const DevArgs& A, const Grid<D, NDF, UNI, FT>& G, int eb,
                                      int i, int li, const uint32_t* kp, const T (&u)[D],
                                      T (&fl)[NDF]) {
  const FT* of = reinterpret_cast<const FT*>(A.oflux);
#pragma unroll
  for (int e = 0; e < 2 * D; ++e) {
    const EntryDesc& ed = A.ents[eb + e];
    const T ud = pick<D>(u, ed.dir);
    const T ru = desc_val<T>(ed.rot, ed.rot_f) * ud;
    const bool own = ed.side ? (ru > T(0)) : (ru <= T(0));
    const T area = desc_val<T>(ed.area, ed.area_f);
    if (own) {
      const FT* b = of + G.off * (NDF * D);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        const T t = area * T(Grid<D, NDF, UNI, FT>::at(b, (ed.dir * NDF + k) * G.n + i));
        fl[k] += ed.side ? -t : t;
      }
    } else if constexpr (SAME) {
      const int64_t no = UNI ? int64_t(ed.oth_cell) * G.n : A.voff[ed.oth_cell];
      const FT* b = of + no * (NDF * D);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        const T t = area * T(Grid<D, NDF, UNI, FT>::at(b, (ed.dir * NDF + k) * G.n + i));
        fl[k] += ed.side ? -t : t;
      }
    } else {
      const Grid<D, NDF, UNI, FT> N(A, ed.oth_cell);
      const FT* b = of + N.off * (NDF * D);
      for_overlap_m<D, NDF, UNI, FT>(A, N, ed.map, i, li, kp,
                                     [&](int j, double den, double rden) {
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          T t = area * T(Grid<D, NDF, UNI, FT>::at(b, (ed.dir * NDF + k) * N.n + j));
          if (den != 1.0) t = t * T(rden);   // exact: den is a power of two
          fl[k] += ed.side ? -t : t;
        }
      });
    }
  }
}

// heat_flux_2D2F / heat_flux_3D1F partial sums of one velocity cell (lib/KitCore/2D2F.jl:68-88,
// 3D1F.jl:41-72): the sums are kept apart as in the reference and combined after reduction.
// In FP64 throughout (q is a small sum of large terms of both signs): the velocities are the
// grid's FP64 ones, f is widened from the arithmetic type.
template <int D, int NDF, int NQ, class T>
__device__ inline void heat_terms(double wt, const double (&u)[D], const T (&fv)[NDF],
                                  const double* p, double (&qa)[NQ]) {
  double f[NDF];
#pragma unroll
  for (int k = 0; k < NDF; ++k) f[k] = fv[k];
  if constexpr (D == 2) {
    double cu = u[0] - p[1], cv = u[1] - p[2];
    double c2 = cu * cu + cv * cv;
    qa[0] += wt * cu * c2 * f[0];
    qa[1] += wt * cu * f[1];
    qa[2] += wt * cv * c2 * f[0];
    qa[3] += wt * cv * f[1];
  } else {
    double cu = u[0] - p[1], cv = u[1] - p[2], cw = u[2] - p[3];
    double c2 = cu * cu + cv * cv + cw * cw;
    qa[0] += wt * cu * c2 * f[0];
    qa[1] += wt * cv * c2 * f[0];
    qa[2] += wt * cw * c2 * f[0];
  }
}
template <int D>
__device__ inline void heat_finish(const double* qs, double* q) {
  if constexpr (D == 2) {
    q[0] = 0.5 * (qs[0] + qs[1]);
    q[1] = 0.5 * (qs[2] + qs[3]);
  } else {
    q[0] = 0.5 * qs[0];
    q[1] = 0.5 * qs[1];
    q[2] = 0.5 * qs[2];
  }
}

// Block minimum of two values per thread (fixed shuffle tree, then the wave partials).
template <int NWV>
__device__ inline void block_min2(double a, double b, double* red, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double x = __shfl_xor(a, o, 64), y = __shfl_xor(b, o, 64);
    a = x < a ? x : a;
    b = y < b ? y : b;
  }
  if (lane == 0) {
    red[2 * wv] = a;
    red[2 * wv + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double m = red[threadIdx.x];
    for (int w = 1; w < NWV; ++w) m = red[2 * w + threadIdx.x] < m ? red[2 * w + threadIdx.x] : m;
    out[threadIdx.x] = m;
  }
  __syncthreads();
}

// The iterative part of solve_I_projection (src/Theory/I-projection.jl:74-141) on h in
// sc[0..n): the shave of negative entries, then Newton on the dual with Armijo backtracking.
// st holds W[0..5) and lam[5..10) (in: W, lam = 0; out: lam), dl[10..15) and the scalars set
// up by thread 0: 16 tolerance, 17 previous |G|, 18 stall count, 19 done, 24 f_min,
// 25 fp_min, 26 d, 27 1/d, 28 shave; 20-23 and 40.. are work space. sc[NDF*n..) receives the
// c_k of the last Newton step.
template <int D, int NDF, bool UNI, class FT, int NWV>
__device__ inline void cip_shave_newton(const Grid<D, NDF, UNI, FT>& G, int n, double* sc,
                                        double* red, double* st) {
  constexpr int M = D + 2, NJ = M * (M + 1) / 2;
  constexpr int BS = NWV * 64;
  double* W = st;
  double* lam = st + 5;
  double* dl = st + 10;
  // shave negative entries of h: (h - f_min)/d*fp_min
  if (st[28] != 0.0) {
    const double f_min = st[24], fp_min = st[25], dd = st[26], rd = st[27];
    for (int p = threadIdx.x; p < n; p += BS) {
      const int i = G.vi(p);
      const double x = sc[i];
      if (x < 0) sc[i] = div_by(x - f_min, dd, rd) * fp_min;
    }
    __syncthreads();
  }
  double* ck = sc + NDF * n;   // c_k = weight*h*exp(lam.psi) of the current Newton step
  for (int it = 0; it < 10; ++it) {
    double lm[M];
#pragma unroll
    for (int m = 0; m < M; ++m) lm[m] = lam[m];
    double acc[NJ + M];
#pragma unroll
    for (int m = 0; m < NJ + M; ++m) acc[m] = 0.0;
    for (int p = threadIdx.x; p < n; p += BS) {
      const int i = G.vi(p);
      double u[D], ps[M];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
      cip_psi<D>(u, ps);
      double dot = 0.0;
#pragma unroll
      for (int m = 0; m < M; ++m) dot += lm[m] * ps[m];
      const double cc = G.wt(i) * sc[i] * exp(dot);
      ck[i] = cc;
      int t = M;
#pragma unroll
      for (int a = 0; a < M; ++a) {
        const double ci = cc * ps[a];
        acc[a] += ci;
#pragma unroll
        for (int b = a; b < M; ++b) acc[t++] += ci * ps[b];
      }
    }
    block_reduce<NJ + M, NWV>(acc, NJ + M, red, &st[40]);
    if (threadIdx.x == 0) {
      double Gv[5], J[5][5], gn = 0.0;
      for (int a = 0; a < M; ++a) {
        Gv[a] = st[40 + a] - W[a];
        gn += Gv[a] * Gv[a];
      }
      gn = sqrt(gn);
      bool done = gn < st[16];
      if (!done) {
        if (gn > 0.9 * st[17]) {
          st[18] += 1.0;
          done = st[18] >= 2.0;
        } else {
          st[18] = 0.0;
        }
      }
      st[19] = done ? 1.0 : 0.0;
      if (!done) {
        st[17] = gn;
        int t = 40 + M;
        for (int a = 0; a < M; ++a)
          for (int b = a; b < M; ++b) J[a][b] = J[b][a] = st[t++];
        double x[5];
        cholesky_solve(M, J, Gv, x);
        double slope = 0.0, dW = 0.0;
        for (int a = 0; a < M; ++a) {
          dl[a] = -x[a];
          slope += Gv[a] * dl[a];
          dW += dl[a] * W[a];
        }
        st[20] = slope;
        st[21] = dW;
        st[22] = 1.0;
      }
    }
    __syncthreads();
    if (st[19] != 0.0) break;
    double dd[M];
#pragma unroll
    for (int m = 0; m < M; ++m) dd[m] = dl[m];
    // Armijo: sum c_k*expm1(alpha*dl.psi_k) - alpha*dl.W <= 1e-4*alpha*slope
    for (int tr = 0; tr < 10; ++tr) {
      const double alpha = st[22];
      double sv[1] = {0.0};
      for (int p = threadIdx.x; p < n; p += BS) {
        const int i = G.vi(p);
        double u[D], ps[M];
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
        cip_psi<D>(u, ps);
        double e = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) e += dd[m] * ps[m];
        sv[0] += ck[i] * expm1(alpha * e);
      }
      block_reduce<1, NWV>(sv, 1, red, &st[40]);
      if (threadIdx.x == 0) {
        const bool ok = st[40] - alpha * st[21] <= 1e-4 * alpha * st[20];
        st[23] = ok ? 1.0 : 0.0;
        if (!ok) st[22] = alpha * 0.5;
      }
      __syncthreads();
      if (st[23] != 0.0) break;
    }
    if (threadIdx.x == 0)
      for (int m = 0; m < M; ++m) lam[m] += st[22] * dl[m];
    __syncthreads();
  }
}

// The CIP part of iterate!(CIP_Marching) (src/Theory/I-projection.jl:161-192) after the
// convection pass, which left f* in sc and the partial sums of the heat flux (w.r.t. prim_c),
// of e_int and the minima of h: g+ = M(prim_c) + Shakhov(q), the conserved I-projection of h
// (conserved_I_porjection!, :144-158, with solve_I_projection, :55-141: shave, Newton on the
// dual, Armijo backtracking evaluated without cancellation), then f = f*tau/(tau+dt) +
// dt/(tau+dt)*g+, written to the next f. The same algorithm as the oracle's cip_projection;
// thread 0 takes the per-cell decisions, all threads loop uniformly on them.
// positivity_preserving_ib! (src/Boundary/Positivity.jl:1-43), the CIP_Marching step of a
// donor cell: the micro flux through IB face ed of the SolidNeighbor's reconstruction
// perturbance, ssdf*(x_s - aux)[dir] + ndx.ssdf e_dir with ndx = x_face - u dt - x_s, times
// u[dir] and the signed face area, on rot*u[dir] > 0 (zero elsewhere).
template <int D, int NDF, bool UNI, class FT>
__device__ inline bool ib_micro(const DevArgs& A, const EntryDesc& ed, int i, const double (&u)[D],
                                double dt, double (&mi)[NDF]) {
  if (!(ed.rot * pick<D>(u, ed.dir) > 0.0)) return false;
  const IbFaceDesc& F = A.ib_faces[ed.oth_cell - A.ib_base];
  const Grid<D, NDF, UNI, FT> PS(A, ed.oth_cell);
  const int n = PS.n;
  double ndx[D];
#pragma unroll
  for (int t = 0; t < D; ++t) ndx[t] = ed.fmid[t] - u[t] * dt - ed.xoth[t];
  const double xsd = ed.xoth[ed.dir];   // a descriptor in memory: a dynamic index is a load
  const double ud = pick<D>(u, ed.dir);
#pragma unroll
  for (int k = 0; k < NDF; ++k) {
    const double sl = A.ib_ssdf[F.row * NDF + int64_t(k) * n + i];
    const double pert = sl * (xsd - F.apd);
    double dot = ndx[0] * (ed.dir == 0 ? sl : 0.0);
#pragma unroll
    for (int t = 1; t < D; ++t) dot += ndx[t] * (ed.dir == t ? sl : 0.0);
    mi[k] = (pert + dot) * ud * ed.area;
  }
  return true;
}

// The limiter of positivity_preserving_ib!: we (the moments of the IB micro fluxes times
// dt/V) is added to w with the largest theta <= 1 that keeps density and internal energy
// positive (margin delta = 1e-3). Updates w in place, returns theta.
template <int D>
__device__ inline double ib_theta(double* w, double (&we)[D + 2]) {
  constexpr int NW = D + 2;
  const double delta = 1e-3;
  const double th_rho = we[0] > 0.0 ? 1.0
                                    : jl_min(1.0, (1.0 - delta) * w[0] / (fabs(we[0]) + MACH_EPS));
  double sb = w[1] * w[1], se = we[1] * we[1], sd = w[1] * we[1];
#pragma unroll
  for (int t = 2; t <= D; ++t) {
    sb += w[t] * w[t];
    se += we[t] * we[t];
    sd += w[t] * we[t];
  }
  const double eb = w[NW - 1] - sb / (2.0 * w[0]);
  const double ee = we[NW - 1] - sd / w[0];
  const double gam = se / (2.0 * w[0]);
  const double th_e = jl_min(1.0, 2.0 * (1.0 - delta) * eb /
                                      (sqrt(ee * ee + 4.0 * gam * (1.0 - delta) * eb) - ee + MACH_EPS));
  const double th = jl_min(th_rho, th_e);
#pragma unroll
  for (int q = 0; q < NW; ++q) w[q] += th * we[q];
  return th;
}

template <int D, int NDF, bool UNI, class FT, int NWV>
__device__ inline void relax_cip(const DevArgs& A, const Grid<D, NDF, UNI, FT>& G, int n,
                                 double* sc, double* red, double* cell_sh,
                                 const double (&qa)[(D == 2) ? 4 : 3], double eint, double fmn,
                                 double fmp, double (&q)[D], double (&pc)[D + 2]) {
  constexpr int NW = D + 2, M = D + 2, NQ = (D == 2) ? 4 : 3, NJ = M * (M + 1) / 2;
  constexpr int BS = NWV * 64;
  __shared__ double st[64];   // W[5] lam[5] dl[5] | scalars from 16 on
  double* W = st;
  double* lam = st + 5;
  double* dl = st + 10;
  // scalars: 16 tol_eff, 17 g_prev, 18 stall, 19 done, 20 slope, 21 dW, 22 alpha, 23 accept,
  // 24 f_min, 25 fp_min, 26 d, 27 1/d, 28 shave, 29.. q, 40.. sums
  {
    double v[NQ + 1];
#pragma unroll
    for (int m = 0; m < NQ; ++m) v[m] = qa[m];
    v[NQ] = eint;
    block_reduce<NQ + 1, NWV>(v, NQ + 1, red, &st[40]);
    block_min2<NWV>(fmn, fmp, red, &st[46]);
  }
  if (threadIdx.x == 0) {
    double qs[NQ], qq[D];
#pragma unroll
    for (int m = 0; m < NQ; ++m) qs[m] = st[40 + m];
    heat_finish<D>(qs, qq);
#pragma unroll
    for (int d = 0; d < D; ++d) st[29 + d] = qq[d];
    double wn = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      W[m] = cell_sh[m];
      lam[m] = 0.0;
    }
    if constexpr (NDF == 2) W[M - 1] -= st[40 + NQ] / 2;   // e_int = dot(weight, b)/2
#pragma unroll
    for (int m = 0; m < M; ++m) wn += W[m] * W[m];
    st[16] = 1e-10 * fmax(1.0, sqrt(wn));
    st[17] = HUGE_VAL;
    st[18] = 0.0;
    st[19] = 0.0;
    const double f_min = 1.1 * st[46];
    st[24] = f_min;
    st[25] = st[47];
    st[26] = st[47] - f_min;
    st[27] = 1.0 / st[26];
    st[28] = f_min < 0 ? 1.0 : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = st[29 + d];
#pragma unroll
  for (int m = 0; m < NW; ++m) pc[m] = cell_sh[NW + m];
  cip_shave_newton<D, NDF, UNI, FT, NWV>(G, n, sc, red, st);
  // h *= exp(lam.psi); g+ = M(prim_c) + Shakhov; f = f*tau/(tau+dt) + dt/(tau+dt)*g+
  double lm[M];
#pragma unroll
  for (int m = 0; m < M; ++m) lm[m] = lam[m];
  const double dt = A.dt;
  const double tau = get_tau<D>(pc, A.mu, A.omega);
  const double a1 = tau / (tau + dt), a2 = dt / (tau + dt);
  const MaxPre<D> mpc = max_pre<D>(pc, A.Pr);
  for (int p = threadIdx.x; p < n; p += BS) {
    const int i = G.vi(p);
    double u[D], ps[M], Fc[2], Sp[2];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
    cip_psi<D>(u, ps);
    double dot = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) dot += lm[m] * ps[m];
    maxwell_pre<D>(u, mpc, A.K, Fc);
    shakhov_pre<D>(u, Fc, mpc, q, A.K, Sp);
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      double fk = sc[k * n + i];
      if (k == 0) fk = fk * exp(dot);
      const double g = Fc[k] + Sp[k];
      double fn = fk * a1;
      fn = fn + a2 * g;
      G.set_out(i, k, fn);
    }
  }
}

// iterate!(CAIDVM_Marching) (src/Theory/Iterate.jl:96-130) or iterate!(Euler) (:131-162)
// for one cell, fed by the gathered flux (MODE_FUSED), by the flux accumulators
// (MODE_ITER), or only accumulating the flux (MODE_FLUX).
// Half flux of entry e's side sd, component q, of a cell: the sum of its partial rows in chunk
// order, compensated (micro_to_macro's 1/2 on the energy row applied after the sum).
template <int D>
__device__ inline double half_total(const DevArgs& A, const EntryDesc& ed, int sd, int q) {
  constexpr int NW = D + 2;
  const int slot = ed.pair * 2 + sd;
  const int64_t off = A.part_off[slot];
  const int cnt = A.part_cnt[slot];
  double sum = 0.0, err = 0.0;
  for (int k = 0; k < cnt; ++k) {
    double e;
    sum = two_sum(sum, A.half_fw_r[(off + k) * NW + q], e);
    err += e;
  }
  sum = sum + err;
  return (q == NW - 1) ? 0.5 * sum : sum;
}

// Macroscopic face flux of cell c in face order (flux += fw*area / flux -= fw*area) and
// w + dt/V*flux (both marchings: src/Theory/Iterate.jl:108, :144-149); returns dt/V.
template <int D, class HF>
__device__ inline double cell_update(const DevArgs& A, int c, int eb, int ne, bool faces,
                                     HF half, double (&wn)[D + 2]) {
  constexpr int NW = D + 2;
  const double* dsc = &A.rec_ds[int64_t(c) * D];
  double V = dsc[0];
#pragma unroll
  for (int d = 1; d < D; ++d) V = V * dsc[d];
  double psf[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) psf[q] = A.psflux[int64_t(c) * NW + q];
  if (faces) {
    for (int e = 0; e < ne; ++e) {
      const EntryDesc& ed = A.ents[eb + e];
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        const double h0 = half(e, 0, q), h1 = half(e, 1, q);
        double t;
        if (ed.oth_cell < 0) t = ed.area * (h0 + h1);    // update_domain_flux!
        else t = (h0 + h1) * ed.area;                     // update_macro_flux!
        psf[q] = ed.side ? psf[q] - t : psf[q] + t;
      }
    }
  }
  const double* wold = &A.w[int64_t(c) * NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) wn[q] = wold[q] + psf[q] * A.dt / V;
  return A.dt / V;
}

// Per-cell prologue of the fused step under CAIDVM_Marching and Euler, one thread per cell:
// the half fluxes of the cell's faces, w + dt/V*flux, prim_c, the Maxwellian pre-factors and
// tau of prim_c and the relaxation weights, into the cell's record (cpre_len(D) doubles).
// k_flux_relax reads the record instead of forming it in one thread of each workgroup while
// the others wait at a barrier; the operations and their order are the same.
template <int D>
__global__ __launch_bounds__(256) void k_cell_pre(DevArgs A, const int32_t* __restrict__ cells,
                                                  int ncells) {
  constexpr int NW = D + 2;
  static_assert(sizeof(MaxPre<D>) == sizeof(double) * (NW + 4), "MaxPre layout");
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ncells) return;
  const int c = cells[t];
  const int eb = A.ent_off[c];
  const int ne = A.ent_off[c + 1] - eb;
  double wn[NW], pc[NW];
  const double dtV = cell_update<D>(
      A, c, eb, ne, A.flux_kind == 0,
      [&](int e, int sd, int q) { return half_total<D>(A, A.ents[eb + e], sd, q); }, wn);
  get_prim<D>(wn, A.gamma, pc);
  const MaxPre<D> pre = max_pre<D>(pc, A.Pr);
  const double tau = get_tau<D>(pc, A.mu, A.omega);
  double* r = A.cpre + int64_t(c) * cpre_len(D);
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    r[q] = wn[q];
    r[NW + q] = pc[q];
  }
  const double* pp = reinterpret_cast<const double*>(&pre);
#pragma unroll
  for (int k = 0; k < NW + 4; ++k) r[2 * NW + k] = pp[k];
  r[3 * NW + 4] = tau;
  r[3 * NW + 5] = tau / (tau + A.dt);
  r[3 * NW + 6] = A.dt / (tau + A.dt);
  r[3 * NW + 7] = dtV;
}

template <int D, int NDF, bool UNI, class FT, int MODE, int MARCH, bool LDS, int NWV>
__global__ __launch_bounds__(NWV * 64) void k_flux_relax(DevArgs A0,
                                                         const int32_t* __restrict__ cells,
                                                         int ncells, KAH_RO_PARAMS,
                                                         const double* __restrict__ ro_sdf,
                                                         const double* __restrict__ ro_half,
                                                         const double* __restrict__ ro_rho) {
  constexpr int NW = D + 2;
  constexpr int NQ = (D == 2) ? 4 : 3;   // heat-flux partial sums
  constexpr int BS = NWV * 64;
  DevArgs A = A0;
  KAH_RO_BIND(A);
  A.sdf_r = ro_sdf;
  A.half_fw_r = ro_half;
  A.rho_w_r = ro_rho;
  const int c = cells[xcd_remap(blockIdx.x, ncells)];
  const Grid<D, NDF, UNI, FT> G(A, c);
  const int n = G.n;
  const int eb = A.ent_off[c];
  const int ne = A.ent_off[c + 1] - eb;
  const double dt = A.dt;
  // the largest block sum: G and J of the I-projection (CIP_Marching), else D+2 values
  constexpr int NRED = red_stride((MARCH == MARCH_CIP) ? (D + 2) * (D + 5) / 2 : 8);
  __shared__ double red[2 * NWV * NRED];
  __shared__ double res[8];
  __shared__ double cell_sh[4 * NW];   // wnew | prim_c | prim | q
  __shared__ double hsum[MAX_ENTRIES * 2 * NW];   // the half fluxes of the cell's faces
  // per-cell constants of the relaxation, computed once: the Maxwellian pre-factors of
  // prim_c and of the moments of f*, tau(prim_c) and the relaxation weights
  __shared__ MaxPre<D> pre_c;
  __shared__ double pre_t[3];
  extern __shared__ double lds[];
  // per-velocity arithmetic type: FP32 in the FP32 form, except for CIP_Marching, whose
  // Newton iteration on the projected distribution stays FP64
  using CT = std::conditional_t<(sizeof(FT) == 4 && MARCH != MARCH_CIP), float, double>;
  // velocity-cell loops of the FP32 form unrolled twice (two cells' loads in flight per
  // thread; measured -2 % to -4 %); the FP64 form's registers leave no room for it (+6 %)
  constexpr int UNR = sizeof(CT) == 4 ? 2 : 1;
  CT* sc = LDS ? reinterpret_cast<CT*>(lds)
               : reinterpret_cast<CT*>(&A.scratch[G.off * (NDF + 1)]);

  // half flux = sum of its partial rows in chunk order (micro_to_macro's 1/2 on the energy
  // row applied after the sum), one thread per (entry, side, component)
  constexpr bool FUSED = (MODE == MODE_FUSED || MODE == MODE_FUSED_REG || MODE == MODE_FUSED_REGX);
  // the record of k_cell_pre (fused CAIDVM_Marching / Euler steps, which the host always
  // precedes by it): copied to LDS below
  constexpr bool pre = FUSED && MARCH != MARCH_CIP;
  if constexpr (pre) {
    const double* r = A.cpre + int64_t(c) * cpre_len(D);
    const int j = threadIdx.x;
    if (j < 2 * NW) cell_sh[j] = r[j];
    else if (j < 3 * NW + 4) reinterpret_cast<double*>(&pre_c)[j - 2 * NW] = r[j];
    else if (j < 3 * NW + 7) pre_t[j - (3 * NW + 4)] = r[j];
    else if (j == 3 * NW + 7) res[0] = r[j];
  } else {
    if (MODE != MODE_ITER && A.flux_kind == 0)
      for (int x = threadIdx.x; x < ne * 2 * NW; x += BS) {
        const int e = x / (2 * NW), s = (x / NW) % 2, q = x % NW;
        hsum[x] = half_total<D>(A, A.ents[eb + e], s, q);
      }
    __syncthreads();
  }

  if (!pre && threadIdx.x == 0) {
    const bool faces = MODE != MODE_ITER && A.flux_kind == 0;
    auto half = [&](int e, int sd, int q) { return hsum[(e * 2 + sd) * NW + q]; };
    if (MODE == MODE_FLUX) {
      // the split flux! call: the macroscopic face flux into PsData.flux
      double psf[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) psf[q] = A.psflux[int64_t(c) * NW + q];
      for (int e = 0; faces && e < ne; ++e) {
        const EntryDesc& ed = A.ents[eb + e];
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          const double h0 = half(e, 0, q), h1 = half(e, 1, q);
          double t;
          if (ed.oth_cell < 0) t = ed.area * (h0 + h1);    // update_domain_flux!
          else t = (h0 + h1) * ed.area;                     // update_macro_flux!
          psf[q] = ed.side ? psf[q] - t : psf[q] + t;
        }
      }
#pragma unroll
      for (int q = 0; q < NW; ++q) A.psflux[int64_t(c) * NW + q] = psf[q];
    } else {
      double wn[NW], pc[NW];
      res[0] = cell_update<D>(A, c, eb, ne, faces, half, wn);
      get_prim<D>(wn, A.gamma, pc);
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        cell_sh[q] = wn[q];
        cell_sh[NW + q] = pc[q];
      }
    }
  }
  if (MODE == MODE_FLUX) {
    for (int p = threadIdx.x; p < n; p += BS) {
      const int i = G.vi(p);
      CT u[D], f[NDF], s[D][NDF], fl[NDF];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = CT(G.u(i, d));
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        f[k] = CT(G.f(i, k));
#pragma unroll
        for (int d = 0; d < D; ++d) s[d][k] = CT(G.s(i, k, d));
        fl[k] = CT(A.vflux[G.off * NDF + int64_t(k) * n + i]);
      }
      gather_flux<D, NDF, UNI, FT, CT>(A, c, n, eb, ne, i, u, f, s, G.lev(i), G.keyp(i), fl);
#pragma unroll
      for (int k = 0; k < NDF; ++k) A.vflux[G.off * NDF + int64_t(k) * n + i] = fl[k];
    }
    return;
  }
  __syncthreads();
  const double dtV = uniform(res[0]);
  // CIP_Marching next to an immersed boundary: positivity_preserving_ib! on w before prim_c
  bool ib_cell = false;
  double ib_th = 0.0;
  if constexpr (MARCH == MARCH_CIP) {
    if (A.ib_faces) {
      for (int e = 0; e < ne; ++e) ib_cell |= A.ents[eb + e].oth_cell >= A.ib_base;
      if (ib_cell) {
        double mw[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) mw[q] = 0.0;
        for (int p = threadIdx.x; p < n; p += BS) {
          const int i = G.vi(p);
          double u[D];
#pragma unroll
          for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
          for (int e = 0; e < ne; ++e) {
            const EntryDesc& ed = A.ents[eb + e];
            if (ed.oth_cell < A.ib_base) continue;
            double mi[NDF], m[NW];
            if (!ib_micro<D, NDF, UNI, FT>(A, ed, i, u, dt, mi)) continue;
            moment_terms<D, NDF, double>(G.wt(i), u, mi, m);
#pragma unroll
            for (int q = 0; q < NW; ++q) mw[q] += m[q];
          }
        }
        block_reduce<NW, NWV>(mw, NW, red, &cell_sh[2 * NW]);
        if (threadIdx.x == 0) {
          double we[NW], wb[NW], pc[NW];
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            we[q] = (q == NW - 1 ? 0.5 * cell_sh[2 * NW + q] : cell_sh[2 * NW + q]) * dtV;
            wb[q] = cell_sh[q];
          }
          res[1] = ib_theta<D>(wb, we);
          get_prim<D>(wb, A.gamma, pc);
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            cell_sh[q] = wb[q];
            cell_sh[NW + q] = pc[q];
          }
        }
        __syncthreads();
        ib_th = uniform(res[1]);
      }
    }
  }
  if (MARCH != MARCH_CIP && !pre && threadIdx.x == 0) {
    // needed after the first reduction; wave 0 computes them before its share of pass 1
    double pc[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) pc[q] = cell_sh[NW + q];
    pre_c = max_pre<D>(pc, A.Pr);
    const double tau = get_tau<D>(pc, A.mu, A.omega);
    pre_t[0] = tau;
    pre_t[1] = tau / (tau + dt);
    pre_t[2] = dt / (tau + dt);
  }

  // pass 1: flux of every velocity cell (gathered or from the accumulators)
  double mw[NW];
  double qa[NQ];
  double eint = 0.0, fmn = HUGE_VAL, fmp = HUGE_VAL;   // CIP: b's energy, min h, min h > 0
#pragma unroll
  for (int q = 0; q < NW; ++q) mw[q] = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) qa[q] = 0.0;
  {
    double pc[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) pc[q] = cell_sh[NW + q];
    const CT dtVc = CT(dtV);
#pragma unroll UNR
    for (int p = threadIdx.x; p < n; p += BS) {
      const int i = G.vi(p);
      CT u[D], f[NDF], fl[NDF];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = CT(G.u(i, d));
#pragma unroll
      for (int k = 0; k < NDF; ++k) f[k] = CT(G.f(i, k));
      if constexpr (MODE == MODE_FUSED_REG || MODE == MODE_FUSED_REGX) {
#pragma unroll
        for (int k = 0; k < NDF; ++k) fl[k] = CT(0);
        if constexpr (MODE == MODE_FUSED_REG)
          gather_regular<D, NDF, UNI, FT, CT, true>(A, G, eb, i, 0, nullptr, u, fl);
        else
          gather_regular<D, NDF, UNI, FT, CT, false>(A, G, eb, i, G.lev(i), G.keyp(i), u, fl);
      } else if (FUSED) {
        CT s[D][NDF];
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fl[k] = CT(0);
#pragma unroll
          for (int d = 0; d < D; ++d) s[d][k] = CT(G.s(i, k, d));
        }
        gather_flux<D, NDF, UNI, FT, CT>(A, c, n, eb, ne, i, u, f, s, G.lev(i), G.keyp(i), fl);
      } else {
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fl[k] = CT(A.vflux[G.off * NDF + int64_t(k) * n + i]);
          A.vflux[G.off * NDF + int64_t(k) * n + i] = 0.0;
        }
      }
      const CT wt = CT(G.wt(i));
      if (MARCH == MARCH_CAIDVM) {
        // convection first: f* = f + dt/V*flux, and the moments of f*
        CT fs[NDF];
        double m[NW];
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fs[k] = f[k] + dtVc * fl[k];
          sc[k * n + i] = fs[k];
        }
        moment_terms<D, NDF, CT>(wt, u, fs, m);
#pragma unroll
        for (int q = 0; q < NW; ++q) mw[q] += m[q];
      } else if (MARCH == MARCH_CIP) {
        // convection first, then the heat flux of f* with prim_c (calc_qf after the
        // convection, src/Theory/I-projection.jl:177-179); e_int and the minima of h for the
        // I-projection
        if (ib_cell) {   // + theta times the IB micro fluxes, face after face
          double ud[D];
#pragma unroll
          for (int d = 0; d < D; ++d) ud[d] = G.u(i, d);
          for (int e = 0; e < ne; ++e) {
            const EntryDesc& ed = A.ents[eb + e];
            if (ed.oth_cell < A.ib_base) continue;
            double mi[NDF];
            if (!ib_micro<D, NDF, UNI, FT>(A, ed, i, ud, dt, mi)) continue;
#pragma unroll
            for (int k = 0; k < NDF; ++k) fl[k] = fl[k] + CT(ib_th * mi[k]);
          }
        }
        CT fs[NDF];
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fs[k] = f[k] + dtVc * fl[k];
          sc[k * n + i] = fs[k];
        }
        double ud[D];
#pragma unroll
        for (int d = 0; d < D; ++d) ud[d] = G.u(i, d);
        heat_terms<D, NDF, NQ, CT>(G.wt(i), ud, fs, pc, qa);
        if constexpr (NDF == 2) eint += wt * fs[1];
        fmn = fs[0] < fmn ? fs[0] : fmn;
        if (fs[0] > 0.0 && fs[0] < fmp) fmp = fs[0];
      } else {
        // Euler: heat flux of f^n with the updated prim
#pragma unroll
        for (int k = 0; k < NDF; ++k) sc[k * n + i] = fl[k];
        double ud[D];
#pragma unroll
        for (int d = 0; d < D; ++d) ud[d] = G.u(i, d);
        heat_terms<D, NDF, NQ, CT>(G.wt(i), ud, f, pc, qa);
      }
    }
  }

  double q[D];
  double pc[NW];
  if constexpr (MARCH == MARCH_CIP) {
    relax_cip<D, NDF, UNI, FT, NWV>(A, G, n, sc, red, cell_sh, qa, eint, fmn, fmp, q, pc);
  } else {
  if (MARCH == MARCH_CAIDVM) {
    block_reduce<NW, NWV>(mw, NW, red, &cell_sh[2 * NW]);
    // prim of f* and its Maxwellian pre-factors, formed by every wave from the reduced
    // moments (the same operations in each; no further barrier)
    MaxPre<D> mpr;
    {
      double wf[NW], pr[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) wf[q] = uniform(cell_sh[2 * NW + q]);
      wf[NW - 1] = 0.5 * wf[NW - 1];
      get_prim<D>(wf, A.gamma, pr);
      mpr = uniform<D>(max_pre<D>(pr, A.Pr));
    }
    // pass 2: conservation correction f += M(prim_c) - M(prim); heat flux of the result
    {
      double pc[NW];
      const MaxPre<D> mpc = uniform<D>(pre_c);
#pragma unroll
      for (int q = 0; q < NW; ++q) pc[q] = mpc.p[q];
      const CT K = CT(A.K);
#pragma unroll UNR
      for (int p = threadIdx.x; p < n; p += BS) {
        const int i = G.vi(p);
        double ud[D];
        CT u[D], Mc[2], M[2];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          ud[d] = G.u(i, d);
          u[d] = CT(ud[d]);
        }
        maxwell_pre<D, CT>(ud, mpc, K, Mc);
        maxwell_pre<D, CT>(ud, mpr, K, M);
        CT fs[NDF];
#pragma unroll
        for (int k = 0; k < NDF; ++k) {
          fs[k] = sc[k * n + i] + (Mc[k] - M[k]);
          sc[k * n + i] = fs[k];
        }
        sc[NDF * n + i] = Mc[0];
        heat_terms<D, NDF, NQ, CT>(G.wt(i), ud, fs, pc, qa);
      }
    }
  }
  double qs_red[NQ];
  block_reduce<NQ, NWV>(qa, NQ, red, &cell_sh[3 * NW]);
#pragma unroll
  for (int m = 0; m < NQ; ++m) qs_red[m] = cell_sh[3 * NW + m];
  heat_finish<D>(qs_red, q);
#pragma unroll
  for (int m = 0; m < D; ++m) q[m] = uniform(q[m]);
  const MaxPre<D> mpc = uniform<D>(pre_c);
#pragma unroll
  for (int m = 0; m < NW; ++m) pc[m] = mpc.p[m];
  const double tau = uniform(pre_t[0]);

  if (MARCH == MARCH_CAIDVM) {
    // pass 3: g+ = M(prim_c) + Shakhov part; f = f*tau/(tau+dt) + dt/(tau+dt)*g+
    const double a1 = uniform(pre_t[1]), a2 = uniform(pre_t[2]);
    const CT K = CT(A.K), a1c = CT(a1), a2c = CT(a2);
    CT qc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) qc[d] = CT(q[d]);
#pragma unroll UNR
    for (int p = threadIdx.x; p < n; p += BS) {
      const int i = G.vi(p);
      CT u[D], Fc[2], Sp[2];
#pragma unroll
      for (int d = 0; d < D; ++d) u[d] = CT(G.u(i, d));
      Fc[0] = sc[NDF * n + i];
      if constexpr (NDF == 2) {
        if constexpr (sizeof(CT) == 8) Fc[1] = div_by(Fc[0] * A.K, mpc.two_l, mpc.r2l);
        else Fc[1] = Fc[0] * K * CT(mpc.r2l);
      }
      shakhov_pre<D, CT>(u, Fc, mpc, qc, K, Sp);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        CT g = Fc[k] + Sp[k];
        CT fn = sc[k * n + i] * a1c;
        fn = fn + a2c * g;
        G.set_out(i, k, fn);
      }
    }
  } else {
    // Euler: f = (f + dt/V*flux)*tau/(tau+dt) + dt/(tau+dt)*F with F = M + Shakhov(M)
    const double tpd = tau + dt, a2 = uniform(pre_t[2]);
    const CT K = CT(A.K), dtVc = CT(dtV), tauc = CT(tau), tpdc = CT(tpd), a2c = CT(a2);
    CT qc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) qc[d] = CT(q[d]);
    for (int p = threadIdx.x; p < n; p += BS) {
      const int i = G.vi(p);
      double ud[D];
      CT u[D], M[2], Sp[2];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        ud[d] = G.u(i, d);
        u[d] = CT(ud[d]);
      }
      maxwell_pre<D, CT>(ud, mpc, K, M);
      shakhov_pre<D, CT>(u, M, mpc, qc, K, Sp);
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        CT F = M[k] + Sp[k];
        CT fx = (CT(G.f(i, k)) + dtVc * sc[k * n + i]) * tauc / tpdc + a2c * F;
        G.set_out(i, k, fx);
      }
    }
  }
  }
  if (threadIdx.x == 0) {
    double* pr = &A.prim[int64_t(c) * NW];
    double* rs = &A.res[int64_t(c) * 2 * NW];
#pragma unroll
    for (int m = 0; m < NW; ++m) {
      double dl = pc[m] - pr[m];
      rs[m] = dl * dl;
      rs[NW + m] = fabs(pc[m]);
      pr[m] = pc[m];
      A.w[int64_t(c) * NW + m] = cell_sh[m];
      A.psflux[int64_t(c) * NW + m] = 0.0;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) A.qf[int64_t(c) * D + d] = q[d];
  }
}

// ------------------------------------------------------------------ mirror (ghost) buffers
// The per-cell buffers of the ghost exchange, concatenated in list order at buf + off[k]:
// DATA (kind 0) is [ds(D) midpoint(D) w(D+2) vs_num bound_enc df(vs_num*NDF)], SLOPE (kind 1)
// is [sdf(vs_num*NDF*D) sw(D*(D+2))], every array column-major as the reference reshapes it
// (src/Parallel/Ghost.jl:1-3, :42-55, :525-579). A cell inside a solid (bound_enc < 0) packs
// as [ds midpoint Inf*(D+2) 0 bound_enc] and [EPS 0 ...]. Unpack (the ghost side,
// src/Parallel/Ghost.jl:688-749) writes w and df, or sdf and sw, of each listed fluid cell.
// One workgroup per listed cell.
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(256) void k_mirror(DevArgs A, const int32_t* __restrict__ cells,
                                                const int64_t* __restrict__ off, int kind,
                                                int unpack, double* __restrict__ buf,
                                                const double* __restrict__ mid,
                                                const int32_t* __restrict__ benc,
                                                double* __restrict__ sw, FT* __restrict__ dfw) {
  constexpr int NW = D + 2;
  const int c = cells[blockIdx.x];
  const bool solid = benc[c] < 0;
  double* b = buf + off[blockIdx.x];
  if (solid) {
    if (unpack || threadIdx.x != 0) return;
    if (kind == 0) {
      for (int d = 0; d < D; ++d) {
        b[d] = A.rec_ds[int64_t(c) * D + d];
        b[D + d] = mid[int64_t(c) * D + d];
      }
      for (int q = 0; q < NW; ++q) b[2 * D + q] = __builtin_inf();
      b[3 * D + 2] = 0.0;
      b[3 * D + 3] = double(benc[c]);
    } else {
      b[0] = 1e-12;   // EPS, src/Abstract/Types.jl:3
      for (int x = 1; x < D * NW; ++x) b[x] = 0.0;
    }
    return;
  }
  const Grid<D, NDF, UNI, FT> G(A, c);
  const int n = G.n;
  if (kind == 0) {
    if (threadIdx.x == 0) {
      if (!unpack) {
        for (int d = 0; d < D; ++d) {
          b[d] = A.rec_ds[int64_t(c) * D + d];
          b[D + d] = mid[int64_t(c) * D + d];
        }
        for (int q = 0; q < NW; ++q) b[2 * D + q] = A.w[int64_t(c) * NW + q];
        b[3 * D + 2] = double(n);
        b[3 * D + 3] = double(benc[c]);
      } else {
        for (int q = 0; q < NW; ++q) A.w[int64_t(c) * NW + q] = b[2 * D + q];
      }
    }
    double* fb = b + 3 * D + 4;
    FT* fo = dfw + G.off * NDF;
    for (int x = threadIdx.x; x < NDF * n; x += 256) {
      if (!unpack) fb[x] = double(G.at(G.fb, x));
      else G.at(fo, x) = FT(fb[x]);
    }
  } else {
    for (int x = threadIdx.x; x < NDF * D * n; x += 256) {
      if (!unpack) b[x] = double(G.at(G.sb, x));
      else G.at(G.swb, x) = FT(b[x]);
    }
    double* sb = b + int64_t(NDF) * D * n;
    for (int x = threadIdx.x; x < D * NW; x += 256) {
      const int d = x / NW, q = x - d * NW;
      if (!unpack) sb[x] = sw[(int64_t(c) * NW + q) * D + d];
      else sw[(int64_t(c) * NW + q) * D + d] = sb[x];
    }
  }
}

// ------------------------------------------------------------------ restart arrays
// The per-rank phase-space arrays of a restart checkpoint (src/IO/Restart.jl:80-117, read back
// at :306-321): ws (N x (D+2)), vs_levels (Nv), vs_midpoints (Nv x D) and vs_df (Nv x NDF),
// every matrix column-major, the velocity cells of the fluid cells concatenated in cell order
// (row[c] = first row of cell c, -1 inside a solid). Save gathers them from the device state;
// load scatters ws and vs_df back and sets prim = get_prim(w) as PsData(...) does.
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(256) void k_restart(DevArgs A, const int64_t* __restrict__ row,
                                                 int64_t N, int64_t Nv, int load,
                                                 double* __restrict__ ws,
                                                 int8_t* __restrict__ lev,
                                                 double* __restrict__ vmid,
                                                 double* __restrict__ vdf, FT* __restrict__ dfw) {
  constexpr int NW = D + 2;
  const int c = blockIdx.x;
  const int64_t r0 = row[c];
  if (r0 < 0) return;
  const Grid<D, NDF, UNI, FT> G(A, c);
  const int n = G.n;
  if (threadIdx.x == 0) {
    if (!load) {
      for (int q = 0; q < NW; ++q) ws[q * N + c] = A.w[int64_t(c) * NW + q];
    } else {
      double w[NW], p[NW];
      for (int q = 0; q < NW; ++q) w[q] = ws[q * N + c];
      get_prim<D>(w, A.gamma, p);
      for (int q = 0; q < NW; ++q) {
        A.w[int64_t(c) * NW + q] = w[q];
        A.prim[int64_t(c) * NW + q] = p[q];
      }
    }
  }
  FT* fo = dfw + G.off * NDF;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int64_t r = r0 + i;
    if (!load) {
      lev[r] = A.glev[G.goff() + i];
      for (int d = 0; d < D; ++d) vmid[d * Nv + r] = G.u(i, d);
      for (int k = 0; k < NDF; ++k) vdf[k * Nv + r] = G.f(i, k);
    } else {
      for (int k = 0; k < NDF; ++k) G.at(fo, k * n + i) = FT(vdf[k * Nv + r]);
    }
  }
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_restart_t(const DevArgs& A, const int64_t* row, int64_t N, int64_t Nv,
                                   int load, double* ws, int8_t* lev, double* vmid, double* vdf,
                                   void* dfw, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_restart<D, NDF, UNI, FT>), dim3(unsigned(N)), dim3(256), 0, st, A, row,
                     N, Nv, load, ws, lev, vmid, vdf, static_cast<FT*>(dfw));
  return hipGetLastError();
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_mirror_t(const DevArgs& A, const int32_t* cells, const int64_t* off,
                                  int n, int kind, int unpack, double* buf, const double* mid,
                                  const int32_t* benc, double* sw, void* dfw, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_mirror<D, NDF, UNI, FT>), dim3(n), dim3(256), 0, st, A, cells, off, kind,
                     unpack, buf, mid, benc, sw, static_cast<FT*>(dfw));
  return hipGetLastError();
}

// ------------------------------------------------------------------ velocity-space adaptation
// vs_adaptive_mesh_refinement! (src/Velocity_space/AMR.jl:133-181) on the device: the
// resolution, the per-velocity-cell criteria (Criteria.jl), the remap of the state onto the
// rebuilt grids (Rebuild.jl) and the conservative I-projection. The face neighbours of the
// Loehner indicator come from a binary search over the cell's keys (the grid is in key
// order), as VsNeighborIndex locates them (Neighbor.jl:126-201).
template <int D>
__device__ inline uint64_t vs_code(const VsArgs& P, const int64_t (&g)[D]) {
  int64_t root = 0;
  uint64_t morton = 0;
  const int64_t mask = (int64_t(1) << P.L) - 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    root += (g[d] >> P.L) * P.stride[d];
    const int64_t q = g[d] & mask;
    for (int b = 0; b < P.L; ++b) morton |= uint64_t((q >> b) & 1) << (b * D + d);
  }
  return (uint64_t(root) << (D * P.L)) | morton;
}

template <int D, int NDF, bool UNI, class FT>
struct VsGrid {
  const Grid<D, NDF, UNI, FT>& G;
  const VsArgs& P;
  __device__ int lev(int j) const { return int(G.A.glev[G.goff() + j]); }
  __device__ double size(int j, int d) const {
    return P.hf[d] * double(int64_t(1) << (P.L - lev(j)));
  }
  __device__ uint64_t key(int j) const {   // _corner_index, Neighbor.jl:38-41
    int64_t g[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
      g[d] = int64_t(rint((G.u(j, d) - 0.5 * size(j, d) - P.vmin[d]) / P.hf[d]));
    return vs_code<D>(P, g);
  }
  __device__ int locate(uint64_t code) const {   // _vs_locate_code, Neighbor.jl:126-139
    int lo = 0, hi = G.n - 1, k = -1;
    while (lo <= hi) {
      const int m = (lo + hi) >> 1;
      if (key(m) <= code) {
        k = m;
        lo = m + 1;
      } else {
        hi = m - 1;
      }
    }
    if (k < 0) return -1;
    const uint64_t kk = key(k), span = uint64_t(1) << (D * (P.L - lev(k)));
    return (code >= kk && code < kk + span) ? k : -1;
  }
  __device__ int face(int i, int dim, int dir) const {   // vs_face_neighbor, :186-201
    const double cell = size(i, dim);
    int64_t g[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double coord = G.u(i, d);
      if (d == dim) coord += double(dir) * (0.5 * cell + 0.5 * P.hf[d]);
      const double x = floor((coord - P.vmin[d]) / P.hf[d]);
      if (x < 0 || x >= double(P.tn[d] << P.L)) return -1;
      g[d] = int64_t(x);
    }
    return locate(vs_code<D>(P, g));
  }
};

__device__ inline double jl_maxd(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return b > a ? b : a;
}

// _lohner_ratio, src/Velocity_space/Criteria.jl:221-227
__device__ inline double lohner_ratio(double L, double C, double R, double dsL, double dsR,
                                      double scale) {
  const double eflt = 1e-2, eabs = 1e-3;
  const double num = fabs(dsR * L - (dsL + dsR) * C + dsL * R);
  const double den = dsR * fabs(L - C) + dsL * fabs(R - C) +
                     eflt * (dsR * fabs(L) + (dsL + dsR) * fabs(C) + dsL * fabs(R)) +
                     eabs * scale * (dsL + dsR);
  return den > 0 ? num / den : 0.0;
}

// vs_resolution of one cell (AMR.jl:153-167): out[2c] = max(df)*w_fine,
// out[2c+1] = 0.5*max(h*c^2 (+ b))*w_fine
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(256) void k_vs_res(DevArgs A, VsArgs P, double* __restrict__ out) {
  constexpr int NW = D + 2;
  __shared__ double red[2 * 4];
  __shared__ double res[2];
  const int c = blockIdx.x;
  const Grid<D, NDF, UNI, FT> G(A, c);
  const double* U = &A.prim[int64_t(c) * NW + 1];
  double dm = -HUGE_VAL, em = -HUGE_VAL;
  for (int i = threadIdx.x; i < G.n; i += 256) {
    double c2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double t = U[d] - G.u(i, d);
      c2 = d == 0 ? t * t : c2 + t * t;
    }
    const double h = G.f(i, 0);
    dm = jl_maxd(dm, h);
    if constexpr (NDF == 2) {
      const double b = G.f(i, 1);
      dm = jl_maxd(dm, b);
      em = jl_maxd(em, h * c2 + b);
    } else {
      em = jl_maxd(em, h * c2);
    }
  }
  block_min2<4>(-dm, -em, red, res);
  if (threadIdx.x == 0) {
    out[2 * c] = -res[0] * P.wfine;
    out[2 * c + 1] = 0.5 * -res[1] * P.wfine;
  }
}

// refine / coarsen flags of every velocity cell of one cell (vs_refine!, vs_coarsen!,
// AMR.jl:24-111, with the criteria of Criteria.jl)
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(256) void k_vs_flags(DevArgs A, VsArgs P,
                                                  uint8_t* __restrict__ flags) {
  constexpr int NW = D + 2;
  __shared__ double red[2 * 4];
  __shared__ double scl[2];
  const int c = blockIdx.x;
  const Grid<D, NDF, UNI, FT> G(A, c);
  const VsGrid<D, NDF, UNI, FT> X{G, P};
  const int n = G.n;
  const bool lohner = P.mode == VS_LOHNER;
  // vs_lohner_scales: peaks of |h| and |b|
  double a1 = 0.0, a2 = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double x = fabs(G.f(i, 0));
    if (x > a1) a1 = x;
    if constexpr (NDF == 2) {
      const double y = fabs(G.f(i, 1));
      if (y > a2) a2 = y;
    }
  }
  block_min2<4>(-a1, -a2, red, scl);
  const double s1 = -scl[0], s2 = -scl[1];
  const double* w = &A.w[int64_t(c) * NW];
  const double* U = &A.prim[int64_t(c) * NW + 1];
  const double* ds = &A.rec_ds[int64_t(c) * D];
  double SU = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) SU = d == 0 ? U[d] * U[d] : SU + U[d] * U[d];
  const double den = w[NW - 1] - 0.5 * w[0] * SU;
  const double P2 = double(1 << D);
  const double tauC = 0.3 * P.tl;   // VS_LOHNER_COARSEN_RATIO * ADAPT_COEFFI_VS_LOHNER
  for (int i = threadIdx.x; i < n; i += 256) {
    double cdf[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NDF; ++k) {   // _criterion_cell!: df + max_d |sdf*ds|
      double m = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double x = fabs(G.s(i, k, d) * ds[d]);
        if (x > m) m = x;
      }
      cdf[k] = G.f(i, k) + m;
    }
    const double wt = G.wt(i);
    double S = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double t = U[d] - G.u(i, d);
      S = d == 0 ? t * t : S + t * t;
    }
    const double c1w = cdf[0] * wt;
    double loc_r, loc_c, e_w;
    if constexpr (NDF == 2) {
      e_w = 0.5 * (S * cdf[0] + cdf[1]) * wt;
      loc_r = jl_maxd(fabs(e_w) / den, c1w / w[0]);
      loc_c = jl_maxd(e_w / den, c1w / w[0]);
    } else {
      loc_r = jl_maxd(fabs(0.5 * S * cdf[0] * wt) / den, c1w / w[0]);
      e_w = 0.5 * (S * cdf[0]) * wt;
      loc_c = jl_maxd(e_w / den, c1w / w[0]);
    }
    const bool local_refine = loc_r > P.cl;
    const bool local_coarsen = loc_c < P.cl / P2;
    const bool global_refine =
        (NDF == 2 ? c1w > P.cg * P.vr[0] : c1w > P.vr[0] * P.cg) || e_w > P.vr[1] * P.cg;
    const bool global_coarsen =
        (NDF == 2 ? c1w < P.cg * P.vr[0] / P2 : c1w < P.vr[0] * P.cg / P2) &&
        e_w < P.vr[1] * P.cg / P2;
    double eta = 0.0;
    if (lohner) {   // vs_lohner_indicator, Criteria.jl:254-283
      const double h0 = G.f(i, 0), b0 = NDF == 2 ? G.f(i, NDF - 1) : 0.0;
      for (int d = 0; d < D; ++d) {
        const double hi = X.size(i, d);
        const int Ln = X.face(i, d, -1), Rn = X.face(i, d, +1);
        const double dsL = Ln < 0 ? hi : 0.5 * (hi + X.size(Ln, d));
        const double dsR = Rn < 0 ? hi : 0.5 * (hi + X.size(Rn, d));
        const double hL = Ln < 0 ? 0.0 : G.f(Ln, 0), hR = Rn < 0 ? 0.0 : G.f(Rn, 0);
        eta = jl_maxd(eta, lohner_ratio(hL, h0, hR, dsL, dsR, s1));
        if constexpr (NDF == 2) {
          const double bL = Ln < 0 ? 0.0 : G.f(Ln, 1), bR = Rn < 0 ? 0.0 : G.f(Rn, 1);
          eta = jl_maxd(eta, lohner_ratio(bL, b0, bR, dsL, dsR, s2));
        }
      }
    }
    const bool base = lohner ? (eta > P.tl || local_refine) : (local_refine || global_refine);
    const bool refine = X.lev(i) < P.L && base;
    const bool coarsen = lohner ? (eta < tauC && local_coarsen) : (local_coarsen && global_coarsen);
    flags[G.off + i] = uint8_t((refine ? VS_REFINE : 0) | (coarsen ? VS_COARSEN : 0));
  }
}

// The state on rebuilt grids: new velocity cell i of cell c takes old cell src >= 0, or the
// mean of the 2^D old cells from -(src+1) (summed in order, then divided, as
// coarsen_grid_stream! does); slopes are copied for unchanged cells and zero otherwise.
template <int D, int NDF, class FT>
__global__ __launch_bounds__(256) void k_vs_remap(DevArgs A, const int64_t* __restrict__ voff_old,
                                                  const FT* __restrict__ df_old,
                                                  const FT* __restrict__ sdf_old,
                                                  const int32_t* __restrict__ src,
                                                  const uint8_t* __restrict__ changed) {
  const int c = blockIdx.x;
  const int64_t on = A.voff[c], nn = A.voff[c + 1] - on;
  const int64_t oo = voff_old[c], no = voff_old[c + 1] - oo;
  FT* df = reinterpret_cast<FT*>(A.df_out);
  FT* sdf = reinterpret_cast<FT*>(A.sdf);
  const bool ch = changed[c] != 0;
  for (int64_t i = threadIdx.x; i < nn; i += 256) {
    const int32_t j = src[on + i];
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      double v;
      if (j >= 0) {
        v = double(df_old[oo * NDF + k * no + j]);
      } else {
        const int64_t j0 = -int64_t(j) - 1;
        double sum = 0.0;
        for (int g = 0; g < (1 << D); ++g) sum += double(df_old[oo * NDF + k * no + j0 + g]);
        v = sum / double(1 << D);
      }
      df[on * NDF + k * nn + i] = FT(v);
#pragma unroll
      for (int d = 0; d < D; ++d)
        sdf[on * NDF * D + (d * NDF + k) * nn + i] =
            ch ? FT(0) : sdf_old[oo * NDF * D + (d * NDF + k) * no + i];
    }
  }
}

// conserved_I_porjection! of one listed cell's h onto its w (src/Theory/I-projection.jl:
// 144-158): e_int removed from the energy for 2D2F, shave and Newton as in CIP_Marching,
// then h *= exp(lam.psi), in place.
template <int D, int NDF, bool UNI, class FT, int NWV>
__global__ __launch_bounds__(NWV * 64) void k_vs_project(DevArgs A, const int32_t* __restrict__ cells,
                                                         double* __restrict__ scratch,
                                                         FT* __restrict__ dfw) {
  constexpr int M = D + 2, BS = NWV * 64;
  __shared__ double red[2 * NWV * red_stride((D + 2) * (D + 5) / 2)];   // G and J of the Newton step
  __shared__ double st[64];
  extern __shared__ double lds[];
  const int c = cells[blockIdx.x];
  const Grid<D, NDF, UNI, FT> G(A, c);
  const int n = G.n;
  double* sc = scratch ? &scratch[G.off * (NDF + 1)] : lds;
  double eint = 0.0, fmn = HUGE_VAL, fmp = HUGE_VAL;
  for (int p = threadIdx.x; p < n; p += BS) {
    const int i = G.vi(p);
    const double h = G.f(i, 0);
    sc[i] = h;
    if constexpr (NDF == 2) eint += G.wt(i) * G.f(i, 1);
    fmn = h < fmn ? h : fmn;
    if (h > 0.0 && h < fmp) fmp = h;
  }
  {
    double v[1] = {eint};
    block_reduce<1, NWV>(v, 1, red, &st[40]);
    block_min2<NWV>(fmn, fmp, red, &st[46]);
  }
  if (threadIdx.x == 0) {
    double wn = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      st[m] = A.w[int64_t(c) * M + m];
      st[5 + m] = 0.0;
    }
    if constexpr (NDF == 2) st[M - 1] -= st[40] / 2;   // e_int = dot(weight, b)/2
#pragma unroll
    for (int m = 0; m < M; ++m) wn += st[m] * st[m];
    st[16] = 1e-10 * fmax(1.0, sqrt(wn));
    st[17] = HUGE_VAL;
    st[18] = 0.0;
    st[19] = 0.0;
    const double f_min = 1.1 * st[46];
    st[24] = f_min;
    st[25] = st[47];
    st[26] = st[47] - f_min;
    st[27] = 1.0 / st[26];
    st[28] = f_min < 0 ? 1.0 : 0.0;
  }
  __syncthreads();
  cip_shave_newton<D, NDF, UNI, FT, NWV>(G, n, sc, red, st);
  double lm[M];
#pragma unroll
  for (int m = 0; m < M; ++m) lm[m] = st[5 + m];
  FT* fo = dfw + G.off * NDF;
  for (int p = threadIdx.x; p < n; p += BS) {
    const int i = G.vi(p);
    double u[D], ps[M];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = G.u(i, d);
    cip_psi<D>(u, ps);
    double dot = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) dot += lm[m] * ps[m];
    G.at(fo, i) = FT(sc[i] * exp(dot));
  }
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_vs_t(const DevArgs& A, int which, const VsArgs& P, int nc, double* out,
                              uint8_t* flags, hipStream_t st) {
  if (nc <= 0) return hipSuccess;
  if (which == 0)
    hipLaunchKernelGGL((k_vs_res<D, NDF, UNI, FT>), dim3(nc), dim3(256), 0, st, A, P, out);
  else
    hipLaunchKernelGGL((k_vs_flags<D, NDF, UNI, FT>), dim3(nc), dim3(256), 0, st, A, P, flags);
  return hipGetLastError();
}
template <int D, int NDF, class FT>
inline hipError_t launch_vs_remap_t(const DevArgs& A, const int64_t* voff_old, const void* df_old,
                                    const void* sdf_old, const int32_t* src,
                                    const uint8_t* changed, int nc, hipStream_t st) {
  if (nc <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_vs_remap<D, NDF, FT>), dim3(nc), dim3(256), 0, st, A, voff_old,
                     static_cast<const FT*>(df_old), static_cast<const FT*>(sdf_old), src,
                     changed);
  return hipGetLastError();
}
template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_vs_project_t(const DevArgs& A, const int32_t* cells, int n, size_t lds,
                                      double* scratch, void* dfw, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_vs_project<D, NDF, UNI, FT, 8>), dim3(n), dim3(512),
                     scratch ? 0 : lds, st, A, cells, scratch, static_cast<FT*>(dfw));
  return hipGetLastError();
}

// ------------------------------------------------------------------ K3: macro slope
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(256) void k_macro_slope(DevArgs A, const int32_t* __restrict__ cells,
                                                     int ncells, double* __restrict__ sw) {
  constexpr int NW = D + 2;
  const int c = cells[blockIdx.x];
  const Grid<D, NDF, UNI, FT> G(A, c);
  __shared__ double red[2 * 4 * red_stride(NW)];
  __shared__ double tot[NW];
  for (int d = 0; d < D; ++d) {
    double acc[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) acc[q] = 0.0;
    for (int i = threadIdx.x; i < G.n; i += 256) {
      double u[D], s[NDF], m[NW];
#pragma unroll
      for (int t = 0; t < D; ++t) u[t] = G.u(i, t);
#pragma unroll
      for (int k = 0; k < NDF; ++k) s[k] = G.s(i, k, d);
      moment_terms<D, NDF>(G.wt(i), u, s, m);
#pragma unroll
      for (int q = 0; q < NW; ++q) acc[q] += m[q];
    }
    block_reduce<NW, 4>(acc, NW, red, tot);
    if (int(threadIdx.x) < NW)
      sw[(int64_t(c) * NW + threadIdx.x) * D + d] =
          (int(threadIdx.x) == NW - 1) ? 0.5 * tot[threadIdx.x] : tot[threadIdx.x];
  }
}

// ------------------------------------------------------------------ launchers
template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_slope_t(const DevArgs& A, const int32_t* cells, int n, int maxh,
                                 hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (maxh <= 4)
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, FT, 4, 4>), dim3(n), dim3(256), 0, st, A, cells, n,
                       KAH_RO_ARGS(A));
  else if (maxh <= 8)
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, FT, 8, 4>), dim3(n), dim3(256), 0, st, A, cells, n,
                       KAH_RO_ARGS(A));
  else if (maxh <= 16)
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, FT, 16, 4>), dim3(n), dim3(256), 0, st, A, cells, n,
                       KAH_RO_ARGS(A));
  else
    hipLaunchKernelGGL((k_slope<D, NDF, UNI, FT, 24, 4>), dim3(n), dim3(256), 0, st, A, cells, n,
                       KAH_RO_ARGS(A));
  return hipGetLastError();
}

template <int D, int NDF, bool UNI, class FT, int VPT, int WPI>
inline hipError_t launch_slope_chunk_w(const DevArgs& A, const int2* work, int n, int maxh,
                                       hipStream_t st) {
  const dim3 grid(n), block(WPI * 64);
  if (maxh == -1)   // regular cells, neighbours on the cell's grid: 2*D half fluxes
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 2 * D, VPT, 1, WPI>), grid, block, 0,
                       st, A, work, n, KAH_RO_ARGS(A));
  else if (maxh == -2)   // regular cells, neighbours on other grids
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 2 * D, VPT, 2, WPI>), grid, block, 0,
                       st, A, work, n, KAH_RO_ARGS(A));
  else if (maxh <= 4)
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 4, VPT, 0, WPI>), grid, block, 0, st,
                       A, work, n, KAH_RO_ARGS(A));
  else if (maxh <= 6)
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 6, VPT, 0, WPI>), grid, block, 0, st,
                       A, work, n, KAH_RO_ARGS(A));
  else if (maxh <= 8)
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 8, VPT, 0, WPI>), grid, block, 0, st,
                       A, work, n, KAH_RO_ARGS(A));
  else if (maxh <= 16)
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 16, VPT, 0, WPI>), grid, block, 0, st,
                       A, work, n, KAH_RO_ARGS(A));
  else
    hipLaunchKernelGGL((k_slope_chunk<D, NDF, UNI, FT, 24, VPT, 0, WPI>), grid, block, 0, st,
                       A, work, n, KAH_RO_ARGS(A));
  return hipGetLastError();
}

template <int D, int NDF, bool UNI, class FT, int VPT>
inline hipError_t launch_slope_chunk_v(const DevArgs& A, const int2* work, int n, int maxh,
                                       int wpi, hipStream_t st) {
  return wpi == 3 ? launch_slope_chunk_w<D, NDF, UNI, FT, VPT, 3>(A, work, n, maxh, st)
                  : launch_slope_chunk_w<D, NDF, UNI, FT, VPT, 4>(A, work, n, maxh, st);
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_slope_chunk_t(const DevArgs& A, const int2* work, int n, int maxh,
                                       int vpt, int wpi, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (vpt) {
    case 1: return launch_slope_chunk_v<D, NDF, UNI, FT, 1>(A, work, n, maxh, wpi, st);
    case 2: return launch_slope_chunk_v<D, NDF, UNI, FT, 2>(A, work, n, maxh, wpi, st);
    case 3: return launch_slope_chunk_v<D, NDF, UNI, FT, 3>(A, work, n, maxh, wpi, st);
    case 4: return launch_slope_chunk_v<D, NDF, UNI, FT, 4>(A, work, n, maxh, wpi, st);
    case 5: return launch_slope_chunk_v<D, NDF, UNI, FT, 5>(A, work, n, maxh, wpi, st);
    case 6: return launch_slope_chunk_v<D, NDF, UNI, FT, 6>(A, work, n, maxh, wpi, st);
    default: return hipErrorInvalidValue;
  }
}

template <int D, int NDF, bool UNI, class FT, int MODE, int MARCH, int NWV>
inline hipError_t launch_fr_t(const DevArgs& A, const int32_t* cells, int n, size_t lds,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (lds > 0)
    hipLaunchKernelGGL((k_flux_relax<D, NDF, UNI, FT, MODE, MARCH, true, NWV>), dim3(n),
                       dim3(NWV * 64), lds, st, A, cells, n, KAH_RO_ARGS(A), A.sdf_r,
                       A.half_fw_r, A.rho_w_r);
  else
    hipLaunchKernelGGL((k_flux_relax<D, NDF, UNI, FT, MODE, MARCH, false, NWV>), dim3(n),
                       dim3(NWV * 64), 0, st, A, cells, n, KAH_RO_ARGS(A), A.sdf_r,
                       A.half_fw_r, A.rho_w_r);
  return hipGetLastError();
}

template <int D, int NDF, bool UNI, class FT, int NWV>
inline hipError_t launch_fr_w(const DevArgs& A, const int32_t* cells, int n, int mode,
                              int march, size_t lds, hipStream_t st) {
  if (mode == MODE_FLUX)
    return launch_fr_t<D, NDF, UNI, FT, MODE_FLUX, MARCH_CAIDVM, NWV>(A, cells, n, 0, st);
  if (mode == MODE_ITER) {
    if (march == MARCH_CAIDVM)
      return launch_fr_t<D, NDF, UNI, FT, MODE_ITER, MARCH_CAIDVM, NWV>(A, cells, n, lds, st);
    if (march == MARCH_CIP)
      return launch_fr_t<D, NDF, UNI, FT, MODE_ITER, MARCH_CIP, NWV>(A, cells, n, lds, st);
    return launch_fr_t<D, NDF, UNI, FT, MODE_ITER, MARCH_EULER, NWV>(A, cells, n, lds, st);
  }
  if (mode == MODE_FUSED_REG) {
    if (march == MARCH_CAIDVM)
      return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED_REG, MARCH_CAIDVM, NWV>(A, cells, n, lds, st);
    if (march == MARCH_CIP)
      return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED_REG, MARCH_CIP, NWV>(A, cells, n, lds, st);
    return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED_REG, MARCH_EULER, NWV>(A, cells, n, lds, st);
  }
  if constexpr (!UNI) {   // on a shared grid every neighbour is on the cell's grid
    if (mode == MODE_FUSED_REGX) {
      if (march == MARCH_CAIDVM)
        return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED_REGX, MARCH_CAIDVM, NWV>(A, cells, n, lds, st);
      if (march == MARCH_CIP)
        return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED_REGX, MARCH_CIP, NWV>(A, cells, n, lds, st);
      return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED_REGX, MARCH_EULER, NWV>(A, cells, n, lds, st);
    }
  }
  if (march == MARCH_CAIDVM)
    return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED, MARCH_CAIDVM, NWV>(A, cells, n, lds, st);
  if (march == MARCH_CIP)
    return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED, MARCH_CIP, NWV>(A, cells, n, lds, st);
  return launch_fr_t<D, NDF, UNI, FT, MODE_FUSED, MARCH_EULER, NWV>(A, cells, n, lds, st);
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_fr_d(const DevArgs& A, const int32_t* cells, int n, int mode,
                              int march, size_t lds, int nwv, hipStream_t st) {
  if (nwv == 2) return launch_fr_w<D, NDF, UNI, FT, 2>(A, cells, n, mode, march, lds, st);
  if (nwv == 3) return launch_fr_w<D, NDF, UNI, FT, 3>(A, cells, n, mode, march, lds, st);
  if (nwv == 4) return launch_fr_w<D, NDF, UNI, FT, 4>(A, cells, n, mode, march, lds, st);
  return launch_fr_w<D, NDF, UNI, FT, 8>(A, cells, n, mode, march, lds, st);
}

// ------------------------------------------------------------------ immersed boundary
// update_solid_cell! and update_solid_neighbor! (src/Boundary/Immersed_boundary.jl:122-148,
// 367-483) between the slopes and the flux: one workgroup per solid cell or IB face, the
// threads striding over the velocity cells of the target grid; sources on other grids are
// read through their key-interval overlap. Arithmetic in FP64.

// threads per workgroup of the immersed-boundary kernels: few workgroups (one per solid cell
// or IB face) over grids of up to ~10^4 velocity cells, so each gets 8 waves (more would cap
// the registers below what the 3D kernels need)
constexpr int IB_THREADS = 512;

// sum of one value over the IB_THREADS-thread workgroup (every thread gets the sum; fixed
// order: wave sums, then the waves' partials in index order)
__device__ inline double block_sum_ib(double v, double* sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = sh[0];
#pragma unroll
  for (int k = 1; k < IB_THREADS / 64; ++k) t += sh[k];
  return t;
}

// sum_j wn_j (f_j + sdf_j . (x - x_j)) over the listed cells at velocity cell i, with
// wn_j = w_j / sum w, w_j = max(0, u.l_j/|u|)^2, l_j = (x - x_j)/|x - x_j| (1/N when the sum
// is zero): vs_extrapolate! with the weights of update_solid_cell! / image_df (:24-47,
// :122-141, :378-395). Sums of products in index order, as Julia's dot and norm.
template <int D, int NDF, bool UNI, class FT>
__device__ inline void ib_extrapolate(const DevArgs& A, const int32_t* cells, int N,
                                      const double* cmid, const double* x, int i, int li,
                                      uint32_t ki, const double (&u)[D], double (&out)[NDF]) {
  double nu = u[0] * u[0];
#pragma unroll
  for (int t = 1; t < D; ++t) nu += u[t] * u[t];
  nu = sqrt(nu);
  double w[8], ws = 0.0;
  for (int j = 0; j < N; ++j) {
    const double* xj = &cmid[int64_t(cells[j]) * D];
    double l[D];
#pragma unroll
    for (int t = 0; t < D; ++t) l[t] = x[t] - xj[t];
    double nl = l[0] * l[0];
#pragma unroll
    for (int t = 1; t < D; ++t) nl += l[t] * l[t];
    nl = sqrt(nl);
#pragma unroll
    for (int t = 0; t < D; ++t) l[t] = l[t] / nl;
    double d = u[0] * l[0];
#pragma unroll
    for (int t = 1; t < D; ++t) d += u[t] * l[t];
    const double r = d / nu;
    const double c = (r != r) ? r : (r > 0.0 ? r : 0.0);
    w[j] = c * c;
    ws = (j == 0) ? w[j] : ws + w[j];
  }
#pragma unroll
  for (int k = 0; k < NDF; ++k) out[k] = 0.0;
  for (int j = 0; j < N; ++j) {
    const double wn = ws == 0.0 ? 1.0 / double(N) : w[j] / ws;
    const Grid<D, NDF, UNI, FT> G(A, cells[j]);
    const double* xj = &cmid[int64_t(cells[j]) * D];
    double dx[D];
#pragma unroll
    for (int t = 0; t < D; ++t) dx[t] = x[t] - xj[t];
    // the source cells over target cell i: same level or coarser (injection) with factor 1,
    // finer ones with 2^-D per level (vs_extrapolate!'s walk)
    for_overlap<D, NDF, UNI, FT>(G, i, li, ki, A.lmax, [&](int jj, double den, double rden) {
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        double ddf = G.s(jj, k, 0) * dx[0];
#pragma unroll
        for (int t = 1; t < D; ++t) ddf += G.s(jj, k, t) * dx[t];
        double v = G.f(jj, k) + ddf;
        if (den != 1.0) v = v * rden;   // exact: den is a power of two
        out[k] += v * wn;
      }
    });
  }
}

// update_solid_cell!: f of solid cell s from its fluid face neighbours, written to both
// state buffers (this step's IB faces read it; the next step keeps it), then w and prim
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(IB_THREADS) void k_ib_solid(DevArgs A, const int32_t* __restrict__ solid,
                                                  const int32_t* __restrict__ nb_off,
                                                  const int32_t* __restrict__ nb,
                                                  const double* __restrict__ cmid, FT* df_a,
                                                  FT* df_b) {
  constexpr int NW = D + 2;
  __shared__ double red[IB_THREADS / 64];
  const int q = blockIdx.x;
  const int s = solid[q];
  const int b = nb_off[q], N = nb_off[q + 1] - b;
  const Grid<D, NDF, UNI, FT> G(A, s);
  const int n = G.n;
  const double* xs = &cmid[int64_t(s) * D];
  double m[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) m[r] = 0.0;
  for (int i = threadIdx.x; i < n; i += IB_THREADS) {
    double u[D];
#pragma unroll
    for (int t = 0; t < D; ++t) u[t] = G.u(i, t);
    double f[NDF];
    ib_extrapolate<D, NDF, UNI, FT>(A, nb + b, N, cmid, xs, i, G.lev(i), G.key(i), u, f);
    double fs[NDF];
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      const FT v = FT(f[k]);
      df_a[G.off * NDF + int64_t(k) * n + i] = v;
      df_b[G.off * NDF + int64_t(k) * n + i] = v;
      fs[k] = double(v);
    }
    double mt[NW];
    moment_terms<D, NDF, double>(G.wt(i), u, fs, mt);
#pragma unroll
    for (int r = 0; r < NW; ++r) m[r] += mt[r];
  }
  double w[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) w[r] = block_sum_ib(m[r], red);
  w[NW - 1] = 0.5 * w[NW - 1];
  if (threadIdx.x == 0) {
    double p[NW];
    get_prim<D>(w, A.gamma, p);
#pragma unroll
    for (int r = 0; r < NW; ++r) {
      A.w[int64_t(s) * NW + r] = w[r];
      A.prim[int64_t(s) * NW + r] = p[r];
    }
  }
}

// update_solid_neighbor! of one IB face: image_df, boundary_slope! and its positivity
// scaling, the wall Maxwellian of zero net mass flux through the wall (cvc_Mu /
// cvc_density) and the cut-cell mixing (cvc_correction!). The SolidNeighbor's df goes to
// the pseudo cell's rows of both state buffers; the face's there half (df .* u[dir] over
// rot*u[dir] > 0, calc_flux with a solid there side, src/Flux/CAIDVM.jl:110-112) goes to
// its half_fw row, which the flux kernel sums like any other.
template <int D, int NDF, bool UNI, class FT>
__global__ __launch_bounds__(IB_THREADS) void k_ib_face(DevArgs A, const IbFaceDesc* __restrict__ faces,
                                                 const int32_t* __restrict__ img,
                                                 const int32_t* __restrict__ cut_idx,
                                                 const double2* __restrict__ cut_w,
                                                 const double* __restrict__ cmid, FT* df_a,
                                                 FT* df_b) {
  constexpr int NW = D + 2;
  __shared__ double red[IB_THREADS / 64];
  const IbFaceDesc F = faces[blockIdx.x];
  const Grid<D, NDF, UNI, FT> G(A, F.donor);
  const Grid<D, NDF, UNI, FT> S(A, F.solid);
  const Grid<D, NDF, UNI, FT> PS(A, F.pseudo);   // the donor's grid, copied
  const int n = G.n;
  FT* pa = df_a + PS.off * NDF;
  FT* pb = df_b + PS.off * NDF;
  FT* ps = reinterpret_cast<FT*>(A.sdf) + PS.off * NDF * D;
  const int32_t* ci = cut_idx + F.cut_beg;
  // the cut-cell slot of velocity cell i, or -1 (the list is ascending)
  auto cut_of = [&](int i) {
    int lo = 0, hi = F.cut_cnt - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = ci[mid];
      if (v == i) return mid;
      if (v < i) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
  };
  // pass 1: the extrapolated df at the wall, and the two flux sums of the wall density
  double sf = 0.0, mur = 0.0;
  for (int i = threadIdx.x; i < n; i += IB_THREADS) {
    double u[D];
#pragma unroll
    for (int t = 0; t < D; ++t) u[t] = G.u(i, t);
    const int li = G.lev(i);
    const uint32_t ki = G.key(i);
    double im[NDF];
    ib_extrapolate<D, NDF, UNI, FT>(A, img + F.img_beg, F.img_cnt, cmid, F.ip, i, li, ki, u, im);
    // boundary_slope!'s one-sided difference with the solid cell, walked across the grids
    double sLv[NDF];
#pragma unroll
    for (int k = 0; k < NDF; ++k) sLv[k] = 0.0;
    for_overlap<D, NDF, UNI, FT>(S, i, li, ki, A.lmax, [&](int jj, double den, double rden) {
#pragma unroll
      for (int k = 0; k < NDF; ++k) {
        double d = G.f(i, k) - S.f(jj, k);
        if (den != 1.0) d = d * rden;
        sLv[k] += d / F.dxs;
      }
    });
    double aux[NDF];
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      const double fc = G.f(i, k);
      const double sL = sLv[k];
      double sl = minmod(sL, (im[k] - fc) / F.dxf);
      sl = jl_min(fabs((im[k] - MACH_EPS) / (sl * F.dxL + MACH_EPS)), 1.0) * sl;
      aux[k] = im[k] + sl * F.dxL;
      pa[k * n + i] = FT(aux[k]);
      if (A.ib_ssdf) A.ib_ssdf[F.row * NDF + int64_t(k) * n + i] = sl;
      // the SolidNeighbor's slopes: the limited boundary slope along the face direction,
      // zero across it (read by the DVM flux's reconstruction of the wall side)
#pragma unroll
      for (int d = 0; d < D; ++d) ps[(d * NDF + k) * n + i] = d == F.dir ? FT(sl) : FT(0);
    }
    double vn = u[0] * F.nrm[0];
#pragma unroll
    for (int t = 1; t < D; ++t) vn += u[t] * F.nrm[t];
    double M[2];
    maxwell<D>(u, F.bc, A.K, M);
    const int j = cut_of(i);
    if (j >= 0) {
      const double2 gw = cut_w[F.cut_beg + j];
      mur += gw.y * vn * M[0];
      sf += gw.x * vn * aux[0];
    } else if (vn >= 0.0) {
      mur += G.wt(i) * vn * M[0];
    } else {
      sf += G.wt(i) * vn * aux[0];
    }
  }
  const double SF = block_sum_ib(sf, red);
  const double MuR = block_sum_ib(mur, red);
  const double rho = -SF / MuR;
  // pass 2: the wall Maxwellian on u.n >= 0, the cut-cell mix, the there-half moments. Each
  // thread revisits the velocity cells it wrote in pass 1, so no barrier is needed here.
  double hm[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) hm[r] = 0.0;
  for (int i = threadIdx.x; i < n; i += IB_THREADS) {
    double u[D];
#pragma unroll
    for (int t = 0; t < D; ++t) u[t] = G.u(i, t);
    double vn = u[0] * F.nrm[0];
#pragma unroll
    for (int t = 1; t < D; ++t) vn += u[t] * F.nrm[t];
    double M[2];
    maxwell<D>(u, F.bc, A.K, M);
    const int j = cut_of(i);
    double out[NDF];
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      const double Mk = M[k] * rho;
      const double g = double(pa[k * n + i]);
      if (j >= 0) {
        const double2 gw = cut_w[F.cut_beg + j];
        out[k] = (gw.x * g + gw.y * Mk) / (gw.x + gw.y);
      } else {
        out[k] = vn >= 0.0 ? Mk : g;
      }
    }
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      const FT v = FT(out[k]);
      pa[k * n + i] = v;
      pb[k * n + i] = v;
      out[k] = double(v);
    }
    const double ud = pick<D>(u, F.dir);
    if (F.rot * ud > 0.0) {
      double mic[NDF], mt[NW];
#pragma unroll
      for (int k = 0; k < NDF; ++k) mic[k] = out[k] * ud;
      moment_terms<D, NDF, double>(G.wt(i), u, mic, mt);
#pragma unroll
      for (int r = 0; r < NW; ++r) hm[r] += mt[r];
    }
  }
  double tot[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) tot[r] = block_sum_ib(hm[r], red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NW; ++r) A.half_fw[A.part_off[F.slot] * NW + r] = tot[r];
  }
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_ib_t(const DevArgs& A, int ns, const int32_t* solid,
                              const int32_t* nb_off, const int32_t* nb, int nf,
                              const IbFaceDesc* faces, const int32_t* img, const int32_t* cut_idx,
                              const double2* cut_w, const double* cmid, void* df_a, void* df_b,
                              hipStream_t st) {
  if (ns > 0) {
    hipLaunchKernelGGL((k_ib_solid<D, NDF, UNI, FT>), dim3(ns), dim3(IB_THREADS), 0, st, A, solid, nb_off, nb,
                       cmid, static_cast<FT*>(df_a), static_cast<FT*>(df_b));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (nf > 0) {
    hipLaunchKernelGGL((k_ib_face<D, NDF, UNI, FT>), dim3(nf), dim3(IB_THREADS), 0, st, A, faces, img, cut_idx,
                       cut_w, cmid, static_cast<FT*>(df_a), static_cast<FT*>(df_b));
    return hipGetLastError();
  }
  return hipSuccess;
}

template <int D, int NDF, bool UNI, class FT>
inline hipError_t launch_macro_slope_t(const DevArgs& A, const int32_t* cells, int n, double* sw,
                                       hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_macro_slope<D, NDF, UNI, FT>), dim3(n), dim3(256), 0, st, A, cells, n, sw);
  return hipGetLastError();
}

}  // namespace kah

// kah_kernels.hip -- the plain copy kernel (bandwidth reference) and the dispatch of the
// kernel launchers over the physical dimension. The kernels themselves are in kah_device.h,
// instantiated per dimension and storage type in kah_slope_{2d,3d}[_f32].hip and
// kah_flux_{2d,3d}[_f32].hip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kah_kernels.h"

namespace kah {

hipError_t launch_overlap_map_2d(const DevArgs&, const OvJob*, int, uint32_t*, hipStream_t);
#define KAH_MIRROR_DECL(name)                                                                 \
  hipError_t name(const DevArgs&, bool, const int32_t*, const int64_t*, int, int, int, double*, \
                  const double*, const int32_t*, double*, void*, hipStream_t);
KAH_MIRROR_DECL(launch_mirror_2d)
KAH_MIRROR_DECL(launch_mirror_3d)
KAH_MIRROR_DECL(launch_mirror_2d_f32)
KAH_MIRROR_DECL(launch_mirror_3d_f32)
#undef KAH_MIRROR_DECL
#define KAH_RESTART_DECL(name)                                                                \
  hipError_t name(const DevArgs&, bool, const int64_t*, int64_t, int64_t, int, double*, int8_t*, \
                  double*, double*, void*, hipStream_t);
KAH_RESTART_DECL(launch_restart_2d)
KAH_RESTART_DECL(launch_restart_3d)
KAH_RESTART_DECL(launch_restart_2d_f32)
KAH_RESTART_DECL(launch_restart_3d_f32)
#undef KAH_RESTART_DECL
#define KAH_IB_DECL(name)                                                                     \
  hipError_t name(const DevArgs&, bool, int, const int32_t*, const int32_t*, const int32_t*, int, \
                  const IbFaceDesc*, const int32_t*, const int32_t*, const double2*,            \
                  const double*, void*, void*, hipStream_t);
KAH_IB_DECL(launch_ib_2d)
KAH_IB_DECL(launch_ib_3d)
KAH_IB_DECL(launch_ib_2d_f32)
KAH_IB_DECL(launch_ib_3d_f32)
#undef KAH_IB_DECL
#define KAH_VS_DECL(sfx)                                                                        \
  hipError_t launch_vs_##sfx(const DevArgs&, bool, int, const VsArgs&, int, double*, uint8_t*,  \
                             hipStream_t);                                                      \
  hipError_t launch_vs_remap_##sfx(const DevArgs&, const int64_t*, const void*, const void*,    \
                                   const int32_t*, const uint8_t*, int, hipStream_t);           \
  hipError_t launch_vs_project_##sfx(const DevArgs&, bool, const int32_t*, int, size_t,         \
                                     double*, void*, hipStream_t);
KAH_VS_DECL(2d)
KAH_VS_DECL(3d)
KAH_VS_DECL(2d_f32)
KAH_VS_DECL(3d_f32)
#undef KAH_VS_DECL
hipError_t launch_overlap_map_3d(const DevArgs&, const OvJob*, int, uint32_t*, hipStream_t);
hipError_t launch_slope_2d(const DevArgs&, bool, const int32_t*, int, int, hipStream_t);
hipError_t launch_slope_chunk_2d(const DevArgs&, bool, const int2*, int, int, int, int, hipStream_t);
hipError_t launch_macro_slope_2d(const DevArgs&, bool, const int32_t*, int, double*, hipStream_t);
hipError_t launch_flux_relax_2d(const DevArgs&, bool, const int32_t*, int, int, int, size_t, int,
                                  hipStream_t);
hipError_t launch_slope_3d(const DevArgs&, bool, const int32_t*, int, int, hipStream_t);
hipError_t launch_slope_chunk_3d(const DevArgs&, bool, const int2*, int, int, int, int, hipStream_t);
hipError_t launch_macro_slope_3d(const DevArgs&, bool, const int32_t*, int, double*, hipStream_t);
hipError_t launch_flux_relax_3d(const DevArgs&, bool, const int32_t*, int, int, int, size_t, int,
                                  hipStream_t);
hipError_t launch_slope_2d_f32(const DevArgs&, bool, const int32_t*, int, int, hipStream_t);
hipError_t launch_slope_chunk_2d_f32(const DevArgs&, bool, const int2*, int, int, int, int, hipStream_t);
hipError_t launch_macro_slope_2d_f32(const DevArgs&, bool, const int32_t*, int, double*, hipStream_t);
hipError_t launch_flux_relax_2d_f32(const DevArgs&, bool, const int32_t*, int, int, int, size_t, int,
                                  hipStream_t);
hipError_t launch_slope_3d_f32(const DevArgs&, bool, const int32_t*, int, int, hipStream_t);
hipError_t launch_slope_chunk_3d_f32(const DevArgs&, bool, const int2*, int, int, int, int, hipStream_t);
hipError_t launch_macro_slope_3d_f32(const DevArgs&, bool, const int32_t*, int, double*, hipStream_t);
hipError_t launch_flux_relax_3d_f32(const DevArgs&, bool, const int32_t*, int, int, int, size_t, int,
                                  hipStream_t);

// ------------------------------------------------------------------ copy reference
// A plain streaming copy used to measure attainable HBM bandwidth and to calibrate hardware
// byte counters against a kernel of known traffic: each thread has COPY_UNROLL 16-byte loads
// in flight (coalesced: consecutive threads read consecutive 16-byte words) before it stores.
constexpr int COPY_UNROLL = 4;
typedef double copy_vec __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy(const copy_vec* __restrict__ src,
                                              copy_vec* __restrict__ dst, int64_t n) {
  const int64_t step = int64_t(gridDim.x) * 256 * COPY_UNROLL;
  for (int64_t b = int64_t(blockIdx.x) * 256 * COPY_UNROLL + threadIdx.x; b < n; b += step) {
    copy_vec v[COPY_UNROLL];
#pragma unroll
    for (int k = 0; k < COPY_UNROLL; ++k) {
      const int64_t i = b + int64_t(k) * 256;
      if (i < n) v[k] = __builtin_nontemporal_load(&src[i]);
    }
#pragma unroll
    for (int k = 0; k < COPY_UNROLL; ++k) {
      const int64_t i = b + int64_t(k) * 256;
      if (i < n) __builtin_nontemporal_store(v[k], &dst[i]);
    }
  }
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st) {
  const int64_t n = bytes / 16;
  const int64_t per_block = 256 * COPY_UNROLL;
  int blocks = int(std::min<int64_t>((n + per_block - 1) / per_block, 256 * 16));
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, static_cast<const copy_vec*>(src),
                     static_cast<copy_vec*>(dst), n);
  return hipGetLastError();
}

hipError_t launch_slope(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                        int n, int maxh, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_slope_2d_f32(A, uni, cells, n, maxh, st)
                    : launch_slope_3d_f32(A, uni, cells, n, maxh, st);
  return dim == 2 ? launch_slope_2d(A, uni, cells, n, maxh, st)
                  : launch_slope_3d(A, uni, cells, n, maxh, st);
}

hipError_t launch_slope_chunk(const DevArgs& A, int dim, bool uni, bool f32, const int2* work,
                              int n, int maxh, int vpt, int wpi, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_slope_chunk_2d_f32(A, uni, work, n, maxh, vpt, wpi, st)
                    : launch_slope_chunk_3d_f32(A, uni, work, n, maxh, vpt, wpi, st);
  return dim == 2 ? launch_slope_chunk_2d(A, uni, work, n, maxh, vpt, wpi, st)
                  : launch_slope_chunk_3d(A, uni, work, n, maxh, vpt, wpi, st);
}

hipError_t launch_flux_relax(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                             int n, int mode, int march, size_t lds, int nwv, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_flux_relax_2d_f32(A, uni, cells, n, mode, march, lds, nwv, st)
                    : launch_flux_relax_3d_f32(A, uni, cells, n, mode, march, lds, nwv, st);
  return dim == 2 ? launch_flux_relax_2d(A, uni, cells, n, mode, march, lds, nwv, st)
                  : launch_flux_relax_3d(A, uni, cells, n, mode, march, lds, nwv, st);
}

hipError_t launch_mirror(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                         const int64_t* off, int n, int kind, int unpack, double* buf,
                         const double* mid, const int32_t* benc, double* sw, void* dfw,
                         hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_mirror_2d_f32(A, uni, cells, off, n, kind, unpack, buf, mid, benc, sw,
                                           dfw, st)
                    : launch_mirror_3d_f32(A, uni, cells, off, n, kind, unpack, buf, mid, benc, sw,
                                           dfw, st);
  return dim == 2 ? launch_mirror_2d(A, uni, cells, off, n, kind, unpack, buf, mid, benc, sw, dfw,
                                     st)
                  : launch_mirror_3d(A, uni, cells, off, n, kind, unpack, buf, mid, benc, sw, dfw,
                                     st);
}

hipError_t launch_ib(const DevArgs& A, int dim, bool uni, bool f32, int ns, const int32_t* solid,
                     const int32_t* nb_off, const int32_t* nb, int nf, const IbFaceDesc* faces,
                     const int32_t* img, const int32_t* cut_idx, const double2* cut_w,
                     const double* cmid, void* df_a, void* df_b, hipStream_t st) {
  auto fn = dim == 2 ? (f32 ? launch_ib_2d_f32 : launch_ib_2d) : (f32 ? launch_ib_3d_f32 : launch_ib_3d);
  return fn(A, uni, ns, solid, nb_off, nb, nf, faces, img, cut_idx, cut_w, cmid, df_a, df_b, st);
}

hipError_t launch_restart(const DevArgs& A, int dim, bool uni, bool f32, const int64_t* row,
                          int64_t N, int64_t Nv, int load, double* ws, int8_t* lev, double* vmid,
                          double* vdf, void* dfw, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_restart_2d_f32(A, uni, row, N, Nv, load, ws, lev, vmid, vdf, dfw, st)
                    : launch_restart_3d_f32(A, uni, row, N, Nv, load, ws, lev, vmid, vdf, dfw, st);
  return dim == 2 ? launch_restart_2d(A, uni, row, N, Nv, load, ws, lev, vmid, vdf, dfw, st)
                  : launch_restart_3d(A, uni, row, N, Nv, load, ws, lev, vmid, vdf, dfw, st);
}

hipError_t launch_vs_res(const DevArgs& A, int dim, bool uni, bool f32, const VsArgs& P, int nc,
                         double* out, hipStream_t st) {
  if (f32) return dim == 2 ? launch_vs_2d_f32(A, uni, 0, P, nc, out, nullptr, st)
                           : launch_vs_3d_f32(A, uni, 0, P, nc, out, nullptr, st);
  return dim == 2 ? launch_vs_2d(A, uni, 0, P, nc, out, nullptr, st)
                  : launch_vs_3d(A, uni, 0, P, nc, out, nullptr, st);
}
// Maxima over the fluid cells (benc >= 0) of the per-cell pairs part[2c], part[2c+1], each
// started from 0 and taken with Julia's max (a NaN operand wins), as kah_vs_resolution takes
// them on the host; one workgroup, the result in out[0..1].
__device__ inline double vs_jmax(double a, double b) {
  return (a != a) ? a : (b != b) ? b : (b > a ? b : a);
}
__global__ __launch_bounds__(1024) void k_vs_res_max(const double* __restrict__ part,
                                                     const int32_t* __restrict__ benc, int nc,
                                                     double* __restrict__ out) {
  __shared__ double sh[2][1024];
  const int t = threadIdx.x;
  double d = 0.0, e = 0.0;
  for (int c = t; c < nc; c += 1024)
    if (benc[c] >= 0) {
      d = vs_jmax(part[2 * int64_t(c)], d);
      e = vs_jmax(part[2 * int64_t(c) + 1], e);
    }
  sh[0][t] = d;
  sh[1][t] = e;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] = vs_jmax(sh[0][t + w], sh[0][t]);
      sh[1][t] = vs_jmax(sh[1][t + w], sh[1][t]);
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sh[0][0];
    out[1] = sh[1][0];
  }
}

hipError_t launch_cell_pre_2d(const DevArgs&, const int32_t*, int, hipStream_t);
hipError_t launch_cell_pre_3d(const DevArgs&, const int32_t*, int, hipStream_t);
hipError_t launch_cell_pre(const DevArgs& A, int dim, const int32_t* cells, int n,
                           hipStream_t st) {
  return dim == 2 ? launch_cell_pre_2d(A, cells, n, st) : launch_cell_pre_3d(A, cells, n, st);
}

hipError_t launch_vs_res_max(const double* part, const int32_t* benc, int nc, double* out,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_vs_res_max, dim3(1), dim3(1024), 0, st, part, benc, nc, out);
  return hipGetLastError();
}

hipError_t launch_vs_flags(const DevArgs& A, int dim, bool uni, bool f32, const VsArgs& P,
                           int nc, uint8_t* flags, hipStream_t st) {
  if (f32) return dim == 2 ? launch_vs_2d_f32(A, uni, 1, P, nc, nullptr, flags, st)
                           : launch_vs_3d_f32(A, uni, 1, P, nc, nullptr, flags, st);
  return dim == 2 ? launch_vs_2d(A, uni, 1, P, nc, nullptr, flags, st)
                  : launch_vs_3d(A, uni, 1, P, nc, nullptr, flags, st);
}
hipError_t launch_vs_remap(const DevArgs& A, int dim, bool f32, const int64_t* voff_old,
                           const void* df_old, const void* sdf_old, const int32_t* src,
                           const uint8_t* changed, int nc, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_vs_remap_2d_f32(A, voff_old, df_old, sdf_old, src, changed, nc, st)
                    : launch_vs_remap_3d_f32(A, voff_old, df_old, sdf_old, src, changed, nc, st);
  return dim == 2 ? launch_vs_remap_2d(A, voff_old, df_old, sdf_old, src, changed, nc, st)
                  : launch_vs_remap_3d(A, voff_old, df_old, sdf_old, src, changed, nc, st);
}
hipError_t launch_vs_project(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                             int n, size_t lds, double* scratch, void* dfw, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_vs_project_2d_f32(A, uni, cells, n, lds, scratch, dfw, st)
                    : launch_vs_project_3d_f32(A, uni, cells, n, lds, scratch, dfw, st);
  return dim == 2 ? launch_vs_project_2d(A, uni, cells, n, lds, scratch, dfw, st)
                  : launch_vs_project_3d(A, uni, cells, n, lds, scratch, dfw, st);
}

hipError_t launch_overlap_map(const DevArgs& A, int dim, const OvJob* jobs, int njobs,
                              uint32_t* map, hipStream_t st) {
  return dim == 2 ? launch_overlap_map_2d(A, jobs, njobs, map, st)
                  : launch_overlap_map_3d(A, jobs, njobs, map, st);
}

hipError_t launch_macro_slope(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                              int n, double* sw, hipStream_t st) {
  if (f32)
    return dim == 2 ? launch_macro_slope_2d_f32(A, uni, cells, n, sw, st)
                    : launch_macro_slope_3d_f32(A, uni, cells, n, sw, st);
  return dim == 2 ? launch_macro_slope_2d(A, uni, cells, n, sw, st)
                  : launch_macro_slope_3d(A, uni, cells, n, sw, st);
}

}  // namespace kah

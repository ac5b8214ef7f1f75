// kah_slope_2d_f32.hip -- slope kernels (K1, K3) of the 2D2F form, FP32 form (f stored and advected in FP32).
#include "kah_device.h"

namespace kah {

hipError_t launch_slope_2d_f32(const DevArgs& A, bool uni, const int32_t* cells, int n,
                          int maxh, hipStream_t st) {
  return uni ? launch_slope_t<2, 2, true, float>(A, cells, n, maxh, st)
             : launch_slope_t<2, 2, false, float>(A, cells, n, maxh, st);
}

hipError_t launch_slope_chunk_2d_f32(const DevArgs& A, bool uni, const int2* work, int n,
                                int maxh, int vpt, int wpi, hipStream_t st) {
  return uni ? launch_slope_chunk_t<2, 2, true, float>(A, work, n, maxh, vpt, wpi, st)
             : launch_slope_chunk_t<2, 2, false, float>(A, work, n, maxh, vpt, wpi, st);
}

hipError_t launch_macro_slope_2d_f32(const DevArgs& A, bool uni, const int32_t* cells, int n,
                                double* sw, hipStream_t st) {
  return uni ? launch_macro_slope_t<2, 2, true, float>(A, cells, n, sw, st)
             : launch_macro_slope_t<2, 2, false, float>(A, cells, n, sw, st);
}

hipError_t launch_mirror_2d_f32(const DevArgs& A, bool uni, const int32_t* cells,
                              const int64_t* off, int n, int kind, int unpack, double* buf,
                              const double* mid, const int32_t* benc, double* sw, void* dfw,
                              hipStream_t st) {
  return uni ? launch_mirror_t<2, 2, true, float>(A, cells, off, n, kind, unpack, buf, mid,
                                                     benc, sw, dfw, st)
             : launch_mirror_t<2, 2, false, float>(A, cells, off, n, kind, unpack, buf, mid,
                                                      benc, sw, dfw, st);
}

hipError_t launch_restart_2d_f32(const DevArgs& A, bool uni, const int64_t* row, int64_t N,
                               int64_t Nv, int load, double* ws, int8_t* lev, double* vmid,
                               double* vdf, void* dfw, hipStream_t st) {
  return uni ? launch_restart_t<2, 2, true, float>(A, row, N, Nv, load, ws, lev, vmid, vdf,
                                                      dfw, st)
             : launch_restart_t<2, 2, false, float>(A, row, N, Nv, load, ws, lev, vmid, vdf,
                                                       dfw, st);
}

hipError_t launch_vs_2d_f32(const DevArgs& A, bool uni, int which, const VsArgs& P, int nc,
                          double* out, uint8_t* flags, hipStream_t st) {
  return uni ? launch_vs_t<2, 2, true, float>(A, which, P, nc, out, flags, st)
             : launch_vs_t<2, 2, false, float>(A, which, P, nc, out, flags, st);
}
hipError_t launch_vs_remap_2d_f32(const DevArgs& A, const int64_t* voff_old, const void* df_old,
                                const void* sdf_old, const int32_t* src, const uint8_t* changed,
                                int nc, hipStream_t st) {
  return launch_vs_remap_t<2, 2, float>(A, voff_old, df_old, sdf_old, src, changed, nc, st);
}
hipError_t launch_vs_project_2d_f32(const DevArgs& A, bool uni, const int32_t* cells, int n,
                                  size_t lds, double* scratch, void* dfw, hipStream_t st) {
  return uni ? launch_vs_project_t<2, 2, true, float>(A, cells, n, lds, scratch, dfw, st)
             : launch_vs_project_t<2, 2, false, float>(A, cells, n, lds, scratch, dfw, st);
}

hipError_t launch_ib_2d_f32(const DevArgs& A, bool uni, int ns, const int32_t* solid,
                     const int32_t* nb_off, const int32_t* nb, int nf, const IbFaceDesc* faces,
                     const int32_t* img, const int32_t* cut_idx, const double2* cut_w,
                     const double* cmid, void* df_a, void* df_b, hipStream_t st) {
  return uni ? launch_ib_t<2, 2, true, float>(A, ns, solid, nb_off, nb, nf, faces, img,
                                                 cut_idx, cut_w, cmid, df_a, df_b, st)
             : launch_ib_t<2, 2, false, float>(A, ns, solid, nb_off, nb, nf, faces, img,
                                                  cut_idx, cut_w, cmid, df_a, df_b, st);
}

}  // namespace kah

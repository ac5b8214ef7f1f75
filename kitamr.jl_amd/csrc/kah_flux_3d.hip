// kah_flux_3d.hip -- flux + relaxation kernel (K2) of the 3D1F form, FP64 storage.
#include "kah_device.h"

namespace kah {

hipError_t launch_flux_relax_3d(const DevArgs& A, bool uni, const int32_t* cells, int n,
                               int mode, int march, size_t lds, int nwv, hipStream_t st) {
  return uni ? launch_fr_d<3, 1, true, double>(A, cells, n, mode, march, lds, nwv, st)
             : launch_fr_d<3, 1, false, double>(A, cells, n, mode, march, lds, nwv, st);
}

hipError_t launch_cell_pre_3d(const DevArgs& A, const int32_t* cells, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_pre<3>, dim3((n + 255) / 256), dim3(256), 0, st, A, cells, n);
  return hipGetLastError();
}

}  // namespace kah

// kah_flux_2d_f32.hip -- flux + relaxation kernel (K2) of the 2D2F form, FP32 form (f stored and advected in FP32).
#include "kah_device.h"

namespace kah {

hipError_t launch_flux_relax_2d_f32(const DevArgs& A, bool uni, const int32_t* cells, int n,
                               int mode, int march, size_t lds, int nwv, hipStream_t st) {
  return uni ? launch_fr_d<2, 2, true, float>(A, cells, n, mode, march, lds, nwv, st)
             : launch_fr_d<2, 2, false, float>(A, cells, n, mode, march, lds, nwv, st);
}

}  // namespace kah

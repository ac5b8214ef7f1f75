// kah_kernels.h -- kernel arguments and launchers (host <-> device interface of the library)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kah_internal.h"

namespace kah {

enum { MODE_FLUX = 0, MODE_ITER = 1, MODE_FUSED = 2 };
enum { MARCH_CAIDVM = 0, MARCH_EULER = 1 };

// Everything a kernel reads, passed by value (device pointers + scalars).
struct DevArgs {
  // velocity grids
  int32_t uni_n;            // cells per grid when all grids are the same one
  int32_t lmax;             // AMR_VS_MAXLEVEL (key width)
  const int64_t* voff;      // [nc+1]
  const double* gu;         // midpoints, per-cell d-major blocks (one block if uniform)
  const double* gw;         // weights
  const int8_t* glev;       // levels
  const uint32_t* gkey;     // cell keys (adapted grids)
  // state
  const double* df;         // f of this step
  double* df_out;           // f of the next step
  double* sdf;
  double* vflux;            // VsData.flux accumulators (split flux/iterate calls)
  double* psflux;           // PsData.flux accumulators
  double* w;
  double* prim;
  double* qf;
  double* half_fw;          // [npairs*2*(D+2)]
  double* rho_w;            // [npairs]
  double* res;              // [nc*2*(D+2)] residual terms of the last iterate
  double* scratch;          // global scratch when the per-cell state exceeds LDS
  // topology
  const SlopeDesc* sl;      // [nc*D]
  const SlopeNb* slnb;
  const int32_t* half_off;  // [nc+1]
  const HalfDesc* halves;
  const int32_t* ent_off;   // [nc+1]
  const EntryDesc* ents;
  const double* rec_ds;     // [nc*D]
  const double* bc_prim;
  // parameters
  double K, gamma, Pr, mu, omega, dt;
  int32_t flux_kind;
  int32_t pad;
};

hipError_t launch_slope(const DevArgs& A, int dim, bool uni, const int32_t* cells, int n,
                        int maxh, hipStream_t st);
hipError_t launch_flux_relax(const DevArgs& A, int dim, bool uni, const int32_t* cells, int n,
                             int mode, int march, size_t lds, int waves_per_block,
                             hipStream_t st);
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);
hipError_t launch_macro_slope(const DevArgs& A, int dim, bool uni, const int32_t* cells, int n,
                              double* sw, hipStream_t st);

}  // namespace kah

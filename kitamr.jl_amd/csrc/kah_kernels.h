// kah_kernels.h -- kernel arguments and launchers (host <-> device interface of the library)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kah_internal.h"

namespace kah {

// MODE_FUSED_REG / MODE_FUSED_REGX: MODE_FUSED for regular cells (2*DIM faces to fluid cells
// on the same level, limited or plain-slope reconstruction), gathered from the stored outgoing
// micro fluxes; REG: every neighbour on the cell's velocity grid, REGX: on any grid
enum { MODE_FLUX = 0, MODE_ITER = 1, MODE_FUSED = 2, MODE_FUSED_REG = 3, MODE_FUSED_REGX = 4 };
enum { MARCH_CAIDVM = 0, MARCH_EULER = 1, MARCH_CIP = 2 };
constexpr int cpre_len(int dim) { return 3 * (dim + 2) + 8; }

// Everything a kernel reads, passed by value (device pointers + scalars).
struct DevArgs {
  // velocity grids
  int32_t uni_n;            // cells per grid when all grids are the same one
  int32_t lmax;             // AMR_VS_MAXLEVEL (key width)
  const int64_t* voff;      // [nc+1]
  const double* gu;         // midpoints, d-major blocks per distinct grid (one if uniform)
  const double* gw;         // weights
  const int8_t* glev;       // levels
  const uint32_t* gkey;     // cell keys (adapted grids)
  // state
  const double* df;         // f of this step
  double* df_out;           // f of the next step
  double* sdf;
  // outgoing micro flux of every velocity cell per direction, written by K1 for the face of
  // the cell's outgoing side (the cell's own reconstruction times u[dir]); the gather of
  // regular cells reads it instead of reconstructing (layout of sdf, storage type of f)
  double* oflux;
  double* vflux;            // VsData.flux accumulators (split flux/iterate calls)
  double* psflux;           // PsData.flux accumulators
  double* w;
  double* prim;
  double* qf;
  double* half_fw;          // partial half-flux sums: row part_off[slot]+chunk, D+2 each
  const int64_t* part_off;  // [npairs*2] first partial row of each half flux
  const int32_t* part_cnt;  // [npairs*2] number of partial rows (velocity chunks)
  double* rho_w;            // [npairs]
  double* res;              // [nc*2*(D+2)] residual terms of the last iterate
  // per-cell prologue of the fused step (k_cell_pre), cpre_len(D) doubles per cell:
  // w + dt/V*flux | prim_c | MaxPre of prim_c | tau, tau/(tau+dt), dt/(tau+dt) | dt/V;
  // null: k_flux_relax forms them itself (CIP_Marching, split calls)
  double* cpre;
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
  // read views of arrays a kernel does not write (bound to __restrict__ kernel arguments
  // so that uniform loads from them can be scalar loads)
  const double* sdf_r;
  const double* half_fw_r;
  const double* rho_w_r;
  // wave tiles of a shared tensor-product velocity grid (vt_mode 2: 8x8, 3: 4x4x4 blocks of
  // velocity cells per 64 consecutive thread positions; 0: identity order)
  const uint32_t* ovmap;    // overlap maps of adapted grids (see OvJob), or null
  const int64_t* map_base;  // first entry of each overlap map
  const int64_t* gof;       // adapted grids: first geometry entry (gu/gw/glev/gkey) of the
                            // grid of each cell (cells with identical grids share one)
  const int32_t* vt_base;   // first velocity cell of each tile
  int32_t vt_mode, vt_n0, vt_n1, vt_pad;
  // immersed boundary under CIP_Marching (positivity_preserving_ib!): the IB faces, whose
  // SolidNeighbor rows are the cells ib_base + k, and the limited boundary slope of each
  // (rows as the SolidNeighbor's df, FP64); null otherwise
  const IbFaceDesc* ib_faces;
  double* ib_ssdf;
  // parameters
  double K, gamma, Pr, mu, omega, dt;
  int32_t flux_kind;
  int32_t ib_base;
};

// Read-only arrays passed to the kernels as separate __restrict__ arguments.
#define KAH_RO_PARAMS                                                                        \
  const double* __restrict__ ro_df, const double* __restrict__ ro_gu,                        \
      const double* __restrict__ ro_gw, const int8_t* __restrict__ ro_glev,                  \
      const uint32_t* __restrict__ ro_gkey, const int64_t* __restrict__ ro_voff,             \
      const SlopeDesc* __restrict__ ro_sl, const SlopeNb* __restrict__ ro_slnb,              \
      const int32_t* __restrict__ ro_hoff, const HalfDesc* __restrict__ ro_halves,           \
      const int32_t* __restrict__ ro_eoff, const EntryDesc* __restrict__ ro_ents,            \
      const double* __restrict__ ro_bc, const double* __restrict__ ro_recds
#define KAH_RO_ARGS(A)                                                                       \
  (A).df, (A).gu, (A).gw, (A).glev, (A).gkey, (A).voff, (A).sl, (A).slnb, (A).half_off,      \
      (A).halves, (A).ent_off, (A).ents, (A).bc_prim, (A).rec_ds
#define KAH_RO_BIND(B)                                                                       \
  do {                                                                                       \
    (B).df = ro_df; (B).gu = ro_gu; (B).gw = ro_gw; (B).glev = ro_glev; (B).gkey = ro_gkey;  \
    (B).voff = ro_voff; (B).sl = ro_sl; (B).slnb = ro_slnb; (B).half_off = ro_hoff;          \
    (B).halves = ro_halves; (B).ent_off = ro_eoff; (B).ents = ro_ents; (B).bc_prim = ro_bc;  \
    (B).rec_ds = ro_recds;                                                                   \
  } while (0)

constexpr int SLOPE_MAX_VPT = 6;   // velocity cells per thread of the chunked slope kernel: 1..6
constexpr int MAX_ENTRIES = 24;    // face sides per cell (2*DIM sides, up to 2^(DIM-1) each)

// f32: f, sdf and the next f stored as float (FP32 form), else double
hipError_t launch_slope(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                        int n, int maxh, hipStream_t st);
// work = (cell, chunk) pairs, one per workgroup of wpi (3 or 4) waves; chunk = vpt*64*wpi
// velocity cells
hipError_t launch_slope_chunk(const DevArgs& A, int dim, bool uni, bool f32, const int2* work,
                              int n, int maxh, int vpt, int wpi, hipStream_t st);
hipError_t launch_flux_relax(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                             int n, int mode, int march, size_t lds, int waves_per_block,
                             hipStream_t st);
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);
// packs (unpack = 0) or unpacks the ghost-exchange buffers of the listed cells (kind 0: data,
// 1: slope), concatenated at buf + off[k]
hipError_t launch_mirror(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                         const int64_t* off, int n, int kind, int unpack, double* buf,
                         const double* mid, const int32_t* benc, double* sw, void* dfw,
                         hipStream_t st);
// velocity-space adaptation: the root lattice and the criteria parameters
struct VsArgs {
  double vmin[3], hf[3];   // velocity-domain lower corner, finest cell size
  int64_t tn[3], stride[3];
  int32_t L;               // vs_maxlevel
  int32_t mode;            // KAH_VS_*
  double cl, cg, tl;       // ADAPT_COEFFI_VS_LOCAL / _GLOBAL / _LOHNER
  double vr[2];            // Velocity_Resolution
  double wfine;            // weight of a finest velocity cell
};
hipError_t launch_vs_res(const DevArgs& A, int dim, bool uni, bool f32, const VsArgs& P, int nc,
                         double* out, hipStream_t st);
// out[0..1] = the maxima over the fluid cells of the per-cell pairs part[2c], part[2c+1]
hipError_t launch_cell_pre(const DevArgs& A, int dim, const int32_t* cells, int n,
                           hipStream_t st);
hipError_t launch_vs_res_max(const double* part, const int32_t* benc, int nc, double* out,
                             hipStream_t st);
hipError_t launch_vs_flags(const DevArgs& A, int dim, bool uni, bool f32, const VsArgs& P,
                           int nc, uint8_t* flags, hipStream_t st);
hipError_t launch_vs_remap(const DevArgs& A, int dim, bool f32, const int64_t* voff_old,
                           const void* df_old, const void* sdf_old, const int32_t* src,
                           const uint8_t* changed, int nc, hipStream_t st);
hipError_t launch_vs_project(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                             int n, size_t lds, double* scratch, void* dfw, hipStream_t st);
// gathers (load = 0) or scatters the per-rank restart arrays of all cells; row[c] is the first
// velocity row of cell c, -1 for a solid cell
hipError_t launch_restart(const DevArgs& A, int dim, bool uni, bool f32, const int64_t* row,
                          int64_t N, int64_t Nv, int load, double* ws, int8_t* lev, double* vmid,
                          double* vdf, void* dfw, hipStream_t st);
// fills the overlap maps of the given (cell, neighbour) jobs (adapted grids)
hipError_t launch_overlap_map(const DevArgs& A, int dim, const OvJob* jobs, int njobs,
                              uint32_t* map, hipStream_t st);
hipError_t launch_macro_slope(const DevArgs& A, int dim, bool uni, bool f32, const int32_t* cells,
                              int n, double* sw, hipStream_t st);
// immersed boundary (shared or adapted velocity grids): the ns solid cells, then the nf IB faces;
// both write f to the two state buffers df_a and df_b
hipError_t launch_ib(const DevArgs& A, int dim, bool uni, bool f32, int ns, const int32_t* solid,
                     const int32_t* nb_off, const int32_t* nb, int nf, const IbFaceDesc* faces,
                     const int32_t* img, const int32_t* cut_idx, const double2* cut_w,
                     const double* cmid, void* df_a, void* df_b, hipStream_t st);

}  // namespace kah
